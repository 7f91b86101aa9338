"""Counterpart of the reference's simulate.py run loop and outputs (simulate.py:199-371): a heartbeat simulation with
the lumped circulation around the 3D segment, sampled every millisecond, written in the reference's result format.

* Boundary conditions (simulate.py:262-291): every end is a pressure opening. The first takes the inlet pressure
  waveform P_in(t) of the lumped circuit (``artery.circuit``), the others relax towards R_k Q_k (``ArterySolver``,
  flow slabs inside each end). Pressures are gauge values relative to P_OUT_MMHG, stored as absolute mmHg.
* Calibration (simulate.py:323-332): on a healthy run the inlet waveform is re-fitted after every beat but the last,
  so that it draws the sourced mean inflow (240 mL/min).
* Sampling (simulate.py:305-318): every SAMPLE_S the flows (mL/min), the opening pressures, the probe pressures
  (``cs_gather``) and the maximum speed; a progress line per beat.
* Frames (simulate.py:345-352): FRAMES snapshots of (p, u) over the last beat, binned to 0.1 mm cells
  (``coarse_map``, simulate.py:199-205) and stored as float16; the binning runs on the device (``cs_bin_fields``)
  into frame buffers that are copied to the host once, at the end.
* Wall shear stress (wss.py:112-136): when the lattice carries wall points ("wall_verts", "wall_normals": the
  zero crossings of its signed distance, ``wss.wall_surface``), the fit weights are built once and ``cs_wss``
  evaluates tau at every stored frame on the solver's live velocity, accumulating TAWSS and OSI on the device;
  ``wss.npz`` gets the reference's keys (verts, tau, tawss, osi, wall, t) except the mesh faces, which belong to
  the renderer.
* Outputs (simulate.py:356-371): ``series.npz`` (t, Q, p_open, p_probe, umax), ``meta.json`` and ``fields.npz``
  (coords int16, origin, dx_mm, wall_mm, t, p_mmhg, u_m_s), the files wss.py and the explorer read, and
  ``wss.npz``.

The geometry comes from the caller (a lattice with ``wall_mm``, its open ends and probes): ``vessel.artery_case``
builds the story artery as the reference's setup does (simulate.py:155-200); ``geometry.artery_y`` is a smaller
synthetic Y artery.
``heart_period`` may be shortened for tests; the reference's is 1 s.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

from . import _native
from .artery import (HEART_PERIOD, MMHG, MU, OUTLET_RELAX_S, P_OUT_MMHG, RHO, SAMPLE_S, SPONGE_MM, SPONGE_TAU,
                     ArterySolver, circuit, flux_slabs, outlet_resistances, units)

BEATS = 3                 # simulate.py:45
FRAMES = 40               # simulate.py:53


def coarse_map(lat, dx_mm):
    """Cells -> 0.1 mm bins (identity at 0.1 mm; 2x2x2 averages at 0.05 mm) for the stored fields
    (simulate.py:199-205): (bin coordinates int16, bin of each cell, cells per bin, factor)."""
    f = int(round(0.1 / dx_mm))
    key = lat["coords"] // max(f, 1)
    uk, inv = np.unique(key, axis=0, return_inverse=True)
    cnt = np.bincount(inv.ravel())
    return uk.astype(np.int16), inv.ravel(), cnt, f


class FrameBinner:
    """Device binning of the newest fields into float16 frame buffers (FRAMES x bins), one launch per frame."""

    def __init__(self, s, cmap, n_frames, p_scale, p_offset, u_scale):
        import torch
        uk, inv, cnt, _ = cmap
        order = np.argsort(inv, kind="stable")
        start = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
        self.s = s
        with torch.cuda.stream(s._tstream):
            self.start = torch.from_numpy(start).to(s.dev)
            self.slots = torch.from_numpy(np.ascontiguousarray(s.backend.slots(order), np.int32)).to(s.dev)
            self.p = torch.zeros((n_frames, len(uk)), dtype=torch.float16, device=s.dev)
            self.u = torch.zeros((n_frames, len(uk), 3), dtype=torch.float16, device=s.dev)
        s.synchronize()
        self.n_bins = len(uk)
        self.scales = (float(p_scale), float(p_offset), float(u_scale))

    def frame(self, k):
        s = self.s
        A = _native.BinArgs()
        A.start, A.slots, A.n_bins = self.start.data_ptr(), self.slots.data_ptr(), self.n_bins
        A.rho, A.rho_elem = s.backend.density(s.cur).data_ptr(), s.backend.density_stride()
        A.vel = s.backend.velocity(s.cur).data_ptr()
        A.vel_elem, A.vel_comp = s.backend.velocity_addressing()
        A.p_scale, A.p_offset, A.u_scale = self.scales
        A.p_out, A.u_out = self.p[k].data_ptr(), self.u[k].data_ptr()
        _native.call("cs_bin_fields", ctypes.byref(A), s.stream_handle)

    def host(self, n):
        self.s.synchronize()
        return self.p[:n].cpu().numpy(), self.u[:n].cpu().numpy()


def wall_shear_setup(s, lat, dx_mm, u_scale, n_frames):
    """Fit weights of the lattice's wall points (``wss.weights`` over the lattice cells, depth = wall distance) and
    the device evaluator; None without wall points."""
    import torch

    from . import wss
    if "wall_verts" not in lat or not len(lat["wall_verts"]):
        return None
    cells = lat["origin"] + lat["coords"] * dx_mm
    idx, c, quad, any_ = wss.weights(lat["wall_verts"], lat["wall_normals"], cells, lat["wall_mm"])
    ws = wss.WallShear(s, idx, c, lat["wall_normals"], u_scale=u_scale, mu=MU)
    with torch.cuda.stream(s._tstream):
        ws.store = torch.zeros((n_frames, len(idx), 3), dtype=torch.float16, device=s.dev)
    s.synchronize()
    ws.any = any_
    ws.quad = quad
    return ws


def probe_cells(lat, dx_mm, probes):
    """Cells within each probe's radius (simulate.py:193-199): ({"idx", "slot"}, probes with their cell counts)."""
    P = lat["origin"] + lat["coords"] * dx_mm
    idx, slot = [], []
    for k, pr in enumerate(probes):
        m = np.flatnonzero(np.linalg.norm(P - np.asarray(pr["point"], float), axis=1) < pr["radius"])
        pr["cells"] = int(len(m))
        idx.append(m)
        slot.append(np.full(len(m), k))
    return {"idx": np.concatenate(idx).astype(np.int32), "slot": np.concatenate(slot).astype(np.int32)}


def run(lat, ends, probes, dx, dt, out_dir, case="case", beats=BEATS, cal=None, flow=1.0, heart_period=HEART_PERIOD,
        log=print, progress_every=None, autotune=False, wall_shear=True, **solver_kw):
    """Simulate `beats` heartbeats and write series.npz, meta.json, fields.npz (and wss.npz when the lattice has wall
    points and wall_shear is set) into out_dir. cal: the circuit
    (``artery.circuit``) to drive with; None calibrates on this run, as the reference's healthy case does. flow
    scales the absolute inlet pressure. progress_every: also log the flows every that many samples. autotune: time
    the launch variants on this lattice first and keep the fastest (the results do not depend on it). Returns
    (series, meta)."""
    import torch
    os.makedirs(out_dir, exist_ok=True)
    un = units(dx, dt)
    dx_mm = dx * 1e3
    calibrate = cal is None
    if calibrate:
        cal = circuit(0.0)
    q_si = un.dx ** 3 / un.dt                      # flow: SI per lattice unit
    p_si = RHO * un.c ** 2                         # pressure: SI per lattice unit
    R_lu = [r * q_si / p_si for r in outlet_resistances(cal, ends)]
    pref_lu = P_OUT_MMHG * MMHG / p_si
    n_open = len(ends)
    flux = flux_slabs(lat, ends, dx_mm)
    s = ArterySolver(lat, un.nu_lu, flux, R_lu, pref_lu, un.dt / OUTLET_RELAX_S, model="regularised", **solver_kw)
    s.set_outlet_pressure([0.0] + [-pref_lu] * (n_open - 1))
    s.set_sponge(SPONGE_MM / dx_mm, SPONGE_TAU, kinds=("outlet",))
    spb = int(round(heart_period / un.dt))
    every = int(round(SAMPLE_S / un.dt))
    ramp = int(0.1 * heart_period / un.dt)
    frame_every = spb // FRAMES
    if spb % every or frame_every % every:
        raise ValueError(f"{case}: {spb} steps per beat / {frame_every} per frame are not multiples of {every}")

    def drive(c_):
        c = np.array(c_["P_in_coef"]) / p_si * flow
        c[0] -= pref_lu                            # relative to the venous pressure
        s.set_drive(c, period=spb, ramp=ramp)

    drive(cal)
    if autotune:
        tune = s.autotune()
        log(f"{case}: launch variant {s.variant} (ms per step: "
            + ", ".join(f"{k} {v:.4f}" for k, v in tune.items()) + ")")
    gather = probe_cells(lat, dx_mm, probes)
    pcount = np.bincount(gather["slot"], minlength=len(probes))
    cmap = coarse_map(lat, dx_mm)
    binner = FrameBinner(s, cmap, FRAMES, RHO * un.c ** 2 / (3.0 * MMHG), P_OUT_MMHG, un.c)
    ws = wall_shear_setup(s, lat, dx_mm, un.c, FRAMES) if wall_shear else None
    series = {"t": [], "Q": [], "p_open": [], "p_probe": [], "umax": []}
    frames_t = []
    T0 = time.perf_counter()
    total = beats * spb
    while s.steps < total:
        s.run(every, chunk=every)
        g_sum, umax = s.gather(gather["idx"], gather["slot"], len(probes))
        Q = s.flows()[:n_open] * q_si * 6e7                    # mL/min
        series["t"].append(s.steps * un.dt)
        series["Q"].append(Q)
        series["p_open"].append(un.p(s.opening_pressures()[:n_open]) / MMHG + P_OUT_MMHG)
        pp = un.p(g_sum / np.maximum(pcount, 1)) / MMHG + P_OUT_MMHG
        pp[pcount == 0] = np.nan
        series["p_probe"].append(pp)
        series["umax"].append(float(un.u(umax)))
        if progress_every and len(series["t"]) % progress_every == 0:
            log(f"{case}: t {s.steps * un.dt:.3f} s ({time.perf_counter() - T0:.1f} s): flows "
                + ", ".join("%.1f" % q for q in Q) + " mL/min")
        if s.steps % spb == 0:
            beat = s.steps // spb
            if not s.ok():
                raise RuntimeError(f"{case}: solver diverged at step {s.steps}")
            tt = np.array(series["t"])
            m = tt > tt[-1] - heart_period
            Qm = np.array(series["Q"])[m].mean(0)
            flows = ", ".join("%s %.1f" % (e["label"].split()[0], q) for e, q in zip(ends, Qm))
            log(f"{case}: beat {beat} done ({time.perf_counter() - T0:.1f} s): mean flow {flows} mL/min")
            if calibrate and beat < beats:
                # re-fit the segment's resistance from this beat: mean inlet pressure over mean inflow, less the
                # outlets' equivalent resistance
                p_in = np.array(series["p_open"])[m][:, 0].mean() * MMHG
                R_tot = p_in / (max(Qm[0], 1e-9) * 1e-6 / 60)
                cal = circuit(float(max(R_tot - cal["R_eq"], 0.0)))
                drive(cal)
        if s.steps > (beats - 1) * spb and (s.steps - (beats - 1) * spb) % frame_every == 0:
            k = len(frames_t)
            binner.frame(k)
            if ws is not None:
                tau = ws.frame(accumulate=True)
                with torch.cuda.stream(s._tstream):      # ordered before the next frame's kernel rewrites tau
                    ws.store[k].copy_(tau)
            frames_t.append((s.steps - (beats - 1) * spb) * un.dt)
    secs = time.perf_counter() - T0
    ser = {k: np.array(v) for k, v in series.items()}
    np.savez_compressed(os.path.join(out_dir, "series.npz"), **ser)
    meta = {"case": case, "done": True, "dx_mm": dx_mm, "dt_us": un.dt * 1e6, "tau": un.tau, "beats": beats,
            "steps_per_beat": spb, "cells": int(lat["nbr"].shape[1]), "links": lat["counts"], "secs": secs,
            "p_out_mmhg": P_OUT_MMHG, "calibration": cal,
            "open_ends": [{"label": e["label"], "centre": np.asarray(e["centre"]).tolist(),
                           "normal": np.asarray(e["normal"]).tolist(), "radius": e["radius"]} for e in ends],
            "probes": [{"name": p["name"], "point": np.asarray(p["point"]).tolist(), "radius": p["radius"],
                        "cells": p["cells"]} for p in probes]}
    with open(os.path.join(out_dir, "meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1, default=float)
    uk, inv, cnt, f = cmap
    wall = np.bincount(inv, weights=lat["wall_mm"], minlength=len(uk)) / cnt
    frames_p, frames_u = binner.host(len(frames_t))
    np.savez_compressed(os.path.join(out_dir, "fields.npz"), coords=uk, origin=lat["origin"],
                        dx_mm=0.1 if f > 1 else lat["dx"], wall_mm=wall.astype(np.float32), t=np.array(frames_t),
                        p_mmhg=frames_p, u_m_s=frames_u)
    if ws is not None:
        # vertices with no lumen cell in reach carry no wall (wss.py:129-134; the open ends' faces were dropped
        # with the points)
        on = np.asarray(ws.any, bool)
        tau = ws.store[:len(frames_t)].cpu().numpy()
        tawss, osi = ws.tawss(), ws.osi()
        tau[:, ~on], tawss[~on], osi[~on] = 0, 0, 0
        np.savez_compressed(os.path.join(out_dir, "wss.npz"), verts=lat["wall_verts"].astype(np.float32),
                            tau=tau.astype(np.float16), tawss=tawss.astype(np.float32),
                            osi=osi.astype(np.float32), wall=on, t=np.array(frames_t))
        meta["wss"] = {"vertices": int(on.sum()), "quadratic_fit_pct": round(100 * float(ws.quad[on].mean()), 1)
                       if on.any() else 0.0, "tawss_mean_pa": float(tawss[on].mean()) if on.any() else 0.0}
        with open(os.path.join(out_dir, "meta.json"), "w") as fh:
            json.dump(meta, fh, indent=1, default=float)
    log(f"{case}: done in {secs:.1f} s ({int(lat['nbr'].shape[1])} cells, {s.steps} steps)")
    return ser, meta


def main(argv=None):
    """Command line of the reference's simulate.py for the story artery:

        python -m cerebrosim_amd.simulate CASE --segment segment.json [--out DIR] [--beats N] [--cal meta.json]

    The healthy case calibrates its inlet; other cases take the calibration from a healthy run's meta.json
    (--cal), as simulate.py:323-332 does. --heart-period shortens the beat (seconds) for quick runs."""
    import argparse

    from .vessel import GRIDS, artery_case
    ap = argparse.ArgumentParser(description="simulate the story artery over heartbeats")
    ap.add_argument("case", choices=sorted(GRIDS))
    ap.add_argument("--segment", required=True, help="the story's segment file")
    ap.add_argument("--out", default=None, help="output folder (default results/CASE)")
    ap.add_argument("--beats", type=int, default=BEATS)
    ap.add_argument("--cal", default=None, help="meta.json of a healthy run (its calibration drives this case)")
    ap.add_argument("--heart-period", type=float, default=HEART_PERIOD)
    ap.add_argument("--storage", default="fp32", choices=("fp32", "fp16"))
    ap.add_argument("--autotune", action="store_true", help="pick the fastest launch variant on this lattice first")
    args = ap.parse_args(argv)
    cal = None
    if args.cal:
        with open(args.cal) as fh:
            cal = json.load(fh)["calibration"]
    elif args.case != "healthy":
        raise SystemExit(f"{args.case}: pass --cal with a healthy run's meta.json")
    lat, ends, probes = artery_case(args.case, segment=args.segment)
    dx, dt = GRIDS[args.case]
    dt = SAMPLE_S / round(SAMPLE_S / dt)
    ser, meta = run(lat, ends, probes, dx, dt, args.out or os.path.join("results", args.case), case=args.case,
                    beats=args.beats, cal=cal, heart_period=args.heart_period, progress_every=100,
                    storage=args.storage, autotune=args.autotune)
    last = ser["t"] > ser["t"][-1] - args.heart_period
    print(f"{args.case}: last beat mean flows " + ", ".join("%.1f" % q for q in ser["Q"][last].mean(0)) + " mL/min")


if __name__ == "__main__":
    main()
