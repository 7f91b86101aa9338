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
  (``coarse_map``, simulate.py:199-205) and stored as float16.
* Outputs (simulate.py:356-371): ``series.npz`` (t, Q, p_open, p_probe, umax), ``meta.json`` and ``fields.npz``
  (coords int16, origin, dx_mm, wall_mm, t, p_mmhg, u_m_s), the files wss.py and the explorer read.

The geometry comes from the caller (a lattice with ``wall_mm``, its open ends and probes): ``vessel.artery_case``
builds the story artery as the reference's setup does (simulate.py:155-200); ``geometry.artery_y`` is a smaller
synthetic Y artery.
``heart_period`` may be shortened for tests; the reference's is 1 s.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from .artery import (HEART_PERIOD, MMHG, OUTLET_RELAX_S, P_OUT_MMHG, RHO, SAMPLE_S, SPONGE_MM, SPONGE_TAU,
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
        log=print, progress_every=None, autotune=False, **solver_kw):
    """Simulate `beats` heartbeats and write series.npz, meta.json and fields.npz into out_dir. cal: the circuit
    (``artery.circuit``) to drive with; None calibrates on this run, as the reference's healthy case does. flow
    scales the absolute inlet pressure. progress_every: also log the flows every that many samples. autotune: time
    the launch variants on this lattice first and keep the fastest (the results do not depend on it). Returns
    (series, meta)."""
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
    series = {"t": [], "Q": [], "p_open": [], "p_probe": [], "umax": []}
    frames_t, frames_p, frames_u = [], [], []
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
            p, u = s.fields()
            uk, inv, cnt, _ = cmap
            pc = np.bincount(inv, weights=p, minlength=len(uk)) / cnt
            uc = np.stack([np.bincount(inv, weights=u[:, k], minlength=len(uk)) / cnt for k in range(3)], 1)
            frames_t.append((s.steps - (beats - 1) * spb) * un.dt)
            frames_p.append((un.p(pc) / MMHG + P_OUT_MMHG).astype(np.float16))
            frames_u.append(un.u(uc).astype(np.float16))
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
    np.savez_compressed(os.path.join(out_dir, "fields.npz"), coords=uk, origin=lat["origin"],
                        dx_mm=0.1 if f > 1 else lat["dx"], wall_mm=wall.astype(np.float32), t=np.array(frames_t),
                        p_mmhg=np.array(frames_p), u_m_s=np.array(frames_u))
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
