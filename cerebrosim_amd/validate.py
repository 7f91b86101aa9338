"""Counterpart of the reference's validate.py, tests 1-4: the solver's known-answer runs at the reference's operating
point, summarised in a solver.json-shaped file with the reference's published values beside each result.

    python -m cerebrosim_amd.validate [--out solver.json] [--layout auto|bricks|indexed] [--storage fp32|fp16]
                                      [--quick] [--tests poiseuille,womersley,open_pipe,stenosis,artery]
                                      [--segment segment.json]

1. Poiseuille: periodic pipes D = 10 / 20 / 40 with interpolated (Bouzidi) walls and D = 20 half-way, driven by a
   body force for 14 decay times of the slowest mode (validate.py:59-83).
2. Womersley: oscillating body force in the D = 20 pipe at Wo 1.4 / 2.8 / 5.6; L2 over the third cycle against the
   Bessel profile, sampled at t = steps - 1/2 (validate.py:86-117).
3. Open pipes, straight and tilted (1, 2, 2): velocity inlet, pressure outlet, sponge; pressure gradient against
   Hagen-Poiseuille (validate.py:125-215).
4. The 70 % cosine stenosis at narrow70 Re 100 / 200 and narrow70_fine Re 200: pressure drop against Young & Tsai
   (validate.py:218-291).

5. The artery (healthy, 0.1 mm; ``vessel.artery_lattice`` from the story's segment file): lattice build, 1000
   warm-up steps, 20000 timed steps, speed, stability and the maximum speed; then the time per heartbeat at the
   three operating points (validate.py:294-342, 368-375). Runs when a segment file is given.

The periodic pipes are 4 nodes long instead of 2: the flow does not vary along the axis, so the per-cell values are
the same, and the brick layout can wrap a multiple of 4.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import lbm
from .geometry import open_pipe, periodic_pipe
from .validation import (MU_BLOOD, NU_BLOOD, OPERATING, RHO_BLOOD, SPONGE_TAU, l2, open_pipe_case, poiseuille_exact,
                         run_steady, section_profile, units, womersley_exact)

# The reference's published results (geometry/solver.json), for the side-by-side comparison.
PUBLISHED = {
    "poiseuille": {(10, True): (0.13532128869870114, -0.14670134092831366),       # solver.json:20-27
                   (20, True): (0.03429092846606971, -0.03803448455849523),       # :28-34
                   (40, True): (0.009187449059085594, -0.010294269016544288),     # :35-41
                   (20, False): (0.20496644204979705, -0.22679752766633754)},     # :42-48
    "womersley": {1.4: (194172, 0.03370675662345126), 2.8: (48540, 0.03302506896494258),   # :50-68
                  5.6: (12132, 0.060357378395948014)},
    "open_pipe": {"straight": (0.08768536633211488, 0.01247263825422971, -0.001167730051564786),   # :70-90
                  "tilted (1,2,2)": (0.058565768905948445, 0.02435777993182073, -0.002742054870268462)},  # :92-112
    "stenosis": {("narrow70", 100): (3556.4159733738034, -0.24349248472805163),   # :114-137
                 ("narrow70", 200): (12536.677903972353, -0.0975507296228646),    # :139-157
                 ("narrow70_fine", 200): (8247.892698620224, -0.4062777392039203)},  # :159-177
}


def young_tsai(U0, D0, D1, Ls):
    """Pressure drop (Pa) across an axisymmetric stenosis in steady flow, Young & Tsai (1973), with the viscous
    coefficient in the Seeley & Young form Kv = 32 (0.83 Ls + 1.64 D1) / D0 (A0/A1)^2 and Kt = 1.52
    (validate.py:218-223)."""
    ar = (D0 / D1) ** 2
    Kv = 32 * (0.83 * Ls + 1.64 * D1) / D0 * ar ** 2
    return Kv * MU_BLOOD * U0 / D0 + 1.52 / 2 * RHO_BLOOD * (ar - 1) ** 2 * U0 ** 2


def _solver(lat, nu, layout, storage, **kw):
    return lbm.Solver(lat, nu, model="regularised", layout=layout, storage=storage, **kw)


def poiseuille(nu, layout="auto", storage="fp32", cases=((10, True), (20, True), (40, True), (20, False))):
    """Test 1 (validate.py:69-83)."""
    out = []
    for D, bz in cases:
        lat, r, R = periodic_pipe(D, nz=4)
        s = _solver(lat, nu, layout, storage, bouzidi=bz)
        F = 4 * nu * 0.02 / R ** 2                       # centreline 0.02 lattice units
        s.set_force([0, 0, F])
        s.run(int(14 * R * R / (5.78 * nu)))             # 14 decay times of the slowest mode
        _, u = s.fields()
        ua = poiseuille_exact(r, R, F, nu)
        res = {"D": D, "bouzidi": bz, "l2": l2(u[:, 2], ua), "flow_err": float(u[:, 2].sum() / ua.sum() - 1),
               "steps": s.steps, "ok": s.ok()}
        ref = PUBLISHED["poiseuille"].get((D, bz))
        if ref:
            res["published"] = {"l2": ref[0], "flow_err": ref[1]}
        out.append(res)
    return out


def womersley(nu, layout="auto", storage="fp32", wos=(1.4, 2.8, 5.6), D=20, phases=12):
    """Test 2 (validate.py:92-117)."""
    out = []
    lat, r, R = periodic_pipe(D, nz=4)
    for Wo_nominal in wos:
        T = int(round(2 * math.pi * R * R / (Wo_nominal * Wo_nominal * nu) / phases)) * phases
        w = 2 * math.pi / T
        Wo = R * math.sqrt(w / nu)
        rr = np.linspace(0, R, 50)
        peak = max(np.abs(womersley_exact(rr, R, Wo, 1.0, w, t)).max() for t in np.linspace(0, T, 24))
        F0 = 0.02 / peak
        s = _solver(lat, nu, layout, storage)
        s.set_force([0, 0, F0])
        s.set_drive([0.0, 1.0, 0.0], period=T)
        s.run(2 * T)
        num = den = 0.0
        for _ in range(phases):
            s.run(T // phases)
            _, u = s.fields()
            ua = womersley_exact(r, R, Wo, F0, w, s.steps - 0.5)
            num += ((u[:, 2] - ua) ** 2).sum()
            den += (ua ** 2).sum()
        res = {"Wo": Wo, "period_steps": T, "l2": float(math.sqrt(num / den)), "ok": s.ok()}
        ref = PUBLISHED["womersley"].get(Wo_nominal)
        if ref:
            res["published"] = {"period_steps": ref[0], "l2": ref[1]}
        out.append(res)
    return out


def open_pipes(layout="auto", storage="fp32"):
    """Test 3 (validate.py:179-215): straight and tilted (1, 2, 2) pipes, 20 cells across, 12 diameters long."""
    out = []
    for name, axis in (("straight", (0, 0, 1)), ("tilted (1,2,2)", (1, 2, 2))):
        res = open_pipe_case(lambda lat, nu: _solver(lat, nu, layout, storage), axis)
        res = {"name": name, **{k: v for k, v in res.items() if k != "counts"},
               "counts": {k: np.asarray(v).tolist() for k, v in res["counts"].items()}}
        g, p, f = PUBLISHED["open_pipe"][name]
        res["published"] = {"grad_err": g, "profile_l2": p, "flow_vs_inlet": f}
        out.append(res)
    return out


def stenosis(key, Re0, layout="auto", storage="fp32", sev=0.7, D0_mm=2.0, Ls_mm=4.0, up_d=4, down_d=18):
    """Test 4 (validate.py:226-291): pressure drop across the 70 % cosine narrowing, steady inflow; the mean over
    the last flow-through (4 chunks) when the jet is unsteady."""
    un = units(key)
    D = D0_mm * 1e-3 / un.dx
    Ls = Ls_mm * 1e-3 / un.dx
    s0 = up_d * D
    L = (up_d + down_d) * D + Ls
    R0 = D / 2

    def radius(sv):
        sv = np.asarray(sv, float)
        x = np.clip((sv - s0) / Ls, 0, 1)
        return R0 * (1 - sev * 0.5 * (1 - np.cos(2 * math.pi * x)))

    lat, sc, rc, ax = open_pipe((0, 0, 1), L, radius)
    U0_phys = Re0 * NU_BLOOD / (D0_mm * 1e-3)
    U = float(un.u_lu(U0_phys))
    s = _solver(lat, un.nu_lu, layout, storage)
    s.set_sponge(D, SPONGE_TAU)
    s.set_drive([U], ramp=int(0.3 * L / U))
    s.set_outlet_pressure([0.0])
    sc_mid = s0 + Ls / 2

    def measure():
        p, u = s.fields()
        sm, pm = section_profile(sc, p, L)
        mu = (sm > 1 * D) & (sm < s0 - 0.5 * D)
        md = (sm > L - 6 * D) & (sm < L - D)
        fu, fd = np.polyfit(sm[mu], pm[mu], 1), np.polyfit(sm[md], pm[md], 1)
        return float(np.polyval(fu, sc_mid) - np.polyval(fd, sc_mid)), u

    t0 = time.perf_counter()
    hist, ok = run_steady(s, measure, chunk=int(0.25 * L / U) // 200 * 200, min_chunks=8, max_chunks=24, tol=2e-3)
    s.synchronize()
    secs = time.perf_counter() - t0
    dp_lu, u = hist[-1]
    spread = float("nan")
    if len(hist) >= 4:                                   # unsteady jets: average the last flow-through
        last = [h[0] for h in hist[-4:]]
        dp_lu = float(np.mean(last))
        spread = float(np.ptp(last) / abs(dp_lu))
    dp = float(un.p(dp_lu))
    yt = young_tsai(U0_phys, D0_mm * 1e-3, D0_mm * 1e-3 * (1 - sev), Ls_mm * 1e-3)
    N = lat["nbr"].shape[1]
    res = {"key": key, "Re0": Re0, "U0": U0_phys, "D_cells": D, "cells": N, "steps": s.steps, "ok": ok,
           "dp_Pa": dp, "dp_mmHg": dp / 133.322, "young_tsai_Pa": yt, "vs_young_tsai": dp / yt - 1,
           "spread_last_flow_through": spread, "u_throat_max": float(un.u(np.abs(u[:, 2]).max())),
           "u_lu_max": float(np.abs(u).max()), "tau": un.tau, "secs": secs, "mlups": N * s.steps / secs / 1e6}
    ref = PUBLISHED["stenosis"].get((key, Re0))
    if ref:
        res["published"] = {"dp_Pa": ref[0], "vs_young_tsai": ref[1]}
    return res


SPONGE_CELLS = 20                  # validate.py:41
PUBLISHED_ARTERY = {"cells": 620028, "steps": 21000, "mlups": 384.8784235294374,            # solver.json:179-194
                    "u_max_m_s": 1.4885234832763672,
                    "counts": {"wall": 747143, "inlet": 11400, "outlet": [2649, 2682]}}


def artery(segment, layout="auto", storage="fp32", steps=20000, warmup=1000):
    """Test 5 (validate.py:311-340): the healthy artery at 0.1 mm with a steady velocity inlet (the sourced mean
    inflow, ramped over 10000 steps), zero outlet pressures and 20-cell sponges; speed over `steps` after `warmup`."""
    from .artery import Q_INLET_MEAN
    from .vessel import artery_lattice
    un = units("healthy")
    t0 = time.perf_counter()
    lat, g = artery_lattice("healthy", un.dx * 1e3, segment=segment)
    build_s = time.perf_counter() - t0
    R_in = lat["openings"][0]["radius"] * 1e-3
    U_mean = Q_INLET_MEAN * 1e-6 / 60 / (math.pi * R_in ** 2)
    s = _solver(lat, un.nu_lu, layout, storage)
    s.set_sponge(SPONGE_CELLS, SPONGE_TAU)
    s.set_drive([float(un.u_lu(U_mean))], ramp=10000)
    s.set_outlet_pressure([0.0] * lat["n_outlets"])
    s.run(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    s.synchronize()
    secs = time.perf_counter() - t0
    _, u = s.fields()
    N = lat["nbr"].shape[1]
    res = {"cells": N, "build_s": build_s, "steps": s.steps, "secs": secs, "mlups": N * steps / secs / 1e6,
           "ok": s.ok(), "u_max_m_s": float(un.u(np.linalg.norm(u, axis=1).max())), "counts": lat["counts"],
           "openings": [{k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in o.items()}
                        for o in lat["openings"]], "U_mean_inlet": U_mean, "published": PUBLISHED_ARTERY}
    beat = {}
    for key in ("healthy", "narrow70", "narrow70_fine"):       # validate.py:368-375
        uk = units(key)
        cells = N * (1e-4 / uk.dx) ** 3
        n = 1.0 / uk.dt                                         # steps per 1 s heartbeat
        beat[key] = {"dx_mm": uk.dx * 1e3, "dt_us": uk.dt * 1e6, "tau": uk.tau, "cells_M": cells / 1e6,
                     "steps_per_beat": n, "minutes_per_beat": cells * n / (res["mlups"] * 1e6) / 60}
    return res, beat


TESTS = ("poiseuille", "womersley", "open_pipe", "stenosis", "artery")


def run(tests=TESTS, layout="auto", storage="fp32", quick=False, log=print, segment=None):
    """Run the selected tests; returns the solver.json-shaped summary. quick: Poiseuille D = 10 and 20 (half-way)
    and Womersley at Wo 5.6 only. The artery runs when `segment` (the story's segment file) is given."""
    unh = units("healthy")
    T0 = time.perf_counter()
    out = {"solver": f"cerebrosim_amd: sparse D3Q27 lattice Boltzmann (HIP, gfx950), regularised, Galilean "
                     f"correction, Bouzidi walls, equilibrium velocity inlet, pressure outlets, sponge tau "
                     f"{SPONGE_TAU}; layout {layout}, storage {storage}",
           "operating": {k: {"dx_mm": dx * 1e3, "dt_us": dt * 1e6, "tau": units(k).tau}
                         for k, (dx, dt) in OPERATING.items()}}
    if "poiseuille" in tests:
        cases = ((10, True), (20, False)) if quick else ((10, True), (20, True), (40, True), (20, False))
        out["poiseuille"] = poiseuille(unh.nu_lu, layout, storage, cases)
        for r in out["poiseuille"]:
            log(f"Poiseuille D={r['D']} {'interpolated' if r['bouzidi'] else 'half-way'}: L2 {r['l2']:.6f} "
                f"(published {r.get('published', {}).get('l2', float('nan')):.6f}), flow {r['flow_err']:+.6f}")
    if "womersley" in tests:
        out["womersley"] = womersley(unh.nu_lu, layout, storage, wos=(5.6,) if quick else (1.4, 2.8, 5.6))
        for r in out["womersley"]:
            log(f"Womersley {r['Wo']:.2f} (period {r['period_steps']}): L2 {r['l2']:.6f} "
                f"(published {r.get('published', {}).get('l2', float('nan')):.6f})")
    if "open_pipe" in tests and not quick:
        out["open_pipe"] = open_pipes(layout, storage)
        for r in out["open_pipe"]:
            log(f"open pipe {r['name']}: grad_err {r['grad_err']:+.5f} (published {r['published']['grad_err']:+.5f}), "
                f"profile L2 {r['profile_l2']:.5f}, flow vs inlet {r['flow_vs_inlet']:+.5f}")
    if "stenosis" in tests and not quick:
        out["stenosis"] = [stenosis(k, re, layout, storage)
                           for k, re in (("narrow70", 100), ("narrow70", 200), ("narrow70_fine", 200))]
        for r in out["stenosis"]:
            log(f"stenosis {r['key']} Re0 {r['Re0']}: dp {r['dp_Pa']:.0f} Pa (published "
                f"{r['published']['dp_Pa']:.0f}), vs Young-Tsai {r['vs_young_tsai']:+.3f}, spread "
                f"{r['spread_last_flow_through']:.2e}, {r['steps']} steps, {r['mlups']:.0f} MLUPS")
    if "artery" in tests and segment is not None and not quick:
        out["artery"], out["per_heartbeat"] = artery(segment, layout, storage)
        r = out["artery"]
        log(f"artery healthy 0.1 mm: {r['cells'] / 1e6:.2f} M cells (published {PUBLISHED_ARTERY['cells']}), "
            f"{r['mlups']:.0f} MLUPS, ok {r['ok']}, max speed {r['u_max_m_s']:.3f} m/s (published "
            f"{PUBLISHED_ARTERY['u_max_m_s']:.3f}); links {r['counts']}")
    out["total_s"] = time.perf_counter() - T0
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="solver.json")
    ap.add_argument("--layout", default="auto")
    ap.add_argument("--storage", default="fp32", choices=("fp32", "fp16"))
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--tests", default=",".join(TESTS))
    ap.add_argument("--segment", default=None, help="the story's segment file (enables test 5, the artery)")
    args = ap.parse_args(argv)
    res = run(tuple(t for t in args.tests.split(",") if t), args.layout, args.storage, args.quick,
              segment=args.segment)
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1, default=float)
    print(f"wrote {args.out} ({res['total_s']:.1f} s)")


if __name__ == "__main__":
    main()
