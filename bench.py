"""Benchmark of the solver step on one MI355X: MLUPS and HBM-roofline fraction on the pulsatile Womersley pipe.

Workload: radius 40, length 480 pipe along z (about 2.41 M fluid cells), plug velocity inlet
u(t) = 0.02 (1 + 0.5 sin(2 pi t / 4000)), pressure outlet at rho = 1, tau = 0.56, regularised BGK with the Galilean
correction, Bouzidi walls, FP32 arithmetic. A "step" is one solver step over all fluid cells (clock + step kernels).
``--storage fp16`` stores the populations as float16 equilibrium deviations (140 B per fluid cell: 27 populations
read + written x 2 B, previous velocity and density read 16 B, velocity + density written 16 B).

Prints one JSON line. ``roofline.achieved`` = algorithmic bytes of the step kernel per launch (244 B per fluid cell:
27 populations read + 27 written x 4 B, previous velocity read 12 B, velocity + density written 16 B) / the kernel's
mean duration measured with HIP events on the solver's stream. ``cpu_baseline`` times the oracle (the C restatement
of the reference step, FP32, all host threads) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_LUP_FP32 = 244
BYTES_PER_LUP = {"fp32": BYTES_PER_LUP_FP32, "fp16": 140}
HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & % of HBM peak (sparse D3Q27); cases/hour + p99 solve latency"


def build_workload():
    from cerebrosim_amd.geometry import womersley_pipe
    lat, sc, rc = womersley_pipe(40.0, 480)
    return lat


def make_solver(lat, layout, storage="fp32"):
    from cerebrosim_amd import lbm
    tau = 0.56
    s = lbm.Solver(lat, (tau - 0.5) / 3, model="regularised", layout=layout, storage=storage)
    s.set_drive([0.02, 0.0, 0.01], period=4000)
    s.set_outlet_pressure([0.0])
    return s


def kernel_times_ms(s, steps=40):
    """Mean duration of each launch of a step (HIP events recorded on the solver's stream around each launch of
    eager steps; the step's launches run back to back as in the graph)."""
    import torch
    st = s._tstream
    parts = s.backend.launch_parts()
    ev = {name: [] for name, _ in parts}
    for _ in range(steps):
        s._clock()
        for k, (name, _) in enumerate(parts):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            s.backend.step_part(s, s.cur, k == 1)
            e1.record(st)
            ev[name].append((e0, e1))
        s.cur = 1 - s.cur
        s.steps += 1
    st.synchronize()
    return {name: (float(np.mean([a.elapsed_time(b) for a, b in ev[name]])), cells) for name, cells in parts}


def cpu_baseline(lat, steps):
    import oracle
    tau = 0.56
    threads = oracle.default_threads()
    o = oracle.OracleSolver(lat, (tau - 0.5) / 3, model="regularised", threads=threads)
    o.set_drive([0.02, 0.0, 0.01], period=4000)
    o.set_outlet_pressure([0.0])
    o.run(2)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    return {"value": round(o.N * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {o.N}-cell pipe (oracle: C restatement of lbm.py k_step, FP32, "
                      f"{threads} threads; the reference's Warp CPU path cannot run here)"}


def load_traffic(layout, storage="fp32"):
    tag = layout if storage == "fp32" else f"{layout}_{storage}"
    path = os.path.join(ROOT, "profiles", f"traffic_step_{tag}_womersley.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def pmc_probe(layout, storage):
    """Short run for hardware-counter passes: the bench workload, 4 eager steps (2 even, 2 odd) and a 1 GiB device
    copy whose counter readings calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch
    s = make_solver(build_workload(), layout, storage)
    s.run(4, chunk=1000)
    s.synchronize()
    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print("pmc probe done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--layout", default="auto")
    ap.add_argument("--storage", default="fp32", choices=sorted(BYTES_PER_LUP))
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true", help="short run for rocprofv3 --pmc passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU (the step of one lattice "
              f"runs on one device; independent cases shard across GPUs without exchange)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_probe:
        pmc_probe(args.layout, args.storage)
        return
    import torch
    t0 = time.perf_counter()
    lat = build_workload()
    N = lat["nbr"].shape[1]
    print(f"bench: lattice {N} cells built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    s = make_solver(lat, args.layout, args.storage)
    s.run(args.warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    if not s.ok():
        raise RuntimeError("solver diverged during the benchmark")
    mlups = N * args.steps / secs / 1e6
    kt = kernel_times_ms(s)
    name, (k_ms, k_cells) = max(kt.items(), key=lambda kv: kv[1][0])   # dominant launch
    bpl = BYTES_PER_LUP[args.storage]
    achieved = bpl * k_cells / (k_ms * 1e-3) / 1e9
    traffic = load_traffic(s.layout, args.storage)
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(secs / args.steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "womersley_pipe_R40_L480", "cells": N, "tau": 0.56, "layout": s.layout,
                   "storage": args.storage, "model": "regularised+galilean, bouzidi"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": name, "kernel_ms": round(k_ms, 5), "kernel_cells": k_cells,
                     "bytes_per_cell": bpl,
                     "launches_ms": {k: round(v[0], 5) for k, v in kt.items()}},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lat, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
