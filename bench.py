"""Benchmark of the solver step on one MI355X: MLUPS and HBM-roofline fraction on the pulsatile Womersley pipe.

Workload: radius 40, length 480 pipe along z (about 2.41 M fluid cells), plug velocity inlet
u(t) = 0.02 (1 + 0.5 sin(2 pi t / 4000)), pressure outlet at rho = 1, tau = 0.56, regularised BGK with the Galilean
correction, Bouzidi walls, FP32 arithmetic. A "step" is one solver step over all fluid cells (clock + step kernels).
``--storage fp16`` stores the populations as float16 equilibrium deviations (140 B per fluid cell: 27 populations
read + written x 2 B, previous velocity and density read 16 B, velocity + density written 16 B).

The chosen forms are measured after the headline (``--forms``, default all three, on reduced samples unless
``--full``):

* precision: FP16 population storage on the same pipe (MLUPS and roofline of its dominant launch), and the stenosed
  bifurcation (half size by default, ~1.8 M cells; full size with ``--full``) in FP32 and FP16 with their difference
  on u, p and wall shear stress;
* batching: a seeded set of stenosed tubes (radius U[8, 48], length U[128, 960], area reduction U[0, 0.95],
  tau U[0.51, 0.9]; 32 tubes at half size by default, the full 96 with ``--full``) stepped together in brick arenas
  (autotuned launch variant), FP32 and FP16, against the same tubes solved one after the other by the plain solver;
* serving: Poisson-arriving stenosis requests (severity, pressure, grid spacing) served with continuous batching in
  FP16 under the HBM budget -- cases/hour with p50/p99 time-to-solution -- against the same requests solved one
  after the other by the plain FP32 solver (both include building each lattice on the host).

``artery`` (skip with ``--no-artery``): the reference's own speed benchmark, test 5 of validate.py on the story
artery (healthy, 0.1 mm, 620,028 cells, 1000 + 20000 steps), beside its published RTX 3090 figures
(geometry/solver.json:179-194, reports/step3_report.md:58-64), with the time per heartbeat derived the same way.

Prints one JSON line. ``roofline.achieved`` = algorithmic bytes of the step kernel per launch (244 B per fluid cell:
27 populations read + 27 written x 4 B, previous velocity read 12 B, velocity + density written 16 B) / the kernel's
mean duration measured with HIP events on the solver's stream; ``step_frac`` is the whole step's algorithmic rate
(MLUPS x 244 B) over the 8 TB/s peak; ``copy_gbs`` is the library's own streaming copy kernel (cs_stream_copy, 1 GiB,
read + write) on the same GPU, the achievable ceiling beside the peak. The timed steps are graph replays, as the
product runs them (the graph is captured before the timed region). ``cpu_baseline`` times the oracle (the C restatement
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
    """Mean duration of the step's launch as the product makes it (the brick layout: one paired launch of the
    boundary and interior lists, ``k_bricks_pair``), and of each list launched alone in sequence
    (``k_bricks_interior``, ``k_bricks_boundary``): HIP events on the solver's stream around each launch of eager
    steps, parities alternating. Returns {name: (ms, cells)}."""
    import torch
    st = s._tstream
    parts = s.backend.launch_parts()
    ev = {name: [] for name, _ in parts}
    out = {}

    def timed(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        return e0, e1

    if s.layout == "bricks" and len(parts) == 2 and all(c for _, c in parts):
        pair = []
        for _ in range(steps):
            s._clock()
            pair.append(timed(lambda: s.backend.step(s, s.cur)))
            s.cur = 1 - s.cur
            s.steps += 1
        st.synchronize()
        out["k_bricks_pair"] = (float(np.mean([a.elapsed_time(b) for a, b in pair])), s.N)
    for _ in range(steps):
        s._clock()
        for k, (name, _) in enumerate(parts):
            ev[name].append(timed(lambda: s.backend.step_part(s, s.cur, k == 1)))
        s.cur = 1 - s.cur
        s.steps += 1
    st.synchronize()
    out.update({name: (float(np.mean([a.elapsed_time(b) for a, b in ev[name]])), cells) for name, cells in parts})
    return out


def dominant_launch(kt):
    """The product's launch of a step: the paired brick launch when there is one, else the longest launch."""
    if "k_bricks_pair" in kt:
        return "k_bricks_pair", kt["k_bricks_pair"]
    return max(kt.items(), key=lambda kv: kv[1][0])


def copy_bandwidth_gbs(nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth on this device: the library's streaming copy kernel (cs_stream_copy, 16-byte loads
    and stores) over nbytes, (read + write) bytes over its mean duration (HIP events on the stream it runs on)."""
    import ctypes

    import torch
    from cerebrosim_amd import _native
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    st = torch.cuda.current_stream()
    h = ctypes.c_void_p(st.cuda_stream)
    cp = lambda: _native.call("cs_stream_copy", ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                              nbytes, h)
    cp()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        cp()
    e1.record(st)
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) * 1e-3 / reps
    ok = bool(torch.equal(x, y))
    del x, y
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("cs_stream_copy: copy differs from its source")
    return 2.0 * nbytes / secs / 1e9


def oracle_checks():
    """The CPU-baseline leg's checks, with the oracle as the checker (FP32 and FP64 restatements of the step):

    * the steady Poiseuille box (24 x 24 x 64, R 10, 20,160 cells, tau 0.8, force 1e-6, 10,000 steps): the oracle
      on all host threads (timed: the full CPU run of this workload) against the GPU's FP32 and FP16 storage;
    * a served sample of the seed-2 request stream (4 requests, 0.2 mm grid, cycles of 60 steps): each request
      served in FP16 storage against the FP64 oracle.
    Returns {"poiseuille": {...}, "served_fp16": {...}}."""
    import oracle
    from cerebrosim_amd import lbm, serving, workloads
    from cerebrosim_amd.geometry import poiseuille_box

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out = {}
    lat, r, R = poiseuille_box()
    nu, steps = 0.1, 10000
    refs = {}
    for dt in (np.float32, np.float64):
        o = oracle.OracleSolver(lat, nu, model="regularised", dtype=dt)
        o.set_force([0, 0, 1e-6])
        t0 = time.perf_counter()
        o.run(steps)
        secs = time.perf_counter() - t0
        refs[np.dtype(dt).name] = (o.get_f(), o.fields())
        if dt is np.float32:
            out_cpu = {"mlups": round(o.N * steps / secs / 1e6, 2), "s": round(secs, 2), "cores": o.threads}
    res = {"workload": "poiseuille_box_R10_24x24x64", "cells": int(lat["nbr"].shape[1]), "steps": steps,
           "oracle_fp32_cpu": out_cpu}
    for storage in ("fp32", "fp16"):
        s = lbm.Solver(lat, nu, model="regularised", layout="bricks", storage=storage)
        s.set_force([0, 0, 1e-6])
        s.run(steps)
        f, (p, u) = s.get_f(), s.fields()
        f32, (p32, u32) = refs["float32"]
        _, (p64, u64) = refs["float64"]
        res[storage] = {"u_vs_oracle_fp64": rel(u, u64), "u_vs_oracle_fp32": rel(u, u32)}
        if storage == "fp32":
            res[storage]["f_vs_oracle_fp32"] = rel(f, f32)
        del s
    out["poiseuille"] = res
    reqs = workloads.request_specs(4, rate_per_s=50.0, seed=2, beats=4, beat_steps=60, grids=(0.2e-3,))
    srv = serving.Server(storage="fp16", chunk=50, arena_bricks=1 << 15, arena_records=1 << 14, keep="fields")
    rep = srv.serve(reqs, workloads.request_case)
    got = {x["name"]: x for x in rep["records"]}
    errs = []
    for q in reqs:
        lat, nu, setup = workloads.request_case(q)
        o = oracle.OracleSolver(lat, nu, model="regularised", dtype=np.float64)
        setup(o)
        o.run(got[q["name"]]["steps"])
        errs.append(rel(got[q["name"]]["fields"][1], o.fields()[1]))
    out["served_fp16"] = {"workload": "stenosis_requests_seed2_n4_0.2mm_beat60", "u_vs_oracle_fp64": errs,
                          "max": max(errs)}
    return out


def timed_run(s, steps, warmup):
    """`warmup` untimed steps, then `steps` steps timed on the wall clock as the product runs them: one captured
    graph of the even part of `steps` and an odd last step launched eagerly. The graph is captured before the
    warm-up (capture launches nothing), so that the timed steps follow the warm-up without an idle gap. Returns
    seconds."""
    import torch
    if warmup % 2:
        s.run(1)                                   # the odd warm-up step first: the graphs below keep the parity
    chunk, wchunk = steps - steps % 2, warmup - warmup % 2
    for c in (chunk, wchunk):
        if c:
            s.prepare(c)
    if wchunk:
        s.run(wchunk, chunk=wchunk)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(steps, chunk=max(2, chunk))
    s.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(lat, steps, steps_1thread=4):
    """The oracle (C restatement of lbm.py k_step, FP32) on `steps` steps of the same pipe with every host thread,
    and on `steps_1thread` steps with one thread (a serial launch loop)."""
    import oracle
    tau = 0.56

    def rate(threads, n):
        o = oracle.OracleSolver(lat, (tau - 0.5) / 3, model="regularised", threads=threads)
        o.set_drive([0.02, 0.0, 0.01], period=4000)
        o.set_outlet_pressure([0.0])
        o.run(2)
        t0 = time.perf_counter()
        o.run(n)
        return o.N * n / (time.perf_counter() - t0) / 1e6, o.N

    threads = oracle.default_threads()
    v, N = rate(threads, steps)
    v1, _ = rate(1, steps_1thread)
    return {"value": round(v, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "value_1thread": round(v1, 3),
            "sample": f"{steps} steps ({threads} threads) and {steps_1thread} steps (1 thread) of the same {N}-cell "
                      f"pipe (oracle: C restatement of lbm.py k_step, FP32; the reference's Warp CPU path cannot "
                      f"run here)"}


def form_precision(lat, steps, warmup):
    """FP16 storage on the headline pipe: MLUPS and the dominant launch's algorithmic bandwidth."""
    import torch
    s = make_solver(lat, "bricks", "fp16")
    s.autotune()
    secs = timed_run(s, steps, warmup)
    ok = s.ok()
    kt = kernel_times_ms(s)
    name, (k_ms, k_cells) = dominant_launch(kt)
    ach = BYTES_PER_LUP["fp16"] * k_cells / (k_ms * 1e-3) / 1e9
    variant = s.variant
    del s
    torch.cuda.empty_cache()
    mlups = lat["nbr"].shape[1] * steps / secs / 1e6
    return {"storage": "fp16", "mlups": round(mlups, 1), "ok": ok,
            "step_frac": round(mlups * BYTES_PER_LUP["fp16"] / 1e3 / HBM_PEAK_GBS, 4),
            "variant": variant,
            "kernel": name, "kernel_ms": round(k_ms, 5), "achieved_gbs": round(ach, 1),
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_cell": BYTES_PER_LUP["fp16"],
            "error_vs_fp64": None}


def form_precision_bifurcation(full, steps=None):
    """The stenosed bifurcation (precision-form workload) with FP32 and FP16 storage on the GPU: MLUPS of each and the
    FP16-vs-FP32 relative L2 of u, p and wall shear stress after the same steps. (The FP64 comparison needs the
    oracle, which only the tests run: tests/test_gpu_fp16.py::test_precision_bifurcation_u_p_wss_vs_fp64.)"""
    import torch
    from cerebrosim_amd import lbm, wss, workloads
    scale = 1.0 if full else 0.5
    t0 = time.perf_counter()
    lat, nu, setup, info = workloads.bifurcation_case(scale)
    idx, c, normals, _ = workloads.wall_shear_weights(lat, info)
    N = lat["nbr"].shape[1]
    print(f"bench: bifurcation scale {scale}: {N} cells, built in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    steps = steps or int(round(4000 * scale))          # one inlet period
    res, out = {}, {}
    for st in ("fp32", "fp16"):
        s = lbm.Solver(lat, nu, model="regularised", layout="bricks", storage=st)
        setup(s)
        s.run(200)
        s.synchronize()
        t1 = time.perf_counter()
        s.run(steps - 200)
        s.synchronize()
        secs = time.perf_counter() - t1
        p, u = s.fields()
        tau = wss.WallShear(s, idx, c, normals).frame(accumulate=False).cpu().numpy()
        res[st] = (p, u, tau, s.ok())
        out[f"mlups_{st}"] = round(N * (steps - 200) / secs / 1e6, 1)
        del s
        torch.cuda.empty_cache()
    rel = lambda a, b: float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))
    (p32, u32, t32, ok32), (p16, u16, t16, ok16) = res["fp32"], res["fp16"]
    out.update({"workload": f"stenosed_bifurcation_scale{scale}", "cells": N, "steps": steps, "ok": ok32 and ok16,
                "fp16_vs_fp32_rel_l2": {"u": rel(u16, u32), "p": rel(p16, p32), "wss": rel(t16, t32)},
                "wall_points": int(len(normals))})
    return out


def form_batching(full, steps=400):
    """Tubes stepped together (FP32 and FP16 arenas) vs solved one after the other (plain FP32 solver)."""
    import torch
    from cerebrosim_amd import batch as B
    from cerebrosim_amd import bricks, lbm, workloads
    specs = workloads.tube_specs(96, seed=1)
    if not full:
        specs = [dict(sp, R=sp["R"] / 2, L=sp["L"] / 2, u_mean=workloads.inlet_speed(sp["R"] / 2, sp["area_frac"],
                                                                                      sp["tau"]))
                 for sp in specs[:32]]
    t0 = time.perf_counter()
    cases = []
    for sp in specs:
        lat, nu, setup = workloads.tube_case(sp)
        cases.append((lat, nu, setup, bricks.build(lat)))
    host_s = time.perf_counter() - t0
    cells = sum(c[0]["nbr"].shape[1] for c in cases)
    print(f"bench: batching sample {len(cases)} tubes, {cells} cells, built in {host_s:.1f} s", file=sys.stderr,
          flush=True)
    out = {"workload": f"tubes_seed1_{'n96' if full else 'n32_half_size'}", "cases": len(cases), "cells": cells,
           "steps": steps, "host_build_s": round(host_s, 1)}
    # brick-size census of the same tubes: storage slots per fluid cell and the share of cells in boundary bricks
    # for the 4^3 bricks the kernels use and for 8^3 bricks
    cen = {}
    for edge in (4, 8):
        cs_ = [bricks.census(c[0], edge) for c in cases]
        tot_cells = sum(c["cells"] for c in cs_)
        cen[f"{edge}^3"] = {"slots_per_cell": round(sum(c["slots"] for c in cs_) / tot_cells, 4),
                            "boundary_cell_fraction": round(sum(c["boundary_cell_fraction"] * c["cells"]
                                                                for c in cs_) / tot_cells, 4)}
    out["brick_census"] = cen
    # sequential plain solves (same host layouts; set-up and stepping timed apart)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_s = 0.0
    for lat, nu, setup, lay in cases:
        s = lbm.Solver(lat, nu, model="regularised", layout="bricks", brick_layout=lay)
        setup(s)
        s.synchronize()
        t1 = time.perf_counter()
        s.run(steps)
        s.synchronize()
        step_s += time.perf_counter() - t1
        del s
    seq = time.perf_counter() - t0
    torch.cuda.empty_cache()
    out["sequential_fp32"] = {"s": round(seq, 3), "setup_s": round(seq - step_s, 3), "step_s": round(step_s, 3),
                              "mlups": round(cells * steps / step_s / 1e6, 1)}
    for storage in ("fp32", "fp16"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bt = B.Batch(storage=storage, arena_bricks=min(B.MAX_BRICKS, 1 << 21), arena_records=1 << 19,
                     max_cases=max(128, len(cases)))
        hs = [bt.add(lat, nu, layout=lay) for lat, nu, setup, lay in cases]
        for h, (lat, nu, setup, lay) in zip(hs, cases):
            setup(h)
        bt.synchronize()
        t_add = time.perf_counter() - t0
        tune = bt.autotune(steps=10, repeats=1)
        t1 = time.perf_counter()
        bt.run(steps)
        bt.synchronize()
        run_s = time.perf_counter() - t1
        ok = all(h.ok() for h in hs)
        tot = t_add + run_s
        out[f"batched_{storage}"] = {"s": round(tot, 3), "admit_s": round(t_add, 3), "step_s": round(run_s, 3),
                                     "mlups": round(cells * steps / run_s / 1e6, 1), "arenas": len(bt.arenas),
                                     "variant": bt.variant, "autotune_ms_per_step": {str(k): round(v, 4)
                                                                                      for k, v in tune.items()},
                                     "ok": ok}
        out[f"speedup_{storage}_vs_sequential"] = round(seq / tot, 2)
        out[f"step_speedup_{storage}_vs_sequential"] = round(step_s / run_s, 2)
        del hs, bt
        torch.cuda.empty_cache()
    return out


def form_serving(full, rate=None, size=None):
    """Stenosis requests served with continuous batching (FP16) vs solved one after the other (FP32). size:
    "default" (96 requests, 0.2 / 0.1 mm grids, 500-step cycles), "mid" (32 requests at 8/s on all three grids, so
    that 0.05 mm requests of about 5 M cells share the batch, 500-step cycles) or "full" (192 requests, all three
    grids, 2000-step cycles)."""
    import torch
    from cerebrosim_amd import serving, workloads
    size = size or ("full" if full else "default")
    n = {"default": 96, "mid": 32, "full": 192}[size]
    grids = (0.2e-3, 0.1e-3) if size == "default" else (0.2e-3, 0.1e-3, 0.05e-3)
    beat = 2000 if size == "full" else 500
    rate = rate or (8.0 if size == "mid" else 32.0)  # arrivals faster than service: the queue builds up
    reqs = workloads.request_specs(n, rate_per_s=rate, seed=2, beats=4, beat_steps=beat, grids=grids)
    # graph chunks of 400 steps: fewer launches and captures; measured faster than 100 and 200 on this stream
    srv = serving.Server(storage="fp16", chunk=400)
    rep = srv.serve(reqs, workloads.request_case)
    del srv
    torch.cuda.empty_cache()
    seq = serving.run_sequential(reqs, workloads.request_case, storage="fp32", chunk=400)
    ok = all(r["result"]["ok"] for r in rep["records"])
    # steady state: completions between the 25th and 75th percentile completion times
    done = np.sort([r["done"] for r in rep["records"]])
    a, b = done[len(done) // 4], done[(3 * len(done)) // 4]
    steady = ((3 * len(done)) // 4 - len(done) // 4) / max(b - a, 1e-9) * 3600.0
    tag = {"default": "grids_0.2_0.1mm", "mid": "grids_0.2_0.1_0.05mm", "full": "full"}[size]
    cells = [r["cells"] for r in rep["records"]]
    return {"workload": f"stenosis_requests_seed2_n{n}_{tag}_beat{beat}", "max_cells": int(max(cells)),
            "requests": n, "rate_per_s": rate, "storage": "fp16", "chunk": rep["chunk"],
            "cases_per_hour": round(rep["cases_per_hour"], 1), "p50_s": round(rep["p50_s"], 3),
            "p99_s": round(rep["p99_s"], 3), "wall_s": round(rep["wall_s"], 2), "mlups": round(rep["mlups"], 1),
            "steady_cases_per_hour": round(steady, 1), "gpu_chunk_s": round(rep["gpu_chunk_s"], 2),
            "host_s": rep["host_s"],
            "ok": ok, "sequential_fp32_cases_per_hour": round(seq["cases_per_hour"], 1),
            "sequential_fp32_wall_s": round(seq["wall_s"], 2),
            "speedup_vs_sequential": round(rep["cases_per_hour"] / seq["cases_per_hour"], 2),
            "steady_speedup_vs_sequential": round(steady / seq["cases_per_hour"], 2)}


def load_traffic(layout, storage="fp32", variant=0):
    """HBM bytes per launch of the dominant kernel from the committed PMC summary (FETCH_SIZE + WRITE_SIZE, scaled
    by the copy kernel's readings), for the launch variant that was timed; None when that variant was not
    profiled."""
    tag = layout if storage == "fp32" else f"{layout}_{storage}"
    path = os.path.join(ROOT, "profiles", f"traffic_step_{tag}_womersley.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    by_variant = rec.get("bytes_per_launch_by_variant")
    if by_variant is not None:
        return by_variant.get(str(variant))
    return rec.get("bytes_per_launch") if rec.get("variant", 0) == variant else None


def pmc_probe(layout, storage, variant=0):
    """Short run for hardware-counter passes: the bench workload in launch variant `variant`, 4 eager steps (2 even,
    2 odd) as the product launches them (one paired launch per step), then 4 with the interior and boundary lists
    launched alone one after the other, and a 1 GiB cs_stream_copy whose counter readings calibrate FETCH_SIZE /
    WRITE_SIZE."""
    import ctypes

    import torch
    from cerebrosim_amd import _native
    s = make_solver(build_workload(), layout, storage)
    s.variant = int(variant)
    for separate in (False, True):
        for _ in range(4):
            s._clock()
            if separate:
                for part in (False, True):
                    s.backend.step_part(s, s.cur, part)
            else:
                s.backend.step(s, s.cur)
            s.cur = 1 - s.cur
            s.steps += 1
    s.synchronize()
    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    torch.cuda.synchronize()
    _native.call("cs_stream_copy", ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1 << 30,
                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    print("pmc probe done", flush=True)


def traffic_summary(fetch_csv, write_csv, storage, variant, out_path):
    """HBM bytes per dispatch of each brick kernel from two rocprofv3 --pmc runs of --pmc-probe (FETCH_SIZE, then
    WRITE_SIZE: KiB per dispatch), each scaled by what the same counter reads for the probe's 1 GiB copy kernel, merged
    into the JSON summary out_path under `variant`. The dominant launch is the paired step kernel (mean of its even
    and odd dispatches)."""
    import csv
    import re

    def per_kernel(path, counter):
        got = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"]
            if "k_stream_copy" in name:
                key = "copy"
            elif "k_bricks" in name:
                key = re.sub(r"^.*?(k_bricks(?:_pair)?<[^>]*>).*$", r"\1", name)
            else:
                continue
            got.setdefault(key, []).append(float(r["Counter_Value"]) * 1024.0)
        return {k: float(np.mean(v)) for k, v in got.items()}

    rd, wr = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    sr, sw = float(1 << 30) / rd["copy"], float(1 << 30) / wr["copy"]
    kernels = {k: {"read": int(round(rd[k] * sr)), "write": int(round(wr[k] * sw))} for k in rd if k != "copy"}
    pair = [v["read"] + v["write"] for k, v in kernels.items() if k.startswith("k_bricks_pair")]
    rec = json.load(open(out_path)) if os.path.exists(out_path) else {}
    N = build_workload()["nbr"].shape[1]
    rec.update({"kernel": "k_bricks_pair (both brick lists of a step in one launch), mean of even and odd dispatches",
                "workload": "womersley_pipe_R40_L480", "storage": storage,
                "algorithmic_bytes_per_launch": BYTES_PER_LUP[storage] * N})
    rec.setdefault("bytes_per_launch_by_variant", {})[str(variant)] = int(round(np.mean(pair))) if pair else None
    rec.setdefault("per_launch_bytes_by_variant", {})[str(variant)] = kernels
    rec.setdefault("scale_by_variant", {})[str(variant)] = {"read": round(sr, 4), "write": round(sw, 4)}
    with open(out_path, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(json.dumps({"variant": variant, "bytes_per_launch": rec["bytes_per_launch_by_variant"][str(variant)],
                      "kernels": kernels}))


# the reference's published speed on the story artery (RTX 3090): solver.json:179-194, step3_report.md:58-64
ARTERY_PUBLISHED = {"mlups": 384.9, "minutes_per_beat": {"healthy": 5.4, "narrow70": 13.4, "narrow70_fine": 214.8}}


def artery_speed():
    """validate.py's test 5 on the story artery (the segment file committed with the tests), FP32."""
    from cerebrosim_amd import validate
    seg = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "segment_story.json")
    res, beat = validate.artery(seg)
    pub = validate.PUBLISHED_ARTERY
    return {"workload": "story_artery_healthy_0.1mm_validate_test5", "cells": res["cells"],
            "steps": res["steps"], "mlups": round(res["mlups"], 1), "ok": res["ok"],
            "published_rtx3090_mlups": ARTERY_PUBLISHED["mlups"],
            "speedup_vs_published": round(res["mlups"] / ARTERY_PUBLISHED["mlups"], 1),
            "counts_match_published": res["counts"] == pub["counts"] and res["cells"] == pub["cells"],
            "u_max_m_s": res["u_max_m_s"], "published_u_max_m_s": pub["u_max_m_s"],
            "minutes_per_beat": {k: round(v["minutes_per_beat"], 3) for k, v in beat.items()},
            "published_minutes_per_beat": ARTERY_PUBLISHED["minutes_per_beat"]}


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
    ap.add_argument("--forms", default="precision,batching,serving",
                    help="comma-separated chosen forms to measure after the headline ('none' to skip; "
                         "'serving_mid' adds the mid-size request sample with 0.05 mm grids)")
    ap.add_argument("--full", action="store_true", help="full-size batching / serving workloads")
    ap.add_argument("--no-autotune", action="store_true", help="keep the launch variant given by --variant")
    ap.add_argument("--variant", type=int, default=0, help="launch variant without autotuning (and of --pmc-probe)")
    ap.add_argument("--no-artery", action="store_true", help="skip the story-artery speed comparison")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the HBM traffic summary of --storage / --variant into profiles/ from two --pmc runs "
                         "of --pmc-probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU (the step of one lattice "
              f"runs on one device; independent cases shard across GPUs without exchange)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_probe:
        pmc_probe(args.layout, args.storage, args.variant)
        return
    if args.traffic_from:
        tag = "bricks" if args.storage == "fp32" else f"bricks_{args.storage}"
        traffic_summary(*args.traffic_from, args.storage, args.variant,
                        os.path.join(ROOT, "profiles", f"traffic_step_{tag}_womersley.json"))
        return
    import torch
    t0 = time.perf_counter()
    lat = build_workload()
    N = lat["nbr"].shape[1]
    print(f"bench: lattice {N} cells built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    s = make_solver(lat, args.layout, args.storage)
    s.variant = args.variant
    tune = s.autotune() if not args.no_autotune else {}
    secs = timed_run(s, args.steps, args.warmup)
    if not s.ok():
        raise RuntimeError("solver diverged during the benchmark")
    mlups = N * args.steps / secs / 1e6
    kt = kernel_times_ms(s)
    name, (k_ms, k_cells) = dominant_launch(kt)   # the step's launch (both brick lists in one kernel)
    bpl = BYTES_PER_LUP[args.storage]
    achieved = bpl * k_cells / (k_ms * 1e-3) / 1e9
    traffic = load_traffic(s.layout, args.storage, s.variant)
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(secs / args.steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "womersley_pipe_R40_L480", "cells": N, "tau": 0.56, "layout": s.layout,
                   "storage": args.storage, "model": "regularised+galilean, bouzidi", "variant": s.variant,
                   "autotune_ms_per_step": {str(k): round(v, 4) for k, v in tune.items()}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "step_achieved": round(mlups * bpl / 1e3, 1),
                     "step_frac": round(mlups * bpl / 1e3 / HBM_PEAK_GBS, 4),
                     "kernel": name, "kernel_ms": round(k_ms, 5), "kernel_cells": k_cells,
                     "bytes_per_cell": bpl,
                     "launches_ms": {k: round(v[0], 5) for k, v in kt.items()}},
    }
    copy_gbs = copy_bandwidth_gbs()
    out["roofline"].update({"copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)})
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lat, args.cpu_steps)
        out["cpu_baseline"]["checks"] = oracle_checks()
        print("bench: oracle checks done", file=sys.stderr, flush=True)
    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if forms:
        del s
        torch.cuda.empty_cache()
        out["forms"] = {}
        if "precision" in forms:
            out["forms"]["precision"] = form_precision(lat, args.steps, args.warmup)
            out["forms"]["precision"]["bifurcation"] = form_precision_bifurcation(args.full)
            print(f"bench: precision form {out['forms']['precision']['mlups']} MLUPS", file=sys.stderr, flush=True)
        if "batching" in forms:
            out["forms"]["batching"] = form_batching(args.full)
            print(f"bench: batching form done", file=sys.stderr, flush=True)
        if "serving" in forms:
            out["forms"]["serving"] = form_serving(args.full)
            print(f"bench: serving form done", file=sys.stderr, flush=True)
        if "serving_mid" in forms:
            out["forms"]["serving_mid"] = form_serving(False, size="mid")
            print(f"bench: mid-size serving sample done", file=sys.stderr, flush=True)
    checks = out.get("cpu_baseline", {}).get("checks")
    if checks and "forms" in out:
        if "precision" in out["forms"]:
            out["forms"]["precision"]["error_vs_fp64"] = {
                "workload": checks["poiseuille"]["workload"], "steps": checks["poiseuille"]["steps"],
                "u_rel_l2_fp16": checks["poiseuille"]["fp16"]["u_vs_oracle_fp64"],
                "u_rel_l2_fp32": checks["poiseuille"]["fp32"]["u_vs_oracle_fp64"]}
        if "serving" in out["forms"]:
            out["forms"]["serving"]["fp16_u_rel_l2_vs_fp64"] = checks["served_fp16"]
    if not args.no_artery:
        torch.cuda.empty_cache()
        out["artery"] = artery_speed()
        print(f"bench: artery {out['artery']['mlups']} MLUPS", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
