"""Serving form: a resident request loop with continuous batching.

Requests (lattice solves with a step count) arrive over time. The server keeps one ``batch.Batch`` on the device:

* arriving requests are built on host worker threads (lattice + brick layout) while the GPU steps the current
  cases;
* built requests are admitted in arrival order as soon as a case slot and room in the HBM budget are free; a
  request that does not fit waits, and so do the ones behind it (no starvation of large requests);
* all admitted cases advance together in graph-captured chunks of ``chunk`` steps; a request's step count is
  rounded up to whole chunks, so it finishes at a chunk boundary, where its results are read and its memory is
  returned to the arena allocators. The host waits for the device only at such boundaries, and otherwise queues at
  most two chunks ahead; joining and leaving cases do not invalidate the captured graphs (fixed-capacity lists);
* per request the server records arrival, admission and completion times: time-to-solution = completion -
  arrival, reported as p50 / p99 with the throughput in cases per hour.

Results do not depend on the mix (each case's arithmetic is that of the case alone; tested bit for bit).
``run_sequential`` is the baseline: the same requests solved one after the other by ``lbm.Solver``.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import time

import numpy as np

from . import batch as B
from . import bricks


def _summary(case_or_solver):
    p, u = case_or_solver.fields()
    sp = np.linalg.norm(u, axis=1)
    return {"u_max": float(sp.max()), "u_mean": float(sp.mean()), "p_mean": float(p.mean()),
            "ok": bool(case_or_solver.ok())}


def _build(make_case, spec):
    """Worker-thread part of a request: lattice, brick layout, and the setup with its host work done here."""
    lat, nu, setup = make_case(spec)
    lay = bricks.build(lat)
    rec = B.SetupRecorder(lat, nu, lay)
    setup(rec)
    return lat, nu, rec, lay


class Server:
    """Continuous-batching solver service on one GPU.

    storage: population storage of the cases ("fp16" or "fp32"); budget_bytes: HBM the arenas may take (default
    288 GB, capped by the device's free memory less 2 GiB); chunk: steps per graph launch; arena_bricks /
    arena_records: arena capacity (an arena grows to fit a larger case); workers: host threads building
    lattices; keep: "summary" or "fields" (also keep the final (p, u) of each request)."""

    def __init__(self, storage="fp16", budget_bytes=None, chunk=200, arena_bricks=1 << 20, arena_records=1 << 18,
                 workers=12, keep="summary", max_cases=256, variant=0, device=None, **solver_kw):
        import torch
        free = None
        if torch.cuda.is_available():
            free, _ = torch.cuda.mem_get_info(torch.device(device or "cuda:0"))
        budget = 288 * 10 ** 9 if budget_bytes is None else int(budget_bytes)
        if free is not None:
            budget = min(budget, max(0, free - 2 * 2 ** 30))
        self.chunk = int(chunk) + int(chunk) % 2
        self.batch = B.Batch(storage=storage, budget_bytes=budget, arena_bricks=arena_bricks,
                             arena_records=arena_records, max_cases=max_cases, variant=variant, device=device,
                             **solver_kw)
        self.workers = int(workers)
        self.keep = keep

    def serve(self, requests, make_case, time_scale=1.0, log=None):
        """Serve `requests` (dicts with "name", "arrival" in seconds, "steps") whose lattices come from
        make_case(spec) -> (lattice, nu, setup). Arrivals are replayed in real time (x time_scale). Returns the
        report: per-request records and the aggregate throughput / latency figures."""
        bt, chunk = self.batch, self.chunk
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        building = {}
        built = []                                        # (arrival, name, spec, build result)
        active = {}                                       # cid -> record
        records = []
        import torch
        lups = 0
        inflight = []
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(self.workers) as pool:
            while pending or building or built or active:
                now = time.perf_counter() - t0
                while pending and pending[0]["arrival"] * time_scale <= now:
                    r = pending.popleft()
                    building[pool.submit(_build, make_case, r)] = r
                for fut in [f for f in building if f.done()]:
                    r = building.pop(fut)
                    built.append((r["arrival"], r["name"], r, fut.result()))
                built.sort(key=lambda x: (x[0], x[1]))
                while built and bt.can_admit(built[0][3][3]):
                    _, _, r, (lat, nu, setup, lay) = built.pop(0)
                    case = bt.add(lat, nu, layout=lay)
                    setup.replay(case)
                    target = -(-int(r["steps"]) // chunk) * chunk
                    active[case.cid] = {"name": r["name"], "spec": r, "case": case, "target": target,
                                        "arrival": r["arrival"] * time_scale,
                                        "admitted": time.perf_counter() - t0, "cells": case.N}
                    if log:
                        log(f"admit {r['name']} ({case.N} cells, {target} steps) at {active[case.cid]['admitted']:.2f} s"
                            f", {len(active)} active")
                if not active:
                    if built:       # the batch is empty and the head of the queue still does not fit
                        raise RuntimeError(f"request {built[0][1]} does not fit the HBM budget even alone")
                    waits = list(building)
                    if waits:
                        cf.wait(waits, timeout=0.01, return_when=cf.FIRST_COMPLETED)
                    elif pending:
                        time.sleep(max(0.0, min(0.01, pending[0]["arrival"] * time_scale - now)))
                    continue
                bt.run(chunk, chunk=chunk)
                lups += chunk * sum(rec["cells"] for rec in active.values())
                ev = torch.cuda.Event()
                ev.record(bt._tstream)
                inflight.append(ev)
                finishing = [c for c, rec in active.items() if rec["case"].steps >= rec["target"]]
                if finishing:
                    bt.synchronize()             # results of the cases done at the end of this chunk
                    inflight.clear()
                elif len(inflight) > 2:
                    inflight.pop(0).synchronize()   # run at most two chunks ahead of the device
                for cid in finishing:
                    rec = active.pop(cid)
                    rec["done"] = time.perf_counter() - t0
                    rec["result"] = _summary(rec["case"])
                    if self.keep == "fields":
                        rec["fields"] = rec["case"].fields()
                    rec["steps"] = rec["case"].steps
                    bt.remove(rec["case"])
                    del rec["case"]
                    records.append(rec)
                    if log:
                        log(f"done {rec['name']} at {rec['done']:.2f} s (latency {rec['done'] - rec['arrival']:.2f} s)")
        wall = time.perf_counter() - t0
        lat = np.array([r["done"] - r["arrival"] for r in records])
        return {"records": records, "cases": len(records), "wall_s": wall,
                "cases_per_hour": len(records) / wall * 3600.0,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mlups": lups / wall / 1e6, "storage": bt.storage, "chunk": chunk}


def run_sequential(requests, make_case, storage="fp32", chunk=200, keep="summary"):
    """Baseline: each request built and solved alone by lbm.Solver on the bricks layout, one after the other (step
    counts rounded to whole chunks as in the server)."""
    from . import lbm
    chunk = int(chunk) + int(chunk) % 2
    records, lups = [], 0
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r["arrival"]):
        ts = time.perf_counter()
        lat, nu, setup = make_case(r)
        s = lbm.Solver(lat, nu, model="regularised", layout="bricks", storage=storage)
        setup(s)
        steps = -(-int(r["steps"]) // chunk) * chunk
        s.run(steps, chunk=chunk)
        s.synchronize()
        rec = {"name": r["name"], "cells": s.N, "steps": steps, "solve_s": time.perf_counter() - ts,
               "result": _summary(s)}
        if keep == "fields":
            rec["fields"] = s.fields()
        records.append(rec)
        lups += s.N * steps
        del s
    wall = time.perf_counter() - t0
    return {"records": records, "cases": len(records), "wall_s": wall, "cases_per_hour": len(records) / wall * 3600,
            "mlups": lups / wall / 1e6, "storage": storage}
