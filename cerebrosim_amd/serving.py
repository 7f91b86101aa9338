"""Serving form: a resident request loop with continuous batching.

Requests (lattice solves with a step count) arrive over time. The server keeps one ``batch.Batch`` on the device:

* arriving requests are built on host worker threads (lattice + brick layout) while the GPU steps the current
  cases;
* built requests are admitted in arrival order as soon as a case slot and room in the HBM budget are free; a
  request that does not fit waits, and so do the ones behind it (no starvation of large requests);
* all admitted cases advance together in graph-captured chunks of ``chunk`` steps; a request's step count is
  rounded up to whole chunks, so it finishes at a chunk boundary. Its results are then copied to pinned host
  memory behind that chunk and its ranges are returned to the arena allocators at once: stream order keeps the
  copies ahead of any reuse, so a completion never drains the device. The host queues at most two chunks ahead;
  joining and leaving cases do not invalidate the captured graphs (fixed-capacity lists). Uploads of admitted
  cases go through a large pinned ring (``staging_bytes``), so admitting a burst does not wait for the device;
* per request the server records arrival, admission and completion times: time-to-solution = completion -
  arrival, reported as p50 / p99 with the throughput in cases per hour.

Results do not depend on the mix (each case's arithmetic is that of the case alone; tested bit for bit).
``run_sequential`` is the baseline: the same requests solved one after the other by ``lbm.Solver``.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import os
import time

import numpy as np

from . import batch as B
from . import bricks


def _summary(p, u, ok):
    sp = np.linalg.norm(u, axis=1)
    return {"u_max": float(sp.max()), "u_mean": float(sp.mean()), "p_mean": float(p.mean()), "ok": bool(ok)}


def _build(make_case, spec):
    """Worker-thread part of a request: lattice, brick layout, and the setup with its host work done here (one
    native builder thread per request: the pool spreads the requests over the CPUs)."""
    from .lattice import builder_threads
    with builder_threads(1):
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
    lattices (default: the process's CPUs less one, at most 15); keep: "summary" or "fields" (also keep the final (p, u) of each request); staging_bytes: pinned
    host ring for the uploads of admitted cases."""

    def __init__(self, storage="fp16", budget_bytes=None, chunk=200, arena_bricks=1 << 20, arena_records=1 << 18,
                 workers=None, keep="summary", max_cases=256, variant=0, device=None, staging_bytes=512 << 20,
                 **solver_kw):
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
                             staging_bytes=staging_bytes, **solver_kw)
        # builder threads: one per CPU the process may use, less one for this loop, at most 15
        self.workers = int(workers) if workers else max(1, min(15, len(os.sched_getaffinity(0)) - 1))
        self.keep = keep

    def serve(self, requests, make_case, time_scale=1.0, log=None):
        """Serve `requests` (dicts with "name", "arrival" in seconds, "steps") whose lattices come from
        make_case(spec) -> (lattice, nu, setup). Arrivals are replayed in real time (x time_scale). Returns the
        report: per-request records and the aggregate throughput / latency figures, with where the host's time went
        (``host_s``) and the device time of the chunks (``gpu_chunk_s``)."""
        import torch
        bt, chunk = self.batch, self.chunk
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        building = {}
        built = []                                        # (arrival, name, spec, build result)
        active = {}                                       # cid -> record
        reading = []                                      # finished cases whose results are on their way
        records = []
        lups = 0
        inflight = []                                     # (end event, records read behind it) of queued chunks
        chunks = []                                       # (start, end) events of every chunk
        host = {"admit": 0.0, "launch": 0.0, "wait": 0.0, "finish": 0.0, "captures": 0}
        last_chunk_s = 0.0                                # device time of the latest chunk known complete
        t0 = time.perf_counter()

        def finish(rec):
            (p, u), ok = rec.pop("read").result()
            reading.remove(rec)
            rec["done"] = time.perf_counter() - t0
            rec["result"] = _summary(p, u, ok)
            if self.keep == "fields":
                rec["fields"] = (p, u)
            records.append(rec)
            if log:
                log(f"done {rec['name']} at {rec['done']:.2f} s (latency {rec['done'] - rec['arrival']:.2f} s)")

        with cf.ThreadPoolExecutor(self.workers) as pool:
            while pending or building or built or active or reading:
                now = time.perf_counter() - t0
                while pending and pending[0]["arrival"] * time_scale <= now:
                    r = pending.popleft()
                    building[pool.submit(_build, make_case, r)] = r
                for fut in [f for f in building if f.done()]:
                    r = building.pop(fut)
                    built.append((r["arrival"], r["name"], r, fut.result()))
                built.sort(key=lambda x: (x[0], x[1]))
                # admission shares the host with the launches: admit for as long as the queued chunks keep the device
                # busy (half their expected time, at least 20 ms), then queue the next chunk
                th = time.perf_counter()
                queued = sum(not c[1].query() for c in chunks[-3:])
                budget = 0.02 + 0.5 * queued * last_chunk_s
                while built and bt.can_admit(built[0][3][3]) and (time.perf_counter() - th < budget or not active):
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
                th2 = time.perf_counter()
                host["admit"] += th2 - th
                for rec in [rec for rec in reading if rec["read"].ready()]:
                    finish(rec)
                host["finish"] += time.perf_counter() - th2
                if not active:
                    if reading:                 # nothing to step: wait for the results on their way
                        th = time.perf_counter()
                        finish(reading[0])
                        host["finish"] += time.perf_counter() - th
                        continue
                    if built:       # the batch is empty and the head of the queue still does not fit
                        raise RuntimeError(f"request {built[0][1]} does not fit the HBM budget even alone")
                    waits = list(building)
                    if waits:
                        cf.wait(waits, timeout=0.01, return_when=cf.FIRST_COMPLETED)
                    elif pending:
                        time.sleep(max(0.0, min(0.01, pending[0]["arrival"] * time_scale - now)))
                    continue
                th = time.perf_counter()
                ng = len(bt.graphs)
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(bt._tstream)
                bt.run(chunk, chunk=chunk)
                host["captures"] += len(bt.graphs) > ng
                lups += chunk * sum(rec["cells"] for rec in active.values())
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(bt._tstream)
                if len(chunks) >= 1 and chunks[-1][1].query():
                    last_chunk_s = chunks[-1][0].elapsed_time(chunks[-1][1]) * 1e-3
                chunks.append((e0, ev, th - t0, len(active), sum(rec["cells"] for rec in active.values())))
                # cases done at the end of this chunk: their results are copied out behind it and they leave the
                # batch at once (stream-ordered), so the next chunk is queued without waiting for the device
                done_now = []
                for cid in [c for c, rec in active.items() if rec["case"].steps >= rec["target"]]:
                    rec = active.pop(cid)
                    rec["read"] = rec["case"].read_async()
                    rec["steps"] = rec["case"].steps
                    bt.remove(rec["case"])
                    del rec["case"]
                    rec["left"] = time.perf_counter() - t0
                    reading.append(rec)
                    done_now.append(rec)
                inflight.append((ev, done_now))
                th2 = time.perf_counter()
                host["launch"] += th2 - th
                if len(inflight) > 2:            # run at most two chunks ahead of the device
                    ev0, recs = inflight.pop(0)
                    recs = [rec for rec in recs if "read" in rec]
                    if recs:
                        for rec in recs:         # their copies follow the chunk: waiting for them waits for it
                            finish(rec)
                    else:
                        ev0.synchronize()
                host["wait"] += time.perf_counter() - th2
        wall = time.perf_counter() - t0
        bt.synchronize()
        gpu = sum(c[0].elapsed_time(c[1]) for c in chunks) * 1e-3
        # per chunk: host time of the launch, cases, cells, device start (relative to the first chunk's), duration
        trace = [(round(h, 3), n, cells, round(chunks[0][0].elapsed_time(a) * 1e-3, 3), round(a.elapsed_time(b) * 1e-3, 4))
                 for a, b, h, n, cells in chunks] if chunks else []
        lat = np.array([r["done"] - r["arrival"] for r in records])
        return {"records": records, "cases": len(records), "wall_s": wall,
                "cases_per_hour": len(records) / wall * 3600.0,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mlups": lups / wall / 1e6, "storage": bt.storage, "chunk": chunk,
                "gpu_chunk_s": gpu, "chunks": len(chunks), "chunk_trace": trace,
                "host_s": dict({k: (round(v, 3) if isinstance(v, float) else v) for k, v in host.items()},
                               capture=round(bt.capture_s, 3))}


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
        p, u = s.fields()
        rec = {"name": r["name"], "cells": s.N, "steps": steps, "solve_s": time.perf_counter() - ts,
               "result": _summary(p, u, s.ok())}
        if keep == "fields":
            rec["fields"] = (p, u)
        records.append(rec)
        lups += s.N * steps
        del s
    wall = time.perf_counter() - t0
    return {"records": records, "cases": len(records), "wall_s": wall, "cases_per_hour": len(records) / wall * 3600,
            "mlups": lups / wall / 1e6, "storage": storage}
