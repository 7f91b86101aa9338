"""Serving form: a resident request loop with continuous batching.

Requests (lattice solves with a step count) arrive over time. The server keeps one ``batch.Batch`` on the device:

* arriving requests are built on host worker threads (lattice + brick layout) while the GPU steps the current
  cases;
* built requests are admitted in arrival order as soon as a case slot and room in the HBM budget are free; a
  request that does not fit waits, and so do the ones behind it (no starvation of large requests);
* all admitted cases advance together in graph-captured chunks of ``chunk`` steps; a request's step count is
  rounded up to whole chunks, so it finishes at a chunk boundary. Its results are then copied to pinned host
  memory behind that chunk and its ranges are returned to the arena allocators at once: stream order keeps the
  copies ahead of any reuse, so a completion never drains the device (the copies land in a pinned pool allocated
  up front: a pinned allocation per read can wait for the device). The host queues at most two chunks ahead;
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
import sys
import time

import numpy as np

from . import batch as B
from . import bricks


def _summary(p, u, ok):
    sp = np.sqrt(np.einsum("ij,ij->i", u, u))
    return {"u_max": float(sp.max()), "u_mean": float(sp.mean()), "p_mean": float(p.mean()), "ok": bool(ok)}


def _build(make_case, spec, threads=1):
    """Worker-thread part of a request: lattice, brick layout, and the setup with its host work done here. threads:
    native builder threads for its lattice (the pool spreads a busy queue over the CPUs, one thread per request; a
    request that arrives to an idle pool takes the idle CPUs too)."""
    from .lattice import builder_threads
    with builder_threads(threads):
        lat, nu, setup = make_case(spec)
    lay = bricks.build(lat)
    rec = B.SetupRecorder(lat, nu, lay)
    setup(rec)
    # the served case needs the layout and the setter calls only: the lattice's large arrays are released here, on
    # the worker, not on the serving thread
    slim = {k: lat[k] for k in ("origin", "dx", "shape", "n_outlets", "counts", "openings") if k in lat}
    rec.lat = slim
    # the admission's uploads, converted and packed here, so that admitting costs the serving thread one staging
    # copy and one host-to-device copy
    rec.packed = B.pack_admission(lay, rec)
    return slim, nu, rec, lay


class Server:
    """Continuous-batching solver service on one GPU.

    storage: population storage of the cases ("fp16" or "fp32"); budget_bytes: HBM the arenas may take (default
    288 GB, capped by the device's free memory less 2 GiB); chunk: steps per graph launch; arena_bricks /
    arena_records: arena capacity (an arena grows to fit a larger case); workers: host threads building
    lattices (default: the process's CPUs less one, at most 10); keep: "summary" or "fields" (also keep the final
    (p, u) of each request); staging_bytes: pinned host ring for the uploads of admitted cases; result_bytes: pinned
    host pool the finished cases' results are copied into; switch_interval: the interpreter's thread switch interval
    while serving (seconds)."""

    def __init__(self, storage="fp16", budget_bytes=None, chunk=200, arena_bricks=1 << 20, arena_records=1 << 18,
                 workers=None, keep="summary", max_cases=256, variant=0, device=None, staging_bytes=1 << 30,
                 result_bytes=512 << 20, switch_interval=2e-4, **solver_kw):
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
                             staging_bytes=staging_bytes, result_bytes=result_bytes, keep_arenas=True,
                             **solver_kw)
        # builder threads: one per CPU the process may use, less one for this loop, at most 10. More threads hold the
        # interpreter more often against this loop (96 requests, 16 CPUs: wall 6.17-6.26 s with 10 workers, 6.25-6.27 s
        # with 6, 6.77-6.83 s with 15)
        self.workers = int(workers) if workers else max(1, min(10, len(os.sched_getaffinity(0)) - 1))
        self.keep = keep
        self.switch_interval = float(switch_interval)

    def warm_up(self, make_case, spec):
        """Bring the server to its running state before requests are timed: one arena opened (``Batch.reserve``),
        and one request built, admitted, stepped for one chunk, read and removed, so that the kernels are loaded and
        the graph, staging and result paths have been used once. The request's results are discarded."""
        bt = self.batch
        bt.reserve()
        lat, nu, rec, lay = _build(make_case, spec)
        case = bt.add(lat, nu, layout=lay, packed=rec.packed)
        rec.replay(case)
        bt.run(self.chunk, chunk=self.chunk)
        read = case.read_async()
        bt.remove(case)
        read.result()
        bt.synchronize()

    def serve(self, requests, make_case, time_scale=1.0, log=None, warm=True):
        """Serve `requests` (dicts with "name", "arrival" in seconds, "steps") whose lattices come from
        make_case(spec) -> (lattice, nu, setup). Arrivals are replayed in real time (x time_scale). warm: first bring
        the server to its running state with the first request's case (``warm_up``, not timed), as a resident
        server is. Returns the report: per-request records and the aggregate throughput / latency figures, with
        where the host's time went (``host_s``) and the device time of the chunks (``gpu_chunk_s``)."""
        import torch
        bt, chunk = self.batch, self.chunk
        if warm and requests:
            self.warm_up(make_case, min(requests, key=lambda r: r["arrival"]))
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        building = {}
        built = []                                        # (arrival, name, spec, build result, time built)
        active = {}                                       # cid -> record
        reading = []                                      # finished cases whose results are on their way
        records = []
        lups = 0
        inflight = []                                     # (start, end) events of the chunks queued on the device
        chunks = []                                       # (start, end) events of every chunk
        host = {"admit": 0.0, "launch": 0.0, "wait": 0.0, "finish": 0.0, "captures": 0}
        t0 = time.perf_counter()

        def finish(rec):
            (p, u), ok = rec.pop("read").result()
            reading.remove(rec)
            rec["done"] = time.perf_counter() - t0
            rec["result"] = pool.submit(_summary, p, u, ok)   # summarised on a worker, off this loop
            if self.keep == "fields":
                rec["fields"] = (p, u)
            records.append(rec)
            if log:
                log(f"done {rec['name']} at {rec['done']:.2f} s (latency {rec['done'] - rec['arrival']:.2f} s)")

        def launch():
            """Queue one chunk for the admitted cases; the cases done at its end are read out behind it and leave the
            batch at once (stream-ordered), so no launch waits for the device."""
            nonlocal lups, t_one
            ng = bt.captures
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record(bt._tstream)
            bt.run(chunk, chunk=chunk)
            host["captures"] += bt.captures - ng
            cells = sum(rec["cells"] for rec in active.values())
            lups += chunk * cells
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(bt._tstream)
            chunks.append((e0, ev, time.perf_counter() - t0, len(active), cells))
            for cid in [c for c, rec in active.items() if rec["case"].steps >= rec["target"]]:
                rec = active.pop(cid)
                rec["read"] = rec["case"].read_async()
                rec["steps"] = rec["case"].steps
                bt.remove(rec["case"])
                del rec["case"]
                rec["left"] = time.perf_counter() - t0
                reading.append(rec)
            inflight.append((e0, ev))
            if len(inflight) == 1:
                t_one = time.perf_counter()

        def admit():
            _, _, r, (lat, nu, setup, lay), t_built = built.pop(0)
            case = bt.add(lat, nu, layout=lay, packed=setup.packed)
            setup.replay(case)
            target = -(-int(r["steps"]) // chunk) * chunk
            active[case.cid] = {"name": r["name"], "spec": r, "case": case, "target": target,
                                "arrival": r["arrival"] * time_scale, "built": t_built,
                                "admitted": time.perf_counter() - t0, "cells": case.N}
            if log:
                log(f"admit {r['name']} ({case.N} cells, {target} steps) at {active[case.cid]['admitted']:.2f} s, "
                    f"{len(active)} active")

        # One event-driven loop. It never lets the device run dry while cases are admitted: with no chunk queued it
        # launches at once; with one queued it admits built requests (in arrival order) for up to half that chunk's
        # expected time, so that they join the next chunk, then queues the next; with two queued it keeps admitting.
        # It collects results as their copies land and sleeps briefly only when none of this has work (the builder
        # threads then have the interpreter).
        last_chunk_s = 0.0                                # device time of the latest chunk known complete
        t_one = None                                      # when the queue last came down to one chunk
        # the loop gives up the interpreter often (device queries, short sleeps) and must get it back at once from
        # the builder threads: a short switch interval while serving (at the default 5 ms every hand-back can wait)
        switch = sys.getswitchinterval()
        sys.setswitchinterval(self.switch_interval)
        try:
            with cf.ThreadPoolExecutor(self.workers) as pool:
                while pending or building or built or active or reading:
                    now = time.perf_counter() - t0
                    while pending and pending[0]["arrival"] * time_scale <= now:
                        r = pending.popleft()
                        threads = max(1, self.workers // (len(building) + 1))
                        building[pool.submit(_build, make_case, r, threads)] = r
                    done_builds = [f for f in building if f.done()]
                    for fut in done_builds:
                        r = building.pop(fut)
                        built.append((r["arrival"], r["name"], r, fut.result(), time.perf_counter() - t0))
                    if done_builds:
                        built.sort(key=lambda x: (x[0], x[1]))
                    while inflight and inflight[0][1].query():
                        e0, e1 = inflight.pop(0)
                        last_chunk_s = e0.elapsed_time(e1) * 1e-3
                        if len(inflight) == 1:
                            t_one = time.perf_counter()
                    th = time.perf_counter()
                    ready = [rec for rec in reading if rec["read"].ready()]
                    for rec in ready:
                        finish(rec)
                    host["finish"] += time.perf_counter() - th
                    th = time.perf_counter()
                    can = bool(built) and bt.can_admit(built[0][3][3])
                    if active and (not inflight or (len(inflight) == 1 and
                                                    (not can or th - t_one > 0.5 * last_chunk_s))):
                        launch()
                        host["launch"] += time.perf_counter() - th
                        continue
                    if can:
                        admit()
                        host["admit"] += time.perf_counter() - th
                        continue
                    if built and not active and not reading and not inflight:
                        # the batch is empty and the head of the queue still does not fit
                        raise RuntimeError(f"request {built[0][1]} does not fit the HBM budget even alone")
                    if ready:
                        continue
                    # nothing to do now: wait a little for a chunk, a copy, a build or an arrival
                    if building and not inflight and not reading:
                        cf.wait(list(building), timeout=0.002, return_when=cf.FIRST_COMPLETED)
                    else:
                        time.sleep(0.0005)
                    host["wait"] += time.perf_counter() - th
        finally:
            sys.setswitchinterval(switch)
        for rec in records:
            rec["result"] = rec["result"].result()
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
