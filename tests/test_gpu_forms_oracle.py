"""The batching and serving workloads against the oracle, on reduced samples of the seeded case sets:

* batching: tubes of the seed-1 set (radius U[8, 48], length U[128, 960], area reduction U[0, 0.95], tau
  U[0.51, 0.9]), scaled to a third so that the FP64 oracle runs them in seconds, packed in one batch and stepped
  together, in FP32 and in FP16 storage;
* serving: requests of the seed-2 stream (stenosis severity, pressure, grid) on the 0.2 mm grid with cardiac cycles
  compressed to 60 steps, served with continuous batching in FP32 and in FP16 storage.

Per case: FP32 storage within 1e-5 (relative L2 of f, rho, u) of the FP32 oracle after the same steps; FP16 storage
within 1e-3 (relative L2 of u) of the FP64 oracle. The measured errors are printed and, when CEREBRO_PARITY_OUT
names a file, written there as JSON."""
import json
import os

import numpy as np
import pytest

import oracle
from cerebrosim_amd import batch as B
from cerebrosim_amd import serving, workloads

pytestmark = pytest.mark.gpu

RESULTS = {}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _record(key, value):
    RESULTS[key] = value
    path = os.environ.get("CEREBRO_PARITY_OUT")
    if path:
        with open(path, "w") as fh:
            json.dump(RESULTS, fh, indent=1)


def tube_sample(n=5, scale=1.0 / 3.0):
    out = []
    for sp in workloads.tube_specs(96, seed=1)[:n]:
        R, L = max(4.0, sp["R"] * scale), max(24.0, sp["L"] * scale)
        out.append(dict(sp, R=R, L=L, u_mean=workloads.inlet_speed(R, sp["area_frac"], sp["tau"])))
    return out


def run_oracle(lat, nu, setup, steps, dtype):
    o = oracle.OracleSolver(lat, nu, model="regularised", dtype=dtype)
    setup(o)
    o.run(steps)
    return o


def test_batched_tube_sample_against_oracle():
    specs = tube_sample()
    steps = 300
    cases = [workloads.tube_case(sp) for sp in specs]
    cells = [c[0]["nbr"].shape[1] for c in cases]
    print(f"batching sample: {len(cases)} tubes, cells {cells}", flush=True)
    got = {}
    for storage in ("fp32", "fp16"):
        bt = B.Batch(storage=storage, arena_bricks=8192, arena_records=8192)
        hs = [bt.add(lat, nu) for lat, nu, setup in cases]
        for h, (lat, nu, setup) in zip(hs, cases):
            setup(h)
        bt.run(steps, chunk=100)
        got[storage] = [(h.get_f(), h.fields()) for h in hs]
        assert all(h.ok() for h in hs)
        del hs, bt
    errs = []
    for k, (lat, nu, setup) in enumerate(cases):
        o32 = run_oracle(lat, nu, setup, steps, np.float32)
        o64 = run_oracle(lat, nu, setup, steps, np.float64)
        f32, (p32, u32) = got["fp32"][k]
        _, (p16, u16) = got["fp16"][k]
        e = {"cells": cells[k], "fp32_vs_oracle_fp32": {"f": rel(f32, o32.get_f()), "rho": rel(p32, o32.fields()[0]),
                                                          "u": rel(u32, o32.fields()[1])},
             "fp16_vs_oracle_fp64": {"u": rel(u16, o64.fields()[1]), "p": rel(p16, o64.fields()[0])},
             "fp32_vs_oracle_fp64": {"u": rel(u32, o64.fields()[1])}}
        print(f"tube {k}: {e}", flush=True)
        errs.append(e)
        # rho is a small deviation in these open tubes; f and u carry the state
        assert max(e["fp32_vs_oracle_fp32"].values()) <= 1e-5, e
        assert e["fp16_vs_oracle_fp64"]["u"] <= 1e-3, e
    _record("batching_tubes_seed1_third_size", {"steps": steps, "cases": errs,
                                                 "max_fp16_u_vs_fp64": max(e["fp16_vs_oracle_fp64"]["u"] for e in errs)})


def request_sample(n=4):
    reqs = workloads.request_specs(n, rate_per_s=50.0, seed=2, beats=4, beat_steps=60, grids=(0.2e-3,))
    return reqs


def test_served_request_sample_against_oracle():
    reqs = request_sample()
    chunk = 50
    out = {}
    for storage in ("fp32", "fp16"):
        srv = serving.Server(storage=storage, chunk=chunk, arena_bricks=1 << 15, arena_records=1 << 14,
                             keep="fields")
        rep = srv.serve(reqs, workloads.request_case)
        out[storage] = {r["name"]: r for r in rep["records"]}
        assert rep["cases"] == len(reqs)
        del srv
    errs = []
    for r in reqs:
        lat, nu, setup = workloads.request_case(r)
        steps = out["fp16"][r["name"]]["steps"]
        assert out["fp32"][r["name"]]["steps"] == steps
        o32 = run_oracle(lat, nu, setup, steps, np.float32)
        o64 = run_oracle(lat, nu, setup, steps, np.float64)
        p32, u32 = out["fp32"][r["name"]]["fields"]
        p16, u16 = out["fp16"][r["name"]]["fields"]
        e = {"name": r["name"], "cells": int(lat["nbr"].shape[1]), "steps": int(steps),
             "fp32_vs_oracle_fp32": {"rho": rel(p32, o32.fields()[0]), "u": rel(u32, o32.fields()[1])},
             "fp16_vs_oracle_fp64": {"u": rel(u16, o64.fields()[1]), "p": rel(p16, o64.fields()[0])}}
        print(f"request {r['name']}: {e}", flush=True)
        errs.append(e)
        assert max(e["fp32_vs_oracle_fp32"].values()) <= 1e-5, e
        assert e["fp16_vs_oracle_fp64"]["u"] <= 1e-3, e
    _record("serving_requests_seed2_0.2mm_beat60", {"cases": errs,
                                                     "max_fp16_u_vs_fp64": max(e["fp16_vs_oracle_fp64"]["u"]
                                                                               for e in errs)})
