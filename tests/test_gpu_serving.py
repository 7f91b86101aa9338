"""Serving form on the GPU: requests arriving over time are admitted into the running batch (continuous batching)
under an HBM budget; each request's result is the same bits as the request solved alone."""
import numpy as np
import pytest

from cerebrosim_amd import serving, workloads

pytestmark = pytest.mark.gpu


def small_requests():
    specs = [dict(R=5.0, L=40.0, area_frac=0.5, tau=0.6), dict(R=7.0, L=24.0, area_frac=0.0, tau=0.8),
             dict(R=4.0, L=64.0, area_frac=0.8, tau=0.55), dict(R=6.0, L=32.0, area_frac=0.3, tau=0.7)]
    out = []
    for k, sp in enumerate(specs):
        sp = dict(sp, name=f"r{k}", arrival=0.04 * k, steps=200 + 100 * (k % 2),
                  u_mean=workloads.inlet_speed(sp["R"], sp["area_frac"], sp["tau"]))
        out.append(sp)
    return out


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_served_results_equal_sequential(storage):
    reqs = small_requests()
    srv = serving.Server(storage=storage, chunk=100, arena_bricks=300, arena_records=300, keep="fields")
    rep = srv.serve(reqs, workloads.tube_case)
    seq = serving.run_sequential(reqs, workloads.tube_case, storage=storage, chunk=100, keep="fields")
    assert rep["cases"] == len(reqs) == seq["cases"]
    got = {r["name"]: r for r in rep["records"]}
    for s in seq["records"]:
        g = got[s["name"]]
        assert g["steps"] == s["steps"]
        assert np.array_equal(g["fields"][1], s["fields"][1]) and np.array_equal(g["fields"][0], s["fields"][0]), \
            s["name"]
        assert g["result"]["ok"] and g["result"]["u_max"] > 1e-4
    assert 0 < rep["p50_s"] <= rep["p99_s"] and rep["cases_per_hour"] > 0
    print(f"{storage}: served {rep['cases_per_hour']:.0f} cases/h p50 {rep['p50_s']:.3f} s p99 {rep['p99_s']:.3f} s; "
          f"sequential {seq['cases_per_hour']:.0f} cases/h", flush=True)


def test_budget_queues_requests():
    """A budget that holds one request at a time: the others wait, all finish, results unchanged."""
    reqs = small_requests()[:3]
    from cerebrosim_amd import batch as B
    from cerebrosim_amd import bricks
    lays = [bricks.build(workloads.tube_case(r)[0]) for r in reqs]
    probe = B.Batch(storage="fp32", arena_bricks=1, arena_records=1)
    big = max(probe.arena_bytes_for(l) for l in lays)
    small = min(probe.arena_bytes_for(l) for l in lays)
    assert 1.2 * big < 2 * small
    srv = serving.Server(storage="fp32", chunk=100, arena_bricks=1, arena_records=1, budget_bytes=int(1.2 * big),
                         keep="fields")
    for r in reqs:
        r["arrival"] = 0.0
    rep = srv.serve(reqs, workloads.tube_case)
    recs = sorted(rep["records"], key=lambda r: r["admitted"])
    for a, b in zip(recs, recs[1:]):
        assert b["admitted"] >= a["left"] - 1e-6          # one at a time: admitted after the previous case left
    seq = serving.run_sequential(reqs, workloads.tube_case, storage="fp32", chunk=100, keep="fields")
    got = {r["name"]: r for r in rep["records"]}
    for s in seq["records"]:
        assert np.array_equal(got[s["name"]]["fields"][1], s["fields"][1])


def test_fine_grid_request_shares_the_budget():
    """A 0.05 mm request (about 5 M cells) arrives with 0.2 / 0.1 mm ones under a budget that holds it and one
    0.1 mm request: the arena bytes of the cases in the batch never exceed the budget, every request finishes, and
    each result is the same bits as the request solved alone (FP16 storage)."""
    from cerebrosim_amd import batch as B
    from cerebrosim_amd import bricks
    reqs = workloads.request_specs(4, rate_per_s=50.0, seed=2, beats=1, beat_steps=40, grids=(0.2e-3, 0.1e-3))
    fine = workloads.request_specs(1, rate_per_s=50.0, seed=5, beats=1, beat_steps=40, grids=(0.05e-3,))[0]
    reqs.insert(1, dict(fine, name="fine"))
    for r in reqs:
        r["arrival"] = 0.0
    print("requests:", [(r["name"], r["dx"], r["steps"]) for r in reqs], flush=True)
    probe = B.Batch(storage="fp16", arena_bricks=1, arena_records=1)
    need = {r["name"]: probe.arena_bytes_for(bricks.build(workloads.request_case(r)[0])) for r in reqs}
    del probe
    coarse = max(v for k, v in need.items() if k != "fine")
    budget = int(need["fine"] + 1.1 * coarse)
    srv = serving.Server(storage="fp16", chunk=100, arena_bricks=1, arena_records=1, budget_bytes=budget,
                         keep="fields")
    rep = srv.serve(reqs, workloads.request_case)
    assert rep["cases"] == len(reqs)
    recs = rep["records"]
    assert max(r["cells"] for r in recs) > 4_000_000
    for r in recs:
        held = sum(need[x["name"]] for x in recs if x["admitted"] <= r["admitted"] < x["left"])
        assert held <= budget, (r["name"], held, budget)
    seq = serving.run_sequential(reqs, workloads.request_case, storage="fp16", chunk=100, keep="fields")
    got = {r["name"]: r for r in recs}
    for s in seq["records"]:
        g = got[s["name"]]
        assert g["steps"] == s["steps"] and g["result"]["ok"]
        assert np.array_equal(g["fields"][1], s["fields"][1]) and np.array_equal(g["fields"][0], s["fields"][0]), \
            s["name"]
