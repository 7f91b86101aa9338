"""The validate.py counterpart on the GPU (quick subset): Poiseuille D = 10 interpolated and D = 20 half-way, and
Womersley at Wo 5.6, against the reference's published results (geometry/solver.json:20-68)."""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_validate_quick_matches_published(tmp_path):
    from cerebrosim_amd import validate
    res = validate.run(quick=True, log=lambda m: print(m, flush=True))
    assert [(r["D"], r["bouzidi"]) for r in res["poiseuille"]] == [(10, True), (20, False)]
    for r in res["poiseuille"]:
        assert r["ok"]
        assert abs(r["l2"] / r["published"]["l2"] - 1) < 5e-3, r
        assert abs(r["flow_err"] - r["published"]["flow_err"]) < 5e-4, r
    (w,) = res["womersley"]
    assert w["ok"] and w["period_steps"] == w["published"]["period_steps"]
    assert abs(w["l2"] / w["published"]["l2"] - 1) < 1e-4, w
    with open(tmp_path / "solver.json", "w") as fh:
        json.dump(res, fh, default=float)
