"""The validate.py counterpart on the GPU: the quick subset (Poiseuille D = 10 interpolated and D = 20 half-way,
Womersley at Wo 5.6) and test 4's stenosis at narrow70 Re 100, against the reference's published results
(geometry/solver.json:20-68, 114-137)."""
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


def test_stenosis_narrow70_re100_matches_published():
    """validate.py test 4 (validate.py:218-291): the 70 % cosine stenosis at narrow70, Re 100 (144,678 cells) runs to
    the same convergence step as the reference and reproduces its published pressure drop, 3556.4 Pa
    (geometry/solver.json:114-137)."""
    from cerebrosim_amd import validate
    r = validate.stenosis("narrow70", 100)
    print(f"stenosis narrow70 Re 100: {r['cells']} cells, {r['steps']} steps, dp {r['dp_Pa']:.2f} Pa "
          f"(published {r['published']['dp_Pa']:.2f}), {r['mlups']:.0f} MLUPS", flush=True)
    assert r["ok"] and r["cells"] == 144678
    assert r["steps"] == 289600
    assert abs(r["dp_Pa"] / r["published"]["dp_Pa"] - 1) < 1e-3
    assert abs(r["vs_young_tsai"] - r["published"]["vs_young_tsai"]) < 1e-3
