"""Case generators of the batching and serving forms (host only)."""
import numpy as np

from cerebrosim_amd import workloads as W


def test_tube_specs_are_seeded_and_in_range():
    a, b = W.tube_specs(96, 1), W.tube_specs(96, 1)
    assert a == b and len(a) == 96
    for s in a:
        assert 8 <= s["R"] <= 48 and 128 <= s["L"] <= 960 and 0 <= s["area_frac"] <= 0.95 and 0.51 <= s["tau"] <= 0.9
        nu = (s["tau"] - 0.5) / 3
        assert s["u_mean"] <= 0.03 + 1e-12 and s["u_mean"] <= 0.08 * (1 - s["area_frac"]) + 1e-12
        assert 2 * s["R"] * s["u_mean"] / nu <= 400 + 1e-9
    cells = sum(np.pi * s["R"] ** 2 * s["L"] for s in a)
    assert 1.0e8 < cells < 2.0e8


def test_stenosed_tube_throat_area():
    R, L, af = 10.0, 80.0, 0.75
    lat = W.stenosed_tube(R, L, af)
    z = lat["coords"][:, 2] - lat["coords"][:, 2].min()
    counts = np.bincount(z)
    inlet = counts[2:6].mean()
    throat = counts[len(counts) // 2 - 2:len(counts) // 2 + 3].min()
    assert abs(throat / inlet - (1 - af)) < 0.08, (throat / inlet)
    assert (lat["nbr"] == -2).any() and (lat["nbr"] <= -3).any()


def test_request_specs():
    r = W.request_specs(200, rate_per_s=2.0, seed=2)
    t = np.array([x["arrival"] for x in r])
    assert np.all(np.diff(t) > 0) and abs(np.mean(np.diff(t)) - 0.5) < 0.1
    assert {x["dx"] for x in r} == {0.2e-3, 0.1e-3, 0.05e-3}
    for x in r:
        assert 70 <= x["map_mmhg"] <= 145 and 0.6 <= x["dt_scale"] <= 1.0 and x["steps"] == 4 * x["period"]
    assert W.dt_scale(95) == 1.0 and W.dt_scale(118) == 0.75 and W.dt_scale(146) == 0.6 and W.dt_scale(70) == 1.0
