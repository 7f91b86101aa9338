"""Brick layout on the CPU: builder invariants, and the AA-pattern design (parity forms, boundary redirects, side
copies, conversions) checked against the FP64 oracle through a numpy model of the stepping (tests/aa_model.py)."""
import numpy as np
import pytest

import oracle
from aa_model import AAModel
from cerebrosim_amd import bricks
from cerebrosim_amd import lattice as L
from cerebrosim_amd.geometry import open_pipe, periodic_pipe, poiseuille_box


def tilted_pipe():
    lat, _, _, _ = open_pipe((1, 2, 2), 10, lambda s: np.full_like(np.asarray(s, float), 3.0))
    return lat


def test_builder_invariants():
    lat = tilted_pipe()
    lay = bricks.build(lat)
    N = lat["nbr"].shape[1]
    slot = lay["slot_of_cell"]
    assert len(np.unique(slot)) == N
    assert (lay["cell_of_slot"][slot] == np.arange(N)).all()
    assert (lay["cell_of_slot"] >= 0).sum() == N
    # brick coordinates and local positions reproduce the cell coordinates
    b, l = slot // 64, slot % 64
    pos = lay["coords"][b] * 4 + np.stack([l % 4, (l // 4) % 4, l // 16], 1)
    assert (pos == lat["coords"]).all()
    # every fluid link is the implicit neighbour through the brick table
    for i in range(27):
        m = lat["nbr"][i] >= 0
        p = lat["coords"][m] - L.C[i]
        bp = np.floor_divide(p, 4)
        o = bp - lay["coords"][b[m]]
        d = (o[:, 0] + 1) + 3 * (o[:, 1] + 1) + 9 * (o[:, 2] + 1)
        nbk = lay["nbr"][b[m], d]
        lp = p - bp * 4
        assert (nbk * 64 + lp[:, 0] + 4 * lp[:, 1] + 16 * lp[:, 2] == slot[lat["nbr"][i][m]]).all()
    # boundary masks agree with the neighbour table
    bidx = lay["bidx"][b]
    bm = bidx >= 0
    lm = lay["lmask"][bidx[bm], l[bm]]
    for i in range(27):
        assert (((lm >> i) & 1).astype(bool) == (lat["nbr"][i][bm] >= 0)).all()
    inl = (lat["nbr"] == L.INLET).any(0)
    assert ((((lm >> 27) & 1).astype(bool)) == inl[bm]).all()
    assert not inl[~bm].any()
    # interior bricks: every cell fluid, every link fluid
    for bk in lay["interior"]:
        cells = lay["cell_of_slot"][bk * 64:(bk + 1) * 64]
        assert (cells >= 0).all() and (lat["nbr"][:, cells] >= 0).all()


def test_periodic_box_wraps():
    lat, r, R = poiseuille_box()
    lay = bricks.build(lat)
    assert lay["periodic"] == [0, 0, 1]
    assert lay["n_cells"] == 20160


def test_short_periodic_axis_is_rejected():
    lat, _, _ = periodic_pipe(10)           # 2 nodes along the periodic axis: not a multiple of the brick edge
    with pytest.raises(bricks.LayoutError):
        bricks.build(lat)


def test_mismatched_table_is_rejected():
    lat = tilted_pipe()
    nbr = lat["nbr"].copy()
    n = np.flatnonzero(nbr[1] >= 0)[0]
    nbr[1, n] = nbr[1, np.flatnonzero(nbr[1] >= 0)[5]]
    bad = dict(lat)
    bad["nbr"] = nbr
    with pytest.raises(bricks.LayoutError):
        bricks.build(bad)


@pytest.mark.parametrize("par", [0, 1])
def test_conversion_roundtrip(par):
    lat = tilted_pipe()
    lay = bricks.build(lat)
    m = AAModel(lat, lay, 0.05)
    f = np.random.default_rng(1).standard_normal((27, m.N))
    m.set_f(f, par)
    assert np.array_equal(m.get_f(par), f)


@pytest.mark.parametrize("start_par", [0, 1])
def test_aa_stepping_matches_oracle(start_par):
    lat = tilted_pipe()
    lay = bricks.build(lat)
    nu, force, p_out = 0.05, (1e-5, 2e-5, -1e-5), [0.001]
    rng = np.random.default_rng(7)
    f0 = 1e-3 * rng.standard_normal((27, lat["nbr"].shape[1]))
    u0 = 0.02 * rng.standard_normal((lat["nbr"].shape[1], 3))
    o = oracle.OracleSolver(lat, nu, model="regularised", dtype=np.float64, threads=1)
    o.set_force(force)
    o.set_outlet_pressure(p_out)
    o.f[o.cur][:] = f0
    o.vels[o.cur][:] = u0
    m = AAModel(lat, lay, nu, force=np.asarray(force, np.float32).astype(float), p_out=p_out)
    m.set_f(f0, start_par)
    m.vel = u0.copy()
    par = start_par
    for step in range(6):
        o.step_once(1.0)
        m.step(par)
        par = 1 - par
        ref = o.get_f()
        got = m.get_f(par)
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()) * 100, step
        assert np.allclose(m.vel, o.vels[o.cur], atol=1e-15)
