"""Batching form on the GPU: several lattices packed in brick arenas and stepped by shared launches. Each case must
give the same bits alone (lbm.Solver / artery.ArterySolver on the bricks layout) and in a mix whose membership
changes while it runs; every launch variant must give the same bits; the autotuner must leave the state as it
found it."""
import numpy as np
import pytest

from cerebrosim_amd import artery, lbm
from cerebrosim_amd import batch as B
from cerebrosim_amd.geometry import open_pipe, poiseuille_box

pytestmark = pytest.mark.gpu


def pipe(axis, L, R):
    lat, _, _, _ = open_pipe(axis, L, lambda s: np.full_like(np.asarray(s, float), R))
    return lat


def setup_a(x):
    x.set_sponge(6, 0.55)
    x.set_drive([0.03, 0.0, 0.01], period=100, ramp=50)
    x.set_outlet_pressure([0.0])


def setup_b(x):
    x.set_drive([0.02], period=1, ramp=30)
    x.set_outlet_pressure([0.001])


def setup_c(x):
    x.set_force([0, 0, 1e-5])


def cases():
    return [("A", pipe((1, 2, 2), 30, 5.0), 0.02, setup_a), ("B", pipe((0, 0, 1), 24, 6.0), 0.05, setup_b),
            ("C", poiseuille_box(R=6.0, nx=16, ny=16, nz=16)[0], 0.1, setup_c)]


def alone(lat, nu, setup, steps, storage):
    s = lbm.Solver(lat, nu, model="regularised", layout="bricks", storage=storage)
    setup(s)
    s.run(steps, chunk=40)
    return s.get_f(), s.fields()


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_case_alone_equals_in_mix(storage):
    (na, la, nua, sa), (nb, lb, nub, sb), (nc, lc, nuc, sc) = cases()
    bt = B.Batch(storage=storage, arena_bricks=600, arena_records=400)
    b = bt.add(lb, nub)
    sb(b)
    bt.run(30, chunk=10)                       # B alone
    a = bt.add(la, nua)                        # A joins at even parity, as a solver of its own starts
    sa(a)
    bt.run(51, chunk=10)
    c = bt.add(lc, nuc)                        # C joins at odd parity; with A and B present it opens a second arena
    sc(c)
    bt.run(40, chunk=20)
    bt.remove(b)                               # B leaves while A and C run on
    bt.run(109, chunk=40)
    assert a.steps == 200 and c.steps == 149 and len(bt.arenas) >= 1
    for name, h, lat, nu, setup in (("A", a, la, nua, sa), ("C", c, lc, nuc, sc)):
        f0, (p0, u0) = alone(lat, nu, setup, h.steps, storage)
        p1, u1 = h.fields()
        f1 = h.get_f()
        err = float(np.linalg.norm(u1 - u0) / np.linalg.norm(u0))
        print(f"{storage} {name}: max |du| {np.abs(u1 - u0).max():.3e} max |df| {np.abs(f1 - f0).max():.3e} "
              f"rel L2 u {err:.3e}", flush=True)
        # A runs the parities it would alone; C runs them shifted by one (the even and odd kernels do the same
        # arithmetic): both give the same bits
        assert np.array_equal(u1, u0) and np.array_equal(p1, p0) and np.array_equal(f1, f0), (name, err)
        assert h.ok()
    assert np.abs(a.fields()[1]).max() > 1e-3


def test_variants_agree_and_autotune_restores():
    (na, la, nua, sa), (nb, lb, nub, sb), _ = cases()
    ref = None
    for v in range(len(B.Batch(storage="fp32", arena_bricks=64, arena_records=64).variants())):
        bt = B.Batch(storage="fp32", arena_bricks=800, arena_records=400, variant=v)
        a, b = bt.add(la, nua), bt.add(lb, nub)
        sa(a)
        sb(b)
        bt.run(30, chunk=10)
        got = (a.get_f(), b.get_f())
        if ref is None:
            ref = got
        assert all(np.array_equal(x, y) for x, y in zip(got, ref)), v
    bt = B.Batch(storage="fp32", arena_bricks=800, arena_records=400)
    a, b = bt.add(la, nua), bt.add(lb, nub)
    sa(a)
    sb(b)
    bt.run(10, chunk=10)
    times = bt.autotune(steps=4, repeats=1)
    assert set(times) == set(range(len(bt.variants()))) and bt.variant in times
    bt.run(20, chunk=10)
    assert all(np.array_equal(x, y) for x, y in zip((a.get_f(), b.get_f()), ref))


def test_flow_measurement_equals_artery_solver():
    """Resistance outlets in the batch (cs_flux_cases) give ArterySolver's bits (cs_flux_bc on the case alone)."""
    from cerebrosim_amd.geometry import OFFSET
    axis = np.array([0.0, 0.0, 1.0])
    Lp, R = 28.0, 5.0
    base, _, _, _ = open_pipe(axis, Lp, lambda s: np.full_like(np.asarray(s, float), R))
    openings = [dict(o, kind="outlet") for o in base["openings"]]
    from cerebrosim_amd import lattice as L
    ext = R + 3
    lo = np.array([-ext, -ext, -ext])
    n = np.ceil(np.array([2 * ext, 2 * ext, Lp + 2 * ext])).astype(int) + 1
    A = -lo + OFFSET
    G = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) - A
    s = G @ axis
    rad = np.linalg.norm(G - s[:, None] * axis, axis=1)
    phi = np.maximum(np.maximum(rad - R, -s), s - Lp).reshape(n)
    lat = L.build_lattice(phi, [0, 0, 0], 1.0, openings=openings)
    flux = artery.flux_slabs(lat, openings, 1.0, slab=(3.0, 5.0))
    R_lu, pref, alpha, nu = [0.0, 0.01], 0.002, 0.05, 0.02
    coef = [0.004, 0.001, 0.0005]

    def setup(x):
        x.set_outlet_pressure([0.0, -pref])
        x.set_sponge(6, 0.55, kinds=("outlet",))
        x.set_drive(coef, period=40, ramp=10)

    s = artery.ArterySolver(lat, nu, flux, R_lu, pref, alpha, model="regularised", layout="bricks")
    setup(s)
    s.run(80, chunk=20)
    bt = B.Batch(storage="fp32", arena_bricks=400, arena_records=300)
    other = bt.add(pipe((0, 0, 1), 16, 4.0), 0.05)
    other.set_drive([0.01])
    c = bt.add(lat, nu)
    c.set_outlets(flux, R_lu, pref, alpha)
    setup(c)
    bt.run(80, chunk=20)
    assert np.array_equal(c.flows(), s.flows()), (c.flows(), s.flows())
    assert np.array_equal(c.opening_pressures()[:2], s.opening_pressures()[:2])
    assert np.array_equal(c.fields()[1], s.fields()[1])
    assert abs(s.flows()[1]) > 0


def test_admission_respects_budget():
    lat = pipe((0, 0, 1), 16, 4.0)
    from cerebrosim_amd import bricks
    lay = bricks.build(lat)
    bt = B.Batch(storage="fp16", arena_bricks=lay["n_bricks"], arena_records=lay["n_boundary"])
    need = bt.arena_bytes_for(lay)
    bt.budget = int(2.5 * need)
    c1 = bt.add(lat, 0.05, layout=lay)
    c2 = bt.add(lat, 0.05, layout=lay)
    assert len(bt.arenas) == 2 and not bt.can_admit(lay)
    with pytest.raises(B.BatchFull):
        bt.add(lat, 0.05, layout=lay)
    bt.remove(c1)
    assert bt.can_admit(lay)
    c3 = bt.add(lat, 0.05, layout=lay)
    c2.set_drive([0.01])
    c3.set_drive([0.01])
    bt.run(20)
    assert c2.steps == c3.steps == 20
    assert np.array_equal(c2.fields()[1], c3.fields()[1])         # the same case twice, in two arenas
    assert c2.ok() and c3.ok()
