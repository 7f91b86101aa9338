"""GPU parity of the simulate.py counterpart: per-step flow measurement and opening-pressure update captured in the
graphs (k_flux + k_bc), and the probe reduction (k_gather), against the oracle stepped with the same sequence."""
import numpy as np
import pytest

import oracle
from cerebrosim_amd import artery
from cerebrosim_amd import lattice as L
from cerebrosim_amd.geometry import open_pipe

pytestmark = pytest.mark.gpu

LAYOUTS = ["indexed", "bricks"]




def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tilted_outlet_pipe():
    """A tilted pipe with two pressure openings, its flow slabs and the resistance coupling (and the grid offset A
    and the axis, for probes)."""
    from cerebrosim_amd.geometry import OFFSET
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    Lp, R = 36.0, 5.0
    base, _, _, _ = open_pipe(axis, Lp, lambda s: np.full_like(np.asarray(s, float), R))
    openings = [dict(o, kind="outlet") for o in base["openings"]]
    # phi is not kept by build_lattice: rebuild the same grid for the outlet-only classification
    ext = R * np.sqrt(np.clip(1 - axis ** 2, 0, 1)) + 3
    ends = np.array([np.zeros(3), axis * Lp])
    lo = ends.min(0) - ext
    n = np.ceil(ends.max(0) + ext - lo).astype(int) + 1
    A = -lo + OFFSET
    G = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) - A
    s = G @ axis
    rad = np.linalg.norm(G - s[:, None] * axis, axis=1)
    phi = np.maximum(np.maximum(rad - R, -s), s - Lp).reshape(n)
    lat = L.build_lattice(phi, [0, 0, 0], 1.0, openings=openings)
    assert lat["n_outlets"] == 2 and lat["counts"]["inlet"] == 0
    flux = artery.flux_slabs(lat, openings, 1.0, slab=(3.0, 5.0))
    R_lu, pref, alpha = [0.0, 0.01], 0.002, 0.05        # stable explicit coupling (R Q of the order of pref)
    return lat, flux, R_lu, pref, alpha, A, axis


@pytest.mark.parametrize("layout", LAYOUTS)
def test_artery_solver_matches_oracle(layout):
    lat, flux, R_lu, pref, alpha, A, axis = tilted_outlet_pipe()
    coef = [0.004, 0.001, 0.0005]
    nu = 0.02
    s = artery.ArterySolver(lat, nu, flux, R_lu, pref, alpha, model="regularised", layout=layout)
    o = oracle.OracleSolver(lat, nu, model="regularised")
    for x in (s, o):
        x.set_outlet_pressure([0.0, -pref])
        x.set_sponge(6, 0.55, kinds=("outlet",))
        x.set_drive(coef, period=40, ramp=10)
    steps = 80
    s.run(steps, chunk=20)
    Qacc = np.zeros(2, np.float32)
    q_last = np.zeros(2, np.float32)
    Rv = np.asarray(R_lu, np.float32)
    for _ in range(steps):
        o.run(1)
        oracle.flux_bc(flux["idx"], flux["w"], flux["slot"], o.vels[o.cur], Qacc, Rv, pref, o.drv, 2, alpha, o.p_out,
                       q_last)
    errs = {"f": rel(s.get_f(), o.get_f()), "u": rel(s.fields()[1], o.fields()[1]),
            "Q": rel(s.flows(), q_last), "p_out": rel(s.opening_pressures(), o.p_out)}
    print(f"{layout} artery: {errs} Q {s.flows()} vs {q_last} p {s.opening_pressures()} vs {o.p_out}", flush=True)
    assert all(v <= 1e-5 for v in errs.values()), errs
    assert abs(q_last[0]) > 0 and abs(q_last[1]) > 0

    # probes: three balls of cells
    P = lat["coords"].astype(float)
    centres = [A + axis * 10, A + axis * 18, A + axis * 26]
    gidx, gslot = [], []
    for k, c in enumerate(centres):
        m = np.flatnonzero(np.linalg.norm(P - c, axis=1) < 3.0)
        gidx.append(m)
        gslot.append(np.full(len(m), k))
    gidx = np.concatenate(gidx).astype(np.int32)
    gslot = np.concatenate(gslot).astype(np.int32)
    out, umax = s.gather(gidx, gslot, 3)
    po, uo = o.fields()
    expect = np.array([(o.rho[gidx[gslot == k]] / 3.0).astype(np.float64).sum() for k in range(3)])
    assert np.allclose(out, expect, rtol=1e-4, atol=1e-9)
    assert abs(umax - np.linalg.norm(uo[gidx], axis=1).max()) <= 1e-5 * umax


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_autotune_restores_the_state(storage):
    """ArterySolver.autotune times every launch variant and then puts back what its timing steps changed: the
    populations, the moments, the clock, the opening pressures and the last flows are bitwise those before it, and
    the steps after it give the bits of a run that never autotuned."""
    import torch
    lat, flux, R_lu, pref, alpha, _, _ = tilted_outlet_pipe()
    runs = []
    for tune in (False, True):
        s = artery.ArterySolver(lat, 0.02, flux, R_lu, pref, alpha, model="regularised", layout="bricks",
                                storage=storage)
        s.set_outlet_pressure([0.0, -pref])
        s.set_drive([0.004, 0.001, 0.0005], period=40, ramp=10)
        s.run(21, chunk=10)
        s.synchronize()
        before = (s.get_f(), s.opening_pressures(), s.flows(), s.tick.cpu().numpy().copy(), s.cur, s.steps)
        if tune:
            times = s.autotune(steps=4, repeats=1)
            assert set(times) and s.variant in times
            after = (s.get_f(), s.opening_pressures(), s.flows(), s.tick.cpu().numpy().copy(), s.cur, s.steps)
            for x, y in zip(before[:4], after[:4]):
                assert np.array_equal(x, y)
            assert before[4:] == after[4:]
        s.run(30, chunk=10)
        s.synchronize()
        runs.append((s.get_f(), s.opening_pressures(), s.flows()))
        del s
        torch.cuda.empty_cache()
    for x, y in zip(*runs):
        assert np.array_equal(x, y)
