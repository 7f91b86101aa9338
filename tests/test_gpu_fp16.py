"""FP16 population storage (bricks layout): float16 deviations from the writing cell's equilibrium, float32
arithmetic. Bar: relative L2 of u within 1e-3 of the FP64 oracle run on the same lattice and step count; the FP32
storage error is printed beside it."""
import os

import numpy as np
import pytest

import oracle
from cerebrosim_amd import lbm
from cerebrosim_amd import lattice as L
from cerebrosim_amd.geometry import open_pipe, poiseuille_box

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL_U = 1e-3


def rel(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor, 1e-300))


def run_pair(lat, nu, steps, setup, **kw):
    """(fp16 solver, fp32 solver, fp64 oracle) after `steps` steps from rest with the same setup."""
    out = []
    for st in ("fp16", "fp32"):
        s = lbm.Solver(lat, nu, layout="bricks", storage=st, **kw)
        setup(s)
        s.run(steps)
        assert s.ok()
        out.append(s)
    o = oracle.OracleSolver(lat, nu, dtype=np.float64, **kw)
    setup(o)
    o.run(steps)
    return out[0], out[1], o


def test_fp16_poiseuille_box_vs_fp64():
    """u ~ 2.5e-4: the populations' deviations from w_i are ~1e-5 (float16-subnormal as f - w); the stored
    non-equilibrium part keeps them normal."""
    lat, r, R = poiseuille_box()
    nu = (0.8 - 0.5) / 3
    h, s, o = run_pair(lat, nu, 10000, lambda x: x.set_force([0, 0, 1e-6]), model="regularised")
    uo = o.fields()[1]
    e16, e32 = rel(h.fields()[1], uo), rel(s.fields()[1], uo)
    print(f"poiseuille box: rel L2 u vs FP64: fp16 {e16:.3e}  fp32 {e32:.3e}", flush=True)
    assert e16 <= TOL_U
    assert e32 <= 1e-5


@pytest.mark.parametrize("model", ["regularised", "trt"])
def test_fp16_open_pipe_vs_fp64(model):
    """Tilted open pipe: Bouzidi walls, velocity inlet with ramp, pressure outlet, sponge; 600 steps."""
    lat, _, _, _ = open_pipe((1, 2, 2), 40, lambda s: np.full_like(np.asarray(s, float), 6.0))

    def setup(x):
        x.set_sponge(6, 0.55)
        x.set_drive([0.03, 0.0, 0.01], period=100, ramp=50)
        x.set_outlet_pressure([0.0])

    h, s, o = run_pair(lat, 0.01, 600, setup, model=model)
    (ph, uh), (ps, us), (po, uo) = h.fields(), s.fields(), o.fields()
    e16, e32 = rel(uh, uo), rel(us, uo)
    p16, p32 = rel(ph, po, floor=np.linalg.norm(uo)), rel(ps, po, floor=np.linalg.norm(uo))
    print(f"open pipe {model}: rel L2 u vs FP64 fp16 {e16:.3e} fp32 {e32:.3e}; p fp16 {p16:.3e} fp32 {p32:.3e}",
          flush=True)
    assert e16 <= TOL_U and p16 <= TOL_U


def test_fp16_state_roundtrip_bound():
    """set_f then get_f: each population comes back within one float16 rounding of its stored deviation."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(HERE, "golden", "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    g = np.load(os.path.join(HERE, "golden", "pipe_small.npz"))
    lat = mf.lattice()
    C = np.asarray(L.C, np.float64)
    W = np.asarray(L.W, np.float64)
    for parity in (0, 1):
        s = lbm.Solver(lat, mf.NU, model="regularised", layout="bricks", storage="fp16")
        s.cur = parity
        s.set_velocity(g["u0"])
        s.set_f(g["f0"])
        back = s.get_f().astype(np.float64)
        u = g["u0"].astype(np.float64)
        cu = C @ u.T                                           # (Q, N)
        rp = (-1.5 * (u * u).sum(1))[None, :]
        eq = cu * (3 + 4.5 * cu)
        dev = g["f0"] / W[:, None] - rp - eq
        # half an ulp of the float16 h = 1024 dev (normal or subnormal), plus float32 rounding of the encode/decode
        eps32 = 2e-7 * (np.abs(g["f0"]) / W[:, None] + np.abs(rp) + np.abs(eq))
        bound = W[:, None] * (np.abs(dev) * 2.0 ** -11 + 2.0 ** -25 / 1024 + eps32)
        err = np.abs(back - g["f0"])
        assert np.all(err <= bound * 1.01 + 1e-9), (parity, float((err / (bound + 1e-30)).max()))


def test_fp16_set_velocity_keeps_populations():
    lat, _, _, _ = open_pipe((0, 0, 1), 12, lambda s: np.full_like(np.asarray(s, float), 4.0))
    s = lbm.Solver(lat, 0.05, model="regularised", layout="bricks", storage="fp16")
    s.set_drive([0.02], period=1)
    s.run(30)
    f0 = s.get_f()
    u = s.fields()[1] * 1.5
    s.set_velocity(u)
    assert np.array_equal(s.fields()[1], u.astype(np.float32))
    assert rel(s.get_f(), f0) <= 1e-3


def test_fp16_graph_replay_equals_eager():
    lat, _, _, _ = open_pipe((0, 0, 1), 12, lambda s: np.full_like(np.asarray(s, float), 4.0))
    runs = []
    for chunk in (10, 1000):
        s = lbm.Solver(lat, 0.05, model="regularised", layout="bricks", storage="fp16")
        s.set_drive([0.02, 0.0, 0.01], period=40)
        s.set_outlet_pressure([0.0])
        s.run(40, chunk=chunk)
        runs.append((s.get_f(), s.fields()[1], s.cur))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def test_fp16_needs_bricks():
    lat, _, _, _ = open_pipe((0, 0, 1), 12, lambda s: np.full_like(np.asarray(s, float), 4.0))
    with pytest.raises(ValueError, match="fp16"):
        lbm.Solver(lat, 0.05, layout="indexed", storage="fp16")
    with pytest.raises(ValueError, match="storage"):
        lbm.Solver(lat, 0.05, storage="bf16")
    s = lbm.Solver(lat, 0.05, storage="fp16")
    assert s.layout == "bricks" and s.f.dtype.itemsize == 2


@pytest.mark.parametrize("scale,steps", [(0.15, 800), (0.3, 1600)])
def test_precision_bifurcation_u_p_wss_vs_fp64(scale, steps):
    """The precision-form workload at 0.15 scale (48,593 cells, 800 steps) and at 0.3 scale (one inlet period of
    1600 steps): stenosed bifurcation, pulsatile 8-harmonic inlet (peak 0.08, throat speed ~0.22), tau 0.52. FP16 and
    FP32 storage against the FP64 oracle on velocity, pressure and wall shear stress (device cs_wss vs the numpy
    restatement on the FP64 field)."""
    from oracle.wss import shear
    from cerebrosim_amd import wss, workloads
    lat, nu, setup, info = workloads.bifurcation_case(scale, period=steps)
    idx, c, normals, _ = workloads.wall_shear_weights(lat, info)
    print(f"bifurcation scale {scale}: {lat['nbr'].shape[1]} cells, {steps} steps", flush=True)
    o = oracle.OracleSolver(lat, nu, model="regularised", dtype=np.float64)
    setup(o)
    o.run(steps)
    po, uo = o.fields()
    to = shear(uo, normals, idx, c)
    errs = {}
    for st in ("fp32", "fp16"):
        s = lbm.Solver(lat, nu, model="regularised", layout="bricks", storage=st)
        setup(s)
        s.run(steps)
        assert s.ok()
        p, u = s.fields()
        ws = wss.WallShear(s, idx, c, normals)
        tau = ws.frame(accumulate=False).cpu().numpy()
        errs[st] = {"u": rel(u, uo), "p": rel(p, po), "wss": rel(tau, to)}
    print(f"bifurcation scale {scale} vs FP64: {errs}; |u|max {np.abs(uo).max():.3f}", flush=True)
    assert errs["fp32"]["u"] <= 1e-5 and errs["fp32"]["p"] <= 1e-5 and errs["fp32"]["wss"] <= 1e-5
    assert errs["fp16"]["u"] <= TOL_U and errs["fp16"]["p"] <= TOL_U and errs["fp16"]["wss"] <= TOL_U
