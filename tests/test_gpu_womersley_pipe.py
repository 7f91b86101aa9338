"""The pulsatile pipe at full size (R 40, L 480, 2,412,960 cells, plug velocity inlet, pressure outlet, tau 0.56)
against analytic solutions.

* Window: cross-sections at z = 360, 408, 456 (0.75 L, 0.85 L, 0.95 L), past the steady entrance length
  (about 0.06 Re D = 384 cells at the mean Re of 80).
* The pipe is weakly compressible (sound speed 1/sqrt(3)) and 480 cells long, so at a period of 4000 steps it is
  close to a quarter-wave resonance: the oscillating flow rate grows from 0.010 at the inlet to about 0.044 at the
  outlet. The flow-rate-driven Womersley solution therefore applies to each cross-section with that section's own
  flow rate. The comparison fits the flow-rate harmonics of each section over the last period and compares the
  radial profiles with ``womersley_flow_profile`` for those harmonics.
* The startup transient of the resonance takes about six periods to die out (the flow-rate fit residual falls
  from 1e-3 after three periods to 1e-4 after six), so the runs are eight periods long.
* The variation of the flow rate along the pipe is checked against Womersley's rigid-tube propagation constant
  (``womersley_line_flow``) in the linear regime (a drive amplitude of 0.001 and no mean flow), where that theory
  holds: at the 0.02 mean flow of the workload the oscillation is 10 % weaker than the linear theory, from its
  interaction with the developing mean flow.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R, L, TAU, T = 40.0, 480, 0.56, 4000
NU = (TAU - 0.5) / 3
WINDOW = (360, 408, 456)


@pytest.fixture(scope="module")
def pipe():
    from cerebrosim_amd.geometry import womersley_pipe
    lat, sc, rc = womersley_pipe(R, L)
    zi = np.round(sc - sc.min()).astype(np.int64)
    return lat, zi, rc


def _solver(lat, coef, period, storage="fp32"):
    from cerebrosim_amd import lbm
    s = lbm.Solver(lat, NU, model="regularised", layout="bricks", storage=storage)
    s.set_drive(coef, period=period)
    s.set_outlet_pressure([0.0])
    return s


def _harmonics(t, y, period, n):
    w = 2 * math.pi / period
    cols = [np.ones_like(t)]
    for k in range(1, n + 1):
        cols += [np.cos(k * w * t), np.sin(k * w * t)]
    X = np.stack(cols, 1)
    co, *_ = np.linalg.lstsq(X, y, rcond=None)
    return co, np.sqrt(((y - X @ co) ** 2).mean(0))


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_window_profiles_follow_the_local_flow_rate(pipe, storage):
    from cerebrosim_amd.validation import womersley_flow_profile
    lat, zi, rc = pipe
    s = _solver(lat, [0.02, 0.0, 0.01], T, storage)
    s.run(7 * T)
    cells = [np.flatnonzero(zi == z) for z in WINDOW]
    ts, prof = [], [[] for _ in WINDOW]
    while s.steps < 8 * T:
        s.run(100)
        _, u = s.fields()
        ts.append(s.steps - 0.5)                      # the clock runs before the step (validate.py:110)
        for k, c in enumerate(cells):
            prof[k].append(u[c, 2].astype(np.float64))
    assert s.ok()
    ts = np.array(ts)
    for z, c, U in zip(WINDOW, cells, prof):
        U = np.array(U)                               # (40 samples, cells of the section)
        co, res = _harmonics(ts, U.mean(1), T, 2)
        assert res < 3e-4, (z, res)                   # periodic state reached
        assert abs(co[0] - 0.02) < 2e-4, (z, co[0])   # mean flow = the inlet's mean (mass conservation)
        harm = [(co[0],), (co[1], co[2]), (co[3], co[4])]
        ua = np.stack([womersley_flow_profile(rc[c], R, NU, harm, T, t) for t in ts])
        err = np.sqrt(((U - ua) ** 2).sum() / (ua ** 2).sum())
        mean_err = np.sqrt(((U.mean(0) - ua.mean(0)) ** 2).sum() / (ua.mean(0) ** 2).sum())
        print(f"{storage} z={z}: |Q1| {math.hypot(co[1], co[2]):.5f}, profile rel L2 {err:.4f}, "
              f"period-mean profile {mean_err:.4f}", flush=True)
        assert err < 0.02, (z, err)
        assert mean_err < 0.02, (z, mean_err)


@pytest.mark.parametrize("period,periods,tol", [(4000, 8, 0.04), (8000, 5, 0.01)])
def test_linear_acoustics_along_the_pipe(pipe, period, periods, tol):
    from cerebrosim_amd.validation import womersley_line_flow
    lat, zi, _ = pipe
    cnt = np.bincount(zi)
    s = _solver(lat, [0.0, 0.0, 0.001], period)
    s.run((periods - 1) * period)
    ts, Q = [], []
    while s.steps < periods * period:
        s.run(25)
        _, u = s.fields()
        ts.append(s.steps - 0.5)
        Q.append(np.bincount(zi, weights=u[:, 2], minlength=len(cnt)) / cnt)
    assert s.ok()
    co, res = _harmonics(np.array(ts), np.array(Q), period, 1)
    assert res.max() < 1e-4, res.max()
    sim = (co[1] - 1j * co[2]) / (-0.001j)            # Q1(z) / Q1(0): the inlet imposes 0.001 sin(w t)
    assert abs(sim[0] - 1) < 1e-3
    for z in (100, 240) + WINDOW:
        th = complex(womersley_line_flow(z + 0.5, float(L), R, NU, period))
        rel = abs(sim[z] - th) / abs(th)
        print(f"T={period} z={z}: sim {abs(sim[z]):.3f}@{math.degrees(np.angle(sim[z])):.1f}, "
              f"theory {abs(th):.3f}@{math.degrees(np.angle(th)):.1f}, rel {rel:.4f}", flush=True)
        assert rel < tol, (z, rel)
