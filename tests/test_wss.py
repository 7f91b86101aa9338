"""Wall shear stress: the fit method on the reference's known-answer pipes (wss.py:61-84: exact 2 mu u_max / R for
Poiseuille, 4 mu u_max / R for the blunt profile), and the device kernel against the numpy restatement of shear()."""
import ctypes

import numpy as np
import pytest

from cerebrosim_amd import _native, wss
from oracle.wss import shear


def pipe_case(R, pw, rng, h=0.1, umax=0.5):
    """Synthetic pipe fields of wss.py:65-78 (jittered 0.1 mm grid, wall vertices at random angles)."""
    g = np.arange(-R - h, R + h + 1e-9, h)
    x, y, z = np.meshgrid(g + rng.uniform(0, h), g + rng.uniform(0, h), np.arange(0, 3, h), indexing="ij")
    r = np.hypot(x, y)
    m = r < R
    cells = np.column_stack([x[m], y[m], z[m]])
    depth = R - r[m]
    U = np.zeros_like(cells)
    U[:, 2] = umax * (1 - (r[m] / R) ** pw)
    th = rng.uniform(0, 2 * np.pi, 400)
    verts = np.column_stack([R * np.cos(th), R * np.sin(th), rng.uniform(1, 2, 400)])
    normals = -np.column_stack([np.cos(th), np.sin(th), np.zeros_like(th)])
    return cells, depth, U, verts, normals, pw * wss.MU * umax / (R * 1e-3)


def test_known_answers():
    rng = np.random.default_rng(0)
    for pw in (2, 4):
        for R in (0.4, 0.65, 1.0, 1.5, 2.0):
            cells, depth, U, verts, normals, exact = pipe_case(R, pw, rng)
            idx, c, quad, any_ = wss.weights(verts, normals, cells, depth)
            t = np.linalg.norm(shear(U, normals, idx, c), axis=1)
            assert quad.all() and any_.all()
            err = t.mean() / exact - 1
            if pw == 2:      # the profile is exactly quadratic in the wall distance: the fit recovers it
                assert abs(err) < 1e-3 and t.std() / exact < 1e-3, (R, err)
            else:            # the quartic profile is resolved as R grows
                assert err < 0 and abs(err) < {0.4: 0.2, 0.65: 0.08, 1.0: 0.04, 1.5: 0.02, 2.0: 0.01}[R], (R, err)


@pytest.mark.gpu
def test_device_shear_and_time_averages():
    import torch
    rng = np.random.default_rng(0)
    cells, depth, U, verts, normals, exact = pipe_case(1.0, 4, rng)
    idx, c, quad, _ = wss.weights(verts, normals, cells, depth)
    dev = torch.device("cuda:0")
    V, K = idx.shape
    t_idx = torch.from_numpy(idx.astype(np.int32)).to(dev)
    t_c = torch.from_numpy(c).to(dev)
    t_n = torch.from_numpy(normals.astype(np.float32)).to(dev)
    tau = torch.zeros((V, 3), dtype=torch.float32, device=dev)
    acc_t = torch.zeros((V, 3), dtype=torch.float32, device=dev)
    acc_m = torch.zeros(V, dtype=torch.float32, device=dev)
    scales = (1.0, -1.0, 0.5)
    for sf in scales:
        Ud = torch.from_numpy((sf * U).astype(np.float32)).to(dev)
        A = _native.WssArgs()
        A.idx, A.c, A.normal, A.n_vert, A.k = t_idx.data_ptr(), t_c.data_ptr(), t_n.data_ptr(), V, K
        A.vel, A.vel_elem, A.vel_comp, A.scale = Ud.data_ptr(), 3, 1, wss.MU * 1e3
        A.tau, A.acc_tau, A.acc_mag = tau.data_ptr(), acc_t.data_ptr(), acc_m.data_ptr()
        _native.call("cs_wss", ctypes.byref(A), None)
        torch.cuda.synchronize()
        ref = shear((sf * U).astype(np.float32), normals.astype(np.float32), idx, c)
        got = tau.cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    n = len(scales)
    tawss = (acc_m / n).cpu().numpy()
    osi = 0.5 * (1 - (acc_t / n).norm(dim=1).cpu().numpy() / tawss)
    mag = np.linalg.norm(shear(U, normals, idx, c), axis=1)
    assert np.allclose(tawss, np.mean(np.abs(scales)) * mag, rtol=1e-5)
    # tau_f = s_f tau_0: OSI = (1 - |mean s| / mean |s|) / 2 = 0.4 for s = (1, -1, 0.5)
    assert np.allclose(osi, 0.4, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["indexed", "bricks"])
def test_wall_shear_of_solver_poiseuille(layout):
    """Steady Poiseuille box (R = 10, tau = 0.8, force 1e-6, 10,000 steps): device WSS on the solver's own velocity
    storage against the analytic wall shear nu * 2 u_max / R (lattice units)."""
    from cerebrosim_amd import lbm
    from cerebrosim_amd.geometry import OFFSET, poiseuille_box
    lat, r, R = poiseuille_box()
    nu = 0.1
    s = lbm.Solver(lat, nu, model="regularised", layout=layout)
    s.set_force([0, 0, 1e-6])
    s.run(10000)
    cx, cy = 24 / 2 - 0.5 + OFFSET[0], 24 / 2 - 0.5 + OFFSET[1]
    rng = np.random.default_rng(3)
    th = rng.uniform(0, 2 * np.pi, 300)
    verts = np.column_stack([cx + R * np.cos(th), cy + R * np.sin(th), rng.uniform(8, 56, 300)])
    normals = -np.column_stack([np.cos(th), np.sin(th), np.zeros_like(th)])
    cells = lat["coords"].astype(float)
    idx, c, quad, _ = wss.weights(verts, normals, cells, R - r, radius=6.0, d_min=0.2, d_max=4.5, sigma=1.5)
    ws = wss.WallShear(s, idx, c, normals, u_scale=1.0, mu=nu / 1e3)
    tau = ws.frame().cpu().numpy()
    exact = nu * 2 * (1e-6 / (4 * nu) * R ** 2) / R
    mag = np.linalg.norm(tau, axis=1)
    print(f"{layout}: WSS {mag.mean():.4e} vs exact {exact:.4e}", flush=True)
    assert abs(mag.mean() / exact - 1) < 0.05
    assert np.abs(tau[:, 2]).min() > 0 and (np.abs(tau[:, :2]) < 0.05 * exact).all()
