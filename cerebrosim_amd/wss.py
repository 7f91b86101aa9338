"""Wall shear stress on the device (counterpart of wss.py).

``weights`` builds, once per wall surface, the sparse least-squares coefficients of the wall-normal velocity
derivative on the host (the method of wss.py:35-50: K nearest lumen cells per wall vertex, Gaussian side weight,
depth window, quadratic fit u_t = a d + b d^2 with a linear fallback below 6 cells), solved here from the fit's
normal equations. ``WallShear`` then evaluates tau = mu * 1e3 * sum c u_t for every frame with the
``cs_wss`` kernel on the solver's live velocity (no field copies to the host) and accumulates the time-averaged WSS
and the oscillatory shear index (wss.py:120-122).
"""
from __future__ import annotations

import ctypes

import numpy as np
from scipy.spatial import cKDTree

from . import _native

MU = 3.5e-3       # Pa s (geometry/physics.json: blood viscosity)
K = 48            # nearest lumen cells per wall vertex (wss.py:30)
D_MIN, D_MAX = 0.02, 0.45      # mm: cells used in the fit, by distance to the wall (wss.py:31)
SIGMA = 0.15      # mm: sideways fall-off of the weights (wss.py:32)


def weights(verts, normals, cells, depth, k=K, radius=0.8, d_min=D_MIN, d_max=D_MAX, sigma=SIGMA):
    """Per wall vertex, the linear functional that maps the tangential velocities u_t of its k nearest lumen cells
    (within `radius` mm) to the wall-normal derivative a of the fit u_t(d) = a d + b d^2 (no slip at d = 0): returns
    (idx (V, k), c (V, k) float32 with a = sum_j c_j u_t,j per mm, quadratic (V,) bool, any (V,) bool).

    Each cell's weight is a Gaussian of its sideways distance from the vertex (sigma) times the depth window
    d_min < d < d_max. The weighted least-squares problem min sum_j w_j (u_j - a d_j - b d_j^2)^2 has the normal
    equations M [a, b]^T = B^T W u, with the basis B_j = (d_j, d_j^2) and M = B^T W B; so c = e_1^T M^-1 B^T W. Where
    M is near-singular or fewer than 6 cells carry weight (> 0.05), the one-term fit u_t = a d is used instead:
    c = w d / sum w d^2 (wss.py:35-50)."""
    dist, idx = cKDTree(cells).query(verts, k=k, distance_upper_bound=radius)
    found = np.isfinite(dist)
    idx = np.where(found, idx, 0)
    d = depth[idx]                                                    # (V, k)
    offset = cells[idx] - verts[:, None, :]
    along = np.einsum("vkx,vx->vk", offset, normals)
    lateral = np.linalg.norm(offset - along[..., None] * normals[:, None, :], axis=-1)
    w = np.exp(-(lateral / sigma) ** 2) * (found & (d > d_min) & (d < d_max))
    basis = np.stack([d, d * d], axis=-1)                             # (V, k, 2)
    M = np.einsum("vk,vki,vkj->vij", w, basis, basis)                 # (V, 2, 2)
    det = M[:, 0, 0] * M[:, 1, 1] - M[:, 0, 1] * M[:, 1, 0]
    quadratic = (det > 1e-12 * np.maximum(M[:, 0, 0] * M[:, 1, 1], 1e-30)) & ((w > 0.05).sum(1) >= 6)
    # first row of M^-1 (M symmetric), applied to the weighted basis
    row = np.stack([M[:, 1, 1], -M[:, 0, 1]], axis=-1) / np.where(quadratic, det, 1.0)[:, None]
    c_quad = w * np.einsum("vki,vi->vk", basis, row)
    c_lin = w * d / np.maximum(M[:, 0, 0], 1e-30)[:, None]
    c = np.where(quadratic[:, None], c_quad, c_lin)
    c[M[:, 0, 0] <= 0] = 0.0
    return idx, c.astype(np.float32), quadratic, M[:, 0, 0] > 0


def wall_surface(phi, origin, dx, ends=(), clearance=0.15):
    """Points on the lumen wall of a signed-distance grid (mm, negative inside, node (i, j, k) at origin + (i, j, k)
    dx): where the distance changes sign along an axis edge between two nodes, at the linear-interpolation zero (the
    vertices marching cubes places on this grid, which the reference's wall mesh has). Normals point into the lumen:
    minus the distance's gradient, by central differences at the edge's two nodes, interpolated like the point.
    Points on an open end's cut face (within `clearance` mm of its plane and 1.2 radii + 0.2 mm of its centre,
    wss.py:130-133) are dropped. Returns (verts (V, 3) float64 mm, normals (V, 3) float64)."""
    phi = np.asarray(phi, np.float32)
    shape = np.array(phi.shape)
    flat = phi.ravel()
    strides = np.array([shape[1] * shape[2], shape[2], 1])
    inside = phi < 0
    verts, normals = [], []

    def grad_at(node):
        """(n, 3) central-difference gradient at flat node indices (one-sided at the grid's faces)."""
        ijk = np.stack(np.unravel_index(node, phi.shape), 1)
        g = np.empty((len(node), 3))
        for a in range(3):
            lo = np.maximum(ijk[:, a] - 1, 0)
            hi = np.minimum(ijk[:, a] + 1, shape[a] - 1)
            span = np.maximum(hi - lo, 1) * dx
            g[:, a] = (flat[node + (hi - ijk[:, a]) * strides[a]].astype(np.float64)
                       - flat[node - (ijk[:, a] - lo) * strides[a]]) / span
        return g

    for a in range(3):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[a] = slice(0, shape[a] - 1)
        sl1[a] = slice(1, shape[a])
        cross = inside[tuple(sl0)] != inside[tuple(sl1)]
        ijk = np.stack(np.nonzero(cross), 1)
        if not len(ijk):
            continue
        n0 = ijk @ strides
        n1 = n0 + strides[a]
        p0, p1 = flat[n0].astype(np.float64), flat[n1].astype(np.float64)
        t = p0 / (p0 - p1)
        pos = ijk.astype(np.float64)
        pos[:, a] += t
        verts.append(np.asarray(origin, float) + pos * dx)
        g = (1.0 - t)[:, None] * grad_at(n0) + t[:, None] * grad_at(n1)
        normals.append(-g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-12))
    verts = np.concatenate(verts) if verts else np.zeros((0, 3))
    normals = np.concatenate(normals) if normals else np.zeros((0, 3))
    keep = np.ones(len(verts), bool)
    for e in ends:
        off = verts - np.asarray(e["centre"], float)
        keep &= ~((np.abs(off @ np.asarray(e["normal"], float)) < clearance)
                  & (np.linalg.norm(off, axis=1) < 1.2 * e["radius"] + 0.2))
    return verts[keep], normals[keep]


class WallShear:
    """Device WSS over a fixed wall surface for one solver (or any velocity storage with the same addressing).

    idx: (V, K) cell ids of the lattice (converted to storage slots), c: (V, K) coefficients, normals: (V, 3) into
    the lumen, u_scale: physical velocity per solver velocity unit (m/s), mu: dynamic viscosity (Pa s)."""

    def __init__(self, solver, idx, c, normals, u_scale=1.0, mu=MU):
        import torch
        self.s = solver
        dev = solver.dev
        V, k = idx.shape
        slots = solver.backend.slots(np.asarray(idx).ravel()).reshape(V, k)
        with torch.cuda.stream(solver._tstream):
            self.idx = torch.from_numpy(np.ascontiguousarray(slots, np.int32)).to(dev)
            self.c = torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(dev)
            self.n = torch.from_numpy(np.ascontiguousarray(normals, np.float32)).to(dev)
            self.tau = torch.zeros((V, 3), dtype=torch.float32, device=dev)
            self.acc_tau = torch.zeros((V, 3), dtype=torch.float32, device=dev)
            self.acc_mag = torch.zeros(V, dtype=torch.float32, device=dev)
        solver.synchronize()
        self.V, self.k = V, k
        self.scale = float(mu * 1e3 * u_scale)
        self.frames = 0

    def frame(self, accumulate=True):
        """WSS (Pa) of the solver's newest velocity field at every vertex (device tensor (V, 3))."""
        s = self.s
        elem, comp = s.backend.velocity_addressing()
        A = _native.WssArgs()
        A.idx, A.c, A.normal, A.n_vert, A.k = self.idx.data_ptr(), self.c.data_ptr(), self.n.data_ptr(), self.V, self.k
        A.vel, A.vel_elem, A.vel_comp = s.backend.velocity(s.cur).data_ptr(), elem, comp
        A.scale = self.scale
        A.tau = self.tau.data_ptr()
        A.acc_tau = self.acc_tau.data_ptr() if accumulate else None
        A.acc_mag = self.acc_mag.data_ptr() if accumulate else None
        _native.call("cs_wss", ctypes.byref(A), s.stream_handle)
        if accumulate:
            self.frames += 1
        import torch
        torch.cuda.current_stream(s.dev).wait_stream(s._tstream)   # later torch work on tau sees the result
        return self.tau

    def tawss(self):
        self.s.synchronize()
        return (self.acc_mag / max(1, self.frames)).cpu().numpy()

    def osi(self):
        self.s.synchronize()
        n = max(1, self.frames)
        mean = (self.acc_tau / n).norm(dim=1)
        tawss = self.acc_mag / n
        return (0.5 * (1 - mean / tawss.clamp_min(1e-9))).cpu().numpy()
