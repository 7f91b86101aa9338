"""Synthetic vessel geometries as signed distances on the lattice grid, and the lattices built from them.

The pipes reproduce the geometry of CerebroSim's validation harness so that results can be compared with its
published numbers: ``periodic_pipe`` follows validate.py:59-66 and ``open_pipe`` follows validate.py:125-148 (same
grid extent, axis offset ``OFFSET`` and flat end caps), because the published cell and link counts depend on them.
The benchmark workloads (steady Poiseuille box, pulsatile Womersley pipe) are defined here by their sizes.
"""
from __future__ import annotations

import math

import numpy as np

from .lattice import build_lattice

OFFSET = np.array([0.31, 0.17, 0.5])   # pipe axes off the grid's symmetry lines (validate.py:42)


def periodic_pipe(D, nz=2):
    """Pipe of diameter D along z, nz nodes long and periodic in z (the reference's Poiseuille/Womersley pipe has
    nz = 2; flows that do not vary along z give the same per-cell values for any nz). Returns (lattice, radial
    distance per cell, R)."""
    R = D / 2.0
    n = int(math.ceil(D)) + 4
    cx, cy = n / 2 - 0.5 + OFFSET[0], n / 2 - 0.5 + OFFSET[1]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r2d = np.hypot(i - cx, j - cy)
    r = np.repeat(r2d[:, :, None], nz, axis=2)
    lat = build_lattice(r - R, [0, 0, 0], 1.0, periodic=(2,))
    return lat, r[tuple(lat["coords"].T)], R


def poiseuille_box(R=10.0, nx=24, ny=24, nz=64):
    """Steady Poiseuille workload: cylinder of radius R along z in an nx x ny x nz box, periodic in z, axis at the
    reference's offset centre (nx/2 - 0.5 + 0.31, ny/2 - 0.5 + 0.17). Returns (lattice, r per cell, R)."""
    cx, cy = nx / 2 - 0.5 + OFFSET[0], ny / 2 - 0.5 + OFFSET[1]
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    r2d = np.hypot(i - cx, j - cy)
    r = np.broadcast_to(r2d[:, :, None], (nx, ny, nz))
    lat = build_lattice(r - R, [0, 0, 0], 1.0, periodic=(2,))
    return lat, np.ascontiguousarray(r[tuple(lat["coords"].T)]), R


def open_pipe(axis, L, radius, pad=3):
    """Straight pipe from A along unit `axis`, length L (cells), radius a function of the axial coordinate, with a
    velocity inlet at A and a pressure outlet at A + L axis. Returns (lattice, axial coordinate per cell, radial
    coordinate per cell, unit axis)."""
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    rmax = float(max(radius(np.linspace(0, L, 400))))
    ends = np.array([np.zeros(3), axis * L])
    ext = rmax * np.sqrt(np.clip(1 - axis ** 2, 0, 1)) + pad
    lo = ends.min(0) - ext
    hi = ends.max(0) + ext
    n = np.ceil(hi - lo).astype(int) + 1
    A = -lo + OFFSET                                     # pipe start in grid coordinates
    a = _grid_axis(axis)
    if a is None:
        G = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) - A
        s = G @ axis
        rad = np.linalg.norm(G - s[:, None] * axis, axis=1)
        phi = np.maximum(np.maximum(rad - radius(s), -s), s - L).reshape(n)
    else:
        # along a grid axis the axial coordinate is that axis' node coordinate and the radial distance depends only
        # on the other two, so both are computed on the grid's lines and planes and broadcast: the same floating
        # point operations on the same values as the general form above (s = G . e_a = G_a exactly; the radial
        # vector's axial component is exactly 0), so phi has the same bits, without the (cells, 3) temporaries
        b, c = [k for k in range(3) if k != a]
        s = np.arange(n[a]) - A[a]
        gb, gc = np.arange(n[b]) - A[b], np.arange(n[c]) - A[c]
        rad = np.sqrt(gb[:, None] * gb[:, None] + gc[None, :] * gc[None, :])          # (n_b, n_c)
        phi = np.maximum(np.maximum(rad[:, :, None] - radius(s)[None, None, :], -s[None, None, :]),
                         (s - L)[None, None, :])                                         # axes (b, c, a)
        phi = np.ascontiguousarray(np.moveaxis(phi, 2, a) if a != 2 else phi)
    R0, R1 = float(radius(np.array([0.0]))[0]), float(radius(np.array([L]))[0])
    openings = [{"kind": "inlet", "centre": A, "normal": -axis, "radius": R0},
                {"kind": "outlet", "centre": A + axis * L, "normal": axis, "radius": R1}]
    lat = build_lattice(phi, [0, 0, 0], 1.0, openings=openings)
    P = lat["coords"] - A
    if a is None:
        sc = P @ axis
        rc = np.linalg.norm(P - sc[:, None] * axis, axis=1)
    else:
        sc = np.ascontiguousarray(P[:, a])
        rc = np.sqrt(P[:, b] * P[:, b] + P[:, c] * P[:, c])
    return lat, sc, rc, axis


def _grid_axis(axis):
    """Index of the grid axis that the unit vector `axis` points along (+ direction only), else None."""
    nz = np.flatnonzero(axis != 0.0)
    return int(nz[0]) if len(nz) == 1 and axis[nz[0]] == 1.0 else None


def plug_inlet(lat):
    """Replace the parabolic inlet profile by a plug of unit speed along the inward normal (inlet cells only)."""
    inlet = lat["openings"][0]
    cells = np.flatnonzero((lat["nbr"] == -2).any(0))
    uin = np.zeros_like(lat["uin"])
    uin[cells] = -np.asarray(inlet["normal"], float)[None].astype(np.float32)
    lat["uin"] = uin
    return lat


def womersley_pipe(R=40.0, L=480):
    """Pulsatile pipe workload: radius R, length L along z, velocity inlet (plug profile) at z = 0 and a pressure
    outlet at z = L; about 2.41 M fluid cells at R = 40, L = 480. Returns (lattice, axial, radial coordinates)."""
    lat, sc, rc, _ = open_pipe((0, 0, 1), L, lambda s: np.full_like(np.asarray(s, float), R))
    return plug_inlet(lat), sc, rc


def stenosis_pipe(dx, sev=0.7, D0_mm=2.0, Ls_mm=4.0, up_d=4, down_d=18):
    """Cosine stenosis of area reduction 1 - (1 - sev)^2 in a straight pipe at grid spacing dx (m), with the
    dimensions of the reference's narrowing test (validate.py:226-239). Returns (lattice, axial, radial, D cells)."""
    D = D0_mm * 1e-3 / dx
    R0 = D / 2
    Ls = Ls_mm * 1e-3 / dx
    s0 = up_d * D
    L = (up_d + down_d) * D + Ls

    def radius(s):
        s = np.asarray(s, float)
        x = np.clip((s - s0) / Ls, 0, 1)
        return R0 * (1 - sev * 0.5 * (1 - np.cos(2 * math.pi * x)))

    lat, sc, rc, _ = open_pipe((0, 0, 1), L, radius)
    return lat, sc, rc, D


def _segment_sdf(P, A, axis, L, radius):
    """Signed distance (approximate, exact on the lateral wall of a straight tube) of points P (n, 3) to a tube from A
    along unit `axis`, length L, radius(s), flat caps."""
    rel = P - A
    s = rel @ axis
    rad = np.linalg.norm(rel - s[:, None] * axis, axis=1)
    return np.maximum(np.maximum(rad - radius(s), -s), s - L)


def bifurcation(scale=1.0, pad=3):
    """Stenosed bifurcation: parent tube along z of radius 60 s narrowed to 30 % area at z = 600 s (Gaussian,
    sigma 40 s), splitting at z = 1100 s into branches of radius 45 s and 35 s at +-30 degrees in the y-z plane, in
    a 240 s x 400 s x 1800 s box (s = scale). The branches end inside the box (length 300 s: at the full 1800 s they
    would leave its 400 s width). Velocity inlet (parabolic profile) at the parent's start, pressure outlets at the
    branch ends. The signed distance is evaluated slab by slab in z, so the full-size grid fits in host memory.

    Returns (lattice, info) with info = {"segments", "sdf" (points -> signed distance), "R", "throat"}."""
    s = float(scale)
    n = np.array([int(round(240 * s)), int(round(400 * s)), int(round(1800 * s))])
    c = np.array([n[0] / 2 - 0.5, n[1] / 2 - 0.5, 0.0]) + OFFSET * np.array([1, 1, 0])
    z_in, z_j = float(pad), 1100.0 * s
    R0, d = 60.0 * s, 1.0 - math.sqrt(0.3)
    zs, sig = 600.0 * s, 40.0 * s

    def parent_r(t):
        z = np.asarray(t, float) + z_in
        return R0 * (1.0 - d * np.exp(-0.5 * ((z - zs) / sig) ** 2))

    ang = math.radians(30.0)
    Lb = 300.0 * s
    A0 = c + np.array([0.0, 0.0, z_in])
    J = c + np.array([0.0, 0.0, z_j])
    segs = [(A0, np.array([0.0, 0.0, 1.0]), z_j - z_in + 0.5 * R0, parent_r, R0)]
    for sgn, Rb in ((1.0, 45.0 * s), (-1.0, 35.0 * s)):
        ax = np.array([0.0, sgn * math.sin(ang), math.cos(ang)])
        segs.append((J, ax, Lb, (lambda Rb: lambda t: np.full_like(np.asarray(t, float), Rb))(Rb), Rb))

    def sdf(P):
        P = np.asarray(P, float)
        out = np.full(len(P), np.inf)
        for A, ax, L, rf, _ in segs:
            out = np.minimum(out, _segment_sdf(P, A, ax, L, rf))
        return out

    phi = np.empty(tuple(n), np.float32)
    I, Jj = np.meshgrid(np.arange(n[0]), np.arange(n[1]), indexing="ij")
    plane = np.stack([I.ravel(), Jj.ravel()], 1).astype(float)
    for k in range(n[2]):
        P = np.concatenate([plane, np.full((len(plane), 1), float(k))], 1)
        phi[:, :, k] = sdf(P).reshape(n[0], n[1])
    openings = [{"kind": "inlet", "centre": A0, "normal": np.array([0.0, 0.0, -1.0]), "radius": float(parent_r(0.0))}]
    for A, ax, L, rf, Rb in segs[1:]:
        openings.append({"kind": "outlet", "centre": A + ax * L, "normal": ax, "radius": Rb})
    lat = build_lattice(phi, [0, 0, 0], 1.0, openings=openings)
    info = {"segments": segs, "sdf": sdf, "R": R0, "throat": R0 * math.sqrt(0.3), "shape": n}
    return lat, info


def wall_points(info, spacing=2.0, end_margin=6.0):
    """Points on the bifurcation's wall (spacing in cells) with unit normals into the lumen, away from the openings
    and from the junction (points inside another segment are dropped). Returns (points (V, 3), normals (V, 3))."""
    sdf = info["sdf"]
    pts, nrm = [], []
    for A, ax, L, rf, Rmax in info["segments"]:
        e1 = np.cross(ax, [1.0, 0.0, 0.0])
        if np.linalg.norm(e1) < 1e-6:
            e1 = np.cross(ax, [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(ax, e1)
        t = np.arange(end_margin, L - end_margin, spacing)
        for tt in t:
            r = float(rf(np.array([tt]))[0])
            dr = (float(rf(np.array([tt + 0.5]))[0]) - float(rf(np.array([tt - 0.5]))[0]))
            th = np.arange(0.0, 2 * math.pi, spacing / max(r, 1.0))
            radial = np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2
            P = A + tt * ax + r * radial
            nvec = -radial + dr * ax
            nvec /= np.linalg.norm(nvec, axis=1, keepdims=True)
            pts.append(P)
            nrm.append(nvec)
    P = np.concatenate(pts)
    Nn = np.concatenate(nrm)
    keep = np.abs(sdf(P)) < 0.25                     # on the union's wall, not inside another segment
    return P[keep], Nn[keep]


def artery_y(dx_mm=0.1, R_parent=1.6, R_a=1.25, R_b=1.0, L_parent=14.0, L_branch=10.0, angle_deg=35.0, pad_mm=0.3):
    """A synthetic Y-shaped artery in millimetres, a small and quick stand-in for the story artery
    (``vessel.artery_case``): a parent tube (the ICA end) splitting into two branches (MCA, ACA) in the y-z plane. Every end is an
    opening of kind "outlet", as in simulate.py (the first is driven by the inlet pressure). Returns (lattice with
    "wall_mm", the distance of each fluid cell to the wall; ends; probes), in the form simulate.py's setup builds
    (simulate.py:162-200)."""
    ang = math.radians(angle_deg)
    ax_a = np.array([0.0, math.sin(ang), math.cos(ang)])
    ax_b = np.array([0.0, -math.sin(ang), math.cos(ang)])
    J = np.array([0.0, 0.0, L_parent])
    A0 = np.zeros(3)
    segs = [(A0, np.array([0.0, 0.0, 1.0]), L_parent + 0.5 * R_parent, R_parent), (J, ax_a, L_branch, R_a),
            (J, ax_b, L_branch, R_b)]

    def sdf(P):
        out = np.full(len(P), np.inf)
        for A, ax, L, R in segs:
            out = np.minimum(out, _segment_sdf(P, A, ax, L, lambda s, R=R: np.full_like(s, R)))
        return out

    ends_pts = [A0] + [A + ax * L for A, ax, L, _ in segs[1:]]
    lo = np.min([p - max(R_parent, R_a, R_b) for p in ends_pts + [J]], axis=0) - pad_mm
    hi = np.max([p + max(R_parent, R_a, R_b) for p in ends_pts + [J]], axis=0) + pad_mm
    lo[2], hi[2] = -pad_mm, max(p[2] for p in ends_pts) + pad_mm
    n = np.ceil((hi - lo) / dx_mm).astype(int) + 1
    origin = lo + OFFSET * dx_mm * np.array([1, 1, 0])
    I, Jg, K = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    P = origin + np.stack([I.ravel(), Jg.ravel(), K.ravel()], 1) * dx_mm
    phi = sdf(P).reshape(tuple(n)).astype(np.float32)
    ends = [{"kind": "outlet", "label": "inlet (ICA)", "centre": A0, "normal": np.array([0.0, 0.0, -1.0]),
             "radius": R_parent}]
    for (A, ax, L, R), label in zip(segs[1:], ("MCA outlet", "ACA outlet")):
        ends.append({"kind": "outlet", "label": label, "centre": A + ax * L, "normal": ax, "radius": R})
    lat = build_lattice(phi, origin, dx_mm, openings=ends)
    lat["wall_mm"] = -phi[tuple(lat["coords"].T)]
    from .wss import wall_surface
    lat["wall_verts"], lat["wall_normals"] = wall_surface(phi, origin, dx_mm, ends)
    probes = [{"name": "ICA", "point": np.array([0.0, 0.0, 0.4 * L_parent]), "radius": 0.5 * R_parent},
              {"name": "junction", "point": J.copy(), "radius": 1.0},
              {"name": "MCA distal", "point": J + ax_a * 0.7 * L_branch, "radius": 0.5 * R_a},
              {"name": "ACA distal", "point": J + ax_b * 0.7 * L_branch, "radius": 0.5 * R_b}]
    return lat, ends, probes
