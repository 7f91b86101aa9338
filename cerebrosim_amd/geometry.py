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
    G = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) - A
    s = G @ axis
    rad = np.linalg.norm(G - s[:, None] * axis, axis=1)
    phi = np.maximum(np.maximum(rad - radius(s), -s), s - L).reshape(n)
    R0, R1 = float(radius(np.array([0.0]))[0]), float(radius(np.array([L]))[0])
    openings = [{"kind": "inlet", "centre": A, "normal": -axis, "radius": R0},
                {"kind": "outlet", "centre": A + axis * L, "normal": axis, "radius": R1}]
    lat = build_lattice(phi, [0, 0, 0], 1.0, openings=openings)
    P = lat["coords"] - A
    sc = P @ axis
    rc = np.linalg.norm(P - sc[:, None] * axis, axis=1)
    return lat, sc, rc, axis


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
