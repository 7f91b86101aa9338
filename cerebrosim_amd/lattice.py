"""D3Q27 lattice tables, unit conversion and the sparse-lattice builder (host side, numpy).

These mirror the module-level API of CerebroSim's ``lbm.py`` so that callers written against it keep working:

* ``C, W, Q, OPP, LATTICE`` and the link codes ``WALL, INLET, OUTLET0`` -- lbm.py:33-50. The direction order must
  be the reference's (rest, +-x, +-y, +-z, 12 edges, 8 corners) because lattices, fixtures and the finite-difference
  gradient (links 1..6) index by it.
* ``Units`` -- lbm.py:63-82.
* ``build_lattice`` -- lbm.py:264-330. Implemented here on a padded cell-id grid (out-of-grid nodes are a sentinel,
  periodic axes are wrapped by padding), which gives the same ``nbr``/``q``/``uin``/``counts`` as the reference.
* ``scatter`` -- lbm.py:333-338.
"""
from __future__ import annotations

import contextlib
import os
import threading

import numpy as np

LATTICE = os.environ.get("ARTERY_LBM_LATTICE", "D3Q27")
_FACES = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
_EDGES = [[1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
          [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]]
_CORNERS = [[1, 1, 1], [-1, -1, -1], [1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, -1], [-1, 1, 1], [1, -1, -1]]
if LATTICE == "D3Q27":
    C = np.array(_FACES + _EDGES + _CORNERS, np.int32)
    W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
elif LATTICE == "D3Q19":
    C = np.array(_FACES + _EDGES, np.int32)
    W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
else:
    raise ValueError(f"ARTERY_LBM_LATTICE={LATTICE}: use D3Q27 or D3Q19")
Q = len(C)
# opposite direction of every link: pairs (1,2), (3,4), ... by construction of the tables above
OPP = np.array([int(np.flatnonzero((C == -c).all(1))[0]) for c in C], np.int32)

WALL, INLET, OUTLET0 = -1, -2, -3   # link codes in the neighbour table; outlet k has code OUTLET0 - k
_OUTSIDE = np.iinfo(np.int32).min   # builder-internal: source node lies outside the grid


class Units:
    """Physical <-> lattice units for spacing dx (m), step dt (s), density rho (kg/m^3), kinematic viscosity nu."""

    def __init__(self, dx, dt, rho, nu):
        self.dx, self.dt, self.rho, self.nu = dx, dt, rho, nu
        self.nu_lu = nu * dt / dx ** 2
        self.tau = 3.0 * self.nu_lu + 0.5
        self.c = dx / dt

    def u(self, u_lu):
        return np.asarray(u_lu) * self.c

    def u_lu(self, u):
        return np.asarray(u) / self.c

    def p(self, p_lu):
        return np.asarray(p_lu) * self.rho * self.c ** 2

    def p_lu(self, p):
        return np.asarray(p) / (self.rho * self.c ** 2)


def _padded(grid, periodic, fill):
    """Grid padded by one node per side: wrapped on periodic axes, `fill` elsewhere."""
    out = grid
    for a in range(grid.ndim):
        width = [(0, 0)] * grid.ndim
        width[a] = (1, 1)
        if a in periodic:
            out = np.pad(out, width, mode="wrap")
        else:
            out = np.pad(out, width, mode="constant", constant_values=fill)
    return out


def _classify_openings(points, code, inlets, outlets, dx, slack):
    """Assign missing links to openings: solid node beyond the opening plane and within radius + slack*dx.
    Inlets are tried before outlets and the first match wins (lbm.py:296-306)."""
    for kind, ops in (("inlet", inlets), ("outlet", outlets)):
        for k, o in enumerate(ops):
            centre = np.asarray(o["centre"], float)
            normal = np.asarray(o["normal"], float)
            rel = points - centre
            along = rel @ normal
            lateral = np.linalg.norm(rel - along[:, None] * normal, axis=1)
            hit = (along >= -1e-6) & (lateral < o["radius"] + slack * dx) & (code == WALL)
            code[hit] = INLET if kind == "inlet" else OUTLET0 - k
    return code


_local = threading.local()


@contextlib.contextmanager
def builder_threads(n):
    """Within the block, lattices built on this thread use n native builder threads (default: every CPU of the
    process). A pool of builder threads sets 1, so that the pool and not each build spreads over the CPUs."""
    old = getattr(_local, "threads", 0)
    _local.threads = int(n)
    try:
        yield
    finally:
        _local.threads = old


class _NativeLattice:
    """Owner of a native lattice handle: the arrays of build_lattice are views of its buffers (no copy of the
    hundreds of megabytes a large lattice holds), and the handle is freed when the last of them is released."""

    def __init__(self, lib, handle):
        self.lib, self.handle = lib, handle

    def view(self, name, shape, dtype):
        import ctypes
        from . import _native
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if nbytes == 0:
            return np.empty(shape, dtype)
        ptr = ctypes.c_void_p()
        _native.call("cs_lattice_view", self.handle, name.encode(), ctypes.byref(ptr))
        buf = (ctypes.c_char * nbytes).from_address(ptr.value)
        buf._owner = self                      # the buffer keeps the handle alive; numpy keeps the buffer
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def __del__(self):
        self.lib.cs_lattice_free(self.handle)


def build_lattice(phi, origin, dx, openings=(), periodic=(), opening_slack=1.0, native=True, threads=None):
    """Sparse fluid-cell lattice from a signed distance sampled on grid nodes (negative = fluid), lbm.py:264-330.

    native=True (default) builds with the threaded C++ builder (cs_lattice_build); native=False with the numpy
    restatement below (build_lattice_numpy), which gives identical arrays (tested). threads: native builder threads
    (default: as set by builder_threads, else every CPU the process may use). Arguments and result as
    build_lattice_numpy."""
    if threads is None:
        threads = getattr(_local, "threads", 0)
    if not native:
        return build_lattice_numpy(phi, origin, dx, openings, periodic, opening_slack)
    import ctypes
    from . import _native
    lib = _native.load()
    phi = np.ascontiguousarray(phi, np.float32)
    origin = np.asarray(origin, float)
    outlets = [o for o in openings if o["kind"] == "outlet"]
    inlets = [o for o in openings if o["kind"] == "inlet"]
    ops = (_native.Opening * max(1, len(openings)))()
    for k, o in enumerate(inlets + outlets):
        ops[k].kind = 0 if o["kind"] == "inlet" else 1
        ops[k].centre = (ctypes.c_double * 3)(*[float(v) for v in o["centre"]])
        ops[k].normal = (ctypes.c_double * 3)(*[float(v) for v in o["normal"]])
        ops[k].radius = float(o["radius"])
    g = _native.GridIn()
    g.phi = phi.ctypes.data
    g.shape = (ctypes.c_int32 * 3)(*phi.shape)
    g.periodic = (ctypes.c_int32 * 3)(*[int(a in tuple(periodic)) for a in range(3)])
    g.origin = (ctypes.c_double * 3)(*origin)
    g.dx, g.openings, g.n_openings = float(dx), ctypes.addressof(ops), len(openings)
    g.slack, g.threads = float(opening_slack), int(threads)
    h = ctypes.c_void_p()
    _native.call("cs_lattice_build", ctypes.byref(g), ctypes.byref(h))
    owner = _NativeLattice(lib, h)
    n = ctypes.c_int64()
    cnt = np.zeros(2 + len(outlets), np.int64)
    lib.cs_lattice_sizes(h, ctypes.byref(n), ctypes.c_void_p(cnt.ctypes.data), len(cnt))
    N = int(n.value)
    shapes = {"coords": ((N, 3), np.int32), "nbr": ((Q, N), np.int32), "q": ((Q, N), np.float32),
              "uin": ((N, 3), np.float32)}
    out = {name: owner.view(name, shape, dt) for name, (shape, dt) in shapes.items()}
    counts = {"wall": int(cnt[0]), "inlet": int(cnt[1]), "outlet": [int(v) for v in cnt[2:]]}
    return {"nbr": out["nbr"], "q": out["q"], "uin": out["uin"], "coords": out["coords"], "origin": origin, "dx": dx,
            "shape": phi.shape, "n_outlets": len(outlets), "counts": counts, "openings": inlets + outlets}


def build_lattice_numpy(phi, origin, dx, openings=(), periodic=(), opening_slack=1.0):
    """Sparse fluid-cell lattice from a signed distance sampled on grid nodes (negative = fluid).

    Nodes sit at origin + (i, j, k) * dx. ``openings`` are dicts {"kind": "inlet"|"outlet", "centre", "normal"
    (outward), "radius"}; ``periodic`` lists the axes that wrap. Returns a dict with ``nbr`` (Q, N) int32 source cell
    or link code, ``q`` (Q, N) float32 wall fraction (0.5 where unused), ``uin`` (N, 3) float32 inlet profile (unit
    mean, parabolic, pointing inward), ``coords`` (N, 3) int32 in C order, and ``origin, dx, shape, n_outlets,
    counts, openings`` -- the reference's contract (lbm.py:264-330).
    """
    phi = np.asarray(phi, np.float32)
    shape = np.array(phi.shape)
    periodic = tuple(int(a) for a in periodic)
    fluid = phi < 0
    coords = np.argwhere(fluid).astype(np.int32)
    N = len(coords)
    ids = np.full(phi.shape, -1, np.int32)
    ids[tuple(coords.T)] = np.arange(N, dtype=np.int32)
    ids_p = _padded(ids, periodic, _OUTSIDE)
    phi_p = _padded(phi, periodic, np.float32(np.nan))
    origin = np.asarray(origin, float)
    outlets = [o for o in openings if o["kind"] == "outlet"]
    inlets = [o for o in openings if o["kind"] == "inlet"]
    nbr = np.empty((Q, N), np.int32)
    qw = np.full((Q, N), 0.5, np.float32)
    uin = np.zeros((N, 3), np.float32)
    counts = {"wall": 0, "inlet": 0, "outlet": [0] * len(outlets)}
    phi_here = phi[tuple(coords.T)]
    shifted = coords.astype(np.int64) + 1            # positions in the padded grids
    for i in range(Q):
        s = shifted - C[i]
        src = ids_p[tuple(s.T)]
        miss = np.flatnonzero(src < 0)
        if len(miss):
            node = coords[miss].astype(np.int64) - C[i]
            for a in periodic:
                node[:, a] %= shape[a]
            code = _classify_openings(origin + node * dx, np.full(len(miss), WALL, np.int32), inlets, outlets, dx,
                                      opening_slack)
            walls = code == WALL
            phi_solid = phi_p[tuple(s[miss[walls]].T)].astype(np.float64)   # NaN when the solid node is off the grid
            ph = phi_here[miss[walls]].astype(np.float64)
            with np.errstate(divide="ignore", invalid="ignore"):
                q = -ph / (phi_solid - ph)                      # FP64 from the FP32 distances, stored as FP32
            qw[i, miss[walls]] = np.where(np.isfinite(q), np.clip(q, 1e-3, 1.0), 0.5).astype(np.float32)
            src[miss] = code
            counts["wall"] += int(walls.sum())
            counts["inlet"] += int((code == INLET).sum())
            for k in range(len(outlets)):
                counts["outlet"][k] += int((code == OUTLET0 - k).sum())
        nbr[i] = src
    if inlets:
        cells = np.flatnonzero((nbr == INLET).any(0))
        pos = origin + coords[cells] * dx
        for o in inlets:                                  # every inlet writes all inlet cells: the last one wins
            centre = np.asarray(o["centre"], float)
            normal = np.asarray(o["normal"], float)
            rel = pos - centre
            lateral = np.linalg.norm(rel - (rel @ normal)[:, None] * normal, axis=1)
            prof = np.clip(2.0 * (1.0 - (lateral / o["radius"]) ** 2), 0.0, None)
            uin[cells] = (-prof[:, None] * normal[None]).astype(np.float32)
    return {"nbr": nbr, "q": qw, "uin": uin, "coords": coords, "origin": origin, "dx": dx, "shape": phi.shape,
            "n_outlets": len(outlets), "counts": counts, "openings": inlets + outlets}


def scatter(lat, values, fill=np.nan):
    """Per-cell values -> dense grid of the lattice's shape, `fill` outside the fluid (lbm.py:333-338)."""
    values = np.asarray(values)
    dtype = values.dtype if values.dtype.kind == "f" else float
    out = np.full(tuple(lat["shape"]) + values.shape[1:], fill, dtype)
    out[tuple(lat["coords"].T)] = values
    return out
