"""Host side of the brick layout: converts a lattice in the reference's indexed layout (lbm.py:264-330) into the
4 x 4 x 4 brick storage used by the AA-pattern kernels (native builder ``cs_bricks_build``)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .lattice import C

BRICK = 4
ARRAYS = {
    # name: (dtype, shape as a function of the sizes)
    "coords": (np.int32, lambda z, N: (z.n_bricks, 3)),
    "nbr": (np.int32, lambda z, N: (z.n_bricks, 27)),
    "bidx": (np.int32, lambda z, N: (z.n_bricks,)),
    "interior": (np.int32, lambda z, N: (z.n_interior,)),
    "boundary": (np.int32, lambda z, N: (z.n_boundary,)),
    "slot_of_cell": (np.int32, lambda z, N: (N,)),
    "cell_of_slot": (np.int32, lambda z, N: (z.n_slots,)),
    "lmask": (np.uint32, lambda z, N: (z.n_boundary, 64)),
    "smask": (np.uint32, lambda z, N: (z.n_boundary, 64)),
    "uin": (np.float32, lambda z, N: (z.n_boundary, 3, 64)),
    "qlink": (np.float32, lambda z, N: (z.n_links + 1,)),
    "qstart": (np.int32, lambda z, N: (z.n_boundary + 1,)),
    "sstart": (np.int32, lambda z, N: (z.n_boundary + 1,)),
    "loff": (np.uint32, lambda z, N: (z.n_boundary, 64)),
    "ldir": (np.uint32, lambda z, N: (z.n_boundary, 2)),
}
DENSE = {"qc": (np.float32, lambda z, N: (z.n_boundary, 27, 64))}


class LayoutError(ValueError):
    """The lattice's neighbour table is not the implicit neighbourhood of a brick grid."""


def infer_periodic(lat):
    """Axes along which some fluid link wraps around the grid (build_lattice's `periodic`, not stored in the dict).
    Only cells on a grid face can have a wrapping link: at coordinate 0 along an axis, a link pulling from -1 has a
    fluid source only if the axis wraps (and likewise at shape - 1)."""
    nbr, coords = lat["nbr"], lat["coords"]
    shape = np.asarray(lat["shape"])
    periodic = [0, 0, 0]
    for a in range(3):
        for edge, sign in ((0, 1), (int(shape[a]) - 1, -1)):
            cells = np.flatnonzero(coords[:, a] == edge)
            if not len(cells):
                continue
            dirs = np.flatnonzero(C[:, a] == sign)            # sources at coordinate edge - sign: off the grid
            if (nbr[np.ix_(dirs, cells)] >= 0).any():
                periodic[a] = 1
    return periodic


def build(lat, dense=False):
    """Brick layout of `lat` as a dict of numpy arrays plus the sizes, all from the native builder (which runs
    outside the GIL). dense: also return ``qc``, the link codes of every boundary brick as (n_boundary, 27, 64)
    (the step reads only their compact form, ``qlink``). Raises LayoutError when the lattice cannot be represented
    (then the indexed layout serves it)."""
    N = int(lat["nbr"].shape[1])
    coords = np.ascontiguousarray(lat["coords"], np.int32)
    nbr = np.ascontiguousarray(lat["nbr"], np.int32)
    q = np.ascontiguousarray(lat["q"], np.float32)
    uin = np.ascontiguousarray(lat["uin"], np.float32)
    inp = _native.LatticeIn()
    inp.n_cells = N
    inp.shape = (ctypes.c_int32 * 3)(*[int(v) for v in lat["shape"]])
    periodic = infer_periodic(lat)
    inp.periodic = (ctypes.c_int32 * 3)(*periodic)
    inp.coords, inp.nbr, inp.q, inp.uin = coords.ctypes.data, nbr.ctypes.data, q.ctypes.data, uin.ctypes.data
    h = ctypes.c_void_p()
    lib = _native.load()
    rc = lib.cs_bricks_build(ctypes.byref(inp), ctypes.byref(h))
    if rc == _native.CS_E_LAYOUT:
        raise LayoutError(lib.cs_last_error().decode(errors="replace"))
    if rc != 0:
        raise _native.NativeError(f"cs_bricks_build returned {rc}: {lib.cs_last_error().decode(errors='replace')}")
    try:
        z = _native.BricksSizes()
        _native.call("cs_bricks_sizes_of", h, ctypes.byref(z))
        out = {"n_bricks": z.n_bricks, "n_boundary": z.n_boundary, "n_interior": z.n_interior,
               "n_slots": z.n_slots, "n_cells": N, "periodic": periodic, "n_links": z.n_links, "n_side": z.n_side}
        for name, (dt, shape) in (dict(ARRAYS, **DENSE) if dense else ARRAYS).items():
            arr = np.empty(shape(z, N), dt)
            if arr.size:
                _native.call("cs_bricks_copy", h, name.encode(), ctypes.c_void_p(arr.ctypes.data))
            out[name] = arr
    finally:
        lib.cs_bricks_free(h)
    return out


def compact_links(lay):
    """Numpy restatement of the native builder's compact boundary data (the tests' checker; needs ``build(lat,
    dense=True)``). The per-step boundary data in compact form (only links that exist):

    * ``qlink``: qc of every missing link of every boundary brick -- brick by brick, lane by lane, each lane's links
      in direction order 1..26 -- with ``qstart`` (n_boundary + 1) the start of each brick's run;
    * the Bouzidi side copies (links with a bit in smask) are stored in the same order, starting at ``sstart``
      (n_boundary + 1); ``n_side`` entries in all;
    * ``loff`` (n_boundary, 64) uint32: the start of each lane's run inside its brick's runs, qlink offset | side
      offset << 16; link i of a lane sits at its offset + the number of its links below i (a popcount);
    * ``ldir`` (n_boundary, 2) uint32: the union over the brick's cells of the missing links and of the smask bits.

    ``qlink`` carries one element of padding after the last run: a lane reads its entry for every direction in
    ``ldir``, also where its own link is not missing, and such an index is at most the end of its run."""
    nb = int(lay["n_boundary"])
    lm = lay["lmask"].reshape(nb, 64)
    sm = lay["smask"].reshape(nb, 64)
    links = np.uint32((1 << 27) - 2)                       # directions 1..26
    miss = np.where(lm != 0, ~lm & links, 0).astype(np.uint32)
    sbit = (sm & links).astype(np.uint32)
    dirs = np.arange(27, dtype=np.uint32)
    mT = ((miss[:, :, None] >> dirs[None, None, :]) & 1) == 1          # (nb, 64, 27): lane, then direction
    sT = ((sbit[:, :, None] >> dirs[None, None, :]) & 1) == 1
    qc = lay["qc"].reshape(nb, 27, 64).transpose(0, 2, 1)
    qlink = np.zeros(int(mT.sum()) + 1, np.float32)
    qlink[:-1] = qc[mT]
    nq, ns = mT.sum(2).astype(np.int64), sT.sum(2).astype(np.int64)
    qrel, srel = np.cumsum(nq, 1) - nq, np.cumsum(ns, 1) - ns
    qstart = np.zeros(nb + 1, np.int64)
    np.cumsum(nq.sum(1), out=qstart[1:])
    sstart = np.zeros(nb + 1, np.int64)
    np.cumsum(ns.sum(1), out=sstart[1:])
    if qstart[-1] >= 2 ** 31 or sstart[-1] >= 2 ** 31:
        raise LayoutError("more than 2^31 boundary links")
    loff = (qrel | (srel << 16)).astype(np.uint32)
    ldir = np.stack([np.bitwise_or.reduce(miss, axis=1) if nb else np.zeros(0, np.uint32),
                     np.bitwise_or.reduce(sbit, axis=1) if nb else np.zeros(0, np.uint32)], 1).astype(np.uint32)
    return {"qlink": qlink, "qstart": qstart.astype(np.int32), "sstart": sstart.astype(np.int32),
            "loff": np.ascontiguousarray(loff), "ldir": np.ascontiguousarray(ldir.reshape(nb, 2)),
            "n_links": int(qstart[-1]), "n_side": int(sstart[-1])}


def slot_table():
    """(27, 64) uint16: offset of population k of cell l inside its brick's block of 1728 floats."""
    out = np.empty((27, 64), np.uint16)
    _native.call("cs_bricks_slot_table", ctypes.c_void_p(out.ctypes.data))
    return out


def stage_table():
    """(27, 64) uint16: staged position of the population an interior brick's lane l pulls for direction i in an
    odd step (own group first, then the neighbour records: faces, edges, corners, each in offset order)."""
    out = np.empty((27, 64), np.uint16)
    _native.call("cs_bricks_stage_table", ctypes.c_void_p(out.ctypes.data))
    return out


def record_table():
    """(728,) uint16: neighbour offset o << 11 | block offset of staged odd-step positions 1000 .. 1727."""
    out = np.empty(27 * 64 - 1000, np.uint16)
    _native.call("cs_bricks_record_table", ctypes.c_void_p(out.ctypes.data))
    return out


def fill_fraction(layout):
    """Fluid cells / storage slots."""
    return layout["n_cells"] / max(1, layout["n_slots"])


def census(lat, edge=BRICK):
    """Brick statistics of a lattice for a brick edge (4 is the layout the kernels use; 8 sizes the alternative):
    bricks holding fluid, storage slots (edge^3 per brick), the fraction of fluid cells in boundary bricks (a brick
    with a solid slot or a missing link, which the slower boundary kernel steps) and storage slots per fluid cell
    (population traffic scales with slots, not with cells). Host-side numpy over the reference layout."""
    coords = np.asarray(lat["coords"], np.int64)
    N = len(coords)
    bc = coords // edge
    nb = bc.max(0) + 1
    key = (bc[:, 0] * nb[1] + bc[:, 1]) * nb[2] + bc[:, 2]
    ukey, inv = np.unique(key, return_inverse=True)
    inv = inv.ravel()
    cells = np.bincount(inv, minlength=len(ukey))
    complete = (np.asarray(lat["nbr"]) >= 0).all(0)               # every link of the cell fluid
    complete_cells = np.bincount(inv, weights=complete.astype(np.float64), minlength=len(ukey))
    interior = (cells == edge ** 3) & (complete_cells == cells)
    slots = len(ukey) * edge ** 3
    return {"edge": int(edge), "cells": int(N), "bricks": int(len(ukey)), "interior_bricks": int(interior.sum()),
            "slots": int(slots), "slots_per_cell": slots / max(N, 1),
            "boundary_cell_fraction": float(cells[~interior].sum()) / max(N, 1),
            "boundary_brick_fill": float(cells[~interior].sum()) / max(1, int((~interior).sum()) * edge ** 3)}
