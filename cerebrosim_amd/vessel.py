"""The story artery's lumen (counterpart of the reference's vessel.py:56-245): three tubes along the story's
centrelines, with typical adult radii, a smooth union at the junction, straight flow extensions and the clot.

* ``build(case, severity, segment)`` -- vessel.py:113-175. Centrelines in mm around the junction, resampled to
  RESAMPLE; bends tighter than BEND_LIMIT relaxed (``limit_bends``, vessel.py:91-110); radii interpolated from the
  DIAM tables; flow extensions of EXT_DIAMETERS end diameters; the clot as a radius scale
  1 - severity * w(s) on the blocked branch (w = 1 over the story's clot body, cosine ramps of CLOT_RAMP).
* ``sdf_grid(g, dx, lo, hi)`` -- vessel.py:234-245: the signed distance (mm, negative inside) on voxel centres,
  evaluated on the GPU by ``cs_sdf_tubes`` (round cones per segment, flat caps near the open ends, smooth minimum
  over the branches).
* ``artery_case(case, dx_mm)`` -- simulate.py:155-200: the lattice with every end a pressure opening, the wall
  distance per cell and the pressure probes (simulate.py:205-228), ready for ``simulate.run``.
* ``artery_lattice(case, dx_mm)`` -- validate.py:294-308: the same lumen with a velocity inlet and two outlets.

The centreline input is the story's segment file (the reference's geometry/segment.json): a path or a dict with
``mm_per_unit``, ``junction_point``, ``branches.{trunk,blocked,sibling}.points`` and ``clot.{face_point,
body_length}``.
"""
from __future__ import annotations

import ctypes
import json
import math

import numpy as np

# geometry choices (vessel.py:27-38)
DIAM = {"trunk": [(0.0, 4.0)], "blocked": [(0.0, 2.2), (24.0, 1.9)], "sibling": [(0.0, 2.0)]}
JUNCTION_BLEND = 0.6
EXT_DIAMETERS = 3.0
CLOT_RAMP = 4.0
BEND_LIMIT = 0.5
RESAMPLE = 0.2
CASES = {"healthy": 0.0, "narrow30": 0.3, "narrow40": 0.4, "narrow50": 0.5, "narrow60": 0.6, "narrow70": 0.7,
         "narrow90": 0.9, "occluded": 1.0}
BRANCHES = ("trunk", "blocked", "sibling")
# simulate.py:40-45, 60: grid (dx, dt) per case and the open ends (all pressure openings there)
GRIDS = {"healthy": (1.0e-4, 5.0e-6), "occluded": (1.0e-4, 5.0e-6), "narrow70": (5.0e-5, 1.0e-6),
         "narrow30": (1.0e-4, 2.0e-6), "narrow50": (1.0e-4, 2.0e-6), "narrow90": (5.0e-5, 1.0e-6),
         "narrow40": (1.0e-4, 2.0e-6), "narrow60": (5.0e-5, 1.0e-6)}
OPEN_ENDS = (("trunk", "start", "inlet (ICA)"), ("blocked", "end", "MCA outlet"), ("sibling", "end", "ACA outlet"))


# ---- centreline helpers -----------------------------------------------------------------------------------------
def arc_length(C):
    """Cumulative arc length at each station of a polyline (n, 3)."""
    return np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(C, axis=0), axis=1))])


def resample(C, h):
    """Stations at (nearly) equal spacing h along the polyline, the ends kept."""
    s = arc_length(C)
    m = max(2, int(math.ceil(s[-1] / h)) + 1)
    u = np.linspace(0.0, s[-1], m)
    return np.column_stack([np.interp(u, s, C[:, k]) for k in range(3)])


def end_tangent(C, at_end, span=1.5):
    """Unit direction of the polyline over its last (or first) `span` mm."""
    s = arc_length(C)
    if at_end:
        k = min(int(np.searchsorted(s, s[-1] - span)), len(C) - 2)
        d = C[-1] - C[k]
    else:
        k = max(int(np.searchsorted(s, span)), 1)
        d = C[k] - C[0]
    return d / np.linalg.norm(d)


def radius_profile(table, s):
    """Radius at arc lengths s from a table of (arc, diameter)."""
    xs = [x for x, _ in table]
    ds = [d for _, d in table]
    return 0.5 * np.interp(s, xs, ds)


def curvature(C):
    s = arc_length(C)
    d1 = np.gradient(C, s, axis=0)
    d2 = np.gradient(d1, s, axis=0)
    speed = np.linalg.norm(d1, axis=1)
    return np.linalg.norm(np.cross(d1, d2), axis=1) / np.maximum(speed ** 3, 1e-12)


def limit_bends(C, table, limit=BEND_LIMIT, max_iter=4000):
    """Relax the path where curvature x radius exceeds `limit` (the inside wall of the bend would fold): half a
    Laplacian step on the offending stations and six neighbours either side, ends fixed, resampled every 25
    iterations (vessel.py:91-110). Returns (path, largest distance moved from the input path in mm)."""
    C0 = C.copy()
    window = np.ones(13)
    for it in range(max_iter):
        bad = curvature(C) * radius_profile(table, arc_length(C)) > limit
        bad[:2] = False
        bad[-2:] = False
        if not bad.any():
            break
        move = np.convolve(bad.astype(float), window, "same") > 0
        move[0] = move[-1] = False
        lap = np.zeros_like(C)
        lap[1:-1] = 0.5 * (C[:-2] + C[2:]) - C[1:-1]
        C = C + 0.5 * lap * move[:, None]
        if it % 25 == 24:
            C = resample(C, RESAMPLE)
    C = resample(C, RESAMPLE)
    moved = np.min(np.linalg.norm(C[:, None, :] - C0[None, :, :], axis=2), axis=1).max()
    return C, float(moved)


def load_segment(segment):
    if isinstance(segment, dict):
        return segment
    with open(segment) as fh:
        return json.load(fh)


# ---- the vessel ---------------------------------------------------------------------------------------------------
def build(case="healthy", severity=None, segment=None):
    """The vessel of one clot case (vessel.py:113-175): {"case", "severity", "branches": {name: {"C", "R0", "R",
    "caps", "segment_from", "segment_to", "path_moved_mm"}}, "clot": {"face_arc", "body_length", "ramp", "weight"},
    "frame"}. segment: the story's segment file (path or dict)."""
    if segment is None:
        raise ValueError("vessel.build: pass the story's segment file (segment=path or dict)")
    seg = load_segment(segment)
    sev = CASES[case] if severity is None else float(severity)
    mm = float(seg["mm_per_unit"])
    J = np.asarray(seg["junction_point"], float)

    def to_mm(p):
        return (np.asarray(p, float) - J) * mm

    origin = np.zeros((1, 3))
    paths = {"trunk": np.vstack([to_mm(seg["branches"]["trunk"]["points"]), origin]),
             "blocked": np.vstack([origin, to_mm(seg["branches"]["blocked"]["points"])]),
             "sibling": np.vstack([origin, to_mm(seg["branches"]["sibling"]["points"])])}
    branches = {}
    for name in BRANCHES:
        C, moved = limit_bends(resample(paths[name], RESAMPLE), DIAM[name])
        R = radius_profile(DIAM[name], arc_length(C))
        if name == "trunk":                        # inlet extension before the trunk's start
            t = end_tangent(C, at_end=False)
            L = EXT_DIAMETERS * 2 * R[0]
            back = np.linspace(L, 0.0, int(L / RESAMPLE) + 1)[:-1]
            ext = C[0] - back[:, None] * t
            C = np.vstack([ext, C])
            R = np.concatenate([np.full(len(ext), R[0]), R])
            caps = {"start": (C[0], -t), "end": None}
            first, last = len(ext), len(C)
        else:                                      # outlet extension after the branch's end
            t = end_tangent(C, at_end=True)
            L = EXT_DIAMETERS * 2 * R[-1]
            fwd = np.linspace(0.0, L, int(L / RESAMPLE) + 1)[1:]
            ext = C[-1] + fwd[:, None] * t
            n_seg = len(C)
            C = np.vstack([C, ext])
            R = np.concatenate([R, np.full(len(ext), R[-1])])
            caps = {"start": None, "end": (C[-1], t)}
            first, last = 0, n_seg
        branches[name] = {"C": C, "R0": R.copy(), "R": R, "caps": caps, "segment_from": first, "segment_to": last,
                          "path_moved_mm": moved}
    # the clot: the story's face point projected on the blocked centreline, its body length, cosine ramps
    Cb = branches["blocked"]["C"]
    sb = arc_length(Cb)
    s_face = float(sb[np.argmin(np.linalg.norm(Cb - to_mm(seg["clot"]["face_point"]), axis=1))])
    body = float(seg["clot"]["body_length"]) * mm
    w = ((sb >= s_face) & (sb <= s_face + body)).astype(float)
    up = (sb > s_face - CLOT_RAMP) & (sb < s_face)
    w[up] = 0.5 - 0.5 * np.cos(np.pi * (sb[up] - (s_face - CLOT_RAMP)) / CLOT_RAMP)
    down = (sb > s_face + body) & (sb < s_face + body + CLOT_RAMP)
    w[down] = 0.5 - 0.5 * np.cos(np.pi * (1.0 - (sb[down] - (s_face + body)) / CLOT_RAMP))
    branches["blocked"]["R"] = branches["blocked"]["R0"] * (1.0 - sev * w)
    return {"case": case, "severity": sev, "branches": branches,
            "clot": {"face_arc": s_face, "body_length": body, "ramp": CLOT_RAMP, "weight": w},
            "frame": {"mm_per_unit": mm, "junction_model": J.tolist()}}


def bounds(g, margin=0.5):
    """Box around every centreline station, padded by the largest radius and `margin` mm (vessel.py:227-230)."""
    P = np.vstack([b["C"] for b in g["branches"].values()])
    rmax = max(float(b["R0"].max()) for b in g["branches"].values())
    return P.min(0) - rmax - margin, P.max(0) + rmax + margin


def packed(g):
    """The tube set in the layout of ``cs_sdf_tubes``: stations (n, 4) float32 = x, y, z, radius; first station and
    count per branch; per branch two caps (start, end): point, outward normal, radius (0 = no cap)."""
    pts, start, count, cap_p, cap_n, cap_r = [], [], [], [], [], []
    n = 0
    for b in g["branches"].values():
        pts.append(np.column_stack([b["C"], b["R"]]))
        start.append(n)
        count.append(len(b["C"]))
        n += len(b["C"])
        for key, r in (("start", b["R"][0]), ("end", b["R"][-1])):
            cap = b["caps"][key]
            cap_p.append(cap[0] if cap is not None else np.zeros(3))
            cap_n.append(cap[1] if cap is not None else np.zeros(3))
            cap_r.append(float(r) if cap is not None else 0.0)
    return (np.ascontiguousarray(np.vstack(pts), np.float32), np.array(start, np.int32), np.array(count, np.int32),
            np.array(cap_p, np.float32), np.array(cap_n, np.float32), np.array(cap_r, np.float32))


def grid_shape(lo, hi, dx):
    return tuple(int(x) for x in np.ceil((np.asarray(hi, float) - np.asarray(lo, float)) / dx).astype(int) + 1)


def sdf_grid(g, dx, lo=None, hi=None, blend=JUNCTION_BLEND):
    """Signed distance (mm, negative inside) on the voxel centres origin + (i, j, k) dx, computed on the GPU by
    ``cs_sdf_tubes``. Returns (float32 array (nx, ny, nz), origin)."""
    import torch

    from . import _native
    if lo is None:
        lo, hi = bounds(g)
    lo = np.asarray(lo, float)
    shape = grid_shape(lo, hi, dx)
    pts, start, count, cap_p, cap_n, cap_r = packed(g)
    dev = torch.device("cuda", torch.cuda.current_device())
    t = {k: torch.from_numpy(v).to(dev) for k, v in
         (("pts", pts), ("start", start), ("count", count), ("cap_p", cap_p), ("cap_n", cap_n), ("cap_r", cap_r))}
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    a = _native.SdfTubesArgs()
    a.pts, a.br_start, a.br_count = t["pts"].data_ptr(), t["start"].data_ptr(), t["count"].data_ptr()
    a.cap_p, a.cap_n, a.cap_r = t["cap_p"].data_ptr(), t["cap_n"].data_ptr(), t["cap_r"].data_ptr()
    a.n_branch = len(start)
    a.blend = float(blend)
    a.origin = (ctypes.c_float * 3)(*[float(x) for x in lo])
    a.dx = float(dx)
    a.n = (ctypes.c_int64 * 3)(*shape)
    a.out = out.data_ptr()
    stream = torch.cuda.current_stream(dev)
    _native.call("cs_sdf_tubes", ctypes.byref(a), ctypes.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return out.cpu().numpy(), lo


# ---- the lattices the reference's callers build from it -----------------------------------------------------------
def probe_points(gh, g):
    """Pressure probes on the healthy centrelines (simulate.py:205-228): ICA, junction, MCA before / in / after the
    clot, the distal ends. Radius half the local healthy radius; the throat probe is limited by the case's lumen."""
    br = gh["branches"]
    Cb = br["blocked"]["C"]
    sb = arc_length(Cb)
    clot = gh["clot"]

    def at(s):
        return Cb[int(np.argmin(np.abs(sb - s)))]

    def r_healthy(s):
        return float(np.interp(s, sb, br["blocked"]["R0"]))

    def r_case(s):
        return float(np.interp(s, sb, g["branches"]["blocked"]["R"]))

    s_up = clot["face_arc"] - clot["ramp"] - 1.0
    s_mid = clot["face_arc"] + clot["body_length"] / 2
    s_dn = clot["face_arc"] + clot["body_length"] + clot["ramp"] + 3.0
    tr, sib, bl = br["trunk"], br["sibling"], br["blocked"]
    pts = [("ICA", tr["C"][tr["segment_from"]], 0.5 * float(tr["R0"][tr["segment_from"]])),
           ("junction", np.zeros(3), 0.5),
           ("MCA before clot", at(s_up), 0.5 * r_healthy(s_up)),
           ("clot throat", at(s_mid), max(min(0.5 * r_healthy(s_mid), 0.8 * r_case(s_mid)), 0.0)),
           ("MCA after clot", at(s_dn), 0.5 * r_healthy(s_dn)),
           ("MCA distal", bl["C"][bl["segment_to"] - 1], 0.5 * float(bl["R0"][bl["segment_to"] - 1])),
           ("ACA distal", sib["C"][sib["segment_to"] - 1], 0.5 * float(sib["R0"][sib["segment_to"] - 1]))]
    return [{"name": n, "point": np.asarray(p, float), "radius": r} for n, p, r in pts]


def _cap_opening(b, key, kind, label=None):
    c, nrm = b["caps"][key]
    r = float(b["R"][0] if key == "start" else b["R"][-1])
    o = {"kind": kind, "centre": np.asarray(c, float), "normal": np.asarray(nrm, float), "radius": r}
    if label is not None:
        o["label"] = label
    return o


def artery_case(case, dx_mm=None, severity=None, segment=None):
    """simulate.py:155-200 for the story artery: the case's lumen on the healthy case's box, every end a pressure
    opening (inlet first), the wall distance per cell (mm), the wall points with their normals (the signed distance's
    zero crossings, "wall_verts" / "wall_normals") and the probes. Returns (lattice, ends, probes)."""
    from .lattice import build_lattice
    from .wss import wall_surface
    if dx_mm is None:
        dx_mm = GRIDS[case][0] * 1e3
    gh = build("healthy", segment=segment)
    g = gh if (case == "healthy" and severity is None) else build(case, severity, segment=segment)
    lo, hi = bounds(gh)
    phi, origin = sdf_grid(g, dx_mm, lo, hi)
    ends = [_cap_opening(g["branches"][name], key, "outlet", label) for name, key, label in OPEN_ENDS]
    lat = build_lattice(phi, origin, dx_mm, openings=ends)
    lat["wall_mm"] = -phi[tuple(lat["coords"].T)]
    lat["wall_verts"], lat["wall_normals"] = wall_surface(phi, origin, dx_mm, ends)
    return lat, ends, probe_points(gh, g)


def artery_lattice(case="healthy", dx_mm=0.1, segment=None):
    """validate.py:294-308: the lumen on its own box with a velocity inlet at the trunk and outlets at the branch
    ends (inlets first). Returns (lattice with "openings", vessel)."""
    from .lattice import build_lattice
    g = build(case, segment=segment)
    phi, origin = sdf_grid(g, dx_mm)
    openings = []
    for name, b in g["branches"].items():
        for key, kind in (("start", "inlet"), ("end", "outlet")):
            if b["caps"][key] is not None:
                o = _cap_opening(b, key, kind)
                o["branch"] = name
                openings.append(o)
    openings.sort(key=lambda o: o["kind"] != "inlet")
    lat = build_lattice(phi, origin, dx_mm, openings=openings)
    lat["openings"] = openings
    return lat, g
