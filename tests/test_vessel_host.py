"""The story artery's geometry (vessel.py counterpart) on the host, against the reference's Step 2 report
(reports/step2_report.md:5-21: path relaxation, curvature, clot position, lumen volume), and the numpy restatement of
the signed distance against closed forms."""
import math
import os

import numpy as np
import pytest

from cerebrosim_amd import vessel as V
from oracle.sdf import sdf_points, smooth_min

SEGMENT = os.path.join(os.path.dirname(__file__), "golden", "segment_story.json")


@pytest.fixture(scope="module")
def healthy():
    return V.build("healthy", segment=SEGMENT)


def test_relaxed_paths_match_the_report(healthy):
    # step2_report.md:7: moved at most 1.17 / 0.04 / 1.20 mm; max curvature x radius 0.65 / 0.48 / 0.49
    moved = {"trunk": 1.17, "blocked": 0.04, "sibling": 1.20}
    kr = {"trunk": 0.65, "blocked": 0.48, "sibling": 0.49}
    for name, b in healthy["branches"].items():
        assert abs(b["path_moved_mm"] - moved[name]) <= 0.006, (name, b["path_moved_mm"])
        assert abs((V.curvature(b["C"]) * b["R0"]).max() - kr[name]) <= 0.006, name
    # step2_report.md:9: the clot 10.15 mm past the junction, body 2.12 mm
    assert abs(healthy["clot"]["face_arc"] - 10.15) < 0.006 and abs(healthy["clot"]["body_length"] - 2.12) < 0.006


def test_extensions_caps_and_radii(healthy):
    br = healthy["branches"]
    tr = br["trunk"]
    assert tr["caps"]["end"] is None and br["blocked"]["caps"]["start"] is None
    # straight extension of 3 end diameters: the inlet cap 12 mm before the trunk's first segment station
    ext = np.linalg.norm(tr["C"][tr["segment_from"]] - tr["caps"]["start"][0])
    assert abs(ext - 12.0) < RESAMPLE_TOL
    assert np.allclose(tr["R0"], 2.0) and np.allclose(br["sibling"]["R0"], 1.0)
    bl = br["blocked"]
    assert abs(bl["R0"][0] - 1.1) < 1e-12 and abs(bl["R0"][-1] - 0.95) < 1e-12
    for b in br.values():                                  # unit outward normals
        for cap in b["caps"].values():
            if cap is not None:
                assert abs(np.linalg.norm(cap[1]) - 1) < 1e-12


RESAMPLE_TOL = V.RESAMPLE + 1e-9


def test_clot_scales_the_blocked_radius():
    g = V.build("narrow70", segment=SEGMENT)
    b = g["branches"]["blocked"]
    ratio = b["R"] / b["R0"]
    assert abs(ratio.min() - 0.3) < 1e-12                  # 70 % diameter reduction over the body
    w = g["clot"]["weight"]
    assert w.max() == 1.0 and np.all((w >= 0) & (w <= 1))
    s = V.arc_length(b["C"])
    span = s[w > 0]
    assert span.max() - span.min() <= g["clot"]["body_length"] + 2 * V.CLOT_RAMP + 1e-9


def test_sdf_oracle_closed_forms():
    # one straight tube of radius 1 along z from 0 to 10, capped at both ends: |r| - 1 on the wall side,
    # the cap planes near the ends
    pts = np.array([[0, 0, 0, 1.0], [0, 0, 10, 1.0]])
    cap_p = np.array([[0, 0, 0], [0, 0, 10.0]])
    cap_n = np.array([[0, 0, -1.0], [0, 0, 1.0]])
    P = np.array([[0.3, 0.4, 5.0], [2.0, 0.0, 5.0], [0.0, 0.0, -0.5], [0.2, 0.0, 10.3], [0.0, 0.0, 5.0]])
    d = sdf_points(P, pts, [0], [2], cap_p, cap_n, [1.0, 1.0], blend=0.6)
    assert np.allclose(d, [0.5 - 1, 1.0, 0.5, 0.3, -1.0])
    # a cone, radius 1 -> 2 over z 0..10: distance to the axis minus the radius at the projection
    cone = np.array([[0, 0, 0, 1.0], [0, 0, 10, 2.0]])
    d = sdf_points(np.array([[0.0, 0.0, 5.0], [1.0, 0.0, 2.0], [0.0, 3.0, 12.0]]), cone, [0], [2], cap_p, cap_n,
                   [0.0, 0.0], blend=0.6)
    assert np.allclose(d, [-1.5, 1.0 - 1.2, math.hypot(3.0, 2.0) - 2.0])
    # smooth minimum: the plain minimum when the two distances differ by more than the blend width
    assert smooth_min(1.0, 3.0, 0.6) == 1.0 and smooth_min(3.0, 1.0, 0.6) == 1.0
    assert abs(smooth_min(1.0, 1.0, 0.6) - (1.0 - 0.15)) < 1e-15


def test_lumen_volume_matches_the_report(healthy):
    """step2_report.md:15-17: 620 mm^3 of fluid at 0.1 mm. Counted here at 0.25 mm on the voxels near a station
    (the others are outside by construction)."""
    from scipy.spatial import cKDTree
    dx = 0.25
    lo, hi = V.bounds(healthy)
    n = V.grid_shape(lo, hi, dx)
    pts, st, cnt, cp, cn, cr = V.packed(healthy)
    grid = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) * dx + lo
    near = cKDTree(pts[:, :3].astype(float)).query(grid, distance_upper_bound=2.6)[0] < np.inf
    d = sdf_points(grid[near], pts, st, cnt, cp, cn, cr, V.JUNCTION_BLEND)
    vol = (d < 0).sum() * dx ** 3
    assert abs(vol - 620.0) < 8.0, vol


def test_openings_match_the_published_artery(healthy):
    """geometry/solver.json "artery".openings: the inlet at the trunk's start and the two outlets, built by the
    reference's validate.py:294-305 from the same vessel."""
    import json
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_data.json")))["artery"]
    mine = []
    for name, b in healthy["branches"].items():
        for key, kind in (("start", "inlet"), ("end", "outlet")):
            if b["caps"][key] is not None:
                r = b["R"][0] if key == "start" else b["R"][-1]
                mine.append((kind, name, b["caps"][key][0], b["caps"][key][1], r))
    mine.sort(key=lambda o: o[0] != "inlet")
    assert len(mine) == len(ref["openings"]) == 3
    for (kind, name, c, n, r), o in zip(mine, ref["openings"]):
        assert kind == o["kind"] and name == o["branch"]
        assert np.abs(np.asarray(c) - o["centre"]).max() < 1e-9, (name, c, o["centre"])
        assert np.abs(np.asarray(n) - o["normal"]).max() < 1e-9
        assert abs(r - o["radius"]) < 1e-12
