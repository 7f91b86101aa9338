"""ORACLE -- TEST INFRASTRUCTURE ONLY. numpy FP64 restatement of the vessel signed distance k_sdf
(vessel.py:183-207), the checker of the cs_sdf_tubes kernel: per branch the minimum over its segments of the
distance to a round cone (radius interpolated along the segment), cut by the flat caps within 2.5 cap radii,
combined over the branches with the polynomial smooth minimum (vessel.py:176-180)."""
import numpy as np


def smooth_min(a, b, k):
    h = np.clip(0.5 + 0.5 * (b - a) / k, 0.0, 1.0)
    return b + (a - b) * h - k * h * (1.0 - h)


def sdf_points(P, pts, br_start, br_count, cap_p, cap_n, cap_r, blend, chunk=65536):
    """Signed distance at points P (m, 3). pts (n, 4) = x, y, z, radius; caps 2 per branch (radius 0 = none)."""
    P = np.asarray(P, np.float64)
    pts = np.asarray(pts, np.float64)
    out = np.empty(len(P))
    for c0 in range(0, len(P), chunk):
        p = P[c0:c0 + chunk]
        d = np.full(len(p), 1.0e9)
        for b in range(len(br_start)):
            S = pts[br_start[b]:br_start[b] + br_count[b]]
            A, B = S[:-1, :3], S[1:, :3]
            ab = B - A
            db = np.full(len(p), 1.0e9)
            for s in range(len(A)):
                ap = p - A[s]
                t = np.clip(ap @ ab[s] / max(ab[s] @ ab[s], 1e-12), 0.0, 1.0)
                r = S[s, 3] + (S[s + 1, 3] - S[s, 3]) * t
                db = np.minimum(db, np.linalg.norm(ap - t[:, None] * ab[s], axis=1) - r)
            for e in range(2):
                c = 2 * b + e
                if cap_r[c] > 0:
                    q = p - np.asarray(cap_p[c], np.float64)
                    near = np.linalg.norm(q, axis=1) < 2.5 * cap_r[c]
                    db = np.where(near, np.maximum(db, q @ np.asarray(cap_n[c], np.float64)), db)
            d = smooth_min(d, db, blend)
        out[c0:c0 + chunk] = d
    return out
