"""NumPy restatement of the self-contact pieces (test infrastructure only).

Follows the reference's collision primitives and IPC barrier
(src/mfem/collision.py:220-546), its scan semantics (src/mfem/scan.py,
tests/test_scan.py: exclusive scan = cumsum - a) and, for the broad phase
(which the reference does not have: its SAP and hash table are unwired,
src/mfem/collision.py:840-895, src/mfem/refinement/geometry_hash.py), the
brute-force O(n^2) candidate enumeration that a broad phase must reproduce.

Pinned by tests/test_oracle_collide.py against the reference's own known
answers (tests/test_collision.py) and closed forms.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def closest_point_on_triangle(x, a, b, c):
    """Closest point of triangle abc to x (Voronoi regions; Ericson 5.1.5),
    src/mfem/collision.py:220-268."""
    x, a, b, c = (np.asarray(v, np.float64) for v in (x, a, b, c))
    ab, ac, ax = b - a, c - a, x - a
    d1, d2 = ab @ ax, ac @ ax
    if d1 <= 0 and d2 <= 0:
        return a
    bx = x - b
    d3, d4 = ab @ bx, ac @ bx
    if d3 >= 0 and d4 <= d3:
        return b
    cx = x - c
    d5, d6 = ab @ cx, ac @ cx
    if d6 >= 0 and d5 <= d6:
        return c
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + (d1 / (d1 - d3)) * ab
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + (d2 / (d2 - d6)) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + (vb * den) * ab + (vc * den) * ac


def point_triangle_dist(x, a, b, c) -> float:
    return float(np.linalg.norm(np.asarray(x, np.float64) - closest_point_on_triangle(x, a, b, c)))


def closest_point_segment_segment(p1, q1, p2, q2):
    """(c1, c2, s) of segments p1q1 and p2q2 (Ericson 5.1.9),
    src/mfem/collision.py:276-322 (epsilon 1e-8, denominator test > eps)."""
    p1, q1, p2, q2 = (np.asarray(v, np.float64) for v in (p1, q1, p2, q2))
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    eps = 1.0e-8
    if a <= eps and e <= eps:
        s = t = 0.0
    elif a <= eps:
        s, t = 0.0, float(np.clip(f / e, 0.0, 1.0))
    else:
        c = d1 @ r
        if e <= eps:
            t, s = 0.0, float(np.clip(-c / a, 0.0, 1.0))
        else:
            b = d1 @ d2
            den = a * e - b * b
            s = float(np.clip((b * f - c * e) / den, 0.0, 1.0)) if den > eps else 0.0
            t = (b * s + f) / e
            if t < 0:
                t, s = 0.0, float(np.clip(-c / a, 0.0, 1.0))
            elif t > 1:
                t, s = 1.0, float(np.clip((b - c) / a, 0.0, 1.0))
    return p1 + s * d1, p2 + t * d2, s


def edge_edge_dist(sa, sb, ra, rb) -> float:
    c1, c2, _ = closest_point_segment_segment(sa, sb, ra, rb)
    return float(np.linalg.norm(c1 - c2))


def soft_rigid_distance(points, tris, rigid_points_count, rigid_tri_count, pairs, pair_count, d_hat, distance):
    """Per soft particle, the minimum over the 15 feature distances of every
    rigid/soft candidate triangle pair below d_hat (distance pre-filled with a
    sentinel >= d_hat, only lowered), src/mfem/collision.py:380-476."""
    distance = np.array(distance, np.float64)
    for k in range(int(pair_count)):
        ta, tb = int(pairs[k][0]), int(pairs[k][1])
        ra_, rb_ = ta < rigid_tri_count, tb < rigid_tri_count
        if ra_ == rb_:
            continue
        tr, ts = (ta, tb) if ra_ else (tb, ta)
        r = [points[i] for i in tris[tr]]
        si = list(tris[ts])
        s = [points[i] for i in si]
        p = [i - rigid_points_count for i in si]

        def put(idx, d):
            if d < d_hat:
                distance[idx] = min(distance[idx], d)
        for j in range(3):
            put(p[j], point_triangle_dist(s[j], *r))
        for j in range(3):
            d = point_triangle_dist(r[j], *s)
            for q in range(3):
                put(p[q], d)
        for i0, i1 in ((0, 1), (1, 2), (2, 0)):
            for j0, j1 in ((0, 1), (1, 2), (2, 0)):
                d = edge_edge_dist(s[i0], s[i1], r[j0], r[j1])
                put(p[i0], d)
                put(p[i1], d)
    return distance


def ipc_barrier(d, d_hat):
    """IPC log barrier b = -(d - dh)^2 ln(d / dh) and its first two
    derivatives in d, for 0 < d < dh (src/mfem/collision.py:529-546)."""
    d = np.asarray(d, np.float64)
    l = np.log(d / d_hat)
    e = -(d - d_hat) ** 2 * l
    de = -2.0 * (d - d_hat) * l - (d - d_hat) ** 2 / d
    d2e = -2.0 * (l + (d - d_hat) / d) - (2.0 * d * (d - d_hat) - (d - d_hat) ** 2) / (d * d)
    return e, de, d2e


def exclusive_scan(a):
    a = np.asarray(a, np.int64)
    return (np.cumsum(a) - a).astype(np.int32)


# ---------------------------------------------------------------------------
# broad phase (brute force): the candidate sets a spatial hash must return
# ---------------------------------------------------------------------------
def surface_edges(tris):
    """Unique edges (i < j) of the triangles, sorted."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0).astype(np.int32)


def surface_points(tris):
    return np.unique(np.asarray(tris, np.int32).reshape(-1))


def _boxes(x, prims, r):
    """FP32 AABBs of primitives (rows of vertex ids) inflated by r, computed
    exactly as the kernels do (min / max then -r / +r in FP32)."""
    x = np.asarray(x, F32)
    pts = x[np.asarray(prims, np.int64)]
    lo = pts.min(axis=1) - F32(r)
    hi = pts.max(axis=1) + F32(r)
    return lo.astype(F32), hi.astype(F32)


def brute_force_pt(x, points, tris, d_hat):
    """Sorted (point, tri) pairs whose boxes inflated by d_hat / 2 overlap,
    the point not a vertex of the triangle."""
    r = F32(d_hat) * F32(0.5)
    plo, phi = _boxes(x, np.asarray(points)[:, None], r)
    tlo, thi = _boxes(x, tris, r)
    tris = np.asarray(tris)
    out = []
    for k, p in enumerate(points):
        ov = np.all((plo[k] <= thi) & (tlo <= phi[k]), axis=1)
        ov &= ~np.any(tris == p, axis=1)
        for t in np.nonzero(ov)[0]:
            out.append((int(p), int(t)))
    return np.array(sorted(out), np.int32).reshape(-1, 2)


def brute_force_ee(x, edges, d_hat):
    """Sorted (e1, e2) edge-index pairs, e1 < e2, no shared vertex, boxes
    inflated by d_hat / 2 overlapping."""
    r = F32(d_hat) * F32(0.5)
    lo, hi = _boxes(x, edges, r)
    edges = np.asarray(edges)
    out = []
    for a in range(len(edges)):
        ov = np.all((lo[a] <= hi[a + 1:]) & (lo[a + 1:] <= hi[a]), axis=1)
        share = np.any(edges[a + 1:][:, :, None] == edges[a][None, None, :], axis=(1, 2))
        for b in np.nonzero(ov & ~share)[0]:
            out.append((a, a + 1 + int(b)))
    return np.array(out, np.int32).reshape(-1, 2)
