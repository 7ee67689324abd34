"""CPU restatement of the surface-tracking loss's point-to-surface query
(test infrastructure only: tests/ compare the GPU query with it).

Follows src/mfem/refinement/surface_loss.py:43-96 (closest point on a
triangle, Ericson 5.1.5 with the reference's region priority) and
:206-245 / :456-502 (k nearest vertices by a KD-tree, projection onto their
incident triangles, nearest-vertex fallback).  scipy's cKDTree stands in
for the reference's own cKDTree call.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def closest_point_on_triangles(p, a, b, c):
    """Closest point to each p on its triangle (a, b, c), broadcasting over
    leading axes; vertex regions take precedence over edges over the face
    (surface_loss.py:43-96)."""
    p, a, b, c = (np.asarray(t, dtype=np.float64) for t in (p, a, b, c))
    ab, ac, ap, bp, cp = b - a, c - a, p - a, p - b, p - c
    dot = lambda u, v: np.einsum("...i,...i->...", u, v)
    d1, d2, d3, d4, d5, d6 = dot(ab, ap), dot(ac, ap), dot(ab, bp), dot(ac, bp), dot(ab, cp), dot(ac, cp)
    va, vb, vc = d3 * d6 - d5 * d4, d5 * d2 - d1 * d6, d1 * d4 - d3 * d2
    den = va + vb + vc
    inv = np.where(den == 0.0, 0.0, 1.0 / np.where(den == 0.0, 1.0, den))
    out = a + ab * (vb * inv)[..., None] + ac * (vc * inv)[..., None]
    ratio = lambda n, d: n / np.where(d == 0.0, 1.0, d)
    cases = [  # lowest priority first: each later case overrides
        ((va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0), lambda: b + (c - b) * ratio(d4 - d3, (d4 - d3) + (d5 - d6))[..., None]),
        ((vb <= 0) & (d2 >= 0) & (d6 <= 0), lambda: a + ac * ratio(d2, d2 - d6)[..., None]),
        ((vc <= 0) & (d1 >= 0) & (d3 <= 0), lambda: a + ab * ratio(d1, d1 - d3)[..., None]),
        ((d6 >= 0) & (d5 <= d6), lambda: np.broadcast_to(c, out.shape)),
        ((d3 >= 0) & (d4 <= d3), lambda: np.broadcast_to(b, out.shape)),
        ((d1 <= 0) & (d2 <= 0), lambda: np.broadcast_to(a, out.shape)),
    ]
    for mask, val in cases:
        out = np.where(mask[..., None], val(), out)
    return out


def vertex_faces(faces, n_vertices):
    """(n_vertices, max degree) incident triangles in triangle order, -1 padded."""
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    lists = [[] for _ in range(n_vertices)]
    for f, tri in enumerate(faces):
        for v in tri:
            lists[v].append(f)
    deg = max([len(l) for l in lists] + [1])
    out = np.full((n_vertices, deg), -1, dtype=np.int64)
    for v, l in enumerate(lists):
        out[v, :len(l)] = l
    return out


def point_mesh_query(query, verts, faces=None, *, valid=None, face_ok=None, k=1, project=True):
    """(sq distance (Q,), face (Q,) or -1, closest point (Q, 3), nearest vertex
    (Q,)) of every query point, as the reference measures them."""
    query = np.asarray(query, dtype=np.float64).reshape(-1, 3)
    verts = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    vmap = np.nonzero(valid)[0] if valid is not None else np.arange(len(verts))
    k = min(int(k), len(vmap))
    dist, loc = cKDTree(verts[vmap]).query(query, k=k)
    dist = np.asarray(dist).reshape(len(query), k)
    near = vmap[np.asarray(loc).reshape(len(query), k)]
    vert_sq = dist[:, 0] ** 2
    if not project or faces is None or len(faces) == 0:
        return vert_sq, np.full(len(query), -1), verts[near[:, 0]], near[:, 0]
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    vf = vertex_faces(faces, len(verts))
    cand = vf[near].reshape(len(query), -1)
    ok = cand >= 0
    if face_ok is not None:
        ok &= np.asarray(face_ok, dtype=bool)[np.where(ok, cand, 0)]
    cs = np.where(ok, cand, 0)
    tri = verts[faces[cs]]
    qexp = np.broadcast_to(query[:, None, :], tri[..., 0, :].shape)
    cp = closest_point_on_triangles(qexp, tri[..., 0, :], tri[..., 1, :], tri[..., 2, :])
    sq = np.where(ok, ((cp - qexp) ** 2).sum(-1), np.inf)
    best = sq.argmin(axis=1)
    rows = np.arange(len(query))
    has = np.isfinite(sq[rows, best])
    return (np.where(has, sq[rows, best], vert_sq), np.where(has, cand[rows, best], -1),
            np.where(has[:, None], cp[rows, best], verts[near[:, 0]]), near[:, 0])
