"""Deterministic sequential specification of one refinement pass (test
infrastructure only).

Restates src/mfem/refinement/refinement.py:914-1322 (refine, "legacy" and
"geometric" scoring) as a sequential algorithm; the reference's own pass is
race-dependent (non-atomic candidate flags, first-inserter ownership of new
vertices, float atomic_add scores), so its output is defined here and the
GPU pass must reproduce this one bit for bit:

1. Candidate edges: every tet edge whose endpoints are not both pinned
   (inv_mass == 0).  Score: if the edge midpoint is inside a shape (signed
   distance < 0) the penetrating score, else the sum over incident tets of
   (tet_score * w_tet + min(vertex_score[a], vertex_score[b]) * w_vertex) * L,
   L = |x_b - x_a| (populate_candidates, refinement.py:74-154).  Each
   contribution is computed in FP32 with round-to-nearest operations and
   accumulated as a 64-bit integer, round(c * 2^40 / threshold), so the sum
   is independent of order; an edge is a candidate when its integer score
   exceeds 2^40 (score > threshold).
2. Selection: edges in descending (score, edge key) order; an edge is
   selected when none of its tets is already split; its tets are then
   split (get_tet_candidate / remove_conflicting_candidates /
   finalize_nonconflicting_candidates / udpate_tet_candidate: at most one
   split edge per tet, highest score first).
3. Capacity: a pass that would exceed max_tets or max_vertices is dropped
   (check_split_capacity, drop_candidates_on_overflow).
4. New vertices: one per selected edge, numbered V + k in the order of the
   lowest-index tet on each edge; x, v and rest position interpolated from
   the edge's lower-id endpoint a to its higher-id endpoint b:
   a + (b - a) * split_t (claim_new_vertices).
5. Tets in order; a split tet becomes two children at consecutive slots:
   the local edge (i, j), i > j, of the split edge found by the reference's
   loop order; child 0 has corner i replaced by the new vertex, child 1
   corner j (scatter_tets).  Children copy stretch, lambda and materials;
   DmInv = inverse of the child's rest edge matrix.  Surface triangles the
   same way (get_tri_candidate, scatter_tris).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
SCORE_ONE = 2.0 ** 40
SCORE_CAP = 2 ** 56


def _rn32(a):
    return np.asarray(a, F32)


def edge_key(a, b):
    a, b = (a, b) if a < b else (b, a)
    return (int(a) << 32) | int(b)


def tet_edges(tet):
    """The 6 edges in the reference's loop order: (t[i], t[j]) for i in 0..3,
    j < i, with their local indices."""
    out = []
    for i in range(4):
        for j in range(i):
            out.append((int(tet[i]), int(tet[j]), i, j))
    return out


def contribution_int(c32, threshold):
    q = SCORE_ONE / float(threshold)
    v = int(np.rint(float(c32) * q))
    return min(max(v, 0), SCORE_CAP)


def edge_contribution(x, a, b, tet_score, vs_a, vs_b, w_tet, w_vert):
    """(tet_score * w_tet + min(vs_a, vs_b) * w_vert) * |x_b - x_a| in FP32,
    each operation rounded to nearest (no fused multiply-add)."""
    d = _rn32(x[b]) - _rn32(x[a])
    # the sums are evaluated left to right in FP32
    L = np.sqrt(F32(F32(F32(d[0] * d[0]) + F32(d[1] * d[1])) + F32(d[2] * d[2])), dtype=F32)
    s = F32(F32(F32(tet_score) * F32(w_tet)) + F32(F32(min(F32(vs_a), F32(vs_b))) * F32(w_vert)))
    return F32(s * L)


def legacy_scores(tets, x, inv_mass, tet_scores, vertex_scores, w_tet, w_vert, pen_score, threshold,
                  midpoint_distance=None):
    """{edge key: integer score} of every candidate edge (step 1).
    midpoint_distance(p) -> signed distance to the shapes (None: no shapes)."""
    scores: dict[int, int] = {}
    pen: set[int] = set()
    for t, tet in enumerate(tets):
        for a, b, _, _ in tet_edges(tet):
            if inv_mass[a] == 0 and inv_mass[b] == 0:
                continue
            k = edge_key(a, b)
            scores.setdefault(k, 0)
            if midpoint_distance is not None:
                mid = F32(0.5) * (_rn32(x[a]) + _rn32(x[b]))
                if midpoint_distance(mid) < 0:
                    pen.add(k)
                    continue
            c = edge_contribution(x, a, b, tet_scores[t], vertex_scores[a], vertex_scores[b], w_tet, w_vert)
            scores[k] = min(scores[k] + contribution_int(c, threshold), SCORE_CAP)
    pen_int = contribution_int(F32(pen_score), threshold)
    for k in pen:
        scores[k] = max(scores[k], pen_int)
    return scores


# --- "geometric" scoring (populate_candidates_geometric,
# populate_tri_candidates_geometric, elastic_density_stats;
# refinement.py:123-317).  The reference combines contributions with
# atomic_max (order-free) but sums the densities with float atomic_add
# (order-dependent); here every quantity is FP64 with each operation rounded
# to nearest, the density sum in the GPU's fixed tree order, and the
# per-edge maximum is taken on the integer priorities round(c 2^40 / threshold).

CAPSULE = 1


def tet_density(tet_score, mu, dminv9):
    """max(psi, 0) / (vol mu), vol = |1 / (6 det DmInv)| (0 for degenerate data)."""
    m = [float(F32(v)) for v in np.asarray(dminv9).reshape(9)]
    c0 = m[4] * m[8] - m[5] * m[7]
    c1 = m[3] * m[8] - m[5] * m[6]
    c2 = m[3] * m[7] - m[4] * m[6]
    det = (m[0] * c0 - m[1] * c1) + m[2] * c2
    vol = abs(1.0 / (6.0 * det)) if det != 0.0 else float("inf")
    mu = float(F32(mu))
    if mu > 0.0 and vol > 0.0 and np.isfinite(vol):
        e = float(F32(tet_score)) if F32(tet_score) > 0 else 0.0
        return e / (vol * mu)
    return 0.0


def density_total(rho, block=256):
    """Sum of the densities in the GPU's order: a pairwise tree inside each
    block of 256 tets, then per-thread strided sums of the block sums and a
    pairwise tree over the 256 threads."""
    rho = [float(v) for v in rho]
    n_blk = (len(rho) + block - 1) // block
    part = []
    for b in range(n_blk):
        sh = rho[b * block:(b + 1) * block] + [0.0] * max(0, (b + 1) * block - len(rho))
        w = block // 2
        while w > 0:
            for i in range(w):
                sh[i] = sh[i] + sh[i + w]
            w //= 2
        part.append(sh[0])
    sh = []
    for t in range(block):
        acc = 0.0
        for i in range(t, n_blk, block):
            acc = acc + part[i]
        sh.append(acc)
    w = block // 2
    while w > 0:
        for i in range(w):
            sh[i] = sh[i] + sh[i + w]
        w //= 2
    return sh[0]


def _len64(x, a, b):
    xa = [float(F32(v)) for v in x[a]]
    xb = [float(F32(v)) for v in x[b]]
    d0, d1, d2 = xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]
    return float(np.sqrt((d0 * d0 + d1 * d1) + d2 * d2))


def _prox(d, d1):
    p = (d1 - float(F32(d))) / d1
    return p if p > 0.0 else 0.0


def _priority(c, threshold):
    v = c * (SCORE_ONE / float(F32(threshold)))
    f = int(np.rint(v)) if v > 0.0 else 0
    return min(f, SCORE_CAP)


def geometric_scores(tets, x, inv_mass, tet_scores, mu, dminv, vertex_dist, vertex_sid, shape_type, h_min, d1,
                     w_e, w_c, threshold, tris=None, tri_dist=None, tri_bary=None, w_t=1.0,
                     elastic_total=None):
    """{edge key: integer priority} of every registered edge (every edge whose
    endpoints are not both pinned; 0 when it may not split).  elastic_total
    = the density sum (None: computed over these tets in the GPU's order;
    0 means absolute densities)."""
    h = float(F32(h_min))
    d1 = float(F32(d1))
    w_e, w_c, w_t = float(F32(w_e)), float(F32(w_c)), float(F32(w_t))
    rho = [tet_density(tet_scores[t], mu[t], dminv[t]) for t in range(len(tets))]
    total = density_total(rho) if elastic_total is None else float(elastic_total)
    scores: dict[int, int] = {}
    for t, tet in enumerate(tets):
        s_el = rho[t]
        if total > 0.0 and len(tets) > 0:
            mean = total / float(len(tets))
            ex = s_el / mean - 1.0
            s_el = ex if ex > 0.0 else 0.0
        l_max = 0.0
        for a, b, _, _ in tet_edges(tet):
            l_max = max(l_max, _len64(x, a, b))
        l_max = max(l_max, 1.0e-20)
        for a, b, _, _ in tet_edges(tet):
            if inv_mass[a] == 0 and inv_mass[b] == 0:
                continue
            k = edge_key(a, b)
            scores.setdefault(k, 0)
            L = _len64(x, a, b)
            if L < 2.0 * h:
                continue
            s_c = 0.0
            for v in (a, b):
                sid = int(vertex_sid[v])
                if 0 <= sid < len(shape_type) and int(shape_type[sid]) == CAPSULE:
                    s_c = max(s_c, _prox(vertex_dist[v], d1))
            c = ((L / h) * (L / l_max)) * (w_e * s_el + w_c * s_c)
            scores[k] = max(scores[k], _priority(c, threshold))
    if tris is not None and tri_dist is not None and tri_bary is not None:
        for f, tv in enumerate(tris):
            d = float(F32(tri_dist[f]))
            if not d < d1:
                continue
            prox = _prox(d, d1)
            l_max = 1.0e-20
            for k in range(3):
                l_max = max(l_max, _len64(x, int(tv[k]), int(tv[(k + 1) % 3])))
            for k in range(3):
                a, b = int(tv[(k + 1) % 3]), int(tv[(k + 2) % 3])
                if inv_mass[a] == 0 and inv_mass[b] == 0:
                    continue
                L = _len64(x, a, b)
                if L < 2.0 * h:
                    continue
                key = edge_key(a, b)
                if key not in scores:
                    continue
                w = 1.0 - float(F32(tri_bary[f][k]))
                c = (((L / h) * (L / l_max)) * w_t) * prox * w
                scores[key] = max(scores[key], _priority(c, threshold))
    return scores


def select_edges(tets, scores):
    """Greedy selection (step 2): {edge key: True} of the split edges."""
    edge_tets: dict[int, list[int]] = {}
    for t, tet in enumerate(tets):
        for a, b, _, _ in tet_edges(tet):
            k = edge_key(a, b)
            edge_tets.setdefault(k, []).append(t)
    cand = [(s, k) for k, s in scores.items() if s > SCORE_ONE]
    cand.sort(key=lambda sk: (-sk[0], -sk[1]))
    split = np.zeros(len(tets), bool)
    chosen = []
    for s, k in cand:
        ts = edge_tets[k]
        if any(split[t] for t in ts):
            continue
        for t in ts:
            split[t] = True
        chosen.append(k)
    return set(chosen)


def refine_pass(tets, tris, x, v, rest, stretch, lam, materials, selected, n_verts, max_tets, max_verts,
                split_t=0.5):
    """Steps 3-5.  Returns a dict of the new arrays (or the inputs when the
    pass is dropped / empty) and the number of new vertices."""
    tets = np.asarray(tets)
    T = len(tets)
    owner: dict[int, int] = {}
    tet_edge = np.full(T, -1, np.int64)
    for t, tet in enumerate(tets):
        for a, b, i, j in tet_edges(tet):
            k = edge_key(a, b)
            if k in selected:
                tet_edge[t] = k
                owner.setdefault(k, t)
    n_split = int((tet_edge >= 0).sum())
    n_new = len(selected)
    dropped = T + n_split > max_tets or n_verts + n_new > max_verts
    if dropped or n_new == 0:
        return dict(tets=tets.copy(), tris=np.asarray(tris).copy(), n_verts=n_verts, dropped=dropped, new_vertex={})
    # new vertex ids in owner-tet order
    order = sorted(owner.items(), key=lambda kv: kv[1])
    new_vertex = {k: n_verts + i for i, (k, _) in enumerate(order)}
    def grown(a):
        out = np.zeros((n_verts + n_new, 3), F32)
        out[:n_verts] = np.asarray(a, F32)[:n_verts]
        return out

    X, Vv, R = grown(x), grown(v), grown(rest)
    ts = F32(split_t)
    for k, nv in new_vertex.items():
        a, b = k >> 32, k & 0xFFFFFFFF
        for arr in (X, Vv, R):
            arr[nv] = arr[a] + (arr[b] - arr[a]) * ts
    new_tets, parent = [], []
    for t, tet in enumerate(tets):
        if tet_edge[t] < 0:
            new_tets.append(tet.copy())
            parent.append(t)
            continue
        k = int(tet_edge[t])
        li = lj = 0
        for a, b, i, j in tet_edges(tet):
            if edge_key(a, b) == k:
                li, lj = i, j
        for loc in (li, lj):
            c = tet.copy()
            c[loc] = new_vertex[k]
            new_tets.append(c)
            parent.append(t)
    new_tets = np.array(new_tets, np.int32)
    parent = np.array(parent)
    new_tris = []
    for tri in np.asarray(tris):
        k = None
        for i in range(3):
            for j in range(i):
                kk = edge_key(tri[i], tri[j])
                if kk in new_vertex:
                    k, li, lj = kk, i, j
        if k is None:
            new_tris.append(tri.copy())
            continue
        for loc in (li, lj):
            c = tri.copy()
            c[loc] = new_vertex[k]
            new_tris.append(c)
    return dict(tets=new_tets, tris=np.array(new_tris, np.int32).reshape(-1, 3), x=X, v=Vv, rest=R,
                stretch=np.asarray(stretch)[parent], lam=np.asarray(lam)[parent],
                materials=np.asarray(materials)[parent], parent=parent, n_verts=n_verts + n_new, dropped=False,
                new_vertex=new_vertex)


def rest_dminv(rest, tets):
    """DmInv of every tet from rest positions (FP64)."""
    R = np.asarray(rest, np.float64)
    t = np.asarray(tets, np.int64)
    Dm = np.stack([R[t[:, 1]] - R[t[:, 0]], R[t[:, 2]] - R[t[:, 0]], R[t[:, 3]] - R[t[:, 0]]], axis=2)
    return np.linalg.inv(Dm)
