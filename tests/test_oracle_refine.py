"""Pins of the refinement specification (oracle/refine.py) on the CPU.

The reference's own invariants (tests/test_refinement.py:10-27, 87-130):
conserved total rest volume, no orphaned or inverted tets, growth only when
a pass fires.  Plus the properties that define the selection: at most one
split edge per tet, every tet on a split edge split (conforming mesh: every
interior face shared by exactly two tets, the split surface triangles equal
to the boundary of the split tets), and greedy maximality by priority.  A
mistake planted in the selection (no conflict check) breaks conformity.
"""
import numpy as np
import pytest

from mfem_amd import mesh as M
from oracle import refine as R


def _grid(cells=(3, 2, 2), h=0.2):
    p, t = M.box_tet_mesh((0, 0, 0), (cells[0] * h, cells[1] * h, cells[2] * h), cells)
    return p.astype(np.float32), t.astype(np.int32)


def _faces(tets):
    f = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    return np.sort(f, axis=1)


def _check_mesh(p, tets, tris, n_verts):
    assert tets.min() >= 0 and tets.max() < n_verts
    used = np.zeros(n_verts, bool)
    used[tets.ravel()] = True
    assert used.all(), "orphaned vertex"
    Dm = np.stack([p[tets[:, k]] - p[tets[:, 0]] for k in (1, 2, 3)], axis=2).astype(np.float64)
    vol = np.linalg.det(Dm) / 6.0
    assert np.all(vol > 0), "inverted tet"
    f, c = np.unique(_faces(tets), axis=0, return_counts=True)
    assert c.max() <= 2, "non-manifold / non-conforming face"
    boundary = {tuple(x) for x in f[c == 1]}
    assert boundary == {tuple(sorted(x)) for x in tris}, "surface triangles != boundary faces"
    return vol.sum()


def _scores(p, t, rng, inv_mass, threshold=1.0):
    vs = rng.uniform(0, 30, len(p)).astype(np.float32)
    ts = rng.uniform(0, 5, len(t)).astype(np.float32)
    return R.legacy_scores(t, p, inv_mass, ts, vs, 0.3, 0.1, 1.0e6, threshold)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pass_conserves_volume_and_conforms(seed):
    p, t = _grid()
    tris = M.surface_tris(t)
    rng = np.random.default_rng(seed)
    inv_mass = np.ones(len(p), np.float32)
    inv_mass[p[:, 0] < 1e-6] = 0  # pinned face
    vol0 = _check_mesh(p, t, tris, len(p))
    sc = _scores(p, t, rng, inv_mass)
    sel = R.select_edges(t, sc)
    assert len(sel) > 0
    out = R.refine_pass(t, tris, p, np.zeros_like(p), p, np.zeros((len(t), 6)), np.zeros((len(t), 6)),
                        np.ones((len(t), 3)), sel, len(p), 10 * len(t), 10 * len(p))
    assert not out["dropped"]
    nv = out["n_verts"]
    vol = _check_mesh(out["rest"][:nv], out["tets"], out["tris"], nv)
    assert vol == pytest.approx(vol0, abs=1e-6)
    assert len(out["tets"]) > len(t) and nv == len(p) + len(sel)
    # no edge between two pinned vertices was split
    for k in sel:
        a, b = k >> 32, k & 0xFFFFFFFF
        assert not (inv_mass[a] == 0 and inv_mass[b] == 0)
    # new vertices at the midpoints of their edges (split_t = 0.5)
    for k, v in out["new_vertex"].items():
        a, b = k >> 32, k & 0xFFFFFFFF
        np.testing.assert_allclose(out["rest"][v], 0.5 * (p[a] + p[b]), atol=1e-7)


def test_selection_is_greedy_and_maximal():
    p, t = _grid((4, 3, 3), 0.1)
    rng = np.random.default_rng(5)
    inv_mass = np.ones(len(p), np.float32)
    sc = _scores(p, t, rng, inv_mass, threshold=2.0)
    sel = R.select_edges(t, sc)
    # split tets: each tet has at most one split edge; tets by edge
    edge_tets = {}
    per_tet = np.zeros(len(t), int)
    for ti, tet in enumerate(t):
        for a, b, _, _ in R.tet_edges(tet):
            k = R.edge_key(a, b)
            edge_tets.setdefault(k, []).append(ti)
            if k in sel:
                per_tet[ti] += 1
    assert per_tet.max() == 1
    split = per_tet > 0
    prio = lambda k: (sc[k], k)
    for k, s in sc.items():
        if s <= R.SCORE_ONE or k in sel:
            continue
        # an unselected candidate has a tet split by a higher-priority edge
        blockers = [k2 for ti in edge_tets[k] if split[ti] for k2 in sel if ti in edge_tets[k2]]
        assert blockers and max(prio(b) for b in blockers) > prio(k)


def test_penetrating_midpoint_and_pinned_edges():
    p, t = _grid()
    inv_mass = np.ones(len(p), np.float32)
    inv_mass[p[:, 0] < 1e-6] = 0
    zero = np.zeros(len(p), np.float32)
    sc = R.legacy_scores(t, p, inv_mass, np.zeros(len(t), np.float32), zero, 0.0, 0.1, 1.0e6, 10.1,
                         midpoint_distance=lambda m: float(m[2]) - 0.05)  # plane at z = 0.05
    pen = R.contribution_int(np.float32(1.0e6), 10.1)
    for tet in t:
        for a, b, _, _ in R.tet_edges(tet):
            k = R.edge_key(a, b)
            if inv_mass[a] == 0 and inv_mass[b] == 0:
                assert k not in sc
            elif 0.5 * (p[a][2] + p[b][2]) < 0.05:
                assert sc[k] == pen
            else:
                assert sc[k] == 0


def test_capacity_guard_drops_the_pass():
    p, t = _grid()
    tris = M.surface_tris(t)
    sc = _scores(p, t, np.random.default_rng(0), np.ones(len(p), np.float32))
    sel = R.select_edges(t, sc)
    out = R.refine_pass(t, tris, p, p, p, np.zeros((len(t), 6)), np.zeros((len(t), 6)), np.ones((len(t), 3)), sel,
                        len(p), len(t) + 1, 10 * len(p))
    assert out["dropped"] and np.array_equal(out["tets"], t) and out["n_verts"] == len(p)


def test_planted_mistake_breaks_conformity(monkeypatch):
    """Selecting every candidate without the one-split-per-tet check leaves
    tets with two split edges; splitting them by one of them only gives
    hanging vertices that the conformity check catches."""
    p, t = _grid()
    tris = M.surface_tris(t)
    sc = _scores(p, t, np.random.default_rng(1), np.ones(len(p), np.float32))
    bad = {k for k, s in sc.items() if s > R.SCORE_ONE}
    out = R.refine_pass(t, tris, p, p, p, np.zeros((len(t), 6)), np.zeros((len(t), 6)), np.ones((len(t), 3)), bad,
                        len(p), 10 * len(t), 10 * len(p))
    with pytest.raises(AssertionError):
        _check_mesh(out["rest"][:out["n_verts"]], out["tets"], out["tris"], out["n_verts"])
