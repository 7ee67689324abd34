"""Pins of the self-contact oracle (oracle/collide.py) on the CPU.

Known answers are the reference's own (tests/test_collision.py: segment
closest points, point-triangle / triangle-point / edge-edge distances with a
d_hat sentinel, the 15-check soft-rigid kernel; tests/test_scan.py:
exclusive scan = cumsum - a).  Closed forms pin the barrier derivatives, and
the brute-force broad phase is pinned by conservativeness (every pair closer
than d_hat is a candidate) on a random scene.
"""
import numpy as np
import pytest

from oracle import collide as C

D_HAT = 2.0
TRI = (np.array([-5.0, -5.0, 0.0]), np.array([5.0, -5.0, 0.0]), np.array([0.0, 5.0, 0.0]))


def test_segment_segment_known_answers():
    c1, c2, s = C.closest_point_segment_segment((0, 0, 0), (1, 0, 0), (0.5, 1, 0.5), (0.5, 1, -0.5))
    np.testing.assert_allclose(c1, [0.5, 0, 0], atol=1e-6)
    np.testing.assert_allclose(c2, [0.5, 1, 0], atol=1e-6)
    assert s == pytest.approx(0.5, abs=1e-6)
    c1, c2, _ = C.closest_point_segment_segment((0, 0, 0), (1, 0, 0), (0.25, 1, 0), (0.75, 1, 0))
    assert np.linalg.norm(c1 - c2) == pytest.approx(1.0, abs=1e-6)
    c1, c2, s = C.closest_point_segment_segment((0, 0, 0), (1, 0, 0), (3, 1, 0), (4, 1, 0))
    np.testing.assert_allclose(c1, [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(c2, [3, 1, 0], atol=1e-6)
    assert s == pytest.approx(1.0, abs=1e-6)


def test_point_triangle_known_answers():
    assert C.point_triangle_dist([0, 0, 1], *TRI) == pytest.approx(1.0, abs=1e-6)
    # beyond d_hat: the sentinel stays; two writes keep the smaller
    d = np.full(1, D_HAT)
    for x in ([0, 0, D_HAT + 1.0], [0, 0, 0.3], [0, 0, 1.0]):
        dd = C.point_triangle_dist(x, *TRI)
        if dd < D_HAT:
            d[0] = min(d[0], dd)
    assert d[0] == pytest.approx(0.3, abs=1e-6)
    # edge-edge: both endpoints get the shared distance
    assert C.edge_edge_dist((0, 1, 0.5), (0, 1, -0.5), (-1, 0, 0), (1, 0, 0)) == pytest.approx(1.0, abs=1e-6)
    # closest point in each Voronoi region equals a brute-force dense sample
    rng = np.random.default_rng(0)
    a, b, c = rng.normal(size=(3, 3))
    uv = rng.uniform(size=(20000, 2))
    uv = np.where(uv.sum(1, keepdims=True) > 1, 1 - uv, uv)
    pts = a + uv[:, :1] * (b - a) + uv[:, 1:] * (c - a)
    for _ in range(20):
        x = rng.normal(size=3) * 2
        d_dense = np.min(np.linalg.norm(pts - x, axis=1))
        d = C.point_triangle_dist(x, a, b, c)
        assert d <= d_dense + 1e-12 and d >= d_dense - 0.05


def _soft_rigid(rigid_tri, soft_tri, rigid_tri_count=1, d_hat=D_HAT, count=1):
    pts = np.vstack([rigid_tri, soft_tri])
    tris = np.array([[0, 1, 2], [3, 4, 5]])
    return C.soft_rigid_distance(pts, tris, 3, rigid_tri_count, [[0, 1]], count, d_hat, np.full(3, d_hat))


def test_soft_rigid_distance_known_answers():
    rigid = np.array([[-10.0, -10.0, 0.0], [10.0, -10.0, 0.0], [0.0, 10.0, 0.0]])
    h = 0.3
    soft = np.array([[-0.2, 0.0, h], [0.2, 0.0, h], [0.0, 0.3, h]])
    np.testing.assert_allclose(_soft_rigid(rigid, soft, d_hat=0.5), [h, h, h], atol=1e-5)
    soft = np.array([[-0.2, 0.0, 0.1], [0.2, 0.0, 0.1], [0.0, 0.3, 0.1]])
    np.testing.assert_allclose(_soft_rigid(rigid, soft, rigid_tri_count=2), [D_HAT] * 3, atol=1e-9)
    np.testing.assert_allclose(_soft_rigid(rigid, soft, count=0), [D_HAT] * 3, atol=1e-9)
    soft = np.array([[0.0, 0.0, 0.1], [0.3, 0.0, 0.4], [0.0, 0.3, 0.4]])
    d = _soft_rigid(rigid, soft, d_hat=0.5)
    assert d[0] == pytest.approx(0.1, abs=1e-5)
    assert np.all(d <= 0.5)


def test_ipc_barrier_derivatives():
    dh = 1e-3
    d = np.linspace(1e-5, dh * 0.999, 50)
    e, de, d2e = C.ipc_barrier(d, dh)
    h = 1e-9
    ep, dep, _ = C.ipc_barrier(d + h, dh)
    em, dem, _ = C.ipc_barrier(d - h, dh)
    np.testing.assert_allclose(de, (ep - em) / (2 * h), rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(d2e, (dep - dem) / (2 * h), rtol=1e-4, atol=1e-9)
    assert np.all(e > 0) and np.all(de < 0) and np.all(d2e > 0)
    # C2 contact with zero at d_hat
    e1, de1, _ = C.ipc_barrier(np.array([dh * (1 - 1e-9)]), dh)
    assert abs(e1[0]) < 1e-20 and abs(de1[0]) < 1e-11


def test_exclusive_scan_known_answer():
    np.testing.assert_array_equal(C.exclusive_scan([3, 0, 2, 5, 1]), [0, 3, 3, 5, 10])
    a = np.random.default_rng(1).integers(0, 10, 1000).astype(np.int32)
    np.testing.assert_array_equal(C.exclusive_scan(a), np.cumsum(a) - a)


def _blob(rng, n=60):
    """Random triangle soup (well separated ids) in a unit box."""
    x = rng.uniform(0, 1, (3 * n, 3)).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    # shrink each triangle around its centroid so boxes are small
    cen = x.reshape(n, 3, 3).mean(axis=1, keepdims=True)
    x = (cen + 0.08 * (x.reshape(n, 3, 3) - cen)).reshape(-1, 3).astype(np.float32)
    return x, tris


def test_brute_force_candidates_are_conservative():
    rng = np.random.default_rng(2)
    x, tris = _blob(rng)
    d_hat = 0.05
    pts = C.surface_points(tris)
    pt = {tuple(p) for p in C.brute_force_pt(x, pts, tris, d_hat)}
    for p in pts:
        for t, tri in enumerate(tris):
            if p in tri:
                continue
            if C.point_triangle_dist(x[p], *x[tri]) < d_hat:
                assert (int(p), t) in pt
    edges = C.surface_edges(tris)
    ee = {tuple(e) for e in C.brute_force_ee(x, edges, d_hat)}
    for a in range(len(edges)):
        for b in range(a + 1, len(edges)):
            if set(edges[a]) & set(edges[b]):
                continue
            if C.edge_edge_dist(*x[edges[a]], *x[edges[b]]) < d_hat:
                assert (a, b) in ee
    assert len(pt) > 0 and len(ee) > 0
