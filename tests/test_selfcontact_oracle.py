"""Pins of the oracle's self-contact restatement (oracle/mfem_oracle.c,
orc_sc_*): the IPC barrier on point-triangle and edge-edge pairs of surface
primitives inside the Newton solve, its gradient and Gauss-Newton diagonal
blocks, and the additive-CCD cut of the line search's first step.

* distances: the reference's known answers (tests/test_collision.py:56-232:
  perpendicular skew, parallel and clamped segments, a point above a
  triangle, crossing edges);
* gradient = central differences of the energy along random directions, on
  the stacked-slab scene deformed into contact;
* Newton's third law: the pair forces sum to zero and so do their moments
  (a distance is invariant under rigid motions), which a wrong closest-point
  weight fails (planted below);
* the Gauss-Newton block of a point above a triangle's interior equals the
  exact Hessian block there (the distance is linear in the point), by finite
  differences of the gradient;
* CCD: a point driven through a face and an edge driven through an edge stop
  before the true time of impact, keep at least the CCD fraction of their
  distance, and do not stop far short of it; pairs moving apart are not cut;
* in the solver: steps of the stacked slabs converge every solve, keep the
  self-contact energy finite (every pair at a positive distance) and leave
  the upper slab resting on the lower one inside dhat.
"""
import ctypes

import numpy as np
import pytest

from mfem_amd import scene
from oracle import Oracle, _lib


def _dist(kind, X):
    lib = _lib(64)
    X = np.ascontiguousarray(np.asarray(X, np.float64).reshape(12))
    cw = np.zeros(4)
    n = np.zeros(3)
    d = lib.orc_sc_distance(kind, X.ctypes.data_as(ctypes.c_void_p), cw.ctypes.data_as(ctypes.c_void_p),
                            n.ctypes.data_as(ctypes.c_void_p))
    return d, cw, n


def _toi(kind, X, P, s=0.1):
    lib = _lib(64)
    X = np.ascontiguousarray(np.asarray(X, np.float64).reshape(12))
    P = np.ascontiguousarray(np.asarray(P, np.float64).reshape(12))
    return lib.orc_sc_accd(kind, X.ctypes.data_as(ctypes.c_void_p), P.ctypes.data_as(ctypes.c_void_p), s)


TRI = [(-5.0, -5.0, 0.0), (5.0, -5.0, 0.0), (0.0, 5.0, 0.0)]


def test_reference_known_distances():
    # segments (test_collision.py:56-96): perpendicular skew, parallel offset, clamped
    d, cw, _ = _dist(1, [(0, 0, 0), (1, 0, 0), (0.5, 1, 0.5), (0.5, 1, -0.5)])
    assert d == pytest.approx(1.0, abs=1e-12) and cw[1] == pytest.approx(0.5) and cw[3] == pytest.approx(-0.5)
    d, _, _ = _dist(1, [(0, 0, 0), (1, 0, 0), (0.25, 1, 0), (0.75, 1, 0)])
    assert d == pytest.approx(1.0, abs=1e-12)
    d, cw, _ = _dist(1, [(0, 0, 0), (1, 0, 0), (3, 1, 0), (4, 1, 0)])
    assert d == pytest.approx(np.hypot(2.0, 1.0), abs=1e-12) and cw[1] == pytest.approx(1.0)  # s = 1
    # point above the triangle (test_collision.py:127-131) and beyond its plane region
    d, cw, n = _dist(0, [(0.0, 0.0, 1.0)] + TRI)
    assert d == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(n, [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(cw[1:].sum(), -1.0, atol=1e-12)
    d, _, _ = _dist(0, [(0.0, -8.0, 0.0)] + TRI)  # nearest the edge AB at y = -5
    assert d == pytest.approx(3.0, abs=1e-12)
    # crossing edges one apart (test_collision.py:213-232)
    d, _, _ = _dist(1, [(0.0, 1.0, 0.5), (0.0, 1.0, -0.5), (-1.0, 0.0, 0.0), (1.0, 0.0, 0.0)])
    assert d == pytest.approx(1.0, abs=1e-12)


@pytest.fixture(scope="module")
def touching():
    """The stacked slabs with the upper one lowered into the barrier range
    and jittered (many active point-triangle and edge-edge pairs)."""
    m = scene.stacked_boxes_scene()
    o = Oracle(m, 64)
    nb = m.particle_count // 2
    x = m.particle_q.copy()
    rng = np.random.default_rng(0)
    x[nb:, 2] -= 2.0e-4
    x[nb:] += 1.0e-5 * rng.normal(size=(nb, 3))
    return m, o, x


def test_gradient_matches_finite_differences(touching):
    m, o, x = touching
    e, g, _ = o.sc_gradient(x)
    assert e > 0 and np.abs(g).max() > 0
    rng = np.random.default_rng(1)
    for _ in range(4):
        v = rng.normal(size=x.shape)
        v /= np.abs(v).max()
        h = 1e-8
        fd = (o.sc_energy(x + h * v) - o.sc_energy(x - h * v)) / (2 * h)
        assert (g * v).sum() == pytest.approx(fd, rel=2e-4, abs=1e-9 * np.abs(g).sum())


def test_forces_and_moments_balance(touching):
    m, o, x = touching
    _, g, _ = o.sc_gradient(x)
    scale = np.abs(g).sum()
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-12 * scale)
    np.testing.assert_allclose(np.cross(x, g).sum(axis=0), 0.0, atol=1e-12 * scale)


def test_planted_weight_mistake_breaks_the_balance():
    # one pair's gradient with the triangle's weights flipped in sign: its
    # net force is 2 kappa b' n, not zero
    X = np.array([(0.1, 0.2, 4e-4)] + TRI)
    d, cw, n = _dist(0, X)
    good = np.outer(cw, n).sum(axis=0)
    bad_cw = cw.copy()
    bad_cw[1:] *= -1
    bad = np.outer(bad_cw, n).sum(axis=0)
    assert np.allclose(good, 0.0, atol=1e-12) and not np.allclose(bad, 0.0, atol=1e-3)


def test_gauss_newton_block_of_a_face_point_is_exact():
    """A one-pair scene: a point 0.4 dhat above the interior of a triangle;
    the point's diagonal block equals the finite-difference Jacobian of its
    gradient (the distance to the face is linear in the point)."""
    # a tet (the triangle's) and a far tet holding the point, in one model
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, -1],
                  [0.2, 0.2, 4e-4], [0.2, 0.2, 1.0], [0.7, 0.2, 1.0], [0.2, 0.7, 1.0]], float)
    t = np.array([[0, 2, 1, 3], [4, 5, 6, 7]])
    m = scene.build_model(p, t, np.tile([1e4, 1e4, 0.0], (2, 1)), 1000.0)
    m.params.update(self_contact=True, self_contact_dhat=1e-3, self_contact_kappa=10.0)
    o = Oracle(m, 64)
    e, g, h = o.sc_gradient(p)
    assert e > 0
    H = np.array([[h[4, 0], h[4, 3], h[4, 4]], [h[4, 3], h[4, 1], h[4, 5]], [h[4, 4], h[4, 5], h[4, 2]]])
    J = np.zeros((3, 3))
    for c in range(3):
        dp = np.zeros_like(p)
        dp[4, c] = 1e-9
        J[:, c] = (o.sc_gradient(p + dp)[1][4] - o.sc_gradient(p - dp)[1][4]) / 2e-9
    np.testing.assert_allclose(H, J, rtol=1e-5, atol=1e-6 * np.abs(J).max())
    assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()


def test_ccd_stops_before_impact():
    s = 0.1
    # point 1e-3 above the face, moving 2e-3 down through it: impact at 0.5
    X = [(0.1, 0.2, 1e-3)] + TRI
    P = [(0.0, 0.0, -2e-3), (0, 0, 0), (0, 0, 0), (0, 0, 0)]
    t = _toi(0, X, P, s)
    assert 0.3 < t < 0.5
    Xt = np.array(X) + t * np.array(P)
    assert _dist(0, Xt)[0] >= s * 1e-3 * (1 - 1e-9)
    # edge crossing another edge (moving in z through it): impact at 0.25
    X = [(0, 1.0, 0.5), (0, 1.0, -0.5), (-1, 0, 0), (1, 0, 0)]
    P = [(0, -4.0, 0), (0, -4.0, 0), (0, 0, 0), (0, 0, 0)]
    t = _toi(1, X, P, s)
    assert 0.15 < t < 0.25
    # moving apart, and a pair that misses: no cut
    assert _toi(0, [(0.1, 0.2, 1e-3)] + TRI, [(0, 0, 1e-3), (0, 0, 0), (0, 0, 0), (0, 0, 0)], s) == 1.0
    assert _toi(0, [(20.0, 0.0, 1e-3)] + TRI, [(1.0, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0)], s) == 1.0


def test_solver_steps_keep_pairs_apart():
    m = scene.stacked_boxes_scene()
    o = Oracle(m, 64)
    nb = m.particle_count // 2
    e0 = o.sc_energy()
    for _ in range(4):
        o.step(0.01)
        st = o.stats()
        assert np.all(st[:, 1] <= 1e-6)
        assert np.isfinite(o.sc_energy()) and o.sc_energy() < 1e299
        # the upper slab sits on the lower one, inside dhat
        gap = o.x[nb:, 2].min() - o.x[:nb, 2].max()
        assert 0.0 < gap < 1e-3
    assert e0 > 0 and o.sc_energy() > 0
