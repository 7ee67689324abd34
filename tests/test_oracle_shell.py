"""Pins of the oracle's membrane shell (oracle/mfem_oracle.c: orc_tri_rest_frame,
orc_membrane, the shell terms of orc_update_system / orc_assemble / orc_merit),
restating src/mfem/refinement/shell_energy.py of the reference.

Each check comes from a route the oracle does not take:
  * rest frame: at the rest configuration, and at any rigid motion of it,
    the energy and its gradient vanish (a transposed or mis-inverted W fails);
    the area against Heron's formula and a 3-4-5 triangle;
  * energy: mu A |F - polar(F)|^2 / 2 with F built from a QR frame of the
    rest edges and the polar factor from numpy's SVD, and the closed form
    mu A (k - 1)^2 of a uniform in-plane scaling by k;
  * gradient: central differences of the energy; the reference's zero stress
    for a collapsed (rank-deficient) triangle;
  * assembled system: H(shell) - H(no shell) against the autograd Hessian of
    sum mu A |F|^2 / 2 (the Gauss-Newton part: R held fixed), g and the merit
    against the autograd gradient and the sum of the numpy energies.
"""
import numpy as np
import pytest
import torch

from mfem_amd import scene
from oracle import Kernels, Oracle


def _qr_F(x, rest):
    """3x2 deformation gradient of triangle x (3, 3) w.r.t. rest (3, 3) in
    the orthonormal frame of numpy's QR of the rest edges."""
    E = (rest[1:] - rest[0]).T
    _, R = np.linalg.qr(E)
    return (x[1:] - x[0]).T @ np.linalg.inv(R)


def _energy_np(x, rest, mu):
    F = _qr_F(x, rest)
    U, _, Vt = np.linalg.svd(F, full_matrices=False)
    area = 0.5 * np.linalg.norm(np.cross(rest[1] - rest[0], rest[2] - rest[0]))
    return mu * area * 0.5 * np.sum((F - U @ Vt) ** 2)


def _rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@pytest.fixture(scope="module")
def K():
    return Kernels(64)


def test_rest_frame_area_and_zero_energy_at_rest(K):
    rng = np.random.default_rng(5)
    W, area = K.tri_rest_frame([0, 0, 0], [3, 0, 0], [0, 4, 0])
    assert area == pytest.approx(6.0, rel=1e-14)
    for _ in range(50):
        r = rng.normal(size=(3, 3))
        a, b, c = (np.linalg.norm(r[i] - r[j]) for i, j in ((0, 1), (1, 2), (2, 0)))
        s = 0.5 * (a + b + c)
        heron = np.sqrt(s * (s - a) * (s - b) * (s - c))
        W, area = K.tri_rest_frame(*r)
        assert area == pytest.approx(heron, rel=1e-10)
        # rest configuration and rigid motions of it: no energy, no force
        for x in (r, r @ _rot(rng).T + rng.normal(size=3)):
            E, g = K.membrane(x[0], x[1], x[2], W, area, 7.0)
            assert abs(E) <= 1e-12 * 7.0 * area
            assert np.abs(g).max() <= 1e-9 * 7.0 * area / np.abs(r).max()


def test_degenerate_rest_triangle_contributes_nothing(K):
    W, area = K.tri_rest_frame([0, 0, 0], [1, 1, 1], [2, 2, 2])
    assert area == 0.0 and np.all(W == 0.0)
    E, g = K.membrane([0, 0, 0], [1, 0, 0], [0, 1, 0], W, area, 3.0)
    assert E == 0.0 and np.all(g == 0.0)


def test_energy_matches_polar_decomposition_and_uniform_scaling(K):
    rng = np.random.default_rng(11)
    mu = 2.5
    for _ in range(100):
        rest = rng.normal(size=(3, 3))
        x = rest @ (np.eye(3) + 0.3 * rng.normal(size=(3, 3))).T + rng.normal(size=3)
        W, area = K.tri_rest_frame(*rest)
        E, _ = K.membrane(x[0], x[1], x[2], W, area, mu)
        assert E == pytest.approx(_energy_np(x, rest, mu), rel=1e-9, abs=1e-14)
    rest = rng.normal(size=(3, 3))
    W, area = K.tri_rest_frame(*rest)
    for k in (0.5, 0.9, 1.1, 2.0):
        x = k * rest @ _rot(rng).T
        E, _ = K.membrane(x[0], x[1], x[2], W, area, mu)
        assert E == pytest.approx(mu * area * (k - 1.0) ** 2, rel=1e-10)


def test_gradient_matches_central_differences(K):
    rng = np.random.default_rng(12)
    mu, h = 1.7, 1e-6
    for _ in range(30):
        rest = rng.normal(size=(3, 3))
        x = rest @ (np.eye(3) + 0.4 * rng.normal(size=(3, 3))).T
        W, area = K.tri_rest_frame(*rest)
        _, g = K.membrane(x[0], x[1], x[2], W, area, mu)
        fd = np.zeros((3, 3))
        for a in range(3):
            for k in range(3):
                xp, xm = x.copy(), x.copy()
                xp[a, k] += h
                xm[a, k] -= h
                fd[a, k] = (K.membrane(*xp, W, area, mu)[0] - K.membrane(*xm, W, area, mu)[0]) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-7 * mu * area)
        # translation invariance: the corner forces sum to zero
        np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-12 * np.abs(g).max())


def test_collapsed_triangle_has_zero_stress(K):
    # rank-deficient F: the reference returns zero stress (shell_energy.py:112-115)
    rest = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    W, area = K.tri_rest_frame(*rest)
    x = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0]])
    E, g = K.membrane(x[0], x[1], x[2], W, area, 1.0)
    assert E > 0.0 and np.all(g == 0.0)


def _torch_energy(X, tris, rinv, area, mu, gauss_newton):
    D = torch.stack([X[tris[:, 1]] - X[tris[:, 0]], X[tris[:, 2]] - X[tris[:, 0]]], dim=2)  # (F, 3, 2)
    F = D @ rinv
    if gauss_newton:
        return (mu * area * 0.5 * (F ** 2).sum(dim=(1, 2))).sum()
    U, _, Vh = torch.linalg.svd(F, full_matrices=False)
    return (mu * area * 0.5 * ((F - U @ Vh) ** 2).sum(dim=(1, 2))).sum()


def test_assembled_shell_terms_match_autograd():
    m = scene.drop_box_scene(cells=(3, 2, 2))
    mu_s = 40.0
    dt = 0.01
    rng = np.random.default_rng(3)
    x = m.particle_q + 0.01 * rng.normal(size=m.particle_q.shape)
    x[:, 2] = np.maximum(x[:, 2], 0.02)  # well above the barrier range
    res = {}
    for on in (False, True):
        o = Oracle(m, 64, params={"shell_mu": mu_s} if on else None)
        o.buf("x")[:] = x.ravel()
        o.begin_step(dt)
        o.update_system(dt)
        o.assemble(dt)
        res[on] = (o.csr().toarray(), o.buf("g").copy(), o.merit(x, o.buf("s"), dt))
    tris = np.asarray(m.tri_indices)
    rest = np.asarray(m.particle_q, np.float64)
    rinv = np.stack([np.linalg.inv(np.linalg.qr((rest[t[1:]] - rest[t[0]]).T)[1]) for t in tris])
    area = 0.5 * np.linalg.norm(np.cross(rest[tris[:, 1]] - rest[tris[:, 0]], rest[tris[:, 2]] - rest[tris[:, 0]]),
                                axis=1)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)
    X = T(x).requires_grad_(True)
    E = _torch_energy(X, T(tris).long(), T(rinv), T(area), mu_s, False)
    (grad,) = torch.autograd.grad(E, X)
    Hgn = torch.autograd.functional.hessian(
        lambda y: _torch_energy(y.reshape(-1, 3), T(tris).long(), T(rinv), T(area), mu_s, True), T(x).reshape(-1))
    dH = res[True][0] - res[False][0]
    np.testing.assert_allclose(dH, Hgn.numpy(), atol=1e-10 * np.abs(Hgn.numpy()).max())
    np.testing.assert_allclose(res[False][1] - res[True][1], grad.numpy().ravel(), rtol=1e-8,
                               atol=1e-10 * np.abs(grad.numpy()).max())
    assert res[True][2] - res[False][2] == pytest.approx(float(E.detach()), rel=1e-9)
    assert float(E.detach()) > 0.0
