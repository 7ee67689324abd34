"""The workgroup-local preconditioner's algebra, on the CPU.

z = D^-1 r followed by s damped block-Jacobi sweeps z <- z + w D^-1 (r - A_loc z)
on the rows of each workgroup (A_loc = A restricted to the workgroup's own
rows) is z = P r with P = p(w D^-1 A_loc) D^-1 for a polynomial p: symmetric,
and positive definite for every even s (p > 0 on the positive spectrum),
which is why the kernel accepts only even sweep counts.  Odd counts can give
an indefinite P (shown on the same matrix), and the even-count preconditioner
cuts PCG iterations against block Jacobi on an elasticity-like system.
"""
import numpy as np
import pytest


def _system(n_blocks=96, seed=0):
    # SPD 3x3-block matrix of a 1-D chain with 3 neighbours each side (a
    # stiff, mesh-like stencil), plus a small mass term
    rng = np.random.default_rng(seed)
    n = 3 * n_blocks
    A = np.zeros((n, n))
    for i in range(n_blocks):
        for d in (1, 2, 3):
            j = i + d
            if j >= n_blocks:
                continue
            R = rng.normal(0, 1, (3, 3))
            K = R @ R.T + 3 * np.eye(3)
            K *= 1.0 / d
            for a, b in ((i, i), (j, j)):
                A[3 * a:3 * a + 3, 3 * b:3 * b + 3] += K
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] -= K
            A[3 * j:3 * j + 3, 3 * i:3 * i + 3] -= K
    A += 1e-3 * np.eye(n)
    return A


def _precond(A, rows_per_group, sweeps, w):
    n = A.shape[0]
    D = np.zeros_like(A)
    for i in range(0, n, 3):
        D[i:i + 3, i:i + 3] = A[i:i + 3, i:i + 3]
    Dinv = np.linalg.inv(D)
    g = np.arange(n) // (3 * rows_per_group)
    A_loc = np.where(g[:, None] == g[None, :], A, 0.0)
    P = Dinv.copy()  # z = P r, column by column
    for _ in range(sweeps):
        P = P + w * Dinv @ (np.eye(n) - A_loc @ P)
    return P


def _pcg_iters(A, P, b, tol=1e-6):
    x = np.zeros_like(b)
    r = b.copy()
    z = P @ r
    p = z.copy()
    rz = r @ z
    for k in range(10 * len(b)):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return k + 1
        z = P @ r
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return 10 * len(b)


@pytest.mark.parametrize("sweeps", [2, 4, 6])
@pytest.mark.parametrize("w", [0.3, 0.5, 1.0])
def test_even_sweeps_are_spd(sweeps, w):
    A = _system()
    P = _precond(A, 16, sweeps, w)
    np.testing.assert_allclose(P, P.T, atol=1e-9 * np.abs(P).max())
    assert np.linalg.eigvalsh(0.5 * (P + P.T)).min() > 0


def test_odd_sweeps_can_be_indefinite():
    A = _system()
    P = _precond(A, 16, 1, 2.0)
    np.testing.assert_allclose(P, P.T, atol=1e-9 * np.abs(P).max())
    assert np.linalg.eigvalsh(0.5 * (P + P.T)).min() < 0


def test_local_sweeps_cut_pcg_iterations():
    A = _system()
    b = np.random.default_rng(1).normal(0, 1, A.shape[0])
    it_bj = _pcg_iters(A, _precond(A, 16, 0, 0.5), b)
    it_2 = _pcg_iters(A, _precond(A, 16, 2, 0.5), b)
    it_4 = _pcg_iters(A, _precond(A, 16, 4, 0.5), b)
    assert it_2 < it_bj and it_4 <= it_2
