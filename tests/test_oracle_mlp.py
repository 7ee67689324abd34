"""Pins of the learned warm-start oracle (oracle/mlp.py) on the CPU.

Each check is fixed by mathematics or by an independent library routine,
chosen so that a dropped term, a wrong constant or a transposed operand in
the oracle fails it: finite differences of the unrounded network, the erf
form of GELU, torch's own BF16 rounding, Adam's closed-form step under a
constant gradient, zero loss at an exact blend, hand-computed features.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mlp


def _params(rng, scale=1.0):
    w = np.zeros(mlp.N_PARAMS)
    w[mlp.O_W1:mlp.O_B1] = rng.normal(0, scale / 4.0, mlp.N_IN * mlp.N_HID)
    w[mlp.O_B1:mlp.O_W2] = rng.normal(0, 0.1, mlp.N_HID)
    w[mlp.O_W2:mlp.O_B2] = rng.normal(0, scale / 8.0, mlp.N_HID * mlp.N_HID)
    w[mlp.O_B2:mlp.O_W3] = rng.normal(0, 0.1, mlp.N_HID)
    w[mlp.O_W3:mlp.O_B3] = rng.normal(0, scale / 8.0, mlp.N_HID * mlp.N_OUT)
    w[mlp.O_B3:] = rng.normal(0, 0.1, mlp.N_OUT)
    return w


def test_layout_size():
    assert mlp.N_PARAMS == 6288
    from mfem_amd import _ffi
    assert _ffi.MFEM_MLP_NPARAMS == mlp.N_PARAMS


def test_gelu_matches_erf_form():
    x = np.linspace(-6, 6, 2001)
    exact = 0.5 * x * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in x]))
    assert np.max(np.abs(mlp.gelu(x) - exact)) < 1e-3
    # derivative against central differences of the tanh form itself
    h = 1e-6
    fd = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(mlp.gelu_grad(x) - fd)) < 1e-7


def test_bf16_rounding_known_answers_and_torch():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    cases = {
        1.0: 0x3F80,
        float(one + ulp / 2): 0x3F80,          # tie -> even (down)
        float(one + 3 * ulp / 2): 0x3F82,      # tie -> even (up)
        float(one + ulp / 2 + 2.0 ** -20): 0x3F81,
        -2.0: 0xC000,
        0.0: 0x0000,
    }
    for v, bits in cases.items():
        assert int(mlp.bf16_bits(np.array([v], np.float32))[0]) == bits, v
    import torch
    a = np.random.default_rng(0).normal(0, 10, 4096).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(mlp.bf16_bits(a), ref)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = _params(rng)
    n = 9
    X = rng.normal(0, 1, (n, 16))
    aux = rng.normal(0, 1, (n, 12))
    _, g = mlp.loss_and_grad(w, X, aux, quantize=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False) for a, b in
                          [(mlp.O_W1, mlp.O_B1), (mlp.O_B1, mlp.O_W2), (mlp.O_W2, mlp.O_B2), (mlp.O_B2, mlp.O_W3),
                           (mlp.O_W3, mlp.O_W3 + 64 * 16)]] + [np.arange(mlp.O_B3, mlp.O_B3 + 3)])
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_and_grad(wp, X, aux, False)[0] - mlp.loss_and_grad(wm, X, aux, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # unused output columns get no gradient
    assert np.all(g[mlp.O_B3 + 3:] == 0)
    W3g = g[mlp.O_W3:mlp.O_B3].reshape(64, 16)
    assert np.all(W3g[:, 3:] == 0)


def test_quantized_pass_tracks_unrounded_network():
    rng = np.random.default_rng(2)
    w = _params(rng)
    X = rng.normal(0, 1, (256, 16))
    aux = rng.normal(0, 1, (256, 12))
    yq, _ = mlp.forward(w, X, True)
    y, _ = mlp.forward(w, X, False)
    assert np.max(np.abs(yq - y)) < 2e-2 * np.max(np.abs(y))
    lq, gq = mlp.loss_and_grad(w, X, aux, True)
    l, g = mlp.loss_and_grad(w, X, aux, False)
    assert abs(lq - l) < 2e-2 * l
    assert np.linalg.norm(gq - g) < 3e-2 * np.linalg.norm(g)


def test_zero_loss_at_exact_blend():
    rng = np.random.default_rng(3)
    w = _params(rng)
    w[mlp.O_W3:mlp.O_B3] = 0.0
    coef = np.array([0.75, -0.5, 0.25])
    w[mlp.O_B3:mlp.O_B3 + 3] = coef
    n = 32
    X = rng.normal(0, 1, (n, 16))
    C = rng.normal(0, 1, (n, 3, 3))
    t = np.einsum("k,nkd->nd", coef, C)
    aux = np.concatenate([C.reshape(n, 9), t], axis=1)
    for q in (False, True):
        loss, g = mlp.loss_and_grad(w, X, aux, q)
        assert loss < 1e-20
        assert np.max(np.abs(g)) < 1e-12


def test_adam_closed_form_under_constant_gradient():
    rng = np.random.default_rng(4)
    w0 = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    lr = 1e-3
    for t in range(1, 6):
        w_new, m, v = mlp.adam(w, m, v, g, lr, t)
        # bias-corrected m/sqrt(v) = g/|g| for a constant gradient
        np.testing.assert_allclose(w_new - w, -lr * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
        w = w_new


def test_features_on_a_two_vertex_system():
    # H = [[A, B], [B^T, C]] with hand-picked blocks
    A = np.diag([2.0, 4.0, 8.0])
    C = np.diag([1.0, 1.0, 2.0])
    B = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.5, 0.0, 0.0]])
    H = sp.bsr_matrix(np.block([[A, B], [B.T, C]]), blocksize=(3, 3))
    g = np.array([[1.0, 2.0, 3.0], [0.5, -1.0, 0.0]])
    xt = np.array([[0.1, 0.0, 0.0], [0.0, 0.2, 0.0]])
    x = np.zeros((2, 3))
    dxp = np.array([[0.0, 0.0, 0.3], [0.4, 0.0, 0.0]])
    Hv = np.zeros((2, 9))
    inv_mass = np.array([1.0, 0.0])
    F, aux = mlp.ws_features(H, g, xt, x, dxp, Hv, inv_mass, dt=0.1, attach_k=10.0, newton_iter=1, iterations=4,
                             alpha_prev=0.25)
    c1 = xt - x
    r = g - (np.block([[A, B], [B.T, C]]) @ c1.reshape(-1)).reshape(2, 3)
    c2 = np.stack([np.linalg.solve(A, r[0]), np.linalg.solve(C, r[1])])
    s = np.linalg.norm(c1, axis=1) + np.linalg.norm(c2, axis=1) + np.linalg.norm(dxp, axis=1)
    np.testing.assert_allclose(aux[:, 0:3] * s[:, None], c1, rtol=1e-12)
    np.testing.assert_allclose(aux[:, 3:6] * s[:, None], c2, rtol=1e-12)
    np.testing.assert_allclose(aux[:, 6:9] * s[:, None], dxp, rtol=1e-12)
    np.testing.assert_allclose(aux[:, 9], s, rtol=1e-12)
    np.testing.assert_allclose(F[:, 0] + F[:, 1] + F[:, 2], 1.0, rtol=1e-12)
    assert F[0, 8] == 0.0 and F[1, 8] == 1.0
    np.testing.assert_allclose(F[:, 9], 0.25)
    np.testing.assert_allclose(F[0, 10], np.linalg.norm(B) / np.linalg.norm(A))
    np.testing.assert_allclose(F[:, 13], 1.0 / 16.0)
    np.testing.assert_allclose(F[:, 15], 0.25)
    # mass term: m / dt^2 relative to |D|_F / sqrt(3); pinned vertex uses 1 + k
    np.testing.assert_allclose(F[0, 7], min(1.0 / 0.01 / (np.linalg.norm(A) / np.sqrt(3)), 10.0))
    np.testing.assert_allclose(F[1, 7], min(11.0 / 0.01 / (np.linalg.norm(C) / np.sqrt(3)), 10.0))
    # y = 0 gives the previous Newton step; y = (1, 0, -1) the plain guess x~ - x
    np.testing.assert_allclose(mlp.ws_apply(np.zeros((2, 3)), aux), dxp, rtol=1e-12)
    np.testing.assert_allclose(mlp.ws_apply(np.tile([1.0, 0.0, -1.0], (2, 1)), aux), c1, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(mlp.ws_apply(np.tile([0.0, 1.0, -1.0], (2, 1)), aux), c2, rtol=1e-12, atol=1e-15)


def _fd_mismatch(w, X, aux, idx, h=1e-6):
    _, g = mlp.loss_and_grad(w, X, aux, quantize=False)
    worst = 0.0
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_and_grad(wp, X, aux, False)[0] - mlp.loss_and_grad(wm, X, aux, False)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1.0, abs(fd)))
    return worst


def test_planted_mistakes_fail_the_gradient_pin(monkeypatch):
    """A dropped term in GELU' or a transposed W2 in the backward pass makes
    the finite-difference pin fail (mistakes planted in the oracle only)."""
    rng = np.random.default_rng(5)
    w = _params(rng)
    X = rng.normal(0, 1, (5, 16))
    aux = rng.normal(0, 1, (5, 12))
    idx = [mlp.O_W1 + 7, mlp.O_W1 + 300, mlp.O_B1 + 5, mlp.O_W2 + 1000]
    assert _fd_mismatch(w, X, aux, idx) < 1e-6
    k = 0.7978845608028654
    with monkeypatch.context() as mp:
        mp.setattr(mlp, "gelu_grad", lambda x: 0.5 * (1.0 + np.tanh(k * (x + 0.044715 * x ** 3))))
        assert _fd_mismatch(w, X, aux, idx) > 1e-4
    with monkeypatch.context() as mp:
        # the forward pass keeps the true W2, the backward pass sees its transpose
        orig_forward = mlp.forward

        def forward_true(wv, Xv, quantize=True):
            y, c = orig_forward(wv, Xv, quantize)
            c["W2"] = np.asarray(c["W2"]).T
            return y, c

        mp.setattr(mlp, "forward", forward_true)
        assert _fd_mismatch(w, X, aux, idx) > 1e-4


def test_residual_objective_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    V = 12
    A = rng.normal(0, 1, (3 * V, 3 * V))
    H = A @ A.T + 3 * V * np.eye(3 * V)  # SPD like the system matrix
    cand = rng.normal(0, 1, (V, 9))
    g = rng.normal(0, 1, (V, 3))
    y = rng.normal(0, 0.3, (V, 3))
    L, dy = mlp.residual_loss_and_dy(y, cand, g, H)
    h = 1e-6
    for n, k in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        yp, ym = y.copy(), y.copy()
        yp[n, k] += h
        ym[n, k] -= h
        fd = (mlp.residual_loss_and_dy(yp, cand, g, H)[0] - mlp.residual_loss_and_dy(ym, cand, g, H)[0]) / (2 * h)
        assert abs(fd - dy[n, k]) <= 1e-6 * max(1.0, abs(fd)), (n, k, fd, dy[n, k])
    # y = 0 is the previous Newton step c3; the exact solution gives r = 0
    dx = np.linalg.solve(H, g.reshape(-1)).reshape(V, 3)
    c = cand.reshape(V, 3, 3).copy()
    c[:, 2] = dx
    L0, _ = mlp.residual_loss_and_dy(np.zeros((V, 3)), c.reshape(V, 9), g, H)
    assert L0 < np.log(1e-20)


def test_energy_objective_gradient_matches_finite_differences():
    rng = np.random.default_rng(8)
    V = 12
    A = rng.normal(0, 1, (3 * V, 3 * V))
    H = A @ A.T + 3 * V * np.eye(3 * V)
    cand = rng.normal(0, 1, (V, 9))
    g = rng.normal(0, 1, (V, 3))
    dx = np.linalg.solve(H, g.reshape(-1)).reshape(V, 3)
    y = rng.normal(0, 0.3, (V, 3))
    L, dy = mlp.energy_loss_and_dy(y, cand, dx, H)
    h = 1e-6
    for n, k in [(0, 0), (4, 1), (9, 2), (11, 1)]:
        yp, ym = y.copy(), y.copy()
        yp[n, k] += h
        ym[n, k] -= h
        fd = (mlp.energy_loss_and_dy(yp, cand, dx, H)[0] - mlp.energy_loss_and_dy(ym, cand, dx, H)[0]) / (2 * h)
        assert abs(fd - dy[n, k]) <= 1e-6 * max(1.0, abs(fd)), (n, k, fd, dy[n, k])
    # e^T H e = 2 Phi(dx0) + g . dx_star, Phi the quadratic model energy
    C = cand.reshape(V, 3, 3)
    dx0 = (C[:, 2] + np.einsum("nk,nkd->nd", y, C)).reshape(-1)
    phi = 0.5 * dx0 @ H @ dx0 - g.reshape(-1) @ dx0
    assert np.exp(L) == pytest.approx(2 * phi + g.reshape(-1) @ dx.reshape(-1), rel=1e-9)


def test_backward_from_output_gradient_matches_chain_rule():
    """backward(w, X, dL/dy) equals the gradient of L(y(w)) for a linear L."""
    rng = np.random.default_rng(7)
    w = _params(rng)
    X = rng.normal(0, 1, (6, 16))
    a = rng.normal(0, 1, (6, 3))  # L = sum a * y
    g = mlp.backward(w, X, a, quantize=False)
    h = 1e-6
    for i in [mlp.O_W1 + 5, mlp.O_B1 + 2, mlp.O_W2 + 77, mlp.O_W3 + 64 * 2 + 1, mlp.O_B3 + 1]:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (np.sum(a * mlp.forward(wp, X, False)[0]) - np.sum(a * mlp.forward(wm, X, False)[0])) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
