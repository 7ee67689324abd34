"""NumPy restatement of the learned CG warm start (test infrastructure only).

The reference has no learned component (its CG starts from x~ - x,
src/mfem/refinement/solver.py:530-531 and src/mfem/refinement/kernels.py
``compute_initial_guess``), so this module restates the network, its
gradient, the Adam update and the per-vertex features exactly as the HIP
kernels define them (warp-mfem_amd/csrc/mfem_mlp.hip and the warm-start
kernels of mfem_kernels.hip), including the BF16 rounding points, and is
pinned by tests/test_oracle_mlp.py against what mathematics fixes: central
finite differences of an unrounded FP64 copy of the same network, the
erf form of GELU, Adam's closed-form step under a constant gradient, zero
loss at an exact blend, the energy-norm objective's identity
e^T H e = 2 Phi(dx0) + g . dx* (Phi the quadratic model), and hand-computed
features on tiny systems.

Parameter layout (FP32, 6288 values): W1 [16][64], b1 [64], W2 [64][64],
b2 [64], W3 [64][16], b3 [16]; only the first 3 output columns are used.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 16
O_W1 = 0
O_B1 = O_W1 + N_IN * N_HID
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_B3 = O_W3 + N_HID * N_OUT
N_PARAMS = O_B3 + N_OUT

_K = 0.7978845608028654  # sqrt(2 / pi)


def bf16_bits(a) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(a) -> np.ndarray:
    """Round to BF16 and back to FP32."""
    return bf16_value(bf16_bits(a))


def gelu(x):
    """GELU, tanh form (Hendrycks & Gimpel 2016)."""
    return 0.5 * x * (1.0 + np.tanh(_K * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = _K * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _K * (1.0 + 3.0 * 0.044715 * x * x)


def unpack(w):
    w = np.asarray(w)
    return (w[O_W1:O_B1].reshape(N_IN, N_HID), w[O_B1:O_W2], w[O_W2:O_B2].reshape(N_HID, N_HID), w[O_B2:O_W3],
            w[O_W3:O_B3].reshape(N_HID, N_OUT), w[O_B3:N_PARAMS])


def forward(w, X, quantize: bool = True):
    """y (n, 3) and the intermediates.  quantize=True follows the kernels:
    BF16 inputs / weights / hidden activations, FP32 accumulation, FP32
    biases; quantize=False is the same network in FP64 without rounding."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    if quantize:
        f = np.float32
        X, W1, W2, W3 = bf16(X), bf16(W1), bf16(W2), bf16(W3)
        b1, b2, b3 = (np.asarray(b, f) for b in (b1, b2, b3))
        z1 = (X.astype(np.float64) @ W1.astype(np.float64)).astype(f) + b1
        h1 = bf16(gelu(z1))
        z2 = (h1.astype(np.float64) @ W2.astype(np.float64)).astype(f) + b2
        h2 = bf16(gelu(z2))
        y = (h2.astype(np.float64) @ W3.astype(np.float64)).astype(f) + b3
    else:
        X = np.asarray(X, np.float64)
        z1 = X @ W1 + b1
        h1 = gelu(z1)
        z2 = h1 @ W2 + b2
        h2 = gelu(z2)
        y = h2 @ W3 + b3
    return y[:, :3], dict(X=X, W1=W1, W2=W2, W3=W3, z1=z1, h1=h1, z2=z2, h2=h2)


def backward(w, X, dy3, quantize: bool = True):
    """Gradient with respect to the parameter vector for an output gradient
    dy3 = dL/dy (n, 3): the kernels' backward pass (BF16 dy, dz2, dz1 when
    quantize; the b3 gradient sums the unrounded dy)."""
    dy3 = np.asarray(dy3, np.float64)
    n = dy3.shape[0]
    _, c = forward(w, X, quantize)
    dy = np.zeros((n, N_OUT))
    dy[:, :3] = dy3
    q = bf16 if quantize else (lambda a: a)
    dyq = np.asarray(q(dy), np.float64)
    h1, h2 = np.asarray(c["h1"], np.float64), np.asarray(c["h2"], np.float64)
    gW3 = h2.T @ dyq
    gb3 = np.zeros(N_OUT)
    gb3[:3] = dy3.sum(axis=0)
    dz2 = np.asarray(q(dyq @ np.asarray(c["W3"], np.float64).T * gelu_grad(np.asarray(c["z2"], np.float64))),
                     np.float64)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(axis=0)
    dz1 = np.asarray(q(dz2 @ np.asarray(c["W2"], np.float64).T * gelu_grad(np.asarray(c["z1"], np.float64))),
                     np.float64)
    gW1 = np.asarray(c["X"], np.float64).T @ dz1
    gb1 = dz1.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def loss_and_grad(w, X, aux, quantize: bool = True):
    """Loss (1/n) sum_n |y0 c1 + y1 c2 + y2 c3 - t|^2 with aux[n] = (c1, c2,
    c3, t) and its gradient with respect to the parameter vector."""
    aux = np.asarray(aux, np.float64).reshape(-1, 12)
    n = aux.shape[0]
    y, _ = forward(w, X, quantize)
    C = aux[:, :9].reshape(n, 3, 3)  # C[n, k] = candidate k
    t = aux[:, 9:12]
    e = np.einsum("nk,nkd->nd", y.astype(np.float64), C) - t
    loss = float(np.sum(e * e) / n)
    dy3 = (2.0 / n) * np.einsum("nd,nkd->nk", e, C)
    return loss, backward(w, X, dy3, quantize)


def residual_loss_and_dy(y, cand, g, H):
    """Log of the squared initial CG residual of the corrected guess,
    L = log |g - H dx0|^2 with dx0 = c3 + y0 c1 + y1 c2 + y2 c3 (absolute
    candidates cand (V, 9)), and dL/dy (V, 3) = c_k . (-2 H r / |r|^2)."""
    V = g.shape[0]
    C = np.asarray(cand, np.float64).reshape(V, 3, 3)
    y = np.asarray(y, np.float64)
    dx0 = C[:, 2] + np.einsum("nk,nkd->nd", y, C)
    r = np.asarray(g, np.float64) - (H @ dx0.reshape(-1)).reshape(V, 3)
    rr = float(np.sum(r * r))
    gd = (-2.0 / rr) * (H @ r.reshape(-1)).reshape(V, 3)
    return float(np.log(rr)), np.einsum("nkd,nd->nk", C, gd)


def energy_loss_and_dy(y, cand, dx_star, H):
    """Log of the squared energy-norm error of the corrected guess,
    L = log (e^T H e) with e = dx0 - dx_star (dx_star: the system's
    solution), dx0 = c3 + y0 c1 + y1 c2 + y2 c3, and dL/dy (V, 3) =
    c_k . (2 H e / e^T H e).  CG's error bound is in this norm."""
    V = dx_star.shape[0]
    C = np.asarray(cand, np.float64).reshape(V, 3, 3)
    y = np.asarray(y, np.float64)
    dx0 = C[:, 2] + np.einsum("nk,nkd->nd", y, C)
    e = dx0 - np.asarray(dx_star, np.float64)
    He = (H @ e.reshape(-1)).reshape(V, 3)
    eHe = float(np.sum(e * He))
    return float(np.log(eHe)), np.einsum("nkd,nd->nk", C, (2.0 / eHe) * He)


def adam(w, m, v, g, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns (w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def ws_features(H, g, x_tilde, x, dx_prev, Hv, inv_mass, *, precond=True, sing_thr=1e-20, attach_k=10.0, dt=0.01,
                newton_iter=0, iterations=1, alpha_prev=0.0):
    """Per-vertex features and normalised candidates of the learned warm start.

    H: scipy BSR (3V x 3V) system matrix; g: (V, 3) right-hand side.
    Candidates: c1 = x~ - x, c2 = D^-1 (g - H c1) (one block-Jacobi step),
    c3 = dx_prev; s = |c1| + |c2| + |c3|.  Returns (X (V, 16) FP32 before
    BF16 rounding, aux (V, 10) = c1/s, c2/s, c3/s, s).  alpha_prev is the
    line-search step of the previous Newton iteration (feature 15, zero at
    the first iteration).
    """
    import scipy.sparse as sp

    V = g.shape[0]
    Hb = sp.bsr_matrix(H, blocksize=(3, 3))
    c1 = np.asarray(x_tilde, np.float64) - np.asarray(x, np.float64)
    r1 = np.asarray(g, np.float64) - (Hb @ c1.reshape(-1)).reshape(V, 3)
    D = np.zeros((V, 3, 3))
    ring = np.zeros(V)
    valence = np.zeros(V)
    for i in range(V):
        for p in range(Hb.indptr[i], Hb.indptr[i + 1]):
            j = Hb.indices[p]
            blk = Hb.data[p]
            if j == i:
                D[i] = blk
            else:
                fro = np.sqrt(np.sum(blk * blk))
                if fro > 0:
                    ring[i] += fro
                    valence[i] += 1
    Minv = np.tile(np.eye(3), (V, 1, 1))
    if precond:
        det = np.linalg.det(D)
        ok = np.abs(det) >= sing_thr
        Minv[ok] = np.linalg.inv(D[ok])
    c2 = np.einsum("nab,nb->na", Minv, r1)
    c3 = np.asarray(dx_prev, np.float64)
    n1, n2, n3 = (np.linalg.norm(c, axis=1) for c in (c1, c2, c3))
    s = n1 + n2 + n3 + 1e-30

    def cos(u, v):
        d = np.linalg.norm(u, axis=1) * np.linalg.norm(v, axis=1)
        return np.where(d > 0, np.sum(u * v, axis=1) / np.where(d > 0, d, 1.0), 0.0)

    dfro = np.sqrt(np.sum(D * D, axis=(1, 2))) + 1e-30
    hfro = np.sqrt(np.sum(np.asarray(Hv, np.float64).reshape(V, 9) ** 2, axis=1))
    im = np.asarray(inv_mass, np.float64)
    mh = np.where(im == 0, 1.0 + attach_k, 1.0 / np.where(im == 0, 1.0, im)) / (dt * dt)
    ng = np.linalg.norm(g, axis=1)
    F = np.zeros((V, 16))
    F[:, 0], F[:, 1], F[:, 2] = n1 / s, n2 / s, n3 / s
    F[:, 3], F[:, 4], F[:, 5] = cos(c1, c2), cos(c1, c3), cos(c2, c3)
    F[:, 6] = hfro / dfro
    F[:, 7] = np.minimum(mh / (dfro / np.sqrt(3.0)), 10.0)
    F[:, 8] = (im == 0).astype(np.float64)
    F[:, 9] = newton_iter / max(iterations, 1)
    F[:, 10] = ring / dfro
    F[:, 11] = np.minimum(ng / (dfro * s), 10.0)
    F[:, 12] = np.minimum(np.linalg.norm(r1, axis=1) / (ng + 1e-30), 10.0)
    F[:, 13] = valence / 16.0
    F[:, 14] = 1.0
    F[:, 15] = alpha_prev if newton_iter > 0 else 0.0
    aux = np.concatenate([c1 / s[:, None], c2 / s[:, None], c3 / s[:, None], s[:, None]], axis=1)
    return F, aux


def ws_apply(y, aux):
    """guess = c3 + s (y0 c1 + y1 c2 + y2 c3) from normalised candidates:
    the previous Newton step plus the network's correction."""
    aux = np.asarray(aux, np.float64)
    y = np.asarray(y, np.float64)
    return aux[:, 9:10] * ((1.0 + y[:, 2:3]) * aux[:, 6:9] + y[:, 0:1] * aux[:, 0:3] + y[:, 1:2] * aux[:, 3:6])
