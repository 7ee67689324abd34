"""CPU checks of the coarse-weight network's training machinery
(mfem_amd/coarse.py) on the beam's assembled FP64 system (from the oracle):

* the block products equal scipy's H x and the aggregate restriction H_aa x;
* the two-level preconditioner M(w)^-1 equals a dense numpy restatement,
  D^-1 + sum_a P'_a (P'_a^T H_aa P'_a)^-1 P'_a^T with P' = P - w D^-1 H_aa P,
  and is symmetric positive definite for any weights;
* the loss at one step equals the closed form of one optimally damped
  preconditioned Richardson step, and at k steps the energy-norm error of
  k preconditioned CG iterations computed with numpy;
* its gradient with respect to the weights (torch autograd) matches central
  finite differences;
* a planted sign mistake in the smoothing fails the dense comparison.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from mfem_amd import coarse, pattern, scene
from oracle import Oracle

AGG = coarse.AGG


@pytest.fixture(scope="module")
def beam_system():
    m = scene.beam_scene(cells=(10, 3, 3), energy="neohookean")
    o = Oracle(m, 64)
    dt = 1.0 / 60.0
    rng = np.random.default_rng(3)
    o.buf("x")[:] = (m.particle_q + 0.003 * rng.normal(size=m.particle_q.shape)).ravel()
    o.begin_step(dt)
    o.update_system(dt)
    o.assemble(dt)
    H = o.csr()
    n = m.particle_count
    pat = pattern.build_system_pattern(n, m.tet_indices)
    Hb = H.tobsr(blocksize=(3, 3))
    assert np.array_equal(Hb.indptr, pat.rowptr) and np.array_equal(Hb.indices, pat.cols)
    sell = np.zeros(pat.n_slots * 576)
    lane = np.asarray(pat.csr_lane, np.int64)
    for q in range(9):
        sell[((lane // 64) * 9 + q) * 64 + lane % 64] = Hb.data[:, q // 3, q % 3]
    x = o.x.copy()
    g = o.buf("g").copy()
    dx = np.linalg.solve(H.toarray(), g)
    solve = {"values": torch.as_tensor(sell), "x": torch.as_tensor(x), "x0": torch.as_tensor(0.3 * dx.reshape(-1, 3)),
             "dx_star": torch.as_tensor(dx.reshape(-1, 3)), "n": n}
    bs = coarse.BlockSystem(pat, n, "cpu")
    return H.tocsr(), x, bs, coarse.CoarseSystem(bs, solve), solve


def _dense_M(H, x, w, sign=-1.0):
    """Dense numpy restatement of M(w)^-1."""
    n = len(x)
    Hd = H.toarray()
    agg = np.arange(n) // AGG
    D = np.zeros((n, 3, 3))
    for i in range(n):
        D[i] = Hd[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    Dinv = np.linalg.inv(D)
    Minv = np.zeros_like(Hd)
    for i in range(n):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Dinv[i]
    for a in range(agg.max() + 1):
        idx = np.nonzero(agg == a)[0]
        c = x[idx].mean(axis=0)
        d = x[idx] - c
        rad = np.sqrt((d * d).sum() / len(idx))
        e = (d / rad).astype(np.float32).astype(np.float64)
        P = np.zeros((3 * len(idx), 6))
        for k, i in enumerate(idx):
            blk = np.zeros((3, 6))
            blk[:, :3] = np.eye(3)
            for j in range(3):  # rotation about axis j: e_j x e
                ej = np.zeros(3)
                ej[j] = 1.0
                blk[:, 3 + j] = np.cross(ej, e[k])
            P[3 * k:3 * k + 3] = blk
        dof = np.concatenate([3 * idx[:, None] + np.arange(3)[None, :]]).ravel()
        Haa = Hd[np.ix_(dof, dof)]
        Dblk = np.zeros_like(Haa)
        for k, i in enumerate(idx):
            Dblk[3 * k:3 * k + 3, 3 * k:3 * k + 3] = Dinv[i]
        W = np.repeat(w[idx], 3)
        Pp = P + sign * W[:, None] * (Dblk @ Haa @ P)
        Ac = Pp.T @ Haa @ Pp
        Minv[np.ix_(dof, dof)] += Pp @ np.linalg.inv(Ac) @ Pp.T
    return Minv


def test_block_products_match_scipy(beam_system):
    H, x, bs, cs, _ = beam_system
    rng = np.random.default_rng(0)
    v = rng.normal(size=(bs.n, 3, 2))
    y = bs.apply(cs.B, torch.as_tensor(v)).numpy()
    ref = np.stack([H @ v[:, :, k].ravel() for k in range(2)], axis=1).reshape(bs.n, 3, 2)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())
    # restriction to aggregates
    agg3 = np.repeat(np.arange(bs.n) // AGG, 3)
    Hc = H.tocoo()
    keep = agg3[Hc.row] == agg3[Hc.col]
    Haa = sp.csr_matrix((Hc.data[keep], (Hc.row[keep], Hc.col[keep])), shape=H.shape)
    yl = bs.apply(cs.B, torch.as_tensor(v), local=True).numpy()
    refl = np.stack([Haa @ v[:, :, k].ravel() for k in range(2)], axis=1).reshape(bs.n, 3, 2)
    np.testing.assert_allclose(yl, refl, rtol=1e-12, atol=1e-9 * np.abs(refl).max())


@pytest.mark.parametrize("kind", ["zero", "uniform", "random"])
def test_preconditioner_matches_dense_restatement_and_is_spd(beam_system, kind):
    H, x, bs, cs, _ = beam_system
    rng = np.random.default_rng(1)
    w = {"zero": np.zeros(bs.n), "uniform": np.full(bs.n, 0.75), "random": rng.uniform(0.0, 1.5, bs.n)}[kind]
    Md = _dense_M(H, x, w)
    V = rng.normal(size=(3 * bs.n, 4))
    Z = np.stack([cs.apply_M(torch.as_tensor(w), torch.as_tensor(V[:, k]).view(-1, 3, 1)).numpy().ravel()
                  for k in range(4)], axis=1)
    np.testing.assert_allclose(Z, Md @ V, rtol=1e-8, atol=1e-10 * np.abs(Md @ V).max())
    np.testing.assert_allclose(Md, Md.T, atol=1e-9 * np.abs(Md).max())
    assert np.linalg.eigvalsh(0.5 * (Md + Md.T)).min() > 0


def test_planted_smoothing_sign_fails_the_dense_restatement(beam_system):
    H, x, bs, cs, _ = beam_system
    w = np.full(bs.n, 0.75)
    Md_wrong = _dense_M(H, x, w, sign=+1.0)
    v = np.random.default_rng(2).normal(size=3 * bs.n)
    z = cs.apply_M(torch.as_tensor(w), torch.as_tensor(v).view(-1, 3, 1)).numpy().ravel()
    assert not np.allclose(z, Md_wrong @ v, rtol=1e-6)


def test_loss_closed_forms(beam_system):
    H, x, bs, cs, solve = beam_system
    w = np.full(bs.n, 0.6)
    Md = _dense_M(H, x, w)
    e = (solve["dx_star"] - solve["x0"]).numpy().ravel()
    He = H @ e
    z = Md @ He
    rich = 1.0 - (z @ He) ** 2 / ((e @ He) * (z @ (H @ z)))
    assert float(cs.loss(torch.as_tensor(w), steps=1)) == pytest.approx(np.log(rich), rel=1e-9)
    # k-step PCG energy error from y = 0 on H y = H e
    y = np.zeros_like(e)
    r = He.copy()
    zz = Md @ r
    p = zz.copy()
    rz = r @ zz
    for _ in range(5):
        Hp = H @ p
        a = rz / (p @ Hp)
        y += a * p
        r -= a * Hp
        zz = Md @ r
        rz2 = r @ zz
        p = zz + rz2 / rz * p
        rz = rz2
    err = e - y
    assert float(cs.loss(torch.as_tensor(w), steps=5)) == pytest.approx(np.log((err @ (H @ err)) / (e @ He)),
                                                                       rel=1e-8)


def test_loss_gradient_matches_finite_differences(beam_system):
    H, x, bs, cs, _ = beam_system
    rng = np.random.default_rng(4)
    w0 = rng.uniform(0.2, 1.0, bs.n)
    w = torch.as_tensor(w0, dtype=torch.float64).requires_grad_(True)
    L = cs.loss(w, steps=3)
    L.backward()
    g = w.grad.numpy()
    for i in rng.choice(bs.n, 6, replace=False):
        h = 1e-6
        wp, wm = w0.copy(), w0.copy()
        wp[i] += h
        wm[i] -= h
        fd = (float(cs.loss(torch.as_tensor(wp), steps=3)) - float(cs.loss(torch.as_tensor(wm), steps=3))) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-9 * max(1.0, np.abs(g).max()))
