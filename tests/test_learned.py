"""Learned CG warm start on the GPU against the NumPy oracle (oracle/mlp.py).

Tolerances:
  * MLP forward: BF16 operands and FP32 accumulation on both sides; only the
    summation order differs, which can flip a BF16 rounding of a hidden
    activation now and then: max |y_gpu - y_orc| <= 4e-3 * max|y_orc| and
    mean <= 1e-4 * max|y_orc|.
  * MLP gradient: |g_gpu - g_orc| <= 1e-2 |g_orc| (same BF16 rounding
    points), loss within 1e-4 relative.
  * Adam: the FP32 master weights within 1e-7 absolute + 3e-7 relative
    (their FP32 rounding).
  * features: BF16 values within one BF16 step (2^-7 relative) + 1e-3;
    candidates within 1e-4 of their scale.
  * steps with the warm start: positions within 1e-4 relative of the FP64
    oracle (the same bar as plain PCG), relative residual <= 1e-6.
"""
import numpy as np
import pytest
import torch

from mfem_amd import scene
from oracle import Oracle, mlp

pytestmark = pytest.mark.gpu


def _make(model, device, **extra):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = {k: v for k, v in model.params.items() if k != "iterations"}
    solver = RefinementSolver(model, model.params["iterations"], max_tets=model.tet_count, device=device,
                              enable_refinement=False, **kw, **extra)
    return solver, model_state(model, device)


def _net(device, seed=0, jitter=0.3):
    from mfem_amd.learned import WarmStartNet, init_params

    w = init_params(seed)
    rng = np.random.default_rng(seed + 100)
    w += (jitter * rng.normal(0, 1, w.shape) * (np.abs(w) + 0.05)).astype(np.float32)
    return WarmStartNet(device, weights=w)


def _bf16_tensor(a, device):
    return torch.from_numpy(mlp.bf16_bits(a).view(np.int16).copy()).to(device)


def test_mlp_forward_matches_oracle(gpu):
    net = _net(gpu)
    rng = np.random.default_rng(7)
    n = 1000  # not a multiple of the 64-sample tile
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    y = net.forward(_bf16_tensor(X, gpu)).cpu().numpy()
    y_orc, _ = mlp.forward(net.weights(), mlp.bf16(X), quantize=True)
    err = np.abs(y - y_orc)
    scale = np.abs(y_orc).max()
    print("forward max err", err.max(), "mean", err.mean(), "scale", scale)
    assert err.max() <= 4e-3 * scale
    assert err.mean() <= 1e-4 * scale
    # a layout mistake in the MFMA fragments would show at O(1): also check
    # the unrounded FP64 network agrees to BF16 accuracy
    y64, _ = mlp.forward(net.weights(), mlp.bf16(X), quantize=False)
    assert np.abs(y - y64).max() <= 3e-2 * scale


def test_mlp_gradient_and_loss_match_oracle(gpu):
    net = _net(gpu, seed=1)
    rng = np.random.default_rng(8)
    n = 3000
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    aux = rng.normal(0, 0.5, (n, 12)).astype(np.float32)
    grad = torch.zeros(mlp.N_PARAMS, dtype=torch.float32, device=gpu)
    loss = torch.zeros(1, dtype=torch.float32, device=gpu)
    w0 = net.weights()
    net.train_step(_bf16_tensor(X, gpu), torch.from_numpy(aux).to(gpu), 0.0, loss_out=loss, grad_out=grad)
    assert np.array_equal(net.weights(), w0)  # lr = 0 leaves the weights alone
    l_orc, g_orc = mlp.loss_and_grad(w0, mlp.bf16(X), aux, quantize=True)
    g = grad.cpu().numpy()
    rel = np.linalg.norm(g - g_orc) / np.linalg.norm(g_orc)
    print("loss", float(loss.item()), l_orc, "grad rel err", rel)
    assert abs(float(loss.item()) - l_orc) <= 1e-4 * l_orc
    assert rel <= 1e-2
    for name, a, b in (("W1", mlp.O_W1, mlp.O_B1), ("W2", mlp.O_W2, mlp.O_B2), ("W3", mlp.O_W3, mlp.O_B3),
                       ("b1", mlp.O_B1, mlp.O_W2), ("b2", mlp.O_B2, mlp.O_W3), ("b3", mlp.O_B3, mlp.N_PARAMS)):
        r = np.linalg.norm(g[a:b] - g_orc[a:b]) / max(np.linalg.norm(g_orc[a:b]), 1e-30)
        print(name, "rel err", r)
        assert r <= 2e-2, name


def test_adam_step_matches_oracle(gpu):
    net = _net(gpu, seed=2)
    rng = np.random.default_rng(9)
    n = 640
    Xt = _bf16_tensor(rng.normal(0, 1, (n, 16)).astype(np.float32), gpu)
    auxt = torch.from_numpy(rng.normal(0, 0.5, (n, 12)).astype(np.float32)).to(gpu)
    w0 = net.weights()
    grad = torch.zeros(mlp.N_PARAMS, dtype=torch.float32, device=gpu)
    m = np.zeros(mlp.N_PARAMS)
    v = np.zeros(mlp.N_PARAMS)
    w_ref = w0.astype(np.float64)
    for t in range(1, 4):
        net.train_step(Xt, auxt, 1e-3, grad_out=grad)
        w_ref, m, v = mlp.adam(w_ref, m, v, grad.cpu().numpy().astype(np.float64), 1e-3, t)
        np.testing.assert_allclose(net.weights(), w_ref, rtol=3e-7, atol=1e-7)  # FP32 master weights
    # the BF16 copy follows the master weights
    np.testing.assert_array_equal(net.wbf.cpu().numpy().view(np.uint16), mlp.bf16_bits(net.weights()))


def test_training_reduces_loss_on_solver_data(gpu):
    from mfem_amd.learned import WarmStartNet, collect

    m = scene.drop_box_scene()
    solver, state = _make(m, gpu)
    X, aux = collect(solver, state, 0.01, steps=6)
    print("samples", X.shape[0])
    assert X.shape[0] > 1000
    net = WarmStartNet(gpu)
    l0 = net.loss(X, aux)
    hist = net.fit(X, aux, epochs=30, batch=4096, lr=3e-3)
    print("loss", l0, "->", hist["train"][-1], "val", hist["val"][:1], hist["val"][-1:])
    assert hist["val"][-1] < 0.9 * hist["val"][0]
    assert hist["train"][-1] < l0


def test_features_match_oracle(gpu):
    m = scene.beam_scene()
    solver, state = _make(m, gpu)
    dt = 1.0 / 60.0
    out = state.clone()
    solver.step(state, out, None, None, dt)  # a previous step, so c3 != 0
    solver.bind(out)
    solver.begin_step(dt)
    solver.update_system(dt)
    solver.assemble(dt)
    solver.global_solve(0)
    solver.local_solve()
    solver.line_search(dt, 0)
    solver.update_system(dt)
    solver.assemble(dt)
    solver.ws_features(dt, 1)
    b = solver.buffers
    torch.cuda.synchronize()
    F, aux = mlp.ws_features(solver.system_csr(), b["g"].double().cpu().numpy(), b["x_tilde"].double().cpu().numpy(),
                             out.particle_q.double().cpu().numpy(), b["dx"].cpu().numpy(),
                             b["Hv"].double().cpu().numpy(), m.particle_inv_mass, precond=True,
                             attach_k=m.params.get("attachment_stiffness", 10.0), dt=dt, newton_iter=1,
                             iterations=solver.iterations, alpha_prev=float(solver.last_stats()[0, 4]))
    Fg = mlp.bf16_value(b["ws_x"].cpu().numpy().view(np.uint16))
    ag = b["ws_aux"].double().cpu().numpy()
    err = np.abs(Fg - F)
    tol = 2.0 ** -7 * np.abs(F) + 1e-3
    bad = np.argwhere(err > tol)
    print("feature max err", err.max(axis=0))
    assert len(bad) == 0, bad[:10]
    np.testing.assert_allclose(ag[:, 9], aux[:, 9], rtol=1e-4)
    assert np.abs(ag[:, :9] - aux[:, :9]).max() <= 1e-4


def _run(m, gpu, steps, dt, net=None):
    solver, state = _make(m, gpu, warm_start=net)
    out = state.clone()
    stats = []
    for k in range(steps):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        stats.append(solver.last_stats())
        assert solver.status() == 0
    return state.particle_q.double().cpu().numpy(), np.stack(stats)


def test_warm_start_steps_match_fp64_oracle(gpu):
    m = scene.drop_box_scene()
    net = _net(gpu, seed=3, jitter=0.05)
    x, st = _run(m, gpu, 6, 0.01, net)
    o = Oracle(m, 64)
    for k in range(6):
        o.step(0.01)
    assert np.all(st[:, :, 1] <= 1e-6)
    scale = np.abs(o.x).max()
    err = np.abs(x - o.x).max()
    print("learned guess used", st[:, :, 14].astype(int).tolist(), "ratio", np.round(st[:, :, 15], 3).tolist())
    print("positions rel err", err / scale)
    assert err <= 1e-4 * scale


def test_poisoned_network_gets_no_weight(gpu):
    """Weights whose guess is far off: the Galerkin start is used, its
    residual stays near the plain guess's (the far-off direction gets next to
    no weight), and the steps still meet the FP64 oracle's positions within
    1e-4."""
    from mfem_amd.learned import WarmStartNet

    m = scene.drop_box_scene()
    w = np.zeros(mlp.N_PARAMS, np.float32)
    w[mlp.O_B3:mlp.O_B3 + 3] = 1e4
    net = WarmStartNet(gpu, weights=w)
    x_l, st_l = _run(m, gpu, 4, 0.01, net)
    assert np.all(st_l[:, :, 14] == 3)
    assert np.all(np.isfinite(st_l[:, :, 15])) and np.all(st_l[:, :, 15] < 10.0)
    assert np.all(st_l[:, :, 1] <= 1e-6)
    o = Oracle(m, 64)
    for k in range(4):
        o.step(0.01)
    assert np.abs(x_l - o.x).max() <= 1e-4 * np.abs(o.x).max()


def _galerkin_host(H, g, cands, thr=1e-10):
    """Lowest quadratic-model energy over span(cands): x0 = C a with
    (C^T H C) a = C^T g, by Cholesky with pivots <= thr * diagonal dropped
    (candidates already in the span of the earlier ones).  Returns x0 and
    the smallest margin of a pivot ratio from thr (decision robustness)."""
    C = np.stack(cands, 1)
    G = C.T @ (H @ C)
    f = C.T @ g
    K = len(cands)
    L = np.zeros((K, K))
    kept = np.zeros(K, bool)
    margin = np.inf
    for j in range(K):
        d = G[j, j] - sum(L[j, q] ** 2 for q in range(j) if kept[q])
        if G[j, j] > 0:
            margin = min(margin, abs(np.log10(max(d / G[j, j], 1e-300)) - np.log10(thr)))
        kept[j] = G[j, j] > 0 and d > thr * G[j, j]
        L[j, j] = np.sqrt(d) if kept[j] else 1.0
        for i in range(j + 1, K):
            v = G[j, i] - sum(L[i, q] * L[j, q] for q in range(j) if kept[q])
            L[i, j] = v / L[j, j] if kept[j] else 0.0
    z = np.zeros(K)
    for j in range(K):
        z[j] = (f[j] - L[j, :j] @ z[:j]) / L[j, j] if kept[j] else 0.0
    a = np.zeros(K)
    for j in reversed(range(K)):
        a[j] = (z[j] - L[j + 1:, j] @ a[j + 1:]) / L[j, j] if kept[j] else 0.0
    return C @ a, margin


@pytest.mark.parametrize("jitter", [None, 0.0, 0.05, 3.0])
def test_galerkin_start_matches_host(gpu, jitter):
    """The PCG starts from the lowest quadratic-model energy over the span of
    x~ - x, the learned guess, the last three Newton steps and the same Newton
    iteration's solution one step earlier (stats[14] = 3): the start's
    residual ratio |g - H x0| / |g - H (x~ - x)|
    (stats[15]) equals the one of the same Galerkin start computed here in
    FP64 on the host from the assembled matrix and the solver's buffers, and
    that start's energy is at most every candidate's; nets from exact
    previous-step (zero output layer) to far off, and no network at all (the
    kernel's schedule without a learned guess: that candidate is zero)."""
    m = scene.drop_box_scene()
    if jitter is None:
        solver, state = _make(m, gpu)
        solver.set_galerkin_start(True)
    else:
        net = _net(gpu, seed=5, jitter=jitter)
        solver, state = _make(m, gpu, warm_start=net)
    out = state.clone()
    for k in range(2):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
    out.assign(state)
    solver.bind(out)
    solver.begin_step(0.01)
    b = solver.buffers
    n = m.particle_count
    rp = solver._S.n_chunks * 64
    checked = 0
    for it in range(solver.iterations):
        solver.update_system(0.01)
        solver.assemble(0.01)
        if jitter is not None:
            solver.warm_start(0.01, it)
        torch.cuda.synchronize()
        H = solver.system_csr().tocsr()
        g = b["g"][:n].double().cpu().numpy().reshape(-1)
        hist = b["dx_hist"].cpu().numpy()
        same = b["dx_same"].cpu().numpy()[3 * rp * it: 3 * rp * it + 3 * n]
        c0 = (b["x_tilde"][:n].double() - out.particle_q[:n].double()).cpu().numpy().reshape(-1)
        guess = np.zeros_like(c0) if jitter is None else b["guess"][:n].double().cpu().numpy().reshape(-1)
        cands = [c0, guess, b["dx"][:n].cpu().numpy().reshape(-1),
                 hist[: 3 * n], hist[3 * rp: 3 * rp + 3 * n], same]
        x0, margin = _galerkin_host(H, g, cands)
        phi = lambda c: 0.5 * c @ (H @ c) - g @ c
        assert phi(x0) <= min(phi(c) for c in cands) + 1e-9 * abs(phi(x0))
        ratio = np.linalg.norm(g - H @ x0) / np.linalg.norm(g - H @ c0)
        solver.global_solve(it)
        st = solver.last_stats()[it]
        assert int(st[14]) == 3
        if margin > 0.5:  # no pivot within a factor 3 of the drop threshold
            assert abs(st[15] - ratio) <= 2e-2 * ratio + 1e-12, (it, st[15], ratio)
            checked += 1
        solver.local_solve()
        solver.line_search(0.01, it)
    solver.end_step(0.01)
    print("jitter", jitter, "iterations compared", checked)
    assert checked >= solver.iterations // 2
    assert solver.status() == 0


def test_backward_from_output_gradient_matches_oracle(gpu):
    net = _net(gpu, seed=4)
    rng = np.random.default_rng(10)
    n = 2000
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    dy = rng.normal(0, 1e-3, (n, 3)).astype(np.float32)
    grad = torch.zeros(mlp.N_PARAMS, dtype=torch.float32, device=gpu)
    w0 = net.weights()
    net.backward_step(_bf16_tensor(X, gpu), torch.from_numpy(dy).to(gpu), 0.0, grad_out=grad)
    assert np.array_equal(net.weights(), w0)
    g_orc = mlp.backward(w0, mlp.bf16(X), dy, quantize=True)
    g = grad.cpu().numpy()
    rel = np.linalg.norm(g - g_orc) / np.linalg.norm(g_orc)
    print("grad rel err", rel)
    assert rel <= 1e-2


def test_residual_training_on_solver_solves(gpu):
    """The residual objective (log |g - H dx0|^2 per solve, H applied by the
    project's SpMV) decreases with training, its gradient agrees with the
    oracle's on one solve, and the trained guess cuts the initial residual."""
    from mfem_amd.learned import WarmStartNet, collect_solves

    m = scene.drop_box_scene()
    solver, state = _make(m, gpu)
    out = state.clone()
    for k in range(2):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
    solves = collect_solves(solver, state, 0.01, steps=3)
    # gradient of one solve against the oracle (H from the stored values)
    import scipy.sparse as sp
    from mfem_amd.pattern import sell_to_csr_values

    net = WarmStartNet(gpu)
    sv = solves[5]
    y = net.forward(sv["X"]).cpu().numpy()
    vals = sell_to_csr_values(solver.pattern, sv["values"].cpu().numpy())
    n = m.particle_count
    H = sp.bsr_matrix((vals, solver.pattern.cols, solver.pattern.rowptr), shape=(3 * n, 3 * n))
    L_orc, dy_orc = mlp.residual_loss_and_dy(y, sv["cand"].cpu().numpy(), sv["g"].cpu().numpy(), H)
    C = sv["cand"].view(-1, 3, 3)
    dx0 = C[:, 2] + torch.einsum("nk,nkd->nd", torch.from_numpy(y).to(gpu), C)
    r = sv["g"] - solver.spmv_with(sv["values"], dx0)
    L = float(torch.log(torch.sum(r.double() ** 2)).item())
    print("residual objective", L, L_orc)
    assert abs(L - L_orc) <= 1e-3 * max(1.0, abs(L_orc))
    hist = net.fit_residual(solves, solver, epochs=15)
    print("base", hist["base"], "loss", [round(v, 3) for v in hist["loss"]], "kept", hist["best_epoch"])
    # the untrained network (zero output layer) is exactly the previous
    # Newton step; the kept weights are the best epoch, never worse
    assert abs(hist["loss"][0] - hist["base"]) <= 1e-6 * abs(hist["base"]) + 1e-9
    final = net.residual_objective(solves, solver)
    assert abs(final - min(hist["loss"])) <= 1e-6 * abs(final) + 1e-9
    assert final <= hist["loss"][0]
    assert solver.status() == 0


def test_energy_training_and_cg_validation(gpu):
    """Energy-norm objective (log e^T H e, e = dx0 - dx*): the GPU loss and
    output gradient agree with the oracle's on one logged solve (H from the
    stored values, dx* the logged PCG solution), training never returns a
    worse objective than it started from, and validate_cg keeps whichever of
    the trained correction and the previous Newton step took fewer CG
    iterations, with the solver state restored."""
    import scipy.sparse as sp
    from mfem_amd.learned import WarmStartNet, collect_solves
    from mfem_amd.pattern import sell_to_csr_values

    m = scene.drop_box_scene()
    solver, state = _make(m, gpu)
    out = state.clone()
    for k in range(2):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
    solves = collect_solves(solver, state, 0.01, steps=3)
    n = m.particle_count
    net = _net(gpu, seed=3, jitter=0.05)
    sv = solves[4]
    y = net.forward(sv["X"])
    vals = sell_to_csr_values(solver.pattern, sv["values"].cpu().numpy())
    H = sp.bsr_matrix((vals, solver.pattern.cols, solver.pattern.rowptr), shape=(3 * n, 3 * n))
    L_orc, dy_orc = mlp.energy_loss_and_dy(y.cpu().numpy()[:n], sv["cand"].cpu().numpy()[:n],
                                           sv["dx_star"].cpu().numpy()[:n], H)
    from mfem_amd.learned import _blend

    C = sv["cand"].view(-1, 3, 3)
    L, gd = net._solve_loss(sv, solver, "energy", _blend(C, y), True)
    dy = (C * gd[:, None, :]).sum(dim=2).cpu().numpy()[:n]
    print("energy objective", float(L), L_orc)
    assert abs(float(L) - L_orc) <= 1e-3 * max(1.0, abs(L_orc))
    assert np.linalg.norm(dy - dy_orc) <= 1e-2 * np.linalg.norm(dy_orc)
    net = WarmStartNet(gpu)
    hist = net.fit_solves(solves, solver, "energy", epochs=5)
    print("energy objective per epoch", [round(v, 3) for v in hist["loss"]], "base", hist["base"])
    assert abs(hist["loss"][0] - hist["base"]) <= 1e-6 * abs(hist["base"]) + 1e-9
    assert min(hist["loss"]) == pytest.approx(net.objective(solves, solver, "energy"), rel=1e-6, abs=1e-9)
    before = {k: t.clone() for k, t in solver.snapshot().items()}
    val = net.validate_cg(solver, state, 0.01, steps=2)
    print("validation", val)
    after = solver.snapshot()
    for k, t in before.items():
        assert torch.equal(t, after[k]), k
    c = val["cg_iters"]
    assert val["kept"] == ("trained" if c["trained"] < c["prev_newton_step"] else "prev_newton_step")
    if val["kept"] == "prev_newton_step":
        assert float(net.w[-(64 * 16 + 16):].abs().max()) == 0.0
    assert solver.status() == 0
