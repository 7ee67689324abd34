"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (FP32 bulk arithmetic on the GPU vs the FP64 oracle):
  * per-tet / assembled quantities: |gpu - orc| <= 2e-4 * max|orc| (FP32
    element math incl. a 3x3 SVD, summed over up to ~24 tets per block)
  * PCG: relative residual of the GPU's own system <= 1e-6 (computed in FP64
    on the GPU and re-checked here with scipy in FP64)
  * positions after full steps: max|x_gpu - x_orc| <= 1e-4 * max|x_orc|
    (the bar BASELINE sets for positions vs the FP64 oracle)
"""
import numpy as np
import pytest
import scipy.sparse.linalg
import torch

from mfem_amd import scene
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _solver_kwargs(model):
    kw = {k: v for k, v in model.params.items() if k != "iterations"}
    kw["enable_refinement"] = False
    return kw


def _make(model, device, **extra):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = _solver_kwargs(model)
    kw.update(extra)  # explicit arguments override the scene's parameters
    solver = RefinementSolver(model, model.params["iterations"], max_tets=model.tet_count, device=device, **kw)
    return solver, model_state(model, device)


def _close(a, b, rel, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max()
    print(f"{what}: max abs err {err:.3e}, scale {scale:.3e}, rel {err / scale:.3e}", flush=True)
    assert err <= rel * scale, f"{what}: rel err {err / scale:.3e} > {rel}"


def _unpack21(p):
    n = p.shape[0]
    out = np.zeros((n, 6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            out[:, i, j] = p[:, k]
            out[:, j, i] = p[:, k]
            k += 1
    return out


def _perturbed_scene(name):
    if name == "beam_arap":
        m = scene.beam_scene()
        dt = 1.0 / 60.0
    elif name == "beam_nh":
        m = scene.beam_scene(energy="neohookean")
        dt = 1.0 / 60.0
    elif name == "capsule_poke":
        m = scene.capsule_poke_scene()
        dt = 0.01
    else:
        m = scene.drop_box_scene()
        dt = 0.01
    return m, dt


def _sync_oracle(o, solver, state):
    o.buf("x")[:] = state.particle_q.double().cpu().numpy().ravel()
    o.buf("v")[:] = state.particle_qd.double().cpu().numpy().ravel()
    o.buf("s")[:] = solver.buffers["s"].double().cpu().numpy().ravel()
    o.buf("lam")[:] = solver.buffers["lam"].double().cpu().numpy().ravel()


@pytest.mark.parametrize("name", ["beam_arap", "beam_nh", "drop_box", "capsule_poke"])
def test_newton_stages_match_oracle(gpu, name):
    m, dt = _perturbed_scene(name)
    solver, state = _make(m, gpu, fuse_local_solve=False)  # the stage order: records, then the local solve
    rng = np.random.default_rng(1)
    # a deformed, moving configuration: perturb positions and velocities
    q = m.particle_q + 0.004 * rng.normal(size=m.particle_q.shape)
    q[:, 2] = np.maximum(q[:, 2], 2e-4)  # keep above the ground plane
    state.particle_q.copy_(torch.as_tensor(q, dtype=torch.float32))
    state.particle_qd.copy_(torch.as_tensor(0.05 * rng.normal(size=q.shape), dtype=torch.float32))
    o = Oracle(m, 64)
    _sync_oracle(o, solver, state)

    solver.bind(state)
    solver.begin_step(dt)
    o.begin_step(dt)
    solver.update_system(dt)
    o.update_system(dt)
    solver.assemble(dt)
    o.assemble(dt)
    torch.cuda.synchronize()
    B = {k: v.double().cpu().numpy() for k, v in solver.buffers.items()}
    nt = m.tet_count
    _close(B["x_tilde"].ravel(), o.buf("x_tilde"), 1e-6, "x_tilde")
    _close(B["G"].ravel(), o.buf("G"), 2e-4, "G")
    # g_s = vol mu (SYM s - 1) + ... cancels to FP32 noise near the rest
    # stretch: scale its tolerance by the size of the terms, vol * mu * |s|
    gs_scale = np.max(np.abs(1.0 / (6.0 * np.linalg.det(m.tet_poses))) * m.tet_materials[:, 0]) * 3.0
    err = np.abs(B["gs"].ravel() - o.buf("gs")).max()
    print(f"g_s: max abs err {err:.3e}, term scale {gs_scale:.3e}")
    assert err <= 2e-6 * gs_scale
    _close(_unpack21(B["Hl"]).ravel(), o.buf("Hl"), 2e-4, "H_lambda")
    _close(_unpack21(B["Hsinv"]).ravel(), o.buf("Hsinv"), 2e-4, "H_s^-1")
    _close(B["gl"].ravel(), o.buf("gl"), 2e-4, "g_lambda")
    # psi also cancels to ~0 at the rest stretch: tolerance from the term sizes
    psi_scale = np.max(np.abs(1.0 / (6.0 * np.linalg.det(m.tet_poses))) * (m.tet_materials[:, 0] + m.tet_materials[:, 1]))
    err = np.abs(B["psi"].ravel() - o.buf("psi")).max()
    print(f"psi: max abs err {err:.3e}, term scale {psi_scale:.3e}")
    assert err <= 2e-6 * psi_scale
    _close(B["gB"].ravel(), o.buf("gB"), 2e-4, "barrier grad")
    _close(B["gD"].ravel(), o.buf("gD"), 2e-4, "friction grad")
    Hg = solver.system_csr().toarray()
    Ho = o.csr().toarray()
    _close(Hg, Ho, 2e-4, "H")
    _close(B["g"].ravel(), o.buf("g"), 2e-4, "g")
    np.testing.assert_allclose(Hg, Hg.T, atol=1e-5 * np.abs(Hg).max())

    # SpMV through the C ABI vs scipy on the GPU's own matrix
    xv = rng.normal(size=(m.particle_count, 3))
    y = solver.spmv(torch.as_tensor(xv, dtype=torch.float32, device=gpu)).double().cpu().numpy()
    _close(y.ravel(), solver.system_csr() @ xv.ravel(), 1e-5, "spmv")

    # PCG on the GPU's own system, rel-res <= 1e-6 in FP64
    solver.global_solve(0)
    o.global_solve()
    st = solver.last_stats()
    dxg = solver.buffers["dx"].cpu().numpy().ravel()
    gg = solver.buffers["g"].double().cpu().numpy().ravel()
    H = solver.system_csr().tocsr().astype(np.float64)
    rr = np.linalg.norm(gg - H @ dxg) / np.linalg.norm(gg)
    print("pcg iters", st[0, 0], "relres(dev)", st[0, 1], "relres(host)", rr, "restarts", st[0, 7])
    assert st[0, 1] <= 1e-6 and rr <= 1.5e-6
    ref = scipy.sparse.linalg.spsolve(H.tocsc(), gg)
    _close(dxg, ref, 1e-4, "dx vs direct solve of the GPU system")
    _close(dxg, o.buf("dx"), 1e-3, "dx vs oracle")

    # local solve vs the closed form on the GPU's own inputs
    solver.local_solve()
    torch.cuda.synchronize()
    lam = solver.buffers["lam"].double().cpu().numpy()
    ds = solver.buffers["ds"].double().cpu().numpy()
    G = B["G"].reshape(nt, 6, 12)
    Hl = _unpack21(B["Hl"])
    Hsi = _unpack21(B["Hsinv"])
    de = dxg.reshape(-1, 3)[m.tet_indices].reshape(nt, 12)
    lam_ref = -np.einsum("eij,ej->ei", Hl, B["gl"] - np.einsum("eij,ej->ei", G, de))
    ds_ref = -np.einsum("eij,ej->ei", Hsi, lam_ref * np.array([1, 1, 1, 2, 2, 2]) + B["gs"])
    _close(lam, lam_ref, 2e-4, "lambda")
    _close(ds, ds_ref, 2e-4, "ds")

    # merit (FP64 sums) at the current state vs the oracle's
    o.buf("lam")[:] = lam.ravel()
    f_gpu = solver.merit(dt, 0.0)
    f_orc = o.merit(o.x, o.buf("s"), dt)
    print("merit", f_gpu, f_orc)
    assert abs(f_gpu - f_orc) <= 1e-4 * max(abs(f_orc), 1e-12) + 1e-7


@pytest.mark.parametrize("name,steps", [("beam_arap", 5), ("beam_nh", 5), ("drop_box", 8), ("capsule_poke", 5)])
def test_full_steps_match_fp64_oracle(gpu, name, steps):
    m, dt = _perturbed_scene(name)
    solver, state = _make(m, gpu)
    o = Oracle(m, 64)
    out = state.clone()
    for k in range(steps):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        o.step(dt)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st
        assert solver.status() == 0
    x = state.particle_q.double().cpu().numpy()
    v = state.particle_qd.double().cpu().numpy()
    _close(x, o.x, 1e-4, f"{name} positions after {steps} steps")
    disp = np.abs(o.x - m.particle_q).max()
    print("displacement", disp, "abs err", np.abs(x - o.x).max())
    assert np.all(np.isfinite(v))


@pytest.mark.parametrize("pipelined", [False, True])
@pytest.mark.parametrize("resident", [True, False])
def test_pcg_variants_match_oracle(gpu, pipelined, resident):
    # every persistent-PCG variant: standard / pipelined, rows in VGPRs / streamed
    m, dt = _perturbed_scene("drop_box")
    solver, state = _make(m, gpu, pcg_pipelined=pipelined, pcg_resident_rows=resident)
    o = Oracle(m, 64)
    out = state.clone()
    for k in range(4):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        o.step(dt)
        st = solver.last_stats()
        print("variant", pipelined, resident, "cg", st[:, 0].astype(int).tolist(), "restarts", st[:, 7].astype(int).tolist())
        assert np.all(st[:, 1] <= 1e-6), st
        assert solver.status() == 0
    _close(state.particle_q.double().cpu().numpy(), o.x, 1e-4, f"positions pipelined={pipelined} resident={resident}")


def test_every_pcg_variant_matches_oracle(gpu):
    # each variant of the table forced in turn (chunks per wave 1-8, resident /
    # streamed / low-register rows, standard / pipelined), from the same start
    m, dt = _perturbed_scene("drop_box")
    solver, state0 = _make(m, gpu)
    o = Oracle(m, 64)
    for k in range(3):
        o.step(dt)
    snap = solver.snapshot()
    variants = [v for v in solver.pcg_variants() if v["applicable"]]
    assert {v["mode"] for v in variants} == {"standard", "pipelined", "standard_lowreg"}
    for v in variants:
        solver._S.pcg_variant = v["variant"]
        solver.restore(snap)
        state, out = state0.clone(), state0.clone()
        for k in range(3):
            solver.step(state, out, None, None, dt)
            state, out = out, state
            st = solver.last_stats()
            assert np.all(st[:, 1] <= 1e-6), (v, st)
            assert solver.status() == 0
        print("variant", v, "cg", int(st[:, 0].sum()))
        _close(state.particle_q.double().cpu().numpy(), o.x, 1e-4, f"positions, variant {v}")


@pytest.mark.parametrize("sweeps", [2, 4])
@pytest.mark.parametrize("name", ["beam_arap", "beam_nh", "drop_box", "capsule_poke"])
def test_local_preconditioner_matches_oracle(gpu, name, sweeps):
    # workgroup-local damped block-Jacobi sweeps added to the preconditioner:
    # same solution (positions within 1e-4 of the FP64 oracle, every solve
    # at rel-res <= 1e-6) in fewer CG iterations than block Jacobi
    m, dt = _perturbed_scene(name)
    plain, s_p = _make(m, gpu)
    local, s_l = _make(m, gpu, pcg_local_sweeps=sweeps)
    o = Oracle(m, 64)
    o_p, o_l = s_p.clone(), s_l.clone()
    it_p = it_l = 0
    for k in range(4):
        plain.step(s_p, o_p, None, None, dt)
        s_p, o_p = o_p, s_p
        local.step(s_l, o_l, None, None, dt)
        s_l, o_l = o_l, s_l
        o.step(dt)
        st_p, st_l = plain.last_stats(), local.last_stats()
        assert np.all(st_l[:, 1] <= 1e-6), st_l
        assert local.status() == 0
        it_p += int(st_p[:, 0].sum())
        it_l += int(st_l[:, 0].sum())
    print(name, "sweeps", sweeps, "CG iterations block Jacobi", it_p, "local", it_l)
    _close(s_l.particle_q.double().cpu().numpy(), o.x, 1e-4, f"{name} positions, {sweeps} local sweeps")
    assert it_l < it_p


def test_local_preconditioner_every_standard_variant(gpu):
    # the local sweeps under every standard variant (resident / streamed /
    # low-register rows, 1-8 chunks per wave: different workgroup row sets)
    m, dt = _perturbed_scene("drop_box")
    solver, state0 = _make(m, gpu, pcg_local_sweeps=2)
    o = Oracle(m, 64)
    for k in range(3):
        o.step(dt)
    snap = solver.snapshot()
    variants = [v for v in solver.pcg_variants() if v["applicable"]]
    assert variants and all(v["mode"] != "pipelined" for v in variants)
    assert {v["chunks_per_wave"] for v in variants} == {1, 2, 4, 8}
    for v in variants:
        solver._S.pcg_variant = v["variant"]
        solver.restore(snap)
        state, out = state0.clone(), state0.clone()
        for k in range(3):
            solver.step(state, out, None, None, dt)
            state, out = out, state
            assert np.all(solver.last_stats()[:, 1] <= 1e-6)
            assert solver.status() == 0
        _close(state.particle_q.double().cpu().numpy(), o.x, 1e-4, f"positions, local sweeps, variant {v}")


def test_graph_free_repeatability(gpu):
    # deterministic reductions and gathers: two runs give bit-identical states
    m = scene.drop_box_scene()
    res = []
    for _ in range(2):
        solver, state = _make(m, gpu)
        out = state.clone()
        for k in range(3):
            solver.step(state, out, None, None, 0.01)
            state, out = out, state
        res.append(state.particle_q.cpu().numpy().copy())
    np.testing.assert_array_equal(res[0], res[1])


@pytest.mark.parametrize("pcg,galerkin,tol,restart", [({}, False, 1e-10, True),
                                                      ({"pcg_pipelined": True}, False, 1e-12, True),
                                                      ({"pcg_pipelined": True}, True, 1e-12, False),
                                                      ({"pcg_coarse": True, "pcg_coarse_omega": 0.75}, True, 1e-12,
                                                       False)])
def test_tight_tolerance_solves_and_restarts(gpu, pcg, galerkin, tol, restart):
    # CG far below the default rel-res: the recurrences drift from the true
    # residual there (FP32 directions at 1e-10, the pipelined forms' FP64
    # recurrences near 1e-12), so the FP64 check restarts the Krylov
    # sequence (the restart path, its SpMV inside the reduction's round trip
    # in the pipelined forms); every solve must still end at the tolerance
    # on the host's own FP64 residual of the assembled system
    m = scene.drop_box_scene()
    solver, state = _make(m, gpu, cg_tolerance=tol, cg_max_iterations=4000, **pcg)
    if galerkin:
        solver.set_galerkin_start(True)
    out = state.clone()
    solver.step(state, out, None, None, 0.01)  # (histories for the Galerkin start)
    state, out = out, state
    out.assign(state)
    solver.bind(out)
    solver.begin_step(0.01)
    n = m.particle_count
    restarts = 0
    for it in range(4):
        solver.update_system(0.01)
        solver.assemble(0.01)
        solver.global_solve(it)
        st = solver.last_stats()[it]
        dx = solver.buffers["dx"][:n].cpu().numpy().ravel()
        g = solver.buffers["g"][:n].double().cpu().numpy().ravel()
        H = solver.system_csr().tocsr().astype(np.float64)
        rr = np.linalg.norm(g - H @ dx) / np.linalg.norm(g)
        print(pcg, galerkin, "it", it, "cg", int(st[0]), "relres(dev)", st[1], "relres(host)", rr, "restarts",
              int(st[7]), flush=True)
        assert st[1] <= tol and rr <= 2 * tol
        restarts += int(st[7])
        solver.local_solve()
        solver.line_search(0.01, it)
    assert solver.status() == 0
    print("restarts", restarts)
    if restart:  # (these forms restart at this tolerance on this scene: the restart path ran)
        assert restarts > 0


@pytest.mark.parametrize("pcg", [{}, {"pcg_pipelined": True}, {"pcg_coarse": True, "pcg_coarse_omega": 0.75}])
def test_cooperative_and_ordinary_launches_agree(gpu, pcg):
    # the persistent kernels (PCG, line search) launched cooperatively or as
    # ordinary launches within the resident workgroup count: the same kernels,
    # so bit-identical states, CG iteration counts and line-search results
    m = scene.drop_box_scene()
    res = []
    for coop in (True, False):
        solver, state = _make(m, gpu, cooperative_launch=coop, **pcg)
        out = state.clone()
        stats = []
        for k in range(3):
            solver.step(state, out, None, None, 0.01)
            state, out = out, state
            stats.append(solver.last_stats()[:, :7].copy())
        assert solver.status() == 0
        res.append((state.particle_q.cpu().numpy().copy(), np.stack(stats)))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])


def test_octopus_full_size_step(gpu):
    # full BASELINE size (~2e5 tets): size-independent properties
    m = scene.octopus_scene()
    solver, state = _make(m, gpu)
    out = state.clone()
    for k in range(2):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
        st = solver.last_stats()
        print("octopus step", k, st[:, :4].tolist(), flush=True)
        assert np.all(st[:, 1] <= 1e-6)
        assert solver.status() == 0
    x = state.particle_q.cpu().numpy()
    assert np.all(np.isfinite(x))
    assert x[:, 2].min() > -1e-3  # barrier keeps the body on the ground
    # total mass-weighted displacement is downward (falling)
    assert (m.particle_mass * (x[:, 2] - m.particle_q[:, 2])).sum() < 0


def test_moving_capsule_triangle_contact_matches_oracle(gpu):
    """A capsule driven down into a pinned slab through body_q: triangle and
    vertex barriers against the capsule, friction, per-step poses; positions
    vs the FP64 oracle, and the triangle term is really active."""
    m = scene.capsule_poke_scene()
    dt = 0.01
    solver, state = _make(m, gpu)
    o = Oracle(m, 64)
    out = state.clone()
    pose = m.body_q[0].astype(np.float64).copy()  # the capsule is rigid body 0
    min_tri = []
    for k in range(6):
        pose[2] -= 0.0015  # 1.5 mm per step into the slab's top face
        state.body_q = torch.as_tensor(pose[None, :], dtype=torch.float32, device=gpu)
        o.set_body_q(pose[None, :])
        solver.step(state, out, None, None, dt)
        state, out = out, state
        o.step(dt)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st
        assert solver.status() == 0
        min_tri.append(float(solver.buffers["tri_dist"].min().item()))
    print("min triangle distance per step", min_tri)
    assert min(min_tri) < m.params["contact_d1"]
    x = state.particle_q.double().cpu().numpy()
    _close(x, o.x, 1e-4, "positions with a moving capsule")
    # the same run without triangle contact ends elsewhere
    solver2, state2 = _make(m, gpu, tri_contact=False)
    out2 = state2.clone()
    pose = m.body_q[0].astype(np.float64).copy()
    for k in range(6):
        pose[2] -= 0.0015
        state2.body_q = torch.as_tensor(pose[None, :], dtype=torch.float32, device=gpu)
        solver2.step(state2, out2, None, None, dt)
        state2, out2 = out2, state2
    x2 = state2.particle_q.double().cpu().numpy()
    print("triangle term changes positions by", np.abs(x2 - x).max())
    assert np.abs(x2 - x).max() > 1e-5
