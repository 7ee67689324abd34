"""GPU tests of the element kernel's edge branches, the FP64 barrier
distances, the device-error gating of a step and the rigid-body shape poses,
each against the FP64 oracle or a closed form.

Tolerances: per-tet quantities of well-conditioned tets within 1e-3 of the
tet's own scale (FP32 SVD and GMW81 on strongly deformed / inverted tets);
barrier gradients near d0 within 1e-5 relative (FP64 distances, FP32
storage); positions after full steps within 1e-4 relative.
"""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from mfem_amd import mesh, scene
from oracle import Kernels, Oracle

pytestmark = pytest.mark.gpu

K = Kernels(64)


def _make(model, device, **extra):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = {k: v for k, v in model.params.items() if k != "iterations"}
    kw["enable_refinement"] = False
    kw.update(extra)
    solver = RefinementSolver(model, model.params["iterations"], max_tets=model.tet_count, device=device, **kw)
    return solver, model_state(model, device)


def _unpack21(p):
    out = np.zeros((p.shape[0], 6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            out[:, i, j] = out[:, j, i] = p[:, k]
            k += 1
    return out


def _sync(o, solver, state):
    o.buf("x")[:] = state.particle_q.double().cpu().numpy().ravel()
    o.buf("v")[:] = state.particle_qd.double().cpu().numpy().ravel()
    o.buf("s")[:] = solver.buffers["s"].double().cpu().numpy().ravel()
    o.buf("lam")[:] = solver.buffers["lam"].double().cpu().numpy().ravel()


def test_inverted_and_large_strain_tets_match_oracle(gpu):
    """Neo-Hookean beam with large random displacements: inverted tets (the
    rotation-variant SVD's sign flip) and indefinite stretch Hessians (the
    GMW81 corrected branch) are both present; G, H_s^-1, g_s and psi of every
    well-conditioned tet match the oracle."""
    m = scene.beam_scene(cells=(6, 3, 3), energy="neohookean")
    solver, state = _make(m, gpu, fuse_local_solve=False)  # (stores the element records inspected below)
    rng = np.random.default_rng(11)
    h = 0.2 / 3
    q = m.particle_q + 0.45 * h * rng.normal(size=m.particle_q.shape)
    state.particle_q.copy_(torch.as_tensor(q, dtype=torch.float32))
    o = Oracle(m, 64)
    # a stretch far from the rest value (s from the deformed shape, then
    # scaled): both energy branches see large strains
    o.buf("x")[:] = q.ravel()
    o.lib.orc_init_stretch(o.ctx)
    s = o.buf("s").reshape(-1, 6) * np.where(rng.random((m.tet_count, 1)) < 0.5, 0.7, 1.25)
    o.buf("s")[:] = s.ravel()
    solver.buffers["s"][: m.tet_count] = torch.as_tensor(s, dtype=torch.float32)
    solver.bind(state)
    dt = 1.0 / 60.0
    solver.begin_step(dt)
    o.begin_step(dt)
    solver.update_system(dt)
    o.update_system(dt)
    torch.cuda.synchronize()
    nt = m.tet_count
    # preconditions of the test: inverted tets and indefinite raw Hessians
    F = []
    for e, t in enumerate(m.tet_indices):
        Ds = np.stack([q[t[1]] - q[t[0]], q[t[2]] - q[t[0]], q[t[3]] - q[t[0]]], 1)
        F.append(Ds @ m.tet_poses[e])
    F = np.array(F)
    detF = np.linalg.det(F)
    sv = np.linalg.svd(F, compute_uv=False)
    n_inv = int((detF < 0).sum())
    vol = 1.0 / (6.0 * np.linalg.det(m.tet_poses))
    mu, lam = m.tet_materials[0, 0], m.tet_materials[0, 1]
    indef = sum(np.linalg.eigvalsh(K.neohookean(s[e], mu, lam, vol[e])[4]).min() < 0 for e in range(nt))
    print(f"inverted tets {n_inv}, indefinite stretch Hessians {indef} of {nt}")
    assert n_inv >= 5 and indef >= 20
    # well-conditioned: distinct singular values (the rotation gradient's
    # 2 / (s_a + s_b) and the SVD are then well defined in FP32)
    gap = np.minimum(np.abs(sv[:, 0] - sv[:, 1]), np.abs(sv[:, 1] - sv[:, 2])) / sv[:, 0]
    ok = (gap > 2e-2) & (sv[:, 2] > 0.05 * sv[:, 0])
    ok_inv = ok & (detF < 0)
    print("well-conditioned tets", int(ok.sum()), "of which inverted", int(ok_inv.sum()))
    assert ok_inv.sum() >= 3
    B = {k: v.double().cpu().numpy() for k, v in solver.buffers.items() if k in ("G", "Hsinv", "gs", "psi")}
    Gg, Go = B["G"][:nt].reshape(nt, 72), o.buf("G").reshape(nt, 72)
    Hg, Ho = _unpack21(B["Hsinv"][:nt]).reshape(nt, 36), o.buf("Hsinv").reshape(nt, 36)
    gg, go = B["gs"][:nt], o.buf("gs").reshape(nt, 6)
    pg, po = B["psi"][:nt], o.buf("psi")
    for name, a, b in (("G", Gg, Go), ("H_s^-1", Hg, Ho), ("g_s", gg, go)):
        err = np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-30)
        print(f"{name}: worst per-tet rel err {err[ok].max():.3e} (inverted {err[ok_inv].max():.3e})")
        assert err[ok].max() <= 1e-3
    perr = np.abs(pg - po) / np.maximum(vol * (mu + lam), 1e-30)
    assert perr[ok].max() <= 1e-5
    # the corrected branch really ran on the GPU: H_s^-1 of an indefinite
    # raw Hessian is positive definite
    e = next(e for e in np.nonzero(ok)[0] if np.linalg.eigvalsh(K.neohookean(s[e], mu, lam, vol[e])[4]).min() < 0)
    assert np.linalg.eigvalsh(Hg[e].reshape(6, 6)).min() > 0


def test_barrier_in_fp64_near_d0(gpu):
    """Vertices between d0 and 2 d0 from a tilted, offset plane: the barrier
    gradient, Hessian block and energy match the FP64 oracle to 1e-5 (FP64
    distances); in FP32 the distance alone would be off by ~1e-3 relative."""
    rot = Rotation.from_euler("xy", [0.2, -0.15])
    nrm = rot.apply([0.0, 0.0, 1.0])
    plane = scene.Shape(scene.SHAPE_PLANE, pos=np.array([0.41, -0.37, 0.29]), quat=rot.as_quat(), mu=0.0)
    p, t = mesh.box_tet_mesh((0.0, 0.0, 0.0), (0.1, 0.08, 0.06), (3, 2, 2))
    # rest the box on the plane: its lowest face at d in [d0, 2 d0]
    dhat, d0 = 1e-3, 1e-4
    R = rot.as_matrix()
    x = p @ R.T + plane.pos
    rng = np.random.default_rng(2)
    base = np.abs(p[:, 2]) < 1e-12
    lift = np.where(base, rng.uniform(1.05, 1.95, size=len(p)) * d0, p[:, 2] + 2 * d0)
    x = x + (lift - p[:, 2])[:, None] * nrm[None, :]
    m = scene.build_model(x, t, np.tile([2e4, 8e4, 0.0], (len(t), 1)), 1000.0, shapes=[plane])
    m.params.update(energy="neohookean", iterations=2, cg_max_iterations=2000, cg_tolerance=1e-6, line_search=True,
                    contact_d0=d0, contact_d1=dhat, contact_stiffness=scene.auto_kappa(m, dhat), friction_mu=[0.0])
    solver, state = _make(m, gpu)
    o = Oracle(m, 64)
    d_true = (x - plane.pos) @ nrm
    assert np.all(d_true[base] > d0) and np.all(d_true[base] < 2 * d0)
    # the oracle sees the GPU's inputs exactly: positions and pose rounded to FP32
    o.buf("x")[:] = x.astype(np.float32).astype(np.float64).ravel()
    o.buf("shape_xf")[:] = np.concatenate([plane.pos, plane.quat]).astype(np.float32).astype(np.float64)
    solver.bind(state)
    dt = 0.01
    solver.begin_step(dt)
    o.begin_step(dt)
    solver.update_system(dt)
    o.update_system(dt)
    torch.cuda.synchronize()
    dist = solver.buffers["dist"][: m.particle_count].double().cpu().numpy()
    derr = np.abs(dist[base] - o.buf("dist")[base]).max() / d0
    print("distance rel err (of d0)", derr)
    assert derr < 1e-6
    gB = solver.buffers["gB"][: m.particle_count].double().cpu().numpy()
    Hv = solver.buffers["Hv"][: m.particle_count].double().cpu().numpy()
    eB = solver.buffers["eB"][: m.particle_count].double().cpu().numpy()
    go = o.buf("gB").reshape(-1, 3)
    ho = (o.buf("HB") + o.buf("HD")).reshape(-1, 9)
    eo = o.buf("eB")
    assert np.abs(go[base]).max() > 0
    for name, a, b in (("grad", gB[base], go[base]), ("hess", Hv[base], ho[base]), ("energy", eB[base], eo[base])):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"barrier {name}: rel err {err:.3e}")
        assert err <= 1e-5
    # FP32 distance for comparison (what the FP32 formula would give)
    x32 = x.astype(np.float32)
    Rf, pf = R.astype(np.float32), plane.pos.astype(np.float32)
    d32 = ((x32 - pf) @ Rf)[:, 2]
    print("FP32 distance rel err (of d0)", np.abs(d32[base] - o.buf("dist")[base]).max() / d0)
    # a whole step stays within 1e-4 of the oracle
    solver2, state2 = _make(m, gpu)
    out = state2.clone()
    solver2.step(state2, out, None, None, dt)
    o2 = Oracle(m, 64)
    o2.step(dt)
    xg = out.particle_q[: m.particle_count].double().cpu().numpy()
    err = np.abs(xg - o2.x).max() / np.abs(o2.x).max()
    print("step position rel err", err)
    assert err <= 1e-4 and np.all(solver2.last_stats()[:, 1] <= 1e-6)


def test_device_error_bits_leave_the_state_untouched(gpu):
    """With the error bit set (as a grid-barrier timeout leaves it), every
    kernel of a step returns at once: x, v, s and lambda are unchanged, the
    bit stays visible, and the next step raises until reset_status."""
    from mfem_amd import _ffi

    m = scene.drop_box_scene()
    solver, state = _make(m, gpu)
    out = state.clone()
    solver.step(state, out, None, None, 0.01)  # a normal step first: lambda nonzero
    torch.cuda.synchronize()
    state, out = out, state
    before = {k: solver.buffers[k].clone() for k in ("s", "lam")}
    x0, v0 = state.particle_q.clone(), state.particle_qd.clone()
    assert float(before["lam"].abs().max()) > 0
    solver.buffers["status"][0] = 1  # the bit a timed-out barrier sets
    solver.step(state, state, None, None, 0.01)  # in place: state_out is state_in
    torch.cuda.synchronize()
    assert torch.equal(state.particle_q, x0) and torch.equal(state.particle_qd, v0)
    for k, t in before.items():
        assert torch.equal(solver.buffers[k], t), k
    assert solver.status() == 1
    with pytest.raises(_ffi.MfemError):
        solver.step(state, out, None, None, 0.01)
    solver.reset_status()
    solver.step(state, out, None, None, 0.01)
    assert solver.status() == 0
    assert np.all(solver.last_stats()[:, 1] <= 1e-6)
    assert not torch.equal(out.particle_q, x0)


def test_body_attached_capsule_and_static_plane(gpu):
    """A capsule attached to rigid body 0 with a nonzero local offset and
    rotation, plus a static plane: the solver's world poses are
    body_q * local for the capsule and the plane's own pose, and steps with
    a moving, turning body match the FP64 oracle posed the same way (within
    3x the distance of the oracle's own FP32 restatement from it: the pinned
    slab under a pressing capsule amplifies rounding)."""

    def posed_model():
        m = scene.capsule_poke_scene()
        cap = m.shapes[0]
        cap.pos = np.array([-0.004, -0.02, 0.005])  # 4 mm above the body's origin in the world
        cap.quat = Rotation.from_euler("z", 25, degrees=True).as_quat()
        plane = scene.Shape(scene.SHAPE_PLANE, pos=np.array([0.0, 0.0, -0.05]), mu=0.0)
        m.shapes = [cap, plane]
        m.params["friction_mu"] = [0.2, 0.0]
        return m

    def poses(m):
        body = m.body_q[0].astype(np.float64).copy()
        for k in range(4):
            body[:3] += [0.002, -0.001, -0.0015]
            body[3:] = (Rotation.from_quat(body[3:]) * Rotation.from_euler("x", 1.0, degrees=True)).as_quat()
            yield body.copy()

    m = posed_model()
    solver, state = _make(m, gpu)
    o, o32 = Oracle(m, 64), Oracle(m, 32)
    for body in poses(m):
        state.body_q = torch.as_tensor(body[None, :], dtype=torch.float32, device=gpu)
        for orc in (o, o32):
            orc.set_body_q(body[None, :])
            orc.step(0.01)
        out = state.clone()
        solver.step(state, out, None, None, 0.01)
        want = scene.shape_world_poses(m, body[None, :])
        got = solver.buffers["shape_xf"].double().cpu().numpy().reshape(-1, 7)
        np.testing.assert_allclose(got[:, :3], want[:, :3], atol=1e-6)
        # quaternions up to sign
        assert np.all(np.abs(np.abs(np.sum(got[:, 3:] * want[:, 3:], axis=1)) - 1.0) < 1e-6)
        np.testing.assert_allclose(got[1], [0, 0, -0.05, 0, 0, 0, 1], atol=1e-7)
        state = out
        assert np.all(solver.last_stats()[:, 1] <= 1e-6) and solver.status() == 0
    assert float(solver.buffers["tri_dist"].min()) < m.params["contact_d1"]  # the capsule presses
    x = state.particle_q.double().cpu().numpy()
    scale = np.abs(o.x).max()
    err, err32 = np.abs(x - o.x).max() / scale, np.abs(o32.x - o.x).max() / scale
    print("positions vs FP64 oracle", err, "FP32 oracle vs FP64 oracle", err32)
    assert err <= max(1e-4, 3.0 * err32)
    with pytest.raises(ValueError):
        solver._set_body_q(torch.zeros(0, 7, device=gpu))  # fewer poses than the shapes' bodies
