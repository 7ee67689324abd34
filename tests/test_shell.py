"""GPU parity of the membrane shell (shell_mu > 0; src/mfem/refinement/
shell_energy.py and its solver hooks, refinement/solver.py:291-302, 492-493,
638-639, 789-797, 864-867, 910-911) against the FP64 oracle, through the C ABI.

Tolerances as in test_gpu_parity.py: stage quantities within 2e-4 of the
oracle's scale, merit within 1e-4 relative, positions after full steps within
1e-4 relative, every PCG solve at rel-res <= 1e-6 (FP64, on the GPU).
"""
import numpy as np
import pytest
import torch

from mfem_amd import scene
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _scene(name):
    """(model, dt, shell_mu): shell_mu = the bulk mu times a 1 cm thickness,
    as the reference's driver forms it (sim_octopus.py:202-209, 481-483)."""
    if name == "beam":
        m, dt = scene.beam_scene(cells=(10, 2, 2)), 1.0 / 60.0
    elif name == "capsule_poke":
        m, dt = scene.capsule_poke_scene(), 0.01
    else:
        m, dt = scene.drop_box_scene(), 0.01
    return m, dt, float(m.tet_materials[0, 0]) * 0.01


def _make(m, device, **extra):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    kw["enable_refinement"] = False
    kw.update(extra)
    return RefinementSolver(m, m.params["iterations"], max_tets=m.tet_count, device=device, **kw), model_state(m, device)


def _close(a, b, rel, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max()
    print(f"{what}: max abs err {err:.3e}, scale {scale:.3e}, rel {err / scale:.3e}", flush=True)
    assert err <= rel * scale, f"{what}: rel err {err / scale:.3e} > {rel}"


@pytest.mark.parametrize("name", ["beam", "drop_box"])
def test_shell_stages_match_oracle(gpu, name):
    m, dt, mu_s = _scene(name)
    solver, state = _make(m, gpu, shell_mu=mu_s)
    rng = np.random.default_rng(4)
    q = m.particle_q + 0.004 * rng.normal(size=m.particle_q.shape)
    q[:, 2] = np.maximum(q[:, 2], 2e-4)
    state.particle_q.copy_(torch.as_tensor(q, dtype=torch.float32))
    o = Oracle(m, 64, params={"shell_mu": mu_s})
    o.buf("x")[:] = state.particle_q.double().cpu().numpy().ravel()
    o.buf("v")[:] = state.particle_qd.double().cpu().numpy().ravel()
    o.buf("s")[:] = solver.buffers["s"].double().cpu().numpy().ravel()
    solver.bind(state)
    for f, g in ((solver.begin_step, o.begin_step), (solver.update_system, o.update_system),
                 (solver.assemble, o.assemble)):
        f(dt)
        g(dt)
    torch.cuda.synchronize()
    B = solver.buffers
    nf = m.tri_indices.shape[0]
    _close(B["tri_area"][:nf].double().cpu().numpy(), o.buf("tarea"), 1e-6, "rest areas")
    _close(B["tri_W"][:nf].double().cpu().numpy().ravel(), o.buf("tW"), 1e-5, "rest frames W")
    _close(B["gS"].double().cpu().numpy().ravel(), o.buf("gS"), 2e-4, "shell gradient")
    _close(solver.system_csr().toarray(), o.csr().toarray(), 2e-4, "H with the shell")
    _close(B["g"].double().cpu().numpy().ravel(), o.buf("g"), 2e-4, "g with the shell")
    f_gpu = solver.merit(dt, 0.0)
    f_orc = o.merit(o.x, o.buf("s"), dt)
    print("merit", f_gpu, f_orc)
    assert abs(f_gpu - f_orc) <= 1e-4 * max(abs(f_orc), 1e-12) + 1e-7


@pytest.mark.parametrize("name,steps", [("beam", 5), ("drop_box", 6), ("capsule_poke", 4)])
def test_full_steps_with_shell_match_fp64_oracle(gpu, name, steps):
    m, dt, mu_s = _scene(name)
    solver, state = _make(m, gpu, shell_mu=mu_s)
    o = Oracle(m, 64, params={"shell_mu": mu_s})
    o_off = Oracle(m, 64)
    out = state.clone()
    for k in range(steps):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        o.step(dt)
        o_off.step(dt)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st
        assert solver.status() == 0
    x = state.particle_q.double().cpu().numpy()
    _close(x, o.x, 1e-4, f"{name} positions after {steps} steps with the shell")
    # the shell changes the motion by several times the tolerance
    shift = np.abs(o.x - o_off.x).max() / np.abs(o.x).max()
    print("shell effect (rel)", shift)
    assert shift > 3e-4


def test_shell_after_refinement_matches_oracle(gpu):
    """Refinement grows the triangulation: the rest frames and the shell's
    Hessian are rebuilt from the interpolated rest positions, and a step on
    the refined mesh matches the FP64 oracle built from that mesh and rest."""
    from mfem_amd.scene import Model
    from mfem_amd.solver import RefinementSolver, model_state

    m = scene.beam_scene(cells=(8, 2, 2))
    mu_s = float(m.tet_materials[0, 0]) * 0.01
    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    V = 4 * m.particle_count
    solver = RefinementSolver(scene.pad_particles(m, V), m.params["iterations"], max_tets=4 * m.tet_count,
                              max_particles=V, device=gpu, enable_refinement=True, refine_every_n_steps=3,
                              min_refine_score=1e-3, refine_tet_score_weight=1.0, shell_mu=mu_s, **kw)
    state = model_state(m, gpu, capacity=V)
    out = state.clone()
    dt = 1.0 / 60.0
    # step until a refining pass has split surface triangles (passes fire at
    # steps 0, 3, 6, ...); the step after it is not a refining pass
    for k in range(10):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        assert np.all(solver.last_stats()[:, 1] <= 1e-6)
        if k % 3 == 0 and solver.active_tri_count > m.tri_indices.shape[0]:
            break
    nv, nt, nf = solver.active_particle_count, solver.active_tet_count, solver.active_tri_count
    print("refined: verts", nv, "tets", nt, "tris", nf, "after", k + 1, "steps")
    assert nf > m.tri_indices.shape[0], "no triangle was split"
    b = solver.buffers
    mr = Model(particle_q=state.particle_q[:nv].double().cpu().numpy(),
               particle_qd=state.particle_qd[:nv].double().cpu().numpy(),
               particle_mass=b["mass"][:nv].double().cpu().numpy(),
               particle_inv_mass=b["inv_mass"][:nv].double().cpu().numpy(),
               tet_indices=b["tets"][:nt].cpu().numpy(),
               tet_poses=b["dminv"][:nt].double().cpu().numpy().reshape(nt, 3, 3),
               tet_materials=b["materials"][:nt].double().cpu().numpy(),
               tri_indices=b["tris"][:nf].cpu().numpy(), shapes=[], gravity=np.asarray(m.gravity),
               params=dict(m.params))
    o = Oracle(mr, 64, params={"shell_mu": mu_s}, rest=b["rest"][:nv].double().cpu().numpy())
    _close(b["tri_area"][:nf].double().cpu().numpy(), o.buf("tarea"), 1e-6, "rest areas after refinement")
    o.buf("s")[:] = b["s"][:nt].double().cpu().numpy().ravel()
    o.buf("lam")[:] = b["lam"][:nt].double().cpu().numpy().ravel()
    pins = np.nonzero(mr.particle_inv_mass == 0)[0]
    o.buf("pin_pos")[:] = b["pin_pos"][:nv].double().cpu().numpy()[pins].ravel()
    solver.step(state, out, None, None, dt)  # not a refining pass
    assert not solver.refine_log[-1]["fired"]
    state, out = out, state
    o.step(dt)
    _close(state.particle_q[:nv].double().cpu().numpy(), o.x, 1e-4, "positions after refinement with the shell")


def test_batch_with_shell_matches_oracle(gpu):
    """Two bodies in one batched solve with the shell on: each body against
    its own FP64 oracle run."""
    from mfem_amd.batch import BatchSolver

    models = [scene.drop_box_scene(),
              scene.drop_box_scene(cells=(6, 3, 3), size=(0.3, 0.15, 0.15), E=1.0e5, energy="arap")]
    mu_s = 50.0
    params = {k: v for k, v in models[0].params.items() if k not in ("iterations", "energy")}
    it = models[0].params["iterations"]
    bs = BatchSolver(models, it, device=gpu, shell_mu=mu_s, **params)
    a, b = bs.make_state(), bs.make_state()
    dt = 0.01
    for k in range(4):
        bs.step(a, b, dt)
        a, b = b, a
        assert bs.status() == 0
        for i in range(bs.n_bodies):
            assert np.all(bs.body_stats(i)[:, 1] <= 1e-6)
    for i, m in enumerate(models):
        o = Oracle(m, 64, params=dict(params, energy=m.params["energy"], shell_mu=mu_s))
        for k in range(4):
            o.step(dt)
        xb = a.particle_q[bs.body_slice(i)].double().cpu().numpy()
        _close(xb, o.x, 1e-4, f"batched body {i} with the shell")
