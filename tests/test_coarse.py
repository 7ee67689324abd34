"""The workgroup-local rigid-mode correction on top of block Jacobi
(MFEM_PCG_COARSE) on the GPU: every solve still meets rel-res 1e-6 in FP64,
full steps match the FP64 oracle (whose CG uses block Jacobi: the
preconditioner changes the iterates, not the solution) within 1e-4, a
per-vertex weight tensor equal to the uniform weight gives the same bits, and
on the 204k-tet octopus the correction takes fewer CG iterations than block
Jacobi alone while ending at the same positions."""
import numpy as np
import pytest
import torch

from mfem_amd import scene
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _make(model, device, **extra):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = {k: v for k, v in model.params.items() if k != "iterations"}
    kw["enable_refinement"] = False
    kw.update(extra)
    solver = RefinementSolver(model, model.params["iterations"], max_tets=model.tet_count, device=device, **kw)
    return solver, model_state(model, device)


@pytest.mark.parametrize("name,steps", [("beam_nh", 4), ("drop_box", 6), ("capsule_poke", 4)])
@pytest.mark.parametrize("omega", [0.0, 0.6])
def test_coarse_steps_match_fp64_oracle(gpu, name, steps, omega):
    m = {"beam_nh": lambda: scene.beam_scene(energy="neohookean"), "drop_box": scene.drop_box_scene,
         "capsule_poke": scene.capsule_poke_scene}[name]()
    dt = 1.0 / 60.0 if name.startswith("beam") else 0.01
    solver, state = _make(m, gpu, pcg_coarse=True, pcg_coarse_omega=omega)
    assert any(v["mode"] == "pipelined_coarse" and v["applicable"] for v in solver.pcg_variants())
    o = Oracle(m, 64)
    out = state.clone()
    iters = []
    for k in range(steps):
        solver.step(state, out, None, None, dt)
        state, out = out, state
        o.step(dt)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st[:, :2]
        assert solver.status() == 0
        iters.append(int(st[:, 0].sum()))
    x = state.particle_q.double().cpu().numpy()
    err = np.abs(x - o.x).max() / np.abs(o.x).max()
    print(name, "omega", omega, "cg iters per step", iters, "rel err", err)
    assert err <= 1e-4


def test_coarse_weight_tensor_equals_uniform_weight(gpu):
    m = scene.drop_box_scene()
    runs = []
    for use_tensor in (False, True):
        solver, state = _make(m, gpu, pcg_coarse=True, pcg_coarse_omega=0.5)
        if use_tensor:
            w = torch.full((solver.max_particles,), 0.5, dtype=torch.float32, device=gpu)
            solver.set_coarse_correction(True, omega=0.0, weights=w)
        out = state.clone()
        for _ in range(2):
            solver.step(state, out, None, None, 0.01)
            state, out = out, state
        runs.append((state.particle_q.clone(), solver.last_stats()[:, 0].copy()))
    assert torch.equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_coarse_cuts_octopus_cg_iterations(gpu):
    m = scene.octopus_scene()
    res = {}
    for label, kw in (("block_jacobi", dict(pcg_pipelined=True)), ("coarse", dict(pcg_coarse=True)),
                      ("coarse_smoothed", dict(pcg_coarse=True, pcg_coarse_omega=0.6))):
        solver, state = _make(m, gpu, **kw)
        out = state.clone()
        its = []
        for k in range(2):
            solver.step(state, out, None, None, 0.01)
            state, out = out, state
            st = solver.last_stats()
            assert np.all(st[:, 1] <= 1e-6) and solver.status() == 0, (label, st[:, :2])
            its.append(int(st[:, 0].sum()))
        res[label] = (its, state.particle_q.double().cpu().numpy())
        print(label, "cg iters per step", its, flush=True)
    xb = res["block_jacobi"][1]
    for label in ("coarse", "coarse_smoothed"):
        assert sum(res[label][0]) < 0.8 * sum(res["block_jacobi"][0]), label
        err = np.abs(res[label][1] - xb).max() / np.abs(xb).max()
        print(label, "positions vs block Jacobi", err)
        assert err <= 1e-4


def _bf16(bits: torch.Tensor) -> np.ndarray:
    return (bits.to(torch.int32) << 16).view(torch.float32).cpu().numpy() if bits.dtype == torch.int32 else \
        bits.view(torch.bfloat16).float().cpu().numpy()


def test_coarse_features_match_restatement(gpu):
    """The coarse-weight network's features against a numpy restatement from
    the assembled system (BF16 rounding: 1/128 relative)."""
    m = scene.drop_box_scene(cells=(10, 6, 6))
    solver, state = _make(m, gpu, pcg_coarse=True)
    dt = 0.01
    solver.bind(state)
    solver.begin_step(dt)
    solver.update_system(dt)
    solver.assemble(dt)
    solver.coarse_features(dt)
    torch.cuda.synchronize()
    n = m.particle_count
    F = _bf16(solver.buffers["cw_x"][:n])
    H = solver.system_csr().tobsr(blocksize=(3, 3))
    x = state.particle_q[:n].double().cpu().numpy()
    Hv = solver.buffers["Hv"][:n].double().cpu().numpy().reshape(n, 3, 3)
    dist = solver.buffers["dist"][:n].double().cpu().numpy()
    im = solver.buffers["inv_mass"][:n].double().cpu().numpy()
    md = np.where(im == 0, 1.0 + 10.0, 1.0 / np.where(im == 0, 1.0, im)) / dt ** 2
    agg = np.arange(n) // 256
    ref = np.zeros((n, 16))
    for a in range(agg.max() + 1):
        idx = np.nonzero(agg == a)[0]
        c = x[idx].mean(axis=0)
        rad = np.sqrt(((x[idx] - c) ** 2).sum() / len(idx))
        for i in idx:
            cols = H.indices[H.indptr[i]:H.indptr[i + 1]]
            blks = H.data[H.indptr[i]:H.indptr[i + 1]]
            nb = np.sqrt((blks ** 2).sum(axis=(1, 2)))
            D = blks[cols == i][0]
            off = (cols != i) & (nb > 0)
            out = off & ((cols // 256) != (i // 256))
            trD = np.trace(D)
            ref[i, :11] = [1.0, nb[out].sum() / nb[off].sum(), 3 * md[i] / trD, np.trace(Hv[i]) / trD,
                           np.linalg.norm(x[i] - c) / rad, nb[off].sum() / np.linalg.norm(D), float(im[i] == 0),
                           0.25 * np.log10(trD / (3 * md[i])), (off.sum() + 1) / 16, out.sum() / 16,
                           float(dist[i] < m.params["contact_d1"])]
    assert ref[:, 3].max() > 0 and ref[:, 10].max() == 1.0  # contact present
    np.testing.assert_allclose(F, ref, rtol=1.0 / 64, atol=1e-3)


def test_coarse_net_weights_in_a_step(gpu):
    """With a coarse-weight network attached, a step's weights are
    clamp(network output 0, 0, 2) on its first system's features, and the
    step still meets rel-res 1e-6 and the oracle."""
    from mfem_amd.coarse import CoarseWeightNet
    from mfem_amd.learned import O_B3, O_W3

    m = scene.drop_box_scene()
    solver, state = _make(m, gpu)
    w = CoarseWeightNet(gpu).weights()
    w[O_W3:O_B3] = np.random.default_rng(0).normal(0, 0.5, O_B3 - O_W3)
    w[O_B3] = 0.7
    net = CoarseWeightNet(gpu, weights=w)
    solver.set_coarse_net(net)
    out = state.clone()
    solver.step(state, out, None, None, 0.01)
    torch.cuda.synchronize()
    n = m.particle_count
    y = net.forward(solver.buffers["cw_x"][:n])[:, 0].clamp(0.0, 2.0)
    got = solver.buffers["coarse_w_net"][:n]
    assert torch.equal(got, y)
    assert float(got.std()) > 0.01 and float(got.max()) <= 2.0
    assert np.all(solver.last_stats()[:, 1] <= 1e-6)
    o = Oracle(m, 64)
    o.step(0.01)
    err = np.abs(out.particle_q[:n].double().cpu().numpy() - o.x).max() / np.abs(o.x).max()
    assert err <= 1e-4
