"""The reference's tet-mesh file format (MFEMRefinementModel,
src/mfem/refinement/models/refinement_model.py:13-143) on the reference's
own coarse octopus mesh (models/octopus_initial_coarse.npz, committed as a
fixture; read with allow_pickle=False).

CPU: the file's sizes and parameters, save / load / from_model round trips,
scale and translation, and load_sim_model's padded particle buffer (rotated
90 degrees about x, inactive zero rows, masses summing to density x volume,
the surface of the mesh, mu the given scalar).  GPU: the loaded mesh through
the solver (particle buffer padded to the file's max_particles, as the
reference requires) against the FP64 oracle.
"""
import os

import numpy as np
import pytest

from mfem_amd import mesh, scene
from mfem_amd.refinement_model import MFEMRefinementModel

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "octopus_initial_coarse.npz")


def test_file_sizes_and_parameters():
    rm = MFEMRefinementModel.load(FIXTURE, 1.0, 0.0)
    assert rm.particles.shape == (600, 3) and rm.tet_indices.shape == (2025, 4)
    assert rm.tet_materials.shape == (2025, 3)
    assert (rm.max_particles, rm.max_tets, rm.density, rm.max_tris) == (1800, 8100, 1.0, 2436)
    assert rm.resolve_max_particles(None) == 1800 and rm.resolve_max_particles(2.0) == 1200
    assert rm.resolve_max_tets(None) == 8100 and rm.resolve_max_tets(1.5) == 3038


def test_save_load_round_trip_and_transform(tmp_path):
    rm = MFEMRefinementModel.load(FIXTURE, 1.0, 0.0)
    f = str(tmp_path / "copy.npz")
    rm.save(f)
    back = MFEMRefinementModel.load(f, 1.0, 0.0)
    for k in ("particles", "tet_indices", "tet_materials", "parameters"):
        np.testing.assert_array_equal(getattr(back, k), getattr(rm, k))
    moved = MFEMRefinementModel.load(FIXTURE, 2.0, np.array([0.1, -0.2, 0.3]))
    np.testing.assert_allclose(moved.particles, rm.particles * 2.0 + np.array([0.1, -0.2, 0.3]), atol=1e-6)
    # an older file without max_tris
    old = MFEMRefinementModel(rm.particles, rm.tet_indices, rm.tet_materials, rm.parameters[:3])
    assert old.max_tris == 0


def test_from_model_round_trip():
    m = scene.drop_box_scene()
    rm = MFEMRefinementModel.from_model(m, max_particles=3 * m.particle_count, max_tets=4 * m.tet_count, max_tris=77)
    assert (rm.max_particles, rm.max_tets, rm.density, rm.max_tris) == (3 * m.particle_count, 4 * m.tet_count, 1.0, 77)
    np.testing.assert_array_equal(rm.tet_indices, m.tet_indices)
    np.testing.assert_allclose(rm.particles, m.particle_q, atol=1e-7)


def test_load_sim_model_pads_rotates_and_lumps():
    rm = MFEMRefinementModel.load(FIXTURE, 1.0, 0.0)
    m = rm.load_sim_model(mu=6.0)
    n = rm.particles.shape[0]
    assert m.particle_count == rm.max_particles == 1800
    # rotation by 90 degrees about x: (x, y, z) -> (x, -z, y)
    p = rm.particles.astype(np.float64)
    np.testing.assert_allclose(m.particle_q[:n], np.stack([p[:, 0], -p[:, 2], p[:, 1]], axis=1), atol=1e-7)
    assert np.all(m.particle_q[n:] == 0) and np.all(m.particle_mass[n:] == 0) and np.all(m.particle_inv_mass[n:] == 0)
    # lumped masses: density x the total rest volume, every mesh vertex massive
    _, vol = mesh.rest_pose(m.particle_q[:n], m.tet_indices)
    assert m.particle_mass.sum() == pytest.approx(rm.density * vol.sum(), rel=1e-12)
    assert np.all(m.particle_mass[:n] > 0) and np.all(m.particle_inv_mass[:n] > 0)
    assert np.all(m.tet_materials[:, 0] == 6.0)
    np.testing.assert_array_equal(m.tet_materials[:, 1:], rm.tet_materials[:, 1:])
    # the surface: a closed triangle mesh over the mesh's own vertices
    t = m.tri_indices
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert t.max() < n and np.all(cnt == 2) and len(t) <= rm.max_tris
    assert m.params["refine_density"] == rm.density


def test_solver_requires_the_padded_buffer():
    from mfem_amd.solver import RefinementSolver

    m = scene.drop_box_scene()
    with pytest.raises(ValueError, match="padded"):
        RefinementSolver(m, 2, max_tets=m.tet_count, max_particles=2 * m.particle_count, device="cpu")


@pytest.mark.gpu
def test_reference_mesh_steps_match_fp64_oracle(gpu):
    from mfem_amd.solver import RefinementSolver, model_state
    from oracle import Oracle

    rm = MFEMRefinementModel.load(FIXTURE, 1.0, 0.0)
    m = rm.load_sim_model(mu=6.0)
    n = rm.particles.shape[0]
    # a ground plane just under the rotated mesh
    zmin = float(m.particle_q[:n, 2].min())
    m.shapes = [scene.Shape(scene.SHAPE_PLANE, pos=(0.0, 0.0, zmin - 5.0e-4), mu=0.3)]
    dhat = 1.0e-3
    params = dict(energy="arap", cg_max_iterations=2000, cg_tolerance=1e-6, line_search=True,
                  line_search_max_iterations=20, contact_d0=0.1 * dhat, contact_d1=dhat,
                  contact_stiffness=10.0 * float(m.particle_mass[:n].mean()) * 9.81 /
                  abs(float(scene.barrier_slope(0.5 * dhat, dhat))), friction_mu=[0.3],
                  friction_eps_v=1e-3)
    m.params.update(params)
    kw = {k: v for k, v in m.params.items()}
    solver = RefinementSolver(m, 8, max_tets=rm.max_tets, max_particles=rm.max_particles, max_tris=rm.max_tris,
                              device=gpu, enable_refinement=False, **kw)
    assert solver.active_particle_count == n
    state = model_state(m, gpu)
    out = state.clone()
    o = Oracle(m, 64, params=dict(iterations=8))
    for k in range(4):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
        o.step(0.01)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st[:, :2]
        x = state.particle_q[:n].double().cpu().numpy()
        err = np.abs(x - o.x).max() / np.abs(o.x).max()
        print("step", k, "cg", st[:, 0].astype(int).tolist(), "rel err", err)
    assert err <= 1e-4
    assert np.all(state.particle_q[n:].cpu().numpy() == 0)  # padding rows untouched
