"""GPU: the point-to-surface query of the surface losses (mfem_surface_knn /
mfem_surface_project through the C ABI) against the CPU restatement
(oracle/surface.py, scipy's cKDTree + Ericson), the losses on the reference's
own test cases (src/mfem/refinement/surface_loss.py; tests/test_surface_loss.py
of the reference), and the tool driver + recorder around real solver steps.

Bars: nearest vertex identical, squared distances within 1e-12 relative (FP64
on both sides), the chosen triangle identical except at ties between triangles
equally close (shared edges / vertices), where it is as close within 1e-12;
loss values as the reference's tests state them.
"""
import time

import numpy as np
import pytest
import torch

from mfem_amd import scene, surface_loss
from mfem_amd.driver import MeshRecorder, ToolDriver, ToolTrajectory, capsule_signed_distance, simulate_frame
from oracle import surface as osurf
from test_driver_host import _FlatOverlay, _tool

pytestmark = pytest.mark.gpu


def _surface_case(seed, cells=(10, 8, 6), noise=0.01):
    from mfem_amd import mesh
    p, t = mesh.box_tet_mesh((0, 0, 0), (0.5, 0.4, 0.3), cells)
    rng = np.random.default_rng(seed)
    verts = p + noise * rng.normal(size=p.shape)
    faces = surface_loss.surface_triangles_from_tets(t)
    surf = np.unique(faces)
    q = verts[surf] + 0.02 * rng.normal(size=(len(surf), 3))
    return verts, faces, q


def _same_face_or_tie(q, verts, faces, face, oface, osq):
    """The chosen triangles agree, except where two triangles are equally
    close (a closest point on a shared edge or vertex: the FP64 roundings of
    the two projections decide); there the GPU's triangle is as close as the
    oracle's within 1e-12 relative."""
    diff = np.nonzero(face != oface)[0]
    assert np.all((face[diff] >= 0) & (oface[diff] >= 0)), "a vertex fallback on one side only"
    tri = verts[faces[face[diff]]]
    cp = osurf.closest_point_on_triangles(q[diff], tri[:, 0], tri[:, 1], tri[:, 2])
    np.testing.assert_allclose(((cp - q[diff]) ** 2).sum(-1), osq[diff], rtol=1e-12)
    print(f"triangle ties: {len(diff)} of {len(q)}")


@pytest.mark.parametrize("k", [1, 4])
@pytest.mark.parametrize("masked", [False, True])
def test_query_matches_oracle(gpu, k, masked):
    verts, faces, q = _surface_case(3)
    rng = np.random.default_rng(5)
    valid = rng.random(len(verts)) > 0.2 if masked else None
    face_ok = valid[faces].all(axis=1) if masked else None
    sq, face, cp, near = surface_loss.point_mesh_query(q, verts, faces, valid=valid, face_ok=face_ok, k=k,
                                                       device=gpu)
    osq, oface, ocp, onear = osurf.point_mesh_query(q, verts, faces, valid=valid, face_ok=face_ok, k=k)
    print(f"k={k} masked={masked}: {len(q)} queries, vertex fallbacks {int((face < 0).sum())}")
    np.testing.assert_array_equal(near, onear)
    np.testing.assert_allclose(sq, osq, rtol=1e-12, atol=1e-300)
    _same_face_or_tie(q, verts, faces, face, oface, osq)
    same = face == oface
    np.testing.assert_allclose(cp[same], ocp[same], rtol=1e-12, atol=1e-15)
    # vertex-only form
    sqv, _, _, nv = surface_loss.point_mesh_query(q, verts, None, valid=valid, k=k, project=False, device=gpu)
    osqv, _, _, onv = osurf.point_mesh_query(q, verts, None, valid=valid, k=k, project=False)
    np.testing.assert_array_equal(nv, onv)
    np.testing.assert_allclose(sqv, osqv, rtol=1e-12)


def test_query_octopus_surface_timed(gpu):
    """The octopus body's surface (~4e4 vertices) against a perturbed copy of
    itself: identical to the oracle; the GPU query's time is printed."""
    m = scene.octopus_scene()
    faces = np.asarray(m.tri_indices)
    surf = np.unique(faces)
    rng = np.random.default_rng(1)
    verts = np.asarray(m.particle_q, np.float64)
    q = verts[surf] + 0.002 * rng.normal(size=(len(surf), 3))
    surface_loss.point_mesh_query(q[:64], verts, faces, device=gpu)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sq, face, cp, near = surface_loss.point_mesh_query(q, verts, faces, device=gpu)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    osq, oface, _, onear = osurf.point_mesh_query(q, verts, faces)
    t_cpu = time.perf_counter() - t0
    print(f"octopus surface: {len(q)} queries vs {len(verts)} vertices / {len(faces)} triangles: "
          f"GPU {t_gpu * 1e3:.1f} ms (incl. transfers), oracle {t_cpu * 1e3:.1f} ms")
    np.testing.assert_array_equal(near, onear)
    np.testing.assert_allclose(sq, osq, rtol=1e-12)
    _same_face_or_tie(q, verts, faces, face, oface, osq)


def test_reference_loss_cases(gpu):
    """The reference's tests/test_surface_loss.py cases on the GPU query."""
    ov = _FlatOverlay()
    tool = _tool((0.5, 0.5, 0.0), 0.1)
    sim = ov.verts.copy()
    sim[np.linalg.norm(sim[:, :2] - 0.5, axis=1) < 0.15, 2] = -0.02
    plain = surface_loss.TrackedSurfaceLoss(ov, symmetric=True, device=gpu)
    fair = surface_loss.TrackedSurfaceLoss(ov, symmetric=True, exclude_tool_margin=0.06, device=gpu)
    r0 = plain.evaluate(sim, frame=0, tool_signed_distance=tool, sim_faces=ov.faces)
    r1 = fair.evaluate(sim, frame=0, tool_signed_distance=tool, sim_faces=ov.faces)
    assert r0["mse"] > 0.0 and r0["mse_tracked_to_sim"] > 0.0
    assert r1["n_tracked_excluded"] > 0 and r1["n_sim"] < r1["n_sim_all"]
    assert r1["mse"] == pytest.approx(0.0, abs=1e-12) and r1["mse_tracked_to_sim"] == pytest.approx(0.0, abs=1e-12)
    coarse = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=float)
    cf = np.array([[0, 1, 3], [0, 3, 2]])
    assert plain.evaluate(coarse, frame=0, sim_faces=cf)["mse_tracked_to_sim"] == pytest.approx(0.0, abs=1e-12)
    assert plain.evaluate(coarse, frame=0)["mse_tracked_to_sim"] > 1e-3
    corr = surface_loss.TrackedCorrespondenceLoss(ov, device=gpu)
    info = corr.bind(ov.verts, frame=0)
    assert info["n_bound"] == len(ov.verts) and info["bind_rmse_mm"] == pytest.approx(0.0, abs=1e-9)
    r = corr.evaluate(ov.verts + [0.01, 0.0, 0.0], frame=0)
    assert r["rmse_mm"] == pytest.approx(10.0, abs=1e-9)
    rng = np.random.default_rng(0)
    pts = rng.uniform(0.1, 0.9, size=(200, 3))
    pts[:, 2] = 0.003
    rel = surface_loss.TrackedCorrespondenceLoss(ov, relative=True, device=gpu)
    assert rel.bind(pts, frame=0)["bind_rmse_mm"] == pytest.approx(3.0, abs=1e-6)
    assert np.allclose(rel.bary.sum(axis=1), 1.0)
    assert rel.evaluate(pts, frame=0)["rmse_mm"] == pytest.approx(0.0, abs=1e-9)


def test_tool_driver_and_recorder_around_the_solver(gpu, tmp_path):
    """A plunging tool path drives the capsule of the poke scene through
    ToolDriver (device copies into body_q at every substep); the same poses
    passed by hand give bit-identical positions; the recording round-trips
    and scores zero against itself, and a translated copy scores the
    translation."""
    from mfem_amd.solver import RefinementSolver, model_state

    m = scene.capsule_poke_scene()
    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    kw["enable_refinement"] = False
    p0 = np.asarray(m.body_q[0], np.float64)
    R0 = np.eye(4)
    from scipy.spatial.transform import Rotation
    R0[:3, :3] = Rotation.from_quat(p0[3:]).as_matrix()
    n = 6
    T = np.tile(R0, (n, 1, 1))
    for i in range(n):
        T[i, :3, 3] = p0[:3] - [0.0, 0.0, 0.004 * i]  # plunge 4 mm per recorded frame
    traj = ToolTrajectory(T, np.arange(n) / 30.0, capsule_radius=0.01, capsule_half_height=0.08,
                          tip_along_z=0.09)
    assert traj.center_offset == pytest.approx(0.0)
    substeps, dt = 2, 1.0 / 60.0
    runs = []
    for manual in (False, True):
        solver = RefinementSolver(m, m.params["iterations"], max_tets=m.tet_count, device=gpu, **kw)
        s0 = model_state(m, gpu)
        s0.body_q = torch.as_tensor(m.body_q, dtype=torch.float32, device=gpu)
        s1 = s0.clone()
        drv = ToolDriver(traj, substeps, dt, device=gpu)
        rec = MeshRecorder(tmp_path / f"run{int(manual)}.npz", 30, capsule_radius=0.01, capsule_half_height=0.08)
        sim_time = 0.0
        for frame in range(4):
            if manual:
                for k in range(substeps):
                    s0.body_q[0].copy_(torch.as_tensor(traj.pose7(sim_time + k * dt), device=gpu))
                    solver.step(s0, s1, None, None, dt)
                    s0, s1 = s1, s0
            else:
                drv.refresh(sim_time)
                s0, s1 = simulate_frame(solver, s0, s1, dt, substeps, drv)
            assert solver.status() == 0 and np.all(solver.last_stats()[:, 1] <= 1e-6)
            sim_time += substeps * dt
            rec.capture(solver, s0, sim_time)
        rec.close()
        runs.append(s0.particle_q.cpu().numpy().copy())
    np.testing.assert_array_equal(runs[0], runs[1])
    dz = runs[0][:, 2].min() - np.asarray(m.particle_q)[:, 2].min()
    print("plunge moved the slab by", dz)
    # the recording scores zero against itself; a translated copy scores the shift
    ov = surface_loss.RecordingOverlay(str(tmp_path / "run0.npz"))
    loss = surface_loss.TrackedSurfaceLoss(ov, symmetric=True, device=gpu)
    curve = surface_loss.score_recording(str(tmp_path / "run0.npz"), loss,
                                         corr=surface_loss.TrackedCorrespondenceLoss(ov, device=gpu))
    assert np.allclose(curve["mse"], 0.0, atol=1e-14)
    assert np.allclose(curve["corr_rmse_mm"], 0.0, atol=1e-6)
    # tracked -> sim measures every vertex of the reference recording, its
    # interior ones included (as the reference's RecordingOverlay does), to
    # the sim's surface: the oracle's query gives the same mean
    from mfem_amd.driver import load_recording
    rec = load_recording(tmp_path / "run0.npz")
    pts = rec["positions_at"](3)
    tris = surface_loss.surface_triangles_from_tets(rec["tets_at"](3))
    surf = np.unique(tris)
    osq, _, _, _ = osurf.point_mesh_query(pts, pts[surf], np.searchsorted(surf, tris))
    assert curve["mse_tracked_to_sim"][3] == pytest.approx(osq.mean(), rel=1e-9)
    assert curve["mse_tracked_to_sim"][3] > 0.0
    z = dict(np.load(tmp_path / "run0.npz", allow_pickle=False))
    z["positions"] = z["positions"] + np.array([0.003, 0.0, 0.0], np.float32)
    np.savez(tmp_path / "shift.npz", **z)
    curve = surface_loss.score_recording(str(tmp_path / "shift.npz"), loss,
                                         corr=surface_loss.TrackedCorrespondenceLoss(ov, device=gpu))
    # bound at the first (shifted) frame: no relative error there
    assert curve["corr_rmse_mm"][0] == pytest.approx(0.0, abs=1e-9)
    assert np.all(curve["rmse_mm"] <= 3.0 + 1e-6) and np.all(curve["rmse_mm"] > 0.0)
    # the whole curve again with the oracle's query in place of the GPU's
    orig = surface_loss.point_mesh_query
    try:
        surface_loss.point_mesh_query = (lambda q, v, f=None, vf=None, *, valid=None, face_ok=None, k=1, project=True,
                                         device=None: osurf.point_mesh_query(q, v, f, valid=valid, face_ok=face_ok,
                                                                             k=k, project=project))
        ocurve = surface_loss.score_recording(str(tmp_path / "shift.npz"), loss,
                                              corr=surface_loss.TrackedCorrespondenceLoss(ov))
    finally:
        surface_loss.point_mesh_query = orig
    for key, atol in (("mse", 1e-18), ("mse_tracked_to_sim", 1e-18), ("max_mm", 1e-9), ("corr_mse", 1e-18),
                      ("corr_max_mm", 1e-9)):
        np.testing.assert_allclose(curve[key], ocurve[key], rtol=1e-9, atol=atol, err_msg=key)
