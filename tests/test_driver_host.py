"""Host-side checks (no GPU) of the scene driver, the recording format and the
surface-loss bookkeeping (mfem_amd/driver.py, mfem_amd/surface_loss.py), and
pins of the CPU restatement of the point-to-surface query (oracle/surface.py).

References: src/mfem/refinement/sim_octopus.py:994-1176, 2510-2591 (tool path,
recorder); src/mfem/refinement/surface_loss.py (loss); the reference's own
tests/test_surface_loss.py cases are restated here with the oracle's query
standing in for the GPU's, and in tests/test_surface_loss.py on the GPU.
"""
import numpy as np
import pytest
import scipy.optimize
from scipy.spatial.transform import Rotation

from mfem_amd import mesh, surface_loss
from mfem_amd.driver import MeshRecorder, ToolTrajectory, capsule_signed_distance, load_recording
from oracle import surface as osurf


# -- oracle pins ---------------------------------------------------------------
def _closest_brute(p, a, b, c):
    """min over the triangle of |u a + v b + (1-u-v) c - p|^2 by bounded
    minimisation from several starts (a textbook route)."""
    def f(z):
        u, v = z
        if u < 0 or v < 0 or u + v > 1:
            return 1e9 + (min(u, 0) ** 2 + min(v, 0) ** 2 + max(u + v - 1, 0) ** 2)
        q = u * a + v * b + (1 - u - v) * c
        return float(((q - p) ** 2).sum())
    best = None
    for z0 in ((1 / 3, 1 / 3), (0.9, 0.05), (0.05, 0.9), (0.05, 0.05), (0.5, 0.5), (0.5, 0.0), (0.0, 0.5)):
        r = scipy.optimize.minimize(f, z0, method="Nelder-Mead", options=dict(xatol=1e-12, fatol=1e-16, maxiter=4000))
        if best is None or r.fun < best.fun:
            best = r
    return best.fun


def test_oracle_closest_point_matches_minimisation():
    rng = np.random.default_rng(8)
    for _ in range(60):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 1.5
        cp = osurf.closest_point_on_triangles(p, a, b, c)
        d2 = float(((cp - p) ** 2).sum())
        assert d2 == pytest.approx(_closest_brute(p, a, b, c), rel=1e-6, abs=1e-12)


def test_oracle_closest_point_known_regions():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    f = lambda p: osurf.closest_point_on_triangles(np.asarray(p, float), a, b, c)
    np.testing.assert_allclose(f([0.2, 0.3, 5.0]), [0.2, 0.3, 0.0])   # face
    np.testing.assert_allclose(f([-1.0, -1, 0]), a)                   # vertex A
    np.testing.assert_allclose(f([2.0, -0.5, 1]), b)                  # vertex B
    np.testing.assert_allclose(f([-0.5, 2.0, 0]), c)                  # vertex C
    np.testing.assert_allclose(f([0.5, -1.0, 0]), [0.5, 0, 0])        # edge AB
    np.testing.assert_allclose(f([-1.0, 0.5, 0]), [0, 0.5, 0])        # edge AC
    np.testing.assert_allclose(f([1.0, 1.0, 0]), [0.5, 0.5, 0])       # edge BC


def test_oracle_query_against_exhaustive_projection():
    # with every vertex's faces reachable (k = all vertices) the query is the
    # exact point-to-mesh distance: the minimum over all triangles
    rng = np.random.default_rng(2)
    verts = rng.normal(size=(30, 3))
    faces = np.array([rng.choice(30, 3, replace=False) for _ in range(40)])
    q = rng.normal(size=(50, 3))
    sq, face, cp, near = osurf.point_mesh_query(q, verts, faces, k=30)
    every = osurf.closest_point_on_triangles(q[:, None, :], verts[faces[:, 0]][None], verts[faces[:, 1]][None],
                                             verts[faces[:, 2]][None])
    exact = ((every - q[:, None, :]) ** 2).sum(-1)
    np.testing.assert_allclose(sq, exact.min(axis=1), rtol=1e-12)
    np.testing.assert_allclose(((cp - q) ** 2).sum(-1), sq, rtol=1e-12)


# -- surface extraction and incidence --------------------------------------------
def test_surface_triangles_single_tet_outward_and_box_count():
    tet = np.array([[0, 1, 2, 3]])
    x = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    f = surface_loss.surface_triangles_from_tets(tet)
    assert len(f) == 4
    cen = x.mean(axis=0)
    for tri in f:
        n = np.cross(x[tri[1]] - x[tri[0]], x[tri[2]] - x[tri[0]])
        assert n @ (x[tri].mean(axis=0) - cen) > 0
    p, t = mesh.box_tet_mesh((0, 0, 0), (1, 1, 1), (3, 2, 2))
    f = surface_loss.surface_triangles_from_tets(t)
    # a 3x2x2 box: 2 triangles per boundary quad, 2 (3*2 + 3*2 + 2*2) quads
    assert len(f) == 2 * 2 * (6 + 6 + 4)
    assert set(map(tuple, np.sort(f, axis=1))) == set(map(tuple, np.sort(mesh.surface_tris(t), axis=1)))


def test_vertex_faces_padded_matches_loops():
    rng = np.random.default_rng(1)
    faces = rng.integers(0, 25, size=(60, 3))
    vf, deg = surface_loss.vertex_faces_padded(faces, 25)
    ref = osurf.vertex_faces(faces, 25)
    np.testing.assert_array_equal(vf[:, :ref.shape[1]], ref)
    assert np.all(vf[:, ref.shape[1]:] == -1)
    np.testing.assert_array_equal(deg, (ref >= 0).sum(axis=1))


# -- tool path --------------------------------------------------------------------
def _traj(n=5, **kw):
    times = np.arange(n) / 30.0
    T = np.tile(np.eye(4), (n, 1, 1))
    ang = np.linspace(0.0, 0.8, n)
    for i in range(n):
        T[i, :3, :3] = Rotation.from_rotvec([0.0, ang[i], 0.0]).as_matrix()
        T[i, :3, 3] = [0.1 * i, 0.0, 0.5 - 0.02 * i]
    return ToolTrajectory(T, times, capsule_radius=0.01, capsule_half_height=0.02, **kw), T, times


def test_tool_path_samples_recorded_poses_and_slerp_midpoints():
    tr, T, times = _traj(tip_along_z=0.05)
    assert tr.center_offset == pytest.approx(0.05 - 0.03)
    for i in range(len(times)):
        pos, q = tr.sample(times[i])
        R = Rotation.from_quat(q).as_matrix()
        np.testing.assert_allclose(R, T[i, :3, :3], atol=1e-9)
        np.testing.assert_allclose(pos, T[i, :3, 3] + tr.center_offset * T[i, :3, 2], atol=1e-7)
    # halfway between two samples: half the rotation about the fixed axis, mean translation
    tm = 0.5 * (times[1] + times[2])
    pos, q = tr.sample(tm)
    rv = Rotation.from_quat(q).as_rotvec()
    np.testing.assert_allclose(rv, [0.0, 0.3, 0.0], atol=1e-7)
    z = Rotation.from_quat(q).apply([0, 0, 1.0])
    np.testing.assert_allclose(pos - tr.center_offset * z, 0.5 * (T[1, :3, 3] + T[2, :3, 3]), atol=1e-7)
    # clamped beyond the end; looped when asked
    np.testing.assert_allclose(tr.pose7(10.0), tr.pose7(times[-1]), atol=1e-6)
    trl, _, _ = _traj(loop=True, tip_along_z=0.05)
    span = times[-1] - times[0]
    np.testing.assert_allclose(trl.pose7(0.05 + span), trl.pose7(0.05), atol=1e-5)


def test_tool_path_contact_position_and_time_offset():
    tr, T, times = _traj(contact_position=np.array([[0.0, 0.0, 0.5 + 0.0121]] * 5), time_offset=2 / 30.0)
    assert tr.center_offset == pytest.approx(0.0121 - 0.03, abs=1e-12)
    pos, _ = tr.sample(0.0)  # sim time 0 maps to recorded frame 2
    np.testing.assert_allclose(pos, T[2, :3, 3] + tr.center_offset * T[2, :3, 2], atol=1e-7)


def test_retract_to_standoff_clears_the_mesh():
    tr, T, times = _traj(tip_along_z=0.05)
    verts = np.array([[0.0, 0.0, 0.548], [0.0, 0.003, 0.551]])  # at the capsule's +Z cap
    r = tr.retract_to_standoff(verts, 0.002)
    d = capsule_signed_distance(verts, tr.pose7(0.0).astype(np.float64), tr.capsule_radius, tr.capsule_half_height)
    assert r > 0.0 and d.min() >= 0.002 - 1e-9


def test_capsule_signed_distance_sphere_case():
    # the reference's known answer (tests/test_surface_loss.py:45-48)
    d = capsule_signed_distance(np.array([[0.0, 0.0, 0.3], [0.0, 0.0, 0.0]]), np.array([0, 0, 0, 0, 0, 0, 1.0]),
                                0.1, 0.0)
    np.testing.assert_allclose(d, [0.2, -0.1])


# -- recorder format ------------------------------------------------------------------
class _FakeSolver:
    def __init__(self, tets, n):
        import torch
        self.tet_indices = torch.as_tensor(tets, dtype=torch.int32)
        self.active_tet_count = len(tets)
        self.active_particle_count = n


class _FakeState:
    def __init__(self, q, bq):
        import torch
        self.particle_q = torch.as_tensor(q, dtype=torch.float32)
        self.body_q = torch.as_tensor(bq, dtype=torch.float32)[None]


def test_recording_round_trip_constant_and_ragged(tmp_path):
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    q = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=float)
    cap = np.array([0.1, 0.2, 0.3, 0.0, 0.0, 0.0, 1.0])
    rec = MeshRecorder(tmp_path / "c.npz", 30, capsule_radius=0.01, capsule_half_height=0.02)
    for k in range(3):
        rec.capture(_FakeSolver(tets, 5), _FakeState(q + 0.1 * k, cap), k / 30.0)
    info = rec.close()
    assert info["tets"] == "constant" and not info["positions_ragged"]
    z = np.load(tmp_path / "c.npz", allow_pickle=False)
    assert z["positions"].shape == (3, 5, 3) and z["tet_indices"].shape == (2, 4)
    r = load_recording(tmp_path / "c.npz")
    assert r["frame_count"] == 3 and r["capsule"]["radius"] == pytest.approx(0.01)
    np.testing.assert_allclose(r["positions_at"](2), q + 0.2, atol=1e-6)
    np.testing.assert_array_equal(r["tets_at"](1), tets)
    # refinement changed the sizes: object arrays for the reference's readers,
    # flat + offsets for a reader that never unpickles
    rec = MeshRecorder(tmp_path / "r.npz", 30)
    rec.capture(_FakeSolver(tets[:1], 4), _FakeState(q[:4], cap), 0.0)
    rec.capture(_FakeSolver(tets, 5), _FakeState(q, cap), 1 / 30.0)
    info = rec.close()
    assert info["positions_ragged"] and info["tets"] == "ragged"
    r = load_recording(tmp_path / "r.npz")
    np.testing.assert_allclose(r["positions_at"](0), q[:4], atol=1e-6)
    np.testing.assert_array_equal(r["tets_at"](1), tets)
    # a file whose ragged frames exist only as object arrays is refused
    obj = np.empty(2, dtype=object)
    obj[:] = [q[:4].astype(np.float32), q.astype(np.float32)]
    np.savez(tmp_path / "o.npz", fps=30.0, frame_count=2, times=np.zeros(2), positions=obj, tet_indices=tets)
    with pytest.raises(ValueError, match="unpickle"):
        load_recording(tmp_path / "o.npz")["positions_at"](0)


# -- loss bookkeeping with the oracle's query (the GPU's in tests/test_surface_loss.py)
@pytest.fixture
def oracle_query(monkeypatch):
    def q(query, verts, faces=None, vert_faces=None, *, valid=None, face_ok=None, k=1, project=True, device=None):
        sq, face, cp, near = osurf.point_mesh_query(query, verts, faces, valid=valid, face_ok=face_ok, k=k,
                                                    project=project)
        return sq, face, cp, near
    monkeypatch.setattr(surface_loss, "point_mesh_query", q)


class _FlatOverlay:
    """21x21 grid in z = 0 (the reference's tests/test_surface_loss.py:13-37)."""

    def __init__(self, n=21, size=1.0):
        xs = np.linspace(0.0, size, n)
        X, Y = np.meshgrid(xs, xs, indexing="ij")
        self.verts = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], axis=1)
        idx = np.arange(n * n).reshape(n, n)
        a, b, c, d = idx[:-1, :-1], idx[1:, :-1], idx[:-1, 1:], idx[1:, 1:]
        self.faces = np.concatenate([np.stack([a.ravel(), b.ravel(), d.ravel()], 1),
                                     np.stack([a.ravel(), d.ravel(), c.ravel()], 1)])
        self.traj = self.verts[None]
        self.valid = np.ones((1, n * n), dtype=bool)
        self.active = True

    def positions(self, frame):
        return self.verts

    def frame_for_time(self, t):
        return 0


def _tool(center, radius):
    tf = np.array([*center, 0.0, 0.0, 0.0, 1.0])
    return lambda P: capsule_signed_distance(P, tf, radius, 0.0)


def test_loss_tool_exclusion_and_sim_faces(oracle_query):
    ov = _FlatOverlay()
    tool = _tool((0.5, 0.5, 0.0), 0.1)
    sim = ov.verts.copy()
    sim[np.linalg.norm(sim[:, :2] - 0.5, axis=1) < 0.15, 2] = -0.02
    plain = surface_loss.TrackedSurfaceLoss(ov, symmetric=True)
    fair = surface_loss.TrackedSurfaceLoss(ov, symmetric=True, exclude_tool_margin=0.06)
    r0 = plain.evaluate(sim, frame=0, tool_signed_distance=tool, sim_faces=ov.faces)
    r1 = fair.evaluate(sim, frame=0, tool_signed_distance=tool, sim_faces=ov.faces)
    assert r0["mse"] > 0.0 and r0["mse_tracked_to_sim"] > 0.0
    assert r1["n_tracked_excluded"] > 0 and r1["n_sim"] < r1["n_sim_all"]
    assert r1["mse"] == pytest.approx(0.0, abs=1e-12) and r1["mse_tracked_to_sim"] == pytest.approx(0.0, abs=1e-12)
    assert fair.evaluate(sim, frame=0, sim_faces=ov.faces)["mse"] == pytest.approx(r0["mse"])
    coarse = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=float)
    cf = np.array([[0, 1, 3], [0, 3, 2]])
    assert plain.evaluate(coarse, frame=0, sim_faces=cf)["mse_tracked_to_sim"] == pytest.approx(0.0, abs=1e-12)
    assert plain.evaluate(coarse, frame=0)["mse_tracked_to_sim"] > 1e-3
    bad = ov.verts.copy()
    bad[0, 2] = np.nan
    row = plain.evaluate(bad, frame=0, sim_faces=ov.faces)
    assert np.isnan(row["mse"]) and np.isnan(row["mse_symmetric"])


def test_correspondence_slide_and_bind_residual(oracle_query):
    ov = _FlatOverlay()
    sim = ov.verts.copy()
    corr = surface_loss.TrackedCorrespondenceLoss(ov)
    info = corr.bind(sim, frame=0)
    assert info["n_bound"] == len(sim) and info["bind_rmse_mm"] == pytest.approx(0.0, abs=1e-9)
    r = corr.evaluate(sim + [0.01, 0.0, 0.0], frame=0)
    assert r["rmse_mm"] == pytest.approx(10.0, abs=1e-9) and r["max_mm"] == pytest.approx(10.0, abs=1e-9)
    rng = np.random.default_rng(0)
    pts = rng.uniform(0.1, 0.9, size=(200, 3))
    pts[:, 2] = 0.003
    rel = surface_loss.TrackedCorrespondenceLoss(ov, relative=True)
    ab = surface_loss.TrackedCorrespondenceLoss(ov, relative=False)
    assert rel.bind(pts, frame=0)["bind_rmse_mm"] == pytest.approx(3.0, abs=1e-6)
    ab.bind(pts, frame=0)
    tgt, ok = rel.targets(0)
    assert ok.all() and np.allclose(tgt[:, :2], pts[:, :2], atol=1e-9) and np.allclose(tgt[:, 2], 0.0)
    assert rel.evaluate(pts, frame=0)["rmse_mm"] == pytest.approx(0.0, abs=1e-9)
    assert ab.evaluate(pts, frame=0)["rmse_mm"] == pytest.approx(3.0, abs=1e-6)
