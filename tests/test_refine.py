"""One refinement pass on the GPU against the sequential specification
(oracle/refine.py).

Bars: topology (tets, triangles) and the new vertices' positions, velocities
and rest positions bit-exact (integer work, and FP32 interpolation with
round-to-nearest operations on both sides); per-tet data of the children
copied exactly; DmInv of the children within 1e-5 relative of the FP64
inverse of the same rest positions.
"""
import numpy as np
import pytest
import torch

from mfem_amd import mesh as M
from oracle import refine as R

pytestmark = pytest.mark.gpu


def _case(cells, h, seed, pinned_face=True):
    p, t = M.box_tet_mesh((0, 0, 0), (cells[0] * h, cells[1] * h, cells[2] * h), cells)
    p = p.astype(np.float32)
    t = t.astype(np.int32)
    rng = np.random.default_rng(seed)
    p = (p + rng.normal(0, 0.05 * h, p.shape)).astype(np.float32)  # deformed positions
    rest = M.box_tet_mesh((0, 0, 0), (cells[0] * h, cells[1] * h, cells[2] * h), cells)[0].astype(np.float32)
    inv_mass = np.ones(len(p), np.float32)
    if pinned_face:
        inv_mass[rest[:, 0] < 1e-6] = 0
    vs = rng.uniform(0, 30, len(p)).astype(np.float32)
    ts = rng.uniform(0, 5, len(t)).astype(np.float32)
    return p, rest, t, M.surface_tris(t).astype(np.int32), inv_mass, vs, ts


def _gpu_pass(gpu, p, rest, t, tris, inv_mass, vs, ts, threshold, w_tet, w_vert, max_tets, max_verts, max_tris,
              plane_z=None, split_t=0.5, geo=None, shape_types=None):
    from mfem_amd.refine import Refiner

    V, T, F = len(p), len(t), len(tris)
    dev = gpu

    def cap(a, n, dtype=torch.float32):
        out = torch.zeros((n,) + tuple(np.shape(a)[1:]), dtype=dtype, device=dev)
        out[:len(a)] = torch.as_tensor(np.asarray(a), dtype=dtype, device=dev)
        return out

    rng = np.random.default_rng(99)
    x = cap(p, max_verts)
    v = cap(rng.normal(size=p.shape).astype(np.float32), max_verts)
    rs = cap(rest, max_verts)
    stretch = rng.normal(size=(T, 6)).astype(np.float32)
    lam = rng.normal(size=(T, 6)).astype(np.float32)
    mat = rng.uniform(1, 2, (T, 3)).astype(np.float32)
    dminv = R.rest_dminv(rest, t).reshape(T, 9).astype(np.float32)
    shapes = None
    if plane_z is not None:
        shapes = (torch.tensor([0], dtype=torch.int32, device=dev),
                  torch.tensor([[0, 0, plane_z, 0, 0, 0, 1]], dtype=torch.float32, device=dev),
                  torch.tensor([[0, 0, 0]], dtype=torch.float32, device=dev))
    if shape_types is not None:  # far away: only their types matter (geometric scoring)
        n = len(shape_types)
        shapes = (torch.tensor(shape_types, dtype=torch.int32, device=dev),
                  torch.tensor([[0, 0, -100.0, 0, 0, 0, 1]] * n, dtype=torch.float32, device=dev),
                  torch.tensor([[0.01, 0.05, 0]] * n, dtype=torch.float32, device=dev))
    extra = {}
    if geo is not None:
        extra = dict(geo)
        for k in ("vertex_dist", "tri_dist", "tri_bary"):
            if extra.get(k) is not None:
                extra[k] = torch.as_tensor(extra[k], dtype=torch.float32, device=dev).contiguous()
        extra["vertex_sid"] = torch.as_tensor(extra["vertex_sid"], dtype=torch.int32, device=dev)
    out = {"tets": torch.zeros(max_tets, 4, dtype=torch.int32, device=dev),
           "stretch": torch.zeros(max_tets, 6, device=dev), "lam": torch.zeros(max_tets, 6, device=dev),
           "materials": torch.zeros(max_tets, 3, device=dev), "dminv": torch.zeros(max_tets, 9, device=dev),
           "tris": torch.zeros(max_tris, 3, dtype=torch.int32, device=dev)}
    ref = Refiner(max_tets, max_tris, max_verts, dev)
    res = ref.run(tets=cap(t, T, torch.int32), tris=cap(tris, F, torch.int32), x=x, v=v, rest=rs,
                  inv_mass=cap(inv_mass, V), tet_score=cap(ts, T), vertex_score=cap(vs, V),
                  stretch=cap(stretch, T), lam=cap(lam, T), materials=cap(mat, T), dminv=cap(dminv, T),
                  n_verts=V, n_tets=T, n_tris=F, shapes=shapes, threshold=threshold, w_tet=w_tet, w_vertex=w_vert,
                  split_t=split_t, out=out, **extra)
    host = {k: val.cpu().numpy() for k, val in out.items()}
    return res, host, x.cpu().numpy(), v.cpu().numpy(), rs.cpu().numpy(), stretch, lam, mat, dminv


def _oracle(p, rest, t, tris, inv_mass, vs, ts, threshold, w_tet, w_vert, max_tets, max_verts, v, stretch, lam, mat,
            plane_z=None, split_t=0.5):
    thr = float(np.float32(threshold))
    md = None if plane_z is None else (lambda m: float(np.float32(m[2]) - np.float32(plane_z)))
    sc = R.legacy_scores(t, p, inv_mass, ts, vs, w_tet, w_vert, 1.0e6, thr, midpoint_distance=md)
    sel = R.select_edges(t, sc)
    out = R.refine_pass(t, tris, p, v[:len(p)], rest, stretch, lam, mat, sel, len(p), max_tets, max_verts,
                        split_t=split_t)
    return sel, out


@pytest.mark.parametrize("cells,h,seed,threshold,w_tet,plane", [
    ((6, 3, 3), 0.1, 0, 2.0, 0.3, None),
    ((6, 3, 3), 0.1, 1, 0.5, 0.0, None),
    ((8, 4, 4), 0.05, 2, 1.0, 0.3, 0.06),
    ((4, 4, 4), 0.1, 3, 5.0, 1.0, 0.05),
])
def test_refine_pass_matches_specification(gpu, cells, h, seed, threshold, w_tet, plane):
    p, rest, t, tris, inv_mass, vs, ts = _case(cells, h, seed)
    max_tets, max_verts, max_tris = 4 * len(t), 4 * len(p), 4 * len(tris)
    res, G, x, v, rs, stretch, lam, mat, dminv = _gpu_pass(gpu, p, rest, t, tris, inv_mass, vs, ts, threshold, w_tet,
                                                           0.1, max_tets, max_verts, max_tris, plane_z=plane)
    sel, O = _oracle(p, rest, t, tris, inv_mass, vs, ts, threshold, w_tet, 0.1, max_tets, max_verts, v, stretch, lam,
                     mat, plane_z=plane)
    print("selected", len(sel), "rounds", res["rounds"], "tets", res["n_tets"], "verts", res["n_verts"])
    assert len(sel) > 0 and res["new_verts"] == len(sel)
    nt, nv, nf = res["n_tets"], res["n_verts"], res["n_tris"]
    assert nt == len(O["tets"]) and nv == O["n_verts"] and nf == len(O["tris"])
    np.testing.assert_array_equal(G["tets"][:nt], O["tets"])
    np.testing.assert_array_equal(G["tris"][:nf], O["tris"])
    np.testing.assert_array_equal(x[:nv], O["x"])
    np.testing.assert_array_equal(v[:nv], O["v"])
    np.testing.assert_array_equal(rs[:nv], O["rest"])
    par = O["parent"]
    np.testing.assert_array_equal(G["stretch"][:nt], stretch[par])
    np.testing.assert_array_equal(G["lam"][:nt], lam[par])
    np.testing.assert_array_equal(G["materials"][:nt], mat[par])
    ref_dminv = R.rest_dminv(O["rest"], O["tets"]).reshape(nt, 9)
    err = np.abs(G["dminv"][:nt] - ref_dminv).max() / np.abs(ref_dminv).max()
    print("dminv rel err", err)
    assert err <= 1e-5


@pytest.mark.parametrize("seed,with_tri", [(0, False), (1, True), (2, True)])
def test_geometric_refine_pass_matches_specification(gpu, seed, with_tri):
    """Geometric scoring (edge length, longest-edge ratio, excess energy
    density, capsule proximity of the endpoints, triangle proximity): the
    same topology and new vertices as the specification, bit for bit."""
    p, rest, t, tris, inv_mass, vs, ts = _case((6, 3, 3), 0.1, seed)
    rng = np.random.default_rng(100 + seed)
    V, F = len(p), len(tris)
    shape_types = [1, 0]  # capsule, plane
    geo = dict(scoring="geometric", min_edge_length=0.04, contact_d1=0.2, w_elastic=1.0, w_vertex_contact=0.8,
               w_tri_contact=1.5,
               vertex_dist=rng.uniform(-0.05, 0.3, V).astype(np.float32),
               vertex_sid=rng.integers(-1, 2, V).astype(np.int32))
    if with_tri:
        bary = rng.uniform(0, 1, (F, 3)).astype(np.float32)
        geo.update(tri_dist=rng.uniform(-0.05, 0.4, F).astype(np.float32),
                   tri_bary=(bary / bary.sum(axis=1, keepdims=True)).astype(np.float32))
    threshold = 3.0
    max_tets, max_verts, max_tris = 4 * len(t), 4 * len(p), 4 * len(tris)
    res, G, x, v, rs, stretch, lam, mat, dminv = _gpu_pass(gpu, p, rest, t, tris, inv_mass, vs, ts, threshold, 0.0,
                                                           0.1, max_tets, max_verts, max_tris, geo=geo,
                                                           shape_types=shape_types)
    sc = R.geometric_scores(t, p, inv_mass, ts, mat[:, 0], dminv, geo["vertex_dist"], geo["vertex_sid"], shape_types,
                            0.04, 0.2, 1.0, 0.8, threshold, tris=tris if with_tri else None,
                            tri_dist=geo.get("tri_dist"), tri_bary=geo.get("tri_bary"), w_t=1.5)
    sel = R.select_edges(t, sc)
    O = R.refine_pass(t, tris, p, v[:len(p)], rest, stretch, lam, mat, sel, len(p), max_tets, max_verts)
    print("geometric: candidates", sum(1 for s_ in sc.values() if s_ > R.SCORE_ONE), "selected", len(sel),
          "rounds", res["rounds"])
    assert len(sel) > 0 and res["new_verts"] == len(sel)
    nt, nv, nf = res["n_tets"], res["n_verts"], res["n_tris"]
    assert nt == len(O["tets"]) and nv == O["n_verts"] and nf == len(O["tris"])
    np.testing.assert_array_equal(G["tets"][:nt], O["tets"])
    np.testing.assert_array_equal(G["tris"][:nf], O["tris"])
    np.testing.assert_array_equal(x[:nv], O["x"])
    np.testing.assert_array_equal(rs[:nv], O["rest"])


def test_refine_capacity_guard_drops_pass(gpu):
    p, rest, t, tris, inv_mass, vs, ts = _case((6, 3, 3), 0.1, 0)
    res, G, x, *_ = _gpu_pass(gpu, p, rest, t, tris, inv_mass, vs, ts, 2.0, 0.3, 0.1, len(t) + 2, 4 * len(p),
                              4 * len(tris))
    assert res["dropped"] and res["n_tets"] == len(t) and res["n_verts"] == len(p)
    np.testing.assert_array_equal(x[:len(p)], p)


def test_refine_octopus_scale_deterministic(gpu):
    """Full-size body: two identical passes give identical meshes; the
    result conforms and conserves rest volume."""
    from mfem_amd import scene

    m = scene.octopus_scene()
    p = m.particle_q.astype(np.float32)
    t = np.asarray(m.tet_indices, np.int32)
    tris = np.asarray(m.tri_indices, np.int32)
    rng = np.random.default_rng(0)
    vs = (rng.uniform(0, 1, len(p)) ** 8 * 40000).astype(np.float32)  # a few hot vertices
    ts = np.zeros(len(t), np.float32)
    inv_mass = np.asarray(m.particle_inv_mass, np.float32)
    outs = []
    for _ in range(2):
        res, G, x, v, rs, *_ = _gpu_pass(gpu, p, p, t, tris, inv_mass, vs, ts, 10.1, 0.0, 0.1, 2 * len(t), 2 * len(p),
                                         2 * len(tris))
        outs.append((res, G["tets"][:res["n_tets"]].copy(), G["tris"][:res["n_tris"]].copy(), rs[:res["n_verts"]]))
    (r1, t1, f1, rs1), (r2, t2, f2, rs2) = outs
    print("octopus pass", r1)
    assert r1["new_verts"] > 100
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(f1, f2)
    vol = lambda P, T: np.linalg.det(np.stack([P[T[:, k]] - P[T[:, 0]] for k in (1, 2, 3)], axis=2).astype(np.float64)) / 6
    v0, v1 = vol(p.astype(np.float64), t), vol(rs1.astype(np.float64), t1)
    assert np.all(v1 > 0)
    assert abs(v1.sum() - v0.sum()) <= 1e-6 * abs(v0.sum())
    f, c = np.unique(np.sort(np.concatenate([t1[:, [1, 2, 3]], t1[:, [0, 2, 3]], t1[:, [0, 1, 3]], t1[:, [0, 1, 2]]]),
                             axis=1), axis=0, return_counts=True)
    assert c.max() <= 2
    assert {tuple(x) for x in f[c == 1]} == {tuple(sorted(x)) for x in f1}


def _refining_beam(gpu, every=3):
    from mfem_amd import scene
    from mfem_amd.solver import RefinementSolver, model_state

    m = scene.beam_scene(cells=(8, 2, 2))
    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    # the reference's own refinement test: elastic-energy scoring, a low
    # threshold, refine every few steps (tests/test_refinement.py:68-76, 98-100)
    # (the particle buffer padded for the new vertices, as the reference's
    # caller does; refine_density 1000 from the model)
    solver = RefinementSolver(scene.pad_particles(m, 4 * m.particle_count), m.params["iterations"],
                              max_tets=4 * m.tet_count, max_particles=4 * m.particle_count,
                              device=gpu, enable_refinement=True, refine_every_n_steps=every, min_refine_score=1e-3,
                              refine_tet_score_weight=1.0, **kw)
    return m, solver, model_state(m, gpu, capacity=4 * m.particle_count)


def _rest_volume(solver):
    nt, nv = solver.active_tet_count, solver.active_particle_count
    t = solver.tet_indices.cpu().numpy()
    r = solver.rest_particle_q.double().cpu().numpy()
    assert t.min() >= 0 and t.max() < nv, "orphaned index"
    Dm = np.stack([r[t[:, k]] - r[t[:, 0]] for k in (1, 2, 3)], axis=2)
    vol = np.linalg.det(Dm) / 6.0
    assert np.all(vol > 0), "inverted tet"
    return vol.sum()


def test_solver_refinement_invariants(gpu):
    """Steps with refinement on: conserved rest volume, no orphans / inverted
    tets, finite state, every PCG solve converged, the mesh grows on the
    gated steps only (tests/test_refinement.py:87-130)."""
    m, solver, state = _refining_beam(gpu)
    out = state.clone()
    vol0 = _rest_volume(solver)
    counts = []
    for k in range(9):
        solver.step(state, out, None, None, 1.0 / 60.0)
        state, out = out, state
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st
        assert solver.status() == 0
        nv = solver.active_particle_count
        assert np.all(np.isfinite(state.particle_q[:nv].cpu().numpy()))
        assert abs(_rest_volume(solver) - vol0) <= 1e-6
        counts.append((solver.active_tet_count, nv, solver.active_tri_count))
    print("tets / verts / tris per step", counts, [r["new_verts"] for r in solver.refine_log])
    fired = [r["fired"] for r in solver.refine_log]
    assert fired == [k % 3 == 0 for k in range(9)]
    for k in range(1, 9):
        if not fired[k]:
            assert counts[k] == counts[k - 1]
    assert counts[-1][0] > m.tet_count and counts[-1][1] > m.particle_count


def test_step_after_refinement_matches_fp64_oracle(gpu):
    """After a refining step, one more step on the refined mesh against the
    FP64 oracle built from the solver's own refined topology and state."""
    from mfem_amd.scene import Model
    from oracle import Oracle

    m, solver, state = _refining_beam(gpu, every=3)
    out = state.clone()
    dt = 1.0 / 60.0
    for k in range(4):  # passes fire at steps 0 and 3 (the beam has sagged by then)
        solver.step(state, out, None, None, dt)
        state, out = out, state
    assert solver.active_tet_count > m.tet_count
    b = solver.buffers
    nv, nt, nf = solver.active_particle_count, solver.active_tet_count, solver.active_tri_count
    mr = Model(particle_q=state.particle_q[:nv].double().cpu().numpy(),
               particle_qd=state.particle_qd[:nv].double().cpu().numpy(),
               particle_mass=b["mass"][:nv].double().cpu().numpy(),
               particle_inv_mass=b["inv_mass"][:nv].double().cpu().numpy(),
               tet_indices=b["tets"][:nt].cpu().numpy(),
               tet_poses=b["dminv"][:nt].double().cpu().numpy().reshape(nt, 3, 3),
               tet_materials=b["materials"][:nt].double().cpu().numpy(),
               tri_indices=b["tris"][:nf].cpu().numpy(), shapes=[], gravity=np.asarray(m.gravity),
               params=dict(m.params))
    o = Oracle(mr, 64)
    o.buf("s")[:] = b["s"][:nt].double().cpu().numpy().ravel()
    o.buf("lam")[:] = b["lam"][:nt].double().cpu().numpy().ravel()
    pins = np.nonzero(mr.particle_inv_mass == 0)[0]
    o.buf("pin_pos")[:] = b["pin_pos"][:nv].double().cpu().numpy()[pins].ravel()
    solver.step(state, out, None, None, dt)  # pass 4: not a refining pass
    assert not solver.refine_log[-1]["fired"]
    state, out = out, state
    o.step(dt)
    x = state.particle_q[:nv].double().cpu().numpy()
    err = np.abs(x - o.x).max() / np.abs(o.x).max()
    print("positions after refinement: rel err", err)
    assert err <= 1e-4


def test_solver_geometric_refinement_near_the_tool(gpu):
    """refine_scoring="geometric" in the solver (distances recomputed for the
    current tool pose, solver.py:449-466): with the elastic term off, only
    edges near the capsule split (contact proximity), every new vertex lies
    within the barrier range plus an edge length of the capsule, the rest
    volume is conserved and every solve converges."""
    from mfem_amd import scene
    from mfem_amd.solver import RefinementSolver, model_state

    m = scene.capsule_poke_scene()
    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    solver = RefinementSolver(scene.pad_particles(m, 4 * m.particle_count), m.params["iterations"],
                              max_tets=4 * m.tet_count, max_particles=4 * m.particle_count,
                              device=gpu, enable_refinement=True, refine_every_n_steps=2, refine_scoring="geometric",
                              refine_elastic_weight=0.0, refine_geometric_threshold=1.0, **kw)
    state = model_state(m, gpu, capacity=4 * m.particle_count)
    out = state.clone()
    vol0 = _rest_volume(solver)
    for k in range(3):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
        assert np.all(solver.last_stats()[:, 1] <= 1e-6)
        assert solver.status() == 0
        assert abs(_rest_volume(solver) - vol0) <= 1e-6
    log = solver.refine_log
    print("geometric passes", [(r["fired"], r["new_verts"]) for r in log])
    assert log[0]["fired"] and log[0]["new_verts"] > 0
    nv0, nv = m.particle_count, solver.active_particle_count
    x = solver.rest_particle_q[:nv].double().cpu().numpy()
    cap = m.shapes[0]
    axis_c = np.asarray(m.body_q[0][:3], np.float64)  # the capsule is rigid body 0, at the body's origin
    d = np.hypot(x[nv0:, 1] - axis_c[1], x[nv0:, 2] - axis_c[2]) - cap.scale[0]  # capsule along world x
    edge = 0.3 / 6
    assert np.all(d <= float(kw["contact_d1"]) + edge), d.max()


@pytest.mark.gpu
def test_galerkin_start_across_refinement(gpu):
    """The learned warm start's Galerkin start with refinement passes: the
    Newton-step history restarts empty when the topology changes (its row
    stride is the chunk count), and every solve still starts from the
    Galerkin combination and meets rel-res 1e-6 with no device error."""
    from mfem_amd.learned import WarmStartNet, init_params

    m, solver, state = _refining_beam(gpu, every=3)
    solver.set_warm_start(WarmStartNet(gpu, weights=init_params(3)))
    out = state.clone()
    dt = 1.0 / 60.0
    chunks = []
    for k in range(5):  # passes fire at steps 0 and 3
        solver.step(state, out, None, None, dt)
        state, out = out, state
        st = solver.last_stats()
        assert np.all(st[:, 14] == 3) and np.all(st[:, 1] <= 1e-6), (k, st[:, [0, 1, 14]])
        chunks.append(solver._S.n_chunks)
        if solver.refine_log[-1]["fired"]:
            # the pass ran before this step's solves: the first solve saw an
            # empty history (its two history candidates are dropped)
            assert np.isfinite(st[0, 15])
    assert solver.active_tet_count > m.tet_count and solver.status() == 0
    print("chunks per step", chunks)
