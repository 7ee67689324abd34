"""Self-contact inside the Newton solve on the GPU (mfem_selfcontact.hip)
against the FP64 oracle's restatement (every pair enumerated):

* the candidate lists from the broad phase equal the brute-force sets of
  pairs whose boxes (inflated by dhat / 2) overlap, static and swept by dx
  (boxes over x and x + dx);
* the barrier gradient and diagonal blocks (FP64 sums) equal the oracle's at
  the same positions;
* the CCD bound of the first line-search step equals the oracle's time of
  impact;
* whole steps of the stacked slabs match the FP64 oracle within 1e-4, with
  every solve at rel-res <= 1e-6 and the slabs apart.
"""
import ctypes

import numpy as np
import pytest
import torch

from mfem_amd import scene
from oracle import Oracle
from oracle import collide as C

pytestmark = pytest.mark.gpu


def _make(m, device):
    from mfem_amd.solver import RefinementSolver, model_state

    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    solver = RefinementSolver(m, m.params["iterations"], max_tets=m.tet_count, enable_refinement=False,
                              device=device, **kw)
    return solver, model_state(m, device)


def _touching(m):
    nb = m.particle_count // 2
    x = m.particle_q.copy()
    rng = np.random.default_rng(0)
    x[nb:, 2] -= 2.0e-4
    x[nb:] += 1.0e-5 * rng.normal(size=(nb, 3))
    return x.astype(np.float32)


def _fp32_params(m):
    """The barrier's parameters as the device holds them (FP32)."""
    return {k: float(np.float32(m.params[k])) for k in ("self_contact_dhat", "self_contact_kappa")}


def _sets(pairs):
    return sorted(map(tuple, np.asarray(pairs).tolist()))


def test_candidates_equal_brute_force(gpu):
    m = scene.stacked_boxes_scene()
    solver, state = _make(m, gpu)
    x = _touching(m)
    state.particle_q[: m.particle_count] = torch.as_tensor(x, device=gpu)
    solver.bind(state)
    solver.begin_step(0.01)  # static candidates at x
    c = solver.self_contact_candidates()
    assert c["status"].tolist() == [0, 0]
    dh = m.params["self_contact_dhat"]
    tris = m.tri_indices.astype(np.int32)
    pts = np.unique(tris.reshape(-1)).astype(np.int32)
    np.testing.assert_array_equal(c["points"], pts)
    ref_pt = C.brute_force_pt(x, pts, tris, dh)
    ref_ee = C.brute_force_ee(x, c["edges"], dh)
    assert len(ref_pt) > 0 and len(ref_ee) > 0
    assert _sets(c["pt"]) == _sets(ref_pt)
    assert _sets(c["ee"]) == _sets(ref_ee)
    # swept by a step of the upper slab (down into the lower one and
    # sideways): the pairs whose boxes over x and x + dx overlap
    nb = m.particle_count // 2
    dx = np.zeros((solver.max_particles, 3))
    dx[nb:m.particle_count] = (-0.02, 0.01, -1.5e-3)
    solver.buffers["dx"].copy_(torch.as_tensor(dx, device=gpu))
    solver._call("mfem_self_contact_candidates", 1, solver._stream())
    cs = solver.self_contact_candidates()
    assert cs["status"].tolist() == [0, 0]
    xe = x + dx[: m.particle_count].astype(np.float32)
    sw_pt, sw_ee = _swept_brute_force(x, xe, pts, tris, cs["edges"], dh)
    assert _sets(cs["pt"]) == _sets(sw_pt) and _sets(cs["ee"]) == _sets(sw_ee)
    assert set(_sets(ref_pt)) < set(_sets(sw_pt))


def _swept_brute_force(x, xe, pts, tris, edges, dh):
    """Pairs whose FP32 boxes over the vertices at x and at xe, inflated by
    dhat / 2 each, overlap (no shared vertex; edge pairs e1 < e2)."""
    r = np.float32(dh) * np.float32(0.5)

    def boxes(prims):
        a, b = x[prims], xe[prims]
        return (np.minimum(a.min(axis=1), b.min(axis=1)) - r, np.maximum(a.max(axis=1), b.max(axis=1)) + r)

    plo, phi = boxes(pts[:, None])
    tlo, thi = boxes(tris)
    out_pt = []
    for k, p in enumerate(pts):
        ov = np.all((plo[k] <= thi) & (tlo <= phi[k]), axis=1) & ~np.any(tris == p, axis=1)
        out_pt += [(int(p), int(t)) for t in np.nonzero(ov)[0]]
    lo, hi = boxes(edges)
    out_ee = []
    for a in range(len(edges)):
        ov = np.all((lo[a] <= hi[a + 1:]) & (lo[a + 1:] <= hi[a]), axis=1)
        share = np.any(edges[a + 1:][:, :, None] == edges[a][None, None, :], axis=(1, 2))
        out_ee += [(a, a + 1 + int(b)) for b in np.nonzero(ov & ~share)[0]]
    return np.array(out_pt).reshape(-1, 2), np.array(out_ee).reshape(-1, 2)


def test_barrier_terms_match_oracle(gpu):
    m = scene.stacked_boxes_scene()
    solver, state = _make(m, gpu)
    x = _touching(m)
    state.particle_q[: m.particle_count] = torch.as_tensor(x, device=gpu)
    solver.bind(state)
    solver.begin_step(0.01)
    solver.update_system(0.01)
    torch.cuda.synchronize()
    n = m.particle_count
    g = solver.buffers["sc_g"][:n].cpu().numpy()
    h = solver.buffers["sc_h"][:n].cpu().numpy()
    o = Oracle(m, 64, params=_fp32_params(m))
    e, go, ho = o.sc_gradient(x.astype(np.float64))
    assert e > 0
    np.testing.assert_allclose(g, go, rtol=1e-9, atol=1e-12 * np.abs(go).max())
    np.testing.assert_allclose(h, ho, rtol=1e-9, atol=1e-12 * np.abs(ho).max())


def test_ccd_bound_matches_oracle(gpu):
    m = scene.stacked_boxes_scene()
    solver, state = _make(m, gpu)
    x = _touching(m)
    state.particle_q[: m.particle_count] = torch.as_tensor(x, device=gpu)
    solver.bind(state)
    nb = m.particle_count // 2
    dx = np.zeros((solver.max_particles, 3))
    dx[nb:m.particle_count, 2] = -1.5e-3  # through the lower slab: cut
    solver.buffers["dx"].copy_(torch.as_tensor(dx, device=gpu))
    solver._call("mfem_self_contact_candidates", 1, solver._stream())
    solver._call("mfem_self_contact_ccd", solver._stream())
    torch.cuda.synchronize()
    a = float(solver.buffers["sc_alpha"].item())
    o = Oracle(m, 64, params=_fp32_params(m))
    t = o.sc_toi(x.astype(np.float64), dx[: m.particle_count])
    print("CCD bound", a, "oracle", t)
    assert 0.0 < a < 0.2
    assert a == pytest.approx(t, rel=1e-5)


def test_stacked_boxes_steps_match_fp64_oracle(gpu):
    m = scene.stacked_boxes_scene()
    solver, state = _make(m, gpu)
    o = Oracle(m, 64)
    out = state.clone()
    nb = m.particle_count // 2
    for k in range(6):
        solver.step(state, out, None, None, 0.01)
        state, out = out, state
        o.step(0.01)
        st = solver.last_stats()
        assert np.all(st[:, 1] <= 1e-6), st[:, :2]
        assert solver.status() == 0
        x = state.particle_q[: m.particle_count].double().cpu().numpy()
        err = np.abs(x - o.x).max() / np.abs(o.x).max()
        gap = x[nb:, 2].min() - x[:nb, 2].max()
        print("step", k, "cg", st[:, 0].astype(int).tolist(), "alpha", np.round(st[:, 4], 3).tolist(),
              "rel err", err, "gap", gap)
        assert gap > 0.0
    assert err <= 1e-4
