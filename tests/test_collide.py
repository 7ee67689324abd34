"""Self-contact building blocks on the GPU against the oracle (oracle/collide.py).

Bars: scan outputs and candidate lists bit-exact (integer work); distances
within 1e-5 of the FP64 oracle (FP32 arithmetic on unit-scale inputs, the
reference's own tolerance in tests/test_collision.py); barrier gradients
within 1e-3 relative of central differences of the FP64 oracle energy.
"""
import numpy as np
import pytest
import torch

from oracle import collide as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 5, 1023, 1024, 1025, 16384 - 7, (1 << 20) + 7, 3 * (1 << 20) + 1])
def test_scan_matches_cumsum(gpu, n):
    from mfem_amd.collide import scan_i32

    rng = np.random.default_rng(n)
    a = rng.integers(0, 10, n).astype(np.int32)
    t = torch.from_numpy(a).to(gpu)
    ex, tot = scan_i32(t)
    inc, tot2 = scan_i32(t, inclusive=True)
    np.testing.assert_array_equal(ex.cpu().numpy(), np.cumsum(a) - a)
    np.testing.assert_array_equal(inc.cpu().numpy(), np.cumsum(a))
    assert int(tot.item()) == int(a.sum()) == int(tot2.item())


def _soup(rng, n, spread=0.08):
    x = rng.uniform(0, 1, (3 * n, 3)).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    cen = x.reshape(n, 3, 3).mean(axis=1, keepdims=True)
    x = (cen + spread * (x.reshape(n, 3, 3) - cen)).reshape(-1, 3).astype(np.float32)
    return x, tris


def _surface_scene():
    from mfem_amd import scene

    m = scene.drop_box_scene()
    rng = np.random.default_rng(3)
    x = (m.particle_q + 0.003 * rng.normal(size=m.particle_q.shape)).astype(np.float32)
    return x, np.asarray(m.tri_indices, np.int32)


@pytest.mark.parametrize("case", ["soup", "surface"])
def test_broadphase_matches_brute_force(gpu, case):
    from mfem_amd.collide import BroadPhase

    if case == "soup":
        x, tris = _soup(np.random.default_rng(11), 400)
        d_hat = 0.03
    else:
        x, tris = _surface_scene()
        d_hat = 0.01
    bp = BroadPhase(tris, x.shape[0], gpu)
    xt = torch.from_numpy(x).to(gpu)
    pt = bp.point_triangle(xt, d_hat).cpu().numpy()
    ee = bp.edge_edge(xt, d_hat).cpu().numpy()
    pt_ref = C.brute_force_pt(x, C.surface_points(tris), tris, d_hat)
    ee_ref = C.brute_force_ee(x, C.surface_edges(tris), d_hat)
    print(case, "pt pairs", len(pt), "ee pairs", len(ee))
    assert len(pt_ref) > 0 and len(ee_ref) > 0
    np.testing.assert_array_equal(pt, pt_ref)
    np.testing.assert_array_equal(ee, ee_ref)
    # same lists with a cell size 3x larger (more entries per bucket) and
    # twice again (determinism)
    np.testing.assert_array_equal(bp.point_triangle(xt, d_hat, cell=3 * 0.05).cpu().numpy(), pt_ref)
    np.testing.assert_array_equal(bp.edge_edge(xt, d_hat).cpu().numpy(), ee_ref)


def test_broadphase_octopus_surface_properties(gpu):
    """Full-size surface: deterministic, conservative on sampled pairs."""
    from mfem_amd import scene
    from mfem_amd.collide import BroadPhase

    m = scene.octopus_scene()
    tris = np.asarray(m.tri_indices, np.int32)
    x = m.particle_q.astype(np.float32)
    d_hat = 2e-3
    bp = BroadPhase(tris, x.shape[0], gpu, max_pairs=1 << 22)
    xt = torch.from_numpy(x).to(gpu)
    pt1 = bp.point_triangle(xt, d_hat).cpu().numpy()
    pt2 = bp.point_triangle(xt, d_hat).cpu().numpy()
    ee = bp.edge_edge(xt, d_hat).cpu().numpy()
    print("octopus surface: tris", len(tris), "pt", len(pt1), "ee", len(ee))
    np.testing.assert_array_equal(pt1, pt2)
    # sorted, unique, no incident pairs
    keys = pt1[:, 0].astype(np.int64) * len(tris) + pt1[:, 1]
    assert np.all(np.diff(keys) > 0)
    assert not np.any(np.any(tris[pt1[:, 1]] == pt1[:, 0:1], axis=1))
    # every sampled (point, tri) pair closer than d_hat is a candidate
    rng = np.random.default_rng(0)
    have = set(map(tuple, pt1))
    pts = np.unique(tris)
    cen = x[tris].mean(axis=1)
    for p in rng.choice(pts, 200, replace=False):
        near = np.nonzero(np.linalg.norm(cen - x[p], axis=1) < 0.02)[0]
        for t in near:
            if p in tris[t]:
                continue
            if C.point_triangle_dist(x[p], *x[tris[t]]) < d_hat:
                assert (p, t) in have


def test_soft_rigid_distance_known_answers(gpu):
    from mfem_amd.collide import soft_rigid_distance

    def run(rigid, soft, rigid_tri_count=1, d_hat=2.0, count=1):
        pts = torch.tensor(np.vstack([rigid, soft]), dtype=torch.float32, device=gpu)
        tris = torch.tensor([[0, 1, 2], [3, 4, 5]], dtype=torch.int32, device=gpu)
        pairs = torch.tensor([[0, 1]], dtype=torch.int32, device=gpu)
        cnt = torch.tensor([count], dtype=torch.int32, device=gpu)
        dist = torch.full((3,), d_hat, dtype=torch.float32, device=gpu)
        soft_rigid_distance(pts, tris, 3, rigid_tri_count, pairs, cnt, d_hat, dist)
        out = dist.cpu().numpy()
        ref = C.soft_rigid_distance(np.vstack([rigid, soft]), [[0, 1, 2], [3, 4, 5]], 3, rigid_tri_count, [[0, 1]],
                                    count, d_hat, np.full(3, d_hat))
        np.testing.assert_allclose(out, ref, atol=1e-5)
        return out

    rigid = np.array([[-10.0, -10.0, 0.0], [10.0, -10.0, 0.0], [0.0, 10.0, 0.0]])
    h = 0.3
    np.testing.assert_allclose(run(rigid, [[-0.2, 0.0, h], [0.2, 0.0, h], [0.0, 0.3, h]], d_hat=0.5), [h] * 3, atol=1e-5)
    soft = [[-0.2, 0.0, 0.1], [0.2, 0.0, 0.1], [0.0, 0.3, 0.1]]
    np.testing.assert_allclose(run(rigid, soft, rigid_tri_count=2), [2.0] * 3, atol=1e-9)
    np.testing.assert_allclose(run(rigid, soft, count=0), [2.0] * 3, atol=1e-9)
    d = run(rigid, [[0.0, 0.0, 0.1], [0.3, 0.0, 0.4], [0.0, 0.3, 0.4]], d_hat=0.5)
    assert d[0] == pytest.approx(0.1, abs=1e-5) and np.all(d <= 0.5)
    # a random rigid/soft pair set against the oracle (many features active)
    rng = np.random.default_rng(4)
    for _ in range(20):
        rig = rng.normal(0, 0.3, (3, 3))
        sof = rng.normal(0, 0.3, (3, 3)) + np.array([0, 0, 0.2])
        run(rig, sof, d_hat=0.6)


def _pt_energy(x, d_hat, kappa):
    d = C.point_triangle_dist(x[0], x[1], x[2], x[3])
    return kappa * C.ipc_barrier(d, d_hat)[0] if d < d_hat else 0.0


def _ee_energy(x, d_hat, kappa):
    d = C.edge_edge_dist(x[0], x[1], x[2], x[3])
    return kappa * C.ipc_barrier(d, d_hat)[0] if d < d_hat else 0.0


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_pair_barrier_matches_oracle_and_fd(gpu, kind):
    from mfem_amd.collide import pair_barrier

    rng = np.random.default_rng(5 if kind == "pt" else 6)
    d_hat, kappa = 0.1, 10.0
    n = 64
    X = []
    for _ in range(n):
        a, b, c = rng.normal(0, 0.3, (3, 3))
        if kind == "pt":
            w = rng.dirichlet([1, 1, 1])
            nrm = np.cross(b - a, c - a)
            p = w @ np.array([a, b, c]) + rng.uniform(0.01, 0.08) * nrm / np.linalg.norm(nrm)
            X.append([p, a, b, c])
        else:
            m1 = (a + b) / 2
            dirv = rng.normal(size=3)
            dirv -= dirv @ (b - a) / ((b - a) @ (b - a)) * (b - a)
            off = rng.uniform(0.01, 0.08) * dirv / np.linalg.norm(dirv)
            e = rng.normal(size=3)
            X.append([a, b, m1 + off - 0.2 * e, m1 + off + 0.2 * e])
    X = np.array(X, np.float64)
    x = torch.tensor(X.reshape(-1, 3), dtype=torch.float32, device=gpu)
    if kind == "pt":
        pairs = torch.tensor([[4 * k, k] for k in range(n)], dtype=torch.int32, device=gpu)
        tris = torch.tensor([[4 * k + 1, 4 * k + 2, 4 * k + 3] for k in range(n)], dtype=torch.int32, device=gpu)
        edges = None
        energy_fn = _pt_energy
    else:
        pairs = torch.tensor([[2 * k, 2 * k + 1] for k in range(n)], dtype=torch.int32, device=gpu)
        edges = torch.tensor([[4 * k, 4 * k + 1] if j == 0 else [4 * k + 2, 4 * k + 3] for k in range(n) for j in (0, 1)],
                             dtype=torch.int32, device=gpu)
        tris = None
        energy_fn = _ee_energy
    cnt = torch.tensor([n], dtype=torch.int32, device=gpu)
    dist, energy, grad, vids = pair_barrier(kind, x, pairs, cnt, kappa, d_hat, tris=tris, edges=edges)
    dist, energy, grad = dist.cpu().numpy(), energy.cpu().numpy(), grad.cpu().numpy()
    np.testing.assert_array_equal(vids.cpu().numpy(), np.arange(4 * n).reshape(n, 4))
    for k in range(n):
        d_ref = (C.point_triangle_dist(*X[k]) if kind == "pt" else C.edge_edge_dist(*X[k]))
        assert abs(dist[k] - d_ref) <= 1e-5, (k, dist[k], d_ref)
        e_ref = energy_fn(X[k], d_hat, kappa)
        assert abs(energy[k] - e_ref) <= 1e-3 * abs(e_ref) + 1e-9
        h = 1e-6
        g_fd = np.zeros((4, 3))
        for v in range(4):
            for a in range(3):
                xp, xm = X[k].copy(), X[k].copy()
                xp[v, a] += h
                xm[v, a] -= h
                g_fd[v, a] = (energy_fn(xp, d_hat, kappa) - energy_fn(xm, d_hat, kappa)) / (2 * h)
        scale = np.abs(g_fd).max()
        assert np.abs(grad[k] - g_fd).max() <= 1e-3 * scale + 1e-6, (k, grad[k], g_fd)
