"""Pins of the geometric refinement score (oracle/refine.py) on the CPU.

Restates the reference's own checks of populate_candidates_geometric /
populate_tri_candidates_geometric (tests/test_refine_scoring.py:74-198) on
the oracle: per-edge maxima (no valence bias), edges shorter than 2 h_min
registered but never candidates, the longest edge of a tet scoring
(L / h_min) x energy density, fully pinned edges excluded, vertex contact
counted for capsules only and penetration above surface contact, the
triangle term favouring the edge nearest the closest point, and the elastic
term as the excess over the mean density.  The density sum in the GPU's
tree order equals math.fsum to rounding.
"""
import math

import numpy as np
import pytest

from oracle import refine as R

H_MIN = 0.1
D1 = 0.2
THR = 1.0
Q2 = np.array([(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.3, 0.6, 0.0), (0.3, 0.2, 0.6), (0.3, 0.2, -0.6)], np.float32)
T2 = np.array([(0, 1, 2, 3), (0, 1, 2, 4)], np.int32)


def _dminv(q, tets):
    return np.array([np.linalg.inv(np.stack([q[t[1]] - q[t[0]], q[t[2]] - q[t[0]], q[t[3]] - q[t[0]]], axis=1))
                     for t in tets], np.float32).reshape(-1, 9)


def _run(q, tets, energy, dist=None, sid=None, shape_type=(R.CAPSULE,), inv_mass=None, total=0.0, **kw):
    n = len(q)
    sc = R.geometric_scores(tets, q, np.ones(n, np.float32) if inv_mass is None else inv_mass,
                            np.asarray(energy, np.float32), np.ones(len(tets), np.float32), _dminv(q, tets),
                            np.full(n, 1e8, np.float32) if dist is None else dist,
                            np.full(n, -1, np.int32) if sid is None else sid, shape_type, H_MIN, D1, 1.0, 1.0,
                            THR, elastic_total=total, **kw)
    return {(k >> 32, k & 0xFFFFFFFF): v / R.SCORE_ONE * THR for k, v in sc.items()}


def _vol(q, t):
    return abs(np.linalg.det(np.stack([q[t[1]] - q[t[0]], q[t[2]] - q[t[0]], q[t[3]] - q[t[0]]], axis=1))) / 6.0


def test_scores_are_per_edge_maxima_not_sums():
    one = _run(Q2, T2, [1.0, 0.0])
    both = _run(Q2, T2, [1.0, 1.0])
    for e in [(0, 1), (0, 2), (1, 2)]:
        assert one[e] > 0.0 and both[e] == pytest.approx(one[e], rel=1e-5)
    assert one[(0, 4)] == 0.0 and one[(1, 4)] == 0.0 and one[(2, 4)] == 0.0


def test_longest_edge_scores_length_times_density():
    sc = _run(Q2, T2, [1.0, 1.0])
    lengths = {e: float(np.linalg.norm(Q2[e[0]].astype(float) - Q2[e[1]].astype(float))) for e in sc}
    longest = max(lengths, key=lengths.get)
    assert longest == (0, 1) and sc[longest] == max(sc.values())
    assert sc[longest] == pytest.approx(lengths[longest] / H_MIN / _vol(Q2.astype(float), T2[0]), rel=1e-6)


def test_short_edges_registered_but_never_candidates():
    sc = _run(Q2 * np.float32(0.15), T2, [1.0, 1.0])
    assert len(sc) == 9 and all(v == 0.0 for v in sc.values())


def test_fully_pinned_edges_excluded():
    sc = _run(Q2, T2, [1.0, 1.0], inv_mass=np.array([0, 0, 1, 1, 1], np.float32))
    assert (0, 1) not in sc and (0, 2) in sc


def test_vertex_contact_capsules_only_and_penetration_scores_higher():
    n = len(Q2)
    dist = np.full(n, 1e8, np.float32)
    dist[3] = 0.5 * D1
    sid = np.full(n, -1, np.int32)
    sid[3] = 0
    cap = _run(Q2, T2, [0.0, 0.0], dist, sid, (R.CAPSULE,))
    plane = _run(Q2, T2, [0.0, 0.0], dist, sid, (0,))
    touching = [(0, 3), (1, 3), (2, 3)]
    for e in cap:
        assert (cap[e] > 0.0) == (e in touching)
        assert plane[e] == 0.0
    # closed form on (0, 3): (L / h)(L / L_max) * 0.5 with L_max = |x0 - x1| = 1
    L = float(np.linalg.norm(Q2[3].astype(float) - Q2[0].astype(float)))
    assert cap[(0, 3)] == pytest.approx((L / H_MIN) * (L / 1.0) * 0.5, rel=1e-6)
    dist[3] = -0.5 * D1
    pen = _run(Q2, T2, [0.0, 0.0], dist, sid)
    assert pen[(0, 3)] > cap[(0, 3)]


def test_tri_term_favours_edge_nearest_closest_point():
    q = np.array([(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.5, np.sqrt(3) / 2, 0.0), (0.5, 0.3, 0.8)], np.float32)
    tets = np.array([(0, 1, 2, 3)], np.int32)
    tris = np.array([(0, 1, 2)], np.int32)

    def tri_run(bary, d):
        return _run(q, tets, [0.0], tris=tris, tri_dist=np.array([d], np.float32),
                    tri_bary=np.array([bary], np.float32))

    sc = tri_run((0.0, 0.5, 0.5), 0.5 * D1)
    assert sc[(1, 2)] > sc[(0, 1)] and sc[(1, 2)] > sc[(0, 2)]
    assert sc[(0, 1)] == pytest.approx(sc[(0, 2)], rel=1e-5)
    assert sc[(1, 2)] == pytest.approx((1.0 / H_MIN) * 1.0 * 0.5 * 1.0, rel=1e-4)
    assert sc[(0, 3)] == 0.0
    sc = tri_run((1 / 3, 1 / 3, 1 / 3), 0.5 * D1)
    assert sc[(0, 1)] == pytest.approx(sc[(1, 2)], rel=1e-5) and sc[(0, 1)] == pytest.approx(sc[(0, 2)], rel=1e-5)
    assert all(v == 0.0 for v in tri_run((0.0, 0.5, 0.5), 2.0 * D1).values())
    assert tri_run((0.0, 0.5, 0.5), 0.0)[(1, 2)] == pytest.approx(2.0 * tri_run((0.0, 0.5, 0.5), 0.5 * D1)[(1, 2)],
                                                                   rel=1e-5)


def test_elastic_term_is_excess_over_mean_density():
    dm = _dminv(Q2, T2)
    rho = [R.tet_density(1.0, 1.0, dm[t]) for t in range(2)]
    assert rho[0] == pytest.approx(1.0 / _vol(Q2.astype(float), T2[0]), rel=1e-6)
    sc = _run(Q2, T2, [1.0, 1.0], total=rho[0] + rho[1])  # both tets at the mean: nothing
    assert all(v == 0.0 for v in sc.values())
    sc = _run(Q2, T2, [1.0, 1.0], total=rho[0])  # tet 0 at twice the mean: excess 1
    L = float(np.linalg.norm(Q2[0].astype(float) - Q2[1].astype(float)))
    assert sc[(0, 1)] == pytest.approx(L / H_MIN, rel=1e-6)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 70000])
def test_density_total_matches_fsum(n):
    rho = np.random.default_rng(n).uniform(0, 10, n)
    assert R.density_total(rho) == pytest.approx(math.fsum(rho), rel=1e-13)


def test_planted_sum_instead_of_max_breaks_the_valence_pin(monkeypatch):
    """A restatement that sums contributions (the legacy combination) fails
    the per-edge-maximum pin."""
    orig = max

    def summing_max(a, b, *rest):
        if isinstance(a, int) and isinstance(b, int) and not rest:
            return a + b
        return orig(a, b, *rest)

    monkeypatch.setattr(R, "max", summing_max, raising=False)
    one = _run(Q2, T2, [1.0, 0.0])
    both = _run(Q2, T2, [1.0, 1.0])
    assert both[(0, 1)] != pytest.approx(one[(0, 1)], rel=1e-5)
