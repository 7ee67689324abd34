"""Host side of batching: the seeded batch scene and the union layout."""
import numpy as np

from mfem_amd import mesh as M
from mfem_amd import scene
from mfem_amd.batch import union_model


def test_soft_body_batch_scene():
    models = scene.soft_body_batch()
    assert len(models) == 64
    T = np.array([m.tet_count for m in models])
    # lattice rounding keeps every body within ~15 % of the [2e3, 5e4] draw
    assert T.min() >= 1700 and T.max() <= 57500
    assert 1.1e6 < T.sum() < 1.3e6  # ~1.2e6 tets in all
    energies = {m.params["energy"] for m in models}
    assert energies == {"arap", "neohookean"}
    E = []
    for m in models:
        mu, lam = m.tet_materials[0, 0], m.tet_materials[0, 1]
        E.append(mu * (3 * lam + 2 * mu) / (lam + mu))
        # started above the plane, deformed from rest but not inverted
        assert m.particle_q[:, 2].min() > 1e-3
        Ds = np.stack([m.particle_q[m.tet_indices[:, k]] - m.particle_q[m.tet_indices[:, 0]] for k in (1, 2, 3)], 2)
        F = Ds @ m.tet_poses.reshape(-1, 3, 3)
        J = np.linalg.det(F)
        assert J.min() > 0.3 and np.abs(F - np.eye(3)).max() > 0.01
    assert 1e4 <= min(E) and max(E) <= 1e6
    # one shared barrier stiffness
    assert len({m.params["contact_stiffness"] for m in models}) == 1


def test_union_alignment():
    models = scene.soft_body_batch(n_bodies=6, seed=3, tets_min=200, tets_max=2000)
    for align in (256, 1024):
        u, bv, bt, bf, en = union_model(models, align=align)
        assert np.all(bv % align == 0) and bv[-1] == u.particle_count
        for i, m in enumerate(models):
            assert bv[i + 1] - bv[i] >= m.particle_count
            np.testing.assert_array_equal(u.tet_indices[bt[i]:bt[i + 1]], m.tet_indices + bv[i])
            # padding rows are pinned, massless and far from the plane
            pad = slice(bv[i] + m.particle_count, bv[i + 1])
            assert np.all(u.particle_inv_mass[pad] == 0) and np.all(u.particle_q[pad, 2] > 10)
        assert list(en[:bt[-1]]) == sum(([int(m.params["energy"] != "arap")] * m.tet_count for m in models), [])
