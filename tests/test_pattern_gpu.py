"""The GPU pattern builder (csrc/mfem_pattern.hip) against the host builder
(pattern.build_system_pattern / build_triangle_lists): every array equal,
element for element, on meshes of several shapes, including a refined one
and a 4-body batch with padding vertices."""
import numpy as np
import pytest
import torch

from mfem_amd import mesh as M
from mfem_amd import scene
from mfem_amd.pattern import GpuPatternBuilder, build_system_pattern, build_triangle_lists

pytestmark = pytest.mark.gpu


def _meshes():
    out = []
    m = scene.beam_scene()
    out.append(("beam", m.particle_count, m.tet_indices, m.tri_indices))
    m = scene.octopus_scene(h=0.02)
    out.append(("octopus_h20", m.particle_count, m.tet_indices, m.tri_indices))
    from mfem_amd.batch import union_model

    u = union_model(scene.soft_body_batch(n_bodies=4, seed=3, tets_min=300, tets_max=3000))[0]
    out.append(("batch_union", u.particle_count, u.tet_indices, u.tri_indices))
    # a refined grid (the specification's pass): rows of unequal length
    from oracle import refine as R

    p, t = M.box_tet_mesh((0, 0, 0), (0.6, 0.3, 0.3), (6, 3, 3))
    p, t = p.astype(np.float32), t.astype(np.int32)
    tris = M.surface_tris(t)
    rng = np.random.default_rng(0)
    sc = R.legacy_scores(t, p, np.ones(len(p), np.float32), rng.uniform(0, 5, len(t)).astype(np.float32),
                         rng.uniform(0, 30, len(p)).astype(np.float32), 0.3, 0.1, 1.0e6, 2.0)
    o = R.refine_pass(t, tris, p, p, p, np.zeros((len(t), 6)), np.zeros((len(t), 6)), np.ones((len(t), 3)),
                      R.select_edges(t, sc), len(p), 10 * len(t), 10 * len(p))
    out.append(("refined_grid", o["n_verts"], o["tets"], o["tris"]))
    return out


@pytest.mark.parametrize("case", range(4))
def test_gpu_pattern_equals_host(gpu, case):
    name, nv, tets, tris = _meshes()[case]
    tets = np.asarray(tets, np.int32)
    tris = np.asarray(tris, np.int32).reshape(-1, 3)
    host = build_system_pattern(nv, tets)
    tc_ptr, tc_src, vt_ptr, vt_src = build_triangle_lists(host, tris)
    b = GpuPatternBuilder(nv + 100, len(tets) + 50, len(tris) + 10, gpu)
    tg = torch.as_tensor(tets, device=gpu).contiguous()
    fg = torch.as_tensor(tris, device=gpu).contiguous()
    pat, out = b.build(tg, len(tets), fg, len(tris), nv)
    print(name, "verts", nv, "tets", len(tets), "nnz", len(host.cols), "slots", host.n_slots)
    assert pat.n_slots == host.n_slots and pat.n_chunks == host.n_chunks
    np.testing.assert_array_equal(pat.rowptr, host.rowptr)
    np.testing.assert_array_equal(pat.cols, host.cols)
    np.testing.assert_array_equal(pat.csr_lane, host.csr_lane)
    for k, ref in (("chunk_ptr", host.chunk_ptr), ("sell_col", host.sell_col), ("diag_lane", host.diag_lane),
                   ("contrib_ptr", host.contrib_ptr), ("contrib_src", host.contrib_src), ("vinc_ptr", host.vinc_ptr),
                   ("vinc_src", host.vinc_src), ("tcon_ptr", tc_ptr), ("tcon_src", tc_src), ("vtri_ptr", vt_ptr),
                   ("vtri_src", vt_src)):
        got = out[k].cpu().numpy()[:len(ref)]
        np.testing.assert_array_equal(got, np.asarray(ref), err_msg=f"{name}: {k}")


def test_gpu_pattern_rejects_foreign_triangle(gpu):
    p, t = M.box_tet_mesh((0, 0, 0), (0.2, 0.2, 0.2), (2, 2, 2))
    t = np.asarray(t, np.int32)
    bad = np.array([[0, 1, len(p) - 1]], np.int32)  # not a face of any tet
    b = GpuPatternBuilder(len(p), len(t), 1, gpu)
    with pytest.raises(ValueError):
        b.build(torch.as_tensor(t, device=gpu), len(t), torch.as_tensor(bad, device=gpu), 1, len(p))
