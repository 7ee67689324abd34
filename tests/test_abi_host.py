"""CPU checks: the C-ABI library loads and exports what include/mfem_amd.h
declares, the ctypes mirror matches the compiled struct layout, and the
host-side mesh / pattern logic is right.  No GPU compute is invoked."""
import os
import re

import numpy as np
import pytest

from mfem_amd import _ffi, mesh, pattern, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "mfem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(mfem_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _ffi.load()
    names = _declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(_ffi.ENTRY_POINTS)
    assert lib.mfem_version() >= 1


def test_struct_layout_matches_header():
    _ffi.check_layout()


def test_constants_match_header():
    # every #define the ctypes mirror restates has the header's value
    text = open(os.path.join(ROOT, "include", "mfem_amd.h")).read()
    defs = {k: int(v, 0) for k, v in re.findall(r"^#define\s+(MFEM_\w+)\s+\(?(-?(?:0x)?[0-9A-Fa-f]+)",
                                                 text, flags=re.M)}
    mirror = {"MFEM_STATS_PER_ITER": _ffi.MFEM_STATS_PER_ITER, "MFEM_MAX_STATS_ITERS": _ffi.MFEM_MAX_STATS_ITERS,
              "MFEM_RED_MAX_BLOCKS": _ffi.MFEM_RED_MAX_BLOCKS, "MFEM_PCG_PIPELINED": _ffi.PCG_PIPELINED,
              "MFEM_PCG_STREAM_ROWS": _ffi.PCG_STREAM_ROWS, "MFEM_PCG_LEARNED_GUESS": _ffi.PCG_LEARNED_GUESS,
              "MFEM_PCG_COARSE": _ffi.PCG_COARSE, "MFEM_PCG_COOPERATIVE": _ffi.PCG_COOPERATIVE,
              "MFEM_ERR_BARRIER": _ffi.MFEM_ERR_BARRIER, "MFEM_ERR_CONTACT": _ffi.MFEM_ERR_CONTACT}
    for k, v in mirror.items():
        assert k in defs, k
        assert defs[k] == v, (k, defs[k], v)


def test_errors_are_reported_not_raised_in_c():
    import ctypes
    lib = _ffi.load()
    S = _ffi.MfemSolverStruct()  # all zero: empty mesh
    rc = lib.mfem_step(ctypes.byref(S), ctypes.c_float(0.01), None)
    assert rc == -1
    assert b"empty" in lib.mfem_last_error()


def test_beam_mesh_counts_and_mass():
    m = scene.beam_scene()
    assert m.particle_count == 525 and m.tet_count == 1920
    assert int((m.particle_inv_mass == 0).sum()) == 25
    assert m.particle_mass.sum() == pytest.approx(40.0)
    vol = 1.0 / (6.0 * np.linalg.det(m.tet_poses))
    assert np.all(vol > 0) and vol.sum() == pytest.approx(0.04)
    assert m.tri_indices.shape == (704, 3)


def _surface_volume(p, tris):
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def test_surface_is_closed_and_outward():
    for m in (scene.beam_scene(), scene.drop_box_scene()):
        vol = (1.0 / (6.0 * np.linalg.det(m.tet_poses))).sum()
        assert _surface_volume(m.particle_q, m.tri_indices) == pytest.approx(vol, rel=1e-10)
        e = np.sort(np.concatenate([m.tri_indices[:, [0, 1]], m.tri_indices[:, [1, 2]], m.tri_indices[:, [2, 0]]]), 1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)  # watertight


def test_octopus_mesh_size():
    m = scene.octopus_scene()
    assert 1.9e5 <= m.tet_count <= 2.1e5
    assert 4.0e4 <= m.particle_count <= 5.0e4
    assert m.particle_q[:, 2].min() == pytest.approx(2e-3)
    E = m.tet_materials[:, 0] * 2 * 1.45
    assert 4.0e4 < E.mean() < 6.0e4  # log-normal field with mean 5e4
    assert 0.35 < np.std(np.log(E)) < 0.65


def test_system_pattern_roundtrip():
    m = scene.beam_scene(cells=(5, 3, 2))
    pat = pattern.build_system_pattern(m.particle_count, m.tet_indices)
    nv, T = m.particle_count, m.tet_count
    # SELL columns reproduce the CSR pattern; diagonal lanes point at (i, i)
    rows = np.repeat(np.arange(nv), np.diff(pat.rowptr))
    np.testing.assert_array_equal(pat.sell_col[pat.csr_lane], pat.cols)
    np.testing.assert_array_equal(pat.sell_col[pat.diag_lane], np.arange(nv))
    # every (tet, a, b) block lands exactly once: 16 per tet
    assert pat.contrib_ptr[-1] == 16 * T
    # dense re-assembly through the gather lists equals direct scatter
    rng = np.random.default_rng(0)
    Kb = rng.normal(size=(10, T, 3, 3))
    val = np.zeros((pat.n_slots * 64, 3, 3))
    for l in range(pat.n_slots * 64):
        for k in range(pat.contrib_ptr[l], pat.contrib_ptr[l + 1]):
            src = pat.contrib_src[k]
            B = Kb.reshape(-1, 3, 3)[src >> 1]
            val[l] += B.T if src & 1 else B
    dense = np.zeros((3 * nv, 3 * nv))
    for e, t in enumerate(m.tet_indices):
        for p, (a, b) in enumerate(pattern.PAIRS):
            dense[3 * t[a]:3 * t[a] + 3, 3 * t[b]:3 * t[b] + 3] += Kb[p, e]
            if a != b:
                dense[3 * t[b]:3 * t[b] + 3, 3 * t[a]:3 * t[a] + 3] += Kb[p, e].T
    sell = np.zeros(pat.n_slots * 576)
    sell.reshape(pat.n_slots, 9, 64)[:] = val.reshape(pat.n_slots, 64, 9).transpose(0, 2, 1)
    import scipy.sparse as sp
    csr = sp.bsr_matrix((pattern.sell_to_csr_values(pat, sell), pat.cols, pat.rowptr), shape=dense.shape)
    np.testing.assert_allclose(csr.toarray(), dense, atol=1e-12)
    # vertex incidence covers every tet corner once, in order
    assert pat.vinc_ptr[-1] == 4 * T
    src = pat.vinc_src
    np.testing.assert_array_equal(m.tet_indices.reshape(-1)[src], np.repeat(np.arange(nv), np.diff(pat.vinc_ptr)))


def test_kuhn_mesh_conforming():
    p, t = mesh.box_tet_mesh((0, 0, 0), (1, 1, 1), (3, 3, 3))
    faces = np.sort(np.concatenate([t[:, [0, 1, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 3]], t[:, [1, 2, 3]]]), 1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() <= 2  # every interior face shared by exactly two tets
    assert len(p) == 64 and len(t) == 162
