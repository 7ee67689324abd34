"""ctypes binding of libmfem_amd.so (declared in include/mfem_amd.h).

The library is the only compute path: importing this module on a machine
without the built library raises, and nothing falls back to a CPU
implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfem_amd.so")

c_i32, c_f32, c_ptr = ctypes.c_int32, ctypes.c_float, ctypes.c_void_p

# (name, ctype) in declaration order of struct mfem_solver
FIELDS = [
    ("n_verts", c_i32), ("n_tets", c_i32), ("n_tris", c_i32), ("n_shapes", c_i32), ("n_chunks", c_i32),
    ("n_slots", c_i32), ("energy", c_i32), ("iterations", c_i32), ("cg_maxiter", c_i32), ("precond", c_i32),
    ("line_search", c_i32), ("ls_max", c_i32), ("tri_contact", c_i32), ("friction_on", c_i32),
    ("cg_tol", c_f32), ("ls_alpha0", c_f32), ("ls_tau", c_f32), ("ls_c", c_f32), ("ls_thr", c_f32),
    ("sing_thr", c_f32), ("attach_k", c_f32), ("d0", c_f32), ("d1", c_f32), ("kappa", c_f32), ("eps_v", c_f32),
    ("qb", c_f32 * 3), ("gravity", c_f32 * 3), ("max_width", c_i32), ("pcg_flags", c_i32), ("pcg_variant", c_i32),
] + [(n, c_ptr) for n in (
    "tets", "dminv", "materials", "vol", "mass", "inv_mass", "pin_pos", "tris", "shape_type", "shape_xf",
    "shape_scale", "shape_mu", "chunk_ptr", "sell_col", "sell_val", "diag_lane", "contrib_ptr", "contrib_src",
    "vinc_ptr", "vinc_src", "x", "v", "s", "lam", "x_tilde", "x_start", "lag_n", "lag_fn", "lag_sid",
    "G", "Hl", "gl", "Hsinv", "gs", "psi", "Kb", "fe", "gx", "gB", "gD", "Hv", "dist", "eB", "eD", "g",
    "dx", "ds", "pz", "px", "red", "flags", "status", "stats", "mlp_w", "mlp_wbf", "ws_x", "ws_aux", "ws_y", "guess",
    "tcon_ptr", "tcon_src", "vtri_ptr", "vtri_src", "tri_dist", "tri_bary", "tri_nrm", "tri_sid",
)] + [("n_bodies", c_i32), ("ls_blocks", c_i32)] + [(n, c_ptr) for n in (
    "body_wg", "body_vert", "body_tet", "body_tri", "body_ls_wg", "body_epoch", "tet_energy",
)] + [("body_wg_align", c_i32), ("local_sweeps", c_i32), ("local_omega", c_f32), ("dx_hist", c_ptr), ("dx_same", c_ptr), ("dx_same_iters", c_i32),
   ("coarse_w", c_ptr), ("coarse_omega", c_f32), ("cw_x", c_ptr), ("cw_y", c_ptr), ("cw_w", c_ptr),
   ("cw_wbf", c_ptr)] + [(n, c_i32) for n in (
    "self_contact", "n_sc_points", "n_sc_edges", "sc_max_pairs", "sc_max_span",
)] + [(n, c_f32) for n in ("sc_dhat", "sc_kappa", "sc_cell", "sc_ccd_s")] + [(n, c_ptr) for n in (
    "sc_points", "sc_edges", "sc_pt", "sc_ee", "sc_count", "sc_g", "sc_h", "sc_alpha", "sc_ws",
)] + [("sc_ws_bytes", ctypes.c_int64), ("shell_on", c_i32), ("shell_mu", c_f32)] + [(n, c_ptr) for n in (
    "rest", "tri_W", "tri_area", "shell_h", "gS",
)] + [("fused_local", c_i32)]


class MfemSolverStruct(ctypes.Structure):
    _fields_ = FIELDS


# mirrors of the header's constants
MFEM_STATS_PER_ITER = 16
MFEM_MAX_STATS_ITERS = 64
MFEM_RED_MAX_BLOCKS = 2048
SHAPE_PLANE, SHAPE_CAPSULE = 0, 1
PCG_PIPELINED, PCG_STREAM_ROWS, PCG_LEARNED_GUESS, PCG_COARSE, PCG_COOPERATIVE = 1, 2, 4, 8, 16
MFEM_MLP_NPARAMS = 6288
MFEM_SCAN_NOALLOC_MAX = 1 << 26
MFEM_ERR_BARRIER, MFEM_ERR_CONTACT = 1, 2

ENTRY_POINTS = [
    "mfem_solver_sizeof", "mfem_solver_offsetof", "mfem_version", "mfem_wall_clock_khz", "mfem_last_error",
    "mfem_init_stretch",
    "mfem_begin_step", "mfem_update_system", "mfem_assemble", "mfem_global_solve", "mfem_local_solve",
    "mfem_line_search", "mfem_end_step", "mfem_step", "mfem_spmv", "mfem_merit", "mfem_ws_features",
    "mfem_ws_apply", "mfem_mlp_forward", "mfem_mlp_train_step", "mfem_mlp_backward_step", "mfem_mlp_last_error",
    "mfem_scan_i32", "mfem_scan_workspace_bytes", "mfem_broadphase_pt", "mfem_broadphase_ee",
    "mfem_broadphase_workspace_bytes", "mfem_soft_rigid_distance", "mfem_pair_barrier", "mfem_collide_last_error",
    "mfem_refine_workspace_bytes", "mfem_refine_select", "mfem_refine_count", "mfem_refine_apply",
    "mfem_refine_last_error", "mfem_pcg_variant_count", "mfem_pcg_variant_info", "mfem_line_search_max_blocks",
    "mfem_contact_distance", "mfem_pattern_workspace_bytes", "mfem_pattern_count", "mfem_pattern_fill",
    "mfem_pattern_last_error", "mfem_pose_shapes", "mfem_scan_i32_noalloc", "mfem_coarse_features",
    "mfem_coarse_weights", "mfem_broadphase_swept", "mfem_self_contact_workspace_bytes",
    "mfem_self_contact_candidates", "mfem_self_contact_eval", "mfem_self_contact_ccd",
    "mfem_self_contact_last_error", "mfem_last_pcg_variant", "mfem_shell_refresh",
    "mfem_surface_knn", "mfem_surface_project", "mfem_surface_last_error",
]


class MfemError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libmfem_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MfemError(f"{path} not found: build it with `python warp-mfem_amd/build.py`")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(MfemSolverStruct)
    vp = ctypes.c_void_p
    lib.mfem_solver_sizeof.restype = ctypes.c_int64
    lib.mfem_solver_offsetof.restype = ctypes.c_int64
    lib.mfem_solver_offsetof.argtypes = [ctypes.c_char_p]
    lib.mfem_version.restype = ctypes.c_int
    lib.mfem_last_error.restype = ctypes.c_char_p
    for name in ("mfem_init_stretch", "mfem_local_solve"):
        getattr(lib, name).argtypes = [P, vp]
    for name in ("mfem_begin_step", "mfem_update_system", "mfem_assemble", "mfem_end_step", "mfem_step"):
        getattr(lib, name).argtypes = [P, ctypes.c_float, vp]
    lib.mfem_global_solve.argtypes = [P, ctypes.c_int, vp]
    lib.mfem_line_search.argtypes = [P, ctypes.c_float, ctypes.c_int, vp]
    lib.mfem_spmv.argtypes = [P, vp, vp, vp]
    lib.mfem_merit.argtypes = [P, ctypes.c_float, ctypes.c_float, vp, vp]
    lib.mfem_contact_distance.argtypes = [P, vp, vp, vp]
    i = ctypes.c_int
    lib.mfem_pattern_workspace_bytes.restype = ctypes.c_size_t
    lib.mfem_pattern_workspace_bytes.argtypes = [i, i, i]
    lib.mfem_pattern_count.argtypes = [vp, i, i, i, i, i, vp, ctypes.c_size_t, vp, vp, vp, vp]
    lib.mfem_pattern_fill.argtypes = [vp, i, vp, i, i, i, i, i, vp, ctypes.c_size_t, vp, vp, i] + [vp] * 13
    lib.mfem_pattern_last_error.restype = ctypes.c_char_p
    lib.mfem_ws_features.argtypes = [P, ctypes.c_float, ctypes.c_int, vp]
    lib.mfem_ws_apply.argtypes = [P, vp]
    lib.mfem_coarse_features.argtypes = [P, ctypes.c_float, vp]
    lib.mfem_coarse_weights.argtypes = [P, vp]
    lib.mfem_mlp_forward.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
    lib.mfem_mlp_train_step.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp,
                                        ctypes.c_float, ctypes.c_int, vp]
    lib.mfem_mlp_backward_step.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp, ctypes.c_float,
                                           ctypes.c_int, vp]
    lib.mfem_wall_clock_khz.restype = ctypes.c_int
    for name in ENTRY_POINTS[5:]:
        getattr(lib, name).restype = ctypes.c_int
    lib.mfem_mlp_last_error.restype = ctypes.c_char_p
    i, f, sz = ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    lib.mfem_scan_i32.argtypes = [vp, vp, i, i, vp, vp, sz, vp]
    lib.mfem_scan_workspace_bytes.argtypes = [i]
    lib.mfem_scan_workspace_bytes.restype = sz
    lib.mfem_scan_i32_noalloc.argtypes = [vp, vp, i, i, vp, vp]
    lib.mfem_broadphase_pt.argtypes = [vp, vp, i, vp, i, f, f, i, vp, i, vp, vp, vp, sz, vp]
    lib.mfem_broadphase_ee.argtypes = [vp, vp, i, f, f, i, vp, i, vp, vp, vp, sz, vp]
    lib.mfem_broadphase_workspace_bytes.argtypes = [i, i, i]
    lib.mfem_broadphase_workspace_bytes.restype = sz
    lib.mfem_soft_rigid_distance.argtypes = [vp, vp, i, i, vp, vp, i, f, vp, vp]
    lib.mfem_pair_barrier.argtypes = [i, vp, vp, vp, i, vp, vp, f, f, vp, vp, vp, vp, vp]
    lib.mfem_collide_last_error.restype = ctypes.c_char_p
    lib.mfem_refine_workspace_bytes.argtypes = [i, i]
    lib.mfem_refine_workspace_bytes.restype = sz
    lib.mfem_refine_select.argtypes = [vp, vp, sz, vp, vp, i, vp]
    lib.mfem_refine_count.argtypes = [vp, vp, sz, vp, vp]
    lib.mfem_refine_apply.argtypes = [vp, vp, sz, vp, vp]
    lib.mfem_refine_last_error.restype = ctypes.c_char_p
    lib.mfem_pcg_variant_count.argtypes = []
    lib.mfem_line_search_max_blocks.argtypes = []
    lib.mfem_pcg_variant_info.argtypes = [P, i, vp]
    lib.mfem_pose_shapes.argtypes = [vp, i, vp, vp, i, vp, vp]
    lib.mfem_broadphase_swept.argtypes = [i, vp, vp, vp, i, vp, i, f, f, i, vp, i, vp, vp, vp, sz, vp]
    lib.mfem_self_contact_workspace_bytes.argtypes = [i, i, i, i]
    lib.mfem_self_contact_workspace_bytes.restype = sz
    for name in ("mfem_self_contact_eval", "mfem_self_contact_ccd"):
        getattr(lib, name).argtypes = [P, vp]
    lib.mfem_self_contact_candidates.argtypes = [P, i, vp]
    lib.mfem_self_contact_last_error.restype = ctypes.c_char_p
    lib.mfem_last_pcg_variant.argtypes = []
    lib.mfem_shell_refresh.argtypes = [P, vp]
    lib.mfem_surface_knn.argtypes = [vp, i, vp, vp, i, i, vp, vp, vp]
    lib.mfem_surface_project.argtypes = [vp, i, vp, vp, vp, i, vp, vp, vp, i, vp, vp, vp, vp]
    lib.mfem_surface_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check_layout(lib: ctypes.CDLL | None = None) -> None:
    """Raise if the ctypes mirror disagrees with the compiled struct layout."""
    lib = lib or load()
    if lib.mfem_solver_sizeof() != ctypes.sizeof(MfemSolverStruct):
        raise MfemError("mfem_solver size mismatch")
    for name, _ in FIELDS:
        off = lib.mfem_solver_offsetof(name.encode())
        if off != getattr(MfemSolverStruct, name).offset:
            raise MfemError(f"mfem_solver.{name} offset mismatch ({off})")


def call(fn_name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, fn_name)(*args)
    if rc != 0:
        raise MfemError(f"{fn_name} failed ({rc}): {lib.mfem_last_error().decode()}")
