"""RefinementSolver: drop-in for warp-mfem's per-timestep soft-body solve.

Mirrors ``mfem.refinement.solver.RefinementSolver``
(src/mfem/refinement/solver.py:40-985): same constructor arguments and
keyword names (solver.py:56-128, 185, 241-245, 265, 297), same
``step(state_in, state_out, control, contacts, dt)`` contract (solver.py:429-579)
and the same device-side failure semantics (CG non-convergence and a rejected
line search do not raise; they are visible in ``last_stats()``).

All compute runs in libmfem_amd.so (HIP, gfx950) through the C ABI in
include/mfem_amd.h; torch provides device memory and the stream.
"""
from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _ffi
from .pattern import GpuPatternBuilder

_ENERGIES = {"arap": 0, "neohookean": 1}


@dataclass
class State:
    """Per-step state (newton State fields used by the solver)."""

    particle_q: torch.Tensor   # (V, 3) float32, device
    particle_qd: torch.Tensor  # (V, 3) float32, device
    body_q: torch.Tensor | None = None  # (n_bodies, 7) rigid body poses p, q(xyzw), or None = keep the poses

    def assign(self, other: "State") -> None:
        self.particle_q.copy_(other.particle_q)
        self.particle_qd.copy_(other.particle_qd)
        if other.body_q is not None:
            if self.body_q is None:
                self.body_q = other.body_q.clone()
            else:
                self.body_q.copy_(other.body_q)

    def clone(self) -> "State":
        return State(self.particle_q.clone(), self.particle_qd.clone(),
                     None if self.body_q is None else self.body_q.clone())


def model_state(model, device="cuda", capacity: int | None = None) -> State:
    """A State holding the model's initial positions / velocities, padded to
    ``capacity`` rows (the solver's max_particles when refinement may add
    vertices)."""
    q0 = np.asarray(model.particle_q, dtype=np.float32).reshape(-1, 3)
    qd0 = np.asarray(model.particle_qd, dtype=np.float32).reshape(-1, 3)
    n = len(q0) if capacity is None else int(capacity)
    if n < len(q0):
        raise ValueError("capacity < model.particle_count")
    q = torch.zeros(n, 3, dtype=torch.float32, device=device)
    qd = torch.zeros(n, 3, dtype=torch.float32, device=device)
    q[:len(q0)] = torch.as_tensor(q0, device=device)
    qd[:len(qd0)] = torch.as_tensor(qd0, device=device)
    return State(q, qd)


def _barrier_at(d, d0, d1):
    """barrier(d) and its first two derivatives (src/mfem/refinement/contact.py:16-31),
    evaluated on the log branch; used for the quadratic extension at d0."""
    if not (d0 < d < d1):
        return (0.0, 0.0, 0.0)
    ll = math.log(d / d1)
    e = -(d - d1) ** 2 * ll
    de = -(2.0 * (d - d1) * ll + (d - d1) ** 2 / d)
    d2e = -(2.0 * ll + 4.0 * (d - d1) / d - (d - d1) ** 2 / (d * d))
    return (e, de, d2e)


def _surface_edges(tris: np.ndarray) -> np.ndarray:
    """Unique edges (i < j) of triangles (F, 3), sorted."""
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    if len(tris) == 0:
        return np.zeros((0, 2), np.int32)
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0).astype(np.int32)


class RefinementSolver:
    def __init__(self, model, iterations: int, max_tets: int, max_particles: int | None = None,
                 max_tris: int | None = None, refine_every_n_steps: int = 10, min_refine_score: float = 10.1,
                 device: str = "cuda", **kwargs):
        self._refine_enabled = bool(kwargs.get("enable_refinement", True))
        self._refine_scoring = kwargs.get("refine_scoring", "legacy")
        if self._refine_scoring not in ("legacy", "geometric"):
            raise ValueError(f"Unknown refine_scoring {self._refine_scoring!r} (expected 'legacy' or 'geometric')")
        # membrane shell over the surface triangles (solver.py:291-302): a
        # surface shear modulus in Pa m, 0 disables
        self._shell_mu = float(kwargs.get("shell_mu", 0.0))
        energy = kwargs.get("energy", "arap")
        if energy not in _ENERGIES:
            raise ValueError(f"Unknown energy type: {energy}")
        nt = int(model.tet_count)
        # the model's particle buffer is already padded to max_particles by
        # the caller (solver.py:130-136); the active particles are those up
        # to the largest tet vertex (AdditionalState.from_model, get_active_particles)
        if max_particles is not None and max_particles != model.particle_count:
            raise ValueError(
                f"max_particles={max_particles} does not match model.particle_count={model.particle_count}; the "
                "model's particle buffer must already be padded to max_particles by the caller (e.g. "
                "MFEMRefinementModel.load_sim_model(..., max_particles=...) or scene.pad_particles) before the "
                "solver is constructed")
        nv = int(np.asarray(model.tet_indices).max()) + 1 if nt > 0 else 0
        nv = int(getattr(model, "active_particle_count", nv))  # a layout may declare its active rows
        if max_tets < nt:
            raise ValueError(f"max_tets={max_tets} < model.tet_count={nt}")
        tris = np.zeros((0, 3), np.int32) if model.tri_indices is None else np.asarray(model.tri_indices, np.int32)
        tris = tris.reshape(-1, 3)
        Vc = int(model.particle_count)
        Tc = int(max_tets)
        # surface triangles grow with the tets (same headroom when not given)
        Fc = int(max_tris) if max_tris is not None else int(np.ceil(len(tris) * Tc / max(nt, 1)))
        if Fc < len(tris):
            raise ValueError(f"max_tris={Fc} < the model's {len(tris)} surface triangles")
        if iterations > _ffi.MFEM_MAX_STATS_ITERS:
            raise ValueError(f"iterations must be <= {_ffi.MFEM_MAX_STATS_ITERS}")
        self.lib = _ffi.load()
        _ffi.check_layout(self.lib)
        self.model = model
        self.iterations = int(iterations)
        self.max_tets, self.max_particles, self.max_tris = Tc, Vc, Fc
        self.refine_every_n_steps = refine_every_n_steps
        # the geometric score is dimensionless: its own threshold (solver.py:141-142)
        if self._refine_scoring == "geometric":
            min_refine_score = float(kwargs.get("refine_geometric_threshold", 1.0))
        self.min_refine_score = min_refine_score
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self._step = -1
        self._timings: dict[str, list[float]] = {}

        shapes = list(getattr(model, "shapes", []))
        n_shapes = len(shapes)
        mu = np.asarray(kwargs.get("friction_mu", 0.0), dtype=np.float64)
        if mu.ndim == 0:
            mu = np.full(n_shapes, float(mu))
        elif mu.shape[0] != n_shapes:
            raise ValueError(f"friction_mu has {mu.shape[0]} entries but the model has {n_shapes} shapes")
        tri_contact = bool(kwargs.get("tri_contact", True))
        d0 = float(kwargs.get("contact_d0", 1.0e-2))
        d1 = float(kwargs.get("contact_d1", 1.0e-1))
        qb = _barrier_at(d0, 0.0, d1)

        dev = self.device
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        zf = lambda *shape: torch.zeros(*shape, dtype=torch.float32, device=dev)

        tets = np.asarray(model.tet_indices, dtype=np.int32).reshape(-1, 4)
        dminv = np.asarray(model.tet_poses, dtype=np.float64).reshape(-1, 3, 3)
        vol = 1.0 / (6.0 * np.linalg.det(dminv))
        inv_mass = np.asarray(model.particle_inv_mass, dtype=np.float64)
        pin_pos = np.where((inv_mass == 0)[:, None], np.asarray(model.particle_q, dtype=np.float64), 0.0)

        def pad(a, n, dtype=np.float32, fill=0):
            a = np.asarray(a, dtype=dtype)
            out = np.full((n,) + a.shape[1:], fill, dtype=dtype)
            out[:len(a)] = a
            return out

        # every per-vertex / per-tet / per-triangle buffer is allocated at its
        # capacity (refinement appends); the struct carries the active counts
        b = {}
        b["tets"] = i32(pad(tets, Tc, np.int32))
        b["dminv"] = f32(pad(dminv.reshape(-1, 9), Tc))
        b["materials"] = f32(pad(np.asarray(model.tet_materials).reshape(-1, 3), Tc))
        b["vol"] = f32(pad(vol, Tc))
        b["mass"] = f32(pad(model.particle_mass, Vc))
        b["inv_mass"] = f32(pad(inv_mass, Vc))
        b["pin_pos"] = f32(pad(pin_pos, Vc))
        b["rest"] = f32(pad(np.asarray(model.particle_q).reshape(-1, 3), Vc))  # rest positions (refinement)
        b["tris"] = i32(pad(tris, max(Fc, 1), np.int32))
        b["shape_type"] = i32([sh.type for sh in shapes] or [0])
        # shape poses relative to their rigid bodies (newton's shape_transform
        # / shape_body; body -1 = static, the pose is the world pose), and the
        # world poses for the model's initial body poses
        shape_body = [int(getattr(sh, "body", -1)) for sh in shapes]
        self._n_rigid = max(shape_body, default=-1) + 1
        model_bq = getattr(model, "body_q", None)
        if self._n_rigid > 0 and (model_bq is None or np.asarray(model_bq).reshape(-1, 7).shape[0] < self._n_rigid):
            raise ValueError("the model's shapes are attached to rigid bodies but model.body_q lacks their poses")
        b["shape_body"] = i32(shape_body or [-1])
        b["shape_local"] = f32(np.array([np.concatenate([sh.pos, sh.quat]) for sh in shapes]) if shapes
                               else np.zeros((1, 7)))
        world = np.zeros((1, 7))
        if shapes:
            from .scene import shape_world_poses
            world = shape_world_poses(model)
        b["shape_xf"] = f32(world)
        b["shape_scale"] = f32(np.array([sh.scale for sh in shapes]) if shapes else np.zeros((1, 3)))
        b["shape_mu"] = f32(mu if n_shapes else np.zeros(1))
        b["s"] = zf(Tc, 6)
        b["lam"] = zf(Tc, 6)
        b["x_tilde"] = zf(Vc, 3)
        b["x_start"] = zf(Vc, 3)
        b["lag_n"] = zf(Vc, 3)
        b["lag_fn"] = zf(Vc)
        b["lag_sid"] = torch.full((Vc,), -1, dtype=torch.int32, device=dev)
        b["G"] = zf(Tc, 72)
        b["Hl"] = zf(Tc, 21)
        b["gl"] = zf(Tc, 6)
        b["Hsinv"] = zf(Tc, 21)
        b["gs"] = zf(Tc, 6)
        b["psi"] = zf(Tc)
        b["Kb"] = zf(10 * Tc * 9)  # [10][n_tets][9] for the active n_tets
        b["fe"] = zf(Tc, 12)
        for n in ("gx", "gB", "gD", "g"):
            b[n] = zf(Vc, 3)
        b["Hv"] = zf(Vc, 9)
        for n in ("dist", "eB", "eD"):
            b[n] = zf(Vc)
        # dx and its history: the PCG reads them whole in 64-byte lines over
        # n_chunks * 64 rows, so the allocations are padded to a multiple of
        # 64 rows (dx is a view of its first Vc rows)
        Vpad = -(-Vc // 64) * 64
        b["dx"] = torch.zeros(Vpad, 3, dtype=torch.float64, device=dev)[:Vc]
        b["dx_hist"] = torch.zeros(2 * Vpad * 3, dtype=torch.float64, device=dev)
        b["dx_same"] = torch.zeros(int(iterations) * Vpad * 3, dtype=torch.float64, device=dev)
        b["ds"] = zf(Tc, 6)
        b["red"] = torch.zeros(2 * 2 * _ffi.MFEM_RED_MAX_BLOCKS * 8, dtype=torch.float64, device=dev)
        b["flags"] = torch.zeros(_ffi.MFEM_RED_MAX_BLOCKS, dtype=torch.int32, device=dev)
        b["status"] = torch.zeros(8, dtype=torch.int32, device=dev)
        b["stats"] = torch.zeros(_ffi.MFEM_MAX_STATS_ITERS * _ffi.MFEM_STATS_PER_ITER, dtype=torch.float64, device=dev)
        b["tri_dist"] = torch.full((max(Fc, 1),), 1.0e8, dtype=torch.float32, device=dev)
        b["tri_bary"] = zf(max(Fc, 1), 3)
        b["tri_nrm"] = zf(max(Fc, 1), 3)
        b["tri_sid"] = torch.full((max(Fc, 1),), -1, dtype=torch.int32, device=dev)
        # learned warm start workspace (features, candidates, outputs, guess)
        b["ws_x"] = torch.zeros(Vc, 16, dtype=torch.int16, device=dev)
        b["ws_aux"] = zf(Vc, 10)
        b["ws_y"] = zf(Vc, 3)
        b["guess"] = zf(Vc, 3)
        if self._shell_mu > 0.0:
            # shell rest frames, areas and gradient (shell_energy.py:248-259)
            b["tri_W"] = zf(max(Fc, 1), 4)
            b["tri_area"] = zf(max(Fc, 1))
            b["gS"] = zf(Vc, 3)
        self.buffers = b
        # attachments: the active particles with inv_mass = 0, fixed at
        # construction (padding rows and vertices added by refinement are
        # never pinned; solver.py:320-393, 395-425)
        pinned = np.zeros(Vc, np.bool_)
        pinned[:nv] = inv_mass[:nv] == 0
        self._pinned = torch.as_tensor(pinned, device=dev)
        self._refine_density = float(kwargs.get("refine_density", 1.0))
        self._refine_kw = dict(
            threshold=float(min_refine_score),
            w_tet=float(kwargs.get("refine_tet_score_weight", 0.0)),
            w_vertex=float(kwargs.get("refine_vertex_score_weight", 0.1)),
            penetrating_score=float(kwargs.get("penetrating_edge_score", 1.0e6)),
            split_t=float(kwargs.get("refine_split_position", 0.5)))
        if self._refine_scoring == "geometric":
            # solver.py:71-87: min edge length defaults to the barrier range
            min_edge = kwargs.get("refine_min_edge_length", None)
            self._refine_kw.update(
                scoring="geometric", contact_d1=float(kwargs.get("contact_d1", 1.0e-1)),
                min_edge_length=float(min_edge) if min_edge is not None else float(kwargs.get("contact_d1", 1.0e-1)),
                w_elastic=float(kwargs.get("refine_elastic_weight", 1.0)),
                w_vertex_contact=float(kwargs.get("refine_contact_weight", 1.0)),
                w_tri_contact=float(kwargs.get("refine_tri_contact_weight", 1.0)))
        self._pass_counter = 0
        self._refiner = None
        self.refine_log: list[dict] = []

        S = _ffi.MfemSolverStruct()
        S.n_shapes = n_shapes
        # solver-specific knobs (not reference kwargs): pipelined PCG, resident
        # rows, cooperative launches of the persistent kernels (off: ordinary
        # launches within the resident workgroup count, no idle gap before them)
        S.pcg_flags = (_ffi.PCG_PIPELINED if kwargs.get("pcg_pipelined", False) else 0) | \
                      (0 if kwargs.get("pcg_resident_rows", True) else _ffi.PCG_STREAM_ROWS) | \
                      (_ffi.PCG_COOPERATIVE if kwargs.get("cooperative_launch", False) else 0)
        S.pcg_variant = -1  # chosen from pcg_flags until autotune() picks one
        self.set_local_preconditioner(int(kwargs.get("pcg_local_sweeps", 0)),
                                      float(kwargs.get("pcg_local_omega", 0.5)), S)
        # workgroup-local rigid-mode correction on top of block Jacobi (no
        # counterpart in the reference): off unless asked for
        S.coarse_omega = float(kwargs.get("pcg_coarse_omega", 0.0))
        if kwargs.get("pcg_coarse", False):
            S.pcg_flags |= _ffi.PCG_COARSE | _ffi.PCG_PIPELINED
        S.energy = _ENERGIES[energy]
        S.iterations = self.iterations
        S.dx_same_iters = self.iterations
        S.cg_maxiter = int(kwargs.get("cg_max_iterations", 1000))
        S.precond = int(bool(kwargs.get("preconditioner", True)))
        S.line_search = int(bool(kwargs.get("line_search", True)))
        # the local solve inside the line search's first pass (no counterpart
        # in the reference, whose stages store and re-read the element
        # records): on unless asked off, and only with the line search
        S.fused_local = int(bool(kwargs.get("fuse_local_solve", True)) and S.line_search)
        S.ls_max = int(kwargs.get("line_search_max_iterations", 20))
        # Contact(tri_contact=...) keeps it only with triangles (contact.py:844);
        # it measures triangles against capsules, so without one it is skipped
        self._has_capsule = any(int(sh.type) == _ffi.SHAPE_CAPSULE for sh in shapes)
        S.tri_contact = int(tri_contact and len(tris) > 0 and self._has_capsule)
        self._tri_contact_kw = tri_contact
        S.friction_on = int(n_shapes > 0 and float(mu.max(initial=0.0)) > 0.0)
        S.cg_tol = float(kwargs.get("cg_tolerance", 1.0e-6))
        S.ls_alpha0 = float(kwargs.get("line_search_alpha0", 1.0))
        S.ls_tau = float(kwargs.get("line_search_tau", 0.5))
        S.ls_c = float(kwargs.get("line_search_c", 0.01))
        S.ls_thr = float(kwargs.get("line_search_threshold", 1.0e-8))
        S.sing_thr = float(kwargs.get("preconditioner_singular_threshold", 1.0e-20))
        S.attach_k = float(kwargs.get("attachment_stiffness", 1.0e1))
        S.d0, S.d1 = d0, d1
        S.kappa = float(kwargs.get("contact_stiffness", 1.0e7))
        S.eps_v = float(kwargs.get("friction_eps_v", 1.0e-3))
        S.qb[:] = qb
        S.gravity[:] = [float(g) for g in np.asarray(model.gravity).reshape(3)]
        S.shell_mu = self._shell_mu
        for name, t in b.items():
            if hasattr(S, name):
                setattr(S, name, t.data_ptr())
        self._S = S
        self._init_self_contact(kwargs, tris, Vc, Fc, d1)
        # device error bits as of the last step that has completed, copied
        # into pinned host memory at the end of every step (no host sync)
        self._status_mirror = torch.zeros(1, dtype=torch.int32, pin_memory=self.device.type == "cuda")
        self._pattern_builder = None
        self._set_topology(nv, nt, len(tris))
        self.launch_log = None  # a list: every step's PCG launches (profiling)
        self.warm_start_net = None
        self.set_warm_start(kwargs.get("warm_start"))
        self.coarse_net = None
        self._coarse_w = None

        # AdditionalState.from_model: rest stretch from the model's positions
        x0 = b["rest"].clone()
        self._bind_state(x0, zf(Vc, 3))
        self._call("mfem_init_stretch", self._stream())

    # -- topology -------------------------------------------------------------
    def _set_topology(self, n_verts: int, n_tets: int, n_tris: int) -> None:
        """(Re)builds the system pattern and the gather lists of the active
        mesh (on the GPU, csrc/mfem_pattern.hip: once per topology) and points
        the struct at them."""
        b, dev = self.buffers, self.device
        if self._pattern_builder is None:
            self._pattern_builder = GpuPatternBuilder(self.max_particles, self.max_tets, self.max_tris, dev)
        pat, out = self._pattern_builder.build(b["tets"], n_tets, b["tris"], n_tris, n_verts)
        self.pattern = pat
        for k in ("chunk_ptr", "sell_col", "diag_lane", "contrib_ptr", "contrib_src", "vinc_ptr", "vinc_src",
                  "tcon_ptr", "tcon_src", "vtri_ptr", "vtri_src"):
            b[k] = out[k]
        b["sell_val"] = torch.zeros(pat.n_slots * 576, dtype=torch.float32, device=dev)
        b["pz"] = torch.zeros(2, pat.n_chunks * 64, 3, dtype=torch.float32, device=dev)
        b["px"] = torch.zeros(2, pat.n_chunks * 64, 3, dtype=torch.float64, device=dev)
        S = self._S
        S.n_verts, S.n_tets, S.n_tris = int(n_verts), int(n_tets), int(n_tris)
        S.n_chunks, S.n_slots = pat.n_chunks, pat.n_slots
        # the histories' row stride is n_chunks * 64: they restart empty
        b["dx_hist"].zero_()
        b["dx_same"].zero_()
        S.max_width = int(np.diff(pat.chunk_ptr).max())
        S.tri_contact = int(self._tri_contact_kw and n_tris > 0 and self._has_capsule)
        S.shell_on = int(self._shell_mu > 0.0 and n_tris > 0)
        S.pcg_variant = -1  # block counts change with the topology: back to the automatic choice
        for name in ("chunk_ptr", "sell_col", "sell_val", "diag_lane", "contrib_ptr", "contrib_src", "vinc_ptr",
                     "vinc_src", "pz", "px", "tcon_ptr", "tcon_src", "vtri_ptr", "vtri_src"):
            setattr(S, name, b[name].data_ptr())
        if S.self_contact:
            self._set_self_contact_surface(int(n_tris))
        if S.shell_on:
            # Shell.refresh (shell_energy.py:262-293): rest frames and the
            # constant Gauss-Newton Hessian of the current triangulation
            b["shell_h"] = torch.zeros(pat.n_slots * 64, dtype=torch.float32, device=dev)
            S.shell_h = b["shell_h"].data_ptr()
            self._call("mfem_shell_refresh", self._stream())

    # -- self-contact (no counterpart in the reference's live solve) ----------
    def _init_self_contact(self, kwargs: dict, tris: np.ndarray, Vc: int, Fc: int, d1: float) -> None:
        """IPC point-triangle / edge-edge self-contact among the surface
        triangles inside the Newton solve (``self_contact=True``): barrier
        -(d - dhat)^2 ln(d / dhat) (src/mfem/collision.py:529-546) with
        ``self_contact_dhat`` (default contact_d1) and
        ``self_contact_kappa`` (default contact_stiffness), candidates from
        the spatial-hash broad phase on boxes swept by the Newton step, the
        line search's first step cut by a conservative CCD bound.  The hook
        in the reference's solve is commented out (src/mfem/solver.py:443-447)."""
        S, b, dev = self._S, self.buffers, self.device
        S.self_contact = 0
        if not kwargs.get("self_contact", False):
            return
        if len(tris) == 0:
            raise ValueError("self_contact needs the model's surface triangles (tri_indices)")
        dhat = float(kwargs.get("self_contact_dhat", d1))
        S.sc_dhat = dhat
        S.sc_kappa = float(kwargs.get("self_contact_kappa", S.kappa))
        S.sc_ccd_s = float(kwargs.get("self_contact_ccd_s", 0.1))
        S.sc_max_span = int(kwargs.get("self_contact_max_span", 4))
        x = np.asarray(self.model.particle_q, np.float64)
        e = _surface_edges(tris)
        max_edge = float(np.linalg.norm(x[e[:, 0]] - x[e[:, 1]], axis=1).max()) if len(e) else dhat
        # boxes of a primitive and of its motion over a step span up to
        # max_span cells: a cell of two edge lengths leaves room for motions
        # of several edge lengths per Newton step
        S.sc_cell = float(kwargs.get("self_contact_cell", max(2.0 * max_edge, 4.0 * dhat)))
        n_pts_cap, n_edge_cap = Vc, max(2 * Fc, 1)
        S.sc_max_pairs = int(kwargs.get("self_contact_max_pairs", max(1 << 16, 16 * Vc)))
        i32z = lambda *shape: torch.zeros(*shape, dtype=torch.int32, device=dev)
        b["sc_points"] = i32z(max(n_pts_cap, 1))
        b["sc_edges"] = i32z(n_edge_cap, 2)
        b["sc_pt"] = i32z(S.sc_max_pairs, 2)
        b["sc_ee"] = i32z(S.sc_max_pairs, 2)
        b["sc_count"] = i32z(4)
        b["sc_g"] = torch.zeros(Vc, 3, dtype=torch.float64, device=dev)
        b["sc_h"] = torch.zeros(Vc, 6, dtype=torch.float64, device=dev)
        b["sc_alpha"] = torch.ones(1, dtype=torch.float32, device=dev)
        nbytes = int(self.lib.mfem_self_contact_workspace_bytes(n_pts_cap, max(Fc, 1), n_edge_cap, S.sc_max_span))
        b["sc_ws"] = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        S.sc_ws_bytes = nbytes
        self._sc_caps = (n_pts_cap, n_edge_cap)
        for name in ("sc_points", "sc_edges", "sc_pt", "sc_ee", "sc_count", "sc_g", "sc_h", "sc_alpha", "sc_ws"):
            setattr(S, name, b[name].data_ptr())
        S.self_contact = 1

    def _set_self_contact_surface(self, n_tris: int) -> None:
        """Surface vertices and edges of the active triangles (host, once per
        topology)."""
        S, b = self._S, self.buffers
        tris = b["tris"][:n_tris].cpu().numpy()
        pts = np.unique(tris.reshape(-1)).astype(np.int32)
        e = _surface_edges(tris)
        if len(pts) > self._sc_caps[0] or len(e) > self._sc_caps[1]:
            raise RuntimeError("surface grew past the self-contact capacity")
        if len(pts):
            b["sc_points"][: len(pts)] = torch.as_tensor(pts, device=self.device)
        if len(e):
            b["sc_edges"][: len(e)] = torch.as_tensor(e, device=self.device)
        S.n_sc_points, S.n_sc_edges = int(len(pts)), int(len(e))

    def set_attachment_targets(self, indices: torch.Tensor, positions: torch.Tensor) -> None:
        """Moves the penalty attachments of the pinned vertices ``indices``
        (inv_mass = 0) to ``positions`` (device tensors; no host
        synchronisation, capture-safe): DirichletBoundaryCondition with new
        pin_positions followed by update_b
        (src/mfem/boundary_condition.py:85-92).  Rows of vertices that are not
        pinned are stored but unused (their attachment term is off)."""
        idx = indices.to(device=self.device, dtype=torch.long)
        self.buffers["pin_pos"].index_copy_(0, idx, positions.to(device=self.device, dtype=torch.float32))

    def self_contact_candidates(self) -> dict:
        """The current candidate lists (host copies): point-triangle pairs
        (vertex, triangle), edge-edge pairs (edge, edge), the surface edges
        and the broad-phase status bits."""
        torch.cuda.synchronize(self.device)
        b, S = self.buffers, self._S
        cnt = b["sc_count"].cpu().numpy()
        npt, nee = min(int(cnt[0]), S.sc_max_pairs), min(int(cnt[1]), S.sc_max_pairs)
        return {"pt": b["sc_pt"][:npt].cpu().numpy(), "ee": b["sc_ee"][:nee].cpu().numpy(),
                "edges": b["sc_edges"][: S.n_sc_edges].cpu().numpy(),
                "points": b["sc_points"][: S.n_sc_points].cpu().numpy(), "status": cnt[2:].copy(),
                "alpha": float(b["sc_alpha"].item())}

    def _set_bodies(self, body_vert, body_tet, body_tri, tet_energy=None) -> None:
        """Declares the mesh a batch of independent bodies (vertex ranges
        starting at multiples of 256, contiguous tets / triangles): one PCG
        and one line-search launch per Newton iteration cover all bodies,
        each with its own CG scalars, convergence, line search and stats."""
        dev, b, S = self.device, self.buffers, self._S
        bv = np.asarray(body_vert, np.int64)
        n = len(bv) - 1
        if n < 1 or np.any(bv[:-1] % 256) or bv[-1] != S.n_verts or S.n_verts % 256:
            raise ValueError("body vertex ranges must start at multiples of 256 and end at n_verts (a multiple of 256)")
        # workgroups of up to 8 blocks of 256 rows stay within one body
        align = next(a for a in (8, 4, 2, 1) if not np.any((bv // 256) % a))
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        b["body_vert"] = i32(bv)
        b["body_tet"] = i32(body_tet)
        b["body_tri"] = i32(body_tri)
        b["body_wg"] = i32(bv // 256)
        # line-search workgroups in proportion to each body's work, at least one
        work = np.diff(bv) + np.diff(np.asarray(body_tet, np.int64))
        cap = int(self.lib.mfem_line_search_max_blocks())
        if cap < n:
            raise ValueError(f"{n} bodies need {n} line-search workgroups, {cap} can be resident")
        share = np.maximum(1, np.floor(work / work.sum() * (cap - n)).astype(np.int64))
        while share.sum() > cap:
            share[np.argmax(share)] -= 1
        b["body_ls_wg"] = i32(np.concatenate([[0], np.cumsum(share)]))
        b["body_epoch"] = torch.zeros(n, dtype=torch.int32, device=dev)
        b["stats"] = torch.zeros(n * _ffi.MFEM_MAX_STATS_ITERS * _ffi.MFEM_STATS_PER_ITER, dtype=torch.float64,
                                 device=dev)
        if tet_energy is not None:
            b["tet_energy"] = i32(np.concatenate([np.asarray(tet_energy, np.int32),
                                                  np.zeros(self.max_tets - len(tet_energy), np.int32)]))
            S.tet_energy = b["tet_energy"].data_ptr()
        S.n_bodies, S.ls_blocks, S.body_wg_align = n, int(share.sum()), align
        for name in ("body_vert", "body_tet", "body_tri", "body_wg", "body_ls_wg", "body_epoch", "stats"):
            setattr(S, name, b[name].data_ptr())
        self.n_bodies = n

    def body_stats(self, body: int) -> np.ndarray:
        """last_stats() of one body of a batch."""
        torch.cuda.synchronize(self.device)
        st = self.buffers["stats"].view(-1, _ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)[body]
        return st[: self.iterations].cpu().numpy().copy()

    @property
    def active_particle_count(self) -> int:
        return int(self._S.n_verts)

    @property
    def active_tet_count(self) -> int:
        return int(self._S.n_tets)

    @property
    def active_tri_count(self) -> int:
        return int(self._S.n_tris)

    def _refresh_masses(self) -> None:
        """Lumped masses from the rest volumes and refine_density, as
        refine() rewrites model.particle_mass on every call
        (refinement.py:1253, scatter_tets :733-824).  The inverse masses of
        free vertices follow (pinned vertices keep 0); the reference keeps
        model.particle_inv_mass unchanged in its Hessian mass term."""
        b = self.buffers
        nv, nt = self.active_particle_count, self.active_tet_count
        vol = b["vol"][:nt].double().cpu().numpy()
        tets = b["tets"][:nt].cpu().numpy()
        # (bincount adds the quarter volumes in element order, as add.at would)
        m = np.bincount(tets.reshape(-1), weights=np.repeat(self._refine_density * vol / 4.0, 4), minlength=nv)[:nv]
        pinned = self._pinned[:nv].cpu().numpy()
        b["mass"][:nv] = torch.as_tensor(m, dtype=torch.float32, device=self.device)
        b["inv_mass"][:nv] = torch.as_tensor(np.where(pinned, 0.0, 1.0 / np.maximum(m, 1e-30)), dtype=torch.float32,
                                             device=self.device)

    def _refine_pass(self, state: "State") -> dict:
        """One refine() call (refinement.py:914-1322): gated by
        refine_every_n_steps on a pass counter, as in the reference."""
        from .refine import Refiner

        counter = self._pass_counter
        self._pass_counter += 1
        every = max(int(self.refine_every_n_steps), 1)
        res = {"pass": counter, "fired": counter % every == 0, "new_verts": 0, "dropped": False}
        changed = False
        if res["fired"]:
            b = self.buffers
            if self._refiner is None:
                self._refiner = Refiner(self.max_tets, self.max_tris, self.max_particles, self.device)
                dev = self.device
                self._ref_out = {"tets": torch.zeros(self.max_tets, 4, dtype=torch.int32, device=dev),
                                 "stretch": torch.zeros(self.max_tets, 6, device=dev),
                                 "lam": torch.zeros(self.max_tets, 6, device=dev),
                                 "materials": torch.zeros(self.max_tets, 3, device=dev),
                                 "dminv": torch.zeros(self.max_tets, 9, device=dev),
                                 "tris": torch.zeros(max(self.max_tris, 1), 3, dtype=torch.int32, device=dev)}
            S = self._S
            shapes = (b["shape_type"], b["shape_xf"], b["shape_scale"]) if S.n_shapes > 0 else None
            geo = {}
            if self._refine_scoring == "geometric":
                # distances for the current shape poses (solver.py:449-466)
                if "vdist" not in b:
                    b["vdist"] = torch.zeros(self.max_particles, dtype=torch.float32, device=self.device)
                    b["vsid"] = torch.zeros(self.max_particles, dtype=torch.int32, device=self.device)
                _ffi.call("mfem_contact_distance", ctypes.byref(S), ctypes.c_void_p(b["vdist"].data_ptr()),
                          ctypes.c_void_p(b["vsid"].data_ptr()), self._stream())
                tri_on = bool(S.tri_contact and S.n_tris > 0)
                geo = dict(vertex_dist=b["vdist"], vertex_sid=b["vsid"],
                           tri_dist=b["tri_dist"] if tri_on else None, tri_bary=b["tri_bary"] if tri_on else None)
            r = self._refiner.run(tets=b["tets"], tris=b["tris"], x=state.particle_q, v=state.particle_qd,
                                  rest=b["rest"], inv_mass=b["inv_mass"], tet_score=b["psi"], vertex_score=b["eB"],
                                  stretch=b["s"], lam=b["lam"], materials=b["materials"], dminv=b["dminv"],
                                  n_verts=S.n_verts, n_tets=S.n_tets, n_tris=S.n_tris, shapes=shapes,
                                  out=self._ref_out, **self._refine_kw, **geo)
            res.update(r)
            if r["new_verts"] > 0 and not r["dropped"]:
                o, nt, nf = self._ref_out, r["n_tets"], r["n_tris"]
                b["tets"][:nt] = o["tets"][:nt]
                b["s"][:nt] = o["stretch"][:nt]
                b["lam"][:nt] = o["lam"][:nt]
                b["materials"][:nt] = o["materials"][:nt]
                b["dminv"][:nt] = o["dminv"][:nt]
                d = o["dminv"][:nt].double()
                det = (d[:, 0] * (d[:, 4] * d[:, 8] - d[:, 5] * d[:, 7]) - d[:, 1] * (d[:, 3] * d[:, 8] - d[:, 5] * d[:, 6])
                       + d[:, 2] * (d[:, 3] * d[:, 7] - d[:, 4] * d[:, 6]))
                b["vol"][:nt] = (1.0 / (6.0 * det)).float()
                if nf:
                    b["tris"][:nf] = o["tris"][:nf]
                nv0 = self.active_particle_count
                b["dx"][nv0:r["n_verts"]] = 0.0
                self._set_topology(r["n_verts"], nt, nf)
                changed = True
        if changed or counter == 0:
            self._refresh_masses()
        self.refine_log.append(res)
        return res

    # -- plumbing ---------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _call(self, fn, *args):
        _ffi.call(fn, ctypes.byref(self._S), *args)

    def _bind_state(self, q: torch.Tensor, qd: torch.Tensor) -> None:
        for t in (q, qd):
            if t.dtype != torch.float32 or t.device != self.device or not t.is_contiguous() or t.shape != (self.max_particles, 3):
                raise ValueError("state tensors must be contiguous float32 (V, 3) on the solver's device")
        self._S.x = q.data_ptr()
        self._S.v = qd.data_ptr()
        self._bound = (q, qd)

    def _set_body_q(self, body_q) -> None:
        """World poses of the shapes for this step's rigid body poses:
        body_q[shape_body] * shape_transform, static shapes unchanged
        (src/mfem/refinement/contact.py:74-82)."""
        if body_q is None or self._S.n_shapes == 0:
            return
        bq = torch.as_tensor(body_q, dtype=torch.float32, device=self.device).reshape(-1, 7).contiguous()
        if bq.shape[0] < self._n_rigid:
            raise ValueError(f"body_q holds {bq.shape[0]} poses but the shapes use {self._n_rigid} rigid bodies")
        b = self.buffers
        self._body_q = bq  # (kept alive until the launch has read it)
        _ffi.call("mfem_pose_shapes", ctypes.c_void_p(bq.data_ptr()), ctypes.c_int(bq.shape[0]),
                  ctypes.c_void_p(b["shape_body"].data_ptr()), ctypes.c_void_p(b["shape_local"].data_ptr()),
                  ctypes.c_int(self._S.n_shapes), ctypes.c_void_p(b["shape_xf"].data_ptr()), self._stream())

    def set_galerkin_start(self, on: bool = True) -> None:
        """The Galerkin start without a network: the lowest quadratic-model
        energy over x~ - x, the last three Newton steps and the same Newton
        iteration's solution one step earlier (no learned guess, no MLP
        launch).  A warm-start network attached later replaces it."""
        self.warm_start_net = None
        S = self._S
        S.mlp_w = S.mlp_wbf = None
        S.guess = None
        if on:
            S.pcg_flags |= _ffi.PCG_LEARNED_GUESS
        else:
            S.pcg_flags &= ~_ffi.PCG_LEARNED_GUESS

    def set_warm_start(self, net) -> None:
        """Attach a trained ``learned.WarmStartNet`` (or None for plain PCG).
        With a network, every PCG solve first evaluates it and starts from the
        lowest quadratic-model energy over the span of x~ - x, the learned
        guess and the last three Newton steps (a Galerkin start)."""
        self.warm_start_net = net
        self._S.guess = self.buffers["guess"].data_ptr()
        if net is None:
            self._S.pcg_flags &= ~_ffi.PCG_LEARNED_GUESS
            self._S.mlp_w = self._S.mlp_wbf = None
            return
        if net.device != self.device:
            raise ValueError("the warm-start network must live on the solver's device")
        self._S.mlp_w = net.w.data_ptr()
        self._S.mlp_wbf = net.wbf.data_ptr()
        self._S.pcg_flags |= _ffi.PCG_LEARNED_GUESS

    # -- API ----------------------------------------------------------------
    def step(self, state_in: State, state_out: State, control=None, contacts=None, dt: float = 1.0 / 60.0) -> None:
        """One implicit step (solver.py:429-579): state_out <- state_in, then
        predictor, ``iterations`` Newton iterations (update, assemble, PCG,
        local solve, line search) and the velocity update, all on the device
        without a host synchronisation."""
        self._check_device_status()
        self._step += 1
        if state_out is not state_in:
            state_out.assign(state_in)
        self._set_body_q(state_out.body_q)
        self._bind_state(state_out.particle_q, state_out.particle_qd)
        if self._refine_enabled:
            self._refine_pass(state_out)
        self._call("mfem_step", ctypes.c_float(dt), self._stream())
        self._mirror_status()
        self._log_launches()

    def _log_launches(self, newton_iter: int | None = None) -> None:
        """With ``launch_log`` a list (profiling), appends the PCG launches of
        the last step (or of Newton iteration ``newton_iter``): the variant
        and the CG iterations of each, in launch order (synchronises)."""
        if self.launch_log is not None:
            v = int(self.lib.mfem_last_pcg_variant())
            it = self.last_stats()[:, 0].astype(int).tolist()
            self.launch_log += [(v, n) for n in (it if newton_iter is None else [it[newton_iter]])]

    def _check_device_status(self) -> None:
        """Raises when an earlier step ended with a device error (a grid
        barrier timed out): every kernel of a step then does nothing, so the
        state stops advancing until ``reset_status``.  Reads the pinned
        mirror of the last completed step: no host synchronisation."""
        bits = int(self._status_mirror[0])
        if bits != 0:
            what = []
            if bits & _ffi.MFEM_ERR_BARRIER:
                what.append("a grid barrier timed out")
            if bits & _ffi.MFEM_ERR_CONTACT:
                what.append("self-contact candidates overflowed (raise self_contact_max_pairs / "
                            "self_contact_max_span or self_contact_cell)")
            raise _ffi.MfemError(f"{' and '.join(what) or 'device error'} in an earlier step (device error bits "
                                 f"{bits}); the state was left untouched since; call reset_status() to continue")

    def _mirror_status(self) -> None:
        self._status_mirror.copy_(self.buffers["status"][:1], non_blocking=True)

    def step_instrumented(self, state_in: State, state_out: State, dt: float, timers: dict) -> None:
        """Same kernels and order as ``step`` (mfem_step), issued stage by stage
        so that torch.cuda.Events on the launching stream bracket each phase.
        ``timers[phase]`` collects (start, end) event pairs."""
        self._check_device_status()
        self._step += 1
        if state_out is not state_in:
            state_out.assign(state_in)
        self._set_body_q(state_out.body_q)
        self._bind_state(state_out.particle_q, state_out.particle_qd)
        if self._refine_enabled:
            self._refine_pass(state_out)

        def timed(name, fn, *args):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn(*args)
            b.record()
            timers.setdefault(name, []).append((a, b))

        timed("begin_step", self.begin_step, dt)
        for it in range(self.iterations):
            timed("update_system", self.update_system, dt)
            timed("assemble", self.assemble, dt)
            if self.warm_start_net is not None:
                timed("warm_start", self.warm_start, dt, it)
            if it == 0 and self.coarse_net is not None and (self._S.pcg_flags & _ffi.PCG_COARSE):
                timed("coarse_net", self.coarse_weights_from_net, dt)
            timed("global_solve", self.global_solve, it)
            timed("local_solve", self.local_solve)
            timed("line_search", self.line_search, dt, it)
        timed("end_step", self.end_step, dt)
        self._mirror_status()

    # stage entry points (used by tests and profiling) ------------------------
    def bind(self, state: State) -> None:
        self._set_body_q(state.body_q)
        self._bind_state(state.particle_q, state.particle_qd)

    def begin_step(self, dt):
        self._call("mfem_begin_step", ctypes.c_float(dt), self._stream())

    def update_system(self, dt):
        self._call("mfem_update_system", ctypes.c_float(dt), self._stream())

    def assemble(self, dt):
        self._call("mfem_assemble", ctypes.c_float(dt), self._stream())

    def ws_features(self, dt, newton_iter=0):
        """Warm-start features and candidates into buffers ws_x / ws_aux."""
        self._call("mfem_ws_features", ctypes.c_float(dt), ctypes.c_int(newton_iter), self._stream())

    def warm_start(self, dt, newton_iter=0):
        """Features, network and blend: the learned initial guess (buffer
        ``guess``) that the next ``global_solve`` tries first."""
        net = self.warm_start_net
        if net is None:
            raise RuntimeError("no warm-start network attached")
        self.ws_features(dt, newton_iter)
        b = self.buffers
        _ffi_mlp = self.lib.mfem_mlp_forward
        rc = _ffi_mlp(ctypes.c_void_p(b["ws_x"].data_ptr()), ctypes.c_int(self.active_particle_count),
                      ctypes.c_void_p(net.wbf.data_ptr()), ctypes.c_void_p(net.w.data_ptr()),
                      ctypes.c_void_p(b["ws_y"].data_ptr()), self._stream())
        if rc != 0:
            raise RuntimeError(f"mfem_mlp_forward failed ({rc}): {self.lib.mfem_mlp_last_error().decode()}")
        self._call("mfem_ws_apply", self._stream())

    def global_solve(self, newton_iter=0, learned: bool | None = None):
        """PCG on H dx = g.  learned=None follows the attached network (the
        guess from the last ``warm_start``); False forces the plain guess."""
        flags = self._S.pcg_flags
        if learned is not None:
            if learned and self.warm_start_net is None:
                raise RuntimeError("no warm-start network attached")
            self._S.pcg_flags = (flags | _ffi.PCG_LEARNED_GUESS) if learned else (flags & ~_ffi.PCG_LEARNED_GUESS)
        try:
            self._call("mfem_global_solve", ctypes.c_int(newton_iter), self._stream())
        finally:
            self._S.pcg_flags = flags
        self._log_launches(newton_iter)

    def local_solve(self):
        self._call("mfem_local_solve", self._stream())

    def line_search(self, dt, newton_iter=0):
        self._call("mfem_line_search", ctypes.c_float(dt), ctypes.c_int(newton_iter), self._stream())

    def end_step(self, dt):
        self._call("mfem_end_step", ctypes.c_float(dt), self._stream())

    def spmv(self, x: torch.Tensor) -> torch.Tensor:
        x = x.to(device=self.device, dtype=torch.float32).contiguous()
        y = torch.empty_like(x)
        self._call("mfem_spmv", ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), self._stream())
        return y

    def spmv_with(self, values: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """y = H x for stored matrix values in this solver's SELL layout (a
        copy of ``buffers['sell_val']`` taken at another Newton iteration)."""
        if values.shape != self.buffers["sell_val"].shape or values.dtype != torch.float32 or \
                values.device != self.device:
            raise ValueError("values must match buffers['sell_val'] in shape, dtype and device")
        x = x.to(device=self.device, dtype=torch.float32).contiguous()
        y = torch.empty_like(x)
        S2 = type(self._S).from_buffer_copy(self._S)
        S2.sell_val = values.data_ptr()
        _ffi.call("mfem_spmv", ctypes.byref(S2), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                  self._stream())
        return y

    def merit(self, dt: float, alpha: float = 0.0) -> float:
        out = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._call("mfem_merit", ctypes.c_float(dt), ctypes.c_float(alpha), ctypes.c_void_p(out.data_ptr()),
                   self._stream())
        return float(out.item())

    # autotuning and graph capture ----------------------------------------------
    def pcg_variants(self) -> list[dict]:
        """The persistent-PCG variants and whether each applies to the
        current system (chunks per wave, resident row slots, mode,
        workgroups)."""
        n = self.lib.mfem_pcg_variant_count()
        out = []
        info = (ctypes.c_int32 * 5)()
        for k in range(n):
            _ffi.call("mfem_pcg_variant_info", ctypes.byref(self._S), ctypes.c_int(k), info)
            out.append(dict(variant=k, chunks_per_wave=info[0], resident_slots=info[1],
                            mode=("standard", "pipelined", "standard_lowreg", "pipelined_coarse")[info[2]],
                            workgroups=info[3], applicable=bool(info[4])))
        return out

    def set_coarse_correction(self, on: bool, omega: float | None = None, weights: torch.Tensor | None = None) -> None:
        """Block Jacobi plus the workgroup-local rigid-mode correction (on) or
        block Jacobi alone (off).  Each 256-row PCG workgroup is one
        aggregate whose rigid modes, smoothed per vertex i by
        P'_i = P_i - w_i D_i^-1 (H_aa P)_i, span a local coarse space solved
        exactly: M^-1 = D^-1 + P' (P'^T H_aa P')^+ P'^T.  w_i = weights[i]
        (a device float32 tensor of n_verts, e.g. a network's output) or the
        uniform ``omega``.  The choice of PCG variant restarts."""
        S = self._S
        if on:
            S.pcg_flags |= _ffi.PCG_COARSE | _ffi.PCG_PIPELINED
        else:
            S.pcg_flags &= ~_ffi.PCG_COARSE
        if omega is not None:
            S.coarse_omega = float(omega)
        if weights is None:
            self._coarse_w = None
            S.coarse_w = self.buffers["coarse_w_net"].data_ptr() if self.coarse_net is not None else None
        else:
            if weights.dtype != torch.float32 or weights.device != self.device or weights.numel() < S.n_verts:
                raise ValueError("coarse weights must be a float32 tensor of n_verts on the solver's device")
            self._coarse_w = weights
            S.coarse_w = weights.data_ptr()
        S.pcg_variant = -1

    def _ensure_coarse_buffers(self) -> None:
        b, dev, Vc = self.buffers, self.device, self.max_particles
        if "cw_x" not in b:
            b["cw_x"] = torch.zeros(Vc, 16, dtype=torch.int16, device=dev)
            b["cw_y"] = torch.zeros(Vc, 3, dtype=torch.float32, device=dev)
            b["coarse_w_net"] = torch.zeros(Vc, dtype=torch.float32, device=dev)
            self._S.cw_x = b["cw_x"].data_ptr()
            self._S.cw_y = b["cw_y"].data_ptr()

    def coarse_features(self, dt: float) -> None:
        """The coarse-weight network's features of the assembled system into
        buffers['cw_x']."""
        self._ensure_coarse_buffers()
        self._call("mfem_coarse_features", ctypes.c_float(dt), self._stream())

    def set_coarse_net(self, net) -> None:
        """Attach a ``coarse.CoarseWeightNet`` (or None): with a network, the
        coarse correction is on and its smoothing weights are the network's,
        evaluated by every step at its first Newton iteration; None leaves
        the correction's switch as it is with the uniform weight."""
        S = self._S
        self.coarse_net = net
        if net is None:
            S.cw_w = S.cw_wbf = None
            if self._coarse_w is None:
                S.coarse_w = None
            return
        if net.device != self.device:
            raise ValueError("the coarse-weight network must live on the solver's device")
        self._ensure_coarse_buffers()
        S.cw_w = net.w.data_ptr()
        S.cw_wbf = net.wbf.data_ptr()
        S.coarse_w = self.buffers["coarse_w_net"].data_ptr()
        S.pcg_flags |= _ffi.PCG_COARSE | _ffi.PCG_PIPELINED

    def coarse_weights_from_net(self, dt: float) -> None:
        """Features, network and weights (what mfem_step does at the first
        Newton iteration of a step with a coarse-weight network)."""
        net = self.coarse_net
        self.coarse_features(dt)
        b = self.buffers
        rc = self.lib.mfem_mlp_forward(ctypes.c_void_p(b["cw_x"].data_ptr()), ctypes.c_int(self.active_particle_count),
                                       ctypes.c_void_p(net.wbf.data_ptr()), ctypes.c_void_p(net.w.data_ptr()),
                                       ctypes.c_void_p(b["cw_y"].data_ptr()), self._stream())
        if rc != 0:
            raise RuntimeError(f"mfem_mlp_forward failed ({rc}): {self.lib.mfem_mlp_last_error().decode()}")
        self._call("mfem_coarse_weights", self._stream())

    def set_local_preconditioner(self, sweeps: int, omega: float = 0.5, _S=None) -> None:
        """Preconditioner of the standard PCG: sweeps = 0 is the reference's
        block Jacobi; an even sweeps > 0 adds that many damped block-Jacobi
        sweeps (weight omega) over the rows of each PCG workgroup, a
        symmetric positive definite polynomial preconditioner with no
        synchronisation beyond the workgroup.  Pipelined PCG variants are
        not used while sweeps > 0."""
        S = self._S if _S is None else _S
        if sweeps < 0 or sweeps > 16 or sweeps % 2 or (sweeps and not omega > 0.0):
            raise ValueError("pcg_local_sweeps must be even in [0, 16] and pcg_local_omega > 0")
        S.local_sweeps, S.local_omega = int(sweeps), float(omega)
        if _S is None and S.pcg_variant >= 0 and sweeps:
            info = (ctypes.c_int32 * 5)()
            _ffi.call("mfem_pcg_variant_info", ctypes.byref(S), ctypes.c_int(S.pcg_variant), info)
            if not info[4]:
                S.pcg_variant = -1

    _STEP_KEYS = ("s", "lam", "dx", "dx_hist", "dx_same", "stats", "x_start", "x_tilde", "lag_n", "lag_fn", "lag_sid")

    def snapshot(self) -> dict:
        """Copies of the solver-held per-tet / per-vertex state that a step
        reads besides the State (stretch, multipliers, last Newton step,
        lagged friction, stats), for ``restore``."""
        return {k: self.buffers[k].clone() for k in self._STEP_KEYS}

    def restore(self, snap: dict) -> None:
        for k, t in snap.items():
            self.buffers[k].copy_(t)

    def autotune(self, state: "State", dt: float, repeats: int = 2, local_sweeps=None) -> dict:
        """Times every applicable PCG variant, under each workgroup-local
        preconditioner setting in ``local_sweeps`` (default: the current one
        only), on whole steps from ``state`` (all Newton iterations, with the
        warm start if one is attached: the candidates differ in iteration
        counts as well as in cost per iteration) and keeps the fastest pair.
        The state and the solver's per-tet / per-vertex state are restored
        afterwards; refinement passes are not run while tuning.  The choice
        holds until the topology changes."""
        saved = self.snapshot()
        scratch = state.clone()
        refine, self._refine_enabled = self._refine_enabled, False
        step, self._step = self._step, self._step
        omega = float(self._S.local_omega) if self._S.local_omega > 0 else 0.5
        sweeps_list = [int(self._S.local_sweeps)] if local_sweeps is None else [int(x) for x in local_sweeps]
        table = []
        try:
            for sw in sweeps_list:
                self._S.pcg_variant = -1
                self.set_local_preconditioner(sw, omega)
                for v in self.pcg_variants():
                    if not v["applicable"]:
                        continue
                    self._S.pcg_variant = v["variant"]
                    ms, iters = [], []
                    best_ms = min((t["ms_per_step"] for t in table), default=float("inf"))
                    for r in range(repeats + 1):  # the first run warms up
                        self.restore(saved)
                        a = torch.cuda.Event(enable_timing=True)
                        e = torch.cuda.Event(enable_timing=True)
                        a.record()
                        self.step(state, scratch, None, None, dt)
                        e.record()
                        torch.cuda.synchronize(self.device)
                        if self.status() != 0:
                            raise RuntimeError("device error while autotuning")
                        if r > 0 or a.elapsed_time(e) > 3.0 * best_ms:
                            # (a warm-up run 3x slower than the best is final)
                            ms.append(a.elapsed_time(e))
                            iters.append(int(self.last_stats()[:, 0].sum()))
                        if r == 0 and ms:
                            break
                    table.append(dict(v, local_sweeps=sw, ms_per_step=float(np.median(ms)),
                                      cg_iters=int(np.median(iters)),
                                      us_per_iter=float(np.median(ms)) * 1e3 / max(int(np.median(iters)), 1)))
        finally:
            self._refine_enabled = refine
            self._step = step
            self.restore(saved)
        best = min(table, key=lambda r: r["ms_per_step"])
        self._S.pcg_variant = -1
        self.set_local_preconditioner(best["local_sweeps"], omega)
        self._S.pcg_variant = best["variant"]
        return dict(chosen=best, table=table)

    def capture_step(self, state_in: "State", state_out: "State", dt: float):
        """A HIP graph of ``step(state_in, state_out, ...)`` on these exact
        tensors (refinement must be off: a pass synchronises with the host).
        Replay with ``graph.replay()``; the caller alternates two graphs
        (a->b, b->a) to advance."""
        if self._refine_enabled:
            raise RuntimeError("capture_step needs enable_refinement=False")
        self.step(state_in, state_out, None, None, dt)  # warm-up outside the capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        log, self.launch_log = self.launch_log, None  # (no host synchronisation under capture)
        try:
            with torch.cuda.graph(g):
                self.step(state_in, state_out, None, None, dt)
        finally:
            self.launch_log = log
        return g

    # diagnostics ------------------------------------------------------------
    def last_stats(self) -> np.ndarray:
        """Per Newton iteration of the last step: cg_iters, cg_relres (FP64),
        ls_trials, ls_accepted, alpha, f0, f1, cg_restarts, PCG phase ticks
        (6), learned guess used, learned / plain initial residual ratio.
        A batch: the maximum over its bodies of each entry (cg_iters is then
        the launch's iteration count, cg_relres the worst body's)."""
        torch.cuda.synchronize(self.device)
        st = self.buffers["stats"].cpu().numpy().reshape(-1, _ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)
        return st[:, : self.iterations].max(axis=0)

    def status(self) -> int:
        """Device error bits (MFEM_ERR_BARRIER: a grid barrier timed out;
        MFEM_ERR_CONTACT: self-contact candidates overflowed; every kernel
        of a step then returns at once until ``reset_status``; the next
        ``step`` raises)."""
        torch.cuda.synchronize(self.device)
        return int(self.buffers["status"][0].item())

    def reset_status(self) -> None:
        """Clears the error bits, the barrier epoch and the epoch-tagged
        reduction partials (their tags restart with the epoch)."""
        torch.cuda.synchronize(self.device)
        self.buffers["status"].zero_()
        self.buffers["flags"].zero_()
        self.buffers["red"].zero_()
        torch.cuda.synchronize(self.device)
        self._status_mirror.zero_()

    def system_csr(self):
        """Assembled H as a scipy.sparse.bsr_matrix (host copy)."""
        import scipy.sparse as sp
        from .pattern import sell_to_csr_values
        vals = sell_to_csr_values(self.pattern, self.buffers["sell_val"].cpu().numpy())
        n = self.active_particle_count
        return sp.bsr_matrix((vals, self.pattern.cols, self.pattern.rowptr), shape=(3 * n, 3 * n))

    @property
    def tet_stretch(self) -> torch.Tensor:
        return self.buffers["s"][: self.active_tet_count]

    @property
    def tet_lambda(self) -> torch.Tensor:
        return self.buffers["lam"][: self.active_tet_count]

    @property
    def elastic_energy(self) -> torch.Tensor:
        return self.buffers["psi"][: self.active_tet_count]

    @property
    def tet_indices(self) -> torch.Tensor:
        return self.buffers["tets"][: self.active_tet_count]

    @property
    def tri_indices(self) -> torch.Tensor:
        return self.buffers["tris"][: self.active_tri_count]

    @property
    def rest_particle_q(self) -> torch.Tensor:
        return self.buffers["rest"][: self.active_particle_count]

    def write_timings(self, filename: str) -> None:
        with open(filename, "w") as f:
            json.dump(self._timings, f)
