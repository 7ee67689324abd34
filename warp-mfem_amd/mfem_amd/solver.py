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
from .pattern import build_system_pattern, build_triangle_lists

_ENERGIES = {"arap": 0, "neohookean": 1}


@dataclass
class State:
    """Per-step state (newton State fields used by the solver)."""

    particle_q: torch.Tensor   # (V, 3) float32, device
    particle_qd: torch.Tensor  # (V, 3) float32, device
    body_q: torch.Tensor | None = None  # (n_shapes, 7) world poses p, q(xyzw), or None = model poses

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


def model_state(model, device="cuda") -> State:
    """A State holding the model's initial positions / velocities."""
    q = torch.as_tensor(np.asarray(model.particle_q, dtype=np.float32), device=device).contiguous()
    qd = torch.as_tensor(np.asarray(model.particle_qd, dtype=np.float32), device=device).contiguous()
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


class RefinementSolver:
    def __init__(self, model, iterations: int, max_tets: int, max_particles: int | None = None,
                 max_tris: int | None = None, refine_every_n_steps: int = 10, min_refine_score: float = 10.1,
                 device: str = "cuda", **kwargs):
        if kwargs.get("enable_refinement", True):
            raise NotImplementedError(
                "adaptive refinement is not implemented by mfem_amd; pass enable_refinement=False")
        if kwargs.get("shell_mu", 0.0) > 0.0:
            raise NotImplementedError("membrane shell energy (shell_mu > 0) is not implemented by mfem_amd")
        energy = kwargs.get("energy", "arap")
        if energy not in _ENERGIES:
            raise ValueError(f"Unknown energy type: {energy}")
        nv, nt = int(model.particle_count), int(model.tet_count)
        if max_particles is not None and max_particles != nv:
            raise ValueError(f"max_particles={max_particles} does not match model.particle_count={nv}")
        if max_tets < nt:
            raise ValueError(f"max_tets={max_tets} < model.tet_count={nt}")
        if iterations > _ffi.MFEM_MAX_STATS_ITERS:
            raise ValueError(f"iterations must be <= {_ffi.MFEM_MAX_STATS_ITERS}")
        self.lib = _ffi.load()
        _ffi.check_layout(self.lib)
        self.model = model
        self.iterations = int(iterations)
        self.max_tets, self.max_particles = int(max_tets), nv
        self.refine_every_n_steps = refine_every_n_steps
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
        pat = build_system_pattern(nv, tets)
        self.pattern = pat
        inv_mass = np.asarray(model.particle_inv_mass, dtype=np.float64)
        pin_pos = np.where((inv_mass == 0)[:, None], np.asarray(model.particle_q, dtype=np.float64), 0.0)

        b = {}
        b["tets"] = i32(tets)
        b["dminv"] = f32(dminv.reshape(-1, 9))
        b["materials"] = f32(np.asarray(model.tet_materials).reshape(-1, 3))
        b["vol"] = f32(vol)
        b["mass"] = f32(model.particle_mass)
        b["inv_mass"] = f32(inv_mass)
        b["pin_pos"] = f32(pin_pos)
        tris = np.zeros((0, 3), np.int32) if model.tri_indices is None else np.asarray(model.tri_indices, np.int32)
        b["tris"] = i32(tris.reshape(-1, 3)) if len(tris) else torch.zeros(1, dtype=torch.int32, device=dev)
        b["shape_type"] = i32([sh.type for sh in shapes] or [0])
        b["shape_xf"] = f32(np.array([np.concatenate([sh.pos, sh.quat]) for sh in shapes]) if shapes else np.zeros((1, 7)))
        b["shape_scale"] = f32(np.array([sh.scale for sh in shapes]) if shapes else np.zeros((1, 3)))
        b["shape_mu"] = f32(mu if n_shapes else np.zeros(1))
        b["chunk_ptr"] = i32(pat.chunk_ptr)
        b["sell_col"] = i32(pat.sell_col)
        b["sell_val"] = zf(pat.n_slots * 576)
        b["diag_lane"] = i32(pat.diag_lane)
        b["contrib_ptr"] = i32(pat.contrib_ptr)
        b["contrib_src"] = i32(pat.contrib_src) if len(pat.contrib_src) else torch.zeros(1, dtype=torch.int32, device=dev)
        b["vinc_ptr"] = i32(pat.vinc_ptr)
        b["vinc_src"] = i32(pat.vinc_src)
        b["s"] = zf(nt, 6)
        b["lam"] = zf(nt, 6)
        b["x_tilde"] = zf(nv, 3)
        b["x_start"] = zf(nv, 3)
        b["lag_n"] = zf(nv, 3)
        b["lag_fn"] = zf(nv)
        b["lag_sid"] = torch.full((nv,), -1, dtype=torch.int32, device=dev)
        b["G"] = zf(nt, 72)
        b["Hl"] = zf(nt, 21)
        b["gl"] = zf(nt, 6)
        b["Hsinv"] = zf(nt, 21)
        b["gs"] = zf(nt, 6)
        b["psi"] = zf(nt)
        b["Kb"] = zf(10, nt, 9)
        b["fe"] = zf(nt, 12)
        for n in ("gx", "gB", "gD", "g"):
            b[n] = zf(nv, 3)
        b["Hv"] = zf(nv, 9)
        for n in ("dist", "eB", "eD"):
            b[n] = zf(nv)
        b["dx"] = torch.zeros(nv, 3, dtype=torch.float64, device=dev)
        b["ds"] = zf(nt, 6)
        b["pz"] = zf(2, pat.n_chunks * 64, 3)
        b["px"] = torch.zeros(2, pat.n_chunks * 64, 3, dtype=torch.float64, device=dev)
        b["red"] = torch.zeros(2 * _ffi.MFEM_RED_MAX_BLOCKS * 4, dtype=torch.float64, device=dev)
        b["flags"] = torch.zeros(_ffi.MFEM_RED_MAX_BLOCKS, dtype=torch.int32, device=dev)
        b["status"] = torch.zeros(8, dtype=torch.int32, device=dev)
        b["stats"] = torch.zeros(_ffi.MFEM_MAX_STATS_ITERS * _ffi.MFEM_STATS_PER_ITER, dtype=torch.float64, device=dev)
        # triangle-capsule contact: gather lists and per-triangle state
        n_tris = len(tris)
        if n_tris:
            tcon_ptr, tcon_src, vtri_ptr, vtri_src = build_triangle_lists(pat, tris)
        else:
            tcon_ptr = np.zeros(pat.n_slots * 64 + 1, np.int32)
            tcon_src = vtri_src = np.zeros(1, np.int32)
            vtri_ptr = np.zeros(nv + 1, np.int32)
        b["tcon_ptr"] = i32(tcon_ptr)
        b["tcon_src"] = i32(tcon_src if len(tcon_src) else np.zeros(1, np.int32))
        b["vtri_ptr"] = i32(vtri_ptr)
        b["vtri_src"] = i32(vtri_src if len(vtri_src) else np.zeros(1, np.int32))
        b["tri_dist"] = torch.full((max(n_tris, 1),), 1.0e8, dtype=torch.float32, device=dev)
        b["tri_bary"] = zf(max(n_tris, 1), 3)
        b["tri_nrm"] = zf(max(n_tris, 1), 3)
        b["tri_sid"] = torch.full((max(n_tris, 1),), -1, dtype=torch.int32, device=dev)
        # learned warm start workspace (features, candidates, outputs, guess)
        b["ws_x"] = torch.zeros(nv, 16, dtype=torch.int16, device=dev)
        b["ws_aux"] = zf(nv, 10)
        b["ws_y"] = zf(nv, 3)
        b["guess"] = zf(nv, 3)
        self.buffers = b

        S = _ffi.MfemSolverStruct()
        S.n_verts, S.n_tets, S.n_tris, S.n_shapes = nv, nt, len(tris), n_shapes
        S.n_chunks, S.n_slots = pat.n_chunks, pat.n_slots
        S.max_width = int(np.diff(pat.chunk_ptr).max())
        # solver-specific knobs (not reference kwargs): pipelined PCG, resident rows
        S.pcg_flags = (_ffi.PCG_PIPELINED if kwargs.get("pcg_pipelined", False) else 0) | \
                      (0 if kwargs.get("pcg_resident_rows", True) else _ffi.PCG_STREAM_ROWS)
        S.energy = _ENERGIES[energy]
        S.iterations = self.iterations
        S.cg_maxiter = int(kwargs.get("cg_max_iterations", 1000))
        S.precond = int(bool(kwargs.get("preconditioner", True)))
        S.line_search = int(bool(kwargs.get("line_search", True)))
        S.ls_max = int(kwargs.get("line_search_max_iterations", 20))
        # Contact(tri_contact=...) keeps it only with triangles (contact.py:844)
        S.tri_contact = int(tri_contact and len(tris) > 0)
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
        for name, t in b.items():
            setattr(S, name, t.data_ptr())
        self._S = S
        self.warm_start_net = None
        self.set_warm_start(kwargs.get("warm_start"))

        # AdditionalState.from_model: rest stretch from the model's positions
        x0 = f32(np.asarray(model.particle_q).reshape(-1, 3))
        self._rest_q = x0
        self._bind_state(x0, zf(nv, 3))
        self._call("mfem_init_stretch", self._stream())

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
        if body_q is None:
            return
        bq = torch.as_tensor(body_q, dtype=torch.float32, device=self.device).reshape(-1, 7)
        if bq.shape[0] != self._S.n_shapes:
            raise ValueError("body_q must hold one (p, q) pose per shape")
        self.buffers["shape_xf"].copy_(bq)

    def set_warm_start(self, net) -> None:
        """Attach a trained ``learned.WarmStartNet`` (or None for plain PCG).
        With a network, every PCG solve first evaluates it and starts from the
        learned guess when that has the smaller residual."""
        self.warm_start_net = net
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
        self._step += 1
        if state_out is not state_in:
            state_out.assign(state_in)
        self._set_body_q(state_out.body_q)
        self._bind_state(state_out.particle_q, state_out.particle_qd)
        self._call("mfem_step", ctypes.c_float(dt), self._stream())

    def step_instrumented(self, state_in: State, state_out: State, dt: float, timers: dict) -> None:
        """Same kernels and order as ``step`` (mfem_step), issued stage by stage
        so that torch.cuda.Events on the launching stream bracket each phase.
        ``timers[phase]`` collects (start, end) event pairs."""
        self._step += 1
        if state_out is not state_in:
            state_out.assign(state_in)
        self._set_body_q(state_out.body_q)
        self._bind_state(state_out.particle_q, state_out.particle_qd)

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
            timed("global_solve", self.global_solve, it)
            timed("local_solve", self.local_solve)
            timed("line_search", self.line_search, dt, it)
        timed("end_step", self.end_step, dt)

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
        rc = _ffi_mlp(ctypes.c_void_p(b["ws_x"].data_ptr()), ctypes.c_int(self.max_particles),
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

    # diagnostics ------------------------------------------------------------
    def last_stats(self) -> np.ndarray:
        """Per Newton iteration of the last step: cg_iters, cg_relres (FP64),
        ls_trials, ls_accepted, alpha, f0, f1, cg_restarts, PCG phase ticks
        (6), learned guess used, learned / plain initial residual ratio."""
        torch.cuda.synchronize(self.device)
        st = self.buffers["stats"].cpu().numpy().reshape(_ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)
        return st[: self.iterations].copy()

    def status(self) -> int:
        """Device error bits (bit 0: a grid barrier timed out; the persistent
        kernels then do nothing until ``reset_status``)."""
        torch.cuda.synchronize(self.device)
        return int(self.buffers["status"][0].item())

    def reset_status(self) -> None:
        """Clears the error bits, the barrier epoch and the arrival flags."""
        torch.cuda.synchronize(self.device)
        self.buffers["status"].zero_()
        self.buffers["flags"].zero_()
        torch.cuda.synchronize(self.device)

    def system_csr(self):
        """Assembled H as a scipy.sparse.bsr_matrix (host copy)."""
        import scipy.sparse as sp
        from .pattern import sell_to_csr_values
        vals = sell_to_csr_values(self.pattern, self.buffers["sell_val"].cpu().numpy())
        n = self.max_particles
        return sp.bsr_matrix((vals, self.pattern.cols, self.pattern.rowptr), shape=(3 * n, 3 * n))

    @property
    def tet_stretch(self) -> torch.Tensor:
        return self.buffers["s"]

    @property
    def tet_lambda(self) -> torch.Tensor:
        return self.buffers["lam"]

    @property
    def elastic_energy(self) -> torch.Tensor:
        return self.buffers["psi"]

    def write_timings(self, filename: str) -> None:
        with open(filename, "w") as f:
            json.dump(self._timings, f)
