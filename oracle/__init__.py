"""CPU oracle for the per-timestep soft-body solve (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
load this package, and only as the checker / CPU baseline: the product path
(``mfem_amd``) never imports it.

``Oracle`` wraps ``liboracle64.so`` (FP64) or ``liboracle32.so`` (FP32), both
built from ``mfem_oracle.c`` by ``oracle/Makefile``; ``Oracle(..., threads=n)``
loads the OpenMP builds (``liboracle{64,32}_omp.so``) instead, which time the
CPU baseline (their sums and scatters run in another order, so they are not
the bit-reproducible checker).  Each stage function
mirrors one phase of ``RefinementSolver.step``
(src/mfem/refinement/solver.py:429-579 of the reference).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[tuple, ctypes.CDLL] = {}


def build() -> None:
    """Compile both oracle variants (gcc, seconds)."""
    subprocess.run(["make", "-s", "-B", "-C", _DIR], check=True)


def _lib(bits: int, omp: bool = False) -> ctypes.CDLL:
    if (bits, omp) in _LIBS:
        return _LIBS[(bits, omp)]
    path = os.path.join(_DIR, f"liboracle{bits}{'_omp' if omp else ''}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    real = ctypes.c_double if bits == 64 else ctypes.c_float
    P = ctypes.c_void_p
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [ctypes.c_int] * 5
    lib.orc_set_param.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
    lib.orc_set_param.restype = ctypes.c_int
    lib.orc_get_param.argtypes = [P, ctypes.c_char_p]
    lib.orc_get_param.restype = ctypes.c_double
    lib.orc_buf.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
    lib.orc_buf.restype = P
    lib.orc_finalize.argtypes = [P]
    lib.orc_finalize.restype = ctypes.c_int
    for name in ("orc_init_stretch", "orc_local_solve", "orc_destroy"):
        getattr(lib, name).argtypes = [P]
        getattr(lib, name).restype = None
    for name in ("orc_begin_step", "orc_update_system", "orc_assemble", "orc_step"):
        getattr(lib, name).argtypes = [P, real]
        getattr(lib, name).restype = None
    lib.orc_global_solve.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.orc_global_solve.restype = ctypes.c_int
    lib.orc_line_search.argtypes = [P, real, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    lib.orc_line_search.restype = ctypes.c_int
    lib.orc_merit.argtypes = [P, P, P, real]
    lib.orc_merit.restype = ctypes.c_double
    lib.orc_spmv.argtypes = [P, P, P]
    lib.orc_spmv.restype = None
    lib.orc_real_size.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    lib.orc_set_surface.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P]
    lib.orc_set_surface.restype = None
    lib.orc_sc_distance.argtypes = [ctypes.c_int, P, P, P]
    lib.orc_sc_distance.restype = ctypes.c_double
    lib.orc_sc_accd.argtypes = [ctypes.c_int, P, P, ctypes.c_double]
    lib.orc_sc_accd.restype = ctypes.c_double
    for name in ("orc_sc_energy", "orc_sc_gradient"):
        getattr(lib, name).argtypes = [P, P]
        getattr(lib, name).restype = ctypes.c_double
    lib.orc_sc_toi.argtypes = [P, P, P]
    lib.orc_sc_toi.restype = ctypes.c_double
    lib.orc_shell_refresh.argtypes = [P]
    lib.orc_shell_refresh.restype = None
    lib.orc_tri_rest_frame.argtypes = [P, P, P, P]
    lib.orc_tri_rest_frame.restype = real
    lib.orc_membrane.argtypes = [P, P, P, P, real, real, P]
    lib.orc_membrane.restype = real
    _LIBS[(bits, omp)] = lib
    return lib


# RefinementSolver keyword (src/mfem/refinement/solver.py:56-128, 241-245) -> oracle parameter
_PARAM_NAMES = {
    "cg_max_iterations": "cg_maxiter", "cg_tolerance": "cg_tol", "preconditioner": "precond",
    "line_search_max_iterations": "ls_max", "line_search_alpha0": "ls_alpha0", "line_search_tau": "ls_tau",
    "line_search_c": "ls_c", "line_search_threshold": "ls_thr", "attachment_stiffness": "attach_k",
    "contact_d0": "d0", "contact_d1": "d1", "contact_stiffness": "kappa", "friction_eps_v": "eps_v",
    "preconditioner_singular_threshold": "sing_thr",
    "self_contact_dhat": "sc_dhat", "self_contact_kappa": "sc_kappa", "self_contact_ccd_s": "sc_ccd_s",
}
# solver keywords with no oracle counterpart (GPU data-structure sizes,
# refinement: the oracle steps a fixed mesh)
_GPU_ONLY = {"self_contact_max_pairs", "self_contact_cell", "self_contact_max_span", "refine_tet_score_weight",
             "refine_vertex_score_weight", "refine_density"}


def _real(bits: int):
    return np.float64 if bits == 64 else np.float32


class Kernels:
    """Per-element oracle functions (no solver context)."""

    def __init__(self, bits: int = 64):
        self.bits = bits
        self.lib = _lib(bits)
        self.dt = _real(bits)

    def _arr(self, a, n=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dt).reshape(-1))
        if n is not None:
            assert a.size == n
        return a

    def _p(self, a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def svd3(self, F):
        F = self._arr(F, 9)
        U, s, V = np.zeros(9, self.dt), np.zeros(3, self.dt), np.zeros(9, self.dt)
        self.lib.orc_svd3(self._p(F), self._p(U), self._p(s), self._p(V))
        return U.reshape(3, 3), s, V.reshape(3, 3)

    def tet_constraints(self, x12, dminv, s6):
        x12, D, s6 = self._arr(x12, 12), self._arr(dminv, 9), self._arr(s6, 6)
        c, G, S6 = np.zeros(6, self.dt), np.zeros(72, self.dt), np.zeros(6, self.dt)
        self.lib.orc_tet_constraints(self._p(x12), self._p(D), self._p(s6), self._p(c), self._p(G), self._p(S6))
        return c, G.reshape(6, 12), S6

    def stretch_gradient(self, F):
        F = self._arr(F, 9)
        U, s, V = self.svd3(F)
        U, s, V = self._arr(U), self._arr(s), self._arr(V)
        out = np.zeros(54, self.dt)
        self.lib.orc_stretch_gradient(self._p(F), self._p(U), self._p(s), self._p(V), self._p(out))
        return out.reshape(6, 9)

    def psd_fix(self, A):
        A = np.asarray(A)
        n = A.shape[0]
        a = self._arr(A, n * n)
        fixed, inv = np.zeros(n * n, self.dt), np.zeros(n * n, self.dt)
        self.lib.orc_psd_fix(self._p(a), ctypes.c_int(n), self._p(fixed), self._p(inv))
        return fixed.reshape(n, n), inv.reshape(n, n)

    def gmw81(self, A):
        A = np.asarray(A)
        n = A.shape[0]
        a = self._arr(A, n * n)
        L = np.zeros(n * n, self.dt)
        self.lib.orc_gmw81(self._p(a), ctypes.c_int(n), self._p(L))
        return L.reshape(n, n)

    def arap(self, s6, mu, vol):
        s6 = self._arr(s6, 6)
        E = np.zeros(1, self.dt)
        g, H, Hi = np.zeros(6, self.dt), np.zeros(36, self.dt), np.zeros(36, self.dt)
        real = ctypes.c_double if self.bits == 64 else ctypes.c_float
        self.lib.orc_arap(self._p(s6), real(mu), real(vol), self._p(E), self._p(g), self._p(H), self._p(Hi))
        return float(E[0]), g, H.reshape(6, 6), Hi.reshape(6, 6)

    def neohookean(self, s6, mu, lam, vol):
        s6 = self._arr(s6, 6)
        E = np.zeros(1, self.dt)
        g, H, Hi, Hr = (np.zeros(n, self.dt) for n in (6, 36, 36, 36))
        real = ctypes.c_double if self.bits == 64 else ctypes.c_float
        self.lib.orc_neohookean(self._p(s6), real(mu), real(lam), real(vol), self._p(E), self._p(g),
                                self._p(H), self._p(Hi), self._p(Hr))
        return float(E[0]), g, H.reshape(6, 6), Hi.reshape(6, 6), Hr.reshape(6, 6)

    def barrier(self, d, d0, d1):
        real = ctypes.c_double if self.bits == 64 else ctypes.c_float
        q = np.zeros(3, self.dt)
        self.lib.orc_barrier(real(d0), real(0.0), real(d1), self._p(q), self._p(q))
        out = np.zeros(3, self.dt)
        self.lib.orc_barrier(real(d), real(d0), real(d1), self._p(q), self._p(out))
        return out

    def _real_t(self):
        return ctypes.c_double if self.bits == 64 else ctypes.c_float

    def friction(self, x, x_start, n, lamn, mu, ev):
        """Lagged friction of one vertex: (energy, gradient (3,), Hessian (3, 3))."""
        x, xs, n = (self._arr(v, 3) for v in (x, x_start, n))
        g, H = np.zeros(3, self.dt), np.zeros(9, self.dt)
        real = self._real_t()
        self.lib.orc_friction.restype = real
        E = self.lib.orc_friction(self._p(x), self._p(xs), self._p(n), real(lamn), real(mu), real(ev), self._p(g),
                                  self._p(H))
        return float(E), g, H.reshape(3, 3)

    def shape_sdf(self, shape_type, xf, scale, x):
        """Signed distance of x to one shape posed by xf = p(3), q(4) (xyzw):
        (d, gradient (3,), Hessian (3, 3))."""
        xf, sc, x = self._arr(xf, 7), self._arr(scale, 3), self._arr(x, 3)
        g, H = np.zeros(3, self.dt), np.zeros(9, self.dt)
        self.lib.orc_shape_sdf.restype = self._real_t()
        d = self.lib.orc_shape_sdf(ctypes.c_int(shape_type), self._p(xf), self._p(sc), self._p(x), self._p(g),
                                   self._p(H))
        return float(d), g, H.reshape(3, 3)

    def tri_barrier(self, verts, shapes, d0, d1, kappa):
        """One triangle against capsules ``shapes`` = [(type, xf(7), scale(3))]:
        (energy, per-corner gradient (3, 3), Hessian blocks (3, 3, 3, 3),
        distance, barycentric weights, normal)."""
        v = self._arr(verts, 9)
        types = np.ascontiguousarray([s[0] for s in shapes], dtype=np.int32)
        xf = self._arr(np.concatenate([np.asarray(s[1], float) for s in shapes]))
        sc = self._arr(np.concatenate([np.asarray(s[2], float) for s in shapes]))
        q = self.barrier(d0, 0.0, d1)
        q = self._arr(q, 3)
        d = np.zeros(1, self.dt)
        w, n, g, H = np.zeros(3, self.dt), np.zeros(3, self.dt), np.zeros(9, self.dt), np.zeros(81, self.dt)
        real = self._real_t()
        self.lib.orc_tri_barrier.restype = real
        E = self.lib.orc_tri_barrier(self._p(v[0:3].copy()), self._p(v[3:6].copy()), self._p(v[6:9].copy()),
                                     ctypes.c_int(len(shapes)), types.ctypes.data_as(ctypes.c_void_p), self._p(xf),
                                     self._p(sc), real(d0), real(d1), real(kappa), self._p(q), self._p(d),
                                     self._p(w), self._p(n), self._p(g), self._p(H))
        return float(E), g.reshape(3, 3), H.reshape(3, 3, 3, 3), float(d[0]), w, n

    def tri_rest_frame(self, r0, r1, r2):
        """(W (2, 2) inverse in-plane rest-shape matrix, rest area) of one
        rest triangle (src/mfem/refinement/shell_energy.py:37-63)."""
        r0, r1, r2 = (self._arr(v, 3) for v in (r0, r1, r2))
        W = np.zeros(4, self.dt)
        area = self.lib.orc_tri_rest_frame(self._p(r0), self._p(r1), self._p(r2), self._p(W))
        return W.reshape(2, 2), float(area)

    def membrane(self, x0, x1, x2, W, area, mu):
        """(energy, gradient (3, 3) per corner) of one membrane triangle
        (shell_energy.py:66-125, 190-229)."""
        x0, x1, x2 = (self._arr(v, 3) for v in (x0, x1, x2))
        W = self._arr(W, 4)
        g = np.zeros(9, self.dt)
        real = self._real_t()
        E = self.lib.orc_membrane(self._p(x0), self._p(x1), self._p(x2), self._p(W), real(area), real(mu), self._p(g))
        return float(E), g.reshape(3, 3)

    def closest_point_triangle_bary(self, x, a, b, c):
        x, a, b, c = (self._arr(v, 3) for v in (x, a, b, c))
        w = np.zeros(3, self.dt)
        self.lib.orc_closest_point_triangle_bary(self._p(x), self._p(a), self._p(b), self._p(c), self._p(w))
        return w

    def closest_point_triangle_segment(self, a, b, c, p0, p1):
        a, b, c, p0, p1 = (self._arr(v, 3) for v in (a, b, c, p0, p1))
        bary, q = np.zeros(3, self.dt), np.zeros(3, self.dt)
        hit = ctypes.c_int(0)
        self.lib.orc_closest_point_triangle_segment.restype = ctypes.c_double if self.bits == 64 else ctypes.c_float
        d = self.lib.orc_closest_point_triangle_segment(self._p(a), self._p(b), self._p(c), self._p(p0),
                                                         self._p(p1), self._p(bary), self._p(q), ctypes.byref(hit))
        return bary, q, float(d), hit.value


class Oracle:
    """Full-solver oracle context.  Construct from a ``mfem_amd`` Scene-like
    object (numpy fields), then call ``step(dt)`` or the stage methods."""

    INT_FIELDS = {"tets", "pin_idx", "tris", "shape_type", "rowptr", "col", "diag", "lag_sid"}

    def __init__(self, scene, bits: int = 64, params: dict | None = None, threads: int | None = None,
                 rest=None):
        self.bits = bits
        self.lib = _lib(bits, omp=threads is not None)
        self.threads = self.lib.orc_set_threads(threads) if threads is not None else 1
        self.dt = _real(bits)
        # the active particles (a model padded for refinement headroom carries
        # unreferenced rows after them)
        nt = scene.tet_indices.shape[0]
        nv = int(np.asarray(scene.tet_indices).max()) + 1 if nt else scene.particle_q.shape[0]
        ntris = 0 if scene.tri_indices is None else scene.tri_indices.shape[0]
        nshapes = len(scene.shapes)
        pins = np.nonzero(np.asarray(scene.particle_inv_mass)[:nv] == 0)[0].astype(np.int32)
        self.ctx = self.lib.orc_create(nv, nt, ntris, nshapes, len(pins))
        self._pins = pins
        self.nv, self.nt = nv, nt
        self.buf("tets")[:] = scene.tet_indices.reshape(-1)
        self.buf("dminv")[:] = scene.tet_poses.reshape(-1)
        self.buf("materials")[:] = scene.tet_materials.reshape(-1)
        self.buf("mass")[:] = np.asarray(scene.particle_mass)[:nv]
        self.buf("inv_mass")[:] = np.asarray(scene.particle_inv_mass)[:nv]
        if len(pins):
            self.buf("pin_idx")[:] = pins
            self.buf("pin_pos")[:] = scene.particle_q[pins].reshape(-1)
        if ntris:
            self.buf("tris")[:] = scene.tri_indices.reshape(-1)
        for k, sh in enumerate(scene.shapes):
            self.buf("shape_type")[k] = sh.type
            self.buf("shape_scale")[3 * k:3 * k + 3] = sh.scale
        if nshapes:
            self.set_body_q(getattr(scene, "body_q", None), scene)
        self.buf("x")[:] = np.asarray(scene.particle_q)[:nv].reshape(-1)
        self.buf("v")[:] = np.asarray(scene.particle_qd)[:nv].reshape(-1)
        p = dict(scene.params)
        if params:
            p.update(params)
        mu = np.asarray(p.pop("friction_mu", 0.0), dtype=np.float64)
        mu = np.full(nshapes, float(mu)) if mu.ndim == 0 else mu
        if nshapes:
            self.buf("shape_mu")[:] = mu
        op = dict(gx=float(scene.gravity[0]), gy=float(scene.gravity[1]), gz=float(scene.gravity[2]))
        if p.get("self_contact", False):
            # the solver's defaults: the vertex barrier's range and stiffness
            p.setdefault("self_contact_dhat", p.get("contact_d1", 1.0e-1))
            p.setdefault("self_contact_kappa", p.get("contact_stiffness", 1.0e7))
            tri = np.asarray(scene.tri_indices, np.int32).reshape(-1, 3)
            pts = np.unique(tri.reshape(-1)).astype(np.int32)
            e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
            e = np.unique(e, axis=0).astype(np.int32)
            self.lib.orc_set_surface(self.ctx, len(pts), pts.ctypes.data_as(ctypes.c_void_p), len(e),
                                     np.ascontiguousarray(e).ctypes.data_as(ctypes.c_void_p))
            self.surface_points, self.surface_edges = pts, e
        for k in _GPU_ONLY:
            p.pop(k, None)
        for k, v in p.items():
            if k == "energy":
                v = {"arap": 0, "neohookean": 1}.get(v, v)
            op[_PARAM_NAMES.get(k, k)] = v
        for k, v in op.items():
            if self.lib.orc_set_param(self.ctx, k.encode(), float(v)) != 0:
                raise KeyError(k)
        if self.lib.orc_finalize(self.ctx) != 0:
            raise ValueError("bad topology")
        # rest positions (AdditionalState.rest_particle_q): the model's
        # positions unless given (a refined mesh's interpolated rest)
        self.buf("rest")[:] = np.asarray(scene.particle_q if rest is None else rest)[:nv].reshape(-1)
        if ntris:
            self.lib.orc_shell_refresh(self.ctx)
        self.lib.orc_init_stretch(self.ctx)

    def set_body_q(self, body_q, scene=None) -> None:
        """World poses of the shapes for rigid body poses body_q (n_bodies, 7):
        body_q[shape.body] * (shape.pos, shape.quat), or the shape's own pose
        for a static shape (src/mfem/refinement/contact.py:74-82)."""
        scene = scene if scene is not None else self._scene
        self._scene = scene
        out = []
        for sh in scene.shapes:
            p, q = np.asarray(sh.pos, float), np.asarray(sh.quat, float)
            b = getattr(sh, "body", -1)
            if b >= 0:
                t = np.asarray(body_q, float).reshape(-1, 7)[b]
                u, w = t[3:6], t[6]
                p = t[:3] + p + 2.0 * np.cross(u, np.cross(u, p) + w * p)
                x1, y1, z1, w1 = t[3:]
                x2, y2, z2, w2 = q
                q = np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                              w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
            out.append(np.concatenate([p, q]))
        self.buf("shape_xf")[:] = np.concatenate(out)

    def __del__(self):
        try:
            self.lib.orc_destroy(self.ctx)
        except Exception:
            pass

    def buf(self, name: str) -> np.ndarray:
        n = ctypes.c_int64(0)
        k = ctypes.c_int(0)
        ptr = self.lib.orc_buf(self.ctx, name.encode(), ctypes.byref(n), ctypes.byref(k))
        if not ptr:
            raise KeyError(name)
        kinds = {0: (self.dt, ctypes.c_double if self.bits == 64 else ctypes.c_float), 1: (np.int32, ctypes.c_int32),
                 2: (np.float64, ctypes.c_double)}
        npt, ctype = kinds[k.value]
        if n.value == 0:
            return np.zeros(0, npt)
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n.value,))

    def param(self, name: str) -> float:
        return self.lib.orc_get_param(self.ctx, name.encode())

    def set_param(self, name: str, value: float) -> None:
        if self.lib.orc_set_param(self.ctx, name.encode(), float(value)) != 0:
            raise KeyError(name)

    # stages -------------------------------------------------------------
    def begin_step(self, dt):
        self.lib.orc_begin_step(self.ctx, dt)

    def update_system(self, dt):
        self.lib.orc_update_system(self.ctx, dt)

    def assemble(self, dt):
        self.lib.orc_assemble(self.ctx, dt)

    def global_solve(self):
        rr = ctypes.c_double(0)
        it = self.lib.orc_global_solve(self.ctx, ctypes.byref(rr))
        return it, rr.value

    def local_solve(self):
        self.lib.orc_local_solve(self.ctx)

    def line_search(self, dt):
        acc = ctypes.c_int(0)
        alpha = ctypes.c_double(0)
        it = self.lib.orc_line_search(self.ctx, dt, ctypes.byref(acc), ctypes.byref(alpha))
        return it, bool(acc.value), alpha.value

    def merit(self, x, s, dt):
        x = np.ascontiguousarray(x, dtype=self.dt).reshape(-1)
        s = np.ascontiguousarray(s, dtype=self.dt).reshape(-1)
        return self.lib.orc_merit(self.ctx, x.ctypes.data_as(ctypes.c_void_p), s.ctypes.data_as(ctypes.c_void_p), dt)

    def step(self, dt):
        self.lib.orc_step(self.ctx, dt)

    def set_attachment_targets(self, indices, positions) -> None:
        """New pin positions of the pinned vertices ``indices`` (penalty
        attachments; src/mfem/boundary_condition.py:85-92)."""
        slot = {int(v): k for k, v in enumerate(self._pins)}
        pp = self.buf("pin_pos").reshape(-1, 3)
        for v, p in zip(np.asarray(indices).tolist(), np.asarray(positions, np.float64)):
            pp[slot[int(v)]] = p

    # self-contact (surface point-triangle / edge-edge barrier), all pairs
    def _ptr(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dt).reshape(-1))
        return a, a.ctypes.data_as(ctypes.c_void_p)

    def sc_energy(self, x=None) -> float:
        """kappa sum b(d) over every surface pair below dhat at x (default: the state)."""
        a, p = self._ptr(self.x if x is None else x)
        return float(self.lib.orc_sc_energy(self.ctx, p))

    def sc_gradient(self, x=None):
        """(energy, gradient (V, 3), Gauss-Newton diagonal blocks (V, 6) in sym6 order)."""
        a, p = self._ptr(self.x if x is None else x)
        e = float(self.lib.orc_sc_gradient(self.ctx, p))
        return e, self.buf("sc_g").reshape(-1, 3).copy(), self.buf("sc_h").reshape(-1, 6).copy()

    def sc_toi(self, x, dx) -> float:
        """Smallest additive-CCD time of impact (<= 1) over the pairs whose
        boxes swept by dx overlap."""
        a, p = self._ptr(x)
        b, q = self._ptr(dx)
        return float(self.lib.orc_sc_toi(self.ctx, p, q))

    def stats(self):
        n = int(self.param("iterations"))
        return self.buf("stats")[: 4 * n].reshape(n, 4).copy()

    def csr(self):
        """Assembled H as a scipy.sparse.bsr_matrix (3x3 blocks)."""
        import scipy.sparse as sp
        rowptr = self.buf("rowptr").copy()
        col = self.buf("col").copy()
        val = self.buf("val").reshape(-1, 3, 3).copy()
        return sp.bsr_matrix((val, col, rowptr), shape=(3 * self.nv, 3 * self.nv))

    @property
    def x(self):
        return self.buf("x").reshape(-1, 3)

    @property
    def v(self):
        return self.buf("v").reshape(-1, 3)
