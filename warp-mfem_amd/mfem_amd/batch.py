"""Many independent soft bodies solved together on one GPU.

The bodies are concatenated into one mesh whose vertex ranges start at
multiples of 256 * a rows (a persistent-PCG workgroup holds 4 waves of 1, 2,
4 or 8 chunks of 64 rows; with a = the chunks per wave every workgroup
belongs to one body; unused rows are inert pinned padding vertices).  The
smallest a whose rows fit one co-resident grid is chosen.  Every per-element kernel runs once over all bodies; the
persistent PCG and line-search kernels run one launch per Newton iteration
for the whole batch, each body with its own grid reductions (its own
workgroups), CG scalars, convergence test, line-search step and stats, so a
converged body stops while the others continue.  The reference solves one
body per solver; this has no counterpart there.

The bodies share the time step, the Newton iteration count, the solver
parameters and the rigid shapes (the first model's); energy model and
material vary per body.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .scene import Model
from .solver import RefinementSolver, State

def union_model(models: list, far: float = 1.0e3, align: int = 256):
    """(union Model, body_vert, body_tet, body_tri, tet_energy); every body's
    vertex range starts at a multiple of ``align``."""
    qs, qds, ms, ims, ts, ps, mats, trs, en = [], [], [], [], [], [], [], [], []
    bv, bt, bf = [0], [0], [0]
    nv = 0
    for k, m in enumerate(models):
        V = m.particle_count
        pad = (-V) % align
        q = np.asarray(m.particle_q, np.float64)
        qs.append(q)
        qds.append(np.asarray(m.particle_qd, np.float64))
        ms.append(np.asarray(m.particle_mass, np.float64))
        ims.append(np.asarray(m.particle_inv_mass, np.float64))
        if pad:  # inert pinned padding vertices, far from every shape
            qs.append(np.tile([0.0, 0.0, far], (pad, 1)))
            qds.append(np.zeros((pad, 3)))
            ms.append(np.zeros(pad))
            ims.append(np.zeros(pad))
        ts.append(np.asarray(m.tet_indices, np.int64) + nv)
        ps.append(np.asarray(m.tet_poses, np.float64).reshape(-1, 3, 3))
        mats.append(np.asarray(m.tet_materials, np.float64).reshape(-1, 3))
        tri = np.zeros((0, 3)) if m.tri_indices is None else np.asarray(m.tri_indices, np.int64)
        trs.append(tri.reshape(-1, 3) + nv)
        en.append(np.full(m.tet_count, 0 if m.params.get("energy", "arap") == "arap" else 1, np.int32))
        nv += V + pad
        bv.append(nv)
        bt.append(bt[-1] + m.tet_count)
        bf.append(bf[-1] + len(tri))
    m0 = models[0]
    u = Model(particle_q=np.concatenate(qs), particle_qd=np.concatenate(qds), particle_mass=np.concatenate(ms),
              particle_inv_mass=np.concatenate(ims), tet_indices=np.concatenate(ts).astype(np.int32),
              tet_poses=np.concatenate(ps), tet_materials=np.concatenate(mats),
              tri_indices=np.concatenate(trs).astype(np.int32), shapes=list(m0.shapes), gravity=np.asarray(m0.gravity),
              params=dict(m0.params), name=f"batch_of_{len(models)}")
    # the alignment rows after the last body are part of the layout: every
    # row is active (not only those up to the largest tet vertex)
    u.active_particle_count = nv
    return u, np.array(bv), np.array(bt), np.array(bf), np.concatenate(en)


def pick_alignment(vert_counts, device="cuda") -> int:
    """Smallest a in (1, 2, 4, 8) for which the bodies, padded to multiples
    of 256 * a rows, fit some streamed-row persistent-PCG variant with a's
    chunks per wave (or fewer) in one co-resident grid."""
    from . import _ffi

    lib = _ffi.load()
    torch.cuda.set_device(torch.device(device))
    info = (ctypes.c_int32 * 5)()
    for a in (1, 2, 4, 8):
        rows = sum(v + (-v) % (256 * a) for v in vert_counts)
        S = _ffi.MfemSolverStruct()
        S.n_chunks, S.max_width, S.n_bodies, S.body_wg_align = rows // 64, 1 << 30, len(vert_counts), a
        for k in range(int(lib.mfem_pcg_variant_count())):
            if lib.mfem_pcg_variant_info(ctypes.byref(S), k, info) != 0:
                raise _ffi.MfemError(lib.mfem_last_error().decode())
            if info[4]:
                return a
    raise ValueError(f"{sum(vert_counts)} vertices in {len(vert_counts)} bodies do not fit one persistent grid")


class BatchSolver:
    """RefinementSolver-style stepping of a list of models as one batch
    (refinement off).  ``state`` tensors hold the union's vertices;
    ``body_slice(b)`` gives body b's real rows."""

    def __init__(self, models: list, iterations: int, device="cuda", body_wg_align: int | None = None, **kwargs):
        self.models = list(models)
        if body_wg_align is None:
            body_wg_align = pick_alignment([m.particle_count for m in self.models], device)
        self.body_wg_align = int(body_wg_align)
        self.union, self.body_vert, self.body_tet, self.body_tri, tet_energy = union_model(
            self.models, align=256 * self.body_wg_align)
        kw = dict(kwargs)
        kw["enable_refinement"] = False
        self.solver = RefinementSolver(self.union, iterations, max_tets=self.union.tet_count, device=device, **kw)
        self.solver._set_bodies(self.body_vert, self.body_tet, self.body_tri, tet_energy)
        self.device = self.solver.device
        self.iterations = self.solver.iterations

    @property
    def n_bodies(self) -> int:
        return len(self.models)

    @property
    def tet_count(self) -> int:
        return int(self.body_tet[-1])

    def body_slice(self, b: int) -> slice:
        return slice(int(self.body_vert[b]), int(self.body_vert[b]) + self.models[b].particle_count)

    def make_state(self) -> State:
        u = self.union
        q = torch.as_tensor(np.asarray(u.particle_q, np.float32), device=self.device).contiguous()
        qd = torch.as_tensor(np.asarray(u.particle_qd, np.float32), device=self.device).contiguous()
        return State(q, qd)

    def step(self, state_in: State, state_out: State, dt: float) -> None:
        self.solver.step(state_in, state_out, None, None, dt)

    def body_stats(self, b: int) -> np.ndarray:
        return self.solver.body_stats(b)

    def status(self) -> int:
        return self.solver.status()
