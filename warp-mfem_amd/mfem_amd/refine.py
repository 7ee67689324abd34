"""One adaptive-refinement pass on the GPU (csrc/mfem_refine.hip).

Replaces refine() of src/mfem/refinement/refinement.py:914-1322 ("legacy"
and "geometric" scoring) with a deterministic pass: candidate edges scored per incident tet
(contact barrier energy, elastic energy, edge length; a penetrating midpoint
always wins), split edges selected greedily by (score, edge key) with at
most one split edge per tet, every tet on a split edge split in two, new
vertices interpolated at ``split_t``, surface triangles split alike.  The
capacity guard drops a pass that would exceed max_tets / max_verts.
"""
from __future__ import annotations

import ctypes

import torch

from . import _ffi

vp = ctypes.c_void_p


class RefineArgs(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("tets", "tris", "x", "inv_mass", "tet_score", "vertex_score", "shape_type",
                                  "shape_xf", "shape_scale", "stretch", "lam", "materials", "dminv")] + \
               [(n, ctypes.c_int32) for n in ("n_verts", "n_tets", "n_tris", "n_shapes", "max_tets", "max_tris",
                                              "max_verts")] + \
               [(n, ctypes.c_float) for n in ("threshold", "w_tet", "w_vertex", "penetrating_score", "split_t")] + \
               [("scoring", ctypes.c_int32)] + \
               [(n, ctypes.c_float) for n in ("min_edge_length", "contact_d1", "w_elastic", "w_vertex_contact",
                                              "w_tri_contact")] + \
               [(n, vp) for n in ("vertex_dist", "vertex_sid", "tri_dist", "tri_bary")]


class RefineOut(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("x", "v", "rest", "tets", "stretch", "lam", "materials", "dminv", "tris")]


def _p(t):
    return None if t is None else t.data_ptr()


class Refiner:
    """Workspace for passes on meshes up to the given capacities."""

    MAX_ROUNDS = 4096

    def __init__(self, max_tets: int, max_tris: int, max_verts: int, device="cuda"):
        self.lib = _ffi.load()
        self.device = torch.device(device)
        self.max_tets, self.max_tris, self.max_verts = int(max_tets), int(max(max_tris, 1)), int(max_verts)
        nb = int(self.lib.mfem_refine_workspace_bytes(self.max_tets, self.max_tris))
        self.ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.changed = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.totals = torch.zeros(4, dtype=torch.int32, device=self.device)

    def _stream(self):
        return vp(torch.cuda.current_stream(self.device).cuda_stream)

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            raise _ffi.MfemError(f"{name} failed ({rc}): {self.lib.mfem_refine_last_error().decode()}")

    def run(self, *, tets, tris, x, v, rest, inv_mass, tet_score, vertex_score, stretch, lam, materials, dminv,
            n_verts: int, n_tets: int, n_tris: int, shapes=None, threshold: float = 10.1, w_tet: float = 0.0,
            w_vertex: float = 0.1, penetrating_score: float = 1.0e6, split_t: float = 0.5, out: dict,
            scoring: str = "legacy", min_edge_length: float = 0.0, contact_d1: float = 0.0,
            w_elastic: float = 1.0, w_vertex_contact: float = 1.0, w_tri_contact: float = 1.0,
            vertex_dist=None, vertex_sid=None, tri_dist=None, tri_bary=None) -> dict:
        """One pass.  Inputs are device tensors of the current mesh (x, v,
        rest sized max_verts: new vertices are appended in place); ``out``
        holds tets / stretch / lam / materials / dminv / tris buffers of the
        capacities.  ``shapes`` = (shape_type, shape_xf, shape_scale) device
        tensors or None.  ``scoring="geometric"`` scores edges as
        populate_candidates_geometric (refinement.py:196-317) from the tet
        energies, the per-vertex distances / nearest shapes and, when given,
        the triangle distances / barycentric weights, with ``threshold`` the
        geometric threshold.  Returns the new counts, the number of split
        edges ("new_verts"), whether the pass was dropped and the rounds
        taken."""
        if scoring not in ("legacy", "geometric"):
            raise ValueError(f"Unknown refinement scoring {scoring!r} (expected 'legacy' or 'geometric')")
        if scoring == "geometric":
            if vertex_dist is None or vertex_sid is None:
                raise ValueError("scoring='geometric' needs particle_distance and particle_shape_id")
            if not (min_edge_length > 0.0 and contact_d1 > 0.0):
                raise ValueError("scoring='geometric' needs min_edge_length > 0 and contact_d1 > 0")
        A = RefineArgs()
        A.scoring = 1 if scoring == "geometric" else 0
        A.min_edge_length, A.contact_d1 = float(min_edge_length), float(contact_d1)
        A.w_elastic, A.w_vertex_contact, A.w_tri_contact = float(w_elastic), float(w_vertex_contact), float(w_tri_contact)
        A.vertex_dist, A.vertex_sid, A.tri_dist, A.tri_bary = _p(vertex_dist), _p(vertex_sid), _p(tri_dist), _p(tri_bary)
        A.tets, A.tris, A.x, A.inv_mass = _p(tets), _p(tris), _p(x), _p(inv_mass)
        A.tet_score, A.vertex_score = _p(tet_score), _p(vertex_score)
        if shapes is not None and shapes[0].numel() > 0:
            A.shape_type, A.shape_xf, A.shape_scale = _p(shapes[0]), _p(shapes[1]), _p(shapes[2])
            A.n_shapes = int(shapes[0].numel())
        A.stretch, A.lam, A.materials, A.dminv = _p(stretch), _p(lam), _p(materials), _p(dminv)
        A.n_verts, A.n_tets, A.n_tris = int(n_verts), int(n_tets), int(n_tris)
        A.max_tets, A.max_tris, A.max_verts = self.max_tets, self.max_tris, self.max_verts
        A.threshold, A.w_tet, A.w_vertex = float(threshold), float(w_tet), float(w_vertex)
        A.penetrating_score, A.split_t = float(penetrating_score), float(split_t)
        st = self._stream()
        ws, nb = vp(self.ws.data_ptr()), self.ws.numel()
        self._call("mfem_refine_select", ctypes.byref(A), ws, nb, vp(self.status.data_ptr()),
                   vp(self.changed.data_ptr()), 0, st)
        rounds = 0
        while True:
            self._call("mfem_refine_select", ctypes.byref(A), ws, nb, vp(self.status.data_ptr()),
                       vp(self.changed.data_ptr()), 1, st)
            rounds += 1
            if int(self.changed.item()) == 0 or rounds >= self.MAX_ROUNDS:
                break
        if int(self.status.item()) != 0:
            raise _ffi.MfemError("refinement edge table full")
        self._call("mfem_refine_count", ctypes.byref(A), ws, nb, vp(self.totals.data_ptr()), st)
        n_tets_new, n_new, n_tris_new = (int(v_) for v_ in self.totals[:3].tolist())
        res = dict(rounds=rounds, new_verts=n_new, n_verts=int(n_verts), n_tets=int(n_tets), n_tris=int(n_tris),
                   dropped=False)
        if n_new == 0:
            return res
        if n_tets_new > self.max_tets or n_verts + n_new > self.max_verts or n_tris_new > self.max_tris:
            res["dropped"] = True  # check_split_capacity: the whole pass is dropped
            return res
        O = RefineOut()
        O.x, O.v, O.rest = _p(x), _p(v), _p(rest)
        for name in ("tets", "stretch", "lam", "materials", "dminv", "tris"):
            setattr(O, name, _p(out[name]))
        self._call("mfem_refine_apply", ctypes.byref(A), ws, nb, ctypes.byref(O), st)
        res.update(n_verts=int(n_verts) + n_new, n_tets=n_tets_new, n_tris=n_tris_new)
        return res
