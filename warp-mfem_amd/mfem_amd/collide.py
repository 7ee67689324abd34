"""Self-contact building blocks on the GPU (csrc/mfem_collide.hip).

* ``scan_i32``: exclusive / inclusive int32 prefix sum (the reference's
  TiledScan / wp.utils.array_scan for int32, src/mfem/scan.py).
* ``broadphase_pt`` / ``broadphase_ee``: spatial-hash candidate pairs
  (point-triangle, edge-edge) whose boxes inflated by d_hat / 2 overlap,
  sorted and duplicate-free (the reference has no live broad phase).
* ``soft_rigid_distance``: the reference's rigid/soft 15-feature minimum
  distance kernel (src/mfem/collision.py:380-476).
* ``pair_barrier``: IPC barrier energy and gradient on candidate pairs
  (src/mfem/collision.py:529-546).
* ``surface_points`` / ``surface_edges``: the primitives of a triangle
  surface.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _ffi

MFEM_BP_SPAN, MFEM_BP_ENTRIES, MFEM_BP_PAIRS = 1, 2, 4


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check(lib, name, rc):
    if rc != 0:
        raise _ffi.MfemError(f"{name} failed ({rc}): {lib.mfem_collide_last_error().decode()}")


def _i32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


def surface_points(tris) -> np.ndarray:
    return np.unique(np.asarray(tris, np.int32).reshape(-1)).astype(np.int32)


def surface_edges(tris) -> np.ndarray:
    """Unique edges (i < j) of a triangle surface, sorted."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0).astype(np.int32)


def scan_i32(a: torch.Tensor, inclusive: bool = False):
    """(prefix sums, total) of an int32 device tensor."""
    lib = _ffi.load()
    a = a.to(torch.int32).contiguous()
    n = a.numel()
    out = torch.empty_like(a)
    total = torch.zeros(1, dtype=torch.int32, device=a.device)
    ws = torch.empty(int(lib.mfem_scan_workspace_bytes(n)), dtype=torch.uint8, device=a.device)
    _check(lib, "mfem_scan_i32", lib.mfem_scan_i32(_p(a), _p(out), n, int(inclusive), _p(total), _p(ws),
                                                   ws.numel(), _stream(a.device)))
    return out, total


def default_cell(x: torch.Tensor, prims: torch.Tensor, d_hat: float) -> float:
    """Largest inflated box extent of the primitives: with this cell size no
    primitive spans more than 2 cells per axis."""
    p = x[prims.long()]
    ext = (p.amax(dim=1) - p.amin(dim=1)).max().item() if prims.numel() else 0.0
    return float(max(ext + d_hat, d_hat, 1e-6)) * (1.0 + 1e-4)


class BroadPhase:
    """Candidate lists for a fixed triangle surface (reusable workspace)."""

    def __init__(self, tris, n_verts: int, device="cuda", max_pairs: int = 1 << 20, max_span: int = 2):
        self.lib = _ffi.load()
        self.device = torch.device(device)
        self.tris = _i32(np.asarray(tris).reshape(-1, 3), self.device)
        self.points = _i32(surface_points(tris), self.device)
        self.edges = _i32(surface_edges(tris), self.device)
        self.n_verts = int(n_verts)
        self.max_pairs = int(max_pairs)
        self.max_span = int(max_span)
        nt, npnt, ne = self.tris.shape[0], self.points.shape[0], self.edges.shape[0]
        ws = max(self.lib.mfem_broadphase_workspace_bytes(nt, npnt, self.max_span),
                 self.lib.mfem_broadphase_workspace_bytes(ne, ne, self.max_span))
        self.ws = torch.empty(int(ws), dtype=torch.uint8, device=self.device)
        self.pairs = torch.empty(self.max_pairs, 2, dtype=torch.int32, device=self.device)
        self.n_pairs = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)

    def _result(self):
        n = int(self.n_pairs.item())
        st = int(self.status.item())
        if st:
            raise _ffi.MfemError(f"broad phase overflow (status {st}, {n} pairs, capacity {self.max_pairs})")
        return self.pairs[:n].clone()

    def point_triangle(self, x: torch.Tensor, d_hat: float, cell: float | None = None) -> torch.Tensor:
        """(vertex, tri) pairs, sorted by vertex then triangle."""
        x = x.to(self.device, torch.float32).contiguous()
        cell = default_cell(x, self.tris, d_hat) if cell is None else cell
        rc = self.lib.mfem_broadphase_pt(_p(x), _p(self.points), self.points.shape[0], _p(self.tris),
                                         self.tris.shape[0], ctypes.c_float(d_hat), ctypes.c_float(cell),
                                         self.max_span, _p(self.pairs), self.max_pairs, _p(self.n_pairs),
                                         _p(self.status), _p(self.ws), self.ws.numel(), _stream(self.device))
        _check(self.lib, "mfem_broadphase_pt", rc)
        return self._result()

    def edge_edge(self, x: torch.Tensor, d_hat: float, cell: float | None = None) -> torch.Tensor:
        """(edge, edge) index pairs into ``self.edges``, e1 < e2, sorted."""
        x = x.to(self.device, torch.float32).contiguous()
        cell = default_cell(x, self.edges, d_hat) if cell is None else cell
        rc = self.lib.mfem_broadphase_ee(_p(x), _p(self.edges), self.edges.shape[0], ctypes.c_float(d_hat),
                                         ctypes.c_float(cell), self.max_span, _p(self.pairs), self.max_pairs,
                                         _p(self.n_pairs), _p(self.status), _p(self.ws), self.ws.numel(),
                                         _stream(self.device))
        _check(self.lib, "mfem_broadphase_ee", rc)
        return self._result()


def soft_rigid_distance(points, tris, rigid_points_count, rigid_tri_count, pairs, pair_count, d_hat, distance):
    """In-place per-soft-particle minimum distance (device tensors)."""
    lib = _ffi.load()
    dev = distance.device
    rc = lib.mfem_soft_rigid_distance(_p(points), _p(tris), int(rigid_points_count), int(rigid_tri_count), _p(pairs),
                                      _p(pair_count), int(pairs.shape[0]), ctypes.c_float(d_hat), _p(distance),
                                      _stream(dev))
    _check(lib, "mfem_soft_rigid_distance", rc)
    return distance


def pair_barrier(kind: str, x, pairs, n_pairs, kappa: float, d_hat: float, tris=None, edges=None):
    """Per candidate pair: distance, kappa * b(d) (FP64) and its gradient
    (n, 4, 3) with the four vertex ids (n, 4)."""
    lib = _ffi.load()
    dev = x.device
    m = int(pairs.shape[0])
    dist = torch.zeros(m, dtype=torch.float32, device=dev)
    energy = torch.zeros(m, dtype=torch.float64, device=dev)
    grad = torch.zeros(m, 4, 3, dtype=torch.float32, device=dev)
    vids = torch.zeros(m, 4, dtype=torch.int32, device=dev)
    k = {"pt": 0, "ee": 1}[kind]
    rc = lib.mfem_pair_barrier(k, _p(x), _p(pairs), _p(n_pairs), m, _p(tris), _p(edges), ctypes.c_float(kappa),
                               ctypes.c_float(d_hat), _p(dist), _p(energy), _p(grad), _p(vids), _stream(dev))
    _check(lib, "mfem_pair_barrier", rc)
    return dist, energy, grad, vids
