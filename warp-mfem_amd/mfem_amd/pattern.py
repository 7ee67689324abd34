"""Host-side construction of the system-matrix pattern (numpy, once per topology).

The reference assembles H with warp.sparse (bsr_set_from_triplets + bsr_mm +
bsr_axpy into a padded BSR, src/mfem/refinement/solver.py:641-799).  Here the
pattern is fixed per topology and laid out for the GPU:

* SELL-64: rows (vertices) are grouped in chunks of 64 = one wavefront; a
  chunk stores max-row-length "slots", and slot s of chunk c holds, for each of
  the 64 lanes, one 3x3 block as 9 lane-minor floats
  (value[(slot*9 + q)*64 + lane]).  Every load in the SpMV is a contiguous
  256-byte wave access.  Padding blocks are zero with column 0.
* Gather lists for a deterministic assembly: every lane-slot (block (i, j))
  owns the ordered list of tet-block sources that sum into it; each vertex owns
  the ordered list of (tet, corner) forces that sum into its right-hand side.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import vertex_adjacency_pattern

# the 10 unique vertex pairs of a tet (4 diagonal, then the 6 upper pairs);
# must match PA / PB in csrc/mfem_kernels.hip
PAIRS = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


@dataclass
class SystemPattern:
    n_verts: int
    n_chunks: int
    n_slots: int
    rowptr: np.ndarray      # CSR of the block pattern (for reference / tests)
    cols: np.ndarray
    chunk_ptr: np.ndarray   # int32 [n_chunks + 1]
    sell_col: np.ndarray    # int32 [n_slots * 64]
    diag_lane: np.ndarray   # int32 [n_verts]
    csr_lane: np.ndarray    # int64 [nnz] lane-slot of each CSR block
    contrib_ptr: np.ndarray  # int32 [n_slots * 64 + 1]
    contrib_src: np.ndarray  # int32
    vinc_ptr: np.ndarray    # int32 [n_verts + 1]
    vinc_src: np.ndarray    # int32

    @property
    def padding_ratio(self) -> float:
        return self.n_slots * 64 / max(1, len(self.cols))


def build_system_pattern(n_verts: int, tets: np.ndarray) -> SystemPattern:
    tets = np.asarray(tets, dtype=np.int64)
    T = tets.shape[0]
    if T and (tets.min() < 0 or tets.max() >= n_verts):
        raise ValueError("tet index out of range")
    rowptr, cols = vertex_adjacency_pattern(n_verts, tets)
    nnz = len(cols)
    n_chunks = (n_verts + 63) // 64
    lens = np.diff(rowptr.astype(np.int64))
    lens_p = np.zeros(n_chunks * 64, dtype=np.int64)
    lens_p[:n_verts] = lens
    width = lens_p.reshape(n_chunks, 64).max(axis=1)
    chunk_ptr = np.zeros(n_chunks + 1, dtype=np.int64)
    chunk_ptr[1:] = np.cumsum(width)
    n_slots = int(chunk_ptr[-1])

    rows = np.repeat(np.arange(n_verts, dtype=np.int64), lens)
    k_in_row = np.arange(nnz, dtype=np.int64) - rowptr[rows].astype(np.int64)
    csr_lane = (chunk_ptr[rows // 64] + k_in_row) * 64 + rows % 64
    sell_col = np.zeros(n_slots * 64, dtype=np.int32)
    sell_col[csr_lane] = cols
    diag_mask = cols == rows
    diag_lane = np.zeros(n_verts, dtype=np.int64)
    diag_lane[rows[diag_mask]] = csr_lane[diag_mask]

    # tet block -> lane-slot contributions
    keys = rows * n_verts + cols.astype(np.int64)  # sorted (row-major, cols ascending)
    srcs, dests = [], []
    for p, (a, b) in enumerate(PAIRS):
        blk = (p * T + np.arange(T, dtype=np.int64))
        ia, ib = tets[:, a], tets[:, b]
        k = np.searchsorted(keys, ia * n_verts + ib)
        dests.append(csr_lane[k])
        srcs.append(blk * 2)
        if a != b:  # block (b, a) is the transpose of (a, b)
            k = np.searchsorted(keys, ib * n_verts + ia)
            dests.append(csr_lane[k])
            srcs.append(blk * 2 + 1)
    dest = np.concatenate(dests)
    src = np.concatenate(srcs)
    order = np.lexsort((src, dest))
    dest, src = dest[order], src[order]
    contrib_ptr = np.zeros(n_slots * 64 + 1, dtype=np.int64)
    np.add.at(contrib_ptr, dest + 1, 1)
    contrib_ptr = np.cumsum(contrib_ptr)

    # vertex -> (tet, corner) incidence
    vs = tets.reshape(-1)
    vsrc = np.arange(4 * T, dtype=np.int64)
    order = np.lexsort((vsrc, vs))
    vinc_ptr = np.zeros(n_verts + 1, dtype=np.int64)
    np.add.at(vinc_ptr, vs + 1, 1)
    vinc_ptr = np.cumsum(vinc_ptr)

    for name, arr in (("contrib_src", src), ("contrib_ptr", contrib_ptr), ("lane", csr_lane)):
        if arr.size and arr.max() >= 2**31:
            raise ValueError(f"{name} exceeds int32")
    return SystemPattern(
        n_verts=n_verts,
        n_chunks=n_chunks,
        n_slots=n_slots,
        rowptr=rowptr,
        cols=cols,
        chunk_ptr=chunk_ptr.astype(np.int32),
        sell_col=sell_col,
        diag_lane=diag_lane.astype(np.int32),
        csr_lane=csr_lane,
        contrib_ptr=contrib_ptr.astype(np.int32),
        contrib_src=src.astype(np.int32),
        vinc_ptr=vinc_ptr.astype(np.int32),
        vinc_src=vsrc[order].astype(np.int32),
    )


def build_triangle_lists(pat: SystemPattern, tris: np.ndarray):
    """Gather lists of the triangle-capsule contact terms (once per topology).

    Returns (tcon_ptr, tcon_src, vtri_ptr, vtri_src): for every SELL
    lane-slot, the ordered triangle vertex pairs (tri * 9 + 3 a + b) whose
    Hessian block w_a w_b k b'' n n^T lands on it; for every vertex, its
    triangle corners (tri * 3 + a) in triangle order.  Surface triangles are
    faces of tets, so every pair is a block of the tet pattern; a triangle
    that is not raises ValueError.
    """
    tris = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    n, F = pat.n_verts, tris.shape[0]
    if F and (tris.min() < 0 or tris.max() >= n):
        raise ValueError("triangle index out of range")
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(pat.rowptr.astype(np.int64)))
    keys = rows * n + pat.cols.astype(np.int64)
    dests, srcs = [], []
    for a in range(3):
        for b in range(3):
            k = np.searchsorted(keys, tris[:, a] * n + tris[:, b])
            k = np.minimum(k, len(keys) - 1)
            if F and not np.all(keys[k] == tris[:, a] * n + tris[:, b]):
                raise ValueError("a surface triangle edge is not an edge of the tet mesh")
            dests.append(pat.csr_lane[k])
            srcs.append(np.arange(F, dtype=np.int64) * 9 + 3 * a + b)
    dest = np.concatenate(dests) if F else np.zeros(0, np.int64)
    src = np.concatenate(srcs) if F else np.zeros(0, np.int64)
    order = np.lexsort((src, dest))
    dest, src = dest[order], src[order]
    tcon_ptr = np.zeros(pat.n_slots * 64 + 1, dtype=np.int64)
    np.add.at(tcon_ptr, dest + 1, 1)
    tcon_ptr = np.cumsum(tcon_ptr)
    vs = tris.reshape(-1)
    vsrc = np.arange(3 * F, dtype=np.int64)
    order = np.lexsort((vsrc, vs))
    vtri_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(vtri_ptr, vs + 1, 1)
    vtri_ptr = np.cumsum(vtri_ptr)
    return (tcon_ptr.astype(np.int32), src.astype(np.int32), vtri_ptr.astype(np.int32),
            vsrc[order].astype(np.int32))


def sell_to_csr_values(pat: SystemPattern, sell_val: np.ndarray) -> np.ndarray:
    """(nnz, 3, 3) CSR block values from the lane-minor SELL value array."""
    sv = np.asarray(sell_val).reshape(pat.n_slots, 9, 64)
    lane = pat.csr_lane
    return sv[lane // 64, :, lane % 64].reshape(-1, 3, 3)


class GpuPatternBuilder:
    """build_system_pattern + build_triangle_lists on the GPU
    (csrc/mfem_pattern.hip), for meshes up to the given capacities: the same
    arrays, in the same order.  ``build`` returns the device tensors and a
    host SystemPattern view (rowptr, cols, chunk_ptr, csr_lane, sell_col,
    diag_lane copied back; the gather lists stay on the device)."""

    def __init__(self, max_verts: int, max_tets: int, max_tris: int, device):
        import torch

        from . import _ffi

        self.lib = _ffi.load()
        self.torch = torch
        self.device = torch.device(device)
        self.caps = (int(max_verts), int(max_tets), int(max(max_tris, 1)))
        nb = int(self.lib.mfem_pattern_workspace_bytes(*self.caps))
        self.ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        self.counts = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)

    def _call(self, name, *args):
        from . import _ffi

        if getattr(self.lib, name)(*args) != 0:
            raise _ffi.MfemError(f"{name}: {self.lib.mfem_pattern_last_error().decode()}")

    def build(self, tets, n_tets: int, tris, n_tris: int, n_verts: int):
        import ctypes

        torch = self.torch
        vp = ctypes.c_void_p
        dev = self.device
        i32 = lambda n: torch.zeros(max(int(n), 1), dtype=torch.int32, device=dev)
        st = vp(torch.cuda.current_stream(dev).cuda_stream)
        n_chunks = (n_verts + 63) // 64
        out = {"rowptr": i32(n_verts + 1), "chunk_ptr": i32(n_chunks + 1)}
        mv, mt, mf = self.caps
        ws, nb = vp(self.ws.data_ptr()), ctypes.c_size_t(self.ws.numel())
        self._call("mfem_pattern_count", vp(tets.data_ptr()), n_tets, n_verts, mv, mt, mf, ws, nb,
                   vp(out["rowptr"].data_ptr()), vp(out["chunk_ptr"].data_ptr()), vp(self.counts.data_ptr()), st)
        nnz, n_slots = (int(v) for v in self.counts.tolist())
        n_lanes = n_slots * 64
        out.update(cols=i32(nnz), sell_col=i32(n_lanes), diag_lane=i32(n_verts), contrib_ptr=i32(n_lanes + 1),
                   contrib_src=i32(16 * n_tets), vinc_ptr=i32(n_verts + 1), vinc_src=i32(4 * n_tets),
                   tcon_ptr=i32(n_lanes + 1), tcon_src=i32(9 * n_tris), vtri_ptr=i32(n_verts + 1),
                   vtri_src=i32(3 * n_tris))
        names = ("cols", "sell_col", "diag_lane", "contrib_ptr", "contrib_src", "vinc_ptr", "vinc_src", "tcon_ptr",
                 "tcon_src", "vtri_ptr", "vtri_src")
        self._call("mfem_pattern_fill", vp(tets.data_ptr()), n_tets, vp(tris.data_ptr() if n_tris else 0), n_tris,
                   n_verts, mv, mt, mf, ws, nb, vp(out["rowptr"].data_ptr()), vp(out["chunk_ptr"].data_ptr()),
                   n_slots, *[vp(out[k].data_ptr()) for k in names], vp(self.status.data_ptr()), st)
        if int(self.status.item()) != 0:
            raise ValueError("a surface triangle edge is not an edge of the tet mesh")
        rowptr = out["rowptr"].cpu().numpy()
        cols = out["cols"][:nnz].cpu().numpy()
        chunk_ptr = out["chunk_ptr"].cpu().numpy()
        rows = np.repeat(np.arange(n_verts, dtype=np.int64), np.diff(rowptr.astype(np.int64)))
        k_in_row = np.arange(nnz, dtype=np.int64) - rowptr[rows].astype(np.int64)
        csr_lane = (chunk_ptr[rows // 64].astype(np.int64) + k_in_row) * 64 + rows % 64
        empty = np.zeros(0, np.int32)
        pat = SystemPattern(n_verts=n_verts, n_chunks=n_chunks, n_slots=n_slots, rowptr=rowptr, cols=cols,
                            chunk_ptr=chunk_ptr, sell_col=out["sell_col"].cpu().numpy(),
                            diag_lane=out["diag_lane"].cpu().numpy(), csr_lane=csr_lane, contrib_ptr=empty,
                            contrib_src=empty, vinc_ptr=empty, vinc_src=empty)
        return pat, out
