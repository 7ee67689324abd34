"""Surface-tracking and correspondence losses of a simulated surface against
a tracked (or a reference simulation's) surface.

Mirrors src/mfem/refinement/surface_loss.py of the reference: the same
classes, arguments, result dictionaries and edge-case behaviour
(``TrackedSurfaceLoss`` :146-367, ``TrackedCorrespondenceLoss`` :370-553,
``RecordingOverlay`` :569-624, ``score_recording`` :684-726,
``surface_triangles_from_tets`` :119-138).  The point-to-surface distance
(k nearest vertices, then projection onto their incident triangles) runs
on the GPU in FP64 (``mfem_surface_knn`` / ``mfem_surface_project`` in
libmfem_amd.so); masks, bookkeeping and the means stay on the host.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _ffi
from .driver import capsule_signed_distance, load_recording

__all__ = ["surface_triangles_from_tets", "vertex_faces_padded", "point_mesh_query", "TrackedSurfaceLoss",
           "TrackedCorrespondenceLoss", "RecordingOverlay", "score_recording", "merge_correspondence_row",
           "capsule_signed_distance", "CORR_PREFIX"]

CORR_PREFIX = "corr_"


def surface_triangles_from_tets(tets) -> np.ndarray:
    """Faces that belong to exactly one tet, wound outward (faces (0,2,1),
    (0,1,3), (0,3,2), (1,2,3) of each tet), in the order of their sorted
    vertex triples (surface_loss.py:119-138)."""
    tets = np.asarray(tets, dtype=np.int64).reshape(-1, 4)
    faces = np.concatenate([tets[:, [0, 2, 1]], tets[:, [0, 1, 3]], tets[:, [0, 3, 2]], tets[:, [1, 2, 3]]])
    key = np.sort(faces, axis=1)
    order = np.lexsort(key.T[::-1])
    key, faces = key[order], faces[order]
    keep = np.ones(len(key), dtype=bool)
    if len(key) > 1:
        same = np.all(key[1:] == key[:-1], axis=1)
        keep[:-1] &= ~same
        keep[1:] &= ~same
    return faces[keep].astype(np.int32)


def vertex_faces_padded(faces, n_vertices: int):
    """(vert_faces (n_vertices, max degree) padded with -1, degree): every
    vertex's incident triangles in triangle order (surface_loss.py:99-116)."""
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    flat = faces.reshape(-1)
    fidx = np.repeat(np.arange(faces.shape[0]), 3)
    order = np.argsort(flat, kind="stable")
    v_sorted, f_sorted = flat[order], fidx[order]
    degree = np.bincount(v_sorted, minlength=n_vertices)
    max_deg = int(degree.max()) if degree.size else 0
    vf = np.full((n_vertices, max(max_deg, 1)), -1, dtype=np.int64)
    starts = np.zeros(n_vertices, dtype=np.int64)
    starts[1:] = np.cumsum(degree)[:-1]
    vf[v_sorted, np.arange(v_sorted.shape[0]) - starts[v_sorted]] = f_sorted
    return vf, degree


def _dev(device):
    d = torch.device(device if device is not None else "cuda")
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def _surf_call(lib, fn, *args):
    rc = getattr(lib, fn)(*args)
    if rc != 0:
        raise _ffi.MfemError(f"{fn} failed ({rc}): {lib.mfem_surface_last_error().decode()}")


def point_mesh_query(query, verts, faces=None, vert_faces=None, *, valid=None, face_ok=None, k: int = 1,
                     project: bool = True, device=None):
    """Squared distance of every query point (Q, 3) to the surface
    (verts, faces): the k nearest (valid) vertices, then the smallest
    distance over their incident (usable) triangles, else the nearest
    vertex's (surface_loss.py:206-245).  Returns numpy arrays
    (sq (Q,), face (Q,) int (-1: vertex fallback), closest point (Q, 3),
    nearest vertex (Q,) int).  Runs on the GPU (FP64)."""
    lib = _ffi.load()
    dev = _dev(device)
    query = np.asarray(query, dtype=np.float64).reshape(-1, 3)
    verts = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    nq = len(query)
    vmap = None if valid is None else np.nonzero(np.asarray(valid, dtype=bool))[0].astype(np.int32)
    m = len(verts) if vmap is None else len(vmap)
    if nq == 0:
        return np.zeros(0), np.zeros(0, np.int64), np.zeros((0, 3)), np.zeros(0, np.int64)
    if m == 0:
        raise ValueError("no candidate vertices to measure against")
    k = max(1, min(int(k), m, 8))
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    q_t, v_t = T(query, torch.float64), T(verts, torch.float64)
    vm_t = T(vmap, torch.int32) if vmap is not None else None
    idx = torch.empty(nq, k, dtype=torch.int32, device=dev)
    d2 = torch.empty(nq, k, dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    _surf_call(lib, "mfem_surface_knn", ptr(q_t), nq, ptr(v_t), ptr(vm_t), m, k, ptr(idx), ptr(d2), stream)
    use_faces = project and faces is not None and len(faces) > 0
    if use_faces:
        faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
        if vert_faces is None:
            vert_faces, _ = vertex_faces_padded(faces, len(verts))
        f_t, vf_t = T(faces, torch.int32), T(vert_faces, torch.int32)
        deg = int(vert_faces.shape[1])
        ok_t = T(np.asarray(face_ok, dtype=np.uint8), torch.uint8) if face_ok is not None else None
    else:
        f_t = vf_t = ok_t = None
        deg = 0
    sq = torch.empty(nq, dtype=torch.float64, device=dev)
    fo = torch.empty(nq, dtype=torch.int32, device=dev)
    cp = torch.empty(nq, 3, dtype=torch.float64, device=dev)
    _surf_call(lib, "mfem_surface_project", ptr(q_t), nq, ptr(v_t), ptr(f_t), ptr(vf_t), deg, ptr(ok_t), ptr(idx),
               ptr(d2), k, ptr(sq), ptr(fo), ptr(cp), stream)
    return (sq.cpu().numpy(), fo.cpu().numpy().astype(np.int64), cp.cpu().numpy(),
            idx[:, 0].cpu().numpy().astype(np.int64))


class TrackedSurfaceLoss:
    """Mean squared distance of the sim surface vertices to the tracked
    surface at the matching frame (sim -> tracked), and with ``symmetric``
    also tracked -> sim and their average (surface_loss.py:146-367).

    ``overlay`` provides ``active``, ``faces``, ``traj`` (frames, N, 3),
    ``valid`` (frames, N) or None, ``positions(frame)`` and
    ``frame_for_time(t)``.  ``exclude_tool_margin`` (m): with a tool signed
    distance given to ``evaluate``, tracked vertices closer than this count
    as untracked and such sim vertices are left out."""

    def __init__(self, overlay, *, project_to_faces=True, use_valid_mask=True, symmetric=False, k_neighbors=1,
                 exclude_tool_margin=None, device=None):
        self.overlay = overlay
        self.exclude_tool_margin = float(exclude_tool_margin) if exclude_tool_margin is not None else None
        self.active = bool(getattr(overlay, "active", False))
        self.project_to_faces = bool(project_to_faces)
        self.symmetric = bool(symmetric)
        self.k_neighbors = max(int(k_neighbors), 1)
        self.device = device
        if not self.active:
            self.faces = None
            return
        self.faces = np.asarray(overlay.faces, dtype=np.int64)
        n_vertices = int(overlay.traj.shape[1])
        self.use_valid_mask = bool(use_valid_mask) and getattr(overlay, "valid", None) is not None
        self.vert_faces, _ = vertex_faces_padded(self.faces, n_vertices)

    def _tracked_frame(self, frame):
        pos = np.asarray(self.overlay.positions(frame), dtype=np.float64)
        valid = np.asarray(self.overlay.valid[frame], dtype=bool) if self.use_valid_mask else None
        return pos, valid

    def _sq_dist(self, query, verts, valid, faces, vert_faces):
        face_ok = None
        if valid is not None and faces is not None:
            face_ok = valid[faces].all(axis=1)
        sq, _, _, _ = point_mesh_query(query, verts, faces, vert_faces, valid=valid, face_ok=face_ok,
                                       k=self.k_neighbors, project=self.project_to_faces, device=self.device)
        return sq

    def evaluate(self, sim_points, sim_time=None, *, frame=None, tool_signed_distance=None, sim_faces=None):
        """Loss dict for one sim surface point cloud (surface_loss.py:248-367):
        ``frame``, ``n_sim``, ``n_sim_all``, ``n_tracked_excluded``, ``mse``,
        ``rmse``, ``rmse_mm``, ``max_mm`` (+ ``mse_tracked_to_sim``,
        ``mse_symmetric``, ``rmse_symmetric_mm`` with ``symmetric``); None when
        the overlay is inactive; NaN entries for non-finite sim points."""
        if not self.active:
            return None
        if frame is None:
            if sim_time is None:
                raise ValueError("pass sim_time or frame")
            frame = self.overlay.frame_for_time(sim_time)
        frame = int(frame)
        verts, valid = self._tracked_frame(frame)
        sim_points = np.asarray(sim_points, dtype=np.float64).reshape(-1, 3)
        if not np.isfinite(sim_points).all():
            out = {"frame": frame, "n_sim": int(len(sim_points)), "n_sim_all": int(len(sim_points)),
                   "n_tracked_excluded": 0, "mse": math.nan, "rmse": math.nan, "rmse_mm": math.nan,
                   "max_mm": math.nan}
            if self.symmetric:
                out.update(mse_tracked_to_sim=math.nan, mse_symmetric=math.nan, rmse_symmetric_mm=math.nan)
            return out
        sim_all = sim_points
        n_sim_all = int(len(sim_points))
        n_tracked_excluded = 0
        if self.exclude_tool_margin is not None and tool_signed_distance is not None:
            margin = self.exclude_tool_margin
            near_trk = np.asarray(tool_signed_distance(verts), dtype=np.float64) < margin
            n_tracked_excluded = int(near_trk.sum())
            valid = ~near_trk if valid is None else valid & ~near_trk
            sim_points = sim_points[np.asarray(tool_signed_distance(sim_points), dtype=np.float64) >= margin]
        if len(sim_points):
            sq = self._sq_dist(sim_points, verts, valid, self.faces, self.vert_faces)
            mse, max_sq = float(sq.mean()), float(sq.max())
        else:
            mse, max_sq = 0.0, 0.0
        out = {"frame": frame, "n_sim": int(len(sim_points)), "n_sim_all": n_sim_all,
               "n_tracked_excluded": n_tracked_excluded, "mse": mse, "rmse": math.sqrt(mse),
               "rmse_mm": math.sqrt(mse) * 1e3, "max_mm": math.sqrt(max_sq) * 1e3}
        if self.symmetric:
            keep = valid if valid is not None else np.ones(len(verts), dtype=bool)
            tracked = verts[keep]
            if len(sim_all) and len(tracked):
                if sim_faces is not None and len(sim_faces):
                    sf = np.asarray(sim_faces, dtype=np.int64)
                    svf, _ = vertex_faces_padded(sf, len(sim_all))
                    rev = self._sq_dist(tracked, sim_all, None, sf, svf)
                else:
                    rev, _, _, _ = point_mesh_query(tracked, sim_all, None, None, k=self.k_neighbors, project=False,
                                                    device=self.device)
                mse_rev = float(rev.mean())
            else:
                mse_rev = 0.0
            out["mse_tracked_to_sim"] = mse_rev
            out["mse_symmetric"] = 0.5 * (mse + mse_rev)
            out["rmse_symmetric_mm"] = math.sqrt(out["mse_symmetric"]) * 1e3
        return out


def _barycentric_on_triangle(p, a, b, c):
    """Barycentric weights of p in the plane of abc, clipped to the simplex
    and renormalised (surface_loss.py:373-393)."""
    ab, ac, ap = b - a, c - a, p - a
    d00 = np.einsum("...i,...i->...", ab, ab)
    d01 = np.einsum("...i,...i->...", ab, ac)
    d11 = np.einsum("...i,...i->...", ac, ac)
    d20 = np.einsum("...i,...i->...", ap, ab)
    d21 = np.einsum("...i,...i->...", ap, ac)
    den = d00 * d11 - d01 * d01
    safe = np.where(den == 0.0, 1.0, den)
    v = np.where(den == 0.0, 0.0, (d11 * d20 - d01 * d21) / safe)
    w = np.where(den == 0.0, 0.0, (d00 * d21 - d01 * d20) / safe)
    bary = np.clip(np.stack([1.0 - v - w, v, w], axis=-1), 0.0, 1.0)
    return bary / bary.sum(axis=-1, keepdims=True)


class TrackedCorrespondenceLoss:
    """Material-point loss: each sim surface vertex is bound once to the
    closest point (triangle + barycentric weights) of the tracked surface and
    compared, frame by frame, with where that point went; on displacement
    from the bind frame with ``relative`` (surface_loss.py:396-553)."""

    def __init__(self, overlay, *, use_valid_mask=True, exclude_tool_margin=None, k_neighbors=4, relative=True,
                 device=None):
        self.overlay = overlay
        self.active = bool(getattr(overlay, "active", False))
        self.exclude_tool_margin = float(exclude_tool_margin) if exclude_tool_margin is not None else None
        self.k_neighbors = max(int(k_neighbors), 1)
        self.relative = bool(relative)
        self.device = device
        self.tri = self.bary = self.sim0 = self.target0 = self.bind_frame = None
        if not self.active:
            self.faces = None
            return
        self.faces = np.asarray(overlay.faces, dtype=np.int64)
        n_vertices = int(overlay.traj.shape[1])
        self.use_valid_mask = bool(use_valid_mask) and getattr(overlay, "valid", None) is not None
        self.vert_faces, _ = vertex_faces_padded(self.faces, n_vertices)

    @property
    def bound(self):
        return self.tri is not None

    def _tracked_frame(self, frame):
        pos = np.asarray(self.overlay.positions(frame), dtype=np.float64)
        valid = np.asarray(self.overlay.valid[frame], dtype=bool) if self.use_valid_mask else None
        return pos, valid

    def _resolve_frame(self, sim_time, frame):
        if frame is None:
            if sim_time is None:
                raise ValueError("pass sim_time or frame")
            frame = self.overlay.frame_for_time(sim_time)
        return int(frame)

    def bind(self, sim_points, sim_time=None, *, frame=None):
        """Fixes the correspondence at the matching frame; returns ``frame``,
        ``n_bound``, ``n_unbound``, ``bind_rmse_mm``, ``bind_max_mm``."""
        if not self.active:
            return None
        frame = self._resolve_frame(sim_time, frame)
        verts, valid = self._tracked_frame(frame)
        sim_points = np.asarray(sim_points, dtype=np.float64).reshape(-1, 3)
        if not np.isfinite(sim_points).all():
            raise ValueError("cannot bind a correspondence to non-finite sim points")
        face_ok = valid[self.faces].all(axis=1) if valid is not None else None
        sq, face, cp, _ = point_mesh_query(sim_points, verts, self.faces, self.vert_faces, valid=valid,
                                           face_ok=face_ok, k=self.k_neighbors, device=self.device)
        has = face >= 0
        tri_v = verts[self.faces[np.where(has, face, 0)]]
        bary = _barycentric_on_triangle(cp, tri_v[:, 0], tri_v[:, 1], tri_v[:, 2])
        bary[~has] = 0.0
        self.tri = np.where(has, face, -1)
        self.bary = bary
        self.sim0 = sim_points.copy()
        self.target0 = np.where(has[:, None], cp, np.nan)
        self.bind_frame = frame
        resid = np.sqrt(sq[has]) if has.any() else np.zeros(0)
        return {"frame": frame, "n_bound": int(has.sum()), "n_unbound": int((~has).sum()),
                "bind_rmse_mm": float(np.sqrt((resid ** 2).mean()) * 1e3) if resid.size else 0.0,
                "bind_max_mm": float(resid.max() * 1e3) if resid.size else 0.0}

    def targets(self, frame):
        """(targets (N, 3) with NaN for unbound points, usable mask (N,))."""
        verts, valid = self._tracked_frame(int(frame))
        tri_safe = np.where(self.tri >= 0, self.tri, 0)
        tgt = np.einsum("nk,nki->ni", self.bary, verts[self.faces[tri_safe]])
        ok = self.tri >= 0
        if valid is not None:
            ok &= valid[self.faces[tri_safe]].all(axis=1)
        return np.where(ok[:, None], tgt, np.nan), ok

    def evaluate(self, sim_points, sim_time=None, *, frame=None, tool_signed_distance=None):
        """Per-frame dict: ``frame``, ``n_sim``, ``n_sim_all``, ``n_excluded``,
        ``mse``, ``rmse``, ``rmse_mm``, ``max_mm`` (surface_loss.py:504-553)."""
        if not self.active:
            return None
        if not self.bound:
            raise RuntimeError("TrackedCorrespondenceLoss.bind() must be called first")
        frame = self._resolve_frame(sim_time, frame)
        n = len(self.tri)
        sim_points = np.asarray(sim_points, dtype=np.float64).reshape(-1, 3)
        if len(sim_points) < n:
            raise ValueError(f"expected at least {n} sim points (bound set), got {len(sim_points)}")
        sim_points = sim_points[:n]
        if not np.isfinite(sim_points).all():
            return {"frame": frame, "n_sim": 0, "n_sim_all": int(n), "n_excluded": int(n), "mse": math.nan,
                    "rmse": math.nan, "rmse_mm": math.nan, "max_mm": math.nan}
        tgt, ok = self.targets(frame)
        if self.exclude_tool_margin is not None and tool_signed_distance is not None:
            m = self.exclude_tool_margin
            d_sim = np.asarray(tool_signed_distance(sim_points), dtype=np.float64)
            d_tgt = np.asarray(tool_signed_distance(np.where(ok[:, None], tgt, 0.0)), dtype=np.float64)
            ok &= (d_sim >= m) & (d_tgt >= m)
        err = (sim_points - self.sim0) - (tgt - self.target0) if self.relative else sim_points - tgt
        sq = (err ** 2).sum(axis=1)[ok]
        mse, max_sq = (float(sq.mean()), float(sq.max())) if sq.size else (0.0, 0.0)
        return {"frame": frame, "n_sim": int(ok.sum()), "n_sim_all": int(n), "n_excluded": int(n - ok.sum()),
                "mse": mse, "rmse": math.sqrt(mse), "rmse_mm": math.sqrt(mse) * 1e3,
                "max_mm": math.sqrt(max_sq) * 1e3}


def merge_correspondence_row(row, corr):
    """Adds a correspondence result to a surface-loss row under ``corr_*`` keys."""
    if corr is None:
        return row
    for k in ("mse", "rmse", "rmse_mm", "max_mm", "n_sim", "n_excluded"):
        row[CORR_PREFIX + k] = corr[k]
    return row


class RecordingOverlay:
    """A recording standing in for the tracked surface (a coarse run scored
    against a converged fine run): constant topology required; faces are its
    boundary triangles (surface_loss.py:569-624)."""

    def __init__(self, path):
        self.path = path
        rec = load_recording(path)
        tets0 = rec["tets_at"](0)
        n0 = len(rec["positions_at"](0))
        for i in range(rec["frame_count"]):
            t = rec["tets_at"](i)
            if t.shape != tets0.shape or not np.array_equal(t, tets0):
                raise ValueError(f"{path}: reference recording must have constant topology")
        faces = surface_triangles_from_tets(tets0)
        self.faces = faces[(faces < n0).all(axis=1)]
        self.traj = np.stack([rec["positions_at"](i)[:n0] for i in range(rec["frame_count"])])
        self.times = np.asarray(rec["times"], dtype=np.float64)
        self.T = int(self.traj.shape[0])
        self.valid = None
        self.active = True
        self.capsule = rec["capsule"]

    def positions(self, frame):
        return self.traj[int(frame)]

    def frame_for_time(self, t):
        return int(np.argmin(np.abs(self.times - float(t))))


def score_recording(rec_path, loss, *, playback_rate=1.0, corr=None):
    """Per-frame loss curve of a recording (dict of stacked arrays), the
    recorded capsule feeding the tool exclusion; with ``corr`` bound to the
    first frame's surface vertices (surface_loss.py:684-726)."""
    rec = load_recording(rec_path)
    tris_cache, rows, surf0 = {}, [], None
    for i in range(rec["frame_count"]):
        pts = rec["positions_at"](i)
        tets = rec["tets_at"](i)
        tris = tris_cache.get(tets.shape)
        if tris is None:
            tris = tris_cache[tets.shape] = surface_triangles_from_tets(tets)
        surf = np.unique(tris)
        surf = surf[surf < len(pts)]
        tool = None
        cap = rec["capsule"]
        if cap is not None and i < len(cap["transform"]):
            tf = cap["transform"][i]
            tool = lambda P, tf=tf: capsule_signed_distance(P, tf, cap["radius"], cap["half_height"])
        local_faces = np.searchsorted(surf, tris[(tris < len(pts)).all(axis=1)])
        t = rec["times"][i] * playback_rate
        res = loss.evaluate(pts[surf], t, tool_signed_distance=tool, sim_faces=local_faces)
        if corr is not None and corr.active:
            if surf0 is None:
                surf0 = surf
                corr.bind(pts[surf0], t)
            merge_correspondence_row(res, corr.evaluate(pts[surf0], t, tool_signed_distance=tool))
        res["time"] = float(rec["times"][i])
        rows.append(res)
    return {k: np.array([r[k] for r in rows]) for k in rows[0].keys()}
