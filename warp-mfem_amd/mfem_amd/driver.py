"""Scene driver around the solve: a recorded tool trajectory that poses the
kinematic capsule at every substep, the frame / substep loop, and the mesh
recorder whose ``.npz`` the reference's replay and scoring scripts read.

Reference (src/mfem/refinement/sim_octopus.py):
  * tool path: loading ``tool_transform`` / ``timestamps`` (:994-1013), the
    capsule-centre offset along the tool axis (:1019-1030), the start
    standoff retract (:1032-1058), sampling with linear translation + slerp
    (:1110-1136), the per-substep device pose sequence (:1138-1176,
    ``_set_tool_pose_kernel`` :309-324);
  * frame loop: ``simulate`` / ``step`` (:1523-1560);
  * recorder: ``_MeshRecorder`` (:2510-2591); recording layout
    (replay_octopus.py:14-37, surface_loss.py:627-655).

The solve itself runs in libmfem_amd.so; this module only moves poses into
``State.body_q`` (a device copy, capturable in a HIP graph) and copies
frames to the host for the recorder.
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.spatial.transform import Rotation, Slerp


def capsule_signed_distance(points, transform7, radius, half_height):
    """Signed distance of ``points`` (N, 3) to a capsule with world pose
    ``transform7`` = (px, py, pz, qx, qy, qz, qw), axis along local +Z;
    negative inside (surface_loss.py:669-681)."""
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    tf = np.asarray(transform7, dtype=np.float64)
    axis = Rotation.from_quat(tf[3:7]).apply([0.0, 0.0, 1.0])
    a = tf[:3] - axis * half_height
    ab = 2.0 * half_height * axis
    t = np.clip((P - a) @ ab / max(float(ab @ ab), 1e-20), 0.0, 1.0)
    return np.linalg.norm(P - (a + t[:, None] * ab), axis=1) - radius


class ToolTrajectory:
    """A recorded rigid tool path sampled at any time.

    ``transforms`` (N, 4, 4) world-from-tool matrices at ``times`` (N,)
    seconds.  The capsule's long axis is the tool's local +Z; its centre sits
    ``center_offset`` along that axis from the tool origin.  ``time_offset``
    is the trajectory time that simulation time 0 maps to (the start frame
    divided by the capture rate); ``loop`` wraps time over the recorded span,
    otherwise it is clamped to it (sim_octopus.py:1110-1136)."""

    def __init__(self, transforms, times=None, *, capsule_radius: float, capsule_half_height: float,
                 tip_along_z: float | None = None, contact_position=None, time_offset: float = 0.0,
                 loop: bool = False, fps: float = 30.0, contact_active=None):
        T = np.asarray(transforms, dtype=np.float64).reshape(-1, 4, 4)
        if len(T) == 0:
            raise ValueError("an empty tool trajectory")
        t = None if times is None else np.asarray(times, dtype=np.float64)
        if t is None or t.ndim != 1 or t.shape[0] != T.shape[0]:
            t = np.arange(T.shape[0], dtype=np.float64) / float(fps)
        # strictly increasing stamps (Slerp needs them): running maximum plus
        # a 1e-9 s ramp (sim_octopus.py:1003-1006)
        t = np.maximum.accumulate(t) + 1e-9 * np.arange(t.shape[0])
        self.times = t
        self.pos = T[:, :3, 3].copy()
        self.rot = Rotation.from_matrix(T[:, :3, :3])
        self._slerp = Slerp(t, self.rot) if len(t) > 1 else None
        self.contact_active = (np.ones(len(t), dtype=bool) if contact_active is None
                               else np.asarray(contact_active, dtype=bool))
        # the capsule's +Z cap on the tool tip at the first sample
        # (sim_octopus.py:1019-1030)
        local_z0 = T[0, :3, 2]
        if contact_position is not None:
            cp0 = np.asarray(contact_position, dtype=np.float64).reshape(-1, 3)[0]
            tip_along_z = float(np.dot(cp0 - T[0, :3, 3], local_z0))
        elif tip_along_z is None:
            tip_along_z = capsule_half_height + capsule_radius
        self.capsule_radius = float(capsule_radius)
        self.capsule_half_height = float(capsule_half_height)
        self.center_offset = float(tip_along_z) - (self.capsule_half_height + self.capsule_radius)
        self.time_offset = float(time_offset)
        self.loop = bool(loop)
        self.start_retract = 0.0
        self.frame_index = 0

    @classmethod
    def from_npz(cls, path, **kw) -> "ToolTrajectory":
        """Reads ``tool_transform`` (N, 4, 4), ``timestamps`` (N,) and the
        optional ``contact_active`` / ``contact_position`` keys
        (sim_octopus.py:994-1030); numeric arrays only (no pickle)."""
        with np.load(path, allow_pickle=False) as z:
            extra = {}
            if "contact_active" in z.files:
                extra["contact_active"] = z["contact_active"]
            if "contact_position" in z.files:
                extra["contact_position"] = z["contact_position"]
            return cls(z["tool_transform"], z["timestamps"] if "timestamps" in z.files else None, **extra, **kw)

    def sample(self, t: float):
        """(position (3,), quaternion xyzw (4,)) of the capsule at simulation-
        relative trajectory time ``t`` seconds: translation linearly
        interpolated, rotation slerped, centre shifted along the local +Z."""
        t = float(t) + self.time_offset
        t0, t1 = float(self.times[0]), float(self.times[-1])
        if self.loop and t1 > t0:
            t = t0 + ((t - t0) % (t1 - t0))
        tc = float(np.clip(t, t0, t1))
        self.frame_index = min(int(np.searchsorted(self.times, tc)), len(self.times) - 1)
        rot = self._slerp(tc) if self._slerp is not None else self.rot[0]
        pos = np.array([np.interp(tc, self.times, self.pos[:, k]) for k in range(3)])
        pos = pos + self.center_offset * rot.apply([0.0, 0.0, 1.0])
        return pos, rot.as_quat()

    def pose7(self, t: float) -> np.ndarray:
        """The capsule pose at time ``t`` as (px, py, pz, qx, qy, qz, qw) float32."""
        pos, q = self.sample(t)
        return np.concatenate([pos, q]).astype(np.float32)

    def retract_to_standoff(self, vertices, standoff: float, max_iters: int = 40) -> float:
        """Moves the whole path back along -tool-axis until the capsule at
        time 0 clears ``vertices`` by ``standoff`` (sim_octopus.py:1032-1058);
        returns the total retract (m)."""
        V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        if len(V) == 0 or standoff < 0.0:
            return 0.0
        for _ in range(max_iters):
            d0 = float(capsule_signed_distance(V, self.pose7(0.0).astype(np.float64), self.capsule_radius,
                                               self.capsule_half_height).min())
            if d0 >= standoff:
                break
            step = max(standoff - d0, 1.0e-3)
            self.center_offset -= step
            self.start_retract += step
        return self.start_retract


class ToolDriver:
    """Poses rigid body ``body`` (the capsule's) at every substep of a frame
    from a ``ToolTrajectory``: ``refresh(sim_time)`` samples the substeps'
    poses on the host into a device tensor once per frame; ``apply(state, k)``
    copies pose k into ``state.body_q`` on the device (a graph-capturable
    copy; sim_octopus.py:309-324, 1138-1176, 1523-1553)."""

    def __init__(self, trajectory: ToolTrajectory, substeps: int, sim_dt: float, playback_rate: float = 1.0,
                 device="cuda", body: int = 0):
        self.traj = trajectory
        self.substeps = max(int(substeps), 1)
        self.sim_dt = float(sim_dt)
        self.rate = float(playback_rate)
        self.body = int(body)
        self.pose_seq = torch.as_tensor(np.tile(trajectory.pose7(0.0), (self.substeps, 1)), device=device)

    def refresh(self, sim_time: float) -> None:
        seq = np.stack([self.traj.pose7((sim_time + k * self.sim_dt) * self.rate) for k in range(self.substeps)])
        self.pose_seq.copy_(torch.as_tensor(seq))

    def apply(self, state, k: int) -> None:
        if state.body_q is None:
            raise ValueError("the state carries no body_q to pose")
        state.body_q[self.body].copy_(self.pose_seq[int(k)])


def simulate_frame(solver, state_0, state_1, sim_dt: float, substeps: int, driver: ToolDriver | None = None):
    """One frame: ``substeps`` solver steps, the tool posed before each and
    the states swapped after each (sim_octopus.py:1523-1553).  Returns the
    (state_0, state_1) pair after the swaps; state_0 holds the result."""
    for k in range(max(int(substeps), 1)):
        if driver is not None:
            driver.apply(state_0, k)
        solver.step(state_0, state_1, None, None, sim_dt)
        state_0, state_1 = state_1, state_0
    return state_0, state_1


class MeshRecorder:
    """Accumulates the active mesh (positions, tets) and the capsule pose at
    every captured frame and writes them into one ``.npz`` in the layout the
    reference's ``replay_octopus`` and ``surface_loss`` read
    (sim_octopus.py:2510-2591): ``fps``, ``episode``, ``start_frame``,
    ``frame_count``, ``times``, ``positions`` ((F, N, 3) float32),
    ``tet_indices`` ((M, 4) when the topology never changed, else per
    frame), ``capsule_transform`` ((F, 7)), ``capsule_radius``,
    ``capsule_half_height``.  When refinement changed the sizes, the
    reference stores per-frame object arrays; this recorder writes those too
    (for the reference's readers) and, beside them, flat arrays with frame
    offsets (``positions_flat`` / ``positions_offsets``,
    ``tet_indices_flat`` / ``tet_indices_offsets``), which ``load_recording``
    reads without unpickling anything."""

    def __init__(self, path, fps: float, *, episode: str = "octopus", start_frame: int = 0,
                 capsule_radius: float = 0.0, capsule_half_height: float = 0.0):
        self.path = str(path)
        self.fps = float(fps)
        self.episode = str(episode)
        self.start_frame = int(start_frame)
        self.capsule_radius = float(capsule_radius)
        self.capsule_half_height = float(capsule_half_height)
        self._times: list[float] = []
        self._positions: list[np.ndarray] = []
        self._tets: list[np.ndarray] = []
        self._capsule: list[np.ndarray] = []

    def capture(self, solver, state, sim_time: float, body: int = 0) -> None:
        """Snapshot of the active mesh and the capsule pose (sim_octopus.py:2535-2548)."""
        n, m = solver.active_particle_count, solver.active_tet_count
        self._positions.append(state.particle_q[:n].detach().cpu().numpy().astype(np.float32))
        self._tets.append(solver.tet_indices[:m].detach().cpu().numpy().astype(np.int32))
        cap = np.zeros(7, np.float32)
        cap[6] = 1.0
        if state.body_q is not None and state.body_q.shape[0] > body:
            cap = state.body_q[body].detach().cpu().numpy().astype(np.float32)
        self._capsule.append(cap)
        self._times.append(float(sim_time))

    @staticmethod
    def _pack(frames):
        if len({a.shape for a in frames}) == 1:
            return np.stack(frames), False
        packed = np.empty(len(frames), dtype=object)
        packed[:] = frames
        return packed, True

    def close(self) -> dict:
        """Writes the file; returns a summary (frames, ragged / constant flags)."""
        if not self._positions:
            return {"frames": 0, "written": False}
        positions, ragged_p = self._pack(self._positions)
        tets, ragged_t = self._pack(self._tets)
        constant = not ragged_t and all(np.array_equal(t, self._tets[0]) for t in self._tets[1:])
        if constant:
            tets = self._tets[0]
        out = dict(fps=np.float64(self.fps), episode=np.str_(self.episode), start_frame=np.int64(self.start_frame),
                   frame_count=np.int64(len(self._positions)), times=np.asarray(self._times, dtype=np.float64),
                   positions=positions, tet_indices=tets, capsule_transform=np.stack(self._capsule),
                   capsule_radius=np.float64(self.capsule_radius),
                   capsule_half_height=np.float64(self.capsule_half_height))
        if ragged_p:
            out["positions_flat"] = np.concatenate(self._positions)
            out["positions_offsets"] = np.cumsum([0] + [len(p) for p in self._positions]).astype(np.int64)
        if ragged_t:
            out["tet_indices_flat"] = np.concatenate(self._tets)
            out["tet_indices_offsets"] = np.cumsum([0] + [len(t) for t in self._tets]).astype(np.int64)
        np.savez_compressed(self.path, **out)
        return {"frames": len(self._positions), "written": True, "positions_ragged": ragged_p,
                "tets": "constant" if constant else ("ragged" if ragged_t else "per-frame")}


def load_recording(path) -> dict:
    """Reads a recording without unpickling: dense arrays directly, ragged
    frames from the flat / offsets keys.  A file whose ragged frames exist
    only as object arrays is refused (loading them would run pickle).
    Returns the fields of ``surface_loss._load_recording``
    (surface_loss.py:627-655): ``fps``, ``episode``, ``start_frame``,
    ``frame_count``, ``times``, ``capsule`` (or None), ``positions_at(i)``,
    ``tets_at(i)``."""
    z = np.load(path, allow_pickle=False)
    files = set(z.files)

    def frames(key):
        if key + "_offsets" in files:
            flat, off = z[key + "_flat"], z[key + "_offsets"]
            return lambda i: flat[off[i]:off[i + 1]]
        try:
            arr = z[key]
        except ValueError as e:  # object array: would need pickle
            raise ValueError(f"{path}: '{key}' holds per-frame object arrays and no flat copy; "
                             "refusing to unpickle it") from e
        return arr

    pos = frames("positions")
    tets = frames("tet_indices")
    n = int(z["frame_count"]) if "frame_count" in files else (len(pos) if not callable(pos) else 0)
    times = (np.asarray(z["times"], dtype=np.float64) if "times" in files
             else np.arange(n, dtype=np.float64) / float(z["fps"]))
    capsule = None
    if "capsule_transform" in files:
        capsule = {"transform": np.asarray(z["capsule_transform"], dtype=np.float64),
                   "radius": float(z["capsule_radius"]), "half_height": float(z["capsule_half_height"])}
    if callable(pos):
        positions_at = lambda i: np.asarray(pos(i), dtype=np.float64)
    else:
        positions_at = lambda i: np.asarray(pos[i], dtype=np.float64)
    if callable(tets):
        tets_at = lambda i: np.asarray(tets(i), dtype=np.int64)
    elif tets.ndim == 2:
        tets_at = lambda i: np.asarray(tets, dtype=np.int64)
    else:
        tets_at = lambda i: np.asarray(tets[i], dtype=np.int64)
    return {"fps": float(z["fps"]), "episode": str(z["episode"]) if "episode" in files else "octopus",
            "start_frame": int(z["start_frame"]) if "start_frame" in files else 0, "frame_count": n,
            "times": times, "capsule": capsule, "positions_at": positions_at, "tets_at": tets_at}
