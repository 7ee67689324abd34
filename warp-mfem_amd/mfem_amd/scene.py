"""Model / State containers and the synthetic benchmark scenes.

``Model`` carries the fields ``RefinementSolver`` reads from a newton Model
(src/mfem/refinement/solver.py:41-304; src/mfem/refinement/contact.py:778-862):
particle_q/qd/mass/inv_mass, tet_indices/poses/materials, tri_indices, the
rigid shapes and gravity.  Arrays are numpy on the host; the solver copies them
to the device once.  ``State`` holds the per-step device tensors
(particle_q, particle_qd, body_q).

Scenes:

* ``beam_scene``      cantilever beam 1.0 x 0.2 x 0.2 m, 20x4x4 cubes x 6 tets
                      (1,920 tets, 525 vertices), ARAP, E=1e5 Pa, nu=0.45,
                      rho=1000, x=0 face pinned, g=(0,0,-9.81), dt=1/60.
* ``octopus_scene``   ellipsoid mantle (0.3x0.2x0.2 m) plus 8 tapered arms
                      0.6 m long, lattice-meshed at h=7.3 mm (~2e5 tets),
                      Neo-Hookean with a seeded log-normal stiffness field
                      (mean E=5e4 Pa, sigma_log=0.5, correlation length
                      0.05 m), dropped onto the plane z=0 with an IPC barrier
                      (dhat=1e-3) and friction mu=0.3, dt=1/100.
* ``drop_box_scene``  small Neo-Hookean box over a plane, for parity tests.
* ``soft_body_batch`` 64 seeded boxes / ellipsoids, 2e3 to 5e4 tets each
                      (log-uniform, ~1.2e6 in all), mixed ARAP / Neo-Hookean,
                      E log-uniform in [1e4, 1e6] Pa, started deformed over
                      the plane z=0 with IPC, dt=1/100 (a batch).

``Model.params`` holds RefinementSolver keyword arguments (reference names,
src/mfem/refinement/solver.py:56-128, 241-245) plus ``iterations``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as _mesh

SHAPE_PLANE = 0
SHAPE_CAPSULE = 1


@dataclass
class Shape:
    """Rigid collider: a plane (local +Z normal through the origin) or a
    capsule (local Z axis, scale = (radius, half_height))
    (src/mfem/ipc/distance.py:4-57).  ``pos`` and the quaternion ``quat`` =
    (x, y, z, w) are the shape's transform relative to its rigid body
    ``body`` (newton's shape_transform / shape_body), or its world pose when
    ``body`` is -1 (a static shape); the world pose is
    body_q[body] * (pos, quat) (src/mfem/refinement/contact.py:74-82)."""

    type: int
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))
    quat: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 0.0, 1.0]))
    scale: np.ndarray = field(default_factory=lambda: np.zeros(3))
    mu: float = 0.0
    body: int = -1


def quat_mul(q1, q2) -> np.ndarray:
    """Hamilton product q1 * q2 of (x, y, z, w) quaternions (q2 applied first)."""
    x1, y1, z1, w1 = q1
    x2, y2, z2, w2 = q2
    return np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])


def quat_rotate(q, v) -> np.ndarray:
    """Rotation of v by the unit quaternion q = (x, y, z, w)."""
    u, w = np.asarray(q[:3], float), float(q[3])
    v = np.asarray(v, float)
    return v + 2.0 * np.cross(u, np.cross(u, v) + w * v)


def shape_world_poses(model, body_q=None) -> np.ndarray:
    """(n_shapes, 7) world poses p, q of the model's shapes for the rigid
    body poses ``body_q`` (n_bodies, 7) (default: the model's body_q)."""
    bq = getattr(model, "body_q", None) if body_q is None else body_q
    out = []
    for sh in model.shapes:
        p, q = np.asarray(sh.pos, float), np.asarray(sh.quat, float)
        if sh.body >= 0:
            t = np.asarray(bq, float).reshape(-1, 7)[sh.body]
            p, q = t[:3] + quat_rotate(t[3:], p), quat_mul(t[3:], q)
        out.append(np.concatenate([p, q]))
    return np.array(out).reshape(-1, 7)


@dataclass
class Model:
    particle_q: np.ndarray
    particle_qd: np.ndarray
    particle_mass: np.ndarray
    particle_inv_mass: np.ndarray
    tet_indices: np.ndarray
    tet_poses: np.ndarray
    tet_materials: np.ndarray
    tri_indices: np.ndarray | None
    shapes: list
    gravity: np.ndarray
    params: dict = field(default_factory=dict)
    name: str = ""
    body_q: np.ndarray | None = None  # (n_bodies, 7) initial rigid body poses p, q (xyzw)

    @property
    def particle_count(self) -> int:
        return int(self.particle_q.shape[0])

    @property
    def tet_count(self) -> int:
        return int(self.tet_indices.shape[0])


def build_model(particles, tets, materials, density, *, pinned=None, shapes=(), gravity=(0.0, 0.0, -9.81),
                velocity=None, params=None, name="") -> Model:
    """Rest pose, lumped masses and surface triangles for a tet mesh."""
    particles = np.asarray(particles, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int32)
    dminv, vol = _mesh.rest_pose(particles, tets)
    mass = _mesh.lumped_mass(particles.shape[0], tets, vol, density)
    inv_mass = np.where(mass > 0, 1.0 / np.maximum(mass, 1e-300), 0.0)
    if pinned is not None:
        inv_mass[np.asarray(pinned)] = 0.0
    qd = np.zeros_like(particles) if velocity is None else np.broadcast_to(velocity, particles.shape).copy()
    return Model(
        particle_q=particles,
        particle_qd=qd,
        particle_mass=mass,
        particle_inv_mass=inv_mass,
        tet_indices=tets,
        tet_poses=dminv,
        tet_materials=np.asarray(materials, dtype=np.float64),
        tri_indices=_mesh.surface_tris(tets),
        shapes=list(shapes),
        gravity=np.asarray(gravity, dtype=np.float64),
        # a refinement pass re-lumps the masses with refine_density
        # (refinement.py:1253): the density the model was built with
        params={"refine_density": float(density), **dict(params or {})},
        name=name,
    )


def pad_particles(model: Model, max_particles: int) -> Model:
    """The model with its particle buffer padded to ``max_particles`` rows
    (zero position, velocity and mass; inactive), as a caller pads it for
    refinement headroom before constructing the solver
    (src/mfem/refinement/models/refinement_model.py:104-143)."""
    n = model.particle_count
    if max_particles < n:
        raise ValueError(f"max_particles={max_particles} < model.particle_count={n}")
    import dataclasses

    def pad(a):
        a = np.asarray(a)
        out = np.zeros((max_particles,) + a.shape[1:], dtype=a.dtype)
        out[:n] = a
        return out

    out = dataclasses.replace(model, particle_q=pad(model.particle_q), particle_qd=pad(model.particle_qd),
                              particle_mass=pad(model.particle_mass), particle_inv_mass=pad(model.particle_inv_mass),
                              params=dict(model.params))
    # attributes a scene added beside the fields (drive settings and the like)
    names = {f.name for f in dataclasses.fields(model)}
    out.__dict__.update({k: v for k, v in model.__dict__.items() if k not in names})
    return out


def _materials(mu, lam, n):
    m = np.zeros((n, 3))
    m[:, 0] = mu
    m[:, 1] = lam
    return m


def beam_scene(cells=(20, 4, 4), energy="arap") -> Model:
    """Cantilever beam, x=0 face pinned (25 vertices at the default size)."""
    p, t = _mesh.box_tet_mesh((0.0, 0.0, 0.0), (1.0, 0.2, 0.2), cells)
    mu, lam = _mesh.lame(1.0e5, 0.45)
    pinned = np.nonzero(np.abs(p[:, 0]) < 1e-9)[0]
    return build_model(
        p, t, _materials(mu, lam, len(t)), 1000.0, pinned=pinned,
        params=dict(energy=energy, iterations=8, cg_max_iterations=1000, cg_tolerance=1e-6,
                    attachment_stiffness=10.0, line_search=True, line_search_max_iterations=20),
        name="beam_%dx%dx%d" % tuple(cells),
    )


def gaussian_field(points: np.ndarray, length: float, seed: int, n_features: int = 256) -> np.ndarray:
    """Unit-variance Gaussian random field with squared-exponential covariance
    of correlation length ``length``, by random Fourier features."""
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 1.0 / length, size=(n_features, 3))
    b = rng.uniform(0.0, 2.0 * np.pi, size=n_features)
    return np.sqrt(2.0 / n_features) * np.cos(points @ w.T + b).sum(axis=1)


def barrier_slope(d, d1):
    d = np.asarray(d, dtype=np.float64)
    return -(2.0 * (d - d1) * np.log(d / d1) + (d - d1) ** 2 / d)


def auto_kappa(model: Model, dhat: float, factor: float = 10.0) -> float:
    """Barrier stiffness that holds ``factor`` x the mean vertex weight at
    half the activation distance: kappa |b'(dhat/2)| = factor m g."""
    w = float(np.mean(model.particle_mass)) * float(np.linalg.norm(model.gravity))
    return factor * w / abs(float(barrier_slope(0.5 * dhat, dhat)))


def octopus_inside(points: np.ndarray, arms: int = 8) -> np.ndarray:
    """Ellipsoid mantle (semi-axes 0.15, 0.10, 0.10 m, centre 0.10 m above the
    arm plane) united with ``arms`` tapered arms (radius 0.03 -> 0.01 m, length
    0.6 m) radiating horizontally, their undersides flush with z=0."""
    p = points
    c = np.array([0.0, 0.0, 0.13])
    inside = ((p[:, 0] - c[0]) / 0.15) ** 2 + ((p[:, 1] - c[1]) / 0.10) ** 2 + ((p[:, 2] - c[2]) / 0.10) ** 2 <= 1.0
    r0, r1, L = 0.03, 0.01, 0.6
    for k in range(arms):
        th = 2.0 * np.pi * (k + 0.5) / arms
        d = np.array([np.cos(th), np.sin(th), 0.0])
        a = np.array([0.08 * np.cos(th), 0.06 * np.sin(th), 0.0])
        rel = p - a
        t = np.clip(rel @ d / L, 0.0, 1.0)
        r = r0 + (r1 - r0) * t
        axis = a[None, :] + (t * L)[:, None] * d[None, :]
        axis[:, 2] = r  # underside on z = 0
        inside |= np.linalg.norm(p - axis, axis=1) <= r
    return inside


def octopus_scene(h: float = 7.3e-3, seed: int = 7, clearance: float = 2.0e-3, dhat: float = 1.0e-3,
                  friction_mu: float = 0.3, iterations: int = 12) -> Model:
    lo = np.array([-0.72, -0.72, -0.001])
    hi = np.array([0.72, 0.72, 0.24])
    p, t = _mesh.implicit_tet_mesh(octopus_inside, lo, hi, h)
    p, t, _ = _mesh.spatial_order(p, t)  # compact 256-vertex blocks (the PCG's local aggregates)
    p[:, 2] += clearance - p[:, 2].min()
    cen = p[t].mean(axis=1)
    g = gaussian_field(cen, 0.05, seed)
    E = 5.0e4 * np.exp(0.5 * g - 0.125)
    mu, lam = _mesh.lame(E, 0.45)
    ground = Shape(SHAPE_PLANE, mu=friction_mu)
    m = build_model(p, t, np.stack([mu, lam, np.zeros_like(mu)], axis=1), 1000.0, shapes=[ground],
                    name="octopus_h%.1fmm" % (h * 1e3))
    m.params.update(energy="neohookean", iterations=iterations, cg_max_iterations=5000, cg_tolerance=1e-6,
                    line_search=True, line_search_max_iterations=30, contact_d0=0.1 * dhat, contact_d1=dhat,
                    contact_stiffness=auto_kappa(m, dhat), friction_mu=[friction_mu], friction_eps_v=1e-3,
                    attachment_stiffness=10.0)
    return m


class TipDrive:
    """Arm tips driven as moving penalty attachments: each arm's tip vertices
    (pinned: inv_mass = 0) follow base + a_k (sin(2 pi f_k t + phi_k) -
    sin(phi_k)) (at rest at t = 0), with
    amplitude vector a_k (``amplitude`` along a seeded random horizontal
    direction),
    frequency f_k ~ U[0.5, 2] Hz and phase phi_k ~ U[0, 2 pi) per arm.  The
    targets move the attachments' pin positions, as
    DirichletBoundaryCondition.update_b does for new pin positions
    (src/mfem/boundary_condition.py:85-92)."""

    def __init__(self, indices, arm, base, amplitude, freq, phase):
        self.indices = np.asarray(indices, np.int64)
        self.arm = np.asarray(arm, np.int64)
        self.base = np.asarray(base, np.float64)
        self.amplitude = np.asarray(amplitude, np.float64)  # (arms, 3)
        self.freq = np.asarray(freq, np.float64)
        self.phase = np.asarray(phase, np.float64)

    def targets(self, t: float) -> np.ndarray:
        """Pin positions at time t (the rest positions at t = 0)."""
        w = np.sin(2.0 * np.pi * self.freq * t + self.phase) - np.sin(self.phase)  # (arms,)
        return self.base + self.amplitude[self.arm] * w[self.arm, None]


def octopus_drive_scene(h: float = 7.3e-3, seed: int = 11, tip_radius: float = 0.025, amplitude: float = 0.1,
                        refine_every: int = 10, self_contact: bool = True, arms: int = 8):
    """The octopus of ``octopus_scene`` with self-contact between its surface
    primitives (point-triangle and edge-edge barrier, dhat = 1e-3, the
    ground's stiffness), its arm tips (vertices within ``tip_radius`` of each
    arm's end) driven along seeded sinusoids (``TipDrive``: amplitude 0.1 m,
    0.5 to 2 Hz, random phase and horizontal direction), and adaptive refinement every
    ``refine_every`` steps on a strain-energy threshold: a tet's edges are
    split when its strain energy exceeds tau * (mean rest tet volume), tau
    drawn log-uniformly in [20, 200] J/m^3 from the seed.  Returns (model,
    drive)."""
    m = octopus_scene(h=h)
    rng = np.random.default_rng(seed)
    r0, r1, L = 0.03, 0.01, 0.6
    x = m.particle_q
    idx, arm = [], []
    for k in range(arms):
        th = 2.0 * np.pi * (k + 0.5) / arms
        d = np.array([np.cos(th), np.sin(th), 0.0])
        a = np.array([0.08 * np.cos(th), 0.06 * np.sin(th), 0.0])
        tip = a + L * d
        tip[2] = x[:, 2].min() + r1
        near = np.nonzero(np.linalg.norm(x - tip, axis=1) < tip_radius)[0]
        idx.append(near)
        arm.append(np.full(len(near), k))
    idx = np.concatenate(idx)
    arm = np.concatenate(arm)
    # horizontal directions: the tips slide over the ground (a downward
    # component would drive the attachments into the plane)
    ang = rng.uniform(0.0, 2.0 * np.pi, arms)
    dirs = np.stack([np.cos(ang), np.sin(ang), np.zeros(arms)], axis=1)
    drive = TipDrive(idx, arm, x[idx].copy(), amplitude * dirs, rng.uniform(0.5, 2.0, arms),
                     rng.uniform(0.0, 2.0 * np.pi, arms))
    m.particle_inv_mass = m.particle_inv_mass.copy()
    m.particle_inv_mass[idx] = 0.0
    vol = 1.0 / (6.0 * np.linalg.det(m.tet_poses.reshape(-1, 3, 3)))
    tau = float(np.exp(rng.uniform(np.log(20.0), np.log(200.0))))
    m.params.update(self_contact=self_contact, self_contact_dhat=m.params["contact_d1"],
                    self_contact_kappa=m.params["contact_stiffness"], refine_tet_score_weight=1.0,
                    refine_vertex_score_weight=0.1)
    m.refine_every = refine_every
    m.min_refine_score = tau * float(vol.mean())
    m.strain_energy_threshold = tau
    m.name = "octopus_drive_h%.1fmm" % (h * 1e3)
    return m, drive


def drop_box_scene(cells=(4, 3, 3), size=(0.2, 0.15, 0.15), clearance=5.0e-4, E=2.0e4, energy="neohookean",
                   friction_mu=0.3, dhat=1.0e-3, velocity=(0.3, 0.0, -0.2)) -> Model:
    """Small box falling and sliding onto the ground plane (contact + friction)."""
    p, t = _mesh.box_tet_mesh((0.0, 0.0, clearance), (size[0], size[1], size[2] + clearance), cells)
    mu, lam = _mesh.lame(E, 0.4)
    m = build_model(p, t, _materials(mu, lam, len(t)), 1000.0, shapes=[Shape(SHAPE_PLANE, mu=friction_mu)],
                    velocity=np.asarray(velocity), name="drop_box")
    m.params.update(energy=energy, iterations=8, cg_max_iterations=2000, cg_tolerance=1e-6,
                    line_search=True, line_search_max_iterations=20, contact_d0=0.1 * dhat, contact_d1=dhat,
                    contact_stiffness=auto_kappa(m, dhat), friction_mu=[friction_mu], friction_eps_v=1e-3)
    return m


def stacked_boxes_scene(cells=(4, 4, 2), size=(0.12, 0.12, 0.04), gap=4.0e-4, twist_deg=30.0, E=5.0e4,
                        energy="neohookean", dhat=1.0e-3, friction_mu=0.3, velocity=(0.0, 0.0, -0.15)) -> Model:
    """Self-contact: two slabs in one mesh, the lower one resting on the
    plane z = 0, the upper one twisted about z by ``twist_deg`` (so that
    their edges cross) and falling onto it from ``gap`` above (inside dhat)
    at ``velocity``.  Self-contact on (point-triangle and edge-edge barrier
    between the slabs, same dhat and stiffness as the ground's)."""
    p0, t0 = _mesh.box_tet_mesh((-0.5 * size[0], -0.5 * size[1], 5.0e-4),
                                (0.5 * size[0], 0.5 * size[1], 5.0e-4 + size[2]), cells)
    th = np.deg2rad(twist_deg)
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    p1 = p0 @ R.T + np.array([0.01, 0.005, size[2] + gap])
    p = np.concatenate([p0, p1])
    t = np.concatenate([t0, t0 + len(p0)])
    mu, lam = _mesh.lame(E, 0.4)
    vel = np.zeros_like(p)
    vel[len(p0):] = velocity
    m = build_model(p, t, _materials(mu, lam, len(t)), 1000.0, shapes=[Shape(SHAPE_PLANE, mu=friction_mu)],
                    name="stacked_boxes")
    m.particle_qd = vel
    kappa = auto_kappa(m, dhat)
    m.params.update(energy=energy, iterations=8, cg_max_iterations=2000, cg_tolerance=1e-6,
                    line_search=True, line_search_max_iterations=30, contact_d0=0.1 * dhat, contact_d1=dhat,
                    contact_stiffness=kappa, friction_mu=[friction_mu], friction_eps_v=1e-3,
                    self_contact=True, self_contact_dhat=dhat, self_contact_kappa=kappa)
    return m


def soft_body_batch(n_bodies: int = 64, seed: int = 360, tets_min: float = 2.0e3, tets_max: float = 5.0e4,
                    dhat: float = 1.0e-3, friction_mu: float = 0.3, iterations: int = 12,
                    spacing: float = 1.0) -> list:
    """``n_bodies`` independent soft bodies, each a box or an ellipsoid
    (equal odds) with seeded random proportions (edges / diameters 0.15 to
    0.4 m), lattice-meshed to a tet count drawn log-uniformly from
    [tets_min, tets_max]; ARAP or Neo-Hookean (equal odds), E log-uniform in
    [1e4, 1e6] Pa, nu = 0.45, rho = 1000.  Each starts deformed (a random
    stretch / shear of up to 15 % from its rest shape), 2 mm above the plane
    z = 0 (IPC barrier dhat, friction mu) with a small random velocity, on an
    8-wide grid ``spacing`` apart.  The log-uniform draw of 64 counts in
    [2e3, 5e4] averages ~0.95e6 tets; the default seed is one whose draw
    totals ~1.2e6 (1,209,342 tets, 248,279 vertices, bodies of 2,052 to
    46,872 tets).  Solver parameters are shared (a batch
    shares them): those of ``octopus_scene`` with the barrier stiffness of
    the heaviest-vertex body, so every body's contact holds."""
    rng = np.random.default_rng(seed)
    models = []
    for b in range(n_bodies):
        target = float(np.exp(rng.uniform(np.log(tets_min), np.log(tets_max))))
        dims = rng.uniform(0.15, 0.4, size=3)
        ellipsoid = bool(rng.integers(2))
        vol = np.prod(dims) * (np.pi / 6.0 if ellipsoid else 1.0)
        h = float((6.0 * vol / target) ** (1.0 / 3.0))  # 6 tets per lattice cube
        if ellipsoid:
            inside = lambda p, r=0.5 * dims: ((p / r) ** 2).sum(axis=1) <= 1.0
            p, t = _mesh.implicit_tet_mesh(inside, -0.5 * dims, 0.5 * dims, h)
        else:
            cells = np.maximum(np.round(dims / h).astype(int), 1)
            p, t = _mesh.box_tet_mesh(-0.5 * dims, 0.5 * dims, cells)
        p, t, _ = _mesh.spatial_order(p, t)
        energy = "neohookean" if rng.integers(2) else "arap"
        E = float(np.exp(rng.uniform(np.log(1.0e4), np.log(1.0e6))))
        mu, lam = _mesh.lame(E, 0.45)
        F = np.eye(3) + rng.uniform(-0.15, 0.15, size=(3, 3))
        x = p @ F.T
        offset = np.array([spacing * (b % 8), spacing * (b // 8), 0.0])
        x += offset - np.array([0.0, 0.0, x[:, 2].min() - 2.0e-3])
        vel = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.5, 0.0)])
        m = build_model(p + offset, t, _materials(mu, lam, len(t)), 1000.0,
                        shapes=[Shape(SHAPE_PLANE, mu=friction_mu)], velocity=vel,
                        name=f"{'ellipsoid' if ellipsoid else 'box'}_{len(t)}_{energy}_E{E:.3g}")
        m.particle_q = x
        m.params.update(energy=energy)
        models.append(m)
    kappa = max(auto_kappa(m, dhat) for m in models)
    for m in models:
        m.params.update(iterations=iterations, cg_max_iterations=5000, cg_tolerance=1e-6, line_search=True,
                        line_search_max_iterations=30, contact_d0=0.1 * dhat, contact_d1=dhat,
                        contact_stiffness=kappa, friction_mu=[friction_mu], friction_eps_v=1e-3,
                        attachment_stiffness=10.0)
    return models


def capsule_poke_scene(cells=(6, 6, 2), energy="arap") -> Model:
    """Slab pinned at its base, poked from above by a tilted capsule (triangle
    and vertex contact against a capsule, src/mfem/refinement/contact.py)."""
    p, t = _mesh.box_tet_mesh((0.0, 0.0, 0.0), (0.3, 0.3, 0.1), cells)
    mu, lam = _mesh.lame(5.0e4, 0.4)
    pinned = np.nonzero(p[:, 2] < 1e-9)[0]
    # capsule axis along world x (local Z rotated 90 degrees about y), its
    # surface 2 mm above the slab's top face
    # the capsule is rigid body 0 (its shape at the body's origin), posed by
    # State.body_q each step
    cap = Shape(SHAPE_CAPSULE, scale=np.array([0.01, 0.08, 0.0]), mu=0.2, body=0)
    m = build_model(p, t, _materials(mu, lam, len(t)), 1000.0, pinned=pinned, shapes=[cap], gravity=(0, 0, -9.81),
                    name="capsule_poke")
    m.body_q = np.array([[0.15, 0.15, 0.1 + 0.012, 0.0, np.sin(np.pi / 4), 0.0, np.cos(np.pi / 4)]])
    m.params.update(energy=energy, iterations=6, cg_max_iterations=2000, cg_tolerance=1e-6,
                    line_search=True, line_search_max_iterations=20, contact_d0=1e-3, contact_d1=5e-3,
                    contact_stiffness=50.0, friction_mu=[0.2], friction_eps_v=1e-3, attachment_stiffness=10.0,
                    tri_contact=True)
    return m
