"""Tet-mesh ingestion: rest pose, lumped mass, surface triangles, meshers.

These are the duties newton's ``ModelBuilder.add_soft_mesh`` performs for the
reference (consumed by ``MFEMRefinementModel.load_sim_model``,
src/mfem/refinement/models/refinement_model.py:104-143):

* ``tet_poses`` = DmInv with Dm = [x1-x0, x2-x0, x3-x0] as columns, so the
  solver's signed rest volume is 1/(6 det DmInv)
  (src/mfem/refinement/elastic_energy.py:143-145);
* ``particle_mass`` = sum over incident tets of density*volume/4 and
  ``particle_inv_mass`` = 1/mass (0 marks a pinned particle, which the solver
  turns into a penalty attachment, src/mfem/refinement/kernels.py:7-30);
* ``tri_indices`` = boundary faces (faces used by exactly one tet), oriented
  outward (the surface-extraction idea of
  src/mfem/refinement/mesh_topology.py:686-741).

The meshers build conforming Kuhn (Freudenthal) tetrahedralisations of a
lattice: every cube is split into the same six tets around its main diagonal,
so shared faces match between neighbouring cubes.
"""
from __future__ import annotations

import itertools

import numpy as np

# Kuhn split of the unit cube: one tet per permutation of the axes, walking
# 000 -> e_a -> e_a + e_b -> 111.
_CORNER = {c: i for i, c in enumerate(itertools.product((0, 1), repeat=3))}  # (ix, iy, iz) -> 0..7


def _kuhn_local() -> np.ndarray:
    tets = []
    for perm in itertools.permutations(range(3)):
        p = [0, 0, 0]
        path = [tuple(p)]
        for ax in perm:
            p[ax] = 1
            path.append(tuple(p))
        v = [_CORNER[c] for c in path]
        a = np.array(path, dtype=float)
        if np.linalg.det(np.stack([a[1] - a[0], a[2] - a[0], a[3] - a[0]], axis=1)) < 0:
            v[1], v[2] = v[2], v[1]
        tets.append(v)
    return np.array(tets, dtype=np.int64)


KUHN = _kuhn_local()


def lattice_tet_mesh(cell_mask: np.ndarray, origin, h: float):
    """Kuhn tets of every lattice cell where ``cell_mask`` is True.

    ``cell_mask`` has shape (nx, ny, nz); returns (particles (V,3) float64,
    tets (T,4) int32) with unused lattice vertices removed, vertices in
    lattice (x-major) order and tets in cell order.
    """
    nx, ny, nz = cell_mask.shape
    cells = np.argwhere(cell_mask)  # x-major order
    corners = np.array(list(itertools.product((0, 1), repeat=3)))  # matches _CORNER order
    vid = lambda ijk: (ijk[..., 0] * (ny + 1) + ijk[..., 1]) * (nz + 1) + ijk[..., 2]
    cell_corner_ids = vid(cells[:, None, :] + corners[None, :, :])  # (C, 8)
    tets = cell_corner_ids[:, KUHN].reshape(-1, 4)
    used, inverse = np.unique(tets.reshape(-1), return_inverse=True)
    tets = inverse.reshape(-1, 4).astype(np.int32)
    iz = used % (nz + 1)
    iy = (used // (nz + 1)) % (ny + 1)
    ix = used // ((nz + 1) * (ny + 1))
    particles = np.asarray(origin, dtype=np.float64)[None, :] + h * np.stack([ix, iy, iz], axis=1)
    return particles, tets


def box_tet_mesh(lo, hi, cells):
    """Box [lo, hi] split into cells[0] x cells[1] x cells[2] cubes of 6 tets."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    cells = tuple(int(c) for c in cells)
    h = (hi - lo) / np.array(cells)
    if not np.allclose(h, h[0]):
        # anisotropic cells: mesh the unit lattice and scale the axes
        p, t = lattice_tet_mesh(np.ones(cells, dtype=bool), np.zeros(3), 1.0)
        return lo + p * h, t
    return lattice_tet_mesh(np.ones(cells, dtype=bool), lo, float(h[0]))


def implicit_tet_mesh(inside, lo, hi, h: float):
    """Lattice-clipped Kuhn mesh of the solid {p : inside(p)} inside [lo, hi]:
    a cube is kept when its centre is inside."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    n = np.maximum(np.ceil((hi - lo) / h).astype(int), 1)
    gx, gy, gz = (lo[k] + h * (np.arange(n[k]) + 0.5) for k in range(3))
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    centres = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    mask = inside(centres).reshape(tuple(n))
    return lattice_tet_mesh(mask, lo, h)


def morton_order(points: np.ndarray, bits: int = 10) -> np.ndarray:
    """Permutation sorting points along a Z-order (Morton) curve of a
    2^bits lattice over their bounding cube (ties keep their order)."""
    p = np.asarray(points, dtype=np.float64)
    lo = p.min(axis=0)
    ext = float((p.max(axis=0) - lo).max()) or 1.0
    q = np.minimum(((p - lo) / ext * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    code = np.zeros(len(p), dtype=np.int64)
    for b in range(bits):
        for k in range(3):
            code |= ((q[:, k] >> b) & 1) << (3 * b + k)
    return np.argsort(code, kind="stable")


def spatial_order(particles: np.ndarray, tets: np.ndarray):
    """Renumbers a tet mesh for locality: vertices along a Morton curve, tets
    by their lowest new vertex (corner order, hence orientation, kept).
    Consecutive blocks of vertices are then compact in space, which the
    solver's workgroup-local coarse correction uses as aggregates (each PCG
    workgroup's 256 rows).  Returns (particles, tets, perm) with
    new particles = old particles[perm]."""
    particles = np.asarray(particles)
    perm = morton_order(particles)
    inv = np.empty(len(perm), dtype=np.int64)
    inv[perm] = np.arange(len(perm))
    t = inv[np.asarray(tets, dtype=np.int64)]
    t = t[np.argsort(t.min(axis=1), kind="stable")]
    return particles[perm], t.astype(np.int32), perm


def rest_pose(particles: np.ndarray, tets: np.ndarray):
    """DmInv (T,3,3) and signed rest volume (T,) of each tet."""
    x = particles[tets]  # (T,4,3)
    Dm = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)  # columns = edges
    vol = np.linalg.det(Dm) / 6.0
    if np.any(vol <= 0):
        raise ValueError(f"{int(np.sum(vol <= 0))} inverted or degenerate tets")
    return np.linalg.inv(Dm), vol


def lumped_mass(n_particles: int, tets: np.ndarray, vol: np.ndarray, density) -> np.ndarray:
    """particle mass = sum over incident tets of density * volume / 4."""
    dens = np.broadcast_to(np.asarray(density, dtype=np.float64), vol.shape)
    m = np.zeros(n_particles, dtype=np.float64)
    np.add.at(m, tets.reshape(-1), np.repeat(dens * vol / 4.0, 4))
    return m


def surface_tris(tets: np.ndarray) -> np.ndarray:
    """Boundary faces of a positively oriented tet mesh, oriented outward."""
    faces = np.concatenate(
        [tets[:, [0, 2, 1]], tets[:, [0, 1, 3]], tets[:, [0, 3, 2]], tets[:, [1, 2, 3]]], axis=0
    )
    key = np.sort(faces, axis=1)
    _, idx, counts = np.unique(key, axis=0, return_index=True, return_counts=True)
    return faces[np.sort(idx[counts == 1])].astype(np.int32)


def lame(E, nu):
    """Lame parameters (mu, lambda) from Young's modulus and Poisson ratio."""
    E = np.asarray(E, dtype=np.float64)
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def vertex_adjacency_pattern(n_particles: int, tets: np.ndarray):
    """CSR pattern (rowptr, cols) of the 3x3-block system matrix: vertex i's
    row holds i and every vertex sharing a tet with it, columns ascending."""
    a = np.repeat(tets, 4, axis=1).reshape(-1)
    b = np.tile(tets, (1, 4)).reshape(-1)
    keys = np.unique(a.astype(np.int64) * n_particles + b.astype(np.int64))
    diag = np.arange(n_particles, dtype=np.int64) * (n_particles + 1)
    keys = np.union1d(keys, diag)
    rows = (keys // n_particles).astype(np.int32)
    cols = (keys % n_particles).astype(np.int32)
    rowptr = np.zeros(n_particles + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), cols
