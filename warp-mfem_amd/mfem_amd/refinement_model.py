"""Tet-mesh files of the reference (``MFEMRefinementModel``,
src/mfem/refinement/models/refinement_model.py:13-143) without newton.

An ``.npz`` holds ``particles`` (N, 3) float32, ``tet_indices`` (M, 4) int32,
``tet_materials`` (M, 3) float32 [mu, lambda, damping] and ``parameters``
float64 [max_particles, max_tets, density, max_tris] (max_tris may be
missing in older files: 0).  ``load`` reads it with ``allow_pickle=False``
(plain arrays only); ``load_sim_model`` builds the solver's ``Model`` the way
the reference's ``load_sim_model`` drives newton's ``add_soft_mesh``:
vertices rotated by 90 degrees about x, the particle buffer padded to
max_particles (zero rows, no mass, inactive), lumped masses from the file's
density, the surface triangles of the mesh, mu the given scalar and lambda /
damping per tet from the file.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mesh as _mesh
from .scene import Model, build_model

PARAM_MAX_PARTICLES = 0
PARAM_MAX_TETS = 1
PARAM_DENSITY = 2
PARAM_MAX_TRIS = 3


@dataclass
class MFEMRefinementModel:
    particles: np.ndarray
    tet_indices: np.ndarray
    tet_materials: np.ndarray
    parameters: np.ndarray

    @classmethod
    def from_model(cls, model: Model, max_particles: int, max_tets: int, max_tris: int = 0) -> "MFEMRefinementModel":
        """refinement_model.py:19-31 (density stored as 1.0, as there)."""
        n = int(np.asarray(model.tet_indices).max()) + 1
        return cls(particles=np.asarray(model.particle_q, np.float32)[:n].copy(),
                   tet_indices=np.asarray(model.tet_indices, np.int32).copy(),
                   tet_materials=np.asarray(model.tet_materials, np.float32).copy(),
                   parameters=np.array([max_particles, max_tets, 1.0, max_tris], np.float64))

    def save(self, file: str) -> None:
        np.savez(file, particles=self.particles, tet_indices=self.tet_indices, tet_materials=self.tet_materials,
                 parameters=self.parameters)

    @staticmethod
    def load(file: str, scale: float = 1.0, translation=0.0) -> "MFEMRefinementModel":
        """refinement_model.py:42-55: positions scaled then translated."""
        with np.load(file, allow_pickle=False) as z:
            particles = np.asarray(z["particles"])
            tets = np.asarray(z["tet_indices"])
            mats = np.asarray(z["tet_materials"])
            params = np.asarray(z["parameters"], np.float64)
        particles = (particles * np.float32(scale) + np.asarray(translation, np.float32)).astype(np.float32)
        return MFEMRefinementModel(particles=particles, tet_indices=tets.astype(np.int32),
                                   tet_materials=mats.astype(np.float32), parameters=params)

    @property
    def max_particles(self) -> int:
        return int(self.parameters[PARAM_MAX_PARTICLES])

    @property
    def max_tets(self) -> int:
        return int(self.parameters[PARAM_MAX_TETS])

    @property
    def max_tris(self) -> int:
        """0 when the file predates the field (the solver then derives it)."""
        if len(self.parameters) <= PARAM_MAX_TRIS:
            return 0
        return int(self.parameters[PARAM_MAX_TRIS])

    @property
    def density(self) -> float:
        return float(self.parameters[PARAM_DENSITY])

    def resolve_max_particles(self, mult: float | None) -> int:
        return self.max_particles if mult is None else math.ceil(mult * self.particles.shape[0])

    def resolve_max_tets(self, mult: float | None) -> int:
        return self.max_tets if mult is None else math.ceil(mult * self.tet_indices.shape[0])

    def resolve_max_tris(self, mult: float | None, model: Model) -> int:
        return self.max_tris if mult is None else math.ceil(mult * model.tri_indices.shape[0])

    def load_sim_model(self, mu: float, max_particles: int | None = None, gravity=(0.0, 0.0, -9.81), shapes=(),
                       rotate: bool = True) -> Model:
        """The solver's Model (refinement_model.py:104-143): positions rotated
        90 degrees about x (add_soft_mesh's rot), the particle buffer padded
        to ``max_particles`` (default: the file's), masses lumped with the
        file's density, materials [mu, lambda, damping] with the scalar mu.
        Pass the same ``max_particles`` to RefinementSolver."""
        if max_particles is None:
            max_particles = self.max_particles
        n = self.particles.shape[0]
        if max_particles < n:
            raise ValueError(f"max_particles={max_particles} < the mesh's {n} particles")
        p = np.zeros((max_particles, 3), np.float64)
        p[:n] = self.particles
        if rotate:
            R = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])  # quat_from_euler((pi/2, 0, 0))
            p = p @ R.T
        mats = np.asarray(self.tet_materials, np.float64).copy()
        mats[:, 0] = mu
        tets = np.asarray(self.tet_indices, np.int32)
        # rest poses, volumes and masses over the mesh's own particles; the
        # padding rows stay massless (inv_mass 0, inactive)
        m = build_model(p[:n], tets, mats, self.density, shapes=shapes, gravity=gravity, name="npz_mesh")
        out = Model(particle_q=p, particle_qd=np.zeros_like(p),
                    particle_mass=np.concatenate([m.particle_mass, np.zeros(max_particles - n)]),
                    particle_inv_mass=np.concatenate([m.particle_inv_mass, np.zeros(max_particles - n)]),
                    tet_indices=tets, tet_poses=m.tet_poses, tet_materials=mats, tri_indices=m.tri_indices,
                    shapes=list(shapes), gravity=np.asarray(gravity, np.float64), params=dict(m.params),
                    name="npz_mesh")
        return out


def surface_triangles(tet_indices: np.ndarray) -> np.ndarray:
    """Outward boundary triangles of a tet mesh (faces used by one tet)."""
    return _mesh.surface_tris(np.asarray(tet_indices, np.int32))
