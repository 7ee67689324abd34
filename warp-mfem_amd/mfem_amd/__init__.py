"""mfem_amd: MI355X-native per-timestep soft-body solve (warp-mfem drop-in).

Host code (numpy / torch) mirrors mfem.refinement's solver, mesh and step API;
all per-step compute runs in libmfem_amd.so (hand-written HIP for gfx950)
through the C ABI declared in include/mfem_amd.h.
"""
from . import mesh, pattern, scene
from .scene import Model, Shape, build_model, SHAPE_CAPSULE, SHAPE_PLANE

__all__ = ["mesh", "pattern", "scene", "Model", "Shape", "build_model", "SHAPE_CAPSULE", "SHAPE_PLANE",
           "RefinementSolver", "State", "model_state"]


def __getattr__(name):
    # the solver imports torch and loads the HIP library; keep the mesh/scene
    # helpers importable without either
    if name in ("RefinementSolver", "State", "model_state"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
