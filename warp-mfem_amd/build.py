"""Builds libmfem_amd.so (gfx950) next to the Python package.

    python warp-mfem_amd/build.py

hipcc cross-compiles for gfx950 without a GPU.  Packed-FP32 instructions are
disabled (-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "mfem_amd", "libmfem_amd.so")
SOURCES = ["mfem_kernels.hip", "mfem_mlp.hip", "mfem_collide.hip", "mfem_refine.hip", "mfem_pattern.hip",
           "mfem_selfcontact.hip", "mfem_surface.hip"]
ARCH = os.environ.get("MFEM_AMD_ARCH", "gfx950")


def hipcc_command(out: str = OUT, extra: list[str] | None = None) -> list[str]:
    return [
        "hipcc",
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
        *(extra or []),
        "-o", out,
        *[os.path.join(CSRC, s) for s in SOURCES],
    ]


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = hipcc_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
