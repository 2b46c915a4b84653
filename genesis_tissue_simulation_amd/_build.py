"""Compiles the HIP library libgts_xpbd.so in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIB = PACKAGE / "libgts_xpbd.so"
SOURCES = ("gts_topology.hip", "gts_kernels.hip", "gts_batched.hip", "gts_adjoint.hip", "gts_surface.hip")

# -ffp-contract=off: no multiply-add fusion, so kernels round like the host code and the CPU checker;
# correctly rounded f32 division and square root; packed FP32 instructions disabled.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def build_library(verbose=False):
    sources = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp), *sources]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=str(PACKAGE))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
