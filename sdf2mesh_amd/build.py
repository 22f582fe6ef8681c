"""Build libsdf2mesh.so in-tree with hipcc for gfx950.

    python -m sdf2mesh_amd.build        # or sdf2mesh_amd.build.build()

All device code is compiled with -ffp-contract=off and correctly rounded fp32 division
and square root, so fp32 results follow the written operation order exactly; packed
FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsdf2mesh.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("S2M_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["kernels.hip", "sparse.hip", "api.cpp", "writer.cpp"]


def _compile(src: str) -> Path:
    obj = BUILD / (src + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += DEVICE
    else:
        cmd += ["-x", "c++", f"-I{ROCM / 'include'}", "-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
