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
CLI = PKG / "sdf2mesh"         # command line tool (cli.cpp) over the C ABI
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
SOURCES = ["kernels.hip", "sparse.hip", "api.cpp", "hostmem.cpp", "tune.cpp", "server.cpp", "writer.cpp", "compile.cpp",
           "frontend_parse.cpp", "frontend_check.cpp", "frontend_emit.cpp"]

# Device headers a program compiled at run time includes, in order; written into one
# string (rtc_source.inc) that hiprtc receives ahead of the generated code.
RTC_HEADERS = [PKG.parent / "include" / "sdf2mesh.h", CSRC / "device_sdf.h", CSRC / "device_dense.h",
               CSRC / "rtc_prelude.h"]
RTC_TYPES = """typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int32_t int32_t;
"""


def _write_rtc_source() -> None:
    """Concatenate the device headers (without their local #include lines and include
    guards' #pragma once) into a C++ raw string literal."""
    parts = [RTC_TYPES]
    for h in RTC_HEADERS:
        for line in h.read_text().splitlines():
            t = line.strip()
            if t == "#pragma once" or (t.startswith("#include \"") and t.endswith(".h\"")):
                continue
            parts.append(line + "\n")
    text = "".join(parts)
    assert ")s2mrtc\"" not in text
    out = BUILD / "rtc_source.inc"
    new = 'R"s2mrtc(' + text + ')s2mrtc"\n'
    if not out.exists() or out.read_text() != new:
        out.write_text(new)


def _compile(src: str) -> Path:
    obj = BUILD / (src + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += DEVICE
    else:
        cmd += ["-x", "c++", f"-I{ROCM / 'include'}", f"-I{BUILD}", "-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    _write_rtc_source()
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), f"-L{ROCM / 'lib'}", "-lhiprtc", "-lpthread",
           f"-Wl,-rpath,{ROCM / 'lib'}"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    tmp_cli = CLI.with_suffix(".tmp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", str(tmp_cli), str(CSRC / "cli.cpp"),
                    f"-L{PKG}", "-lsdf2mesh", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(tmp_cli, CLI)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
