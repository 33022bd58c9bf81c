"""Build the HIP library (gfx950) in-tree.

`python -m minirender_amd.build` compiles every source under csrc/ with hipcc
and links minirender_amd/libminirender_amd.so.  Device code is compiled with
-ffp-contract=off (no fused multiply-add: the reference's fp32 operation
order is the parity contract) and without packed-fp32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJDIR = PKG / "_build"
LIB = PKG / "libminirender_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
SOURCES = ["kernels.hip", "context.cpp", "projection.cpp", "cpp_api.cpp", "io.cpp",
           "x3d.cpp", "server.cpp", "tune.cpp"]


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.name + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    return obj


def build_library(verbose: bool = True) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), 8)) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs),
           "-o", str(LIB), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build_library()
    sys.exit(0)
