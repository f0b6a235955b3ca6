"""Build libfp3d_hip.so (gfx950) in-tree with hipcc.

Usage: python -m up3d_amd.csrc.build   (or up3d_amd.csrc.build.build())
The library lands in up3d_amd/lib/ next to the Python package, so it
travels with the source tree and is the one the package loads.
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
LIB_DIR = PKG / "lib"
LIB_NAME = "libfp3d_hip.so"

HIP_SOURCES = ["fp3d_elem.hip", "fp3d_linalg.hip", "fp3d_adjoint.hip", "fp3d_sheet.hip"]
HOST_SOURCES = ["fp3d_host.cpp"]

# -ffp-contract=off: no FMA contraction, so every element's arithmetic is the
# written IEEE sequence (bit-comparable with the oracle apart from pow()).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 ops.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libfp3d_hip.so)")


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    objs = []
    for src in HIP_SOURCES + HOST_SOURCES:
        obj = obj_dir / (src + ".o")
        cmd = [hipcc, *HIP_FLAGS, "-c", str(HERE / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off",
                   "-pthread", "-c", str(HERE / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    out = LIB_DIR / LIB_NAME
    tmp = LIB_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-pthread",
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
