"""Compile the HIP step library for gfx950 (hipcc, in-tree).

``python -m zhenmode_amd.build`` writes ``zhenmode_amd/libzhenmode_hip.so``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libzhenmode_hip.so"
SOURCES = ["zm_kernels.hip", "zm_column.hip", "zm_runtime.hip", "zm_project.hip", "zm_ledger.hip", "zm_host.hip"]
ARCH = os.environ.get("ZHENMODE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP step library cannot be built")
    return path


def compile_flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 instructions stay disabled in device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value", "-I", str(CSRC), "-I", str(HERE.parent / "include")]


def build(verbose=False):
    """Compile every source to an object, then link the shared library."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [hipcc(), *compile_flags(), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
