"""Build the gfx950 shared library ``libffsreml.so`` in-tree with hipcc.

Usage: ``python -m fastfuncstuff_amd.build`` (or ``build()`` from Python).
The library is written next to this file so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libffsreml.so"
# reml_api.hip includes reml_kernels.hip (one translation unit).
SOURCES = ["reml_api.hip", "mlp_kernels.hip"]

# Packed FP32 instructions (v_pk_*_f32) are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(cand).exists():
        raise RuntimeError("hipcc not found; a ROCm toolchain is required to build libffsreml.so")
    return cand


def build(verbose: bool = False, arch: str = "gfx950") -> Path:
    """Compile every HIP source for ``arch`` and link ``libffsreml.so``."""
    objs = []
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = build_dir / (path.stem + ".o")
        cmd = [hipcc(), f"--offload-arch={arch}", "-O3", "-fPIC", "-std=c++17", "-I", str(INCLUDE),
               "-I", str(CSRC), *DEVICE_FLAGS, "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={arch}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
