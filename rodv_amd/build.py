"""Builds the engine's shared library (HIP for gfx950) in-tree.

    python -m rodv_amd.build        -> rodv_amd/librodv_phmm.so
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librodv_phmm.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("phmm_kernels.hip", "phmm_adjoint.hip", "phmm_host.cpp")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: only the explicit fmas of the kernels fuse (bit-exact
# FP32 parity).  Denormals stay IEEE like the reference library.  Packed FP32
# is disabled for all GPU code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
         "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}"]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *SOURCES, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (res.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
