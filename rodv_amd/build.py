"""Builds the engine's shared library (HIP for gfx950) in-tree.

    python -m rodv_amd.build        -> rodv_amd/librodv_phmm.so
                                       rodv_amd/librodv_bam.so (BAM input, host code)
                                       rodv_amd/libvalu_ceiling.so (benchmark only)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librodv_phmm.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("phmm_kernels.hip", "phmm_adjoint.hip", "phmm_host.cpp",
                                                   "phmm_schedule.cpp", "phmm_noalloc.hip")]
# BAM reader/writer over zlib (host code, built with the host compiler)
BAM_LIB = os.path.join(HERE, "librodv_bam.so")
BAM_SOURCES = [os.path.join(HERE, "csrc", "bam_io.cpp")]
CXX = os.environ.get("CXX", "g++")
BAM_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-fvisibility=hidden",
             f"-I{os.path.join(ROOT, 'include')}"]
# FP32 VALU issue-ceiling probe used by bench.py's roofline (not part of the engine)
CEILING_LIB = os.path.join(HERE, "libvalu_ceiling.so")
CEILING_SOURCES = [os.path.join(HERE, "csrc", "valu_ceiling.hip")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: only the explicit fmas of the kernels fuse (bit-exact
# FP32 parity).  Denormals stay IEEE like the reference library.  Packed FP32
# is disabled for all GPU code.  The atomic optimizer is off: it turns a
# one-lane atomicAdd into a form that waits for the result right away, where
# the table kernel fetches its next work item early and reads it late.
# Unaligned access mode is off so that the table kernel's per-step byte loads
# of haplotype codes are not merged into unaligned wider reads.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
         "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Xclang", "-target-feature", "-Xclang", "-unaligned-access-mode",
         "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
         "-Wall", "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}"]


# the engine exports its C ABI and schedule_pairhmm only
ENGINE_FLAGS = ["-fvisibility=hidden", "-fvisibility-inlines-hidden",
                f"-Wl,--version-script={os.path.join(HERE, 'csrc', 'phmm.map')}"]


def build(verbose: bool = False) -> str:
    jobs = []
    cmds = [(LIB, [HIPCC, *FLAGS, *ENGINE_FLAGS, *SOURCES]),
            (CEILING_LIB, [HIPCC, *FLAGS, *CEILING_SOURCES]),
            (BAM_LIB, [CXX, *BAM_FLAGS, *BAM_SOURCES, "-lz",
                       f"-Wl,--version-script={os.path.join(HERE, 'csrc', 'bam_io.map')}"])]
    for out, cmd in cmds:
        jobs.append((out, subprocess.Popen(cmd + ["-o", out + ".tmp"], stdout=subprocess.PIPE,
                                           stderr=subprocess.PIPE, text=True)))
    failed = []
    for out, proc in jobs:
        _, stderr = proc.communicate()
        errs = "\n".join(l for l in (stderr or "").splitlines()
                         if "packed-fp32-ops" not in l and "unaligned-access-mode" not in l)
        if proc.returncode != 0:
            failed.append(f"hipcc -> {os.path.basename(out)} failed ({proc.returncode}):\n{errs}")
            continue
        if verbose and errs.strip():
            print(errs, file=sys.stderr)
        os.replace(out + ".tmp", out)
    if failed:
        raise RuntimeError("\n".join(failed))
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
