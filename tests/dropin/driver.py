"""Build and run tests/dropin/schedule_driver: a C++ caller that declares
pairhmm::schedule::schedule_pairhmm as the reference's header does
(src/core/pairhmm_schedule.h:14-25) and links against librodv_phmm.so."""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "schedule_driver.cpp")
EXE = os.path.join(HERE, "schedule_driver")
LIBDIR = os.path.join(ROOT, "rodv_amd")


def build(out: str = EXE) -> str:
    """g++ with the library found next to the binary's tree ($ORIGIN)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", SRC, "-o", out, f"-L{LIBDIR}", "-lrodv_phmm",
           "-Wl,-rpath,$ORIGIN/../../rodv_amd", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def ensure_built() -> str:
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < os.path.getmtime(SRC):
        build()
    return EXE


def write_input(path, haps, reads, quals, ins, dels, gcps, use_double=False, idle_f=0.1, idle_d=0.1,
                verbose=False):
    with open(path, "wb") as f:
        f.write(struct.pack("<IIBddB", len(haps), len(reads), int(use_double), idle_f, idle_d,
                            int(verbose)))
        for h in haps:
            h = bytes(h)
            f.write(struct.pack("<I", len(h)) + h)
        for r, q, i, d, c in zip(reads, quals, ins, dels, gcps):
            r = bytes(r)
            f.write(struct.pack("<I", len(r)) + r)
            for a in (q, i, d, c):
                a = np.asarray(a, np.uint8).tobytes()
                assert len(a) == len(r)
                f.write(a)


def run(haps, reads, quals, ins, dels, gcps, reps=1, timeout=120, **kw):
    """Runs the driver as a child process: (ok, M x N result, per-call us)."""
    exe = ensure_built()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fin, haps, reads, quals, ins, dels, gcps, **kw)
        p = subprocess.run([exe, fin, fout, str(reps)], capture_output=True, text=True, timeout=timeout)
        if p.returncode not in (0, 1):
            raise RuntimeError(f"schedule_driver exited {p.returncode}: {p.stderr[-2000:]}")
        with open(fout, "rb") as f:
            data = f.read()
    ok, M, N = struct.unpack_from("<BII", data, 0)
    off = 9
    res = np.frombuffer(data, np.float64, M * N, off).reshape(M, N)
    off += 8 * M * N
    (nr,) = struct.unpack_from("<I", data, off)
    us = np.frombuffer(data, np.float64, nr, off + 4)
    return bool(ok), res, us


def batch_lists(b, region=0):
    """One region of a rodv_amd Batch as the reference call's argument lists."""
    h0, m, r0, n = (int(x) for x in b.regions[region])
    haps = [b.hap(k) for k in range(h0, h0 + m)]
    reads = [b.read(k) for k in range(r0, r0 + n)]
    qs = [b.quals(k) for k in range(r0, r0 + n)]
    return haps, reads, [q[0] for q in qs], [q[1] for q in qs], [q[2] for q in qs], [q[3] for q in qs]
