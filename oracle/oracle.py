"""ctypes bindings for the CPU oracle (oracle/libphmm_oracle.so) and, when
built, the reference kernels compiled into oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py.  The engine under rodv_amd/ never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphmm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MODE_INTRA_API, MODE_RAW_F32, MODE_RAW_F64, MODE_EXTENDED = 0, 1, 2, 3
INTRA, INTER = 0, 1
MM_SIZE = 32640

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)

_lib = None


def build() -> None:
    """Compile the oracle (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/src/core"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.phmm_oracle_batch.restype = C.c_int
        L.phmm_oracle_batch.argtypes = [
            C.c_uint32, _u8p, _u64p, _u32p,
            C.c_uint32, _u8p, _u8p, _u8p, _u8p, _u8p, _u64p, _u32p,
            C.c_uint64, _u32p, _u32p, C.c_int, C.c_int, C.c_int,
            _f64p, _f32p, _f64p, _u8p]
        L.phmm_oracle_tables_f32.argtypes = [_f32p, _f32p]
        L.phmm_oracle_tables_f64.argtypes = [_f64p, _f64p]
        L.phmm_oracle_schedule.restype = C.c_int
        L.phmm_oracle_schedule.argtypes = [
            C.c_uint32, C.POINTER(_u8p), _u32p, C.c_uint32, C.POINTER(_u8p),
            C.POINTER(_u8p), C.POINTER(_u8p), C.POINTER(_u8p), C.POINTER(_u8p), _u32p,
            C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_uint32, _f64p, _u8p, _f32p, _u8p]
        L.phmm_oracle_param_batch.restype = C.c_int
        L.phmm_oracle_param_batch.argtypes = [
            C.c_uint32, _u8p, _u64p, _u32p, C.c_uint32, _u8p, _u8p, _u64p, _u32p,
            C.c_uint64, _u32p, _u32p, C.c_void_p, _f64p, C.c_int, _f64p, _f64p, C.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def tables_f32():
    ph = np.zeros(128, np.float32)
    mm = np.zeros(MM_SIZE, np.float32)
    lib().phmm_oracle_tables_f32(_p(ph, _f32p), _p(mm, _f32p))
    return ph, mm


def tables_f64():
    ph = np.zeros(128, np.float64)
    mm = np.zeros(MM_SIZE, np.float64)
    lib().phmm_oracle_tables_f64(_p(ph, _f64p), _p(mm, _f64p))
    return ph, mm


def run_batch(batch, pairs=None, mode=MODE_INTRA_API, semantics=INTRA, nthreads=1):
    """batch: rodv_amd.batch.Batch-like object (flat arrays). pairs: (hap, read)
    index arrays; defaults to the hap-major cross product."""
    if pairs is None:
        ph, pr = batch.cross_pairs()
    else:
        ph, pr = pairs
    ph = np.ascontiguousarray(ph, np.uint32)
    pr = np.ascontiguousarray(pr, np.uint32)
    n = len(ph)
    out = np.zeros(n, np.float64)
    r32 = np.zeros(n, np.float32)
    r64 = np.zeros(n, np.float64)
    fl = np.zeros(n, np.uint8)
    rc = lib().phmm_oracle_batch(
        len(batch.hap_len), _p(batch.hap_bases, _u8p), _p(batch.hap_off, _u64p),
        _p(batch.hap_len, _u32p), len(batch.read_len), _p(batch.read_bases, _u8p),
        _p(batch.base_q, _u8p), _p(batch.ins_q, _u8p), _p(batch.del_q, _u8p),
        _p(batch.gcp_q, _u8p), _p(batch.read_off, _u64p), _p(batch.read_len, _u32p),
        n, _p(ph, _u32p), _p(pr, _u32p), mode, semantics, nthreads,
        _p(out, _f64p), _p(r32, _f32p), _p(r64, _f64p), _p(fl, _u8p))
    if rc != 0:
        raise RuntimeError("oracle batch failed")
    return out, r32, r64, fl


ROUTE_INTRA, ROUTE_INTER32, ROUTE_INTER64, ROUTE_INTRA64 = 0, 1, 2, 3


def schedule(haps, reads, quals, ins, dels, gcps, use_double=False, idle_f=0.1, idle_d=0.1,
             wf=16, wd=8, details=False):
    """schedule_pairhmm restatement over lists of uint8 arrays; returns the
    M x N result matrix and the flags (and, with details, the deciding FP32
    raw sums and the ROUTE_* of every pair)."""
    M, N = len(haps), len(reads)
    keep = []

    def arr(seqs):
        ptrs = (_u8p * max(len(seqs), 1))()
        for k, s in enumerate(seqs):
            a = np.ascontiguousarray(s, np.uint8)
            keep.append(a)
            ptrs[k] = a.ctypes.data_as(_u8p)
        return ptrs

    hl = np.array([len(h) for h in haps], np.uint32)
    rl = np.array([len(r) for r in reads], np.uint32)
    res = np.zeros(max(M * N, 1), np.float64)
    fl = np.zeros(max(M * N, 1), np.uint8)
    r32 = np.zeros(max(M * N, 1), np.float32)
    route = np.zeros(max(M * N, 1), np.uint8)
    ok = lib().phmm_oracle_schedule(
        M, arr(haps), _p(hl, _u32p), N, arr(reads), arr(quals), arr(ins), arr(dels), arr(gcps),
        _p(rl, _u32p), int(use_double), idle_f, idle_d, wf, wd, _p(res, _f64p), _p(fl, _u8p),
        _p(r32, _f32p), _p(route, _u8p))
    out = (bool(ok), res[:M * N].reshape(M, N), fl[:M * N].reshape(M, N))
    if details:
        out += (r32[:M * N].reshape(M, N), route[:M * N].reshape(M, N))
    return out


def schedule_batch(batch, use_double=False, idle_f=0.1, idle_d=0.1, wf=16, wd=8, nthreads=8):
    """One schedule_pairhmm restatement call per region of `batch`; flat
    arrays in the engine's output order (region by region, hap-major):
    (log10, flags, raw32, route)."""
    from concurrent.futures import ThreadPoolExecutor
    regs = batch.regions if len(batch.regions) else \
        np.array([[0, len(batch.hap_len), 0, len(batch.read_len)]], np.int64)

    def one(reg):
        h0, m, r0, n = (int(x) for x in reg)
        haps = [np.frombuffer(batch.hap(k), np.uint8) for k in range(h0, h0 + m)]
        reads = [np.frombuffer(batch.read(k), np.uint8) for k in range(r0, r0 + n)]
        qs = [batch.quals(k) for k in range(r0, r0 + n)]
        _, res, fl, r32, route = schedule(haps, reads, [q[0] for q in qs], [q[1] for q in qs],
                                          [q[2] for q in qs], [q[3] for q in qs],
                                          use_double=use_double, idle_f=idle_f, idle_d=idle_d,
                                          wf=wf, wd=wd, details=True)
        return res.ravel(), fl.ravel(), r32.ravel(), route.ravel()

    with ThreadPoolExecutor(max(1, nthreads)) as ex:
        parts = list(ex.map(one, regs))
    return tuple(np.concatenate([p[k] for p in parts]) for k in range(4))


class Params(C.Structure):
    _fields_ = [("gap_open", C.c_double), ("gap_cont", C.c_double),
                ("scale", C.c_double * 8)]


def param_batch(batch, theta, weights=None, pairs=None, nthreads=1, renorm=True):
    """FP64 forward + reverse-mode gradient of the parametrised model.
    theta: 10 values (gap_open, gap_cont, scale[8]).  Returns (log10L, grad).
    renorm: rows renormalised by powers of two (finite for long reads);
    False: the plain form at K = 2^1020."""
    if pairs is None:
        ph, pr = batch.cross_pairs()
    else:
        ph, pr = pairs
    ph = np.ascontiguousarray(ph, np.uint32)
    pr = np.ascontiguousarray(pr, np.uint32)
    n = len(ph)
    th = Params()
    th.gap_open, th.gap_cont = float(theta[0]), float(theta[1])
    for k in range(8):
        th.scale[k] = float(theta[2 + k])
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    out = np.zeros(n, np.float64)
    grad = np.zeros(10, np.float64)
    lib().phmm_oracle_param_batch(
        len(batch.hap_len), _p(batch.hap_bases, _u8p), _p(batch.hap_off, _u64p),
        _p(batch.hap_len, _u32p), len(batch.read_len), _p(batch.read_bases, _u8p),
        _p(batch.base_q, _u8p), _p(batch.read_off, _u64p), _p(batch.read_len, _u32p),
        n, _p(ph, _u32p), _p(pr, _u32p), C.byref(th), _p(w, _f64p), nthreads,
        _p(out, _f64p), _p(grad, _f64p), int(bool(renorm)))
    return out, grad


# ------------------------------------------------------------ reference build

def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


class RefLib:
    """The reference's own intra (and AVX2 inter) kernels, compiled from its
    sources into oracle/_ref/.  isa: 'avx512' or 'avx2'."""

    def __init__(self, isa):
        path = os.path.join(REF_DIR, f"libphmm_ref_{isa}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        need = {"avx512": {"avx512f", "avx512dq", "avx512bw", "avx512vl"},
                "avx2": {"avx2", "fma"}}[isa]
        if not need <= _cpu_flags():
            raise RuntimeError(f"host CPU lacks {isa}")
        self.isa = isa
        L = C.CDLL(path)
        pre = f"phmm_ref_{isa}_"
        args = [C.c_uint32, C.c_uint32] + [_u8p] * 6
        self._raw32 = getattr(L, pre + "intra_raw_f32")
        self._raw32.restype = C.c_float
        self._raw32.argtypes = args
        self._raw64 = getattr(L, pre + "intra_raw_f64")
        self._raw64.restype = C.c_double
        self._raw64.argtypes = args
        self._log10 = getattr(L, pre + "intra_log10")
        self._log10.restype = C.c_double
        self._log10.argtypes = args + [C.c_int]
        self._batch = getattr(L, pre + "intra_log10_batch")
        self._batch.restype = C.c_int
        self._batch.argtypes = [_u8p, _u64p, _u32p, _u8p, _u8p, _u8p, _u8p, _u8p, _u64p, _u32p,
                                C.c_uint64, _u32p, _u32p, C.c_int, _f64p]
        self._t32 = getattr(L, pre + "tables_f32")
        self._t32.argtypes = [_f32p, _f32p, C.c_int]
        self._t64 = getattr(L, pre + "tables_f64")
        self._t64.argtypes = [_f64p, _f64p, C.c_int]
        self._inter32 = self._inter64 = None
        if isa == "avx2":
            gargs = [_u32p, _u32p] + [C.POINTER(_u8p)] * 6 + [_f64p]
            self._inter32 = L.phmm_ref_avx2_inter_raw_f32
            self._inter32.argtypes = gargs
            self._inter64 = L.phmm_ref_avx2_inter_raw_f64
            self._inter64.argtypes = gargs

    @staticmethod
    def _pair_args(batch, h, r):
        ho, ro = int(batch.hap_off[h]), int(batch.read_off[r])
        rl, hl = int(batch.read_len[r]), int(batch.hap_len[h])

        def at(a, o):
            return C.cast(C.c_void_p(a.ctypes.data + o), _u8p)
        return (rl, hl, at(batch.hap_bases, ho), at(batch.read_bases, ro), at(batch.base_q, ro),
                at(batch.ins_q, ro), at(batch.del_q, ro), at(batch.gcp_q, ro))

    def raw_f32(self, batch, h, r):
        return self._raw32(*self._pair_args(batch, h, r))

    def raw_f64(self, batch, h, r):
        return self._raw64(*self._pair_args(batch, h, r))

    def log10(self, batch, h, r, use_double=False):
        return self._log10(*self._pair_args(batch, h, r), int(use_double))

    def log10_batch(self, batch, pairs, nthreads):
        """computeLikelihoods over the (hap, read) index arrays on nthreads."""
        ph = np.ascontiguousarray(pairs[0], np.uint32)
        pr = np.ascontiguousarray(pairs[1], np.uint32)
        out = np.zeros(len(ph), np.float64)
        self._batch(_p(batch.hap_bases, _u8p), _p(batch.hap_off, _u64p), _p(batch.hap_len, _u32p),
                    _p(batch.read_bases, _u8p), _p(batch.base_q, _u8p), _p(batch.ins_q, _u8p),
                    _p(batch.del_q, _u8p), _p(batch.gcp_q, _u8p), _p(batch.read_off, _u64p),
                    _p(batch.read_len, _u32p), len(ph), _p(ph, _u32p), _p(pr, _u32p),
                    int(nthreads), _p(out, _f64p))
        return out

    def tables_f32(self):
        ph = np.zeros(128, np.float32)
        mm = np.zeros(MM_SIZE, np.float32)
        self._t32(_p(ph, _f32p), _p(mm, _f32p), MM_SIZE)
        return ph, mm

    def tables_f64(self):
        ph = np.zeros(128, np.float64)
        mm = np.zeros(MM_SIZE, np.float64)
        self._t64(_p(ph, _f64p), _p(mm, _f64p), MM_SIZE)
        return ph, mm

    def inter_raw(self, batch, hs, rs, double=False):
        W = 4 if double else 8
        assert len(hs) == W and len(rs) == W
        fn = self._inter64 if double else self._inter32
        cols = [self._pair_args(batch, h, r) for h, r in zip(hs, rs)]
        rl = np.array([c[0] for c in cols], np.uint32)
        hl = np.array([c[1] for c in cols], np.uint32)
        ptrs = [(_u8p * W)(*[c[2 + k] for c in cols]) for k in range(6)]
        out = np.zeros(W, np.float64)
        fn(_p(rl, _u32p), _p(hl, _u32p), *ptrs, _p(out, _f64p))
        return out
