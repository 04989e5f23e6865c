"""ctypes front end of the C ABI (include/rodv_phmm.h).

Mirrors the reference's batch API for this path:
``schedule_pairhmm(haplotypes, reads, result, quality, insertion_qualities,
deletion_qualities, gap_contiguous_qualities, use_double, max_idle_ratio_float,
max_idle_ratio_double, verbose)`` (pairhmm_schedule.h:14-25) -- same argument
meaning, same hap-major result matrix, same ``False`` for an empty batch.

The HIP library must be built (``python -m rodv_amd.build``); there is no
CPU fallback: without the library every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .batch import Batch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librodv_phmm.so")

OK, ERR_INVALID_ARG, ERR_OOM, ERR_DEVICE, ERR_UNSUPPORTED, ERR_EMPTY, ERR_TOO_LARGE = range(7)
NOALLOC_SCRATCH_BYTES = 32 << 20  # RODV_PHMM_NOALLOC_SCRATCH_BYTES
FLAG_RESCUED, FLAG_UNDERFLOW, FLAG_AMBIGUOUS, FLAG_INTER, FLAG_EXTENDED = 1, 2, 4, 8, 0x20
SEMANTICS_INTRA, SEMANTICS_SCHEDULE = 0, 1

# exported symbols of include/rodv_phmm.h
EXPORTS = (
    "rodv_phmm_init", "rodv_phmm_strerror", "rodv_phmm_config_default",
    "rodv_phmm_ctx_create", "rodv_phmm_ctx_destroy", "rodv_phmm_forward",
    "rodv_phmm_forward_noalloc",
    "rodv_phmm_plan_create", "rodv_phmm_plan_execute", "rodv_phmm_plan_execute_plain",
    "rodv_phmm_plan_profile", "rodv_phmm_plan_read", "rodv_phmm_plan_stats",
    "rodv_phmm_plan_num_pairs", "rodv_phmm_plan_destroy", "rodv_phmm_tables_f32",
    "rodv_phmm_tables_f64", "rodv_phmm_adjoint", "rodv_phmm_autotune",
    "rodv_phmm_set_tuning", "rodv_phmm_get_tuning",
)


class PhmmError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {_strerror(code)} ({code})")


class Region(C.Structure):
    _fields_ = [("hap_begin", C.c_uint32), ("num_haps", C.c_uint32),
                ("read_begin", C.c_uint32), ("num_reads", C.c_uint32)]


class CBatch(C.Structure):
    _fields_ = [
        ("num_haps", C.c_uint32), ("hap_bases", C.c_void_p), ("hap_offsets", C.c_void_p),
        ("hap_lengths", C.c_void_p), ("num_reads", C.c_uint32), ("read_bases", C.c_void_p),
        ("base_quals", C.c_void_p), ("ins_quals", C.c_void_p), ("del_quals", C.c_void_p),
        ("gcp_quals", C.c_void_p), ("read_offsets", C.c_void_p), ("read_lengths", C.c_void_p),
        ("num_regions", C.c_uint32), ("regions", C.c_void_p)]


class Config(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("semantics", C.c_int32), ("use_double", C.c_int32),
        ("max_idle_ratio_float", C.c_double), ("max_idle_ratio_double", C.c_double),
        ("simd_width_float", C.c_uint32), ("simd_width_double", C.c_uint32),
        ("use_graph", C.c_int32), ("waves_per_cu", C.c_int32), ("short_lanes", C.c_int32),
        ("extended_range", C.c_int32), ("prior_table", C.c_int32), ("use_tuning", C.c_int32),
        ("haplotype_chains", C.c_int32), ("shared_products", C.c_int32)]


TUNE_CLASSES = 18


class TuneEntry(C.Structure):
    _fields_ = [
        ("max_read_len", C.c_uint32), ("pairs_per_wave", C.c_uint32),
        ("max_rows_per_lane", C.c_uint32), ("waves_per_cu", C.c_uint32),
        ("sample_pairs", C.c_uint32), ("gcups", C.c_double), ("default_gcups", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("num_classes", C.c_uint32),
                ("cls", TuneEntry * TUNE_CLASSES)]

    def entries(self):
        """The classes as dicts (pairs_per_wave 0: built-in shape)."""
        return [{f: getattr(e, f) for f, _ in TuneEntry._fields_} for e in self.cls]


class Stats(C.Structure):
    _fields_ = [
        ("num_pairs", C.c_uint64), ("cells", C.c_uint64), ("rescued_pairs", C.c_uint64),
        ("rescued_cells", C.c_uint64), ("launches", C.c_uint64), ("num_buckets", C.c_uint32),
        ("fp32_kernel_ms", C.c_double), ("fp64_kernel_ms", C.c_double),
        ("total_ms", C.c_double), ("stacked_pairs", C.c_uint64), ("shared_pairs", C.c_uint64)]


class Params(C.Structure):
    _fields_ = [("gap_open", C.c_double), ("gap_cont", C.c_double),
                ("scale", C.c_double * 8)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP engine.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build the HIP engine with `python -m rodv_amd.build`")
    L = C.CDLL(path)
    vp, i32, u64 = C.c_void_p, C.c_int, C.c_uint64
    L.rodv_phmm_init.restype = i32
    L.rodv_phmm_strerror.restype = C.c_char_p
    L.rodv_phmm_strerror.argtypes = [i32]
    L.rodv_phmm_config_default.argtypes = [C.POINTER(Config)]
    L.rodv_phmm_ctx_create.argtypes = [i32, C.POINTER(vp)]
    L.rodv_phmm_ctx_destroy.argtypes = [vp]
    L.rodv_phmm_forward.argtypes = [vp, C.POINTER(CBatch), C.POINTER(Config), vp, vp]
    L.rodv_phmm_forward_noalloc.argtypes = [vp, C.POINTER(CBatch), C.POINTER(Config), vp, vp]
    L.rodv_phmm_plan_create.argtypes = [vp, C.POINTER(CBatch), C.POINTER(Config), C.POINTER(vp)]
    L.rodv_phmm_plan_execute.argtypes = [vp, vp]
    L.rodv_phmm_plan_execute_plain.argtypes = [vp, vp]
    L.rodv_phmm_plan_profile.argtypes = [vp, vp]
    L.rodv_phmm_plan_read.argtypes = [vp, vp, vp, vp, vp]
    L.rodv_phmm_plan_stats.argtypes = [vp, C.POINTER(Stats)]
    L.rodv_phmm_plan_num_pairs.restype = u64
    L.rodv_phmm_plan_num_pairs.argtypes = [vp]
    L.rodv_phmm_plan_destroy.argtypes = [vp]
    L.rodv_phmm_tables_f32.argtypes = [vp, vp, i32]
    L.rodv_phmm_tables_f64.argtypes = [vp, vp, i32]
    L.rodv_phmm_adjoint.argtypes = [vp, C.POINTER(CBatch), C.POINTER(Params), vp, i32, vp, vp, vp]
    L.rodv_phmm_autotune.argtypes = [vp, C.POINTER(CBatch), C.POINTER(Config), C.POINTER(Tuning)]
    L.rodv_phmm_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.rodv_phmm_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    for name in EXPORTS:
        getattr(L, name)  # every declared entry point must be exported
    _lib = L
    return L


def _strerror(code):
    try:
        return load_library().rodv_phmm_strerror(code).decode()
    except Exception:  # library missing while formatting an error
        return "error"


def _check(code, what):
    if code != OK:
        raise PhmmError(code, what)


def _ptr(a):
    return None if a is None else a.ctypes.data


def default_config(**kw) -> Config:
    cfg = Config()
    load_library().rodv_phmm_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def tables_f32(mm_n=8256):
    ph = np.zeros(128, np.float32)
    mm = np.zeros(mm_n, np.float32)
    _check(load_library().rodv_phmm_tables_f32(_ptr(ph), _ptr(mm), mm_n), "tables_f32")
    return ph, mm


def tables_f64(mm_n=8256):
    ph = np.zeros(128, np.float64)
    mm = np.zeros(mm_n, np.float64)
    _check(load_library().rodv_phmm_tables_f64(_ptr(ph), _ptr(mm), mm_n), "tables_f64")
    return ph, mm


class _CBatchHolder:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, b: Batch):
        self.b = b
        self.arrays = [np.ascontiguousarray(x) for x in (
            b.hap_bases, b.hap_off, b.hap_len, b.read_bases, b.base_q, b.ins_q, b.del_q,
            b.gcp_q, b.read_off, b.read_len)]
        (hb, ho, hl, rb, q, i, d, c, ro, rl) = self.arrays
        self.regions = np.ascontiguousarray(b.regions, np.uint32).reshape(-1, 4)
        self.c = CBatch(len(hl), _ptr(hb), _ptr(ho), _ptr(hl), len(rl), _ptr(rb), _ptr(q),
                        _ptr(i), _ptr(d), _ptr(c), _ptr(ro), _ptr(rl), len(self.regions),
                        _ptr(self.regions) if len(self.regions) else None)


class Engine:
    """One HIP device context (rodv_phmm_ctx)."""

    def __init__(self, device: int = 0):
        L = load_library()
        _check(L.rodv_phmm_init(), "init")
        h = C.c_void_p()
        _check(L.rodv_phmm_ctx_create(device, C.byref(h)), "ctx_create")
        self._h = h
        self._lib = L

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rodv_phmm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, batch: Batch, config: Config | None = None) -> "Plan":
        return Plan(self, batch, config)

    def forward(self, batch: Batch, config: Config | None = None):
        """Synchronous batch forward: (log10 likelihoods, flags), flat in
        output order (region by region, hap-major)."""
        hold = _CBatchHolder(batch)
        n = batch.num_pairs()
        out = np.zeros(n, np.float64)
        fl = np.zeros(n, np.uint8)
        cfg = config or default_config()
        _check(self._lib.rodv_phmm_forward(self._h, C.byref(hold.c), C.byref(cfg),
                                           _ptr(out), _ptr(fl)), "forward")
        return out, fl

    def forward_noalloc(self, batch: Batch, config: Config | None = None):
        """forward() without device allocation (rodv_phmm_forward_noalloc):
        the same results, from the library's fixed device scratch; raises
        PhmmError(ERR_TOO_LARGE) for a batch that does not fit it."""
        hold = _CBatchHolder(batch)
        n = batch.num_pairs()
        out = np.zeros(n, np.float64)
        fl = np.zeros(n, np.uint8)
        cfg = config or default_config()
        _check(self._lib.rodv_phmm_forward_noalloc(self._h, C.byref(hold.c), C.byref(cfg),
                                                   _ptr(out), _ptr(fl)), "forward_noalloc")
        return out, fl

    def schedule_pairhmm(self, haplotypes, reads, result, quality, insertion_qualities,
                         deletion_qualities, gap_contiguous_qualities, use_double=False,
                         max_idle_ratio_float=0.1, max_idle_ratio_double=0.1, verbose=False):
        """Same call shape and meaning as pairhmm::schedule::schedule_pairhmm
        (pairhmm_schedule.cpp:116-340): fills `result` (a list) with M rows of
        N log10 likelihoods and returns True, or False for an empty batch."""
        M, N = len(haplotypes), len(reads)
        if M == 0 or N == 0:
            return False
        b = Batch.from_lists(haplotypes, reads, quality, insertion_qualities,
                             deletion_qualities, gap_contiguous_qualities)
        cfg = default_config(use_double=int(bool(use_double)),
                             max_idle_ratio_float=max_idle_ratio_float,
                             max_idle_ratio_double=max_idle_ratio_double)
        out, _ = self.forward(b, cfg)
        result[:] = [list(row) for row in out.reshape(M, N)]
        return True

    def autotune(self, sample: Batch, config: Config | None = None) -> Tuning:
        """Times every candidate shape per read-length class on `sample`
        (rodv_phmm_autotune) and keeps the fastest in this context."""
        hold = _CBatchHolder(sample)
        t = Tuning()
        cfg = config if config is not None else default_config()
        _check(self._lib.rodv_phmm_autotune(self._h, C.byref(hold.c), C.byref(cfg), C.byref(t)),
               "autotune")
        return t

    def set_tuning(self, tuning: Tuning | None):
        """Installs a saved tuning; None restores the built-in shapes."""
        _check(self._lib.rodv_phmm_set_tuning(self._h, None if tuning is None else C.byref(tuning)),
               "set_tuning")

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(self._lib.rodv_phmm_get_tuning(self._h, C.byref(t)), "get_tuning")
        return t

    def adjoint(self, batch: Batch, theta, weights=None, ckpt_stride=32, timing=False):
        """log10 L per pair and grad[10] = sum_k w_k d log10 L_k / d theta for
        the parametrised model (include/rodv_phmm.h).  theta = (gap_open,
        gap_cont, scale[0..7]).  With timing=True also the kernel ms."""
        hold = _CBatchHolder(batch)
        n = batch.num_pairs()
        p = Params()
        p.gap_open, p.gap_cont = float(theta[0]), float(theta[1])
        for k in range(8):
            p.scale[k] = float(theta[2 + k])
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        if w is not None and len(w) != n:
            raise ValueError("one weight per pair")
        out = np.zeros(n, np.float64)
        grad = np.zeros(10, np.float64)
        ms = C.c_double(0.0)
        _check(self._lib.rodv_phmm_adjoint(self._h, C.byref(hold.c), C.byref(p), _ptr(w),
                                           int(ckpt_stride), _ptr(out), _ptr(grad),
                                           C.byref(ms)), "adjoint")
        return (out, grad, ms.value) if timing else (out, grad)


class Plan:
    """rodv_phmm_plan: a validated, bucketed batch resident in HBM."""

    def __init__(self, engine: Engine, batch: Batch, config: Config | None = None):
        self._lib = engine._lib
        self.engine = engine
        hold = _CBatchHolder(batch)
        cfg = config or default_config()
        h = C.c_void_p()
        _check(self._lib.rodv_phmm_plan_create(engine._h, C.byref(hold.c), C.byref(cfg),
                                               C.byref(h)), "plan_create")
        self._h = h
        self.num_pairs = int(self._lib.rodv_phmm_plan_num_pairs(h))

    def execute(self, stream=None):
        _check(self._lib.rodv_phmm_plan_execute(self._h, stream), "plan_execute")

    def execute_plain(self, stream=None):
        _check(self._lib.rodv_phmm_plan_execute_plain(self._h, stream), "plan_execute_plain")

    def profile(self, stream=None):
        _check(self._lib.rodv_phmm_plan_profile(self._h, stream), "plan_profile")
        return self.stats()

    def read(self, raw=False):
        n = self.num_pairs
        out = np.zeros(n, np.float64)
        fl = np.zeros(n, np.uint8)
        r32 = np.zeros(n, np.float32) if raw else None
        r64 = np.zeros(n, np.float64) if raw else None
        _check(self._lib.rodv_phmm_plan_read(self._h, _ptr(out), _ptr(fl), _ptr(r32), _ptr(r64)),
               "plan_read")
        return (out, fl, r32, r64) if raw else (out, fl)

    def stats(self) -> dict:
        s = Stats()
        _check(self._lib.rodv_phmm_plan_stats(self._h, C.byref(s)), "plan_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rodv_phmm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
