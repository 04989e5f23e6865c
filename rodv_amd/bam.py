"""BAM input for the PairHMM engine: ctypes binding of include/rodv_bam.h.

`BamReader` and `BamWriter` mirror the reference's `BamHandle`
(src/iostream/bam_handle.h:6-88): open a BAM for reading (with its `.bai`
index when one sits next to it) or for writing, iterate the records
(`next` / `has_next`), query a region (`query_from_region`, by string or by
reference index and 0-based interval, and `query_from_multi_region`), write
records and close.  Errors raise `BamError` (the reference throws
std::runtime_error).  `region_batch` turns the reads of a region into one
`schedule_pairhmm` region against given haplotypes (a `Batch`).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from .batch import Batch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librodv_bam.so")

OK, ERR_IO, ERR_FORMAT, ERR_ARG, ERR_NO_INDEX, ERR_REGION = 0, -1, -2, -3, -4, -5
_NAMES = {ERR_IO: "I/O error", ERR_FORMAT: "not BGZF/BAM/BAI or corrupt", ERR_ARG: "invalid argument",
          ERR_NO_INDEX: "no BAI index", ERR_REGION: "unknown reference or bad interval"}
CIGAR_OPS = "MIDNSHP=X"
# GATK's default read filters for the PairHMM's input: unmapped, secondary,
# QC-failed, duplicate and supplementary records are left out
SKIP_FLAGS = 0x4 | 0x100 | 0x200 | 0x400 | 0x800

EXPORTS = ("rodv_bam_open", "rodv_bam_load_index", "rodv_bam_num_targets", "rodv_bam_target",
           "rodv_bam_header_text", "rodv_bam_next", "rodv_bam_query", "rodv_bam_query_region",
           "rodv_bam_query_multi", "rodv_bam_create", "rodv_bam_write", "rodv_bam_close",
           "rodv_bam_aux_z", "rodv_bam_read_inputs")


class BamError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {_NAMES.get(code, code)} ({code})")
        self.code = code


class CRecord(C.Structure):
    _fields_ = [("tid", C.c_int32), ("pos", C.c_int32), ("end", C.c_int32), ("mapq", C.c_uint32),
                ("flag", C.c_uint32), ("next_tid", C.c_int32), ("next_pos", C.c_int32),
                ("tlen", C.c_int32), ("l_seq", C.c_uint32), ("n_cigar", C.c_uint32),
                ("l_aux", C.c_uint32), ("name", C.c_char_p), ("cigar", C.POINTER(C.c_uint32)),
                ("seq", C.POINTER(C.c_uint8)), ("qual", C.POINTER(C.c_uint8)),
                ("aux", C.POINTER(C.c_uint8))]


@dataclasses.dataclass
class BamRecord:
    name: str
    tid: int = -1
    pos: int = -1
    mapq: int = 255
    flag: int = 0
    cigar: list = dataclasses.field(default_factory=list)  # [(op letter, length)]
    seq: bytes = b""
    qual: bytes = b""   # Phred values; b"\xff" * len(seq) when absent
    next_tid: int = -1
    next_pos: int = -1
    tlen: int = 0
    aux: bytes = b""    # raw auxiliary fields
    end: int = -1       # reference end (exclusive), filled on reading

    @staticmethod
    def aux_z(tag: str, value: str) -> bytes:
        """A Z-type auxiliary field, ready to concatenate into `aux`."""
        return tag.encode() + b"Z" + value.encode() + b"\0"


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is not built (python -m rodv_amd.build)")
        L = C.CDLL(path)
        vp, cp, i32, i64 = C.c_void_p, C.c_char_p, C.c_int32, C.c_int64
        L.rodv_bam_open.argtypes = [cp, C.POINTER(vp)]
        L.rodv_bam_load_index.argtypes = [vp, cp]
        L.rodv_bam_num_targets.argtypes = [vp]
        L.rodv_bam_target.argtypes = [vp, i32, C.POINTER(cp), C.POINTER(i64)]
        L.rodv_bam_header_text.argtypes = [vp]
        L.rodv_bam_header_text.restype = cp
        L.rodv_bam_next.argtypes = [vp, C.POINTER(CRecord)]
        L.rodv_bam_query.argtypes = [vp, i32, i64, i64]
        L.rodv_bam_query_region.argtypes = [vp, cp]
        L.rodv_bam_query_multi.argtypes = [vp, C.POINTER(cp), C.c_int]
        L.rodv_bam_create.argtypes = [cp, cp, i32, C.POINTER(cp), C.POINTER(i64), C.c_int, C.POINTER(vp)]
        L.rodv_bam_write.argtypes = [vp, C.POINTER(CRecord)]
        L.rodv_bam_close.argtypes = [vp]
        L.rodv_bam_aux_z.argtypes = [C.POINTER(CRecord), cp]
        L.rodv_bam_aux_z.restype = C.c_void_p
        u8p = C.POINTER(C.c_uint8)
        L.rodv_bam_read_inputs.argtypes = [C.POINTER(CRecord), C.c_int, C.c_uint8, C.c_uint8, C.c_uint8,
                                           C.c_uint8, u8p, u8p, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def _check(code, what):
    if code < 0:
        raise BamError(code, what)
    return code


def _from_c(r: CRecord) -> BamRecord:
    n, k = r.l_seq, r.n_cigar
    cig = [(CIGAR_OPS[r.cigar[i] & 0xf] if (r.cigar[i] & 0xf) < 9 else "?", r.cigar[i] >> 4) for i in range(k)]
    return BamRecord(name=r.name.decode(), tid=r.tid, pos=r.pos, mapq=r.mapq, flag=r.flag, cigar=cig,
                     seq=C.string_at(r.seq, n), qual=C.string_at(r.qual, n), next_tid=r.next_tid,
                     next_pos=r.next_pos, tlen=r.tlen, aux=C.string_at(r.aux, r.l_aux) if r.l_aux else b"",
                     end=r.end)


class _Keep:
    """Buffers a CRecord points into, kept alive for the call."""

    def __init__(self, rec: BamRecord):
        self.name = rec.name.encode()
        self.cigar = (C.c_uint32 * max(1, len(rec.cigar)))(
            *[(ln << 4) | CIGAR_OPS.index(op) for op, ln in rec.cigar])
        self.seq = (C.c_uint8 * max(1, len(rec.seq))).from_buffer_copy(rec.seq or b"\0")
        qual = rec.qual if rec.qual else b"\xff" * len(rec.seq)
        self.qual = (C.c_uint8 * max(1, len(qual))).from_buffer_copy(qual or b"\0")
        self.aux = (C.c_uint8 * max(1, len(rec.aux))).from_buffer_copy(rec.aux or b"\0")
        self.c = CRecord(tid=rec.tid, pos=rec.pos, end=-1, mapq=rec.mapq, flag=rec.flag,
                         next_tid=rec.next_tid, next_pos=rec.next_pos, tlen=rec.tlen, l_seq=len(rec.seq),
                         n_cigar=len(rec.cigar), l_aux=len(rec.aux), name=self.name,
                         cigar=self.cigar, seq=self.seq, qual=self.qual, aux=self.aux)


class BamReader:
    def __init__(self, path: str):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.rodv_bam_open(path.encode(), C.byref(h)), f"open {path}")
        self._h = h
        self._rec = CRecord()
        self._pending = None  # a record read by has_next() and not yet returned
        n = self._lib.rodv_bam_num_targets(h)
        self.targets = []
        for t in range(n):
            nm, ln = C.c_char_p(), C.c_int64()
            _check(self._lib.rodv_bam_target(h, t, C.byref(nm), C.byref(ln)), "target")
            self.targets.append((nm.value.decode(), ln.value))
        self.header_text = self._lib.rodv_bam_header_text(h).decode()

    def load_index(self, bai_path: str | None = None):
        _check(self._lib.rodv_bam_load_index(self._h, bai_path.encode() if bai_path else None),
               "load_index")

    def _read(self):
        r = _check(self._lib.rodv_bam_next(self._h, C.byref(self._rec)), "next")
        return _from_c(self._rec) if r == 1 else None

    def next(self) -> BamRecord | None:
        """The next record of the file or the current query; None at the end."""
        if self._pending is not None:
            rec, self._pending = self._pending, None
            return rec
        return self._read()

    def has_next(self) -> bool:
        if self._pending is None:
            self._pending = self._read()
        return self._pending is not None

    def __iter__(self):
        while True:
            rec = self.next()
            if rec is None:
                return
            yield rec

    def query(self, region=None, tid: int | None = None, beg: int = 0, end: int | None = None):
        """query_from_region: a region string ("chr", "chr:beg", "chr:beg-end",
        1-based inclusive) or tid with a 0-based half-open interval."""
        self._pending = None
        if region is not None:
            _check(self._lib.rodv_bam_query_region(self._h, region.encode()), f"query {region}")
        else:
            e = end if end is not None else self.targets[tid][1]
            _check(self._lib.rodv_bam_query(self._h, tid, beg, e), f"query {tid}:{beg}-{e}")
        return self

    def query_multi(self, regions):
        """query_from_multi_region: records overlapping any of the regions, once each."""
        self._pending = None
        arr = (C.c_char_p * max(1, len(regions)))(*[r.encode() for r in regions])
        _check(self._lib.rodv_bam_query_multi(self._h, arr, len(regions)), "query_multi")
        return self

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rodv_bam_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BamWriter:
    def __init__(self, path: str, targets, header_text: str = "", level: int = -1):
        self._lib = load_library()
        names = (C.c_char_p * max(1, len(targets)))(*[t[0].encode() for t in targets])
        lens = (C.c_int64 * max(1, len(targets)))(*[int(t[1]) for t in targets])
        h = C.c_void_p()
        _check(self._lib.rodv_bam_create(path.encode(), header_text.encode(), len(targets), names, lens,
                                         level, C.byref(h)), f"create {path}")
        self._h = h

    def write(self, rec: BamRecord):
        k = _Keep(rec)
        _check(self._lib.rodv_bam_write(self._h, C.byref(k.c)), "write")

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(self._lib.rodv_bam_close(h), "close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_inputs(rec: BamRecord, clip_soft: bool = True, default_q: int = 20, default_ins: int = 45,
                default_del: int = 45, default_gcp: int = 10):
    """PairHMM read inputs of a record (rodv_bam_read_inputs): bases (ACGT,
    others N), base qualities, insertion / deletion qualities from BI / BD or
    the defaults, gap-continuation qualities; soft clips dropped if asked."""
    lib = load_library()
    k = _Keep(rec)
    n = len(rec.seq)
    bufs = [np.zeros(max(1, n), np.uint8) for _ in range(5)]
    ln = C.c_uint32()
    ptrs = [b.ctypes.data_as(C.POINTER(C.c_uint8)) for b in bufs]
    _check(lib.rodv_bam_read_inputs(C.byref(k.c), int(clip_soft), default_q, default_ins, default_del,
                                    default_gcp, *ptrs, C.byref(ln)), "read_inputs")
    m = ln.value
    return tuple(b[:m].copy() for b in bufs)


def region_batch(reader: BamReader, region: str, haplotypes, skip_flags: int = SKIP_FLAGS,
                 min_mapq: int = 0, **inputs) -> tuple[Batch, list]:
    """The reads of `region` (a query of `reader`) that pass the flag and
    mapping-quality filters, as one region against `haplotypes` (bytes or
    str): a Batch whose pairs are haplotype-major, and the records used."""
    reads, q, ins, dl, gcp, used = [], [], [], [], [], []
    for rec in reader.query(region):
        if rec.flag & skip_flags or rec.mapq < min_mapq:
            continue
        b, bq, bi, bd, bg = read_inputs(rec, **inputs)
        if len(b) == 0:
            continue
        reads.append(b.tobytes())
        q.append(bq)
        ins.append(bi)
        dl.append(bd)
        gcp.append(bg)
        used.append(rec)
    if not reads:
        raise BamError(ERR_REGION, f"no usable reads in {region}")
    return Batch.from_lists(list(haplotypes), reads, q, ins, dl, gcp), used
