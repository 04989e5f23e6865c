"""Host-side batch container for the PairHMM engine.

A batch holds M haplotypes and N reads as concatenated byte arrays with
offsets and lengths, exactly the flat layout of the C-ABI struct
``rodv_phmm_batch`` (include/rodv_phmm.h).  Each read carries the four
per-base Phred arrays of the reference's TestCase (pairhmm/common/common.h:27-36):
base quality, insertion quality, deletion quality and gap-continuation
quality, all of read length, raw Phred integers (not +33).

Regions (a block of haps x a block of reads, scored as a full cross
product) follow schedule_pairhmm's call shape (pairhmm_schedule.h:14-25): one
call per region.  A batch may hold many regions; results are laid out
region by region, hap-major inside a region (result[hap][read],
pairhmm_schedule.cpp:135-138).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import re

import numpy as np

VALID_BASES = frozenset(b"ACGTN")


def _concat(seqs):
    lens = np.array([len(s) for s in seqs], dtype=np.uint32)
    offs = np.zeros(len(seqs), dtype=np.uint64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    data = np.frombuffer(b"".join(bytes(s) for s in seqs), dtype=np.uint8).copy() \
        if len(seqs) else np.zeros(0, np.uint8)
    return data, offs, lens


@dataclass
class Batch:
    hap_bases: np.ndarray
    hap_off: np.ndarray
    hap_len: np.ndarray
    read_bases: np.ndarray
    base_q: np.ndarray
    ins_q: np.ndarray
    del_q: np.ndarray
    gcp_q: np.ndarray
    read_off: np.ndarray
    read_len: np.ndarray
    # regions: rows of (hap_begin, num_haps, read_begin, num_reads)
    regions: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.uint32))

    # ------------------------------------------------------------ builders
    @classmethod
    def from_lists(cls, haps, reads, quals, ins, dels, gcps, regions=None):
        """haps/reads: sequences of bytes/str; quality lists: sequences of int
        arrays (one per read).  Without `regions` the batch is one region."""
        def tob(s):
            return s.encode() if isinstance(s, str) else bytes(s)
        hb, ho, hl = _concat([tob(h) for h in haps])
        rb, ro, rl = _concat([tob(r) for r in reads])

        def qcat(qs):
            d, _, ln = _concat([np.asarray(q, np.uint8).tobytes() for q in qs])
            if not np.array_equal(ln, rl):
                raise ValueError("quality length differs from read length")
            return d
        if regions is None:
            regions = np.array([[0, len(haps), 0, len(reads)]], np.uint32)
        return cls(hb, ho, hl, rb, qcat(quals), qcat(ins), qcat(dels), qcat(gcps), ro, rl,
                   np.asarray(regions, np.uint32).reshape(-1, 4))

    @classmethod
    def concat(cls, batches):
        """Stack several batches (each with its own regions) into one."""
        haps, reads, quals, ins, dels, gcps, regs = [], [], [], [], [], [], []
        nh = nr = 0
        for b in batches:
            for k in range(len(b.hap_len)):
                haps.append(b.hap(k))
            for k in range(len(b.read_len)):
                reads.append(b.read(k))
                quals.append(b.quals(k)[0])
                ins.append(b.quals(k)[1])
                dels.append(b.quals(k)[2])
                gcps.append(b.quals(k)[3])
            for (h0, m, r0, n) in b.regions:
                regs.append((h0 + nh, m, r0 + nr, n))
            nh += len(b.hap_len)
            nr += len(b.read_len)
        return cls.from_lists(haps, reads, quals, ins, dels, gcps, np.array(regs, np.uint32))

    # ------------------------------------------------------------ accessors
    @property
    def num_haps(self):
        return int(len(self.hap_len))

    @property
    def num_reads(self):
        return int(len(self.read_len))

    def hap(self, k):
        o = int(self.hap_off[k])
        return self.hap_bases[o:o + int(self.hap_len[k])].tobytes()

    def read(self, k):
        o = int(self.read_off[k])
        return self.read_bases[o:o + int(self.read_len[k])].tobytes()

    def quals(self, k):
        o, n = int(self.read_off[k]), int(self.read_len[k])
        return tuple(a[o:o + n].copy() for a in (self.base_q, self.ins_q, self.del_q, self.gcp_q))

    def cross_pairs(self):
        """(pair_hap, pair_read) for all regions, in output order."""
        hs, rs = [], []
        for (h0, m, r0, n) in self.regions:
            h = np.repeat(np.arange(h0, h0 + m, dtype=np.uint32), n)
            r = np.tile(np.arange(r0, r0 + n, dtype=np.uint32), m)
            hs.append(h)
            rs.append(r)
        if not hs:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        return np.concatenate(hs), np.concatenate(rs)

    def num_pairs(self):
        return int(sum(int(m) * int(n) for (_, m, _, n) in self.regions))

    def cells(self):
        """Sum over pairs of R*H (the DP interior)."""
        tot = 0
        for (h0, m, r0, n) in self.regions:
            tot += int(self.hap_len[h0:h0 + m].astype(np.int64).sum()) * \
                int(self.read_len[r0:r0 + n].astype(np.int64).sum())
        return tot

    def region_slices(self):
        """Offsets of each region's block in the flat output."""
        sizes = self.regions[:, 1].astype(np.int64) * self.regions[:, 3].astype(np.int64)
        offs = np.zeros(len(sizes) + 1, np.int64)
        offs[1:] = np.cumsum(sizes)
        return offs


def parse_gkl_testdata(path):
    """resouces/pairhmm-testdata.txt: one pair per line, qualities Phred+33
    (parsed as pairhmm_unittest.cpp:33-92).  Returns (batch, expected): the
    batch has one single-pair region per line."""
    haps, reads, qs, ins, dels, gcps, exp = [], [], [], [], [], [], []
    with open(path) as f:
        next(f)
        for line in f:
            parts = line.split()
            if len(parts) < 7:
                continue
            h, r, q, i, d, c, e = parts[:7]
            haps.append(h)
            reads.append(r)
            qs.append([ord(x) - 33 for x in q])
            ins.append([ord(x) - 33 for x in i])
            dels.append([ord(x) - 33 for x in d])
            gcps.append([ord(x) - 33 for x in c])
            exp.append(float(e))
    n = len(haps)
    regions = np.array([[k, 1, k, 1] for k in range(n)], np.uint32)
    return Batch.from_lists(haps, reads, qs, ins, dels, gcps, regions), np.array(exp)


def parse_assemble_log(path, with_names=False):
    """resouces/pairhmm_debug.txt region dump (test/assemble_log_parser.cpp:
    13-149): '=== Region: <name> ==' headers, 'H<k>: bases', 'R<k>: bases'
    followed by integer read-qual / read-ins-qual / read-del-qual / gcp lines.
    Returns one Batch per region (with with_names, (name, Batch) pairs)."""
    batches = []
    names = []
    haps, reads, qs, ins, dels, gcps = [], [], [], [], [], []
    cur = None
    name = [""]

    def flush():
        if haps and reads:
            batches.append(Batch.from_lists(list(haps), list(reads), list(qs), list(ins),
                                            list(dels), list(gcps)))
            names.append(name[0])
        for lst in (haps, reads, qs, ins, dels, gcps):
            lst.clear()

    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if line.startswith("=== Region"):
                flush()
                m = re.search(r"=== Region: (.+) ==", line)
                name[0] = m.group(1) if m else ""
                continue
            if not line or ":" not in line:
                continue
            key, val = line.split(":", 1)
            key, val = key.strip(), val.strip()
            if key.startswith("H") and key[1:].isdigit():
                haps.append(val)
            elif key.startswith("R") and key[1:].isdigit():
                reads.append(val)
                cur = len(reads) - 1
            elif key == "read-qual":
                qs.append([int(x) for x in val.split()])
            elif key == "read-ins-qual":
                ins.append([int(x) for x in val.split()])
            elif key == "read-del-qual":
                dels.append([int(x) for x in val.split()])
            elif key == "gcp":
                gcps.append([int(x) for x in val.split()])
    flush()
    del cur
    return list(zip(names, batches)) if with_names else batches
