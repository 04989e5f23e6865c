"""Deterministic synthetic workloads for the PairHMM engine.

Shapes follow the five benchmark configurations of this path (read length,
haplotype length, mutation rates, quality distributions), interpreted as
follows.  Each region draws one uniform ACGT sequence; its
haplotypes are prefixes of it (of their own lengths) carrying independent
0.5% substitutions, as candidate haplotypes of one locus are (the reference's
own generator builds mutation variants of one base sequence,
pairhmm_schedule_unittest.cpp:39-200).  A read is a substring of a uniformly
chosen haplotype of its region, inside the prefix all its haplotypes share,
with substitutions (to a different base) and indel events of length U{1..3}
(50/50 insertion/deletion), trimmed or padded back to its target length.
Insertion/deletion qualities are 45 and gap-continuation quality 10 unless
stated.  Seeds are 1000 + config.
"""
from __future__ import annotations

import numpy as np

from .batch import Batch

ACGT = np.frombuffer(b"ACGT", np.uint8)


def _offsets(lens):
    off = np.zeros(len(lens), np.uint64)
    if len(lens) > 1:
        off[1:] = np.cumsum(lens[:-1].astype(np.uint64))
    return off


def _ranges(starts, lens):
    """Concatenation of arange(starts[k], starts[k] + lens[k])."""
    lens = lens.astype(np.int64)
    tot = int(lens.sum())
    if tot == 0:
        return np.zeros(0, np.int64)
    rep = np.repeat(starts.astype(np.int64) - np.concatenate(([0], np.cumsum(lens)[:-1])), lens)
    return rep + np.arange(tot, dtype=np.int64)


def _mutate_indels(rng, read, hap_tail, p_indel):
    """Apply indel events to one read (uint8 codes 0..3); keep its length by
    trimming or by extending from the haplotype continuation."""
    R = len(read)
    n_ev = rng.binomial(R, p_indel)
    if n_ev == 0:
        return read
    out = list(read)
    for _ in range(n_ev):
        pos = int(rng.integers(0, max(1, len(out))))
        ln = int(rng.integers(1, 4))
        if rng.random() < 0.5:
            out[pos:pos] = list(rng.integers(0, 4, ln))
        else:
            del out[pos:pos + ln]
    if len(out) >= R:
        return np.array(out[:R], np.uint8)
    pad = list(hap_tail[: R - len(out)])
    if len(out) + len(pad) < R:
        pad += list(rng.integers(0, 4, R - len(out) - len(pad)))
    return np.array(out + pad, np.uint8)


def make_workload(hap_len_per_hap, hap_region, read_len_per_read, read_region, region_haps,
                  region_reads, *, p_sub, p_indel, qual_fn, rng, p_unrelated=0.0,
                  ins_q=45, del_q=45, gcp_q=10, p_hapvar=0.005):
    """Generic generator.  region_haps/region_reads: (begin, count) per region."""
    H = np.asarray(hap_len_per_hap, np.int64)
    hoff = _offsets(H).astype(np.int64)
    # one base sequence per region; haplotypes are variant prefixes of it
    n_reg = len(region_haps)
    reg_max = np.zeros(n_reg, np.int64)
    reg_min = np.full(n_reg, np.iinfo(np.int64).max)
    np.maximum.at(reg_max, hap_region, H)
    np.minimum.at(reg_min, hap_region, H)
    base_off = _offsets(reg_max).astype(np.int64)
    base = rng.integers(0, 4, int(reg_max.sum()), dtype=np.uint8)
    hap_codes = base[_ranges(base_off[hap_region], H)]
    var = rng.random(len(hap_codes)) < p_hapvar
    hap_codes[var] = (hap_codes[var] + rng.integers(1, 4, int(var.sum()), dtype=np.uint8)) % 4
    R = np.asarray(read_len_per_read, np.int64)
    nr = len(R)
    # pick a source hap within the read's region, and a window inside the
    # prefix every haplotype of the region has
    rh0 = region_haps[read_region, 0]
    rhn = region_haps[read_region, 1]
    src = rh0 + (rng.random(nr) * rhn).astype(np.int64)
    Hs = H[src]
    span = np.minimum(reg_min[read_region], Hs)
    start = (rng.random(nr) * np.maximum(span - R + 1, 1)).astype(np.int64)
    take = np.minimum(R, Hs - start)
    idx = _ranges(hoff[src] + start, take)
    roff = _offsets(R).astype(np.int64)
    read_codes = rng.integers(0, 4, int(R.sum()), dtype=np.uint8)  # fills any short tail
    dst = _ranges(roff, take)
    read_codes[dst] = hap_codes[idx]
    # substitutions to a different base
    sub = rng.random(len(read_codes)) < p_sub
    read_codes[sub] = (read_codes[sub] + rng.integers(1, 4, int(sub.sum()), dtype=np.uint8)) % 4
    # indels (rare): per read
    if p_indel > 0:
        exp = rng.random(nr) < 1.0 - (1.0 - p_indel) ** R
        for k in np.nonzero(exp)[0]:
            o, n = int(roff[k]), int(R[k])
            tail_start = int(hoff[src[k]] + start[k] + take[k])
            tail = hap_codes[tail_start:int(hoff[src[k]] + Hs[k])]
            read_codes[o:o + n] = _mutate_indels(rng, read_codes[o:o + n].copy(), tail, p_indel)
    # unrelated reads
    if p_unrelated > 0:
        unrel = np.nonzero(rng.random(nr) < p_unrelated)[0]
        for k in unrel:
            o, n = int(roff[k]), int(R[k])
            read_codes[o:o + n] = rng.integers(0, 4, n, dtype=np.uint8)
    q = qual_fn(rng, int(R.sum())).astype(np.uint8)
    tot = int(R.sum())
    b = Batch(hap_bases=ACGT[hap_codes], hap_off=_offsets(H), hap_len=H.astype(np.uint32),
              read_bases=ACGT[read_codes], base_q=q,
              ins_q=np.full(tot, ins_q, np.uint8), del_q=np.full(tot, del_q, np.uint8),
              gcp_q=np.full(tot, gcp_q, np.uint8), read_off=_offsets(R),
              read_len=R.astype(np.uint32),
              regions=np.stack([region_haps[:, 0], region_haps[:, 1], region_reads[:, 0],
                                region_reads[:, 1]], axis=1).astype(np.uint32))
    return b


def _uniform_q(lo, hi):
    return lambda rng, n: rng.integers(lo, hi + 1, n)


def _peaked_q(rng, n):
    # 95% U{30..40}, 5% tail U{2..10}
    q = rng.integers(30, 41, n)
    tail = rng.random(n) < 0.05
    q[tail] = rng.integers(2, 11, int(tail.sum()))
    return q


def _single_region(n_haps, hap_lens, n_reads, read_lens):
    return (np.zeros(n_haps, np.int64), np.zeros(n_reads, np.int64),
            np.array([[0, n_haps]], np.int64), np.array([[0, n_reads]], np.int64))


def config(cfg: int, scale: float = 1.0) -> Batch:
    """Synthetic batch of configuration `cfg` (0..4).  `scale` < 1 shrinks the
    number of regions (config 2) or pairs (configs 3, 4) for tests."""
    rng = np.random.default_rng(1000 + cfg)
    if cfg == 0:
        H = np.full(8, 150)
        R = np.full(64, 100)
        hr, rr, rh, rd = _single_region(8, H, 64, R)
        return make_workload(H, hr, R, rr, rh, rd, p_sub=0.01, p_indel=0.001,
                             qual_fn=_uniform_q(10, 40), rng=rng)
    if cfg == 1:
        H = rng.integers(300, 501, 32)
        R = np.full(128, 150)
        hr, rr, rh, rd = _single_region(32, H, 128, R)
        return make_workload(H, hr, R, rr, rh, rd, p_sub=0.02, p_indel=0.001,
                             qual_fn=_peaked_q, rng=rng)
    if cfg == 2:
        n_reg = max(1, int(round(10000 * scale)))
        nreads = np.clip(rng.geometric(1.0 / 80.0, n_reg), 1, 500)
        nhaps = rng.integers(2, 65, n_reg)
        rlen = rng.integers(50, 251, n_reg)
        hap_region = np.repeat(np.arange(n_reg), nhaps)
        H = rlen[hap_region] + rng.integers(50, 401, len(hap_region))
        read_region = np.repeat(np.arange(n_reg), nreads)
        R = rlen[read_region]
        rh = np.stack([np.concatenate(([0], np.cumsum(nhaps)[:-1])), nhaps], 1)
        rd = np.stack([np.concatenate(([0], np.cumsum(nreads)[:-1])), nreads], 1)
        return make_workload(H, hap_region, R, read_region, rh, rd, p_sub=0.01, p_indel=0.001,
                             qual_fn=_uniform_q(10, 40), rng=rng, p_unrelated=0.005)
    if cfg == 3:
        nh, nr = 32, 128
        if scale < 1.0:
            nh = max(1, int(round(32 * scale)))
        H = np.full(nh, 400)
        R = np.full(nr, 150)
        hr, rr, rh, rd = _single_region(nh, H, nr, R)
        return make_workload(H, hr, R, rr, rh, rd, p_sub=0.01, p_indel=0.005,
                             qual_fn=_uniform_q(10, 40), rng=rng)
    if cfg == 4:
        n = max(1, int(round(256 * scale)))
        R = rng.integers(1000, 10001, n)
        H = np.minimum(R + rng.integers(200, 2001, n), 12000)
        idx = np.arange(n)
        rh = np.stack([idx, np.ones(n, np.int64)], 1)
        return make_workload(H, idx, R, idx, rh, rh.copy(), p_sub=0.05, p_indel=0.02,
                             qual_fn=_uniform_q(5, 30), rng=rng)
    raise ValueError(f"unknown configuration {cfg}")


def random_batch(rng, n_haps, n_reads, hap_len=(20, 200), read_len=(10, 150), p_sub=0.02,
                 p_indel=0.005, n_frac=0.0, qual=(2, 45), ins=(10, 50), gcp=(5, 20)):
    """Small single-region batch with varied lengths and qualities for parity
    tests.  n_frac > 0 sprinkles N bases into reads and haps."""
    H = rng.integers(hap_len[0], hap_len[1] + 1, n_haps)
    R = rng.integers(read_len[0], read_len[1] + 1, n_reads)
    hr, rr, rh, rd = _single_region(n_haps, H, n_reads, R)
    b = make_workload(H, hr, R, rr, rh, rd, p_sub=p_sub, p_indel=p_indel,
                      qual_fn=_uniform_q(*qual), rng=rng)
    tot = int(R.sum())
    b.ins_q = rng.integers(ins[0], ins[1] + 1, tot).astype(np.uint8)
    b.del_q = rng.integers(ins[0], ins[1] + 1, tot).astype(np.uint8)
    b.gcp_q = rng.integers(gcp[0], gcp[1] + 1, tot).astype(np.uint8)
    if n_frac > 0:
        b.read_bases = b.read_bases.copy()
        b.read_bases[rng.random(tot) < n_frac] = ord("N")
        b.hap_bases = b.hap_bases.copy()
        b.hap_bases[rng.random(len(b.hap_bases)) < n_frac] = ord("N")
    return b


def region(rng, hap_lens, read_lens, unrelated=(), n_read=(), n_hap=(), qual=30):
    """One region (len(hap_lens) haplotypes x len(read_lens) reads) with the
    exact lengths given.  Haplotypes are prefixes of one random sequence;
    read k is a substring of the shortest of them (padded with random bases
    where it is longer), or random bases when k is in `unrelated`; reads in
    `n_read` and haplotypes in `n_hap` get one N.  Qualities: base `qual`,
    insertion/deletion 45, gap continuation 10."""
    base = np.frombuffer(b"ACGT", np.uint8)
    seq = base[rng.integers(0, 4, int(max(hap_lens)))]
    haps = [seq[:int(h)].copy() for h in hap_lens]
    hmin = int(min(hap_lens))
    reads = []
    for k, r in enumerate(read_lens):
        r = int(r)
        if k in unrelated:
            rd = base[rng.integers(0, 4, r)]
        else:
            s = int(rng.integers(0, max(1, hmin - r + 1)))
            rd = seq[s:min(s + r, hmin)]
            if len(rd) < r:
                rd = np.concatenate([rd, base[rng.integers(0, 4, r - len(rd))]])
        reads.append(rd.copy())
    for k in n_read:
        reads[k][len(reads[k]) // 2] = ord("N")
    for k in n_hap:
        haps[k][len(haps[k]) // 2] = ord("N")
    q = [np.full(len(r), qual, np.uint8) for r in reads]
    i = [np.full(len(r), 45, np.uint8) for r in reads]
    c = [np.full(len(r), 10, np.uint8) for r in reads]
    return Batch.from_lists([h.tobytes() for h in haps], [r.tobytes() for r in reads], q, i, i, c)
