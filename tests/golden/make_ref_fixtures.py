#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own
kernels, compiled from its sources into oracle/_ref/ (make -C oracle ref).

Fixtures are data only: inputs (flat arrays of a rodv_amd.batch.Batch) and
the reference's outputs for them.
  ref_tables.npz   Context<float>/Context<double> ph2pr and match-to-match tables
  ref_pairs.npz    inputs + intra raw FP32/FP64 sums (AVX-512 and AVX2 builds),
                   intra log10 (computeLikelihoods), inter raw sums (AVX2 group
                   kernel, groups of 8 / 4) for GKL vectors, random pairs with
                   N bases, unrelated reads and pairs near the FP32 rescue
                   threshold (raw sum within a few percent of 1e-28f)
Run from the repository root: python tests/golden/make_ref_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from rodv_amd import synth  # noqa: E402
from rodv_amd.batch import Batch, parse_gkl_testdata  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def flat_pairs(batches):
    """One region per pair: returns a Batch whose cross_pairs are the pairs."""
    haps, reads, qs, ins, dels, gcps = [], [], [], [], [], []
    for b in batches:
        ph, pr = b.cross_pairs()
        for h, r in zip(ph, pr):
            haps.append(b.hap(h))
            reads.append(b.read(r))
            q = b.quals(r)
            qs.append(q[0]); ins.append(q[1]); dels.append(q[2]); gcps.append(q[3])
    n = len(haps)
    regions = np.array([[k, 1, k, 1] for k in range(n)], np.uint32)
    return Batch.from_lists(haps, reads, qs, ins, dels, gcps, regions)


def near_threshold(rng, ref, want=24, tries=4000):
    """Unrelated reads of length ~55-90 whose FP32 raw sum lands near 1e-28."""
    picked = []
    for _ in range(tries):
        b = synth.random_batch(rng, 1, 1, hap_len=(60, 160), read_len=(50, 95), p_sub=0.75,
                               qual=(8, 40))
        r = ref.raw_f32(b, 0, 0)
        if r > 0 and abs(np.log10(r) + 28.0) < 0.15:
            picked.append(b)
            if len(picked) >= want:
                break
    return picked


def main():
    ref = O.RefLib("avx512")
    ref2 = O.RefLib("avx2")
    ph, mm = ref.tables_f32()
    phd, mmd = ref.tables_f64()
    np.savez_compressed(os.path.join(OUT, "ref_tables.npz"), ph2pr_f=ph, mm_f=mm, ph2pr_d=phd,
                        mm_d=mmd)
    rng = np.random.default_rng(20261203)
    gkl, _ = parse_gkl_testdata(os.path.join(OUT, "pairhmm-testdata.txt"))
    parts = [gkl]
    parts.append(synth.random_batch(rng, 4, 10, hap_len=(1, 60), read_len=(1, 40)))
    parts.append(synth.random_batch(rng, 4, 10, hap_len=(100, 400), read_len=(100, 250)))
    parts.append(synth.random_batch(rng, 3, 8, n_frac=0.05))
    parts.append(synth.random_batch(rng, 3, 10, hap_len=(150, 300), read_len=(40, 200), p_sub=0.75))
    parts.extend(near_threshold(rng, ref))
    b = flat_pairs(parts)
    n = b.num_pairs()
    ph_, pr_ = b.cross_pairs()
    raw32 = np.array([ref.raw_f32(b, h, r) for h, r in zip(ph_, pr_)], np.float32)
    raw32_avx2 = np.array([ref2.raw_f32(b, h, r) for h, r in zip(ph_, pr_)], np.float32)
    raw64 = np.array([ref.raw_f64(b, h, r) for h, r in zip(ph_, pr_)], np.float64)
    lg = np.array([ref.log10(b, h, r) for h, r in zip(ph_, pr_)], np.float64)
    # inter groups over N-free pairs (the scheduler never sends N to inter)
    nfree = [k for k in range(n) if b"N" not in b.read(pr_[k]) and b"N" not in b.hap(ph_[k])]
    g8 = nfree[: (len(nfree) // 8) * 8]
    inter32 = np.zeros(len(g8), np.float64)
    for s in range(0, len(g8), 8):
        ks = g8[s:s + 8]
        inter32[s:s + 8] = ref2.inter_raw(b, ph_[ks], pr_[ks])
    g4 = nfree[: (len(nfree) // 4) * 4]
    inter64 = np.zeros(len(g4), np.float64)
    for s in range(0, len(g4), 4):
        ks = g4[s:s + 4]
        inter64[s:s + 4] = ref2.inter_raw(b, ph_[ks], pr_[ks], double=True)
    np.savez_compressed(
        os.path.join(OUT, "ref_pairs.npz"),
        hap_bases=b.hap_bases, hap_off=b.hap_off, hap_len=b.hap_len, read_bases=b.read_bases,
        base_q=b.base_q, ins_q=b.ins_q, del_q=b.del_q, gcp_q=b.gcp_q, read_off=b.read_off,
        read_len=b.read_len, regions=b.regions, raw32=raw32, raw32_avx2=raw32_avx2,
        raw64=raw64, log10=lg, inter_idx8=np.array(g8, np.int64), inter32=inter32,
        inter_idx4=np.array(g4, np.int64), inter64=inter64)
    print(f"wrote {n} pairs ({int((raw32 < np.float32(1e-28)).sum())} below the rescue threshold)")


if __name__ == "__main__":
    main()
