"""Chained items (reads of a region swept across up to 16 of their
haplotypes in one item, fwd_chain_kernel, short reads two per lane group):
bit-identical to one pair per slot, and to the oracle, across the
hand-over's and the two-read layout's edge cases.  The tests force
chains (haplotype_chains=2); the default forms them only in buckets large
enough to keep the device full, which the last test checks."""
import ctypes

import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch
from rodv_amd.engine import TUNE_CLASSES, Tuning, default_config

from test_forward_gpu import check_against_oracle
from test_schedule_gpu import check_schedule

pytestmark = pytest.mark.gpu


def _run(engine, batch, chains, stacked=False, **cfg):
    plan = engine.plan(batch, default_config(haplotype_chains=chains, **cfg))
    plan.execute()
    got = plan.read(raw=True)
    st = plan.stats()
    plan.close()
    return got, (st["stacked_pairs"] if stacked else st["num_buckets"])


def _same(engine, batch, expect_chains=True, mode=2, **cfg):
    on, nb_on = _run(engine, batch, mode, **cfg)
    off, nb_off = _run(engine, batch, 0, **cfg)
    if expect_chains:
        assert nb_on > nb_off, "no chained items formed"
    np.testing.assert_array_equal(on[2].view(np.uint32), off[2].view(np.uint32))
    np.testing.assert_array_equal(on[1], off[1])
    np.testing.assert_array_equal(on[0], off[0])


def _edge_batch(rng, L=16):
    # haplotypes at the hand-over bound (L + 2) and just below it (L + 1,
    # which leaves its reads unchained), reads whose length fills the lanes
    # exactly (the last lane carries the row-0 boundary) or leaves an
    # identity slot in lane 0, G reads and G + 1, unrelated reads (FP64
    # rescue inside a chain), nine reads of one length (one chain of eight)
    regs = [
        synth.region(rng, [L + 2, L + 2, 300], [16 * 9, 16 * 9 - 1, 100, 100, 150]),
        synth.region(rng, [L + 1, 200, 250], [120, 120, 120, 120]),
        synth.region(rng, [400, 390, 380, 60, 50], [240, 241, 250, 200, 200, 233, 90, 91, 92],
                     unrelated=[0, 4, 6]),
        synth.region(rng, [35] * 7, [5, 6, 7, 8, 30, 31, 32, 33]),
        synth.region(rng, [500, 450], list(rng.integers(50, 250, 23))),
        synth.region(rng, [120, 110, 100], [70] * 9),
    ]
    return Batch.concat(regs)


@pytest.mark.parametrize("semantics", [0, 1])
def test_chains_equal_single_pairs(engine, semantics):
    rng = np.random.default_rng(41)
    _same(engine, _edge_batch(rng), semantics=semantics)
    _same(engine, synth.config(2, scale=0.003), semantics=semantics)
    _same(engine, synth.config(1), semantics=semantics)


def test_chains_against_oracle(engine, oracle_mod):
    rng = np.random.default_rng(42)
    check_against_oracle(oracle_mod, engine, _edge_batch(rng), haplotype_chains=2)
    check_schedule(engine, oracle_mod, _edge_batch(np.random.default_rng(43)), haplotype_chains=2)


def _stack_batch(rng):
    # short reads two per 16-lane group: equal and unequal lengths in one
    # group, the tightest fits (two reads of 127 rows in 16 rows per lane, of
    # 63 in 8), lengths across the one-read shapes' boundaries, unrelated
    # reads (FP64 rescue) and an N in a stacked read, groups left over
    regs = [
        synth.region(rng, [300, 280, 260], [127] * 8),
        synth.region(rng, [200, 190], [63] * 8 + [64] * 8),
        synth.region(rng, [150, 140, 130, 120], list(rng.integers(20, 128, 16))),
        synth.region(rng, [90, 80, 70], [50] * 12, unrelated=[1, 5]),
        synth.region(rng, [400, 300], [100] * 9, n_read=[3]),
        synth.region(rng, [260, 250], [96] * 8 + [31] * 8),
    ]
    return Batch.concat(regs)


@pytest.mark.parametrize("semantics", [0, 1])
def test_stacked_reads(engine, oracle_mod, semantics):
    b = _stack_batch(np.random.default_rng(61 + semantics))
    _, stacked = _run(engine, b, 2, stacked=True, semantics=semantics)
    _, none = _run(engine, b, 4, stacked=True, semantics=semantics)
    assert stacked > 0 and none == 0
    _same(engine, b, semantics=semantics)          # two reads per group where it pays
    _same(engine, b, mode=4, semantics=semantics)  # one read per group
    if semantics == 0:
        check_against_oracle(oracle_mod, engine, b, haplotype_chains=2)
    else:
        check_schedule(engine, oracle_mod, b, haplotype_chains=2)


@pytest.mark.parametrize("ppw", [8, 4, 2, 1])
def test_chains_every_group_width(engine, oracle_mod, ppw):
    # 8, 4, 2 and 1 pairs per wave (8-, 16-, 32- and 64-lane chains)
    t = Tuning()
    t.struct_size = ctypes.sizeof(Tuning)
    t.num_classes = TUNE_CLASSES
    for k in range(TUNE_CLASSES):
        t.cls[k].pairs_per_wave = ppw
    engine.set_tuning(t)
    try:
        rng = np.random.default_rng(50 + ppw)
        L = 64 // ppw
        b = _edge_batch(rng, L=L)
        # per-pair semantics keep each region's pairs together, so the nine
        # equal reads always chain; schedule routing may split them
        _same(engine, b, semantics=0)
        _same(engine, b, expect_chains=False, semantics=1)
        check_against_oracle(oracle_mod, engine, b, haplotype_chains=2)
        check_schedule(engine, oracle_mod, b, haplotype_chains=2)
    finally:
        engine.set_tuning(None)


def test_default_chains_large_batch(engine):
    # the uneven batched workload at 30 %: large enough for chained items in
    # the default mode, which must then give the one-pair-per-slot bits
    b = synth.config(2, scale=0.3)
    _same(engine, b, mode=1, semantics=1)
    _, stacked = _run(engine, b, 1, stacked=True, semantics=1)
    assert stacked > 0
    # a small batch keeps one pair per slot by default
    _, nb1 = _run(engine, synth.config(1), 1, semantics=1)
    _, nb0 = _run(engine, synth.config(1), 0, semantics=1)
    assert nb1 == nb0


def test_chain_mode_out_of_range_is_rejected(engine):
    from rodv_amd.engine import ERR_INVALID_ARG, PhmmError
    b = synth.random_batch(np.random.default_rng(3), 2, 5)
    for bad in (-1, 5):
        with pytest.raises(PhmmError) as e:
            engine.plan(b, default_config(haplotype_chains=bad))
        assert e.value.code == ERR_INVALID_ARG


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_chains_random_multi_region_batches(engine, seed):
    # 300 regions of random shapes (equal or mixed read lengths per region,
    # haplotypes of 1..999 columns, unrelated reads, N bases), about 260,000
    # pairs: forced chains give the one-pair-per-slot bits
    rng = np.random.default_rng(4000 + seed)
    regs = []
    for _ in range(300):
        nh, nr = int(rng.integers(1, 40)), int(rng.integers(1, 90))
        R = int(rng.integers(1, 300))
        rl = [R] * nr if rng.random() < 0.7 else list(rng.integers(1, 300, nr))
        hl = list(rng.integers(1, 1000, nh))
        regs.append(synth.region(rng, hl, rl,
                                 unrelated=[i for i in range(nr) if rng.random() < 0.05],
                                 n_read=[i for i in range(nr) if rng.random() < 0.02],
                                 n_hap=[i for i in range(nh) if rng.random() < 0.02],
                                 qual=int(rng.integers(5, 41))))
    _same(engine, Batch.concat(regs), semantics=seed % 2)
