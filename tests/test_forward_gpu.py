"""Parity of the HIP forward against the CPU oracle (which is itself pinned
against the reference's compiled kernels and golden vectors)."""
import os

import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch, parse_assemble_log, parse_gkl_testdata
from rodv_amd.engine import (ERR_EMPTY, ERR_INVALID_ARG, FLAG_EXTENDED, FLAG_RESCUED,
                             FLAG_UNDERFLOW, PhmmError, default_config)

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-4  # |d log10 L| bound of this path
FP64_STRIP = 63 * 7  # rows of an FP64 strip of the long-read rescue (7 rows per lane)
UNIT = 4 * FP64_STRIP  # rows of an FP64 unit of 4 strips
UNIT8 = 8 * FP64_STRIP  # with 8-wave units (RODV_PHMM_RESCUE_LONG_WAVES=8)


def run(engine, batch, **cfg):
    # per-pair intra semantics: the oracle's run_batch restates the intra API;
    # schedule_pairhmm routing has its own tests (test_schedule_gpu.py)
    cfg.setdefault("semantics", 0)
    plan = engine.plan(batch, default_config(**cfg))
    plan.execute()
    out, fl, r32, r64 = plan.read(raw=True)
    plan.close()
    return out, fl, r32, r64


def check_against_oracle(oracle_mod, engine, batch, raw_exact=True, **cfg):
    out, fl, r32, r64 = run(engine, batch, **cfg)
    o_out, o32, o64, ofl = oracle_mod.run_batch(batch, nthreads=8)
    # FP32 raw sums bit-exact (same rounding as the reference's intra kernel)
    if raw_exact:
        bad = np.nonzero(r32.view(np.uint32) != o32.view(np.uint32))[0]
        assert len(bad) == 0, f"{len(bad)} raw FP32 mismatches, first {bad[:5]}: {r32[bad[:5]]} vs {o32[bad[:5]]}"
    # rescue set identical
    np.testing.assert_array_equal(fl & FLAG_RESCUED, ofl & FLAG_RESCUED)
    # the FP64 sums of the rescued pairs bit-exact too (the extended-range
    # pass rescales the ones it scores)
    if raw_exact:
        resc = ((ofl & FLAG_RESCUED) != 0) & ((fl & FLAG_EXTENDED) == 0)
        bad = np.nonzero(resc & (r64.view(np.uint64) != o64.view(np.uint64)))[0]
        assert len(bad) == 0, f"{len(bad)} raw FP64 mismatches, first {bad[:5]}: {r64[bad[:5]]} vs {o64[bad[:5]]}"
    np.testing.assert_array_equal(fl & FLAG_UNDERFLOW, ofl & FLAG_UNDERFLOW)
    fin = np.isfinite(o_out)
    np.testing.assert_array_equal(np.isfinite(out), fin)
    assert np.all(out[~fin] == o_out[~fin])
    if fin.any():
        err = np.abs(out[fin] - o_out[fin])
        assert err.max() <= TOL, f"max |dlog10L| {err.max()}"
    return out, fl, o_out


def test_dpp_lane_layout_and_kats(engine, oracle_mod):
    # the two known answers of pairhmm_unittest.cpp:103-136
    q4 = [43] * 4
    q5 = [43] * 5
    b = Batch.from_lists(["ACGT", "ACGTA"], ["ACGT", "ACGTA"], [q4, q5], [q4, q5], [q4, q5],
                         [q4, q5], regions=np.array([[0, 1, 0, 1], [1, 1, 1, 1]]))
    out, fl, r32, _ = run(engine, b)
    assert abs(out[0] - (-6.022797e-01)) < 1e-5
    assert abs(out[1] - (-0.69925308227539062)) < 1e-5
    assert not (fl & FLAG_RESCUED).any()


def test_gkl_golden_vectors(engine, oracle_mod):
    b, exp = parse_gkl_testdata(os.path.join(GOLDEN, "pairhmm-testdata.txt"))
    out, fl, _ = check_against_oracle(oracle_mod, engine, b)
    # pairhmm_unittest.cpp:347-396: relative 1e-5 against the expected column
    assert np.max(np.abs(out - exp) / np.abs(exp)) < 1e-5


def test_assemble_log_regions(engine, oracle_mod):
    regs = parse_assemble_log(os.path.join(GOLDEN, "pairhmm_debug.txt"))
    b = Batch.concat(regs)
    check_against_oracle(oracle_mod, engine, b)


@pytest.mark.parametrize("table", [1, 0])
@pytest.mark.parametrize("lanes", [16, 64])
@pytest.mark.parametrize("seed,rlo,rhi", [(1, 1, 20), (2, 10, 120), (3, 100, 240), (4, 200, 400),
                                          (5, 230, 260)])
def test_random_lengths(engine, oracle_mod, seed, rlo, rhi, lanes, table):
    rng = np.random.default_rng(seed)
    b = synth.random_batch(rng, 7, 23, hap_len=(1, 420), read_len=(rlo, rhi))
    check_against_oracle(oracle_mod, engine, b, short_lanes=lanes, prior_table=table)


def test_every_rows_per_lane_shape(engine, oracle_mod):
    # one read length per (L=16, MR) shape and both ends of each bucket
    lens = sorted({max(1, 15 * m - d) for m in range(1, 17) for d in (0, 14)} | {241, 252, 253, 504, 505, 1008})
    rng = np.random.default_rng(11)
    haps = [synth.ACGT[rng.integers(0, 4, n)].tobytes() for n in (1, 2, 17, 64, 333)]
    reads, qs = [], []
    for R in lens:
        reads.append(synth.ACGT[rng.integers(0, 4, R)].tobytes())
        qs.append(rng.integers(2, 45, R))
    b = Batch.from_lists(haps, reads, qs, [np.full(len(q), 45) for q in qs],
                         [np.full(len(q), 45) for q in qs], [np.full(len(q), 10) for q in qs])
    check_against_oracle(oracle_mod, engine, b)


def test_n_bases_are_wildcards(engine, oracle_mod):
    rng = np.random.default_rng(5)
    b = synth.random_batch(rng, 6, 20, n_frac=0.05)
    check_against_oracle(oracle_mod, engine, b)


def test_rescue_unrelated_reads(engine, oracle_mod):
    # unrelated reads of length >= ~70 underflow FP32 (log10 L < -64.12)
    rng = np.random.default_rng(7)
    b = synth.random_batch(rng, 5, 40, hap_len=(150, 300), read_len=(40, 200), p_sub=0.75)
    out, fl, _ = check_against_oracle(oracle_mod, engine, b)
    assert (fl & FLAG_RESCUED).sum() > 10


def rescue_class_batch(seed):
    """Regions of reads at and around every boundary of the FP64 rescue's
    classes (32 lanes, MR = ceil((R + 1) / 32) rows per lane up to 255 rows,
    then the one-pair-per-wave pass): per region two haplotypes (one with an
    N in every third region) and six reads, three of them unrelated so that
    their FP32 sums underflow (from ~17 bases on)."""
    rng = np.random.default_rng(seed)
    lens = [1, 20, 24, 31, 32, 40, 63, 64, 95, 96, 127, 128, 150, 159, 160, 191, 192, 223, 224,
            240, 254, 255, 256, 300, 378]
    regs = []
    for k, R in enumerate(lens):
        hl = [R + int(rng.integers(1, 150)), R + int(rng.integers(1, 150))]
        regs.append(synth.region(rng, hl, [R] * 6, unrelated=(0, 2, 4),
                                 n_hap=(1,) if k % 3 == 0 else (), qual=40))
    b = Batch.concat(regs)
    # gap continuation Phred 40: an unrelated read costs ~4 decades per base
    # even through the insertion state, so reads from ~17 bases underflow
    b.gcp_q[:] = 40
    return b


def test_fp64_rescue_classes(engine, oracle_mod):
    b = rescue_class_batch(21)
    out, fl, _ = check_against_oracle(oracle_mod, engine, b)
    R = np.asarray(b.read_len)[b.cross_pairs()[1]]
    resc = (fl & FLAG_RESCUED) != 0
    assert resc.sum() > 100
    assert len(set((R[resc & (R <= 255)] // 32).tolist())) == 8  # every class has rescued pairs
    assert (resc & (R > 255)).any()
    # the same with haplotype chains and the plain launch order
    check_against_oracle(oracle_mod, engine, b, haplotype_chains=0)


def test_long_haplotypes_short_reads(engine, oracle_mod):
    # haplotypes too long for four or two groups' codes in LDS take the
    # 64-lane shapes; reads of every length class around them
    rng = np.random.default_rng(11)
    b = synth.random_batch(rng, 4, 12, hap_len=(2500, 9000), read_len=(10, 300))
    check_against_oracle(oracle_mod, engine, b)
    check_against_oracle(oracle_mod, engine, b, prior_table=0)


def test_fp64_underflow_long_reads(engine, oracle_mod):
    b = synth.config(4, scale=4 / 256)
    out, fl, o = check_against_oracle(oracle_mod, engine, b)
    assert (fl & FLAG_RESCUED).all()


def test_extended_range_long_reads(engine, oracle_mod):
    # Config-4 long reads, most of which the reference's FP64 path scores
    # -inf (its sum underflows).  With extended_range those pairs are scored
    # again with their strips' input rows renormalised: finite, within 1e-4
    # of the oracle's extended-range forward (itself pinned against an
    # unscaled log-space forward); every other pair keeps its bits, and the
    # default still gives the reference's -inf.
    b = synth.config(4, scale=16 / 256)
    plain, fl0, _, _ = run(engine, b)
    ext, fl1, _, _ = run(engine, b, extended_range=1)
    o_ext, _, _, _ = oracle_mod.run_batch(b, mode=oracle_mod.MODE_EXTENDED, nthreads=8)
    under = (fl0 & FLAG_UNDERFLOW) != 0
    assert under.sum() >= 4 and np.all(np.isneginf(plain[under]))
    np.testing.assert_array_equal(ext[~under], plain[~under])
    np.testing.assert_array_equal((fl1 & FLAG_EXTENDED) != 0, under)
    assert not (fl1 & FLAG_UNDERFLOW).any() and np.all(np.isfinite(ext))
    err = np.abs(ext - o_ext)
    assert err.max() <= TOL, f"max |dlog10L| {err.max()}"
    # graph replays of the same plan give the same results
    plan = engine.plan(b, default_config(semantics=0, extended_range=1, use_graph=1))
    for _ in range(3):
        plan.execute()
        got, gfl = plan.read()
        np.testing.assert_array_equal(got, ext)
        np.testing.assert_array_equal(gfl, fl1)
    plan.close()


def test_extended_range_is_validated(engine):
    b = synth.config(0)
    with pytest.raises(PhmmError) as e:
        engine.plan(b, default_config(extended_range=2))
    assert e.value.code == ERR_INVALID_ARG


@pytest.mark.parametrize("cfg", [0, 1])
def test_configs(engine, oracle_mod, cfg):
    check_against_oracle(oracle_mod, engine, synth.config(cfg))
    check_against_oracle(oracle_mod, engine, synth.config(cfg), prior_table=0)


def test_config2_subsample(engine, oracle_mod):
    b = synth.config(2, scale=0.002)
    check_against_oracle(oracle_mod, engine, b)


def test_plain_equals_batched_and_graph_replays(engine):
    b = synth.config(2, scale=0.003)
    plan = engine.plan(b, default_config(use_graph=1, semantics=0))
    plan.execute_plain()
    ref = plan.read(raw=True)
    for _ in range(3):
        plan.execute()
        got = plan.read(raw=True)
        np.testing.assert_array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[0], ref[0])
    plan.close()


def test_schedule_pairhmm_call_shape(engine, oracle_mod):
    rng = np.random.default_rng(3)
    b = synth.random_batch(rng, 3, 5)
    haps = [b.hap(k) for k in range(3)]
    reads = [b.read(k) for k in range(5)]
    qs = [b.quals(k) for k in range(5)]
    res = []
    ok = engine.schedule_pairhmm(haps, reads, res, [q[0] for q in qs], [q[1] for q in qs],
                                 [q[2] for q in qs], [q[3] for q in qs])
    assert ok and len(res) == 3 and all(len(r) == 5 for r in res)
    assert engine.schedule_pairhmm([], reads, res, [], [], [], []) is False
    o, _, _, _ = oracle_mod.run_batch(b)
    assert np.max(np.abs(np.array(res).ravel() - o)) <= TOL


def test_invalid_inputs_are_rejected(engine):
    q = [[30, 30, 30]]
    bad_base = Batch.from_lists(["ACGT"], ["ACX"], q, q, q, q)
    with pytest.raises(PhmmError) as e:
        engine.plan(bad_base)
    assert e.value.code == ERR_INVALID_ARG
    lower = Batch.from_lists(["acgt"], ["ACG"], q, q, q, q)
    with pytest.raises(PhmmError) as e:
        engine.plan(lower)
    assert e.value.code == ERR_INVALID_ARG
    empty_read = Batch.from_lists(["ACGT"], [""], [[]], [[]], [[]], [[]])
    with pytest.raises(PhmmError) as e:
        engine.plan(empty_read)
    assert e.value.code == ERR_INVALID_ARG
    no_reads = Batch.from_lists(["ACGT"], [], [], [], [], [])
    with pytest.raises(PhmmError) as e:
        engine.plan(no_reads)
    assert e.value.code == ERR_EMPTY


def test_allocation_failure_is_an_error_and_leaves_the_context_usable(engine, oracle_mod):
    # RODV_PHMM_FAIL_ALLOC=1 makes every device allocation of the library
    # fail: plan creation reports ERR_OOM (the partly built plan is freed),
    # and the same context plans and scores normally afterwards
    from rodv_amd.engine import ERR_OOM
    rng = np.random.default_rng(13)
    b = synth.random_batch(rng, 3, 9)
    os.environ["RODV_PHMM_FAIL_ALLOC"] = "1"
    try:
        for _ in range(3):
            with pytest.raises(PhmmError) as e:
                engine.plan(b)
            assert e.value.code == ERR_OOM
    finally:
        del os.environ["RODV_PHMM_FAIL_ALLOC"]
    check_against_oracle(oracle_mod, engine, b)


def test_long_read_wait_bound_is_an_error(engine, oracle_mod):
    # The waits between the units of a long read are bounded.  Forced to run
    # out from the host (RODV_PHMM_SPIN_LIMIT=0: no poll, so every unit that
    # follows another proceeds at once, with the NaN-filled hand-over rows of
    # a fresh plan), the plan's read reports a device error after copying the
    # results, the pairs whose FP64 rescue runs in several units (reads of
    # more than 1,764 rows: 4 strips of 441) are NaN and the others are not,
    # and the same context scores normally afterwards.
    from rodv_amd.engine import ERR_DEVICE, _ptr
    rng = np.random.default_rng(31)
    lens = [100, 500, UNIT + 1, UNIT + 505]
    acgt = np.frombuffer(b"ACGT", np.uint8)
    hap = acgt[rng.integers(0, 4, 3000)].tobytes()
    reads = [acgt[rng.integers(0, 4, n)].tobytes() for n in lens]  # unrelated: all rescued
    quals = [rng.integers(20, 41, n) for n in lens]
    b = Batch.from_lists([hap], reads, quals, quals, quals, [np.full(n, 10) for n in lens])
    multi = np.array(lens) > UNIT
    os.environ["RODV_PHMM_SPIN_LIMIT"] = "0"
    try:
        plan = engine.plan(b, default_config(semantics=0))
    finally:
        del os.environ["RODV_PHMM_SPIN_LIMIT"]
    plan.execute()
    out = np.zeros(plan.num_pairs)
    fl = np.zeros(plan.num_pairs, np.uint8)
    rc = plan._lib.rodv_phmm_plan_read(plan._h, _ptr(out), _ptr(fl), None, None)
    plan.close()
    assert rc == ERR_DEVICE
    assert np.isnan(out[multi]).all() and np.isfinite(out[~multi]).all()
    check_against_oracle(oracle_mod, engine, b)


def test_long_read_units_match_whole_pairs_and_oracle(engine, oracle_mod):
    # Reads longer than one strip: batched launches split each pair into
    # units (8 strips in FP32, 4 in FP64) that run on several workgroups at
    # once and hand boundary rows over through HBM slots (NaN until
    # written); plain launches run every FP32 pair whole in one workgroup.
    # Read lengths sit at the unit boundaries: FP64 strips of 441 rows
    # (1,764 / 1,765 / 3,528 / 3,529 / 7,056 / 7,057 rows: 4 / 5 / 8 / 9 / 16
    # / 17 strips), FP32 strips of 1,008 rows at 16 rows per lane (8,064 /
    # 8,065).  Reads are taken from haplotype 0 with rare substitutions (FP32
    # results), and are unrelated to haplotype 1 (rescued in FP64).
    rng = np.random.default_rng(23)
    lens = [379, 1100, UNIT, UNIT + 1, UNIT8, UNIT8 + 1, 2 * UNIT8, 2 * UNIT8 + 1, 8064, 8065]
    acgt = np.frombuffer(b"ACGT", np.uint8)
    h0 = acgt[rng.integers(0, 4, 8400)]
    h1 = acgt[rng.integers(0, 4, 3000)]
    reads = []
    for n in lens:
        s = int(rng.integers(0, 8400 - n + 1))
        r = h0[s:s + n].copy()
        sub = rng.random(n) < 5e-4
        r[sub] = acgt[(np.searchsorted(acgt, r[sub]) + 1) % 4]
        reads.append(r.tobytes())
    q = [rng.integers(20, 41, n) for n in lens]
    ins = [rng.integers(35, 46, n) for n in lens]
    dels = [rng.integers(35, 46, n) for n in lens]
    gcp = [np.full(n, 10) for n in lens]
    b = Batch.from_lists([h0.tobytes(), h1.tobytes()], reads, q, ins, dels, gcp)
    out, fl, _ = check_against_oracle(oracle_mod, engine, b)  # a fresh plan, batched
    n = len(lens)
    assert not (fl[:n] & FLAG_RESCUED).any()
    assert (fl[n:][np.array(lens) > 378] & FLAG_RESCUED).all()
    plan = engine.plan(b, default_config(semantics=0))
    plan.execute_plain()
    ref = plan.read(raw=True)
    resc = (ref[1] & FLAG_RESCUED) != 0
    for _ in range(2):
        plan.execute()
        got = plan.read(raw=True)
        np.testing.assert_array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
        np.testing.assert_array_equal(got[3][resc].view(np.uint64), ref[3][resc].view(np.uint64))
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[0], ref[0])
    plan.close()


def test_long_read_fp64_units_of_4_and_8_strips(engine):
    # the FP64 long-read rescue in units of 4 strips (two 4-wave workgroups
    # per CU, the default) and of 8 (one 8-wave workgroup per CU): the same
    # strips with the same boundary rows, so the same bits; reads of 2 to 17
    # FP64 strips, all rescued (unrelated to the haplotype), with and without
    # the LDS prior tables (a haplotype with N)
    rng = np.random.default_rng(37)
    lens = [500, UNIT, UNIT + 1, 2 * UNIT + 300, UNIT8 + 1, 2 * UNIT8 + 1]
    acgt = np.frombuffer(b"ACGT", np.uint8)
    for with_n in (False, True):
        hap = acgt[rng.integers(0, 4, 2500)].copy()
        if with_n:
            hap[rng.choice(2500, 5, replace=False)] = ord("N")
        reads = [acgt[rng.integers(0, 4, n)].tobytes() for n in lens]
        q = [rng.integers(20, 41, n) for n in lens]
        gcp = [np.full(n, 10) for n in lens]
        b = Batch.from_lists([hap.tobytes()], reads, q, q, q, gcp)
        got = {}
        for w in ("4", "8"):
            os.environ["RODV_PHMM_RESCUE_LONG_WAVES"] = w
            try:
                plan = engine.plan(b, default_config(semantics=0))
            finally:
                del os.environ["RODV_PHMM_RESCUE_LONG_WAVES"]
            plan.execute()
            got[w] = plan.read(raw=True)
            plan.close()
        assert (got["4"][1] & FLAG_RESCUED).all()
        np.testing.assert_array_equal(got["4"][3].view(np.uint64), got["8"][3].view(np.uint64))
        np.testing.assert_array_equal(got["4"][1], got["8"][1])
        np.testing.assert_array_equal(got["4"][0].view(np.uint64), got["8"][0].view(np.uint64))


def test_long_reads_haplotype_with_n(engine, oracle_mod):
    # Long reads over a haplotype with N: the plan then keeps the long-read
    # kernels' per-cell prior select (the LDS prior tables need N-free
    # haplotypes); batched (units, the sub-strip first launch) and plain
    # launches against the oracle and each other, FP32 results and rescues.
    rng = np.random.default_rng(29)
    lens = [300, 1100, 2300]
    acgt = np.frombuffer(b"ACGT", np.uint8)
    h0 = acgt[rng.integers(0, 4, 2600)].copy()
    h0[rng.choice(2600, 12, replace=False)] = ord("N")
    h1 = acgt[rng.integers(0, 4, 2000)]
    reads = []
    for n in lens:
        st = int(rng.integers(0, 2600 - n + 1))
        r = h0[st:st + n].copy()
        r[r == ord("N")] = ord("A")
        reads.append(r.tobytes())
    q = [rng.integers(20, 41, n) for n in lens]
    ins = [np.full(n, 45) for n in lens]
    gcp = [np.full(n, 10) for n in lens]
    b = Batch.from_lists([h0.tobytes(), h1.tobytes()], reads, q, ins, ins, gcp)
    out, fl, _ = check_against_oracle(oracle_mod, engine, b)
    n = len(lens)
    assert not (fl[:n] & FLAG_RESCUED).any()
    assert (fl[n:][np.array(lens) > 378] & FLAG_RESCUED).all()
    plan = engine.plan(b, default_config(semantics=0))
    plan.execute_plain()
    ref = plan.read(raw=True)
    plan.execute()
    got = plan.read(raw=True)
    plan.close()
    np.testing.assert_array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[0], ref[0])
