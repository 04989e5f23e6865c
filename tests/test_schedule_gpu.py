"""schedule_pairhmm routing on the GPU (RODV_PHMM_SEMANTICS_SCHEDULE) against
the oracle's restatement of it (oracle/phmm_oracle_sched.cpp, pinned by
test_schedule_oracle.py): the same pairs take the inter arithmetic, the FP32
sums that decide the rescue are bit-identical, the same pairs are regrouped
for inter FP64, and every log10 L agrees within 1e-4."""
import dataclasses
import os

import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch, parse_gkl_testdata
from rodv_amd.engine import FLAG_INTER, FLAG_RESCUED, FLAG_UNDERFLOW, default_config

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-4
ROUTE_INTER32, ROUTE_INTER64 = 1, 2
FP64_STRIP = 63 * 7  # rows of an FP64 strip of the long-read rescue (7 rows per lane)
UNIT = 4 * FP64_STRIP  # rows of an FP64 unit of 4 strips


def check_schedule(engine, oracle_mod, batch, wf=16, wd=8, idle_f=0.1, idle_d=0.1,
                   use_double=False, mode="batched", return_out=False, **extra):
    cfg = default_config(semantics=1, simd_width_float=wf, simd_width_double=wd,
                         max_idle_ratio_float=idle_f, max_idle_ratio_double=idle_d,
                         use_double=int(use_double), use_graph=1, **extra)
    plan = engine.plan(batch, cfg)
    if mode == "plain":
        plan.execute_plain()
    else:
        plan.execute()
    out, fl, r32, _ = plan.read(raw=True)
    plan.close()
    o_out, o_fl, o_r32, route = oracle_mod.schedule_batch(
        batch, use_double=use_double, idle_f=idle_f, idle_d=idle_d, wf=wf, wd=wd)
    # the deciding FP32 sums, bit for bit (inter arithmetic where routed)
    bad = np.nonzero(r32.view(np.uint32) != o_r32.view(np.uint32))[0]
    assert len(bad) == 0, f"{len(bad)} raw FP32 mismatches, first {bad[:5]}, routes {route[bad[:5]]}"
    mask = FLAG_RESCUED | FLAG_UNDERFLOW | FLAG_INTER
    np.testing.assert_array_equal(fl & mask, o_fl & mask)
    fin = np.isfinite(o_out)
    np.testing.assert_array_equal(np.isfinite(out), fin)
    if fin.any():
        err = np.abs(out[fin] - o_out[fin])
        assert err.max() <= TOL, f"max |dlog10L| {err.max()}"
    if return_out:
        return fl, route, out
    return fl, route


def test_fp64_rescue_classes_under_routing(engine, oracle_mod):
    # every FP64 rescue class in both arithmetics: 4-wide FP32 groups of
    # equal-length reads, rejected pairs regrouped two at a time
    from test_forward_gpu import rescue_class_batch
    b = rescue_class_batch(22)
    fl, route = check_schedule(engine, oracle_mod, b, wf=4, wd=2)
    assert (route == ROUTE_INTER64).sum() > 20
    assert ((fl & FLAG_RESCUED) != 0).sum() > 100


def test_hand_cases_route_like_the_oracle(engine, oracle_mod):
    rng = np.random.default_rng(1)
    regions = [
        synth.region(rng, [300], [100, 100, 100, 60, 60, 60, 60, 60]),
        synth.region(rng, [300], [53, 100, 51, 55, 52, 54]),
        synth.region(rng, [300, 280], [100] * 5, n_read=[2]),
        synth.region(rng, [200, 300], [90, 100]),
        synth.region(rng, [250], [100, 101, 102, 103, 104, 105, 106, 107], unrelated=range(8)),
    ]
    b = Batch.concat(regions)
    for idle in (0.1, 0.09, 1.0):
        fl, route = check_schedule(engine, oracle_mod, b, wf=4, wd=3, idle_f=idle, idle_d=1.0)
        assert (route == ROUTE_INTER32).any()
    fl, route = check_schedule(engine, oracle_mod, b, wf=4, wd=3, idle_f=1.0, idle_d=1.0)
    assert (route == ROUTE_INTER64).any()


@pytest.mark.parametrize("wf,wd", [(16, 8), (8, 4), (4, 2)])
def test_random_regions(engine, oracle_mod, wf, wd):
    rng = np.random.default_rng(10 + wf)
    regs = []
    for k in range(12):
        nh, nr = int(rng.integers(2, 9)), int(rng.integers(4, 40))
        hl = rng.integers(60, 300, nh)
        rl = rng.integers(30, 200, nr)
        unrelated = [i for i in range(nr) if rng.random() < 0.3]
        regs.append(synth.region(rng, hl, rl, unrelated=unrelated,
                                 n_read=[0] if k % 5 == 0 else []))
    b = Batch.concat(regs)
    fl, route = check_schedule(engine, oracle_mod, b, wf=wf, wd=wd, idle_f=0.5, idle_d=0.5)
    assert (fl & FLAG_INTER).any() and (fl & FLAG_RESCUED).any()


def test_use_double_is_all_intra(engine, oracle_mod):
    rng = np.random.default_rng(3)
    b = synth.region(rng, [300, 250], [100] * 8, unrelated=[0, 1, 2])
    fl, route = check_schedule(engine, oracle_mod, b, wf=4, wd=2, idle_f=1.0, use_double=True)
    assert not (fl & FLAG_INTER).any()


def test_no_simd_widths_is_all_intra(engine, oracle_mod):
    rng = np.random.default_rng(4)
    b = synth.region(rng, [300], [100] * 8)
    fl, _ = check_schedule(engine, oracle_mod, b, wf=0, wd=0, idle_f=1.0)
    assert not (fl & FLAG_INTER).any()


@pytest.mark.parametrize("wf,wd", [(16, 8), (8, 4)])
@pytest.mark.parametrize("cfg", [0, 1])
def test_configs_full_size_schedule(engine, oracle_mod, cfg, wf, wd):
    # the single-region configurations at full size under the default
    # semantics: equal read lengths, so every window of wf pairs sorted by
    # haplotype length has idle ratio 0 and every pair takes the inter
    # arithmetic (config 0: 512 pairs, config 1: 4,096)
    b = synth.config(cfg)
    fl, route = check_schedule(engine, oracle_mod, b, wf=wf, wd=wd)
    assert (route == ROUTE_INTER32).all()
    assert (fl & FLAG_INTER).all()


def test_config2_sample_schedule(engine, oracle_mod):
    # 120 regions (1.2 % of the batched workload's 10,000), 315,193 pairs
    b = synth.config(2, scale=0.012)
    assert len(b.regions) >= 100
    fl, route = check_schedule(engine, oracle_mod, b)
    assert (route == ROUTE_INTER32).sum() > 0.5 * len(route)
    assert (fl & FLAG_RESCUED).any()


@pytest.mark.parametrize("wf,wd", [(16, 8), (8, 4)])
def test_gkl_vectors_through_inter_groups(engine, oracle_mod, wf, wd):
    # The 104 GKL vectors through the inter arithmetic, as the reference's
    # ALLInterMatch tests run them (pairhmm_unittest.cpp:414-495): one region
    # of all 104 haplotypes x 104 reads with idle bound 1.0, so every window
    # of wf sorted pairs is accepted (10,816 pairs, a multiple of 16 and 8);
    # the diagonal pairs are the vectors, within 1e-5 of their expected value.
    # Most off-diagonal pairs are unrelated sequences whose FP32 sums the
    # inter groups reject: they go through the FP64 regroup (inter FP64, or
    # intra FP64 for a short last window).
    b, exp = parse_gkl_testdata(os.path.join(GOLDEN, "pairhmm-testdata.txt"))
    n = len(exp)
    one = dataclasses.replace(b, regions=np.array([[0, n, 0, n]], np.uint32))
    fl, route, out = check_schedule(engine, oracle_mod, one, wf=wf, wd=wd, idle_f=1.0, idle_d=1.0,
                                    return_out=True)
    assert (route != 0).all()  # every pair took an inter FP32 group
    assert (route == ROUTE_INTER64).sum() > 1000
    k = np.arange(n)
    assert (route.reshape(n, n)[k, k] == ROUTE_INTER32).all()
    diag = out.reshape(n, n)[k, k]
    assert np.max(np.abs(diag - exp)) < 1e-5


def test_plain_and_graph_agree_under_routing(engine, oracle_mod):
    rng = np.random.default_rng(5)
    regs = [synth.region(rng, rng.integers(100, 300, 4), rng.integers(60, 160, 24),
                         unrelated=[1, 2, 3, 7, 9, 11]) for _ in range(6)]
    b = Batch.concat(regs)
    check_schedule(engine, oracle_mod, b, wf=8, wd=4, idle_f=0.5, idle_d=0.5, mode="plain")
    cfg = default_config(semantics=1, simd_width_float=8, simd_width_double=4,
                         max_idle_ratio_float=0.5, max_idle_ratio_double=0.5, use_graph=1)
    plan = engine.plan(b, cfg)
    plan.execute_plain()
    ref = plan.read(raw=True)
    for _ in range(3):
        plan.execute()
        got = plan.read(raw=True)
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
        np.testing.assert_array_equal(got[0], ref[0])
    plan.close()


@pytest.mark.parametrize("seed", range(8))
def test_random_shapes_tunings_and_widths(engine, oracle_mod, seed):
    # regions of random lengths (unrelated reads, N in reads and haplotypes),
    # random group widths and idle bounds, and a random forced tuning per
    # length class (8/4/2/1 pairs per wave, strip caps, grids), chained
    # items forced or not: every shape gives the oracle's routing and bits
    import ctypes
    from rodv_amd.engine import TUNE_CLASSES, Tuning
    rng = np.random.default_rng(700 + seed)
    regs = []
    for _ in range(int(rng.integers(1, 5))):
        nh, nr = int(rng.integers(1, 9)), int(rng.integers(1, 30))
        hl = rng.integers(1, int(rng.choice([80, 300, 700, 1500])), nh)
        rl = rng.integers(1, int(rng.choice([60, 250, 320, 600])), nr)
        regs.append(synth.region(rng, hl, rl,
                                 unrelated=[i for i in range(nr) if rng.random() < 0.3],
                                 n_read=[i for i in range(nr) if rng.random() < 0.05],
                                 n_hap=[i for i in range(nh) if rng.random() < 0.05],
                                 qual=int(rng.integers(5, 41))))
    b = Batch.concat(regs)
    t = Tuning()
    t.struct_size = ctypes.sizeof(Tuning)
    t.num_classes = TUNE_CLASSES
    for k in range(TUNE_CLASSES):
        t.cls[k].pairs_per_wave = int(rng.choice([0, 1, 2, 4, 8]))
        t.cls[k].max_rows_per_lane = int(rng.choice([0, 4, 8, 12, 16]))
        t.cls[k].waves_per_cu = int(rng.choice([0, 2, 8, 16]))
    wf, wd = [(16, 8), (8, 4), (4, 2), (4, 3)][seed % 4]
    engine.set_tuning(t)
    try:
        check_schedule(engine, oracle_mod, b, wf=wf, wd=wd,
                       idle_f=float(rng.choice([0.0, 0.1, 0.5, 1.0])),
                       idle_d=float(rng.choice([0.1, 0.5, 1.0])),
                       haplotype_chains=int(rng.choice([1, 2])))
    finally:
        engine.set_tuning(None)


def test_long_reads_in_inter_fp64_groups(engine, oracle_mod):
    # Reads longer than the one-wave FP64 rescue (378 rows), unrelated to their
    # haplotypes: the FP32 pass rejects them, the regroup puts windows of
    # them in inter FP64 groups, and the long-read rescue runs them with the
    # inter arithmetic as units of at most 4 strips (reads over 1,764 rows
    # take more than one unit); batched and plain launches alike.
    rng = np.random.default_rng(29)
    regions = [
        synth.region(rng, [1800, 2300], [3540, 3560, 3600, 3490, 700, 720, 760, 300],
                     unrelated=range(8)),
        synth.region(rng, [900], [600, 610, 640, 650, 4100], unrelated=range(5)),
    ]
    b = Batch.concat(regions)
    _, pr = b.cross_pairs()
    R = b.read_len[pr]
    for mode in ("batched", "plain"):
        fl, route = check_schedule(engine, oracle_mod, b, wf=4, wd=4, idle_f=0.5, idle_d=0.5,
                                   mode=mode)
        assert ((route == ROUTE_INTER64) & (R > UNIT)).any()
        assert ((route == ROUTE_INTER64) & (R > 378) & (R <= UNIT)).any()
        assert (fl[R > 378] & FLAG_RESCUED).all()


def test_extended_range_under_routing(engine, oracle_mod):
    # the long unrelated reads of the test above under extended_range: the
    # inter FP64 groups' underflowed pairs are scored again with the inter
    # prior, the intra ones with the intra prior, each within 1e-4 of the
    # oracle's extended-range forward of its route's arithmetic
    from rodv_amd.engine import FLAG_EXTENDED
    rng = np.random.default_rng(29)
    regions = [
        synth.region(rng, [1800, 2300], [3540, 3560, 3600, 3490, 700, 720, 760, 300],
                     unrelated=range(8)),
        synth.region(rng, [900], [600, 610, 640, 650, 4100], unrelated=range(5)),
    ]
    b = Batch.concat(regions)
    cfg = default_config(semantics=1, simd_width_float=4, simd_width_double=4,
                         max_idle_ratio_float=0.5, max_idle_ratio_double=0.5, extended_range=1)
    plan = engine.plan(b, cfg)
    plan.execute()
    out, fl = plan.read()
    plan.close()
    _, o_fl, _, route = oracle_mod.schedule_batch(b, idle_f=0.5, idle_d=0.5, wf=4, wd=4)
    o_intra, _, _, _ = oracle_mod.run_batch(b, mode=oracle_mod.MODE_EXTENDED, nthreads=8)
    o_inter, _, _, _ = oracle_mod.run_batch(b, mode=oracle_mod.MODE_EXTENDED,
                                           semantics=oracle_mod.INTER, nthreads=8)
    ext = (fl & FLAG_EXTENDED) != 0
    np.testing.assert_array_equal(ext, (o_fl & FLAG_UNDERFLOW) != 0)
    assert ((route == ROUTE_INTER64) & ext).any() and ((route != ROUTE_INTER64) & ext).any()
    want = np.where(route == ROUTE_INTER64, o_inter, o_intra)
    assert np.all(np.isfinite(out[ext]))
    assert np.max(np.abs(out[ext] - want[ext])) <= TOL
