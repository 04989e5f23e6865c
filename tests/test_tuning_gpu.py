"""The autotuner's shapes on the GPU: every shape it can choose (4, 2 or 1
pairs per wavefront, 64-lane strip widths 4..16) computes results
bit-identical to the oracle's, and a tuned context changes speed only."""
import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch
from rodv_amd.engine import TUNE_CLASSES, Tuning, default_config

from test_forward_gpu import check_against_oracle, run

pytestmark = pytest.mark.gpu


def forced(pairs_per_wave, cap=0, wpc=0):
    t = Tuning()
    t.struct_size = __import__("ctypes").sizeof(Tuning)
    t.num_classes = TUNE_CLASSES
    for k in range(TUNE_CLASSES):
        t.cls[k].pairs_per_wave = pairs_per_wave
        t.cls[k].max_rows_per_lane = cap
        t.cls[k].waves_per_cu = wpc
    return t


def shape_batch(seed, lens, hap_lens=(1, 2, 17, 64, 333)):
    rng = np.random.default_rng(seed)
    haps = [synth.ACGT[rng.integers(0, 4, n)].tobytes() for n in hap_lens]
    reads, qs = [], []
    for R in lens:
        reads.append(synth.ACGT[rng.integers(0, 4, R)].tobytes())
        qs.append(rng.integers(2, 45, R))
    return Batch.from_lists(haps, reads, qs, [np.full(len(q), 45) for q in qs],
                            [np.full(len(q), 45) for q in qs], [np.full(len(q), 10) for q in qs])


@pytest.fixture
def tuned_engine(engine):
    yield engine
    engine.set_tuning(None)


def test_two_pairs_per_wave_every_rows_per_lane(tuned_engine, oracle_mod):
    # 32 lanes per pair: rows per lane 2..10 (31-row steps), both ends of each
    lens = sorted({max(1, 31 * m - d) for m in range(1, 11) for d in (0, 30)} | {1, 2, 3, 300, 310})
    b = shape_batch(21, lens)
    tuned_engine.set_tuning(forced(2))
    check_against_oracle(oracle_mod, tuned_engine, b)


@pytest.mark.parametrize("ppw,L", [(8, 8), (4, 16), (2, 32), (1, 64)])
def test_table_strip_uses_every_lane(tuned_engine, oracle_mod, ppw, L):
    # a table strip of L lanes x MR rows holds up to L*MR - 1 rows: the row-0
    # boundary is an identity slot at the top of lane 0 (R % MR != 0) or lane
    # L-1 left without read rows (R % MR == 0); both ends of every step.
    # 8 lanes exist for 6..12 rows per lane; other reads fall back.
    mrs = {8: range(6, 13), 16: range(1, 17), 32: range(2, 11), 64: (4, 8, 12, 16)}[L]
    lens = sorted({x for m in mrs
                   for x in (L * m - 1, L * m - 2, (L - 1) * m, (L - 1) * m + 1, L * (m - 1))
                   if x >= 1} | ({1, 5, 20, 39, 96, 130} if L == 8 else set()))
    b = shape_batch(30 + L, lens, hap_lens=(1, 5, 40, 300))
    tuned_engine.set_tuning(forced(ppw))
    check_against_oracle(oracle_mod, tuned_engine, b)


@pytest.mark.parametrize("cap", [4, 8, 12, 16])
def test_one_pair_per_wave_strip_widths(tuned_engine, oracle_mod, cap):
    # 64 lanes per pair with the strip width capped: longer reads run strip by strip
    lens = [1, 62, 63, 64, 252, 253, 400, 504, 505, 760, 1008, 1009, 1300]
    b = shape_batch(22 + cap, lens, hap_lens=(3, 90, 410))
    tuned_engine.set_tuning(forced(1, cap))
    check_against_oracle(oracle_mod, tuned_engine, b)


def test_random_lengths_two_pairs_per_wave_with_n(tuned_engine, oracle_mod):
    rng = np.random.default_rng(23)
    b = synth.random_batch(rng, 9, 31, hap_len=(1, 500), read_len=(1, 320), n_frac=0.01)
    tuned_engine.set_tuning(forced(2, wpc=6))
    check_against_oracle(oracle_mod, tuned_engine, b)


def test_autotune_changes_speed_not_results(tuned_engine):
    sample = synth.config(2, scale=0.002)
    tuned_engine.set_tuning(None)
    base = run(tuned_engine, sample)
    t = tuned_engine.autotune(sample)
    ents = t.entries()
    assert len(ents) == TUNE_CLASSES
    assert any(e["pairs_per_wave"] for e in ents)
    for e in ents:
        assert e["pairs_per_wave"] in (0, 1, 2, 4, 8)
        if e["pairs_per_wave"]:
            assert e["gcups"] >= e["default_gcups"] > 0
    got = tuned_engine.get_tuning().entries()
    assert [g["pairs_per_wave"] for g in got] == [e["pairs_per_wave"] for e in ents]
    tuned = run(tuned_engine, sample)
    np.testing.assert_array_equal(tuned[2].view(np.uint32), base[2].view(np.uint32))
    np.testing.assert_array_equal(tuned[1], base[1])
    np.testing.assert_array_equal(tuned[0], base[0])


def test_invalid_tuning_is_rejected(tuned_engine):
    from rodv_amd.engine import PhmmError
    bad = forced(3)
    with pytest.raises(PhmmError):
        tuned_engine.set_tuning(bad)
    bad = forced(1, cap=5)
    with pytest.raises(PhmmError):
        tuned_engine.set_tuning(bad)
