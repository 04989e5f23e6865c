"""Shared-product FP32 kernels (config.shared_products): a read whose deletion
quality of each row equals the insertion quality of the row below takes the
16-lane table and chain kernels that compute M(i,j) * mx_{i+1} once for both
I(i+1,j) and D(i,j+1).  The products are the same bits, so every raw FP32
sum, flag and log10 L must equal the unshared kernels' and the oracle's."""
import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch
from rodv_amd.engine import default_config

from test_chain_gpu import _edge_batch, _stack_batch
from test_forward_gpu import check_against_oracle
from test_schedule_gpu import check_schedule

pytestmark = pytest.mark.gpu


def _run(engine, batch, **cfg):
    cfg.setdefault("use_tuning", 0)  # the built-in 16-lane shapes
    plan = engine.plan(batch, default_config(**cfg))
    plan.execute()
    got = plan.read(raw=True)
    st = plan.stats()
    plan.close()
    return got, st


def _same(engine, batch, expect_shared, **cfg):
    on, st_on = _run(engine, batch, shared_products=1, **cfg)
    off, st_off = _run(engine, batch, shared_products=0, **cfg)
    assert st_off["shared_pairs"] == 0
    assert st_on["shared_pairs"] == expect_shared
    np.testing.assert_array_equal(on[2].view(np.uint32), off[2].view(np.uint32))
    np.testing.assert_array_equal(on[1], off[1])
    np.testing.assert_array_equal(on[0], off[0])


def _indel_quals(rng, b, mode):
    """mode 'const': every read 45/45; 'chained': varying qualities with
    del_q[k] == ins_q[k + 1] inside each read (shareable, not constant);
    'random': independent ins/del qualities (not shareable); 'mixed': the
    three kinds by read; 'regions': the three kinds by region."""
    b = Batch(**{k: getattr(b, k) for k in b.__dataclass_fields__})
    tot = int(b.read_len.sum())
    ins = np.full(tot, 45, np.uint8)
    dele = np.full(tot, 45, np.uint8)
    kinds = []
    region_of = np.zeros(len(b.read_len), np.int64)
    for k, rg in enumerate(np.asarray(b.regions, np.uint32).reshape(-1, 4)):
        region_of[int(rg[2]):int(rg[2]) + int(rg[3])] = k
    for r in range(len(b.read_len)):
        o, n = int(b.read_off[r]), int(b.read_len[r])
        kind = mode
        if mode == "mixed":
            kind = ("const", "chained", "random")[r % 3]
        elif mode == "regions":
            kind = ("const", "chained", "random")[region_of[r] % 3]
        kinds.append(kind)
        if kind == "chained":
            q = rng.integers(10, 50, n + 1).astype(np.uint8)
            ins[o:o + n] = q[:n]
            dele[o:o + n - 1] = q[1:n]
            dele[o + n - 1] = q[n]  # the last row's deletion quality is free
        elif kind == "random":
            ins[o:o + n] = rng.integers(10, 50, n).astype(np.uint8)
            dele[o:o + n] = rng.integers(10, 50, n).astype(np.uint8)
            if n > 1 and np.all(dele[o:o + n - 1] == ins[o + 1:o + n]):
                dele[o] ^= 1
    b.ins_q = ins
    b.del_q = dele
    return b, kinds


def _shared_pairs(b, kinds):
    """Pairs whose read is shareable (every read here takes a 16-lane table
    kernel: N-free haplotypes, reads of one strip)."""
    n = 0
    for rg in np.asarray(b.regions, np.uint32).reshape(-1, 4):
        hb, nh, rb, nr = (int(x) for x in rg)
        n += nh * sum(1 for r in range(rb, rb + nr) if kinds[r] != "random")
    return n


@pytest.mark.parametrize("mode", ["const", "chained", "mixed"])
@pytest.mark.parametrize("chains", [0, 2])
def test_shared_equals_unshared(engine, mode, chains):
    rng = np.random.default_rng(11 + chains)
    b, kinds = _indel_quals(rng, _edge_batch(rng), mode)
    _same(engine, b, _shared_pairs(b, kinds), haplotype_chains=chains, semantics=1)
    _same(engine, b, _shared_pairs(b, kinds), haplotype_chains=chains, semantics=0)


@pytest.mark.parametrize("mode", ["chained", "mixed"])
def test_shared_against_oracle(engine, oracle_mod, mode):
    rng = np.random.default_rng(5)
    b, _ = _indel_quals(rng, _edge_batch(rng), mode)
    check_against_oracle(oracle_mod, engine, b, haplotype_chains=2)
    check_schedule(engine, oracle_mod, b, haplotype_chains=2)
    check_schedule(engine, oracle_mod, b, haplotype_chains=0, wf=8, wd=4)


def test_shared_stacked_reads(engine, oracle_mod):
    # short reads two per lane group, in regions of shareable reads (constant
    # or chained qualities) and of reads that are not
    rng = np.random.default_rng(9)
    b, kinds = _indel_quals(rng, _stack_batch(rng), "regions")
    on, st = _run(engine, b, haplotype_chains=2, shared_products=1)
    assert st["stacked_pairs"] > 0
    _same(engine, b, _shared_pairs(b, kinds), haplotype_chains=2)
    check_schedule(engine, oracle_mod, b, haplotype_chains=2)


def test_shared_workload_sample(engine, oracle_mod):
    # the batched workload's own reads (constant indel qualities), 1 %
    b = synth.config(2, scale=0.01)
    on, st = _run(engine, b)
    assert st["shared_pairs"] > 0.9 * b.num_pairs()
    _same(engine, b, st["shared_pairs"])
    check_schedule(engine, oracle_mod, b)
