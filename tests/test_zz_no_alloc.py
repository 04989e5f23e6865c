"""rodv_phmm_forward_noalloc -- the forward without device allocation, from
the library's fixed device scratch -- against rodv_phmm_forward on the same
batches through the C ABI: identical log10 values and flags, on calls in a
row (the second with a smaller batch, so it runs over what the first left
in the scratch); a batch that does not fit the scratch is rejected with
RODV_PHMM_ERR_TOO_LARGE and the next call is unaffected."""
import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import Batch
from rodv_amd.engine import (ERR_TOO_LARGE, FLAG_RESCUED, NOALLOC_SCRATCH_BYTES, PhmmError,
                             default_config)

pytestmark = pytest.mark.gpu


def same(engine, b, cfg=None):
    ref_out, ref_fl = engine.forward(b, cfg)
    out, fl = engine.forward_noalloc(b, cfg)
    np.testing.assert_array_equal(out, ref_out)  # same doubles, -inf included
    np.testing.assert_array_equal(fl, ref_fl)
    return fl


def test_noalloc_matches_forward_on_calls_in_a_row(engine):
    rng = np.random.default_rng(41)
    # 40 regions of the batched workload, plus reads longer than one FP64
    # strip (378 rows) unrelated to their haplotype, which the FP32 pass
    # rejects: their rescue runs in strips with boundary rows in the scratch
    long_reg = synth.region(rng, [4500, 1800], [4000, 1200, 700], unrelated=[0, 1, 2])
    big = Batch.concat([synth.config(2, scale=0.004), long_reg])
    small = synth.config(0)
    fl = same(engine, big)
    assert (fl & FLAG_RESCUED).sum() > 50
    fl = same(engine, small)
    assert len(fl) < len(big.read_len) * 10
    # intra semantics as well, on a batch between the two
    same(engine, synth.config(1), default_config(semantics=0))


def test_noalloc_rejects_a_batch_beyond_its_scratch(engine):
    big = synth.config(2, scale=0.05)  # ~1.3 M pairs
    # per pair at least its (hap, read) indices, FP32 and FP64 sums, log10,
    # flags, a rescue-list entry and an item slot: 37 bytes
    assert big.num_pairs() * 37 > NOALLOC_SCRATCH_BYTES
    with pytest.raises(PhmmError) as e:
        engine.forward_noalloc(big)
    assert e.value.code == ERR_TOO_LARGE
    same(engine, synth.config(0))
