"""The HIP adjoint against the FP64 oracle gradient, central finite
differences of the FP64 forward (h = 1e-3), checkpoint-stride invariance and
the torch.autograd wrapper."""
import dataclasses
import os

import numpy as np
import pytest

from rodv_amd import synth

pytestmark = pytest.mark.gpu

THETA = np.array([45.0, 10.0, 1.0, 0.9, 1.1, 1.0, 0.95, 1.05, 1.0, 1.2])
REL = 1e-3


def assert_grad_close(g, ref, rel=REL):
    scale = np.max(np.abs(ref))
    for k in range(10):
        assert abs(g[k] - ref[k]) <= rel * max(abs(ref[k]), 1e-3 * scale), (k, g[k], ref[k])


def small_batch(seed, **kw):
    rng = np.random.default_rng(seed)
    return synth.random_batch(rng, 3, 5, **kw)


@pytest.mark.parametrize("seed", [1, 2])
def test_adjoint_matches_fp64_oracle(engine, oracle_mod, seed):
    b = small_batch(seed, hap_len=(20, 200), read_len=(5, 150), qual=(5, 45))
    w = np.random.default_rng(seed).normal(size=b.num_pairs())
    out, g = engine.adjoint(b, THETA, weights=w)
    o_out, o_g = oracle_mod.param_batch(b, THETA, weights=w, nthreads=8)
    assert np.max(np.abs(out - o_out)) <= 1e-4
    assert_grad_close(g, o_g)


def test_adjoint_config3_sample(engine, oracle_mod):
    b = synth.config(3, scale=1 / 32)  # 1 hap x 128 reads, R = 150, H = 400
    out, g = engine.adjoint(b, THETA)
    o_out, o_g = oracle_mod.param_batch(b, THETA, nthreads=8)
    assert np.max(np.abs(out - o_out)) <= 1e-4
    assert_grad_close(g, o_g)


def test_adjoint_every_shape(engine, oracle_mod):
    # reads at both ends of every (lanes per pair, rows per lane) shape of the
    # one-strip kernel (16 lanes up to 75 rows, 32 up to 155, 64 up to 378),
    # two and four pairs per wave next to single ones, empty slots
    rng = np.random.default_rng(31)
    lens = [1, 2, 15, 16, 30, 31, 45, 46, 60, 61, 75, 76, 93, 124, 125, 155, 156, 189, 190, 252, 315,
            378]
    regs = [synth.region(rng, [int(R) + int(rng.integers(1, 120)), int(R) + 40], [int(R)] * 3,
                         unrelated=(1,) if R > 80 else (), qual=int(rng.integers(8, 41)))
            for R in lens]
    from rodv_amd.batch import Batch
    b = Batch.concat(regs)
    w = rng.normal(size=b.num_pairs())
    out, g = engine.adjoint(b, THETA, weights=w)
    o_out, o_g = oracle_mod.param_batch(b, THETA, weights=w, nthreads=8)
    fin = np.isfinite(o_out)
    np.testing.assert_array_equal(np.isfinite(out), fin)
    assert np.max(np.abs(out[fin] - o_out[fin])) <= 1e-4
    assert_grad_close(g, o_g)


def test_adjoint_config3_sample_central_differences(engine, oracle_mod):
    # a sample of the 4,096-pair workload (2 haplotypes x 128 reads, R = 150,
    # H = 400), against central differences of the FP64 oracle's forward
    b = synth.config(3, scale=2 / 32)
    w = np.random.default_rng(5).normal(size=b.num_pairs())
    _, g = engine.adjoint(b, THETA, weights=w)
    h = 1e-3
    fd = np.zeros(10)
    for k in range(10):
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        fp, _ = oracle_mod.param_batch(b, tp, nthreads=8)
        fm, _ = oracle_mod.param_batch(b, tm, nthreads=8)
        fd[k] = np.sum(w * (fp - fm)) / (2 * h)
    assert_grad_close(g, fd)


def test_adjoint_vs_central_differences(engine, oracle_mod):
    b = small_batch(7, hap_len=(30, 120), read_len=(20, 100), qual=(5, 45))
    _, g = engine.adjoint(b, THETA)
    h = 1e-3
    fd = np.zeros(10)
    for k in range(10):
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        fp, _ = oracle_mod.param_batch(b, tp, nthreads=8)
        fm, _ = oracle_mod.param_batch(b, tm, nthreads=8)
        fd[k] = np.sum(fp - fm) / (2 * h)
    assert_grad_close(g, fd)


def test_checkpoint_stride_invariance(engine):
    b = small_batch(3, hap_len=(100, 300), read_len=(60, 200))
    ref = None
    for stride in (4, 8, 16, 32, 64):
        out, g = engine.adjoint(b, THETA, ckpt_stride=stride)
        if ref is None:
            ref = (out, g)
        else:
            np.testing.assert_array_equal(out, ref[0])
            np.testing.assert_allclose(g, ref[1], rtol=1e-9)


def test_rescued_pairs_use_fp64(engine, oracle_mod):
    b = small_batch(4, hap_len=(150, 250), read_len=(120, 200), p_sub=0.75)
    out, g = engine.adjoint(b, THETA)
    o_out, o_g = oracle_mod.param_batch(b, THETA, nthreads=8)
    assert np.max(out) < -40  # these pairs underflow FP32 (log10 L < -64 mostly)
    assert np.max(np.abs(out - o_out)) <= 1e-4
    assert_grad_close(g, o_g)


def test_torch_autograd_function(engine, oracle_mod):
    import torch
    from rodv_amd.autograd import PairHMMLogLik
    b = small_batch(5)
    theta = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    ll = PairHMMLogLik.apply(theta, engine, b)
    w = torch.linspace(-1, 1, ll.numel(), dtype=torch.float64)
    (ll * w).sum().backward()
    _, o_g = oracle_mod.param_batch(b, THETA, weights=w.numpy(), nthreads=8)
    assert_grad_close(theta.grad.numpy(), o_g)


def long_pairs(k_shortest):
    """The k shortest pairs of config 4's first six (R 1,007 .. 2,743, one
    pair per region; the pair of R = 2,743 is one whose plain FP64 sum
    underflows)."""
    long = synth.config(4, scale=6 / 256)
    sel = np.sort(np.argsort(long.read_len)[:k_shortest])
    return dataclasses.replace(long, regions=long.regions[sel])


def test_adjoint_long_reads(engine, oracle_mod):
    # reads longer than one strip (378 rows): strips of 252 rows in
    # sequence, each strip's input row renormalised by a power of two, FP64;
    # log10 L and the gradient against the oracle's renormalised reverse
    # mode over full matrices
    b = long_pairs(5)
    w = np.array([1.0, 0.5, -0.7, 1.3, 0.9])
    out, g = engine.adjoint(b, THETA, weights=w)
    o_out, o_g = oracle_mod.param_batch(b, THETA, weights=w, nthreads=5)
    l0, _ = oracle_mod.param_batch(b, THETA, nthreads=5, renorm=False)
    assert np.isneginf(l0).any()  # beyond the plain FP64 range
    assert np.all(np.isfinite(out))
    assert np.max(np.abs(out - o_out)) <= 1e-4
    assert_grad_close(g, o_g)


def test_adjoint_long_reads_central_differences(engine):
    # the engine's own log10 L at theta +- h (h = 1e-3) on two long reads,
    # one beyond the plain FP64 range
    long = synth.config(4, scale=6 / 256)
    order = np.argsort(long.read_len)
    b = dataclasses.replace(long, regions=long.regions[np.sort(order[[0, 4]])])
    out, g = engine.adjoint(b, THETA)
    assert np.all(np.isfinite(out))
    h = 1e-3
    fd = np.zeros(10)
    for k in range(10):
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        fp, _ = engine.adjoint(b, tp)
        fm, _ = engine.adjoint(b, tm)
        fd[k] = np.sum(fp - fm) / (2 * h)
    assert_grad_close(g, fd)


def test_adjoint_long_and_short_reads_together(engine, oracle_mod):
    # one batch with reads on both kernels: per-pair log10 L and the
    # weighted gradient sum over both
    rng = np.random.default_rng(31)
    short = synth.random_batch(rng, 2, 3, hap_len=(100, 300), read_len=(50, 300), qual=(5, 40))
    from rodv_amd.batch import Batch
    b = Batch.concat([short, long_pairs(2)])
    w = np.linspace(-1.0, 1.0, b.num_pairs())
    out, g = engine.adjoint(b, THETA, weights=w)
    o_out, o_g = oracle_mod.param_batch(b, THETA, weights=w, nthreads=8)
    assert np.max(np.abs(out - o_out)) <= 1e-4
    assert_grad_close(g, o_g)


def _adjoint_env(engine, b, env, **kw):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return engine.adjoint(b, THETA, **kw)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def test_adjoint_long_reads_waves_same_bits(engine):
    # by default a long pair's strips run on several workgroups (pass A per
    # pair, pass B one wave per strip handing adjoints up through HBM
    # progress counts, in one launch that starts a pair's pass B once its
    # pass A is done; the sums replayed in strip order by a second); in
    # three launches (A, B, the sums), on one workgroup of 4 waves (LDS
    # progress) or of one wave they are the same bits (the default is
    # checked against the oracle above), and so are the launches run a few
    # pairs at a time.
    # Reads of 2 (fewer strips than waves) to 30 strips of 252 rows, one
    # and two reads per haplotype; the reads unrelated to their haplotypes
    # (R 1,002) lose more per strip than three strips without a fresh shift
    # can hold, so pass A runs again with a shift per strip (finite results).
    rng = np.random.default_rng(77)
    from rodv_amd.batch import Batch
    regs = [synth.region(rng, [R + 300, R + 90], [R, R - 7], unrelated=(1,) if R > 900 else (), qual=q)
            for R, q in ((379, 30), (505, 12), (760, 25), (1009, 35))]
    b = Batch.concat(regs + [long_pairs(6)])
    w = rng.normal(size=b.num_pairs())
    one = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_WAVES": 1}, weights=w)
    four = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_WAVES": 4}, weights=w)
    split = engine.adjoint(b, THETA, weights=w)
    three = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_LAUNCHES": 3}, weights=w)
    # pass B recomputing its checkpoints instead of reading pass A's
    own = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_PASSA_CKPT": 0}, weights=w)
    # the split launches over chunks of 3 long pairs that reuse their buffers
    chunked = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_CHUNK": 3}, weights=w)
    for got in (four, split, three, own, chunked):
        np.testing.assert_array_equal(got[0].view(np.uint64), one[0].view(np.uint64))
        np.testing.assert_array_equal(got[1].view(np.uint64), one[1].view(np.uint64))
    assert np.all(np.isfinite(split[0]))


def test_adjoint_long_wait_bound_is_an_error(engine):
    # the waits between the waves and workgroups are bounded: forced to run
    # out from the host (no poll), every kernel still ends, the call reports
    # a device error, and the next adjoint on the same context is exact again
    # (the default launches, the three launches and the one-workgroup kernel
    # alike)
    from rodv_amd.engine import ERR_DEVICE, PhmmError
    b = long_pairs(2)
    for env in ({"RODV_PHMM_SPIN_LIMIT": 0}, {"RODV_PHMM_SPIN_LIMIT": 0, "RODV_PHMM_ADJ_LONG_LAUNCHES": 3},
                {"RODV_PHMM_SPIN_LIMIT": 0, "RODV_PHMM_ADJ_LONG_WAVES": 4}):
        with pytest.raises(PhmmError) as e:
            _adjoint_env(engine, b, env)
        assert e.value.code == ERR_DEVICE
    ref = _adjoint_env(engine, b, {"RODV_PHMM_ADJ_LONG_WAVES": 1})
    out, g = engine.adjoint(b, THETA)
    np.testing.assert_array_equal(out, ref[0])
    np.testing.assert_array_equal(g, ref[1])
