"""The drop-in as a reference caller uses it: tests/dropin/schedule_driver
declares pairhmm::schedule::schedule_pairhmm as the reference's header does
(src/core/pairhmm_schedule.h:14-25), links against librodv_phmm.so and runs
as a child process.  Its results are the oracle's restatement of
schedule_pairhmm (same routing and widths) within 1e-4, and identical to the
engine's own plan path; the reference's false/true contract holds."""
import dataclasses
import os
import sys

import numpy as np
import pytest

from rodv_amd import synth
from rodv_amd.batch import parse_gkl_testdata
from rodv_amd.engine import default_config

from conftest import GOLDEN

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin"))
import driver as D  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-4


def host_widths():
    c = default_config()
    return int(c.simd_width_float), int(c.simd_width_double)


def check(engine, oracle_mod, b, reps=1, **kw):
    args = D.batch_lists(b)
    ok, res, us = D.run(*args, reps=reps, **kw)
    assert ok
    wf, wd = host_widths()
    haps, reads = ([np.frombuffer(x, np.uint8) for x in seqs] for seqs in args[:2])
    okr, o_res, _ = oracle_mod.schedule(haps, reads, *args[2:], use_double=kw.get("use_double", False),
                                        idle_f=kw.get("idle_f", 0.1), idle_d=kw.get("idle_d", 0.1),
                                        wf=wf, wd=wd)
    assert okr and res.shape == o_res.shape
    fin = np.isfinite(o_res)
    np.testing.assert_array_equal(np.isfinite(res), fin)
    assert np.all(res[~fin] == o_res[~fin])
    if fin.any():
        assert np.max(np.abs(res[fin] - o_res[fin])) <= TOL
    # the same call through the engine's plan path gives the same doubles
    plan = engine.plan(b, default_config(use_double=int(kw.get("use_double", False)),
                                         max_idle_ratio_float=kw.get("idle_f", 0.1),
                                         max_idle_ratio_double=kw.get("idle_d", 0.1)))
    plan.execute()
    out, _ = plan.read()
    plan.close()
    np.testing.assert_array_equal(res.ravel(), out)
    return res, us


@pytest.mark.parametrize("cfg", [0, 1])
def test_configs_through_the_reference_call(engine, oracle_mod, cfg):
    res, us = check(engine, oracle_mod, synth.config(cfg), reps=8)
    # later calls reuse the context's device buffers
    print(f"config {cfg}: schedule_pairhmm per call {np.median(us[1:]):.0f} us "
          f"(first {us[0]:.0f} us)")


def test_gkl_vectors_through_the_reference_call(engine, oracle_mod):
    # all 104 haplotypes x 104 reads in one call; the diagonal pairs are the
    # GKL vectors, within 1e-5 of their expected values
    b, exp = parse_gkl_testdata(os.path.join(GOLDEN, "pairhmm-testdata.txt"))
    n = len(exp)
    one = dataclasses.replace(b, regions=np.array([[0, n, 0, n]], np.uint32))
    res, _ = check(engine, oracle_mod, one)
    k = np.arange(n)
    assert np.max(np.abs(res[k, k] - exp)) < 1e-5


def test_use_double_and_idle_bounds(engine, oracle_mod):
    rng = np.random.default_rng(8)
    b = synth.region(rng, [300, 280, 260], [120, 121, 100, 90, 150, 150, 150, 150, 80],
                     unrelated=[0, 3, 5])
    check(engine, oracle_mod, b, use_double=True)
    check(engine, oracle_mod, b, idle_f=0.5, idle_d=0.5)


def test_false_for_empty_and_invalid_input():
    # no haplotypes: the reference's false; a base outside ACGTN: false
    # instead of undefined behaviour
    ok, _, _ = D.run([], [b"ACGT"], [[30] * 4], [[45] * 4], [[45] * 4], [[10] * 4])
    assert not ok
    ok, _, _ = D.run([b"ACGT"], [b"ACXT"], [[30] * 4], [[45] * 4], [[45] * 4], [[10] * 4])
    assert not ok
