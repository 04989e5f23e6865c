"""The perf CLI (rodv_amd/perftest.py) keeps pairhmm_perftest's arguments and
exit status (test/pairhmm_perftest.cpp:271-488)."""
import os

import pytest

from rodv_amd import perftest

from conftest import GOLDEN

LOG = os.path.join(GOLDEN, "pairhmm_debug.txt")


def test_usage_and_argument_errors(capsys):
    assert perftest.main(["perftest"]) == 1
    assert "Usage:" in capsys.readouterr().out
    assert perftest.main(["perftest", "-h"]) == 0
    assert perftest.main(["perftest", LOG, "0.5"]) == 1
    assert perftest.main(["perftest", LOG, "x", "0.7"]) == 1
    assert perftest.main(["perftest", LOG, "0.5", "0.7", "extra"]) == 1
    assert perftest.main(["perftest", "--bogus", LOG, "0.5", "0.7"]) == 1


def test_compare_follows_the_reference_skips():
    import numpy as np
    from rodv_amd.batch import Batch
    q = [[30] * 4, [30] * 4]
    b = Batch.from_lists(["ACGT", "ANGT"], ["ACGT", "ACNT"], q, q, q, q)
    a = np.array([[-1.0, -2.0], [-3.0, -4.0]])
    ok, _ = perftest.compare(a, a + 5e-6, b, ignore_n=False)
    assert ok
    off = a.copy()
    off[1, 1] += 1.0  # hap 1 and read 1 contain N
    assert not perftest.compare(a, off, b, ignore_n=False)[0]
    assert perftest.compare(a, off, b, ignore_n=True)[0]
    z = a.copy()
    z[0, 0] = 0.0  # |intra| < 1e-10 is skipped
    assert perftest.compare(z, z + np.array([[9.0, 0], [0, 0]]), b, ignore_n=False)[0]


@pytest.mark.gpu
def test_debug_log_regions_pass(capsys):
    assert perftest.main(["perftest", "-v", LOG, "0.5", "0.7"]) == 0
    out = capsys.readouterr().out
    assert "Result check: PASSED" in out and "Total regions: 8" in out
