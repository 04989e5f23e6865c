"""CPU-side checks of the engine: the C-ABI library loads and exports every
declared entry point, its host tables match the reference's, and the host
containers / generators behave (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from rodv_amd import engine as E
from rodv_amd import synth
from rodv_amd.batch import Batch, parse_gkl_testdata

from conftest import GOLDEN, ROOT


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rodv_phmm_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = E.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(E.EXPORTS) <= names


def test_engine_tables_match_reference_dump():
    z = np.load(os.path.join(GOLDEN, "ref_tables.npz"))
    ph, mm = E.tables_f32(32640)
    assert np.array_equal(ph.view(np.uint32), z["ph2pr_f"].view(np.uint32))
    assert np.array_equal(mm.view(np.uint32), z["mm_f"].view(np.uint32))
    phd, mmd = E.tables_f64(32640)
    assert np.array_equal(phd.view(np.uint64), z["ph2pr_d"].view(np.uint64))
    assert np.array_equal(mmd.view(np.uint64), z["mm_d"].view(np.uint64))


def test_strerror_and_config_defaults():
    lib = E.load_library()
    assert lib.rodv_phmm_strerror(0) == b"ok"
    assert b"invalid" in lib.rodv_phmm_strerror(1)
    assert lib.rodv_phmm_strerror(99) == b"unknown status"
    c = E.default_config()
    assert c.struct_size == E.C.sizeof(E.Config)
    assert c.semantics == E.SEMANTICS_SCHEDULE and c.max_idle_ratio_float == 0.1
    # group widths follow this host's CPU like the reference's CpuFeatures
    flags = set()
    with open("/proc/cpuinfo") as f:
        for line in f:
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
    if {"avx512f", "avx512dq", "avx512bw", "avx512vl"} <= flags:
        assert (c.simd_width_float, c.simd_width_double) == (16, 8)
    elif "avx2" in flags:
        assert (c.simd_width_float, c.simd_width_double) == (8, 4)
    assert c.short_lanes == 16 and c.use_graph == 1
    with pytest.raises(E.PhmmError):
        E.tables_f32(40000)


def test_no_device_is_an_error_not_a_fallback():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(E.PhmmError) as e:
        E.Engine(0)
    assert e.value.code == E.ERR_DEVICE


def test_batch_cross_pairs_and_cells():
    q = [[30] * 3, [30] * 5]
    b = Batch.from_lists(["ACGTA", "GG", "TTT"], ["ACG", "ACGTN"], q, q, q, q)
    ph, pr = b.cross_pairs()
    assert list(ph) == [0, 0, 1, 1, 2, 2] and list(pr) == [0, 1, 0, 1, 0, 1]
    assert b.cells() == (5 + 2 + 3) * (3 + 5)
    bb = Batch.concat([b, b])
    assert bb.num_pairs() == 12 and bb.cells() == 2 * b.cells()
    assert list(bb.region_slices()) == [0, 6, 12]
    with pytest.raises(ValueError):
        Batch.from_lists(["A"], ["AC"], [[1]], [[1, 2]], [[1, 2]], [[1, 2]])


def test_gkl_parser():
    b, exp = parse_gkl_testdata(os.path.join(GOLDEN, "pairhmm-testdata.txt"))
    assert b.num_pairs() == 104 and len(exp) == 104
    assert b.base_q.max() <= 45 and b.gcp_q.max() == 10


@pytest.mark.parametrize("cfg,pairs", [(0, 512), (1, 4096), (3, 4096), (4, 256)])
def test_synthetic_configs_are_deterministic(cfg, pairs):
    a = synth.config(cfg, scale=1.0 if cfg != 4 else 0.05)
    b = synth.config(cfg, scale=1.0 if cfg != 4 else 0.05)
    if cfg != 4:
        assert a.num_pairs() == pairs
    assert np.array_equal(a.read_bases, b.read_bases)
    assert set(np.unique(a.read_bases)) <= set(b"ACGT")
    if cfg == 0:
        assert set(a.read_len) == {100} and set(a.hap_len) == {150}
        assert a.base_q.min() >= 10 and a.base_q.max() <= 40
    if cfg == 1:
        assert set(a.read_len) == {150} and a.hap_len.min() >= 300 and a.hap_len.max() <= 500
        assert (a.base_q <= 10).mean() == pytest.approx(0.05, abs=0.01)
    if cfg == 4:
        assert np.all(a.hap_len <= 12000) and a.read_len.min() >= 1000


def test_config2_shape_small():
    b = synth.config(2, scale=0.01)
    assert len(b.regions) == 100
    for (h0, m, r0, n) in b.regions:
        R = set(b.read_len[r0:r0 + n])
        assert len(R) == 1
        r = R.pop()
        assert 50 <= r <= 250 and 2 <= m <= 64 and 1 <= n <= 500
        d = b.hap_len[h0:h0 + m].astype(int) - r
        assert d.min() >= 50 and d.max() <= 400
