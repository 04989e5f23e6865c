"""End to end: reads of a BAM region (rodv_amd.bam) scored against a region's
haplotypes by the engine, under schedule_pairhmm routing, against the
oracle's restatement (the BAM reader itself: tests/test_bam.py)."""
import numpy as np
import pytest

from rodv_amd import bam, synth

from test_schedule_gpu import check_schedule

pytestmark = pytest.mark.gpu


def test_bam_region_reads_score_like_the_oracle(engine, oracle_mod, tmp_path):
    rng = np.random.default_rng(12)
    reg = synth.region(rng, [420, 400, 380], [150] * 40 + [101] * 20, unrelated=(3, 17, 45))
    haps = [reg.hap(k) for k in range(reg.num_haps)]
    recs = []
    for k in range(len(reg.read_len)):
        seq = reg.read(k)
        q = bytes(rng.integers(10, 41, len(seq)).astype(np.uint8))
        aux = b""
        if k % 3 == 0:  # GATK-recalibrated insertion / deletion qualities on some reads
            aux = (bam.BamRecord.aux_z("BI", "".join(chr(33 + int(x)) for x in rng.integers(30, 46, len(seq)))) +
                   bam.BamRecord.aux_z("BD", "".join(chr(33 + int(x)) for x in rng.integers(30, 46, len(seq)))))
        cig = [("S", 5), ("M", len(seq) - 5)] if k % 7 == 0 else [("M", len(seq))]
        recs.append(bam.BamRecord(name=f"read{k}", tid=0, pos=1000 + 3 * k, mapq=60, cigar=cig, seq=seq,
                                  qual=q, aux=aux, flag=0x400 if k == 5 else 0))
    recs.sort(key=lambda r: r.pos)
    p = tmp_path / "region.bam"
    with bam.BamWriter(str(p), [("chr1", 100_000)], "@HD\tVN:1.6\tSO:coordinate\n") as w:
        for r in recs:
            w.write(r)
    with bam.BamReader(str(p)) as rd:
        batch, used = bam.region_batch(rd, "chr1:1001-1200", haps)
    assert 40 < len(used) < len(recs)          # the interval, minus the duplicate
    assert all(r.flag & 0x400 == 0 for r in used)
    fl, route = check_schedule(engine, oracle_mod, batch, wf=16, wd=8)
    assert (route != 0).sum() > 0
