"""BAM input (rodv_amd/bam.py over librodv_bam.so) against the SAM
specification.  The reference's own BamHandle (src/iostream/bam_handle.*)
wraps htslib, which is absent here, and the reference holds no BAM fixture
(src/iostream/tests/test_bam_reader.cpp is empty): parity with it is
unpinned.  What pins the implementation instead:
- BAM bytes built in this file straight from the specification (records
  section 4.2, BGZF 4.1, with Python's zlib) are read back field by field,
  including records split across BGZF blocks;
- what the writer produces decompresses with Python's gzip module (BGZF is
  multi-member gzip) to exactly the bytes this file's encoder gives;
- region queries through the BAI the writer builds return exactly the
  records a brute-force overlap filter of the whole file returns, over many
  random intervals, bins of every level and windows past 16 kb;
- round trips, region strings, corrupt input and the read-input step."""
import gzip
import os
import struct
import zlib

import numpy as np
import pytest

from rodv_amd import bam

SEQ_CODES = "=ACMGRSVTWYHKDBN"
OPS = "MIDNSHP=X"


# ---------------------------------------------------------------- spec encoder
def ref_span(cigar):
    return sum(n for op, n in cigar if op in "MDN=X")


def spec_bin(beg, end):
    # SAM specification 5.3, reg2bin: the smallest bin holding [beg, end)
    end -= 1
    for shift, first in ((14, 4681), (17, 585), (20, 73), (23, 9), (26, 1)):
        if beg >> shift == end >> shift:
            return first + (beg >> shift)
    return 0


def spec_record(r: bam.BamRecord) -> bytes:
    nm = r.name.encode() + b"\0"
    span = ref_span(r.cigar)
    end = r.pos + (span if span > 0 else 1)
    binv = spec_bin(r.pos, end) if r.pos >= 0 else 4680
    body = struct.pack("<iiBBHHHiiii", r.tid, r.pos, len(nm), r.mapq, binv, len(r.cigar), r.flag,
                       len(r.seq), r.next_tid, r.next_pos, r.tlen)
    body += nm + b"".join(struct.pack("<I", n << 4 | OPS.index(op)) for op, n in r.cigar)
    codes = [SEQ_CODES.index(chr(c)) for c in r.seq]
    if len(codes) % 2:
        codes.append(0)
    body += bytes(codes[i] << 4 | codes[i + 1] for i in range(0, len(codes), 2))
    body += (r.qual if r.qual else b"\xff" * len(r.seq)) + r.aux
    return struct.pack("<i", len(body)) + body


def spec_header(text, targets):
    h = b"BAM\1" + struct.pack("<i", len(text)) + text.encode() + struct.pack("<i", len(targets))
    for name, ln in targets:
        h += struct.pack("<i", len(name) + 1) + name.encode() + b"\0" + struct.pack("<i", ln)
    return h


def spec_bgzf(data: bytes, chunk: int) -> bytes:
    out = b""
    for k in range(0, len(data), chunk):
        part = data[k:k + chunk]
        co = zlib.compressobj(6, zlib.DEFLATED, -15)
        cdata = co.compress(part) + co.flush()
        bsize = 18 + len(cdata) + 8 - 1
        out += bytes([31, 139, 8, 4, 0, 0, 0, 0, 0, 255]) + struct.pack("<H", 6) + b"BC" + \
            struct.pack("<HH", 2, bsize) + cdata + struct.pack("<II", zlib.crc32(part), len(part))
    eof = bytes([31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 66, 67, 2, 0, 27, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    return out + eof


# ---------------------------------------------------------------- workloads
TARGETS = [("chr1", 250_000), ("chr2", 90_000), ("chrM", 16_569)]


def random_records(rng, n, sort=True, unplaced=3):
    recs = []
    for k in range(n):
        tid = int(rng.integers(0, len(TARGETS)))
        L = int(rng.integers(1, 160))
        pos = int(rng.integers(0, TARGETS[tid][1] - 1))
        cig = []
        left = L
        if rng.random() < 0.2:
            s = int(rng.integers(1, max(2, L // 3)))
            cig.append(("S", min(s, left)))
            left -= cig[-1][1]
        while left > 0:
            op = rng.choice(list("MMMMIDN=X"))
            if op in "MI=X":
                m = min(left, int(rng.integers(1, 60)))
                left -= m
            else:  # reference-only ops: long ones reach other bins and windows
                m = int(rng.integers(1, 40000 if op == "N" else 30))
            cig.append((op, m))
        if rng.random() < 0.1:
            cig.append(("H", 5))
        seq = bytes(rng.choice(list(b"ACGTN"), L).astype(np.uint8))
        qual = bytes(rng.integers(2, 41, L).astype(np.uint8)) if rng.random() < 0.9 else b"\xff" * L
        aux = b""
        if rng.random() < 0.5:
            aux += bam.BamRecord.aux_z("BI", "".join(chr(33 + int(x)) for x in rng.integers(20, 50, L)))
            aux += b"NMi" + struct.pack("<i", int(rng.integers(0, 9)))
            aux += bam.BamRecord.aux_z("BD", "".join(chr(33 + int(x)) for x in rng.integers(20, 50, L)))
        if rng.random() < 0.2:
            aux += b"XBBc" + struct.pack("<i", 3) + bytes([1, 2, 3])
        recs.append(bam.BamRecord(name=f"r{k}", tid=tid, pos=pos, mapq=int(rng.integers(0, 61)),
                                  flag=int(rng.choice([0, 16, 99, 147, 0x400])), cigar=cig, seq=seq,
                                  qual=qual, next_tid=tid, next_pos=pos + 100, tlen=int(rng.integers(-500, 500)),
                                  aux=aux))
    if sort:
        recs.sort(key=lambda r: (r.tid, r.pos))
    for k in range(unplaced):  # unmapped, no coordinates, last
        recs.append(bam.BamRecord(name=f"u{k}", tid=-1, pos=-1, flag=4, seq=b"ACGTACGT",
                                  qual=bytes([30] * 8)))
    return recs


def write_file(path, recs, level=6, text="@HD\tVN:1.6\tSO:coordinate\n"):
    with bam.BamWriter(str(path), TARGETS, text, level) as w:
        for r in recs:
            w.write(r)


def same(a: bam.BamRecord, b: bam.BamRecord):
    for f in ("name", "tid", "pos", "mapq", "flag", "cigar", "seq", "qual", "next_tid", "next_pos", "tlen", "aux"):
        assert getattr(a, f) == getattr(b, f), (f, getattr(a, f), getattr(b, f))


def expected_end(r):
    span = ref_span(r.cigar)
    return r.pos + (span if span > 0 else 1) if r.pos >= 0 else r.pos


# ---------------------------------------------------------------- tests
def test_library_exports_its_c_abi():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", bam.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = sorted(l.split()[-1] for l in out.splitlines() if l.strip())
    assert syms == sorted(bam.EXPORTS)


@pytest.mark.parametrize("chunk", [1000, 65280])
def test_reader_decodes_spec_encoded_bam(tmp_path, chunk):
    # a file built here from the specification, in blocks of `chunk` bytes
    # (1000: records split across blocks), read field by field
    recs = random_records(np.random.default_rng(1), 300)
    text = "@HD\tVN:1.6\n@SQ\tSN:chr1\tLN:250000\n"
    stream = spec_header(text, TARGETS) + b"".join(spec_record(r) for r in recs)
    p = tmp_path / "spec.bam"
    p.write_bytes(spec_bgzf(stream, chunk))
    with bam.BamReader(str(p)) as rd:
        assert rd.targets == TARGETS
        assert rd.header_text == text
        got = list(rd)
    assert len(got) == len(recs)
    for a, b in zip(got, recs):
        same(a, b)
        assert a.end == expected_end(b)


@pytest.mark.parametrize("level", [0, 1, 6, 9])
def test_written_file_is_gzip_of_the_spec_stream(tmp_path, level):
    recs = random_records(np.random.default_rng(2), 400)
    p = tmp_path / "w.bam"
    text = "@HD\tVN:1.6\tSO:coordinate\n"
    write_file(p, recs, level, text)
    raw = p.read_bytes()
    assert raw[-28:] == spec_bgzf(b"", 1)  # the BGZF end block
    stream = spec_header(text, TARGETS) + b"".join(spec_record(r) for r in recs)
    assert gzip.decompress(raw) == stream
    # every block within the format's 64 KiB limits
    off = 0
    while off < len(raw):
        assert raw[off:off + 4] == bytes([31, 139, 8, 4])
        bsize = struct.unpack("<H", raw[off + 16:off + 18])[0]
        isize = struct.unpack("<I", raw[off + bsize + 1 - 4:off + bsize + 1])[0]
        assert bsize + 1 <= 65536 and isize <= 65536
        off += bsize + 1
    assert off == len(raw)


def test_round_trip(tmp_path):
    recs = random_records(np.random.default_rng(3), 2000)
    p = tmp_path / "rt.bam"
    write_file(p, recs)
    assert os.path.exists(str(p) + ".bai")
    rd = bam.BamReader(str(p))
    got = []
    while rd.has_next():
        got.append(rd.next())
    assert rd.next() is None and not rd.has_next()
    rd.close()
    assert len(got) == len(recs)
    for a, b in zip(got, recs):
        same(a, b)


def brute(recs, ivs):
    return [r.name for r in recs if any(r.tid == t and r.pos >= 0 and r.pos < e and expected_end(r) > b
                                        for t, b, e in ivs)]


def test_region_queries_match_brute_force(tmp_path):
    rng = np.random.default_rng(4)
    recs = random_records(rng, 5000)
    p = tmp_path / "q.bam"
    write_file(p, recs)
    rd = bam.BamReader(str(p))
    for k in range(300):
        tid = int(rng.integers(0, len(TARGETS)))
        ln = TARGETS[tid][1]
        w = int(rng.choice([1, 10, 1000, 20000, 200000]))
        beg = int(rng.integers(0, ln))
        end = min(ln, beg + max(1, int(rng.integers(1, w + 1))))
        got = [r.name for r in rd.query(tid=tid, beg=beg, end=end)]
        assert got == brute(recs, [(tid, beg, end)]), (tid, beg, end)
    # region strings: 1-based inclusive, commas allowed, whole references
    got = [r.name for r in rd.query("chr2:1,001-2,000")]
    assert got == brute(recs, [(1, 1000, 2000)])
    got = [r.name for r in rd.query("chrM")]
    assert got == brute(recs, [(2, 0, TARGETS[2][1])])
    got = [r.name for r in rd.query("chr1:240000")]
    assert got == brute(recs, [(0, 239999, TARGETS[0][1])])
    # several regions at once: each record once, in file order
    regs = ["chr1:1000-5000", "chr1:4000-90000", "chr2:50-60", "chrM:1-16569"]
    ivs = [(0, 999, 5000), (0, 3999, 90000), (1, 49, 60), (2, 0, 16569)]
    assert [r.name for r in rd.query_multi(regs)] == brute(recs, ivs)
    rd.close()


def test_query_errors(tmp_path):
    recs = random_records(np.random.default_rng(5), 50)
    p = tmp_path / "e.bam"
    write_file(p, recs)
    rd = bam.BamReader(str(p))
    for bad in ("chr9:1-10", "chr1:0-10", "chr1:50-10", "chr1:x", ""):
        with pytest.raises(bam.BamError) as e:
            rd.query(bad)
        assert e.value.code == bam.ERR_REGION
    rd.close()
    # unsorted output: no index, so a region query is refused
    u = tmp_path / "u.bam"
    write_file(u, random_records(np.random.default_rng(6), 50, sort=False))
    assert not os.path.exists(str(u) + ".bai")
    rd = bam.BamReader(str(u))
    assert len(list(rd)) == 53
    with pytest.raises(bam.BamError) as e:
        rd.query("chr1:1-100")
    assert e.value.code == bam.ERR_NO_INDEX
    rd.close()


def test_corrupt_input_is_rejected(tmp_path):
    recs = random_records(np.random.default_rng(7), 100)
    p = tmp_path / "c.bam"
    write_file(p, recs)
    raw = bytearray(p.read_bytes())
    q = tmp_path / "bad.bam"
    q.write_bytes(b"not a bam file at all")
    with pytest.raises(bam.BamError) as e:
        bam.BamReader(str(q))
    assert e.value.code == bam.ERR_FORMAT
    # a flipped byte inside the first block's compressed data: CRC or inflate fails
    bad = bytearray(raw)
    bad[30] ^= 0xff
    q.write_bytes(bytes(bad))
    with pytest.raises(bam.BamError) as e:
        bam.BamReader(str(q))
    assert e.value.code == bam.ERR_FORMAT
    # cut in the middle of the records: reading stops with an error
    q.write_bytes(bytes(raw[:len(raw) // 2]))
    with pytest.raises(bam.BamError):
        list(bam.BamReader(str(q)))
    with pytest.raises(bam.BamError) as e:
        bam.BamReader(str(tmp_path / "missing.bam"))
    assert e.value.code == bam.ERR_IO


def test_read_inputs():
    L = 12
    rec = bam.BamRecord(name="x", tid=0, pos=10, cigar=[("H", 3), ("S", 2), ("M", 7), ("S", 3)],
                        seq=b"ACGTRNacgtAC", qual=bytes(range(10, 10 + L)),
                        aux=bam.BamRecord.aux_z("BD", "".join(chr(33 + 40 + k) for k in range(L))) +
                        b"NMi" + struct.pack("<i", 2) +
                        bam.BamRecord.aux_z("BI", "".join(chr(33 + 20 + k) for k in range(L))))
    b, q, ins, dl, gcp = bam.read_inputs(rec, clip_soft=True)
    assert b.tobytes() == b"GTNNACG"          # soft clips off, R and N -> N, lowercase read as bases
    assert list(q) == list(range(12, 19))
    assert list(ins) == list(range(22, 29))
    assert list(dl) == list(range(42, 49))
    assert list(gcp) == [10] * 7
    b, q, ins, dl, gcp = bam.read_inputs(rec, clip_soft=False)
    assert len(b) == L and b.tobytes() == b"ACGTNNACGTAC"
    # no qualities (0xFF), no BI/BD: the defaults
    rec2 = bam.BamRecord(name="y", tid=0, pos=1, cigar=[("M", 4)], seq=b"ACGT", qual=b"\xff" * 4)
    b, q, ins, dl, gcp = bam.read_inputs(rec2, default_q=17, default_ins=44, default_del=43, default_gcp=9)
    assert list(q) == [17] * 4 and list(ins) == [44] * 4 and list(dl) == [43] * 4 and list(gcp) == [9] * 4


def test_region_batch(tmp_path):
    rng = np.random.default_rng(8)
    recs = random_records(rng, 500)
    p = tmp_path / "b.bam"
    write_file(p, recs)
    rd = bam.BamReader(str(p))
    haps = [b"ACGT" * 60, b"ACGTT" * 50]
    batch, used = bam.region_batch(rd, "chr1:1-100000", haps)
    exp = [r for r in recs if r.tid == 0 and r.pos < 100000 and expected_end(r) > 0 and
           not (r.flag & bam.SKIP_FLAGS)]
    assert [r.name for r in used] == [r.name for r in exp]
    assert batch.num_haps == 2 and len(batch.read_len) == len(used)
    for k, r in enumerate(used):
        b, q, ins, dl, gcp = bam.read_inputs(r)
        assert batch.read(k) == b.tobytes()
        bq, bi, bd, bg = batch.quals(k)
        assert np.array_equal(bq, q) and np.array_equal(bi, ins) and np.array_equal(bd, dl)
    rd.close()
