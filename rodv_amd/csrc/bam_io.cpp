// bam_io.cpp -- BGZF/BAM reader and writer with BAI region queries over zlib
// (include/rodv_bam.h).  Host code only.  The formats follow the SAM
// specification (BGZF: section 4.1; BAM records: 4.2; BAI: 5.2); the
// interface replaces the reference's htslib-backed BamHandle
// (src/iostream/bam_handle.h:6-88).
#include "../../include/rodv_bam.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

// the C ABI is the library's only export (it is built with hidden visibility)
#define RODV_BAM_EXPORT __attribute__((visibility("default")))

namespace {

constexpr size_t kBlockMax = 65536;      // a BGZF block's uncompressed and compressed limit
constexpr size_t kWriteChunk = 0xff00;   // uncompressed bytes per written block
// the empty block that ends a BGZF file (SAM specification 4.1.2)
const uint8_t kEof[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43, 2, 0,
                          0x1b, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};
const char kSeqCodes[] = "=ACMGRSVTWYHKDBN";

uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | p[1] << 8); }
uint32_t rd32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint64_t rd64(const uint8_t *p) { return (uint64_t)rd32(p) | (uint64_t)rd32(p + 4) << 32; }
void put16(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)x);
    v.push_back((uint8_t)(x >> 8));
}
void put32(std::vector<uint8_t> &v, uint32_t x) {
    for (int k = 0; k < 4; k++) v.push_back((uint8_t)(x >> (8 * k)));
}
void put64(std::vector<uint8_t> &v, uint64_t x) {
    put32(v, (uint32_t)x);
    put32(v, (uint32_t)(x >> 32));
}

// BAI bins: the smallest bin containing [beg, end) and every bin that
// overlaps it (SAM specification 5.3), levels of 2^29 .. 2^14 bases.
const int kBinShift[6] = {29, 26, 23, 20, 17, 14};
const int kBinFirst[6] = {0, 1, 9, 73, 585, 4681};
int reg2bin(int64_t beg, int64_t end) {
    --end;
    for (int lv = 5; lv >= 1; lv--)
        if ((beg >> kBinShift[lv]) == (end >> kBinShift[lv])) return kBinFirst[lv] + (int)(beg >> kBinShift[lv]);
    return 0;
}
void reg2bins(int64_t beg, int64_t end, std::vector<int> &out) {
    out.clear();
    --end;
    for (int lv = 0; lv < 6; lv++)
        for (int64_t k = beg >> kBinShift[lv]; k <= end >> kBinShift[lv]; k++) out.push_back(kBinFirst[lv] + (int)k);
}
constexpr int kPseudoBin = 37450;  // htslib's per-reference metadata bin

// reference bases consumed by a CIGAR (M, D, N, =, X)
int64_t ref_span(const uint32_t *cigar, uint32_t n) {
    int64_t s = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t op = cigar[k] & 0xf;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) s += cigar[k] >> 4;
    }
    return s;
}

struct Chunk {
    uint64_t beg, end;  // virtual offsets
};

struct RefIndex {
    std::map<int, std::vector<Chunk>> bins;
    std::vector<uint64_t> linear;  // per 16 kb window, the least virtual offset of a record overlapping it
};

}  // namespace

struct rodv_bam {
    bool writing = false;
    std::string path;
    FILE *f = nullptr;
    // ---- BGZF state
    std::vector<uint8_t> cbuf, ubuf;  // current block (compressed, uncompressed)
    uint64_t block_off = 0;           // file offset of the current block
    uint64_t next_off = 0;            // file offset of the block after it
    size_t upos = 0;                  // read position inside ubuf
    int level = -1;
    // ---- header
    std::string text;
    std::vector<std::string> names;
    std::vector<int64_t> lens;
    uint64_t first_record = 0;  // virtual offset of the first record
    // ---- reading
    std::vector<uint8_t> rec;       // the current record's bytes
    std::vector<uint32_t> cig;      // its CIGAR, aligned
    std::vector<uint8_t> seq;       // its decoded bases
    std::vector<RefIndex> index;    // loaded BAI (empty: none)
    bool have_index = false;
    // query: merged chunks and the intervals a record must overlap
    bool query = false;
    std::vector<Chunk> chunks;
    size_t chunk_k = 0;
    struct Iv {
        int32_t tid;
        int64_t beg, end;
    };
    std::vector<Iv> ivs;
    bool done = false;
    // ---- writing: index built as records are written
    bool sorted = true;
    bool seen_unplaced = false;
    int32_t last_tid = -1;
    int64_t last_pos = -1;
    uint64_t n_no_coor = 0;
    struct RefStat {
        uint64_t beg = ~0ull, end = 0, mapped = 0, unmapped = 0;
    };
    std::vector<RefStat> stats;
};

namespace {

// ---------------------------------------------------------------- BGZF read
uint64_t vtell(const rodv_bam *b) { return b->block_off << 16 | (uint64_t)b->upos; }

// Loads the block at file offset `off`; 0 at the end of the file, 1 loaded, <0 error.
int load_block(rodv_bam *b, uint64_t off) {
    if (fseeko(b->f, (off_t)off, SEEK_SET) != 0) return RODV_BAM_ERR_IO;
    uint8_t h[18];
    const size_t got = fread(h, 1, 12, b->f);
    if (got == 0) {
        b->block_off = off;
        b->next_off = off;
        b->ubuf.clear();
        b->upos = 0;
        return 0;
    }
    if (got != 12 || h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || !(h[3] & 4)) return RODV_BAM_ERR_FORMAT;
    const uint32_t xlen = rd16(h + 10);
    std::vector<uint8_t> extra(xlen);
    if (fread(extra.data(), 1, xlen, b->f) != xlen) return RODV_BAM_ERR_FORMAT;
    int64_t bsize = -1;
    for (uint32_t k = 0; k + 4 <= xlen;) {
        const uint32_t slen = rd16(&extra[k + 2]);
        if (extra[k] == 66 && extra[k + 1] == 67 && slen == 2 && k + 6 <= xlen) bsize = rd16(&extra[k + 4]);
        k += 4 + slen;
    }
    if (bsize < 0) return RODV_BAM_ERR_FORMAT;
    const int64_t rest = bsize + 1 - 12 - (int64_t)xlen;  // CDATA + CRC32 + ISIZE
    if (rest < 8) return RODV_BAM_ERR_FORMAT;
    b->cbuf.resize((size_t)rest);
    if (fread(b->cbuf.data(), 1, (size_t)rest, b->f) != (size_t)rest) return RODV_BAM_ERR_FORMAT;
    const uint32_t crc = rd32(&b->cbuf[(size_t)rest - 8]), isize = rd32(&b->cbuf[(size_t)rest - 4]);
    if (isize > kBlockMax) return RODV_BAM_ERR_FORMAT;
    b->ubuf.resize(isize);
    if (isize) {
        z_stream z;
        std::memset(&z, 0, sizeof(z));
        if (inflateInit2(&z, -15) != Z_OK) return RODV_BAM_ERR_FORMAT;
        z.next_in = b->cbuf.data();
        z.avail_in = (uInt)(rest - 8);
        z.next_out = b->ubuf.data();
        z.avail_out = isize;
        const int zr = inflate(&z, Z_FINISH);
        inflateEnd(&z);
        if (zr != Z_STREAM_END || z.avail_out != 0) return RODV_BAM_ERR_FORMAT;
    }
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), b->ubuf.data(), isize) != crc) return RODV_BAM_ERR_FORMAT;
    b->block_off = off;
    b->next_off = off + (uint64_t)bsize + 1;
    b->upos = 0;
    return 1;
}

int vseek(rodv_bam *b, uint64_t voff) {
    const uint64_t coff = voff >> 16;
    if (coff != b->block_off || b->ubuf.empty()) {
        const int r = load_block(b, coff);
        if (r < 0) return r;
    }
    if ((voff & 0xffff) > b->ubuf.size()) return RODV_BAM_ERR_FORMAT;
    b->upos = (size_t)(voff & 0xffff);
    return RODV_BAM_OK;
}

// Reads n bytes; returns n, 0 at a clean end of file before any byte, or <0.
int64_t bread(rodv_bam *b, void *dst, size_t n) {
    uint8_t *d = static_cast<uint8_t *>(dst);
    size_t done = 0;
    while (done < n) {
        if (b->upos >= b->ubuf.size()) {
            // the next block (empty blocks, the end marker among them, are skipped)
            const int r = load_block(b, b->next_off);
            if (r < 0) return r;
            if (r == 0) return done == 0 ? 0 : RODV_BAM_ERR_FORMAT;
            continue;
        }
        const size_t k = std::min(n - done, b->ubuf.size() - b->upos);
        std::memcpy(d + done, b->ubuf.data() + b->upos, k);
        b->upos += k;
        done += k;
    }
    return (int64_t)n;
}

// ---------------------------------------------------------------- BGZF write
int flush_block(rodv_bam *b) {
    if (b->ubuf.empty()) return RODV_BAM_OK;
    const size_t n = b->ubuf.size();
    std::vector<uint8_t> out(kBlockMax);
    z_stream z;
    std::memset(&z, 0, sizeof(z));
    if (deflateInit2(&z, b->level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return RODV_BAM_ERR_IO;
    z.next_in = b->ubuf.data();
    z.avail_in = (uInt)n;
    z.next_out = out.data() + 18;
    z.avail_out = (uInt)(kBlockMax - 18 - 8);
    const int zr = deflate(&z, Z_FINISH);
    const size_t clen = (size_t)z.total_out;
    deflateEnd(&z);
    if (zr != Z_STREAM_END) return RODV_BAM_ERR_IO;  // (kWriteChunk keeps even stored data in bounds)
    const size_t total = 18 + clen + 8;
    const uint8_t hdr[18] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 66, 67, 2, 0,
                             (uint8_t)((total - 1) & 0xff), (uint8_t)((total - 1) >> 8)};
    std::memcpy(out.data(), hdr, 18);
    const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), b->ubuf.data(), (uInt)n);
    uint8_t *t = out.data() + 18 + clen;
    for (int k = 0; k < 4; k++) t[k] = (uint8_t)(crc >> (8 * k));
    for (int k = 0; k < 4; k++) t[4 + k] = (uint8_t)((uint32_t)n >> (8 * k));
    if (fwrite(out.data(), 1, total, b->f) != total) return RODV_BAM_ERR_IO;
    b->block_off += total;
    b->ubuf.clear();
    return RODV_BAM_OK;
}

uint64_t wtell(const rodv_bam *b) { return b->block_off << 16 | (uint64_t)b->ubuf.size(); }

int bwrite(rodv_bam *b, const void *src, size_t n) {
    const uint8_t *s = static_cast<const uint8_t *>(src);
    while (n) {
        const size_t k = std::min(n, kWriteChunk - b->ubuf.size());
        b->ubuf.insert(b->ubuf.end(), s, s + k);
        s += k;
        n -= k;
        if (b->ubuf.size() == kWriteChunk) {
            const int r = flush_block(b);
            if (r) return r;
        }
    }
    return RODV_BAM_OK;
}

// ---------------------------------------------------------------- records
// Decodes b->rec (one record without its block_size) into *out.
int decode(rodv_bam *b, rodv_bam_record *out) {
    const std::vector<uint8_t> &r = b->rec;
    if (r.size() < 32) return RODV_BAM_ERR_FORMAT;
    out->tid = (int32_t)rd32(&r[0]);
    out->pos = (int32_t)rd32(&r[4]);
    const uint32_t l_name = r[8];
    out->mapq = r[9];
    out->n_cigar = rd16(&r[12]);
    out->flag = rd16(&r[14]);
    out->l_seq = rd32(&r[16]);
    out->next_tid = (int32_t)rd32(&r[20]);
    out->next_pos = (int32_t)rd32(&r[24]);
    out->tlen = (int32_t)rd32(&r[28]);
    const uint64_t need = 32ull + l_name + 4ull * out->n_cigar + (out->l_seq + 1ull) / 2 + out->l_seq;
    if (l_name == 0 || need > r.size() || r[32 + l_name - 1] != 0) return RODV_BAM_ERR_FORMAT;
    out->name = reinterpret_cast<const char *>(&r[32]);
    size_t o = 32 + l_name;  // the CIGAR follows the name, at any alignment
    b->cig.resize(out->n_cigar);
    for (uint32_t k = 0; k < out->n_cigar; k++) b->cig[k] = rd32(&r[o + 4ull * k]);
    out->cigar = b->cig.data();
    o += 4ull * out->n_cigar;
    b->seq.resize(out->l_seq + 1);
    for (uint32_t k = 0; k < out->l_seq; k++) {
        const uint8_t byte = r[o + k / 2];
        b->seq[k] = (uint8_t)kSeqCodes[(k & 1) ? (byte & 0xf) : (byte >> 4)];
    }
    b->seq[out->l_seq] = 0;
    out->seq = b->seq.data();
    o += (out->l_seq + 1) / 2;
    out->qual = &r[o];
    o += out->l_seq;
    out->aux = &r[o];
    out->l_aux = (uint32_t)(r.size() - o);
    const int64_t span = ref_span(out->cigar, out->n_cigar);
    out->end = out->pos < 0 ? out->pos : (int32_t)(out->pos + (span > 0 ? span : 1));
    return RODV_BAM_OK;
}

// Reads the next record's bytes at the current position: 1, 0 at the end, <0.
int read_record(rodv_bam *b) {
    uint8_t len4[4];
    const int64_t r = bread(b, len4, 4);
    if (r <= 0) return (int)r;
    const uint32_t n = rd32(len4);
    if (n < 32 || n > (1u << 28)) return RODV_BAM_ERR_FORMAT;
    b->rec.resize(n);
    if (bread(b, b->rec.data(), n) != (int64_t)n) return RODV_BAM_ERR_FORMAT;
    return 1;
}

bool overlaps(const rodv_bam *b, const rodv_bam_record &r) {
    for (const auto &iv : b->ivs)
        if (r.tid == iv.tid && r.pos < iv.end && r.end > iv.beg) return true;
    return false;
}

int parse_region(const rodv_bam *b, const char *region, int32_t *tid, int64_t *beg, int64_t *end) {
    if (!region || !*region) return RODV_BAM_ERR_REGION;
    const std::string s(region);
    auto find = [&](const std::string &name) -> int32_t {
        for (size_t k = 0; k < b->names.size(); k++)
            if (b->names[k] == name) return (int32_t)k;
        return -1;
    };
    int32_t t = find(s);
    if (t >= 0) {  // the whole reference
        *tid = t;
        *beg = 0;
        *end = b->lens[(size_t)t];
        return RODV_BAM_OK;
    }
    const size_t colon = s.rfind(':');
    if (colon == std::string::npos || (t = find(s.substr(0, colon))) < 0) return RODV_BAM_ERR_REGION;
    std::string rest;
    for (char c : s.substr(colon + 1))
        if (c != ',') rest.push_back(c);
    const size_t dash = rest.find('-');
    char *e = nullptr;
    const long long a = std::strtoll(rest.c_str(), &e, 10);
    if (e == rest.c_str() || a < 1) return RODV_BAM_ERR_REGION;
    long long z = (long long)b->lens[(size_t)t];
    if (dash != std::string::npos) {
        const std::string zs = rest.substr(dash + 1);
        z = std::strtoll(zs.c_str(), &e, 10);
        if (e == zs.c_str() || *e) return RODV_BAM_ERR_REGION;
    } else if (*e) {
        return RODV_BAM_ERR_REGION;
    }
    if (z < a) return RODV_BAM_ERR_REGION;
    *tid = t;
    *beg = a - 1;  // 1-based inclusive -> 0-based half-open
    *end = z;
    return RODV_BAM_OK;
}

// Chunks of the index that may hold records overlapping the intervals.
int start_query(rodv_bam *b) {
    if (!b->have_index) return RODV_BAM_ERR_NO_INDEX;
    std::vector<Chunk> cs;
    std::vector<int> bins;
    for (const auto &iv : b->ivs) {
        if (iv.tid < 0 || (size_t)iv.tid >= b->index.size() || iv.end <= iv.beg) continue;
        const RefIndex &ri = b->index[(size_t)iv.tid];
        const size_t w = (size_t)(iv.beg >> 14);
        const uint64_t min_off = ri.linear.empty() ? 0 : ri.linear[std::min(w, ri.linear.size() - 1)];
        reg2bins(iv.beg, iv.end, bins);
        for (int bn : bins) {
            const auto it = ri.bins.find(bn);
            if (it == ri.bins.end()) continue;
            for (const Chunk &c : it->second)
                if (c.end > min_off) cs.push_back(c);
        }
    }
    std::sort(cs.begin(), cs.end(), [](const Chunk &x, const Chunk &y) { return x.beg < y.beg; });
    b->chunks.clear();
    for (const Chunk &c : cs) {
        if (!b->chunks.empty() && c.beg <= b->chunks.back().end)
            b->chunks.back().end = std::max(b->chunks.back().end, c.end);
        else
            b->chunks.push_back(c);
    }
    b->chunk_k = 0;
    b->query = true;
    b->done = b->chunks.empty();
    if (!b->done) return vseek(b, b->chunks[0].beg);
    return RODV_BAM_OK;
}

}  // namespace

extern "C" {

RODV_BAM_EXPORT int rodv_bam_open(const char *path, rodv_bam **out) {
    if (!path || !out) return RODV_BAM_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<rodv_bam> b(new rodv_bam);
    b->path = path;
    b->f = std::fopen(path, "rb");
    if (!b->f) return RODV_BAM_ERR_IO;
    int r = load_block(b.get(), 0);
    if (r <= 0) {
        std::fclose(b->f);
        return r < 0 ? r : RODV_BAM_ERR_FORMAT;
    }
    uint8_t h[8];
    if (bread(b.get(), h, 8) != 8 || std::memcmp(h, "BAM\1", 4) != 0) {
        std::fclose(b->f);
        return RODV_BAM_ERR_FORMAT;
    }
    const uint32_t l_text = rd32(h + 4);
    b->text.resize(l_text);
    uint8_t n4[4];
    bool ok = (l_text == 0 || bread(b.get(), &b->text[0], l_text) == (int64_t)l_text) && bread(b.get(), n4, 4) == 4;
    const uint32_t n_ref = ok ? rd32(n4) : 0;
    if (ok && n_ref > (1u << 24)) ok = false;
    for (uint32_t k = 0; ok && k < n_ref; k++) {
        uint8_t l4[4];
        ok = bread(b.get(), l4, 4) == 4;
        const uint32_t ln = ok ? rd32(l4) : 0;
        ok = ok && ln >= 1 && ln < (1u << 20);
        std::string nm(ln, '\0');
        ok = ok && bread(b.get(), &nm[0], ln) == (int64_t)ln;
        ok = ok && bread(b.get(), l4, 4) == 4;
        if (ok) {
            nm.resize(ln - 1);
            b->names.push_back(nm);
            b->lens.push_back((int64_t)rd32(l4));
        }
    }
    if (!ok) {
        std::fclose(b->f);
        return RODV_BAM_ERR_FORMAT;
    }
    b->text.resize(std::strlen(b->text.c_str()));
    b->first_record = vtell(b.get());
    // the index next to the file, as BamHandle's load_index does by default
    (void)rodv_bam_load_index(b.get(), nullptr);
    *out = b.release();
    return RODV_BAM_OK;
}

RODV_BAM_EXPORT int rodv_bam_load_index(rodv_bam *b, const char *bai_path) {
    if (!b || b->writing) return RODV_BAM_ERR_ARG;
    const std::string p = bai_path ? std::string(bai_path) : b->path + ".bai";
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return RODV_BAM_ERR_NO_INDEX;
    std::vector<uint8_t> d;
    uint8_t tmp[65536];
    size_t n;
    while ((n = fread(tmp, 1, sizeof(tmp), f)) > 0) d.insert(d.end(), tmp, tmp + n);
    std::fclose(f);
    size_t o = 0;
    auto need = [&](size_t k) { return o + k <= d.size(); };
    if (!need(8) || std::memcmp(d.data(), "BAI\1", 4) != 0) return RODV_BAM_ERR_FORMAT;
    o = 4;
    const uint32_t n_ref = rd32(&d[o]);
    o += 4;
    std::vector<RefIndex> idx(n_ref);
    for (uint32_t t = 0; t < n_ref; t++) {
        if (!need(4)) return RODV_BAM_ERR_FORMAT;
        const uint32_t n_bin = rd32(&d[o]);
        o += 4;
        for (uint32_t k = 0; k < n_bin; k++) {
            if (!need(8)) return RODV_BAM_ERR_FORMAT;
            const uint32_t bin = rd32(&d[o]), n_chunk = rd32(&d[o + 4]);
            o += 8;
            if (!need(16ull * n_chunk)) return RODV_BAM_ERR_FORMAT;
            std::vector<Chunk> cs(n_chunk);
            for (uint32_t c = 0; c < n_chunk; c++) cs[c] = {rd64(&d[o + 16ull * c]), rd64(&d[o + 16ull * c + 8])};
            o += 16ull * n_chunk;
            if (bin != (uint32_t)kPseudoBin) idx[t].bins[(int)bin] = std::move(cs);
        }
        if (!need(4)) return RODV_BAM_ERR_FORMAT;
        const uint32_t n_intv = rd32(&d[o]);
        o += 4;
        if (!need(8ull * n_intv)) return RODV_BAM_ERR_FORMAT;
        idx[t].linear.resize(n_intv);
        for (uint32_t k = 0; k < n_intv; k++) idx[t].linear[k] = rd64(&d[o + 8ull * k]);
        o += 8ull * n_intv;
    }
    b->index = std::move(idx);
    b->have_index = true;
    return RODV_BAM_OK;
}

RODV_BAM_EXPORT int rodv_bam_num_targets(const rodv_bam *b) { return b ? (int)b->names.size() : RODV_BAM_ERR_ARG; }

RODV_BAM_EXPORT int rodv_bam_target(const rodv_bam *b, int32_t tid, const char **name, int64_t *len) {
    if (!b || tid < 0 || (size_t)tid >= b->names.size()) return RODV_BAM_ERR_ARG;
    if (name) *name = b->names[(size_t)tid].c_str();
    if (len) *len = b->lens[(size_t)tid];
    return RODV_BAM_OK;
}

RODV_BAM_EXPORT const char *rodv_bam_header_text(const rodv_bam *b) { return b ? b->text.c_str() : nullptr; }

RODV_BAM_EXPORT int rodv_bam_next(rodv_bam *b, rodv_bam_record *rec) {
    if (!b || !rec || b->writing) return RODV_BAM_ERR_ARG;
    for (;;) {
        if (b->done) return 0;
        if (b->query && vtell(b) >= b->chunks[b->chunk_k].end) {
            if (++b->chunk_k >= b->chunks.size()) {
                b->done = true;
                return 0;
            }
            const int r = vseek(b, b->chunks[b->chunk_k].beg);
            if (r) return r;
        }
        const int r = read_record(b);
        if (r <= 0) {
            b->done = true;
            return r;
        }
        const int d = decode(b, rec);
        if (d) return d;
        if (!b->query || overlaps(b, *rec)) return 1;
    }
}

RODV_BAM_EXPORT int rodv_bam_query(rodv_bam *b, int32_t tid, int64_t beg, int64_t end) {
    if (!b || b->writing) return RODV_BAM_ERR_ARG;
    if (tid < 0 || (size_t)tid >= b->names.size() || beg < 0 || end <= beg) return RODV_BAM_ERR_REGION;
    b->ivs = {{tid, beg, end}};
    return start_query(b);
}

RODV_BAM_EXPORT int rodv_bam_query_region(rodv_bam *b, const char *region) {
    if (!b || b->writing) return RODV_BAM_ERR_ARG;
    int32_t tid;
    int64_t beg, end;
    const int r = parse_region(b, region, &tid, &beg, &end);
    if (r) return r;
    b->ivs = {{tid, beg, end}};
    return start_query(b);
}

RODV_BAM_EXPORT int rodv_bam_query_multi(rodv_bam *b, const char *const *regions, int num_regions) {
    if (!b || b->writing || num_regions < 0 || (num_regions && !regions)) return RODV_BAM_ERR_ARG;
    std::vector<rodv_bam::Iv> ivs;
    for (int k = 0; k < num_regions; k++) {
        int32_t tid;
        int64_t beg, end;
        const int r = parse_region(b, regions[k], &tid, &beg, &end);
        if (r) return r;
        ivs.push_back({tid, beg, end});
    }
    b->ivs = std::move(ivs);
    return start_query(b);
}

RODV_BAM_EXPORT int rodv_bam_create(const char *path, const char *header_text, int32_t num_targets,
                    const char *const *target_names, const int64_t *target_lens, int level,
                    rodv_bam **out) {
    if (!path || !out || num_targets < 0 || (num_targets && (!target_names || !target_lens)) ||
        level < -1 || level > 9)
        return RODV_BAM_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<rodv_bam> b(new rodv_bam);
    b->writing = true;
    b->path = path;
    b->level = level;
    for (int32_t k = 0; k < num_targets; k++) {
        if (!target_names[k] || target_lens[k] < 0 || target_lens[k] > INT32_MAX) return RODV_BAM_ERR_ARG;
        b->names.emplace_back(target_names[k]);
        b->lens.push_back(target_lens[k]);
    }
    b->text = header_text ? header_text : "";
    b->f = std::fopen(path, "wb");
    if (!b->f) return RODV_BAM_ERR_IO;
    std::vector<uint8_t> h = {'B', 'A', 'M', 1};
    put32(h, (uint32_t)b->text.size());
    h.insert(h.end(), b->text.begin(), b->text.end());
    put32(h, (uint32_t)num_targets);
    for (int32_t k = 0; k < num_targets; k++) {
        put32(h, (uint32_t)b->names[(size_t)k].size() + 1);
        h.insert(h.end(), b->names[(size_t)k].begin(), b->names[(size_t)k].end());
        h.push_back(0);
        put32(h, (uint32_t)b->lens[(size_t)k]);
    }
    int r = bwrite(b.get(), h.data(), h.size());
    // records start in a block of their own (as htslib writes the header)
    if (!r) r = flush_block(b.get());
    if (r) {
        std::fclose(b->f);
        return r;
    }
    b->index.assign((size_t)num_targets, RefIndex());
    b->stats.assign((size_t)num_targets, rodv_bam::RefStat());
    *out = b.release();
    return RODV_BAM_OK;
}

RODV_BAM_EXPORT int rodv_bam_write(rodv_bam *b, const rodv_bam_record *rec) {
    if (!b || !b->writing || !rec || !rec->name) return RODV_BAM_ERR_ARG;
    const size_t l_name = std::strlen(rec->name) + 1;
    if (l_name > 255 || rec->n_cigar > 0xffff || (rec->n_cigar && !rec->cigar) ||
        (rec->l_seq && (!rec->seq || !rec->qual)) || (rec->l_aux && !rec->aux))
        return RODV_BAM_ERR_ARG;
    if (rec->tid >= (int32_t)b->names.size() || rec->tid < -1) return RODV_BAM_ERR_ARG;
    const int64_t span = ref_span(rec->cigar, rec->n_cigar);
    const int64_t end = rec->pos < 0 ? rec->pos + 1 : rec->pos + (span > 0 ? span : 1);
    std::vector<uint8_t> r;
    r.reserve(36 + l_name + 4 * rec->n_cigar + rec->l_seq * 2 + rec->l_aux);
    put32(r, 0);  // block_size, set below
    put32(r, (uint32_t)rec->tid);
    put32(r, (uint32_t)rec->pos);
    r.push_back((uint8_t)l_name);
    r.push_back((uint8_t)rec->mapq);
    put16(r, rec->pos < 0 ? 4680u : (uint32_t)reg2bin(rec->pos, end));
    put16(r, rec->n_cigar);
    put16(r, rec->flag);
    put32(r, rec->l_seq);
    put32(r, (uint32_t)rec->next_tid);
    put32(r, (uint32_t)rec->next_pos);
    put32(r, (uint32_t)rec->tlen);
    r.insert(r.end(), rec->name, rec->name + l_name);
    for (uint32_t k = 0; k < rec->n_cigar; k++) put32(r, rec->cigar[k]);
    for (uint32_t k = 0; k < rec->l_seq; k += 2) {
        auto code = [&](uint8_t c) -> uint8_t {
            const char *p = std::strchr(kSeqCodes, std::toupper(c));
            return (p && c) ? (uint8_t)(p - kSeqCodes) : 15;
        };
        const uint8_t hi = code(rec->seq[k]), lo = k + 1 < rec->l_seq ? code(rec->seq[k + 1]) : 0;
        r.push_back((uint8_t)(hi << 4 | lo));
    }
    r.insert(r.end(), rec->qual, rec->qual + rec->l_seq);
    if (rec->l_aux) r.insert(r.end(), rec->aux, rec->aux + rec->l_aux);
    const uint32_t bs = (uint32_t)(r.size() - 4);
    for (int k = 0; k < 4; k++) r[(size_t)k] = (uint8_t)(bs >> (8 * k));
    // index: coordinate order is (tid, pos) ascending with unplaced records last
    if (rec->tid < 0)
        b->seen_unplaced = true;
    else if (b->seen_unplaced || rec->tid < b->last_tid || (rec->tid == b->last_tid && rec->pos < b->last_pos))
        b->sorted = false;
    const uint64_t vbeg = wtell(b);
    const int w = bwrite(b, r.data(), r.size());
    if (w) return w;
    const uint64_t vend = wtell(b);
    if (rec->tid >= 0) {
        b->last_tid = rec->tid;
        b->last_pos = rec->pos;
    }
    if (rec->tid >= 0 && rec->pos >= 0 && b->sorted) {
        RefIndex &ri = b->index[(size_t)rec->tid];
        std::vector<Chunk> &cs = ri.bins[reg2bin(rec->pos, end)];
        if (!cs.empty() && cs.back().end == vbeg)
            cs.back().end = vend;
        else
            cs.push_back({vbeg, vend});
        const size_t w0 = (size_t)(rec->pos >> 14), w1 = (size_t)((end - 1) >> 14);
        if (ri.linear.size() <= w1) ri.linear.resize(w1 + 1, ~0ull);
        for (size_t k = w0; k <= w1; k++) ri.linear[k] = std::min(ri.linear[k], vbeg);
        rodv_bam::RefStat &s = b->stats[(size_t)rec->tid];
        s.beg = std::min(s.beg, vbeg);
        s.end = std::max(s.end, vend);
        if (rec->flag & 4)
            s.unmapped++;
        else
            s.mapped++;
    } else if (rec->tid < 0) {
        b->n_no_coor++;
    }
    return RODV_BAM_OK;
}

RODV_BAM_EXPORT int rodv_bam_close(rodv_bam *b) {
    if (!b) return RODV_BAM_ERR_ARG;
    int rc = RODV_BAM_OK;
    if (b->writing) {
        rc = flush_block(b);
        if (!rc && fwrite(kEof, 1, sizeof(kEof), b->f) != sizeof(kEof)) rc = RODV_BAM_ERR_IO;
        if (std::fclose(b->f) != 0 && !rc) rc = RODV_BAM_ERR_IO;
        b->f = nullptr;
        if (!rc && b->sorted) {
            // BAI: per reference its bins' chunks, htslib's metadata bin, and
            // the linear index (windows without records take the last value
            // before them, a lower bound for the records that follow)
            std::vector<uint8_t> d = {'B', 'A', 'I', 1};
            put32(d, (uint32_t)b->index.size());
            for (size_t t = 0; t < b->index.size(); t++) {
                RefIndex &ri = b->index[t];
                const rodv_bam::RefStat &s = b->stats[t];
                const bool any = s.mapped + s.unmapped > 0;
                put32(d, (uint32_t)(ri.bins.size() + (any ? 1 : 0)));
                for (const auto &e : ri.bins) {
                    put32(d, (uint32_t)e.first);
                    put32(d, (uint32_t)e.second.size());
                    for (const Chunk &c : e.second) {
                        put64(d, c.beg);
                        put64(d, c.end);
                    }
                }
                if (any) {
                    put32(d, (uint32_t)kPseudoBin);
                    put32(d, 2);
                    put64(d, s.beg);
                    put64(d, s.end);
                    put64(d, s.mapped);
                    put64(d, s.unmapped);
                }
                uint64_t prev = 0;
                for (uint64_t &v : ri.linear) {
                    if (v == ~0ull) v = prev;
                    prev = v;
                }
                put32(d, (uint32_t)ri.linear.size());
                for (uint64_t v : ri.linear) put64(d, v);
            }
            put64(d, b->n_no_coor);
            FILE *f = std::fopen((b->path + ".bai").c_str(), "wb");
            if (!f || fwrite(d.data(), 1, d.size(), f) != d.size()) rc = RODV_BAM_ERR_IO;
            if (f && std::fclose(f) != 0) rc = RODV_BAM_ERR_IO;
        }
    } else if (b->f) {
        std::fclose(b->f);
    }
    delete b;
    return rc;
}

RODV_BAM_EXPORT const char *rodv_bam_aux_z(const rodv_bam_record *rec, const char tag[2]) {
    if (!rec || !tag) return nullptr;
    const uint8_t *p = rec->aux, *e = rec->aux + rec->l_aux;
    auto fixed = [](uint8_t t) -> int {
        switch (t) {
        case 'A': case 'c': case 'C': return 1;
        case 's': case 'S': return 2;
        case 'i': case 'I': case 'f': return 4;
        default: return 0;
        }
    };
    while (p + 3 <= e) {
        const bool hit = p[0] == (uint8_t)tag[0] && p[1] == (uint8_t)tag[1];
        const uint8_t t = p[2];
        p += 3;
        if (t == 'Z' || t == 'H') {
            const uint8_t *z = static_cast<const uint8_t *>(std::memchr(p, 0, (size_t)(e - p)));
            if (!z) return nullptr;
            if (hit && t == 'Z') return reinterpret_cast<const char *>(p);
            p = z + 1;
        } else if (t == 'B') {
            if (p + 5 > e) return nullptr;
            const int sz = fixed(p[0]);
            if (!sz) return nullptr;
            p += 5 + (size_t)sz * rd32(p + 1);
        } else {
            const int sz = fixed(t);
            if (!sz) return nullptr;
            p += sz;
        }
    }
    return nullptr;
}

RODV_BAM_EXPORT int rodv_bam_read_inputs(const rodv_bam_record *rec, int clip_soft, uint8_t default_q,
                         uint8_t default_ins, uint8_t default_del, uint8_t default_gcp,
                         uint8_t *bases, uint8_t *q, uint8_t *ins, uint8_t *del, uint8_t *gcp,
                         uint32_t *len) {
    if (!rec || !bases || !q || !ins || !del || !gcp || !len) return RODV_BAM_ERR_ARG;
    uint32_t b0 = 0, b1 = rec->l_seq;  // bases [b0, b1) kept
    if (clip_soft && rec->n_cigar) {
        // the first and last operations other than hard clips (H)
        int64_t k = 0, j = (int64_t)rec->n_cigar - 1;
        while (k <= j && (rec->cigar[k] & 0xf) == 5) k++;
        while (j >= k && (rec->cigar[j] & 0xf) == 5) j--;
        if (k <= j && (rec->cigar[k] & 0xf) == 4) b0 = std::min(b1, rec->cigar[k] >> 4);
        if (j > k && (rec->cigar[j] & 0xf) == 4) b1 -= std::min(b1 - b0, rec->cigar[j] >> 4);
    }
    const char *bi = rodv_bam_aux_z(rec, "BI"), *bd = rodv_bam_aux_z(rec, "BD");
    if (bi && std::strlen(bi) != rec->l_seq) bi = nullptr;
    if (bd && std::strlen(bd) != rec->l_seq) bd = nullptr;
    const bool noq = rec->l_seq && rec->qual[0] == 0xff;
    uint32_t n = 0;
    for (uint32_t k = b0; k < b1; k++, n++) {
        const uint8_t c = (uint8_t)std::toupper(rec->seq[k]);
        bases[n] = (c == 'A' || c == 'C' || c == 'G' || c == 'T') ? c : 'N';
        q[n] = noq ? default_q : rec->qual[k];
        ins[n] = bi ? (uint8_t)(bi[k] - 33) : default_ins;
        del[n] = bd ? (uint8_t)(bd[k] - 33) : default_del;
        gcp[n] = default_gcp;
    }
    *len = n;
    return RODV_BAM_OK;
}

}  // extern "C"
