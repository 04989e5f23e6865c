// phmm_internal.h -- types shared by the host planner and the gfx950 kernels.
#pragma once

#include <stdint.h>

namespace rodv {

// Qualities are masked with & 127 (pairhmm_impl.cpp:89-92), so the
// triangular match-to-match table is only read up to index 127*128/2 + 127.
constexpr int kMMQ = 128 * 129 / 2;

// Prior tables in device memory, built on the host exactly as
// Context<float>/Context<double> build them (context.h:90-166).
struct DevTables {
    float ph2pr_f[128];   // 10^(-q/10), powf
    float omp_f[128];     // 1 - ph2pr_f (float subtraction)
    float p3_f[128];      // ph2pr_f / 3 (IEEE division, intra: pairhmm_impl.cpp:154-156)
    float pt_f[128];      // ph2pr_f * (1/3f) (inter: pairhmm_inter.cpp:81,90)
    double ph2pr_d[128];
    double omp_d[128];
    double p3_d[128];
    double pt_d[128];
    float mm_f[kMMQ];
    double mm_d[kMMQ];
};

// One-hot haplotype codes: A=1 C=2 G=4 T=8; N matches every read base.
constexpr uint8_t kHapN = 0x0f;
// Read codes: bit position of the base in the hap code (A0 C1 G2 T3), 4 = N.
constexpr uint8_t kReadN = 4;

// One item slot of the table kernel: the pair and what its set-up needs,
// gathered once per plan so that a wave starts an item with one load.
struct alignas(16) ItemDesc {
    int32_t p;      // pair index, -1 = empty slot (then all fields are 0)
    uint32_t R, H;  // read and haplotype length
    float initd;    // 2^120 / (float)H
    uint64_t ro;    // offset of the read's first base
    uint64_t hoff;  // offset of the haplotype's codes
};

// One haplotype of a chained item (fwd_chain_kernel): the same for every
// group of the item, so the wave keeps it in scalar registers.  A group's
// pair for this haplotype is its pair for the chain's first haplotype plus
// pdelta (pairs are numbered haplotype-major within a region).
struct alignas(16) ChainHap {
    uint32_t H;
    float initd;    // 2^120 / (float)H
    uint64_t hoff;  // offset of the haplotype's codes
    int32_t pdelta;
    uint32_t pad[3];
};

// A unit of a read longer than one strip (the long-read kernels): strips
// [s0, s1) of pair `pair`, one strip per wave of a workgroup.  The units of
// a pair are consecutive in their list, in strip order.  A unit that does
// not end at the pair's last strip hands its last strip's boundary row to
// the next unit through slot xout (a boundary-row buffer and a progress
// counter); xin is the slot of the unit before (kNoSlot: none).
constexpr uint32_t kNoSlot = 0xffffffffu;
struct alignas(16) LongUnit {
    int32_t pair;
    uint32_t strips;  // s0 | s1 << 16
    uint32_t xin, xout;
};

// Arguments of every forward launch (passed by value).
struct FwdArgs {
    const DevTables *tab;
    // haplotype codes; hap h starts at hapc + hapc_off[h], 16-byte aligned,
    // zero-padded to a multiple of 16 bytes
    const uint8_t *hapc;
    const uint64_t *hapc_off;
    const uint32_t *hap_len;
    const float *initd_f;   // 2^120 / (float)H per hap  (pairhmm_impl.cpp:62)
    const double *initd_d;  // 2^1020 / (double)H per hap
    // reads: code per base and the four raw quality bytes
    const uint8_t *rcode;
    const uint8_t *q, *ins, *del, *gcp;
    const uint64_t *read_off;
    const uint32_t *read_len;
    // pairs
    const uint32_t *pair_hap;
    const uint32_t *pair_read;
    const uint8_t *pair_inter;  // 1: inter semantics (may be null)
    // work: items of G pair indices (-1 = empty); or, for the rescue pass,
    // a list of pair indices whose length is *n_items_dev
    const int32_t *items;
    const ItemDesc *desc;  // table kernel: descriptors of the item slots
    // chained items: item i's haplotypes are chain_hap[chain_off[i] ..
    // chain_off[i+1]); desc holds the G slots of each item's first haplotype
    const uint32_t *chain_off;
    const ChainHap *chain_hap;
    uint32_t n_items;
    const uint32_t *n_items_dev;
    uint32_t *counter;  // persistent work counter (zeroed by the reset kernel)
    // outputs
    float *raw32;
    double *raw64;
    uint8_t *flags;
    uint32_t *rescue_list;
    uint32_t *rescue_count;
    // per-block boundary rows for reads longer than one strip
    void *bnd;
    uint64_t bnd_elems;  // elements of T per buffer; 2 buffers per group
    uint32_t lds_stride; // LDS bytes per group
    uint32_t max_strips;
    // long-read kernels: units (null: one whole pair per item of `items`),
    // per slot a progress counter (blocks of the slot's strip written) and
    // a boundary row of bnd_elems elements
    const LongUnit *units;
    uint32_t *prog;
    void *xbnd;
    uint32_t *fault;  // counts the hand-over waits that ran out their bound
    uint32_t spin_limit;  // polls of one hand-over wait
    // extended range: per pair the power of two its FP64 sum carries
    // (written for the pairs the extended-range pass scores)
    int32_t *exp64;
    // long-read FP32 pass in two launches: 1 = strip 0 of each pair, which
    // finds the pairs whose last row of strip 0 is all zero (kFlagDead);
    // 2 = the other strips, skipping those pairs; 0 = one launch
    uint32_t dead_mode;
    // long-read kernels: the rows per lane whose layout (leading identity
    // rows, off = (off_mr - R % off_mr) % off_mr) the strips follow; 0 = the
    // kernel's own.  The FP32 pass's first launch runs the first strip of
    // the 16-row layout as sub-strips of 4 rows per lane in that layout, so
    // that its last row is the row the second launch continues from.
    uint32_t off_mr;
    // FP64 rescue in classes.  Class c < kRescueClasses: reads of the 32-lane
    // table pass with c + 1 rows per lane; its bin_cap[c] entries from
    // bin[bin_base[c]] hold its inter pairs from the front (bin_cnt[c] of
    // them) and its intra pairs from the back (bin_cnt[kRescueClasses + c]).
    // The other rescued pairs are bin[bin_base[kRescueClasses] ..), counted
    // in bin_cnt[2 * kRescueClasses].
    uint32_t *bin;
    uint32_t *bin_cnt;
    const uint8_t *hap_n;  // 1: the haplotype has an N
    uint32_t bin_base[9];
    uint32_t bin_cap[9];
    uint32_t bin_route;    // 1: a pair the FP64 regroup grouped takes the inter arithmetic
};

// The FP64 rescue's table pass: 32 lanes per pair, reads of at most
// 32 * kRescueClasses - 1 rows, haplotypes without N and of at most
// kRescueTabMaxH columns (constexpr: host and device code call it).
constexpr int kRescueClasses = 8;
constexpr uint32_t kRescueTabMaxH = 16384;
constexpr int rescue_class(uint32_t R, uint32_t H, bool hap_n) {
    return (R <= 32u * kRescueClasses - 1 && H <= kRescueTabMaxH && !hap_n) ? (int)(R / 32u) : -1;
}

// Flag bit used between the passes of a plan: the FP64 regroup of
// schedule_pairhmm put this FP32-rejected inter pair in an inter FP64 group.
// The finalisation turns it into RODV_PHMM_FLAG_INTER and clears it.
constexpr uint8_t kFlagGrouped64 = 0x10;
// RODV_PHMM_FLAG_EXTENDED: the FP64 sum underflowed and the extended-range
// pass scored the pair (its raw64 carries 2^exp64 more).
constexpr uint8_t kFlagExtended = 0x20;
// Between the two FP32 launches of long reads: the pair's strip 0 ended in an
// all-zero row, so every later row and its FP32 sum are exactly zero (it is
// rescued); the finalisation clears it.
constexpr uint8_t kFlagDead = 0x40;

// schedule_pairhmm's FP64 stage (pairhmm_schedule.cpp:251-304): per region,
// greedy windows of `wd` FP32-rejected inter pairs in the region's sorted
// order, accepted when their idle ratio is <= idle_d.
struct RegroupArgs {
    const uint32_t *list;     // per region, its FP32 inter pairs in sorted order
    const uint64_t *off;      // n_regions + 1 offsets into list
    uint32_t n_regions;
    uint32_t wd;
    double idle_d;
    const uint32_t *pair_hap;
    const uint32_t *pair_read;
    const uint32_t *hap_len;
    const uint32_t *read_len;
    uint8_t *flags;           // RESCUED from the FP32 pass; kFlagGrouped64 set here
    uint32_t *scratch;        // same layout as list: the region's rejected pairs
};

struct FinalizeArgs {
    const float *raw32;
    const double *raw64;
    const int32_t *exp64;  // extended range (may be null)
    const uint8_t *pair_inter;
    double *out;
    uint8_t *flags;  // in: RESCUED set by the FP32 pass; out: all flags
    uint64_t n;
};

// Arguments of the adjoint launches (parametrised model, include/rodv_phmm.h).
struct AdjArgs {
    const uint8_t *hapc;
    const uint64_t *hapc_off;
    const uint32_t *hap_len;
    const float *initd_f;
    const double *initd_d;
    const uint8_t *rcode;
    const uint8_t *q;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *pair_hap;
    const uint32_t *pair_read;
    const int32_t *items;  // one pair per item
    uint32_t n_items;
    const uint32_t *n_items_dev;
    uint32_t *counter;
    // per base quality: emission priors 1 - e, e / 3, d e / d scale, bin
    const float *xf, *yf;
    const double *xd, *yd, *wq;
    const uint8_t *qbin;
    double delta, gamma, mm_d, g_d;
    float mm_f, g_f, dl_f, gm_f;
    double *out_log10;
    double *out_grad;  // 10 per pair
    uint8_t *flags;
    uint32_t *rescue_list;
    uint32_t *rescue_count;
    void *ck;          // level-1 checkpoints, ck_elems elements of T per block
    uint64_t ck_elems;
    uint32_t lds_stride;  // bytes of hap codes before the level-2 states
    int ckpt;             // level-1 checkpoint stride in steps (multiple of 4)
    // reads longer than one strip (adj_long_kernel), per block: the last row
    // of every strip but the last (lrow doubles each, lbnd_stride per
    // block), and waves + 1 rows of the adjoints a strip hands to the strip
    // above (lab_elems doubles each); ck_elems per wave
    void *lbnd;
    uint64_t lbnd_stride;
    uint64_t lrow;
    void *lab;
    uint64_t lab_elems;
    // hand-over waits between the long-read kernel's waves: polls before a
    // wait gives up, and the count of those that did
    uint32_t spin_limit;
    uint32_t *fault;
    // a long pair's strips over several workgroups, per long pair li
    // (lstrips strips' room each): the shifts, the scaled sum and its exponent from
    // pass A, each strip's published reverse progress, each strip's rows'
    // gradient contributions (4 x 64 doubles per row); the (pair << 8 |
    // strip) items of pass B and their work counter
    int *lsh;
    double *lsum;
    int *lexp;
    int *lprog;
    double *lpart;
    const uint32_t *bitems;
    uint32_t n_bitems;
    uint32_t *bcounter;
    uint32_t lstrips;  // strips' room per long pair in lab, lsh, lprog, lpart
    // passes A and B in one launch: per long pair its strips, whether its
    // pass A is done (published after its rows), and the strips taken
    const uint32_t *lnstr;
    int *lready;
    uint32_t *lnext;
    // pass A's level-1 checkpoints for pass B (null: pass B recomputes
    // them): long pair li's strips from lck1 + lck1_off[li]
    double *lck1;
    const uint64_t *lck1_off;
};

}  // namespace rodv
