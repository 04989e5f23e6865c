// rodv_phmm_schedule.hpp -- header-only C++ shim with the reference's exact
// batch signature, pairhmm::schedule::schedule_pairhmm
// (src/core/pairhmm_schedule.h:14-25), implemented over the C ABI in
// rodv_phmm.h.  Existing callers relink against librodv_phmm.so unchanged.
//
// Contract kept from pairhmm_schedule.cpp:116-340:
//  - returns false when there are no haplotypes or no reads;
//  - result is resized to M rows (haplotypes) of N log10 likelihoods;
//  - use_double / idle ratios keep their meaning: the engine routes pairs
//    between the inter and intra arithmetic exactly as the reference's
//    greedy grouping does, with the group widths of this host's CPU
//    (RODV_PHMM_SEMANTICS_SCHEDULE, the default of rodv_phmm_config_default).
// Differences: invalid input (bases outside ACGTN, zero lengths, quality
// strings of the wrong length) returns false instead of undefined behaviour,
// and nothing is printed unless `verbose`.
#ifndef RODV_PHMM_SCHEDULE_HPP_
#define RODV_PHMM_SCHEDULE_HPP_

#include <cstdint>
#include <cstdio>
#include <vector>

#include "rodv_phmm.h"

namespace pairhmm {
namespace schedule {

namespace detail {
// One engine context per process and device 0, created on first use.
inline rodv_phmm_ctx *default_ctx() {
    static rodv_phmm_ctx *ctx = [] {
        rodv_phmm_ctx *c = nullptr;
        if (rodv_phmm_init() != RODV_PHMM_OK) return c;
        if (rodv_phmm_ctx_create(0, &c) != RODV_PHMM_OK) c = nullptr;
        return c;
    }();
    return ctx;
}
}  // namespace detail

inline bool schedule_pairhmm(const std::vector<std::vector<uint8_t>> &haplotypes,
                             const std::vector<std::vector<uint8_t>> &reads,
                             std::vector<std::vector<double>> &result,
                             const std::vector<std::vector<uint8_t>> &quality,
                             const std::vector<std::vector<uint8_t>> &insertion_qualities,
                             const std::vector<std::vector<uint8_t>> &deletion_qualities,
                             const std::vector<std::vector<uint8_t>> &gap_contiguous_qualities,
                             bool use_double = false, double max_idle_ratio_float = 0.1,
                             double max_idle_ratio_double = 0.1, bool verbose = false) {
    const size_t M = haplotypes.size(), N = reads.size();
    if (M == 0 || N == 0) return false;
    if (quality.size() != N || insertion_qualities.size() != N ||
        deletion_qualities.size() != N || gap_contiguous_qualities.size() != N)
        return false;
    rodv_phmm_ctx *ctx = detail::default_ctx();
    if (!ctx) return false;
    // flatten
    std::vector<uint8_t> hb, rb, q, qi, qd, qc;
    std::vector<uint64_t> ho(M), ro(N);
    std::vector<uint32_t> hl(M), rl(N);
    for (size_t h = 0; h < M; h++) {
        ho[h] = hb.size();
        hl[h] = (uint32_t)haplotypes[h].size();
        hb.insert(hb.end(), haplotypes[h].begin(), haplotypes[h].end());
    }
    for (size_t r = 0; r < N; r++) {
        const size_t n = reads[r].size();
        if (quality[r].size() != n || insertion_qualities[r].size() != n ||
            deletion_qualities[r].size() != n || gap_contiguous_qualities[r].size() != n)
            return false;
        ro[r] = rb.size();
        rl[r] = (uint32_t)n;
        rb.insert(rb.end(), reads[r].begin(), reads[r].end());
        q.insert(q.end(), quality[r].begin(), quality[r].end());
        qi.insert(qi.end(), insertion_qualities[r].begin(), insertion_qualities[r].end());
        qd.insert(qd.end(), deletion_qualities[r].begin(), deletion_qualities[r].end());
        qc.insert(qc.end(), gap_contiguous_qualities[r].begin(), gap_contiguous_qualities[r].end());
    }
    rodv_phmm_batch b{};
    b.num_haps = (uint32_t)M;
    b.hap_bases = hb.data();
    b.hap_offsets = ho.data();
    b.hap_lengths = hl.data();
    b.num_reads = (uint32_t)N;
    b.read_bases = rb.data();
    b.base_quals = q.data();
    b.ins_quals = qi.data();
    b.del_quals = qd.data();
    b.gcp_quals = qc.data();
    b.read_offsets = ro.data();
    b.read_lengths = rl.data();
    b.num_regions = 0;
    b.regions = nullptr;
    rodv_phmm_config cfg;
    rodv_phmm_config_default(&cfg);
    cfg.use_double = use_double ? 1 : 0;
    cfg.max_idle_ratio_float = max_idle_ratio_float;
    cfg.max_idle_ratio_double = max_idle_ratio_double;
    std::vector<double> out(M * N);
    std::vector<uint8_t> flags(M * N);
    const int rc = rodv_phmm_forward(ctx, &b, &cfg, out.data(), flags.data());
    if (rc != RODV_PHMM_OK) {
        if (verbose) std::fprintf(stderr, "rodv_phmm_forward: %s\n", rodv_phmm_strerror(rc));
        return false;
    }
    result.resize(M);
    for (size_t h = 0; h < M; h++) result[h].assign(out.begin() + h * N, out.begin() + (h + 1) * N);
    if (verbose) {
        size_t rescued = 0, inter = 0;
        for (uint8_t f : flags) {
            rescued += (f & RODV_PHMM_FLAG_RESCUED) != 0;
            inter += (f & RODV_PHMM_FLAG_INTER) != 0;
        }
        std::fprintf(stderr, "Total num pairs: %zu, inter: %zu, rescued in FP64: %zu\n", M * N,
                     inter, rescued);
    }
    return true;
}

}  // namespace schedule
}  // namespace pairhmm

#endif  // RODV_PHMM_SCHEDULE_HPP_
