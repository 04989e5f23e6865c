// phmm_schedule.cpp -- roDV's batch call, compiled into librodv_phmm.so with
// the reference's exact declaration and linkage:
//   bool pairhmm::schedule::schedule_pairhmm(...)   (src/core/pairhmm_schedule.h:14-25)
// A caller compiled against the reference's header links against this
// library in place of the reference's pairhmm_schedule.cpp, unchanged.
//
// Contract kept from pairhmm_schedule.cpp:116-340:
//  - returns false when there are no haplotypes or no reads;
//  - result is resized to M rows (haplotypes) of N log10 likelihoods;
//  - use_double / idle ratios keep their meaning: pairs are routed between the
//    inter and intra arithmetic exactly as the reference's greedy grouping
//    routes them, with the group widths of this host's CPU
//    (RODV_PHMM_SEMANTICS_SCHEDULE, the default of rodv_phmm_config_default).
// Differences: invalid input (bases outside ACGTN, zero lengths, quality
// strings of the wrong length) returns false instead of undefined behaviour,
// nothing is printed unless `verbose`, and concurrent calls are safe (they
// take turns on the process's engine context).
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/rodv_phmm.h"

namespace pairhmm {
namespace schedule {

namespace {

// The process's engine context on device 0 (the reference's call names no
// device), created on first use; calls on it take turns.
struct DefaultCtx {
    rodv_phmm_ctx *ctx = nullptr;
    std::mutex mu;
    DefaultCtx() {
        if (rodv_phmm_init() == RODV_PHMM_OK && rodv_phmm_ctx_create(0, &ctx) != RODV_PHMM_OK) ctx = nullptr;
    }
};

DefaultCtx &default_ctx() {
    static DefaultCtx c;  // thread-safe initialisation
    return c;
}

}  // namespace

__attribute__((visibility("default"))) bool schedule_pairhmm(
    const std::vector<std::vector<uint8_t>> &haplotypes, const std::vector<std::vector<uint8_t>> &reads,
    std::vector<std::vector<double>> &result, const std::vector<std::vector<uint8_t>> &quality,
    const std::vector<std::vector<uint8_t>> &insertion_qualities,
    const std::vector<std::vector<uint8_t>> &deletion_qualities,
    const std::vector<std::vector<uint8_t>> &gap_contiguous_qualities, bool use_double,
    double max_idle_ratio_float, double max_idle_ratio_double, bool verbose) {
    const size_t M = haplotypes.size(), N = reads.size();
    if (M == 0 || N == 0) return false;  // pairhmm_schedule.cpp:130-132
    if (quality.size() != N || insertion_qualities.size() != N || deletion_qualities.size() != N ||
        gap_contiguous_qualities.size() != N || M > UINT32_MAX || N > UINT32_MAX)
        return false;
    try {
        // flatten into the C ABI's concatenated layout
        std::vector<uint8_t> hb, rb, q, qi, qd, qc;
        std::vector<uint64_t> ho(M), ro(N);
        std::vector<uint32_t> hl(M), rl(N);
        for (size_t h = 0; h < M; h++) {
            if (haplotypes[h].size() > UINT32_MAX) return false;
            ho[h] = hb.size();
            hl[h] = (uint32_t)haplotypes[h].size();
            hb.insert(hb.end(), haplotypes[h].begin(), haplotypes[h].end());
        }
        for (size_t r = 0; r < N; r++) {
            const size_t n = reads[r].size();
            if (n > UINT32_MAX || quality[r].size() != n || insertion_qualities[r].size() != n ||
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
        b.num_regions = 0;  // one region: every haplotype x every read
        b.regions = nullptr;
        rodv_phmm_config cfg;
        rodv_phmm_config_default(&cfg);
        cfg.use_double = use_double ? 1 : 0;
        cfg.max_idle_ratio_float = max_idle_ratio_float;
        cfg.max_idle_ratio_double = max_idle_ratio_double;
        std::vector<double> out(M * N);
        std::vector<uint8_t> flags(M * N);
        DefaultCtx &dc = default_ctx();
        if (!dc.ctx) {
            if (verbose) std::fprintf(stderr, "schedule_pairhmm: no MI355X engine context\n");
            return false;
        }
        int rc;
        {
            std::lock_guard<std::mutex> lk(dc.mu);
            rc = rodv_phmm_forward(dc.ctx, &b, &cfg, out.data(), flags.data());
        }
        if (rc != RODV_PHMM_OK) {
            if (verbose) std::fprintf(stderr, "schedule_pairhmm: %s\n", rodv_phmm_strerror(rc));
            return false;
        }
        // result[hap][read] (pairhmm_schedule.cpp:135-138)
        result.resize(M);
        for (size_t h = 0; h < M; h++) result[h].assign(out.begin() + h * N, out.begin() + (h + 1) * N);
        if (verbose) {
            size_t rescued = 0, inter = 0;
            for (uint8_t f : flags) {
                rescued += (f & RODV_PHMM_FLAG_RESCUED) != 0;
                inter += (f & RODV_PHMM_FLAG_INTER) != 0;
            }
            std::fprintf(stderr, "Total num pairs: %zu, inter: %zu, rescued in FP64: %zu\n", M * N, inter,
                         rescued);
        }
        return true;
    } catch (...) {  // allocation failure: the reference's bool contract
        return false;
    }
}

}  // namespace schedule
}  // namespace pairhmm
