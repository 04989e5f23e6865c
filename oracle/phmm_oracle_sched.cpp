// phmm_oracle_sched.cpp -- restatement of schedule_pairhmm's routing.
//
// TEST INFRASTRUCTURE ONLY (see phmm_oracle.h).
//
// Follows pairhmm_schedule.cpp:116-340: enumerate M x N pairs, sort by
// (hap_len desc, read_len desc) with std::sort (the same libstdc++ algorithm
// the reference uses, so ties land in the same order), greedily pack runs of
// W unused N-free pairs whose idle ratio is within bound (:46-105), run those
// through the inter kernel in FP32, regroup FP32 failures for the FP64 inter
// kernel, and send every remaining pair to the intra kernel.  The per-lane
// arithmetic of the inter kernel is independent of the other lanes of its
// group, so each grouped pair is evaluated with phmm_oracle_raw_f32/_f64 under
// inter semantics.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "phmm_oracle.h"

namespace {

struct Pair {  // pairhmm_schedule.cpp:23-35
    size_t hap_idx, read_idx;
    uint32_t hap_len, read_len;
    bool contains_n, used;
};

// pairhmm_schedule.cpp:46-105
std::vector<std::vector<size_t>> greedy_grouping(std::vector<Pair> &pairs, double max_idle,
                                                 uint32_t W) {
    std::vector<std::vector<size_t>> groups;
    if (W == 0) return groups;
    const size_t total = pairs.size();
    for (size_t i = 0; i < total; ++i) {
        if (pairs[i].used || pairs[i].contains_n) continue;
        std::vector<size_t> g;
        for (size_t idx = i; idx < total && g.size() < W; ++idx)
            if (!pairs[idx].used && !pairs[idx].contains_n) g.push_back(idx);
        if (g.size() < W) break;
        uint32_t mh = 0, mr = 0;
        uint64_t sum = 0;
        for (size_t k : g) {
            mh = std::max(mh, pairs[k].hap_len);
            mr = std::max(mr, pairs[k].read_len);
            sum += (uint64_t)pairs[k].hap_len * pairs[k].read_len;
        }
        uint64_t cap = (uint64_t)mh * mr * W;
        uint64_t idle = cap > sum ? cap - sum : 0;
        double ratio = cap == 0 ? 0.0 : (double)idle / (double)cap;
        if (ratio <= max_idle) {
            for (size_t k : g) pairs[k].used = true;
            groups.push_back(std::move(g));
        }
    }
    return groups;
}

bool contains_n(const uint8_t *s, uint32_t n) {  // pairhmm_schedule.cpp:107-114
    for (uint32_t i = 0; i < n; i++)
        if (s[i] == 'N' || s[i] == 'n') return true;
    return false;
}

const float kLog10KF = 36.1236000061f;
const double kLog10KD = 307.050595577260822;
const float kMinAccepted = 1e-28f;

}  // namespace

extern "C" int phmm_oracle_schedule(uint32_t M, const uint8_t *const *haps, const uint32_t *hlen,
                                    uint32_t N, const uint8_t *const *reads,
                                    const uint8_t *const *q, const uint8_t *const *ins,
                                    const uint8_t *const *del, const uint8_t *const *gcp,
                                    const uint32_t *rlen, int use_double, double idle_f,
                                    double idle_d, uint32_t Wf, uint32_t Wd, double *result,
                                    uint8_t *flags, float *raw32, uint8_t *route) {
    if (M == 0 || N == 0) return 0;
    auto mkpair = [&](size_t h, size_t r) {
        phmm_oracle_pair p;
        p.hlen = hlen[h];
        p.rlen = rlen[r];
        p.hap = haps[h];
        p.read = reads[r];
        p.q = q[r];
        p.ins = ins[r];
        p.del = del[r];
        p.gcp = gcp[r];
        return p;
    };
    for (size_t k = 0; k < (size_t)M * N; k++) {
        result[k] = 0.0;
        if (flags) flags[k] = 0;
        if (route) route[k] = PHMM_ORACLE_ROUTE_INTRA;
    }
    std::vector<Pair> pairs;
    pairs.reserve((size_t)M * N);
    std::vector<bool> nh(M), nr(N);
    for (uint32_t h = 0; h < M; h++) nh[h] = contains_n(haps[h], hlen[h]);
    for (uint32_t r = 0; r < N; r++) nr[r] = contains_n(reads[r], rlen[r]);
    for (size_t h = 0; h < M; h++)
        for (size_t r = 0; r < N; r++)
            pairs.push_back(Pair{h, r, hlen[h], rlen[r], nh[h] || nr[r], false});
    std::sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) {
        if (a.hap_len == b.hap_len) return a.read_len > b.read_len;
        return a.hap_len > b.hap_len;
    });
    std::vector<bool> done_f((size_t)M * N, false), done_d((size_t)M * N, false);
    if (!use_double && Wf > 0) {
        for (const auto &g : greedy_grouping(pairs, idle_f, Wf)) {
            for (size_t k : g) {
                Pair &pr = pairs[k];
                phmm_oracle_pair p = mkpair(pr.hap_idx, pr.read_idx);
                float raw = phmm_oracle_raw_f32(&p, PHMM_ORACLE_INTER);
                size_t o = pr.hap_idx * N + pr.read_idx;
                if (raw32) raw32[o] = raw;
                if (raw < kMinAccepted) {
                    pr.used = false;
                    if (route) route[o] = PHMM_ORACLE_ROUTE_INTRA64;
                } else {
                    pr.used = true;
                    result[o] = std::log10((double)raw) - (double)kLog10KF;
                    if (flags) flags[o] |= PHMM_ORACLE_FLAG_INTER;
                    if (route) route[o] = PHMM_ORACLE_ROUTE_INTER32;
                    done_d[o] = true;
                }
                done_f[o] = true;
            }
        }
    }
    if (Wd > 0) {
        for (auto &pr : pairs)
            if (!done_f[pr.hap_idx * N + pr.read_idx]) pr.used = true;
        for (const auto &g : greedy_grouping(pairs, idle_d, Wd)) {
            for (size_t k : g) {
                Pair &pr = pairs[k];
                phmm_oracle_pair p = mkpair(pr.hap_idx, pr.read_idx);
                double raw = phmm_oracle_raw_f64(&p, PHMM_ORACLE_INTER);
                size_t o = pr.hap_idx * N + pr.read_idx;
                result[o] = std::log10(raw) - kLog10KD;
                if (route) route[o] = PHMM_ORACLE_ROUTE_INTER64;
                if (flags) {
                    flags[o] |= PHMM_ORACLE_FLAG_INTER | PHMM_ORACLE_FLAG_RESCUED;
                    if (!(raw > 0.0)) flags[o] |= PHMM_ORACLE_FLAG_UNDERFLOW;
                }
                done_d[o] = true;
            }
        }
    }
    for (size_t h = 0; h < M; h++) {
        for (size_t r = 0; r < N; r++) {
            size_t o = h * N + r;
            if (done_d[o]) continue;
            phmm_oracle_pair p = mkpair(h, r);
            uint8_t fl = 0;
            float r32 = 0.0f;
            result[o] = phmm_oracle_intra_log10(&p, done_f[o] ? 1 : 0, &r32, nullptr, &fl);
            if (raw32 && !done_f[o]) raw32[o] = r32;
            if (flags) flags[o] |= fl;
        }
    }
    return 1;
}
