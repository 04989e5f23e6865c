// ref_driver.cpp -- C entry points over the reference's own PairHMM kernels.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the roDV sources where they lie (src/core/pairhmm/{intra,inter}); nothing
// from the reference is copied here.  Compiled once per ISA: with -mavx512f
// it wraps the AVX-512 intra kernel, with -mavx2 the AVX2 intra kernel and
// the AVX2 inter group kernel (the AVX-512 inter traits include a boost header
// this image does not have, so that variant is not built).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "pairhmm/common/common.h"
#include "pairhmm/common/context.h"
#include "pairhmm/intra/pairhmm_api.h"
#include "pairhmm/intra/pairhmm_impl.h"
#if !defined(__AVX512F__)
#include "pairhmm/inter/pairhmm_inter.h"
#include "pairhmm/inter/pairhmm_inter_api.h"
#endif

using pairhmm::common::Context;
using pairhmm::common::ContextBase;
using pairhmm::common::TestCase;

namespace {
TestCase make_tc(uint32_t rlen, uint32_t hlen, const uint8_t *hap, const uint8_t *read,
                 const uint8_t *q, const uint8_t *ins, const uint8_t *del, const uint8_t *gcp) {
    TestCase tc;
    tc.rslen = rlen;
    tc.haplen = hlen;
    tc.hap = hap;
    tc.rs = read;
    tc.q = q;
    tc.i = ins;
    tc.d = del;
    tc.c = gcp;
    return tc;
}
}  // namespace

#if defined(__AVX512F__)
#define REF_SYM(name) phmm_ref_avx512_##name
using FT = pairhmm::intra::AVX512FloatTraits;
using DT = pairhmm::intra::AVX512DoubleTraits;
#else
#define REF_SYM(name) phmm_ref_avx2_##name
using FT = pairhmm::intra::AVX2FloatTraits;
using DT = pairhmm::intra::AVX2DoubleTraits;
#endif

extern "C" {

// Raw FP32 / FP64 sums of the intra kernel (pairhmm_impl.cpp:295-440).
float REF_SYM(intra_raw_f32)(uint32_t rlen, uint32_t hlen, const uint8_t *hap,
                             const uint8_t *read, const uint8_t *q, const uint8_t *ins,
                             const uint8_t *del, const uint8_t *gcp) {
    TestCase tc = make_tc(rlen, hlen, hap, read, q, ins, del, gcp);
    return pairhmm::intra::PairHMMComputer<FT>::compute(tc);
}

double REF_SYM(intra_raw_f64)(uint32_t rlen, uint32_t hlen, const uint8_t *hap,
                              const uint8_t *read, const uint8_t *q, const uint8_t *ins,
                              const uint8_t *del, const uint8_t *gcp) {
    TestCase tc = make_tc(rlen, hlen, hap, read, q, ins, del, gcp);
    return pairhmm::intra::PairHMMComputer<DT>::compute(tc);
}

// The per-pair API (pairhmm_api.cpp:14-35).
double REF_SYM(intra_log10)(uint32_t rlen, uint32_t hlen, const uint8_t *hap, const uint8_t *read,
                            const uint8_t *q, const uint8_t *ins, const uint8_t *del,
                            const uint8_t *gcp, int use_double) {
    TestCase tc = make_tc(rlen, hlen, hap, read, q, ins, del, gcp);
#if defined(__AVX512F__)
    return pairhmm::intra::computeLikelihoodsAVX512(tc, use_double != 0);
#else
    return pairhmm::intra::computeLikelihoodsAVX2(tc, use_double != 0);
#endif
}

// The per-pair API over many pairs on `nthreads` host threads (pairs handed
// out one at a time), for the CPU baseline on every core.  Flat arrays as in
// phmm_oracle_batch.
int REF_SYM(intra_log10_batch)(const uint8_t *hap_bases, const uint64_t *hap_off,
                               const uint32_t *hap_len, const uint8_t *read_bases,
                               const uint8_t *q, const uint8_t *ins, const uint8_t *del,
                               const uint8_t *gcp, const uint64_t *read_off,
                               const uint32_t *read_len, uint64_t n, const uint32_t *pair_hap,
                               const uint32_t *pair_read, int nthreads, double *out) {
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        for (uint64_t k = next++; k < n; k = next++) {
            const uint32_t h = pair_hap[k], r = pair_read[k];
            TestCase tc = make_tc(read_len[r], hap_len[h], hap_bases + hap_off[h],
                                  read_bases + read_off[r], q + read_off[r], ins + read_off[r],
                                  del + read_off[r], gcp + read_off[r]);
#if defined(__AVX512F__)
            out[k] = pairhmm::intra::computeLikelihoodsAVX512(tc, false);
#else
            out[k] = pairhmm::intra::computeLikelihoodsAVX2(tc, false);
#endif
        }
    };
    const int nt = std::max(1, nthreads);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return 0;
}

// Tables as Context<float>/Context<double> build them (context.h:90-166).
void REF_SYM(tables_f32)(float *ph2pr, float *mm, int mm_n) {
    Context<float> ctx;
    (void)ctx;
    std::memcpy(ph2pr, ContextBase<float>::ph2pr, sizeof(float) * 128);
    std::memcpy(mm, ContextBase<float>::matchToMatchProb, sizeof(float) * (size_t)mm_n);
}

void REF_SYM(tables_f64)(double *ph2pr, double *mm, int mm_n) {
    Context<double> ctx;
    (void)ctx;
    std::memcpy(ph2pr, ContextBase<double>::ph2pr, sizeof(double) * 128);
    std::memcpy(mm, ContextBase<double>::matchToMatchProb, sizeof(double) * (size_t)mm_n);
}

#if !defined(__AVX512F__)
// Group of exactly 8 (FP32) or 4 (FP64) pairs through the inter kernel;
// raw sums (islog10 = false), pairhmm_inter_api.cpp:44-73.
int phmm_ref_avx2_inter_raw_f32(const uint32_t *rlen, const uint32_t *hlen,
                                const uint8_t *const *hap, const uint8_t *const *read,
                                const uint8_t *const *q, const uint8_t *const *ins,
                                const uint8_t *const *del, const uint8_t *const *gcp,
                                double *out) {
    TestCase tc[8];
    for (int k = 0; k < 8; k++)
        tc[k] = make_tc(rlen[k], hlen[k], hap[k], read[k], q[k], ins[k], del[k], gcp[k]);
    return pairhmm::inter::compute_inter_pairhmm_AVX2_float(tc, 8, out, false) ? 1 : 0;
}

int phmm_ref_avx2_inter_raw_f64(const uint32_t *rlen, const uint32_t *hlen,
                                const uint8_t *const *hap, const uint8_t *const *read,
                                const uint8_t *const *q, const uint8_t *const *ins,
                                const uint8_t *const *del, const uint8_t *const *gcp,
                                double *out) {
    TestCase tc[4];
    for (int k = 0; k < 4; k++)
        tc[k] = make_tc(rlen[k], hlen[k], hap[k], read[k], q[k], ins[k], del[k], gcp[k]);
    return pairhmm::inter::compute_inter_pairhmm_AVX2_double(tc, 4, out, false) ? 1 : 0;
}
#endif

}  // extern "C"
