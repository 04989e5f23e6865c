/*
 * phmm_oracle.h -- CPU restatement of roDV's PairHMM forward path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * under rodv_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path never calls into it.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the roDV source tree, src/core/...).  Parity status of each function is
 * listed in the header of phmm_oracle.c.
 */
#ifndef RODV_PHMM_ORACLE_H_
#define RODV_PHMM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Size of the triangular match-to-match table for qualities 0..254
 * (pairhmm/common/context.h:33 ((MAX_QUAL + 1) * (MAX_QUAL + 2)) >> 1). */
#define PHMM_ORACLE_MM_SIZE 32640
/* Jacobian log table size (pairhmm/common/context.h:14). */
#define PHMM_ORACLE_JAC_SIZE 80001

/* Semantics of the FP32 pass (which reference kernel is restated). */
enum {
    PHMM_ORACLE_INTRA = 0, /* intra/pairhmm_impl.cpp: eps/3, N is a wildcard   */
    PHMM_ORACLE_INTER = 1  /* inter/pairhmm_inter.cpp: eps*(1/3), raw byte cmp */
};

/* Batch modes of phmm_oracle_batch(). */
enum {
    PHMM_ORACLE_MODE_INTRA_API = 0, /* intra::computeLikelihoods (FP32, FP64 rescue) */
    PHMM_ORACLE_MODE_RAW_F32 = 1,   /* raw FP32 sum only (semantics from `semantics`) */
    PHMM_ORACLE_MODE_RAW_F64 = 2,   /* raw FP64 sum only                               */
    PHMM_ORACLE_MODE_EXTENDED = 3,  /* FP64 with power-of-two row renormalisation      */
    PHMM_ORACLE_MODE_SCHEDULE = 4   /* schedule_pairhmm routing (inter groups + intra) */
};

/* Output flag bits (same meaning as the engine's RODV_PHMM_FLAG_*). */
#define PHMM_ORACLE_FLAG_RESCUED 0x01u  /* FP32 raw < 1e-28f, result from FP64 */
#define PHMM_ORACLE_FLAG_UNDERFLOW 0x02u /* FP64 raw == 0: log10 is -inf        */
#define PHMM_ORACLE_FLAG_INTER 0x08u    /* routed through the inter restatement */

/* Prior tables exactly as Context<float>/Context<double> build them. */
void phmm_oracle_tables_f32(float ph2pr[128], float mm[PHMM_ORACLE_MM_SIZE]);
void phmm_oracle_tables_f64(double ph2pr[128], double mm[PHMM_ORACLE_MM_SIZE]);

/* One pair: hap/read are ASCII bases, q/ins/del/gcp raw Phred bytes. */
typedef struct phmm_oracle_pair {
    uint32_t rlen, hlen;
    const uint8_t *hap, *read, *q, *ins, *del, *gcp;
} phmm_oracle_pair;

float phmm_oracle_raw_f32(const phmm_oracle_pair *p, int semantics);
double phmm_oracle_raw_f64(const phmm_oracle_pair *p, int semantics);
/* log10 likelihood from an FP64 forward that renormalises each row by a power
 * of two; finite wherever the plain FP64 forward underflows to zero.
 * semantics: the prior of PHMM_ORACLE_INTRA or PHMM_ORACLE_INTER. */
double phmm_oracle_log10_extended(const phmm_oracle_pair *p, int semantics);
/* intra::computeLikelihoods semantics; also reports the raw sums. */
double phmm_oracle_intra_log10(const phmm_oracle_pair *p, int use_double, float *raw32,
                               double *raw64, uint8_t *flags);

/*
 * Batch entry: pairs (pair_hap[k], pair_read[k]).  Bases and the four quality
 * strings are concatenated; *_off give the start of each sequence.
 * out_* may be NULL.  nthreads <= 1 runs on the calling thread.
 * MODE_SCHEDULE requires the pairs to form regions (see region_* below).
 */
int phmm_oracle_batch(uint32_t num_haps, const uint8_t *hap_bases, const uint64_t *hap_off,
                      const uint32_t *hap_len, uint32_t num_reads, const uint8_t *read_bases,
                      const uint8_t *q, const uint8_t *ins, const uint8_t *del,
                      const uint8_t *gcp, const uint64_t *read_off, const uint32_t *read_len,
                      uint64_t num_pairs, const uint32_t *pair_hap, const uint32_t *pair_read,
                      int mode, int semantics, int nthreads, double *out_log10,
                      float *out_raw32, double *out_raw64, uint8_t *out_flags);

/* Route a pair takes through schedule_pairhmm (pairhmm_schedule.cpp:185-335). */
#define PHMM_ORACLE_ROUTE_INTRA 0u   /* intra kernel, FP32 first               */
#define PHMM_ORACLE_ROUTE_INTER32 1u /* inter FP32 group, accepted             */
#define PHMM_ORACLE_ROUTE_INTER64 2u /* inter FP32 rejected, inter FP64 group  */
#define PHMM_ORACLE_ROUTE_INTRA64 3u /* inter FP32 rejected, intra FP64 only   */

/*
 * schedule_pairhmm restatement for one region (M haps x N reads), result
 * hap-major like pairhmm_schedule.cpp:135-138.  simd_width_f / _d are the
 * group widths the host CPU would use (16/8 with AVX-512, 8/4 with AVX2).
 * raw32 (may be NULL) receives the FP32 sum that decided the pair's rescue
 * (inter arithmetic for routes 1-3, intra for route 0), route (may be NULL)
 * the PHMM_ORACLE_ROUTE_* of each pair.
 * Returns 0 for M == 0 or N == 0 (the reference's `false`), 1 otherwise.
 */
int phmm_oracle_schedule(uint32_t M, const uint8_t *const *haps, const uint32_t *hlen,
                         uint32_t N, const uint8_t *const *reads, const uint8_t *const *q,
                         const uint8_t *const *ins, const uint8_t *const *del,
                         const uint8_t *const *gcp, const uint32_t *rlen, int use_double,
                         double idle_f, double idle_d, uint32_t simd_width_f,
                         uint32_t simd_width_d, double *result, uint8_t *flags, float *raw32,
                         uint8_t *route);

/*
 * Parametrised model used by the adjoint (the reference has no gradient):
 *   delta = 10^(-gap_open/10), gamma = 10^(-gap_cont/10)
 *   mm = 1 - 2 delta, gapm = 1 - gamma, mx = my = delta, xx = yy = gamma
 *   e_i = scale[bin(q_i)] * 10^(-q_i/10), bin(q) = min(q,47)/6
 *   prior = 1 - e_i on a match (or N), e_i / 3 otherwise.
 * Per-row insertion/deletion/gap qualities are replaced by the two globals.
 */
typedef struct phmm_oracle_params {
    double gap_open;   /* Phred */
    double gap_cont;   /* Phred */
    double scale[8];   /* per-quality-bin multipliers of the error probability */
} phmm_oracle_params;

double phmm_oracle_param_log10(const phmm_oracle_pair *p, const phmm_oracle_params *th);
/* Gradient of log10 L w.r.t. (gap_open, gap_cont, scale[0..7]) by reverse mode
 * over the full FP64 matrices, each row renormalised by a power of two (finite
 * for long reads).  Returns log10 L. */
double phmm_oracle_param_grad(const phmm_oracle_pair *p, const phmm_oracle_params *th,
                              double grad[10]);
/* Batch form: grad_out[10] += sum_k weight[k] * d log10 L_k.  renorm = 1:
 * rows renormalised (as phmm_oracle_param_grad); 0: the plain FP64 form at
 * K = 2^1020, which underflows for long reads (kept to pin the other). */
int phmm_oracle_param_batch(uint32_t num_haps, const uint8_t *hap_bases,
                            const uint64_t *hap_off, const uint32_t *hap_len,
                            uint32_t num_reads, const uint8_t *read_bases, const uint8_t *q,
                            const uint64_t *read_off, const uint32_t *read_len,
                            uint64_t num_pairs, const uint32_t *pair_hap,
                            const uint32_t *pair_read, const phmm_oracle_params *th,
                            const double *weight, int nthreads, double *out_log10,
                            double *grad_out, int renorm);

#ifdef __cplusplus
}
#endif
#endif
