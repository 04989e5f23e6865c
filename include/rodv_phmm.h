/*
 * rodv_phmm.h -- C ABI of the MI355X PairHMM engine.
 *
 * Drop-in boundary for roDV's batch API
 *   bool pairhmm::schedule::schedule_pairhmm(haplotypes, reads, result, quality,
 *        insertion_qualities, deletion_qualities, gap_contiguous_qualities,
 *        use_double, max_idle_ratio_float, max_idle_ratio_double, verbose)
 * (src/core/pairhmm_schedule.h:14-25, implemented at pairhmm_schedule.cpp:116-340).
 * The C++ shim in rodv_phmm_schedule.hpp keeps that exact signature and calls
 * the functions below.  Plain pointers and sizes only; nothing here throws.
 *
 * Encodings follow the reference: bases are ASCII 'A','C','G','T','N'
 * (common.h:42-56); qualities are raw Phred integers masked with & 127
 * (pairhmm_impl.cpp:89-92).  Anything else is rejected with
 * RODV_PHMM_ERR_INVALID_ARG instead of the reference's undefined behaviour.
 */
#ifndef RODV_PHMM_H_
#define RODV_PHMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RODV_PHMM_VERSION 1

/* Status codes.  0 is success. */
enum {
    RODV_PHMM_OK = 0,
    RODV_PHMM_ERR_INVALID_ARG = 1, /* bad pointer, length, base or config     */
    RODV_PHMM_ERR_OOM = 2,         /* host or device allocation failed        */
    RODV_PHMM_ERR_DEVICE = 3,      /* HIP runtime error / no device           */
    RODV_PHMM_ERR_UNSUPPORTED = 4, /* request outside what the engine supports */
    RODV_PHMM_ERR_EMPTY = 5,       /* M == 0 or N == 0 (the reference's false) */
    RODV_PHMM_ERR_TOO_LARGE = 6    /* batch exceeds rodv_phmm_forward_noalloc's
                                      fixed device scratch                     */
};

/* Per-pair output flags. */
#define RODV_PHMM_FLAG_RESCUED 0x01u   /* FP32 raw sum < 1e-28f: FP64 result  */
#define RODV_PHMM_FLAG_UNDERFLOW 0x02u /* FP64 raw sum == 0: log10 is -inf    */
#define RODV_PHMM_FLAG_AMBIGUOUS 0x04u /* FP32 raw within 2^-20 of 1e-28f     */
#define RODV_PHMM_FLAG_INTER 0x08u     /* scored with inter-kernel semantics  */
#define RODV_PHMM_FLAG_EXTENDED 0x20u  /* FP64 sum underflowed; the result is
                                          config.extended_range's (finite
                                          unless UNDERFLOW is set as well)    */

/*
 * A batch: M haplotypes and N reads, concatenated.  The four quality strings
 * of read r start at read_offsets[r] in their arrays, like the read bases.
 * Regions are blocks haps[hap_begin, +num_haps) x reads[read_begin, +num_reads)
 * scored as a full cross product, exactly one schedule_pairhmm call each.
 * num_regions == 0 means one region over all haps x all reads.
 * Output order: region by region, hap-major inside a region
 * (result[hap][read], pairhmm_schedule.cpp:135-138).
 */
typedef struct rodv_phmm_region {
    uint32_t hap_begin, num_haps;
    uint32_t read_begin, num_reads;
} rodv_phmm_region;

typedef struct rodv_phmm_batch {
    uint32_t num_haps;
    const uint8_t *hap_bases;
    const uint64_t *hap_offsets;
    const uint32_t *hap_lengths;
    uint32_t num_reads;
    const uint8_t *read_bases;
    const uint8_t *base_quals;
    const uint8_t *ins_quals;
    const uint8_t *del_quals;
    const uint8_t *gcp_quals;
    const uint64_t *read_offsets;
    const uint32_t *read_lengths;
    uint32_t num_regions;
    const rodv_phmm_region *regions;
} rodv_phmm_batch;

/* Which reference kernel's FP32 arithmetic to reproduce bit for bit. */
enum {
    RODV_PHMM_SEMANTICS_INTRA = 0,   /* intra::computeLikelihoods for every pair
                                        (pairhmm_api.cpp:14-35) */
    RODV_PHMM_SEMANTICS_SCHEDULE = 1 /* schedule_pairhmm routing (default): per
                                        region, the greedy grouping of
                                        pairhmm_schedule.cpp:46-105 sends windows
                                        of simd_width_float pairs to the inter
                                        arithmetic (pairhmm_inter.cpp), regroups
                                        their FP32 rejections in windows of
                                        simd_width_double for inter FP64, and
                                        scores every other pair as intra */
};

typedef struct rodv_phmm_config {
    uint32_t struct_size;         /* sizeof(rodv_phmm_config)                  */
    int32_t semantics;            /* RODV_PHMM_SEMANTICS_*                      */
    int32_t use_double;           /* schedule_pairhmm's use_double flag         */
    double max_idle_ratio_float;  /* schedule_pairhmm's idle bounds (0.1)       */
    double max_idle_ratio_double;
    uint32_t simd_width_float;    /* group widths of the reference's host: the  */
    uint32_t simd_width_double;   /* default is this host's (16/8 AVX-512, 8/4
                                     AVX2, 0/0 neither: cpu_features.h:53-154);
                                     at most 64 */
    int32_t use_graph;            /* capture the batched launch in a hipGraph   */
    int32_t waves_per_cu;         /* persistent grid size; 0 = engine default   */
    int32_t short_lanes;          /* lanes per pair for reads <= 255 (N-free
                                     haplotypes; <= 240 with N): 16 or 64      */
    int32_t extended_range;       /* 1: a pair whose FP64 sum underflows to zero
                                     (the reference's -inf) is scored again by
                                     an extended-range FP64 forward: its strips
                                     of 252 rows in sequence, each strip's input
                                     row scaled by a power of two, the exponent
                                     carried exactly; the result is finite
                                     while a strip's rows lose < ~2 decades per
                                     base, and flagged RODV_PHMM_FLAG_EXTENDED.
                                     Every other pair is unchanged, bit for
                                     bit.  0 (default): the reference's -inf */
    int32_t prior_table;          /* 1: per-pair LDS table of emission priors for
                                     N-free haplotypes (default); 0: select per cell */
    int32_t use_tuning;           /* 1: shapes from rodv_phmm_autotune when the
                                     context has been tuned (default)          */
    int32_t haplotype_chains;     /* reads of a region that share their
                                     haplotypes swept across up to 16 of them in
                                     one item, without a pipeline drain between
                                     haplotypes, short reads two to a lane group
                                     where that packs their rows better: 1
                                     (default) where a length bucket still fills
                                     the device with such items, 2 always, 3
                                     and 4 as 1 and 2 with one read per lane
                                     group, 0 never (one pair per slot)         */
    int32_t shared_products;      /* 1 (default): reads whose deletion quality of
                                     each row equals the insertion quality of
                                     the row below (constant indel qualities)
                                     take the 16-lane FP32 table kernels that
                                     compute M(i,j) * mx_{i+1} once for both
                                     I(i+1,j) and D(i,j+1): 7 FP32 instructions
                                     per cell instead of 8, the same bits;
                                     0: never                                  */
} rodv_phmm_config;

/* Counters of the last execution of a plan. */
typedef struct rodv_phmm_stats {
    uint64_t num_pairs;
    uint64_t cells;          /* sum over pairs of R*H                      */
    uint64_t rescued_pairs;  /* pairs that went through the FP64 pass      */
    uint64_t rescued_cells;  /* sum of R*H over them (rodv_phmm_plan_profile) */
    uint64_t launches;       /* kernel launches of the last execution       */
    uint32_t num_buckets;    /* FP32 kernel shapes used                     */
    double fp32_kernel_ms;   /* summed FP32 forward kernel time (profiled)  */
    double fp64_kernel_ms;   /* FP64 rescue kernel time (profiled)          */
    double total_ms;         /* whole execution (profiled)                  */
    uint64_t stacked_pairs;  /* pairs in chained items with two reads per
                                lane group                                  */
    uint64_t shared_pairs;   /* pairs scored by the shared-product FP32
                                kernels (config.shared_products)            */
} rodv_phmm_stats;

typedef struct rodv_phmm_ctx rodv_phmm_ctx;
typedef struct rodv_phmm_plan rodv_phmm_plan;

/* Builds the prior tables once (thread-safe).  Replaces the lazy
 * Context<float/double> construction of context.h:90-166. */
int rodv_phmm_init(void);
const char *rodv_phmm_strerror(int status);
void rodv_phmm_config_default(rodv_phmm_config *cfg);

/* One context per device; owns device tables and scratch. */
int rodv_phmm_ctx_create(int device, rodv_phmm_ctx **out);
void rodv_phmm_ctx_destroy(rodv_phmm_ctx *ctx);

/*
 * Whole-batch forward, synchronous: host inputs, host outputs.  out_log10
 * holds one double per pair (the schedule_pairhmm result values); out_flags
 * (may be NULL) one RODV_PHMM_FLAG_* byte per pair.
 * Replaces schedule_pairhmm (pairhmm_schedule.cpp:116-340).
 */
int rodv_phmm_forward(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_config *cfg, double *out_log10, uint8_t *out_flags);

/*
 * rodv_phmm_forward without device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as rodv_phmm_forward (the same plan and kernels), but every device
 * buffer of the call comes from a fixed scratch array of
 * RODV_PHMM_NOALLOC_SCRATCH_BYTES in device memory, which each device the
 * library runs on holds until the process ends: no hipMalloc or hipFree.  A
 * batch whose buffers do not fit returns RODV_PHMM_ERR_TOO_LARGE (about
 * 400,000 pairs of 150-base reads; reads longer than 378 bases need far more
 * for their boundary rows).  Synchronous (host inputs and outputs), so not
 * for use under graph capture.  Calls take turns on the one scratch (a
 * process-wide lock): overlapping calls wait rather than race.
 */
#define RODV_PHMM_NOALLOC_SCRATCH_BYTES (32u << 20)
int rodv_phmm_forward_noalloc(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                              const rodv_phmm_config *cfg, double *out_log10, uint8_t *out_flags);

/*
 * Two-phase form for callers that score the same batch shape repeatedly or
 * want inputs resident in HBM: plan_create validates, buckets and uploads;
 * plan_execute runs the batched launch on `stream` (a hipStream_t, NULL for
 * the default stream) without host synchronisation; plan_execute_plain runs
 * one launch per region instead; plan_read synchronises and copies results
 * (RODV_PHMM_ERR_DEVICE, after the copies, if the strips of a read longer
 * than one strip, run on several workgroups, waited past their bound for
 * one another; the results of such pairs are NaN).
 */
int rodv_phmm_plan_create(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                          const rodv_phmm_config *cfg, rodv_phmm_plan **out);
int rodv_phmm_plan_execute(rodv_phmm_plan *plan, void *stream);
int rodv_phmm_plan_execute_plain(rodv_phmm_plan *plan, void *stream);
/* Eager execution with a hipEvent around every launch; fills plan stats. */
int rodv_phmm_plan_profile(rodv_phmm_plan *plan, void *stream);
int rodv_phmm_plan_read(rodv_phmm_plan *plan, double *out_log10, uint8_t *out_flags,
                        float *out_raw32, double *out_raw64);
int rodv_phmm_plan_stats(const rodv_phmm_plan *plan, rodv_phmm_stats *out);
uint64_t rodv_phmm_plan_num_pairs(const rodv_phmm_plan *plan);
void rodv_phmm_plan_destroy(rodv_phmm_plan *plan);

/* The engine's prior tables (for parity checks against context.h:90-166):
 * ph2pr[128] and the triangular match-to-match table, first `mm_n` entries. */
int rodv_phmm_tables_f32(float *ph2pr, float *mm, int mm_n);
int rodv_phmm_tables_f64(double *ph2pr, double *mm, int mm_n);

/*
 * Autotuner (no counterpart in the reference: its SIMD widths are fixed by
 * the host CPU).  Reads fall in RODV_PHMM_TUNE_CLASSES length classes
 * (1..15, 16..31, ..., 240..255 rows, then 256..319, then longer).  For each
 * class present in `sample`, rodv_phmm_autotune times the FP32 forward of
 * (up to 131,072 of) the sample's pairs of that class under every candidate
 * shape -- pairs per wavefront 4, 2 or 1 (16, 32 or 64 lanes per pair, rows
 * per lane as few as the read needs), 8 (8 lanes, 6..12 rows per lane) for
 * the classes from 32..47 to 80..95 rows, for the longest class the 64-lane
 * strip width (rows per lane 4..16, shorter strips for longer reads) -- then
 * the persistent waves per CU of the best shape, and keeps the fastest
 * in the context.  Plans created afterwards use it (config use_tuning).  All
 * shapes compute bit-identical results; only speed changes.  The tuning is
 * per context, i.e. per device; set_tuning installs a saved one (NULL
 * restores the built-in shapes) and get_tuning reads the current one.
 */
#define RODV_PHMM_TUNE_CLASSES 18

typedef struct rodv_phmm_tune_entry {
    uint32_t max_read_len;      /* the class's longest read (UINT32_MAX: open)  */
    uint32_t pairs_per_wave;    /* 8, 4, 2 or 1; 0: class not tuned (defaults)  */
    uint32_t max_rows_per_lane; /* strip width cap of 64-lane shapes; 0 = 16    */
    uint32_t waves_per_cu;      /* persistent grid per CU; 0 = engine default   */
    uint32_t sample_pairs;      /* pairs of the sample timed for this class     */
    double gcups;               /* measured with the chosen shape               */
    double default_gcups;       /* measured with the built-in shape             */
} rodv_phmm_tune_entry;

typedef struct rodv_phmm_tuning {
    uint32_t struct_size;       /* sizeof(rodv_phmm_tuning)                    */
    uint32_t num_classes;       /* RODV_PHMM_TUNE_CLASSES                       */
    rodv_phmm_tune_entry cls[RODV_PHMM_TUNE_CLASSES];
} rodv_phmm_tuning;

int rodv_phmm_autotune(rodv_phmm_ctx *ctx, const rodv_phmm_batch *sample,
                       const rodv_phmm_config *cfg, rodv_phmm_tuning *out);
int rodv_phmm_set_tuning(rodv_phmm_ctx *ctx, const rodv_phmm_tuning *tuning);
int rodv_phmm_get_tuning(const rodv_phmm_ctx *ctx, rodv_phmm_tuning *out);

/*
 * Adjoint (no counterpart in the reference).  Parametrised model:
 *   delta = 10^(-gap_open/10), gamma = 10^(-gap_cont/10),
 *   mm = 1 - 2 delta, gapm = 1 - gamma, mx = my = delta, xx = yy = gamma,
 *   e_i = scale[bin(q_i)] * 10^(-q_i/10) with bin(q) = min(q, 47) / 6,
 *   prior = 1 - e_i on a match (or N), e_i / 3 otherwise.
 * Per-row insertion/deletion/gap-continuation qualities are replaced by the
 * two global parameters.  Computes log10 L per pair (FP32, 2^120 scaling) and
 * grad[10] = sum_k weight[k] * d log10 L_k / d(gap_open, gap_cont, scale[0..7]),
 * checkpointing every `ckpt_stride`-th anti-diagonal and recomputing the rest
 * (0 = 32).  kernel_ms (may be NULL) receives the summed device time of the
 * adjoint launches.  Pairs whose FP32 sum underflows run again in FP64 (2^1020
 * scaling).  Reads longer than 378 bases run in FP64 strips of 252 rows, the
 * row handed from strip to strip renormalised by a power of two, so their
 * log10 L and gradient stay finite where a plain FP64 forward underflows; a
 * read of more than 64 such strips (16,128 bases) returns UNSUPPORTED.
 */
typedef struct rodv_phmm_params {
    double gap_open;
    double gap_cont;
    double scale[8];
} rodv_phmm_params;

int rodv_phmm_adjoint(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_params *params, const double *weights, int ckpt_stride,
                      double *out_log10, double *grad, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* RODV_PHMM_H_ */
