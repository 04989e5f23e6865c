/*
 * phmm_oracle.c -- CPU restatement of roDV's PairHMM forward (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this code.
 *
 * What pins each function (see tests/test_oracle.py):
 *  - phmm_oracle_tables_f32/_f64: context.h:44-166.  Pinned against the
 *    closed form 10^(-q/10), against tables dumped from the compiled
 *    reference (oracle/_ref, tests/golden/ref_tables.npz) and by the
 *    KAT/GKL vectors that depend on them.
 *  - phmm_oracle_raw_f32 (intra): pairhmm_impl.cpp:49-174,244-290,295-440.
 *    Pinned bit-for-bit against raw FP32 sums of the compiled reference
 *    (tests/golden/ref_raw_*.npz), the two KATs (pairhmm_unittest.cpp:115,
 *    132), the 104 GKL vectors (resouces/pairhmm-testdata.txt), a closed
 *    form at R = H = 1 and brute-force alignment enumeration for R,H <= 4.
 *  - phmm_oracle_raw_f32 (inter): pairhmm_inter.cpp:80-98,156-305.  Pinned
 *    bit-for-bit against the compiled reference's inter AVX2 group kernel.
 *  - phmm_oracle_raw_f64 / _intra_log10: pairhmm_api.cpp:14-35.  Pinned
 *    against the reference's FP64 path and the GKL vectors.
 *  - phmm_oracle_log10_extended: no reference counterpart (the reference
 *    underflows).  Pinned against raw_f64 where that is representable and
 *    against the exact scaling identity for long reads.
 *  - phmm_oracle_param_*: no reference counterpart (the reference has no
 *    gradient).  Forward pinned against raw_f64 at the reference transition
 *    values; gradient pinned against central finite differences, and the
 *    renormalised form (rows scaled by powers of two, for long reads)
 *    against the plain one where that is representable and against finite
 *    differences where it is not.
 *
 * Build: oracle/Makefile (gcc -O2 -mfma -ffp-contract=off).  fp-contract is
 * off so that only the explicit fmaf() calls below fuse, matching the FMA
 * split GCC emits for the reference with -O3 -mavx512f (vfmadd231ps).
 */
#include "phmm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define MAX_QUAL 254
#define JAC_STEP 0.0001
#define JAC_INV_STEP (1.0 / JAC_STEP)
#define MAX_JAC_TOL 8.0

/* ------------------------------------------------------------------ tables */

static float g_jac_f[PHMM_ORACLE_JAC_SIZE];
static double g_jac_d[PHMM_ORACLE_JAC_SIZE];
static float g_ph2pr_f[128], g_mm_f[PHMM_ORACLE_MM_SIZE];
static double g_ph2pr_d[128], g_mm_d[PHMM_ORACLE_MM_SIZE];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

/* context.h:62-64 fastRound, in the table's own precision. */
static int fast_round_f(float d) { return d > 0.0f ? (int)(d + 0.5f) : (int)(d - 0.5f); }
static int fast_round_d(double d) { return d > 0.0 ? (int)(d + 0.5) : (int)(d - 0.5); }

/* context.h:66-82 approximateLog10SumLog10 with NUMBER = float. */
static float approx_log10_sum_f(float small, float big) {
    if (small > big) {
        float t = big;
        big = small;
        small = t;
    }
    if (isinf(small) || isinf(big)) return big;
    float diff = big - small;
    if (diff >= (float)MAX_JAC_TOL) return big;
    int ind = fast_round_f((float)(diff * (float)JAC_INV_STEP));
    return big + g_jac_f[ind];
}

/* Same with NUMBER = double. */
static double approx_log10_sum_d(double small, double big) {
    if (small > big) {
        double t = big;
        big = small;
        small = t;
    }
    if (isinf(small) || isinf(big)) return big;
    double diff = big - small;
    if (diff >= MAX_JAC_TOL) return big;
    int ind = fast_round_d(diff * JAC_INV_STEP);
    return big + g_jac_d[ind];
}

static void build_tables(void) {
    /* context.h:44-48 jacobianLogTable, computed in double then cast. */
    for (int k = 0; k < PHMM_ORACLE_JAC_SIZE; k++) {
        double v = log10(1.0 + pow(10.0, -((double)k) * JAC_STEP));
        g_jac_d[k] = v;
        g_jac_f[k] = (float)v;
    }
    /* context.h:50-60 matchToMatchProb: the log-sum goes through the
     * NUMBER-typed approximateLog10SumLog10, everything else in double. */
    const double inv_ln10 = 1.0 / log(10);
    for (int i = 0, offset = 0; i <= MAX_QUAL; offset += ++i) {
        for (int j = 0; j <= i; j++) {
            double s_f = approx_log10_sum_f((float)(-0.1 * i), (float)(-0.1 * j));
            double l_f = log1p(-fmin(1.0, pow(10, s_f))) * inv_ln10;
            g_mm_f[offset + j] = (float)pow(10, l_f);
            double s_d = approx_log10_sum_d(-0.1 * i, -0.1 * j);
            double l_d = log1p(-fmin(1.0, pow(10, s_d))) * inv_ln10;
            g_mm_d[offset + j] = pow(10, l_d);
        }
    }
    /* context.h:97-99 (double pow) and :137-139 (float powf). */
    for (int x = 0; x < 128; x++) {
        g_ph2pr_d[x] = pow(10.0, -((double)x) / 10.0);
        g_ph2pr_f[x] = powf(10.f, -((float)x) / 10.f);
    }
}

static void tables_init(void) { pthread_once(&g_once, build_tables); }

void phmm_oracle_tables_f32(float ph2pr[128], float mm[PHMM_ORACLE_MM_SIZE]) {
    tables_init();
    if (ph2pr) memcpy(ph2pr, g_ph2pr_f, sizeof(g_ph2pr_f));
    if (mm) memcpy(mm, g_mm_f, sizeof(g_mm_f));
}

void phmm_oracle_tables_f64(double ph2pr[128], double mm[PHMM_ORACLE_MM_SIZE]) {
    tables_init();
    if (ph2pr) memcpy(ph2pr, g_ph2pr_d, sizeof(g_ph2pr_d));
    if (mm) memcpy(mm, g_mm_d, sizeof(g_mm_d));
}

/* context.h:115-126 / 155-166 set_mm_prob for qualities already masked &127. */
static int mm_index(int ins, int del) {
    int lo = ins <= del ? ins : del;
    int hi = ins <= del ? del : ins;
    return ((hi * (hi + 1)) >> 1) + lo;
}

/* common.h:42-56 ConvertChar restricted to the five valid bases; anything
 * else is rejected by the engine, so the oracle never sees it. */
static int base_code(uint8_t b) {
    switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'T': return 2;
    case 'G': return 3;
    case 'N': return 4;
    default: return 0; /* the reference's zero-initialised table entry */
    }
}

/* ---------------------------------------------------------- FP32 forward */

/* Per-row priors of the FP32 pass.
 * intra: pairhmm_impl.cpp:88-115 and stripe_initialization :151-156
 *        (distm = eps / 3 by IEEE division, 1 - eps).
 * inter: pairhmm_inter.cpp:80-98 (distm = eps * (1/3f)). */
typedef struct {
    float mm, gapm, mx, xx, my, yy, match, mismatch;
} rowp_f;

typedef struct {
    double mm, gapm, mx, xx, my, yy, match, mismatch;
} rowp_d;

static int is_match(const phmm_oracle_pair *p, uint32_t i, uint32_t j, int semantics) {
    uint8_t r = p->read[i], h = p->hap[j];
    if (semantics == PHMM_ORACLE_INTER) return r == h; /* inter/simd_traits_avx2.h:61-63 */
    /* intra: N in the read or the hap matches everything (pairhmm_impl.cpp:26,37-42) */
    int rc = base_code(r), hc = base_code(h);
    return rc == 4 || hc == 4 || rc == hc;
}

/*
 * The recurrence of pairhmm_impl.cpp:266-289 evaluated cell by cell, in the
 * FMA split GCC gives it (verified bit-exact against the compiled reference):
 *   M = fma(D', g, fma(I', g, M' * mm)) * prior
 *   I = fma(I^, xx, M^ * mx)
 *   D = fma(D<, yy, M< * my)
 * with row 0 = (0, 0, 2^120 / H) for every column 0..H and column 0 = 0.
 * The inter kernel (pairhmm_inter.cpp:279-287) contracts the M sum in a
 * different order, verified bit-exact against its AVX2 build:
 *   M = fma(D', g, fma(M', mm, I' * g)) * prior
 * The result is (sum_j M[R][j]) + (sum_j I[R][j]), each summed left to right
 * (pairhmm_impl.cpp:413-414,434-437).
 */
float phmm_oracle_raw_f32(const phmm_oracle_pair *p, int semantics) {
    tables_init();
    const uint32_t R = p->rlen, H = p->hlen;
    if (R == 0 || H == 0) return 0.0f;
    rowp_f *rp = (rowp_f *)malloc(sizeof(rowp_f) * R);
    float *buf = (float *)malloc(sizeof(float) * 6 * (H + 1));
    float *Mp = buf, *Ip = buf + (H + 1), *Dp = buf + 2 * (H + 1);
    float *Mc = buf + 3 * (H + 1), *Ic = buf + 4 * (H + 1), *Dc = buf + 5 * (H + 1);
    const float third = 1.0f / 3.0f;
    for (uint32_t i = 0; i < R; i++) {
        int qi = p->ins[i] & 127, qd = p->del[i] & 127, qc = p->gcp[i] & 127, qq = p->q[i] & 127;
        float eps = g_ph2pr_f[qq];
        rp[i].mm = g_mm_f[mm_index(qi, qd)];
        rp[i].gapm = 1.0f - g_ph2pr_f[qc];
        rp[i].mx = g_ph2pr_f[qi];
        rp[i].xx = g_ph2pr_f[qc];
        rp[i].my = g_ph2pr_f[qd];
        rp[i].yy = g_ph2pr_f[qc];
        rp[i].match = 1.0f - eps;
        rp[i].mismatch = semantics == PHMM_ORACLE_INTER ? eps * third : eps / 3.0f;
    }
    const float init_d = 0x1p120f / (float)H; /* pairhmm_impl.cpp:62 */
    for (uint32_t j = 0; j <= H; j++) {
        Mp[j] = 0.0f;
        Ip[j] = 0.0f;
        Dp[j] = init_d;
    }
    for (uint32_t i = 1; i <= R; i++) {
        const rowp_f *r = &rp[i - 1];
        Mc[0] = Ic[0] = Dc[0] = 0.0f;
        for (uint32_t j = 1; j <= H; j++) {
            float prior = is_match(p, i - 1, j - 1, semantics) ? r->match : r->mismatch;
            float t;
            if (semantics == PHMM_ORACLE_INTER) {
                t = Ip[j - 1] * r->gapm;
                t = fmaf(Mp[j - 1], r->mm, t);
                t = fmaf(Dp[j - 1], r->gapm, t);
            } else {
                t = Mp[j - 1] * r->mm;
                t = fmaf(Ip[j - 1], r->gapm, t);
                t = fmaf(Dp[j - 1], r->gapm, t);
            }
            Mc[j] = t * prior;
            float u = Mp[j] * r->mx;
            Ic[j] = fmaf(Ip[j], r->xx, u);
            float v = Mc[j - 1] * r->my;
            Dc[j] = fmaf(Dc[j - 1], r->yy, v);
        }
        float *t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ip; Ip = Ic; Ic = t;
        t = Dp; Dp = Dc; Dc = t;
    }
    float sm = 0.0f, si = 0.0f;
    for (uint32_t j = 1; j <= H; j++) sm += Mp[j];
    for (uint32_t j = 1; j <= H; j++) si += Ip[j];
    free(buf);
    free(rp);
    return sm + si;
}

static void row_priors_d(const phmm_oracle_pair *p, int semantics, rowp_d *rp) {
    const double third = 1.0 / 3.0;
    for (uint32_t i = 0; i < p->rlen; i++) {
        int qi = p->ins[i] & 127, qd = p->del[i] & 127, qc = p->gcp[i] & 127, qq = p->q[i] & 127;
        double eps = g_ph2pr_d[qq];
        rp[i].mm = g_mm_d[mm_index(qi, qd)];
        rp[i].gapm = 1.0 - g_ph2pr_d[qc];
        rp[i].mx = g_ph2pr_d[qi];
        rp[i].xx = g_ph2pr_d[qc];
        rp[i].my = g_ph2pr_d[qd];
        rp[i].yy = g_ph2pr_d[qc];
        rp[i].match = 1.0 - eps;
        rp[i].mismatch = semantics == PHMM_ORACLE_INTER ? eps * third : eps / 3.0;
    }
}

/*
 * FP64 forward (PairHMMComputer<...DoubleTraits>, K = 2^1020, context.h:101).
 * If `renorm` is set, each finished row is scaled by a power of two so that
 * its largest entry lies in [1, 2); the exponents are accumulated in *log2s.
 */
static double fwd_f64(const phmm_oracle_pair *p, int semantics, int renorm, double init_d,
                      long *log2s) {
    const uint32_t R = p->rlen, H = p->hlen;
    rowp_d *rp = (rowp_d *)malloc(sizeof(rowp_d) * R);
    row_priors_d(p, semantics, rp);
    double *buf = (double *)malloc(sizeof(double) * 6 * (H + 1));
    double *Mp = buf, *Ip = buf + (H + 1), *Dp = buf + 2 * (H + 1);
    double *Mc = buf + 3 * (H + 1), *Ic = buf + 4 * (H + 1), *Dc = buf + 5 * (H + 1);
    long acc = 0;
    for (uint32_t j = 0; j <= H; j++) {
        Mp[j] = 0.0;
        Ip[j] = 0.0;
        Dp[j] = init_d;
    }
    for (uint32_t i = 1; i <= R; i++) {
        const rowp_d *r = &rp[i - 1];
        Mc[0] = Ic[0] = Dc[0] = 0.0;
        double rowmax = 0.0;
        for (uint32_t j = 1; j <= H; j++) {
            double prior = is_match(p, i - 1, j - 1, semantics) ? r->match : r->mismatch;
            double t;
            if (semantics == PHMM_ORACLE_INTER) {
                t = Ip[j - 1] * r->gapm;
                t = fma(Mp[j - 1], r->mm, t);
                t = fma(Dp[j - 1], r->gapm, t);
            } else {
                t = Mp[j - 1] * r->mm;
                t = fma(Ip[j - 1], r->gapm, t);
                t = fma(Dp[j - 1], r->gapm, t);
            }
            Mc[j] = t * prior;
            double u = Mp[j] * r->mx;
            Ic[j] = fma(Ip[j], r->xx, u);
            double v = Mc[j - 1] * r->my;
            Dc[j] = fma(Dc[j - 1], r->yy, v);
            if (renorm) {
                rowmax = fmax(rowmax, Mc[j]);
                rowmax = fmax(rowmax, Ic[j]);
                rowmax = fmax(rowmax, Dc[j]);
            }
        }
        if (renorm && rowmax > 0.0) {
            int e;
            frexp(rowmax, &e); /* rowmax in [2^(e-1), 2^e) */
            int sh = 1 - e;    /* scale so the max lands in [1, 2) */
            if (sh != 0) {
                for (uint32_t j = 1; j <= H; j++) {
                    Mc[j] = ldexp(Mc[j], sh);
                    Ic[j] = ldexp(Ic[j], sh);
                    Dc[j] = ldexp(Dc[j], sh);
                }
                acc -= sh;
            }
        }
        double *t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ip; Ip = Ic; Ic = t;
        t = Dp; Dp = Dc; Dc = t;
    }
    double sm = 0.0, si = 0.0;
    for (uint32_t j = 1; j <= H; j++) sm += Mp[j];
    for (uint32_t j = 1; j <= H; j++) si += Ip[j];
    free(buf);
    free(rp);
    if (log2s) *log2s = acc;
    return sm + si;
}

double phmm_oracle_raw_f64(const phmm_oracle_pair *p, int semantics) {
    tables_init();
    if (p->rlen == 0 || p->hlen == 0) return 0.0;
    return fwd_f64(p, semantics, 0, 0x1p1020 / (double)p->hlen, NULL);
}

double phmm_oracle_log10_extended(const phmm_oracle_pair *p, int semantics) {
    tables_init();
    if (p->rlen == 0 || p->hlen == 0) return -INFINITY;
    long l2 = 0;
    /* start at 1/H instead of 2^1020/H; the scale is tracked exactly */
    double s = fwd_f64(p, semantics, 1, 1.0 / (double)p->hlen, &l2);
    if (!(s > 0.0)) return -INFINITY;
    return log10(s) + (double)l2 * log10(2.0);
}

/* pairhmm_api.cpp:4,9-10,14-35: threshold 1e-28f and the two conversions. */
#define MIN_ACCEPTED 1e-28f
static const float LOG10_K_F = 36.1236000061f;
static const double LOG10_K_D = 307.050595577260822;

double phmm_oracle_intra_log10(const phmm_oracle_pair *p, int use_double, float *raw32,
                               double *raw64, uint8_t *flags) {
    float r32 = use_double ? 0.0f : phmm_oracle_raw_f32(p, PHMM_ORACLE_INTRA);
    uint8_t fl = 0;
    double out;
    double r64 = 0.0;
    if (r32 < MIN_ACCEPTED) {
        r64 = phmm_oracle_raw_f64(p, PHMM_ORACLE_INTRA);
        fl |= PHMM_ORACLE_FLAG_RESCUED;
        if (!(r64 > 0.0)) fl |= PHMM_ORACLE_FLAG_UNDERFLOW;
        out = log10(r64) - LOG10_K_D;
    } else {
        out = (double)(log10f(r32) - LOG10_K_F);
    }
    if (raw32) *raw32 = r32;
    if (raw64) *raw64 = r64;
    if (flags) *flags = fl;
    return out;
}

/* ------------------------------------------------------------ threading */

typedef struct {
    void (*fn)(void *ctx, uint64_t k);
    void *ctx;
    uint64_t n;
    atomic_ullong next;
} par_job;

static void *par_worker(void *arg) {
    par_job *job = (par_job *)arg;
    for (;;) {
        uint64_t k = atomic_fetch_add(&job->next, 1);
        if (k >= job->n) break;
        job->fn(job->ctx, k);
    }
    return NULL;
}

static void par_for(uint64_t n, int nthreads, void (*fn)(void *, uint64_t), void *ctx) {
    par_job job;
    job.fn = fn;
    job.ctx = ctx;
    job.n = n;
    atomic_init(&job.next, 0);
    if (nthreads <= 1 || n < 2) {
        par_worker(&job);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 0; t < nthreads - 1; t++)
        if (pthread_create(&th[started], NULL, par_worker, &job) == 0) started++;
    par_worker(&job);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

/* ---------------------------------------------------------------- batch */

typedef struct {
    const uint8_t *hap_bases, *read_bases, *q, *ins, *del, *gcp;
    const uint64_t *hap_off, *read_off;
    const uint32_t *hap_len, *read_len, *pair_hap, *pair_read;
    int mode, semantics;
    double *out_log10;
    float *out_raw32;
    double *out_raw64;
    uint8_t *out_flags;
    const phmm_oracle_params *th;
    const double *weight;
    double *grads; /* per-pair 10 entries */
} batch_ctx;

static void make_pair(const batch_ctx *c, uint64_t k, phmm_oracle_pair *p) {
    uint32_t h = c->pair_hap[k], r = c->pair_read[k];
    p->hlen = c->hap_len[h];
    p->rlen = c->read_len[r];
    p->hap = c->hap_bases + c->hap_off[h];
    p->read = c->read_bases + c->read_off[r];
    if (c->q) p->q = c->q + c->read_off[r];
    if (c->ins) p->ins = c->ins + c->read_off[r];
    if (c->del) p->del = c->del + c->read_off[r];
    if (c->gcp) p->gcp = c->gcp + c->read_off[r];
}

static void batch_one(void *vc, uint64_t k) {
    const batch_ctx *c = (const batch_ctx *)vc;
    phmm_oracle_pair p;
    make_pair(c, k, &p);
    float r32 = 0.0f;
    double r64 = 0.0, lg = 0.0;
    uint8_t fl = 0;
    switch (c->mode) {
    case PHMM_ORACLE_MODE_INTRA_API:
        lg = phmm_oracle_intra_log10(&p, 0, &r32, &r64, &fl);
        break;
    case PHMM_ORACLE_MODE_RAW_F32:
        r32 = phmm_oracle_raw_f32(&p, c->semantics);
        lg = log10((double)r32) - (double)LOG10_K_F;
        break;
    case PHMM_ORACLE_MODE_RAW_F64:
        r64 = phmm_oracle_raw_f64(&p, c->semantics);
        lg = log10(r64) - LOG10_K_D;
        break;
    case PHMM_ORACLE_MODE_EXTENDED:
        lg = phmm_oracle_log10_extended(&p, c->semantics);
        break;
    default:
        break;
    }
    if (c->out_log10) c->out_log10[k] = lg;
    if (c->out_raw32) c->out_raw32[k] = r32;
    if (c->out_raw64) c->out_raw64[k] = r64;
    if (c->out_flags) c->out_flags[k] = fl;
}

int phmm_oracle_batch(uint32_t num_haps, const uint8_t *hap_bases, const uint64_t *hap_off,
                      const uint32_t *hap_len, uint32_t num_reads, const uint8_t *read_bases,
                      const uint8_t *q, const uint8_t *ins, const uint8_t *del,
                      const uint8_t *gcp, const uint64_t *read_off, const uint32_t *read_len,
                      uint64_t num_pairs, const uint32_t *pair_hap, const uint32_t *pair_read,
                      int mode, int semantics, int nthreads, double *out_log10,
                      float *out_raw32, double *out_raw64, uint8_t *out_flags) {
    (void)num_haps;
    (void)num_reads;
    tables_init();
    if (mode < 0 || mode > PHMM_ORACLE_MODE_EXTENDED) return -1;
    batch_ctx c;
    memset(&c, 0, sizeof(c));
    c.hap_bases = hap_bases;
    c.hap_off = hap_off;
    c.hap_len = hap_len;
    c.read_bases = read_bases;
    c.q = q;
    c.ins = ins;
    c.del = del;
    c.gcp = gcp;
    c.read_off = read_off;
    c.read_len = read_len;
    c.pair_hap = pair_hap;
    c.pair_read = pair_read;
    c.mode = mode;
    c.semantics = semantics;
    c.out_log10 = out_log10;
    c.out_raw32 = out_raw32;
    c.out_raw64 = out_raw64;
    c.out_flags = out_flags;
    par_for(num_pairs, nthreads, batch_one, &c);
    return 0;
}

/* ------------------------------------------------- parametrised adjoint */

static int qbin(int q) { return (q < 47 ? q : 47) / 6; }

/*
 * FP64 forward of the parametrised model on full (R+1) x (H+1) matrices, then
 * the exact reverse sweep (adjoint of every cell equation), accumulating
 * dS/d(delta, gamma, e_bin).
 * renorm == 0: K = 2^1020 as in the FP64 reference pass, no scaling (the
 * sums underflow for reads of a few thousand bases).
 * renorm == 1: row 0 at K = 1, and every finished row i is scaled by 2^sh_i
 * so that its largest entry lies in [1, 2) (fwd_f64's renormalisation).  The
 * scaled rows satisfy the same recurrence with the cross-row terms of row i
 * multiplied by 2^sh_i (the D recurrence along a row is unchanged), so the
 * reverse sweep runs on the scaled entries with those same factors on the
 * adjoints passed from row i to row i-1; S' = S 2^(sh_1 + ... + sh_R), and
 * d log S = dS' / S' since the shifts are constants.  Exact up to entries
 * that are subnormal in one form and not in the other.
 */
static double param_core(const phmm_oracle_pair *p, const phmm_oracle_params *th,
                         double grad[10], int renorm) {
    tables_init();
    const uint32_t R = p->rlen, H = p->hlen;
    if (R == 0 || H == 0) {
        if (grad) memset(grad, 0, 10 * sizeof(double));
        return -INFINITY;
    }
    const double delta = pow(10.0, -th->gap_open / 10.0);
    const double gamma = pow(10.0, -th->gap_cont / 10.0);
    const double mm = 1.0 - 2.0 * delta, gapm = 1.0 - gamma;
    const double mx = delta, my = delta, xx = gamma, yy = gamma;
    const size_t W = (size_t)H + 1;
    const size_t NC = ((size_t)R + 1) * W;
    double *M = (double *)calloc(NC, sizeof(double));
    double *I = (double *)calloc(NC, sizeof(double));
    double *D = (double *)calloc(NC, sizeof(double));
    double *e = (double *)malloc(sizeof(double) * R);
    int *bin = (int *)malloc(sizeof(int) * R);
    int *sh = (int *)calloc((size_t)R + 1, sizeof(int)); /* shift of row i */
    for (uint32_t i = 0; i < R; i++) {
        int qq = p->q[i] & 127;
        bin[i] = qbin(qq);
        e[i] = th->scale[bin[i]] * g_ph2pr_d[qq];
    }
    const double K = renorm ? 1.0 : 0x1p1020;
    long E = 0; /* sum of the shifts */
    for (uint32_t j = 0; j <= H; j++) D[j] = K / (double)H;
#define AT(a, i, j) a[(size_t)(i) * W + (j)]
    for (uint32_t i = 1; i <= R; i++) {
        double rowmax = 0.0;
        for (uint32_t j = 1; j <= H; j++) {
            int m = is_match(p, i - 1, j - 1, PHMM_ORACLE_INTRA);
            double prior = m ? 1.0 - e[i - 1] : e[i - 1] / 3.0;
            double t = mm * AT(M, i - 1, j - 1) + gapm * (AT(I, i - 1, j - 1) + AT(D, i - 1, j - 1));
            AT(M, i, j) = t * prior;
            AT(I, i, j) = mx * AT(M, i - 1, j) + xx * AT(I, i - 1, j);
            AT(D, i, j) = my * AT(M, i, j - 1) + yy * AT(D, i, j - 1);
            rowmax = fmax(rowmax, fmax(AT(M, i, j), fmax(AT(I, i, j), AT(D, i, j))));
        }
        if (renorm && rowmax > 0.0) {
            int ex;
            frexp(rowmax, &ex); /* rowmax in [2^(ex-1), 2^ex) */
            sh[i] = 1 - ex;
            for (uint32_t j = 1; j <= H; j++) {
                AT(M, i, j) = ldexp(AT(M, i, j), sh[i]);
                AT(I, i, j) = ldexp(AT(I, i, j), sh[i]);
                AT(D, i, j) = ldexp(AT(D, i, j), sh[i]);
            }
            E += sh[i];
        }
    }
    double S = 0.0;
    for (uint32_t j = 1; j <= H; j++) S += AT(M, R, j) + AT(I, R, j);
    double lg = log10(S) - log10(K) - (double)E * log10(2.0);
    if (grad) {
        double *Mb = (double *)calloc(NC, sizeof(double));
        double *Ib = (double *)calloc(NC, sizeof(double));
        double *Db = (double *)calloc(NC, sizeof(double));
        double g_delta = 0.0, g_gamma = 0.0, g_e[8] = {0};
        for (uint32_t j = 1; j <= H; j++) {
            AT(Mb, R, j) = 1.0;
            AT(Ib, R, j) = 1.0;
        }
        for (uint32_t i = R; i >= 1; i--) {
            const double f = ldexp(1.0, sh[i]); /* cross-row factor of row i */
            for (uint32_t j = H; j >= 1; j--) {
                /* D[i][j] = my*M[i][j-1] + yy*D[i][j-1] (same row: no factor) */
                double db = AT(Db, i, j);
                AT(Mb, i, j - 1) += my * db;
                AT(Db, i, j - 1) += yy * db;
                g_delta += db * AT(M, i, j - 1);
                g_gamma += db * AT(D, i, j - 1);
                /* I[i][j] = f * (mx*M[i-1][j] + xx*I[i-1][j]) */
                double ib = AT(Ib, i, j) * f;
                AT(Mb, i - 1, j) += mx * ib;
                AT(Ib, i - 1, j) += xx * ib;
                g_delta += ib * AT(M, i - 1, j);
                g_gamma += ib * AT(I, i - 1, j);
                /* M[i][j] = f * prior * (mm*M' + gapm*(I' + D')) */
                int m = is_match(p, i - 1, j - 1, PHMM_ORACLE_INTRA);
                double prior = m ? 1.0 - e[i - 1] : e[i - 1] / 3.0;
                double mb = AT(Mb, i, j) * f;
                double t = mm * AT(M, i - 1, j - 1) +
                           gapm * (AT(I, i - 1, j - 1) + AT(D, i - 1, j - 1));
                double pb = mb * prior;
                AT(Mb, i - 1, j - 1) += mm * pb;
                AT(Ib, i - 1, j - 1) += gapm * pb;
                AT(Db, i - 1, j - 1) += gapm * pb;
                g_delta += -2.0 * pb * AT(M, i - 1, j - 1);
                g_gamma += -pb * (AT(I, i - 1, j - 1) + AT(D, i - 1, j - 1));
                g_e[bin[i - 1]] += mb * t * (m ? -1.0 : 1.0 / 3.0) *
                                   (th->scale[bin[i - 1]] != 0.0 ? e[i - 1] / th->scale[bin[i - 1]] : 0.0);
            }
        }
        /* chain to Phred parameters and to log10 */
        const double ln10 = log(10.0);
        const double inv = 1.0 / (S * ln10);
        grad[0] = g_delta * (-ln10 / 10.0) * delta * inv;
        grad[1] = g_gamma * (-ln10 / 10.0) * gamma * inv;
        for (int b = 0; b < 8; b++) grad[2 + b] = g_e[b] * inv;
        free(Mb);
        free(Ib);
        free(Db);
    }
#undef AT
    free(M);
    free(I);
    free(D);
    free(e);
    free(bin);
    free(sh);
    return lg;
}

double phmm_oracle_param_log10(const phmm_oracle_pair *p, const phmm_oracle_params *th) {
    return param_core(p, th, NULL, 1);
}

double phmm_oracle_param_grad(const phmm_oracle_pair *p, const phmm_oracle_params *th,
                              double grad[10]) {
    return param_core(p, th, grad, 1);
}

static void param_one(void *vc, uint64_t k) {
    const batch_ctx *c = (const batch_ctx *)vc;
    phmm_oracle_pair p;
    make_pair(c, k, &p);
    double lg = param_core(&p, c->th, c->grads ? c->grads + 10 * k : NULL, c->mode);
    if (c->out_log10) c->out_log10[k] = lg;
}

int phmm_oracle_param_batch(uint32_t num_haps, const uint8_t *hap_bases,
                            const uint64_t *hap_off, const uint32_t *hap_len,
                            uint32_t num_reads, const uint8_t *read_bases, const uint8_t *q,
                            const uint64_t *read_off, const uint32_t *read_len,
                            uint64_t num_pairs, const uint32_t *pair_hap,
                            const uint32_t *pair_read, const phmm_oracle_params *th,
                            const double *weight, int nthreads, double *out_log10,
                            double *grad_out, int renorm) {
    (void)num_haps;
    (void)num_reads;
    tables_init();
    batch_ctx c;
    memset(&c, 0, sizeof(c));
    c.mode = renorm ? 1 : 0;
    c.hap_bases = hap_bases;
    c.hap_off = hap_off;
    c.hap_len = hap_len;
    c.read_bases = read_bases;
    c.q = q;
    c.read_off = read_off;
    c.read_len = read_len;
    c.pair_hap = pair_hap;
    c.pair_read = pair_read;
    c.th = th;
    c.out_log10 = out_log10;
    if (grad_out) c.grads = (double *)calloc(num_pairs * 10 + 1, sizeof(double));
    par_for(num_pairs, nthreads, param_one, &c);
    if (grad_out) {
        for (int t = 0; t < 10; t++) grad_out[t] = 0.0;
        for (uint64_t k = 0; k < num_pairs; k++) {
            double w = weight ? weight[k] : 1.0;
            for (int t = 0; t < 10; t++) grad_out[t] += w * c.grads[10 * k + t];
        }
        free(c.grads);
    }
    return 0;
}
