// phmm_host.cpp -- host side of the MI355X PairHMM engine: prior tables,
// validation, length bucketing, device upload, launch sequencing, hipGraph
// capture and the C ABI declared in include/rodv_phmm.h.
#include <cpuid.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

// the library is built with hidden visibility: the C ABI's declarations
// are the only symbols it exports (with the schedule_pairhmm definition)
#pragma GCC visibility push(default)
#include "../../include/rodv_phmm.h"
#pragma GCC visibility pop
#include "phmm_arena.h"
#include "phmm_internal.h"

namespace rodv {
// defined in phmm_kernels.hip
void *fwd_f32_kernel(int L, int MR, bool inter);
void *fwd_f64_rescue_kernel(bool inter);
void *fwd_tab_kernel_ptr(int L, int MR, bool inter, bool share);
void *fwd_chain_kernel_ptr(int L, int MR, bool inter, int S, bool share);
uint32_t fwd_tab_bytes(int L, int MR);
int fwd_f64_rescue_mr();
int fwd_long_waves();
int rescue_long_waves();
int rescue_long_rows();
int rescue_long_mr();
hipError_t launch_rescue_long(bool route, bool tab, int waves, int grid, size_t lds_bytes, hipStream_t st,
                              const FwdArgs &a);
uint32_t long_tab_bytes(int MR, bool fp64, int waves);
void *fwd_long_kernel_ptr(int MR, bool inter, bool tab);
hipError_t launch_fwd_block(void *kern, int grid, int block, size_t lds_bytes, hipStream_t st, const FwdArgs &a);
hipError_t launch_fwd(void *kern, int grid, size_t lds_bytes, hipStream_t st, const FwdArgs &a);
hipError_t launch_finalize(const FinalizeArgs &f, hipStream_t st);
hipError_t launch_reset(uint32_t *counters, uint32_t n, hipStream_t st);
hipError_t launch_rescued_cells(const FwdArgs &a, unsigned long long *out, hipStream_t st);
hipError_t launch_desc(const FwdArgs &a, const int32_t *items, ItemDesc *out, uint64_t n,
                       hipStream_t st);
hipError_t launch_regroup(const RegroupArgs &r, hipStream_t st);
uint32_t rescue_tab_bytes(int MR);
hipError_t launch_rescue_bin(const FwdArgs &a, int grid, hipStream_t st);
hipError_t launch_rescue_tab(int part, bool share, int grid, size_t lds_bytes, hipStream_t st,
                             const FwdArgs &a);
hipError_t launch_ext(int grid, size_t lds_bytes, hipStream_t st, const FwdArgs &a);
int adj2_max_mr();
void *adj2_kernel_ptr(int L, int MR, bool rescue);
hipError_t launch_adj(void *kern, int grid, size_t lds_bytes, hipStream_t st, const AdjArgs &a);
int adj_long_mr();
int adj_long_max_strips();
int adj_long_waves();
size_t adj_long_lds(uint32_t lds_stride, int K1, int waves);
hipError_t launch_adj_long(int waves, int grid, size_t lds_bytes, hipStream_t st, const AdjArgs &a);
size_t adj_longb_lds(uint32_t lds_stride, int K1);
hipError_t launch_adj_long_split(int grid_a, int grid_b, int grid_c, size_t lds_b, hipStream_t st, const AdjArgs &a);
size_t adj_longab_lds(uint32_t lds_stride, int K1);
int adj_longab_waves();
uint64_t adj_long_ck_elems(uint32_t H, int K1);
hipError_t launch_adj_long_fused(int grid_ab, int grid_c, size_t lds_ab, hipStream_t st, const AdjArgs &a);
}  // namespace rodv

using namespace rodv;

namespace {

// ------------------------------------------------------------------ tables
// Restates Context<float>/Context<double> (context.h:21-166): ph2pr by
// powf/pow, the 80,001-entry Jacobian table in double cast to NUMBER, and the
// match-to-match table whose log-sum runs in NUMBER precision.

constexpr int kMaxQual = 254;
constexpr int kJacSize = 80001;
constexpr int kMMSize = (kMaxQual + 1) * (kMaxQual + 2) / 2;

struct HostTables {
    float ph2pr_f[128];
    double ph2pr_d[128];
    std::vector<float> mm_f;
    std::vector<double> mm_d;
    DevTables dev;
};

HostTables *g_tables = nullptr;
std::once_flag g_tables_once;

template <typename N>
N approx_log10_sum(N small, N big, const std::vector<N> &jac) {
    if (small > big) std::swap(small, big);
    if (std::isinf(small) || std::isinf(big)) return big;
    N diff = big - small;
    if (diff >= (N)8.0) return big;
    N x = (N)(diff * (N)(1.0 / 0.0001));
    int ind = x > (N)0 ? (int)(x + (N)0.5) : (int)(x - (N)0.5);
    return big + jac[ind];
}

void build_tables() {
    HostTables *t = new HostTables();
    std::vector<float> jf(kJacSize);
    std::vector<double> jd(kJacSize);
    for (int k = 0; k < kJacSize; k++) {
        double v = std::log10(1.0 + std::pow(10.0, -((double)k) * 0.0001));
        jd[k] = v;
        jf[k] = (float)v;
    }
    t->mm_f.resize(kMMSize);
    t->mm_d.resize(kMMSize);
    const double inv_ln10 = 1.0 / std::log(10.0);
    for (int i = 0, offset = 0; i <= kMaxQual; offset += ++i) {
        for (int j = 0; j <= i; j++) {
            double sf = approx_log10_sum<float>((float)(-0.1 * i), (float)(-0.1 * j), jf);
            t->mm_f[offset + j] =
                (float)std::pow(10.0, std::log1p(-std::min(1.0, std::pow(10.0, sf))) * inv_ln10);
            double sd = approx_log10_sum<double>(-0.1 * i, -0.1 * j, jd);
            t->mm_d[offset + j] =
                std::pow(10.0, std::log1p(-std::min(1.0, std::pow(10.0, sd))) * inv_ln10);
        }
    }
    const float third_f = 1.0f / 3.0f;
    const double third_d = 1.0 / 3.0;
    for (int x = 0; x < 128; x++) {
        t->ph2pr_f[x] = powf(10.f, -((float)x) / 10.f);
        t->ph2pr_d[x] = std::pow(10.0, -((double)x) / 10.0);
        DevTables &d = t->dev;
        d.ph2pr_f[x] = t->ph2pr_f[x];
        d.omp_f[x] = 1.0f - t->ph2pr_f[x];
        d.p3_f[x] = t->ph2pr_f[x] / 3.0f;
        d.pt_f[x] = t->ph2pr_f[x] * third_f;
        d.ph2pr_d[x] = t->ph2pr_d[x];
        d.omp_d[x] = 1.0 - t->ph2pr_d[x];
        d.p3_d[x] = t->ph2pr_d[x] / 3.0;
        d.pt_d[x] = t->ph2pr_d[x] * third_d;
    }
    for (int k = 0; k < kMMQ; k++) {
        t->dev.mm_f[k] = t->mm_f[k];
        t->dev.mm_d[k] = t->mm_d[k];
    }
    g_tables = t;
}

const HostTables &tables() {
    std::call_once(g_tables_once, build_tables);
    return *g_tables;
}

// ------------------------------------------------------------------ helpers

bool fail_alloc_hook() {
    const char *e = std::getenv("RODV_PHMM_FAIL_ALLOC");
    return e && e[0] == '1';
}

// Polls of one hand-over wait of the long-read kernels: 2^20 (about a
// second; the waits take microseconds to milliseconds).  Tests force the
// bound to run out with RODV_PHMM_SPIN_LIMIT=0, read when a plan is created
// and when an adjoint runs.
uint32_t spin_limit_hook() {
    const char *e = std::getenv("RODV_PHMM_SPIN_LIMIT");
    return e && e[0] ? (uint32_t)std::strtoul(e, nullptr, 10) : (1u << 20);
}

// Waves per workgroup of the FP64 long-read rescue: tests compare the
// 8-wave units (RODV_PHMM_RESCUE_LONG_WAVES=8) with the default 4, bit for
// bit; read when a plan is created.
int rescue_long_waves_hook() {
    const char *e = std::getenv("RODV_PHMM_RESCUE_LONG_WAVES");
    return e && e[0] ? (int)std::strtol(e, nullptr, 10) : 0;
}

// The long-read adjoint on one workgroup per pair of 1 or 4 waves
// (RODV_PHMM_ADJ_LONG_WAVES=1 or 4) instead of the default, a pair's strips
// over several workgroups: tests compare them bit for bit.
int adj_long_waves_hook() {
    const char *e = std::getenv("RODV_PHMM_ADJ_LONG_WAVES");
    return e && e[0] ? (int)std::strtol(e, nullptr, 10) : 0;
}

// Long pairs per chunk of the split long-read adjoint
// (RODV_PHMM_ADJ_LONG_CHUNK=n; default: as many as 8 GiB of buffers hold):
// tests run several small chunks and compare them with one, bit for bit.
uint64_t adj_long_chunk_hook() {
    const char *e = std::getenv("RODV_PHMM_ADJ_LONG_CHUNK");
    return e && e[0] ? (uint64_t)std::strtoull(e, nullptr, 10) : 0;
}

// The split long-read adjoint in three launches (RODV_PHMM_ADJ_LONG_LAUNCHES=3:
// pass A, pass B, the totals) instead of two (passes A and B in one): tests
// compare them bit for bit.
int adj_long_launches_hook() {
    const char *e = std::getenv("RODV_PHMM_ADJ_LONG_LAUNCHES");
    return e && e[0] ? (int)std::strtol(e, nullptr, 10) : 0;
}

// Pass B of the split long-read adjoint recomputes its level-1 checkpoints
// instead of reading pass A's (RODV_PHMM_ADJ_LONG_PASSA_CKPT=0), as it does
// where they would not fit: tests compare the two bit for bit.
bool adj_long_passa_ckpt_hook() {
    const char *e = std::getenv("RODV_PHMM_ADJ_LONG_PASSA_CKPT");
    return !(e && e[0] == '0');
}

template <typename T>
int dev_alloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    if (fail_alloc_hook()) return RODV_PHMM_ERR_OOM;
    if (hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return RODV_PHMM_ERR_OOM;
    }
    return RODV_PHMM_OK;
}

template <typename T>
int upload(T **p, const T *src, size_t n) {
    int rc = dev_alloc(p, n);
    if (rc) return rc;
    if (n && hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return RODV_PHMM_ERR_DEVICE;
    return RODV_PHMM_OK;
}

inline uint8_t hap_code(uint8_t b, bool *ok) {
    switch (b) {
    case 'A': return 1;
    case 'C': return 2;
    case 'G': return 4;
    case 'T': return 8;
    case 'N': return kHapN;
    default: *ok = false; return 0;
    }
}

inline uint8_t read_code(uint8_t b, bool *ok) {
    switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    case 'N': return kReadN;
    default: *ok = false; return 0;
    }
}

constexpr int kUnroll = 8;  // must match kU in phmm_kernels.hip
// persistent waves per CU when the config leaves it to the engine: 12 gave
// 2 % more than 8 on the batched uneven workload (LDS caps the long-read
// shapes near 7, so 16 adds nothing)
constexpr int kDefaultWavesPerCU = 12;

// Read-length classes of the autotuner: class k (0-based) holds reads of
// 16k .. 16k+15 rows for k < 16 (the 16-lane table shapes' rows-per-lane
// steps: 16 lanes of MR rows hold up to 16*MR - 1), then 256..319 (the
// longest 32-lane table shape) and longer.
constexpr uint32_t kClassMax[RODV_PHMM_TUNE_CLASSES] = {
    15, 31, 47, 63, 79, 95, 111, 127, 143, 159, 175, 191, 207, 223, 239, 255, 319, 0xffffffffu};

int tune_class(uint32_t R) {
    int c = 0;
    while (c < RODV_PHMM_TUNE_CLASSES - 1 && R > kClassMax[c]) c++;
    return c;
}

// A class's shape: lanes per pair (16, 32, 64; 0 = engine default), the cap
// on rows per lane of the 64-lane shapes (0 = 16), persistent waves per CU
// (0 = engine default).
struct TuneEntry {
    int L = 0, mr_cap = 0, wpc = 0;
};

struct Bucket {
    int L = 16, MR = 1;
    int wpc = 0;
    bool inter = false;
    bool tab = false;        // table-prior kernel (N-free haplotypes, one strip)
    bool chain = false;      // chained items (fwd_chain_kernel); tab is set too
    bool share = false;      // shared-product kernel variant: every read's
                             // rows have my_i == mx_{i+1} (table kernels)
    int stack = 1;           // chained items: reads per group (1, or 2 stacked in its rows)
    uint32_t tab_bytes = 0;  // LDS prior table per group
    void *kern = nullptr;
    std::vector<int32_t> items;  // G entries per item (G * stack for chained items)
    // chained items: item i's haplotypes are chain_hap[chain_off[i] ..
    // chain_off[i+1]), items holds the G pairs of its first haplotype
    std::vector<uint32_t> chain_off;
    std::vector<ChainHap> chain_hap;
    uint32_t *d_chain_off = nullptr;
    ChainHap *d_chain_hap = nullptr;
    uint32_t n_items = 0;
    uint32_t max_h = 0;
    uint32_t max_strips = 1;
    uint32_t lds_stride = 0;
    int32_t *d_items = nullptr;
    ItemDesc *d_desc = nullptr;  // table kernel: one descriptor per item slot
    float *d_bnd = nullptr;
    uint64_t bnd_elems = 0;
    int grid = 0;
    int grid_cap = 1;  // persistent grid of a full launch (plain launches are capped by it)
    int block = 64;    // threads per workgroup (fwd_long_kernel: several waves)
    uint32_t nbnd = 2; // boundary-row buffers per workgroup
    uint64_t cells = 0;
    // fwd_long_kernel, batched: the items' pairs as units of strips, whose
    // slots follow the grid's boundary rows in d_bnd and take counters
    // prog_base .. prog_base + n_xslots - 1
    std::vector<LongUnit> units;
    LongUnit *d_units = nullptr;
    uint32_t n_xslots = 0;
    uint32_t prog_base = 0;
    // the FP32 pass of long reads in two launches: units1 runs strip 0 of
    // every pair alone (a one-strip pair whole), handing its last row to a
    // slot; a pair whose row is all zero there is rescued with a zero sum,
    // and `units` (strips 1.. of the others, from that slot) skips it
    std::vector<LongUnit> units1;
    LongUnit *d_units1 = nullptr;
    void *kern1 = nullptr;  // units1's kernel: kSubMR rows per lane in the bucket's layout
    uint32_t counter1 = 0;  // work counter of the units1 launch
};

// one launch of the plain (per-region) schedule
struct PlainLaunch {
    int bucket;
    uint32_t item_begin, n_items;
    int grid;
};

}  // namespace

// FP32 buckets run on the launch stream and this many more streams, in turn.
// One: with two, the captured graph's replays slowed down by up to 3 % on
// some plans (more streams than the device's hardware queues comfortably
// serve), while one more stream gives the same overlap with stable replays.
constexpr int kSideStreams = 1;

struct rodv_phmm_ctx {
    int device = 0;
    int num_cus = 256;
    DevTables *d_tab = nullptr;
    // autotuned shapes per read-length class (rodv_phmm_autotune)
    bool tuned = false;
    TuneEntry tune[RODV_PHMM_TUNE_CLASSES];
    // forced entry while the tuner measures a candidate
    int force_class = -1;
    TuneEntry force{};
    // more streams for the FP32 buckets (default priority) and the events
    // that fork them from and join them to the launch stream
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSideStreams] = {};
    // rodv_phmm_forward's plans: device block and pinned mirror kept from
    // call to call (grown on demand, freed with the context)
    Arena fwd_arena;
};

struct rodv_phmm_plan {
    rodv_phmm_ctx *ctx = nullptr;
    rodv_phmm_config cfg{};
    uint64_t n_pairs = 0;
    uint64_t cells = 0;
    // device inputs
    uint8_t *d_hapc = nullptr;
    uint64_t *d_hapc_off = nullptr;
    uint32_t *d_hap_len = nullptr;
    float *d_initd_f = nullptr;
    double *d_initd_d = nullptr;
    uint8_t *d_rcode = nullptr, *d_q = nullptr, *d_ins = nullptr, *d_del = nullptr,
            *d_gcp = nullptr;
    uint64_t *d_read_off = nullptr;
    uint32_t *d_read_len = nullptr;
    uint32_t *d_pair_hap = nullptr, *d_pair_read = nullptr;
    uint8_t *d_pair_inter = nullptr;
    // schedule_pairhmm routing: per region, its FP32 inter pairs in sorted order
    uint32_t *d_sched_list = nullptr;
    uint64_t *d_sched_off = nullptr;
    uint32_t *d_sched_scratch = nullptr;
    uint32_t n_regions = 0;
    // device outputs / scratch
    float *d_raw32 = nullptr;
    double *d_raw64 = nullptr;
    double *d_out = nullptr;
    uint8_t *d_flags = nullptr;
    uint32_t *d_rescue = nullptr;
    uint32_t *d_counters = nullptr;  // [0] rescue count, [1] rescue work, [2..] per launch
    unsigned long long *d_rescued_cells = nullptr;  // plan_profile's sum of R*H over the rescue list
    uint32_t n_counters = 0;
    double *d_bnd64 = nullptr;  // boundary rows of the long-read FP64 rescue
    uint64_t bnd64_elems = 0;
    int rescue_grid = 0;
    int long_grid = 0;          // workgroups of fwd_rescue_long_kernel (0: no long reads)
    int rescue64_waves = 4;     // their waves (rescue_long_waves(), or fwd_long_waves() for tests)
    bool share_all = false;     // every read qualifies for shared products: the
                                // FP64 rescue table kernels use them
    bool long_tab = false;      // every haplotype N-free: the long-read kernels
                                // read their priors from LDS tables
    uint32_t long_counter = 0;  // its work counter
    // its units: every pair whose read is longer than one FP64 strip, the
    // longest pipelines first; slots after the grid's rows in d_bnd64
    std::vector<LongUnit> long_units;
    LongUnit *d_long_units = nullptr;
    uint32_t n_x64 = 0;
    uint32_t prog64_base = 0;
    // extended range: per pair the power of two of its FP64 sum, the
    // extended-range pass's grid, work counter and boundary rows (2 per wave)
    int32_t *d_exp64 = nullptr;
    int ext_grid = 0;
    uint32_t ext_counter = 0;
    double *d_bndx = nullptr;
    uint64_t bndx_elems = 0;
    uint32_t fault_counter = 0;  // hand-over waits of the long-read kernels that ran out
    uint32_t spin_limit = 1u << 20;
    uint32_t rescue_lds = 0;
    // FP64 rescue in classes (FwdArgs::bin): the bin list, per haplotype its
    // N flag, the classes' offsets and capacities, the counters' base in
    // d_counters (2 * kRescueClasses + 1 bin counts, then the two table
    // launches' work counters), the table launches' grids and code stride
    uint32_t *d_bin = nullptr;
    uint8_t *d_hap_n = nullptr;
    uint32_t bin_base[kRescueClasses + 1] = {};
    uint32_t bin_cap[kRescueClasses + 1] = {};
    uint32_t bin_counter = 0;
    int rtab_grid[2] = {0, 0};
    uint32_t rtab_stride = 0;
    // host copies needed at launch
    std::vector<Bucket> buckets;
    std::vector<PlainLaunch> plain;
    std::vector<int32_t> plain_items_all;  // per bucket, region-ordered items
    std::vector<int32_t *> d_plain_items;
    std::vector<size_t> plain_items_n;
    std::vector<ItemDesc *> d_plain_desc;
    std::vector<uint32_t> pair_R, pair_H;
    bool any_inter = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t last_stream = nullptr;
    rodv_phmm_stats stats{};
    // buffers carved from this arena (nullptr: each one allocated)
    Arena *arena = nullptr;
    // no per-region launch list (plans of rodv_phmm_forward, run batched once)
    bool no_plain = false;
};

namespace {

#define HIPCHK(x)                                          \
    do {                                                   \
        if ((x) != hipSuccess) {                           \
            (void)hipGetLastError();                       \
            return RODV_PHMM_ERR_DEVICE;                   \
        }                                                  \
    } while (0)

void free_plan(rodv_phmm_plan *p) {
    if (!p) return;
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->arena) {  // every buffer is arena memory, which its owner keeps
        delete p;
        return;
    }
    void *bufs[] = {p->d_hapc, p->d_hapc_off, p->d_hap_len, p->d_initd_f, p->d_initd_d,
                    p->d_rcode, p->d_q, p->d_ins, p->d_del, p->d_gcp, p->d_read_off,
                    p->d_read_len, p->d_pair_hap, p->d_pair_read, p->d_pair_inter,
                    p->d_raw32, p->d_raw64, p->d_out, p->d_flags, p->d_rescue,
                    p->d_counters, p->d_bnd64, p->d_sched_list, p->d_sched_off, p->d_sched_scratch,
                    p->d_rescued_cells, p->d_exp64, p->d_bndx, p->d_bin, p->d_hap_n};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &b : p->buckets) {
        if (b.d_items) (void)hipFree(b.d_items);
        if (b.d_desc) (void)hipFree(b.d_desc);
        if (b.d_chain_off) (void)hipFree(b.d_chain_off);
        if (b.d_chain_hap) (void)hipFree(b.d_chain_hap);
        if (b.d_bnd) (void)hipFree(b.d_bnd);
        if (b.d_units) (void)hipFree(b.d_units);
        if (b.d_units1) (void)hipFree(b.d_units1);
    }
    if (p->d_long_units) (void)hipFree(p->d_long_units);
    for (auto *d : p->d_plain_items)
        if (d) (void)hipFree(d);
    for (auto *d : p->d_plain_desc)
        if (d) (void)hipFree(d);
    delete p;
}

// A plan's device buffer of n elements: from its arena when it has one
// (256-byte aligned; kArenaFull when the arena is too small), else its own
// allocation.
template <typename T>
int plan_alloc(rodv_phmm_plan *p, T **out, size_t n) {
    if (!p->arena) return dev_alloc(out, n);
    *out = nullptr;
    Arena &A = *p->arena;
    const size_t bytes = ((n ? n : 1) * sizeof(T) + 255) & ~(size_t)255;
    A.need += bytes;
    if (A.used + bytes > A.cap) return kArenaFull;
    *out = reinterpret_cast<T *>(A.dev + A.used);
    A.used += bytes;
    return RODV_PHMM_OK;
}

template <typename T>
int plan_upload(rodv_phmm_plan *p, T **out, const T *src, size_t n) {
    if (!p->arena) return upload(out, src, n);
    int rc = plan_alloc(p, out, n);
    if (rc) return rc;
    if (n) std::memcpy(p->arena->host + (reinterpret_cast<uint8_t *>(*out) - p->arena->dev), src, n * sizeof(T));
    return RODV_PHMM_OK;
}

// bytes of a plan buffer set to `v` (in the arena's mirror, else on the device)
int plan_fill(rodv_phmm_plan *p, void *dst, int v, size_t bytes) {
    if (!bytes) return RODV_PHMM_OK;
    if (p->arena) {
        std::memset(p->arena->host + (reinterpret_cast<uint8_t *>(dst) - p->arena->dev), v, bytes);
        return RODV_PHMM_OK;
    }
    HIPCHK(hipMemset(dst, v, bytes));
    return RODV_PHMM_OK;
}

// the arena mirror's used prefix to the device, in one copy
int plan_flush(rodv_phmm_plan *p) {
    if (!p->arena || !p->arena->used) return RODV_PHMM_OK;
    HIPCHK(hipMemcpy(p->arena->dev, p->arena->host, p->arena->used, hipMemcpyHostToDevice));
    return RODV_PHMM_OK;
}

// (L, MR) of the select-prior FP32 kernel for a read of R rows.  A strip
// holds (L-1)*MR read rows (lane L-1 is the identity/feeder lane); longer
// reads run strip by strip in the L = 64 kernels.
void choose_shape(uint32_t R, int short_lanes, int *L, int *MR) {
    if (short_lanes != 64 && R <= 15u * 16u) {
        *L = 16;
        *MR = std::max(1, (int)((R + 14) / 15));
        return;
    }
    *L = 64;
    int mr = (int)((R + 62) / 63);
    mr = ((mr + 3) / 4) * 4;
    *MR = std::min(16, std::max(4, mr));
}

// Rows per lane of the table-prior kernel with L lanes for a read of R rows,
// or 0 when no instantiated shape (8 lanes: 6..12 rows per lane, 16: 1..16,
// 32: 2..10, 64: 4, 8, 12, 16 up to `cap`) holds it in one strip.  Every lane
// carries read rows and the row-0 boundary needs one identity slot, so a
// strip of L*MR slots holds R <= L*MR - 1 rows.
int tab_rows(uint32_t R, int L, int cap) {
    int mr = (int)((R + (uint32_t)L) / (uint32_t)L);
    if (L == 8) return mr >= 6 && mr <= 12 ? mr : 0;
    if (L == 16) return mr <= 16 ? mr : 0;
    if (L == 32) {
        mr = std::max(2, mr);
        return mr <= 10 ? mr : 0;
    }
    mr = std::max(4, (mr + 3) / 4 * 4);
    return mr <= (cap > 0 ? cap : 16) ? mr : 0;
}

// Lanes per pair of the built-in table shape: 16 while 16 lanes hold the
// read in one strip, else 64.  (8 lanes for reads of 40..95 rows are an
// autotuner candidate only: faster on uniform batches, but on batches with
// mixed haplotype lengths as fast as 16 lanes at best and slower when the
// batch is small, since an item then holds 8 pairs.)
int default_tab_lanes(uint32_t R, int short_lanes) {
    if (short_lanes == 64) return 64;
    return tab_rows(R, 16, 0) ? 16 : 64;
}

// Pairs the autotuner times per length class: enough to give every
// persistent wave of the default grid four items of the largest (8-pair)
// kind.
uint64_t tune_pairs_min(int num_cus) { return (uint64_t)num_cus * kDefaultWavesPerCU * 8u * 4u; }

// The groups of one wave stage their haplotype codes (and prior tables) in
// LDS; a very long haplotype does not fit four or two of them.
bool codes_fit(uint32_t H, int L, int MR);

// Shape and kernel of one pair: the class's tuned lanes per pair when the
// context has one (rows per lane then as few as the read needs), else the
// built-in shape.  The table kernel takes the pair when `tab_ok` (N-free
// haplotype, prior tables on) and the read fits one table strip; otherwise
// the select-prior kernel, whose 32-lane shapes do not exist, takes it.
void shape_for(uint32_t R, uint32_t H, const rodv_phmm_config &cfg, const TuneEntry *te, bool tab_ok,
               int *L, int *MR, bool *tab) {
    const int want = te ? te->L : 0;
    const int cap = want == 64 ? te->mr_cap : 0;
    if (tab_ok) {
        const int l0 = default_tab_lanes(R, cfg.short_lanes);
        int l = want ? want : l0;
        int mr = tab_rows(R, l, cap);
        if (mr == 0 && l != l0) mr = tab_rows(R, l = l0, 0);
        if (mr == 0 && l != 64) mr = tab_rows(R, l = 64, 0);
        if (mr && !codes_fit(H, l, mr)) mr = tab_rows(R, l = 64, cap);
        if (mr) {
            *L = l;
            *MR = mr;
            *tab = true;
            return;
        }
    }
    *tab = false;
    if (want == 16 && R <= 240u) {
        *L = 16;
        *MR = std::max(1, (int)((R + 14) / 15));
    } else if (want == 64) {
        int mr = (int)((R + 62) / 63);
        mr = ((mr + 3) / 4) * 4;
        *L = 64;
        *MR = std::min(cap > 0 ? cap : 16, std::max(4, mr));
    } else {
        choose_shape(R, cfg.short_lanes, L, MR);
    }
    if (!codes_fit(H, *L, *MR)) {
        *L = 64;
        *MR = 4;
        while (*MR < 16 && 63 * *MR < (int)R) *MR += 4;
    }
}

uint32_t lds_stride_for(uint32_t max_h, int L) {
    uint32_t s = max_h + 3u * (uint32_t)L + 2u * kUnroll + 32u;
    return (s + 15u) & ~15u;
}

bool codes_fit(uint32_t H, int L, int MR) {
    if (L == 64) return true;
    const uint64_t per = lds_stride_for(H, L) + 256u + fwd_tab_bytes(L, MR);
    return (64ull / (uint64_t)L) * per <= 60u * 1024u;
}

// Rows per lane of a 16-lane group that holds two reads, of RA rows above
// and RB rows below (fwd_chain_kernel's two-read layout: B ends in the last
// slot, A in the last slot of lane (16 MR - RB - 1) / MR - 1, and the group's
// first slot is an identity slot, which takes what the last lane rotates
// into lane 0); 0 if more than 16 would be needed.  Every MR above the
// returned one fits too.
int stack_rows(uint32_t RA, uint32_t RB) {
    for (int m = 1; m <= 16; m++) {
        const int lo = (16 * m - (int)RB - 1) / m - 1;
        if (lo >= 0 && (lo + 1) * m >= (int)RA + 1) return m;
    }
    return 0;
}

// Issue efficiency of rows read rows in a 16-lane group of MR rows per lane:
// the padding to 16 MR slots times the step's 8 MR cell instructions over
// those plus ~5 per step that do not scale with MR (the code address, the
// row-above D, the two sums).
double group_eff(uint32_t rows, int MR) {
    return (double)rows / (16.0 * MR) * (8.0 * MR) / (8.0 * MR + 5.0);
}

// Chained items of one table bucket: reads of a region whose pairs in the
// bucket cover the same haplotypes (two or more, each at least L + 2 long,
// so that the kernel's hand-overs never overlap), G groups of S reads per
// item, swept across those haplotypes longest first.  A chain is cut into
// segments of at most kseg haplotypes, kseg as large as leaves at least
// `min_items` items (0: no such bound) and at most kChainMaxHaps; a bucket
// too small for kseg >= 2 keeps single pairs (long items in few waves would
// leave the device idle).
// `mem` holds the bucket's pairs; on return only those left out.  S = 1
// fills `bk`: items (the G pairs of each item's first haplotype), chain_off,
// chain_hap, max_h and cells; items are ordered by columns, most first.
// S = 2 (16 lanes): each group stacks two reads, an item takes the rows per
// lane its groups need and goes to the bucket of `stacked` with that MR
// (created from bk's L, inter and wpc); an item is formed only where its
// group_eff beats the reads' one-read shapes by 3 %, the reads of the others
// stay in `mem`.
constexpr uint64_t kChainMaxHaps = 16;
constexpr uint64_t kChainItemsPerWave = 16;
void form_chains(std::vector<uint32_t> &mem, Bucket &bk, const uint32_t *ph, const uint32_t *pr,
                 const std::vector<uint32_t> &pair_region, const uint32_t *hap_len,
                 const uint32_t *read_len, const std::vector<uint64_t> &hapc_off,
                 const std::vector<float> &initd_f, uint64_t min_items, int S = 1,
                 std::vector<Bucket> *stacked = nullptr) {
    const size_t G = (size_t)(64 / bk.L);
    const uint32_t hmin = (uint32_t)bk.L + 2u;
    // LDS of the chain kernel: G prior tables and two code buffers, <= 60 KiB
    const int64_t room = ((61440 - (int64_t)G * fwd_tab_bytes(bk.L, S == 1 ? bk.MR : 16)) / 2) & ~(int64_t)255;
    const int64_t hmax = room - (int64_t)lds_stride_for(0, bk.L) - 15;
    std::sort(mem.begin(), mem.end(), [&](uint32_t x, uint32_t y) {
        if (pair_region[x] != pair_region[y]) return pair_region[x] < pair_region[y];
        if (pr[x] != pr[y]) return pr[x] < pr[y];
        const uint32_t hx = hap_len[ph[x]], hy = hap_len[ph[y]];
        if (hx != hy) return hx > hy;
        return ph[x] < ph[y];
    });
    // runs: one read's pairs, [b, e) in mem; sig hashes the haplotype list
    struct Run {
        uint32_t b, e;
        uint64_t sig;
    };
    std::vector<Run> runs;
    std::vector<uint32_t> single;
    for (size_t i = 0; i < mem.size();) {
        size_t j = i + 1;
        while (j < mem.size() && pr[mem[j]] == pr[mem[i]] && pair_region[mem[j]] == pair_region[mem[i]]) j++;
        bool ok = j - i >= 2;
        uint64_t sig = 1469598103934665603ull ^ (uint64_t)pair_region[mem[i]];
        for (size_t k = i; k < j && ok; k++) {
            ok = hap_len[ph[mem[k]]] >= hmin && (int64_t)hap_len[ph[mem[k]]] <= hmax;
            sig = (sig ^ ph[mem[k]]) * 1099511628211ull;
            sig = (sig ^ (uint32_t)(mem[k] - mem[i])) * 1099511628211ull;
        }
        if (ok)
            runs.push_back({(uint32_t)i, (uint32_t)j, sig});
        else
            single.insert(single.end(), mem.begin() + i, mem.begin() + j);
        i = j;
    }
    std::sort(runs.begin(), runs.end(), [](const Run &x, const Run &y) {
        return x.sig != y.sig ? x.sig < y.sig : x.b < y.b;
    });
    // same haplotypes in the same order, and the same pair offsets
    auto same = [&](const Run &x, const Run &y) {
        if (x.e - x.b != y.e - y.b) return false;
        for (uint32_t k = 0; k < x.e - x.b; k++) {
            if (ph[mem[x.b + k]] != ph[mem[y.b + k]]) return false;
            if (mem[x.b + k] - mem[x.b] != mem[y.b + k] - mem[y.b]) return false;
        }
        return true;
    };
    // items of G * S reads with the same haplotypes: (first run in
    // `chained`, rows per lane)
    const size_t GS = G * (size_t)S;
    std::vector<std::pair<size_t, int>> groups;
    std::vector<Run> chained;
    uint64_t chain_pairs = 0;
    for (size_t i = 0; i < runs.size();) {
        size_t j = i + 1;
        while (j < runs.size() && runs[j].sig == runs[i].sig) j++;
        std::vector<Run> grp;
        for (size_t k = i; k < j; k++) {
            if (same(runs[k], runs[i]))
                grp.push_back(runs[k]);
            else
                single.insert(single.end(), mem.begin() + runs[k].b, mem.begin() + runs[k].e);
        }
        if (S == 2)  // similar lengths stacked together
            std::stable_sort(grp.begin(), grp.end(), [&](const Run &x, const Run &y) {
                return read_len[pr[mem[x.b]]] > read_len[pr[mem[y.b]]];
            });
        const size_t full = grp.size() / GS * GS;
        size_t k = 0;
        for (; k < full; k += GS) {
            int mr = bk.MR;
            if (S == 2) {
                // the item's rows per lane, and whether stacking pays
                double e1 = 0, e2 = 0;
                mr = 0;
                for (size_t q = 0; q < G; q++) {
                    const uint32_t RA = read_len[pr[mem[grp[k + 2 * q].b]]];
                    const uint32_t RB = read_len[pr[mem[grp[k + 2 * q + 1].b]]];
                    const int m = stack_rows(RA, RB);
                    if (m == 0) {
                        mr = 0;
                        break;
                    }
                    mr = std::max(mr, std::max(m, 2));  // (no one-row instance)
                }
                if (mr) {
                    for (size_t q = 0; q < G; q++) {
                        const uint32_t RA = read_len[pr[mem[grp[k + 2 * q].b]]];
                        const uint32_t RB = read_len[pr[mem[grp[k + 2 * q + 1].b]]];
                        e1 += group_eff(RA, (int)(RA + 16) / 16) + group_eff(RB, (int)(RB + 16) / 16);
                        e2 += 2.0 * group_eff(RA + RB, mr);
                    }
                    if (e2 < 1.03 * e1) mr = 0;
                }
                if (!mr) break;  // longest first: the rest of the group would not stack either
            }
            groups.push_back({chained.size(), mr});
            chained.insert(chained.end(), grp.begin() + k, grp.begin() + k + GS);
            chain_pairs += (uint64_t)GS * (grp[0].e - grp[0].b);
        }
        for (; k < grp.size(); k++)
            single.insert(single.end(), mem.begin() + grp[k].b, mem.begin() + grp[k].e);
        i = j;
    }
    // (an item of S reads per group is S items' work: min_items / S of them)
    const uint64_t kseg = min_items ? std::min<uint64_t>(kChainMaxHaps, chain_pairs / ((uint64_t)G * min_items))
                                    : kChainMaxHaps;
    if (kseg < 2) {
        for (const Run &r : chained) single.insert(single.end(), mem.begin() + r.b, mem.begin() + r.e);
        mem.swap(single);
        return;
    }
    // segments: (columns, first run, first and last haplotype), K haplotypes
    // cut into ceil(K / kseg) near-equal parts
    struct Seg {
        uint64_t cols;
        size_t run;
        int mr;
        uint32_t k0, k1;
    };
    std::vector<Seg> segs;
    for (const auto &gi : groups) {
        const Run &r0 = chained[gi.first];
        const uint32_t K = r0.e - r0.b;
        const uint32_t ns = (uint32_t)((K + kseg - 1) / kseg);
        for (uint32_t q = 0; q < ns; q++) {
            Seg sg{0, gi.first, gi.second, (uint32_t)((uint64_t)K * q / ns), (uint32_t)((uint64_t)K * (q + 1) / ns)};
            for (uint32_t k = sg.k0; k < sg.k1; k++) sg.cols += hap_len[ph[mem[r0.b + k]]];
            segs.push_back(sg);
        }
    }
    std::stable_sort(segs.begin(), segs.end(), [](const Seg &x, const Seg &y) { return x.cols > y.cols; });
    for (const Seg &sg : segs) {
        const Run *r = &chained[sg.run];
        Bucket *tb = &bk;
        if (S == 2) {
            tb = nullptr;
            for (Bucket &c : *stacked)
                if (c.MR == sg.mr && c.inter == bk.inter && c.share == bk.share) tb = &c;
            if (!tb) {
                Bucket c;
                c.L = bk.L;
                c.MR = sg.mr;
                c.wpc = bk.wpc;
                c.inter = bk.inter;
                c.share = bk.share;
                c.tab = true;
                c.chain = true;
                c.stack = 2;
                stacked->push_back(std::move(c));
                tb = &stacked->back();
            }
        }
        tb->chain_off.push_back((uint32_t)tb->chain_hap.size());
        for (uint32_t k = sg.k0; k < sg.k1; k++) {
            const uint32_t h = ph[mem[r[0].b + k]];
            ChainHap e{};
            e.H = hap_len[h];
            e.initd = initd_f[h];
            e.hoff = hapc_off[h];
            e.pdelta = (int32_t)(mem[r[0].b + k] - mem[r[0].b + sg.k0]);
            tb->chain_hap.push_back(e);
            tb->max_h = std::max(tb->max_h, e.H);
        }
        for (size_t g = 0; g < GS; g++) {
            tb->items.push_back((int32_t)mem[r[g].b + sg.k0]);
            for (uint32_t k = r[g].b + sg.k0; k < r[g].b + sg.k1; k++)
                tb->cells += (uint64_t)read_len[pr[mem[k]]] * hap_len[ph[mem[k]]];
        }
    }
    // (the stacked buckets take their closing offset from the caller, after
    // every table bucket has added its items)
    if (S == 1 && !bk.chain_off.empty()) bk.chain_off.push_back((uint32_t)bk.chain_hap.size());
    mem.swap(single);
}

// Strips of (L-1)*MR read rows (select-prior kernel), after `off` < MR
// leading identity rows.
int strips_for(uint32_t R, int L, int MR) {
    const int C = (L - 1) * MR;
    const int off = (MR - (int)R % MR) % MR;
    return ((int)R - 1 + off) / C + 1;
}

// A pair's units for the long-read kernels: runs of at most W strips (one
// per wave of a W-wave workgroup) from strip s_begin, each but the last
// handing its boundary row to the next through a new slot; the first reads
// the row above strip s_begin from slot xin (kNoSlot: none).  False if the
// strip count does not fit the units' 16-bit fields.
bool append_units(std::vector<LongUnit> &units, uint32_t &nslots, int32_t pair, int nstr, int W,
                  int s_begin = 0, uint32_t xin = kNoSlot) {
    if (nstr < 1 || nstr > 0xffff) return false;
    for (int s0 = s_begin; s0 < nstr; s0 += W) {
        const int s1 = std::min(nstr, s0 + W);
        LongUnit u;
        u.pair = pair;
        u.strips = (uint32_t)s0 | (uint32_t)s1 << 16;
        u.xin = xin;
        u.xout = s1 < nstr ? nslots++ : kNoSlot;
        xin = u.xout;
        units.push_back(u);
    }
    return true;
}

// Group widths schedule_pairhmm would use on this host (pairhmm_schedule.cpp:
// 176-186 with CpuFeatures, cpu_features.h:53-154): 16/8 with AVX-512
// F+DQ+BW+VL and ZMM state enabled by the OS, 8/4 with AVX2 and YMM state,
// else 0/0.
void host_simd_widths(uint32_t *wf, uint32_t *wd) {
    *wf = *wd = 0;
    unsigned a, b, c, d;
    if (!__get_cpuid_count(1, 0, &a, &b, &c, &d) || !(c & (1u << 27))) return;  // OSXSAVE
    uint32_t xcr0, xcr0_hi;
    __asm__("xgetbv" : "=a"(xcr0), "=d"(xcr0_hi) : "c"(0));
    (void)xcr0_hi;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return;
    const uint32_t skx = (1u << 16) | (1u << 17) | (1u << 30) | (1u << 31);
    const uint32_t zmm = (7u << 5) | (1u << 2) | (1u << 1);
    if ((xcr0 & zmm) == zmm && (b & skx) == skx) {
        *wf = 16;
        *wd = 8;
    } else if ((xcr0 & 6u) == 6u && (b & (1u << 5))) {
        *wf = 8;
        *wd = 4;
    }
}

FwdArgs base_args(rodv_phmm_plan *p);

// ---------------------------------------------------------------- routing
// schedule_pairhmm's FP32 stage (pairhmm_schedule.cpp:141-196): the region's
// pairs in hap-major order, sorted by (hap length desc, read length desc)
// with std::sort -- the algorithm the reference calls, so ties keep its
// order -- then greedy windows of `wf` unused N-free pairs whose idle ratio
// is <= idle_f go to the inter kernel (greedy_grouping, :46-105).
struct RoutePair {
    uint32_t local;  // h * num_reads + r inside the region
    uint32_t hap_len, read_len;
    bool contains_n, used;
};

void route_region(const rodv_phmm_batch *b, const rodv_phmm_region &rg,
                  const std::vector<uint8_t> &hap_n, const std::vector<uint8_t> &read_n,
                  const rodv_phmm_config &cfg, uint8_t *inter32, std::vector<uint32_t> &sorted_inter) {
    sorted_inter.clear();
    const uint32_t m = rg.num_haps, n = rg.num_reads, W = cfg.simd_width_float;
    if (cfg.use_double || W == 0) return;
    std::vector<RoutePair> pairs;
    pairs.reserve((size_t)m * n);
    for (uint32_t h = 0; h < m; h++)
        for (uint32_t r = 0; r < n; r++)
            pairs.push_back(RoutePair{h * n + r, b->hap_lengths[rg.hap_begin + h],
                                      b->read_lengths[rg.read_begin + r],
                                      hap_n[rg.hap_begin + h] || read_n[rg.read_begin + r], false});
    std::sort(pairs.begin(), pairs.end(), [](const RoutePair &x, const RoutePair &y) {
        if (x.hap_len == y.hap_len) return x.read_len > y.read_len;
        return x.hap_len > y.hap_len;
    });
    const size_t total = pairs.size();
    std::vector<size_t> win;
    win.reserve(W);
    for (size_t i = 0; i < total; i++) {
        if (pairs[i].used || pairs[i].contains_n) continue;
        win.clear();
        for (size_t idx = i; idx < total && win.size() < W; idx++)
            if (!pairs[idx].used && !pairs[idx].contains_n) win.push_back(idx);
        if (win.size() < W) break;
        uint32_t mh = 0, mr = 0;
        uint64_t sum = 0;
        for (size_t k : win) {
            mh = std::max(mh, pairs[k].hap_len);
            mr = std::max(mr, pairs[k].read_len);
            sum += (uint64_t)pairs[k].hap_len * pairs[k].read_len;
        }
        const uint64_t cap = (uint64_t)mh * mr * W;
        const uint64_t idle = cap > sum ? cap - sum : 0;
        const double ratio = cap == 0 ? 0.0 : (double)idle / (double)cap;
        if (ratio <= cfg.max_idle_ratio_float)
            for (size_t k : win) pairs[k].used = true;
    }
    // the FP64 stage scans the same sorted order, over the pairs the FP32
    // stage grouped (pairhmm_schedule.cpp:258-263)
    for (const RoutePair &pp : pairs)
        if (pp.used) {
            inter32[pp.local] = 1;
            sorted_inter.push_back(pp.local);
        }
}

int validate(const rodv_phmm_batch *b) {
    if (!b) return RODV_PHMM_ERR_INVALID_ARG;
    if (b->num_haps && (!b->hap_bases || !b->hap_offsets || !b->hap_lengths))
        return RODV_PHMM_ERR_INVALID_ARG;
    if (b->num_reads && (!b->read_bases || !b->base_quals || !b->ins_quals || !b->del_quals ||
                         !b->gcp_quals || !b->read_offsets || !b->read_lengths))
        return RODV_PHMM_ERR_INVALID_ARG;
    if (b->num_regions && !b->regions) return RODV_PHMM_ERR_INVALID_ARG;
    for (uint32_t k = 0; k < b->num_regions; k++) {
        const rodv_phmm_region &r = b->regions[k];
        if ((uint64_t)r.hap_begin + r.num_haps > b->num_haps ||
            (uint64_t)r.read_begin + r.num_reads > b->num_reads)
            return RODV_PHMM_ERR_INVALID_ARG;
    }
    return RODV_PHMM_OK;
}

// No exception leaves the C ABI: a failed host allocation becomes
// RODV_PHMM_ERR_OOM, anything else RODV_PHMM_ERR_DEVICE.
template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return RODV_PHMM_ERR_OOM;
    } catch (...) {
        return RODV_PHMM_ERR_DEVICE;
    }
}

}  // namespace

// ================================================================== C ABI

extern "C" {

static int init_impl(void) {
    (void)tables();
    return RODV_PHMM_OK;
}

int rodv_phmm_init(void) {
    return guarded([&] { return init_impl(); });
}

const char *rodv_phmm_strerror(int status) {
    switch (status) {
    case RODV_PHMM_OK: return "ok";
    case RODV_PHMM_ERR_INVALID_ARG: return "invalid argument";
    case RODV_PHMM_ERR_OOM: return "out of memory";
    case RODV_PHMM_ERR_DEVICE: return "HIP device error";
    case RODV_PHMM_ERR_UNSUPPORTED: return "unsupported request";
    case RODV_PHMM_ERR_EMPTY: return "empty batch (no haplotypes or no reads)";
    case RODV_PHMM_ERR_TOO_LARGE:
        return "batch needs more device memory than rodv_phmm_forward_noalloc's fixed scratch "
               "(RODV_PHMM_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown status";
    }
}

void rodv_phmm_config_default(rodv_phmm_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->struct_size = sizeof(*cfg);
    cfg->semantics = RODV_PHMM_SEMANTICS_SCHEDULE;
    cfg->use_double = 0;
    cfg->max_idle_ratio_float = 0.1;
    cfg->max_idle_ratio_double = 0.1;
    host_simd_widths(&cfg->simd_width_float, &cfg->simd_width_double);
    cfg->use_graph = 1;
    cfg->waves_per_cu = 0;
    cfg->short_lanes = 16;
    cfg->extended_range = 0;
    cfg->prior_table = 1;
    cfg->use_tuning = 1;
    cfg->haplotype_chains = 1;
    cfg->shared_products = 1;
}

static int ctx_create_impl(int device, rodv_phmm_ctx **out) {
    if (!out) return RODV_PHMM_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return RODV_PHMM_ERR_DEVICE;
    }
    HIPCHK(hipSetDevice(device));
    rodv_phmm_ctx *c = new (std::nothrow) rodv_phmm_ctx();
    if (!c) return RODV_PHMM_ERR_OOM;
    c->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        c->num_cus = cus;
    const HostTables &t = tables();
    int rc = upload(&c->d_tab, &t.dev, 1);
    if (!rc && hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    for (int k = 0; k < kSideStreams && !rc; k++)
        if (hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->join_ev[k], hipEventDisableTiming) != hipSuccess)
            rc = RODV_PHMM_ERR_DEVICE;
    if (rc) (void)hipGetLastError();
    if (rc) {
        rodv_phmm_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return RODV_PHMM_OK;
}

int rodv_phmm_ctx_create(int device, rodv_phmm_ctx **out) {
    return guarded([&] { return ctx_create_impl(device, out); });
}

void rodv_phmm_ctx_destroy(rodv_phmm_ctx *ctx) {
    if (!ctx) return;
    for (int k = 0; k < kSideStreams; k++) {
        if (ctx->side[k]) (void)hipStreamDestroy(ctx->side[k]);
        if (ctx->join_ev[k]) (void)hipEventDestroy(ctx->join_ev[k]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->d_tab) (void)hipFree(ctx->d_tab);
    if (ctx->fwd_arena.dev) (void)hipFree(ctx->fwd_arena.dev);
    if (ctx->fwd_arena.host) (void)hipHostFree(ctx->fwd_arena.host);
    delete ctx;
}

static int tables_f32_impl(float *ph2pr, float *mm, int mm_n) {
    const HostTables &t = tables();
    if (mm_n < 0 || mm_n > kMMSize) return RODV_PHMM_ERR_INVALID_ARG;
    if (ph2pr) std::memcpy(ph2pr, t.ph2pr_f, sizeof(t.ph2pr_f));
    if (mm) std::memcpy(mm, t.mm_f.data(), sizeof(float) * (size_t)mm_n);
    return RODV_PHMM_OK;
}

int rodv_phmm_tables_f32(float *ph2pr, float *mm, int mm_n) {
    return guarded([&] { return tables_f32_impl(ph2pr, mm, mm_n); });
}

static int tables_f64_impl(double *ph2pr, double *mm, int mm_n) {
    const HostTables &t = tables();
    if (mm_n < 0 || mm_n > kMMSize) return RODV_PHMM_ERR_INVALID_ARG;
    if (ph2pr) std::memcpy(ph2pr, t.ph2pr_d, sizeof(t.ph2pr_d));
    if (mm) std::memcpy(mm, t.mm_d.data(), sizeof(double) * (size_t)mm_n);
    return RODV_PHMM_OK;
}

int rodv_phmm_tables_f64(double *ph2pr, double *mm, int mm_n) {
    return guarded([&] { return tables_f64_impl(ph2pr, mm, mm_n); });
}

static int plan_create_impl(rodv_phmm_ctx *ctx, const rodv_phmm_batch *b,
                          const rodv_phmm_config *cfg_in, rodv_phmm_plan **out,
                          Arena *arena = nullptr, bool no_plain = false) {
    if (!ctx || !out) return RODV_PHMM_ERR_INVALID_ARG;
    *out = nullptr;
    int rc = validate(b);
    if (rc) return rc;
    rodv_phmm_config cfg;
    rodv_phmm_config_default(&cfg);
    if (cfg_in) {
        if (cfg_in->struct_size != sizeof(rodv_phmm_config)) return RODV_PHMM_ERR_INVALID_ARG;
        cfg = *cfg_in;
    }
    if (cfg.semantics != RODV_PHMM_SEMANTICS_INTRA && cfg.semantics != RODV_PHMM_SEMANTICS_SCHEDULE)
        return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.short_lanes != 16 && cfg.short_lanes != 64) return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.waves_per_cu < 0 || cfg.waves_per_cu > 32) return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.simd_width_float > 64 || cfg.simd_width_double > 64) return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.haplotype_chains < 0 || cfg.haplotype_chains > 4) return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.extended_range != 0 && cfg.extended_range != 1) return RODV_PHMM_ERR_INVALID_ARG;
    if (cfg.shared_products != 0 && cfg.shared_products != 1) return RODV_PHMM_ERR_INVALID_ARG;
    if (!(cfg.max_idle_ratio_float == cfg.max_idle_ratio_float) ||
        !(cfg.max_idle_ratio_double == cfg.max_idle_ratio_double))
        return RODV_PHMM_ERR_INVALID_ARG;
    (void)tables();
    HIPCHK(hipSetDevice(ctx->device));

    // regions (default: one over everything)
    std::vector<rodv_phmm_region> regions;
    if (b->num_regions == 0)
        regions.push_back(rodv_phmm_region{0, b->num_haps, 0, b->num_reads});
    else
        regions.assign(b->regions, b->regions + b->num_regions);
    uint64_t n_pairs = 0;
    for (auto &r : regions) {
        if (r.num_haps == 0 || r.num_reads == 0) return RODV_PHMM_ERR_EMPTY;
        n_pairs += (uint64_t)r.num_haps * r.num_reads;
    }
    if (n_pairs == 0) return RODV_PHMM_ERR_EMPTY;
    if (n_pairs >= (1ull << 31)) return RODV_PHMM_ERR_UNSUPPORTED;

    // ---- encode and validate sequences
    const uint32_t M = b->num_haps, N = b->num_reads;
    std::vector<uint64_t> hapc_off(M);
    uint64_t hap_bytes = 0;
    for (uint32_t h = 0; h < M; h++) {
        if (b->hap_lengths[h] == 0) return RODV_PHMM_ERR_INVALID_ARG;
        hapc_off[h] = hap_bytes;
        hap_bytes += ((uint64_t)b->hap_lengths[h] + 15) & ~15ull;
    }
    std::vector<uint8_t> hapc(hap_bytes, 0);
    std::vector<uint8_t> hap_has_n(M, 0);
    std::vector<float> initd_f(M);
    std::vector<double> initd_d(M);
    bool ok = true;
    for (uint32_t h = 0; h < M; h++) {
        const uint8_t *s = b->hap_bases + b->hap_offsets[h];
        uint8_t *d = hapc.data() + hapc_off[h];
        for (uint32_t j = 0; j < b->hap_lengths[h]; j++) {
            d[j] = hap_code(s[j], &ok);
            hap_has_n[h] |= d[j] == kHapN;
        }
        initd_f[h] = 0x1p120f / (float)b->hap_lengths[h];   // pairhmm_impl.cpp:62
        initd_d[h] = 0x1p1020 / (double)b->hap_lengths[h];
    }
    if (!ok) return RODV_PHMM_ERR_INVALID_ARG;
    std::vector<uint64_t> read_off(N);
    uint64_t read_bytes = 0;
    for (uint32_t r = 0; r < N; r++) {
        if (b->read_lengths[r] == 0) return RODV_PHMM_ERR_INVALID_ARG;
        read_off[r] = read_bytes;
        read_bytes += b->read_lengths[r];
    }
    std::vector<uint8_t> rcode(read_bytes), q(read_bytes), qi(read_bytes), qd(read_bytes),
        qc(read_bytes);
    for (uint32_t r = 0; r < N; r++) {
        const uint64_t so = b->read_offsets[r], dst = read_off[r];
        const uint32_t n = b->read_lengths[r];
        for (uint32_t k = 0; k < n; k++) rcode[dst + k] = read_code(b->read_bases[so + k], &ok);
        std::memcpy(&q[dst], b->base_quals + so, n);
        std::memcpy(&qi[dst], b->ins_quals + so, n);
        std::memcpy(&qd[dst], b->del_quals + so, n);
        std::memcpy(&qc[dst], b->gcp_quals + so, n);
    }
    if (!ok) return RODV_PHMM_ERR_INVALID_ARG;
    std::vector<uint8_t> read_has_n(N, 0);
    for (uint32_t r = 0; r < N; r++)
        for (uint32_t k = 0; k < b->read_lengths[r]; k++)
            read_has_n[r] |= rcode[read_off[r] + k] == kReadN;
    // reads whose deletion prior of each row is the insertion prior of the
    // row below (ph2pr[del_q[k]] == ph2pr[ins_q[k + 1]], the masked
    // qualities of pairhmm_impl.cpp:89-92): the FP32 table kernels then
    // share one product between the D and I recurrences (same bits)
    std::vector<uint8_t> read_share(N, 0);
    if (cfg.shared_products) {
        const float *ph2pr = tables().ph2pr_f;
        for (uint32_t r = 0; r < N; r++) {
            bool ok_r = true;
            const uint64_t o = read_off[r];
            for (uint32_t k = 0; k + 1 < b->read_lengths[r] && ok_r; k++) {
                const float my = ph2pr[qd[o + k] & 127u], mx = ph2pr[qi[o + k + 1] & 127u];
                uint32_t a32, b32;
                std::memcpy(&a32, &my, 4);
                std::memcpy(&b32, &mx, 4);
                ok_r = a32 == b32;
            }
            read_share[r] = ok_r ? 1 : 0;
        }
    }
    const bool share_all = cfg.shared_products && std::all_of(read_share.begin(), read_share.end(),
                                                              [](uint8_t v) { return v != 0; });

    rodv_phmm_plan *p = new (std::nothrow) rodv_phmm_plan();
    if (!p) return RODV_PHMM_ERR_OOM;
    // frees the plan on every way out but success (which clears p), an
    // exception included
    struct Hold {
        rodv_phmm_plan *&p;
        ~Hold() {
            if (p) free_plan(p);
        }
    } hold{p};
    p->ctx = ctx;
    p->cfg = cfg;
    p->n_pairs = n_pairs;
    p->share_all = share_all;
    p->rescue64_waves = rescue_long_waves_hook() == fwd_long_waves() ? fwd_long_waves() : rescue_long_waves();
    {
        // the long-read kernels' LDS prior tables: every haplotype N-free,
        // and the codes of the longest haplotype and the tables fit
        uint32_t mh = 0;
        for (uint32_t h = 0; h < M; h++) mh = std::max(mh, b->hap_lengths[h]);
        int maxlds = 65536;
        (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
        const uint32_t tabs = std::max(long_tab_bytes(16, false, fwd_long_waves()),
                                       long_tab_bytes(rescue_long_mr(), true, p->rescue64_waves));
        p->long_tab = std::none_of(hap_has_n.begin(), hap_has_n.end(), [](uint8_t v) { return v != 0; }) &&
                      (uint64_t)((lds_stride_for(mh, 64) + 255u) & ~255u) + tabs <= (uint64_t)maxlds;
    }
    p->arena = arena;
    p->no_plain = no_plain;
    if (arena) arena->used = arena->need = 0;

    // ---- pairs in output order, and their buckets
    std::vector<uint32_t> ph(n_pairs), pr(n_pairs);
    std::vector<uint32_t> pair_region(n_pairs);
    {
        uint64_t k = 0;
        for (uint32_t ri = 0; ri < regions.size(); ri++) {
            const auto &rg = regions[ri];
            for (uint32_t a = 0; a < rg.num_haps; a++)
                for (uint32_t c = 0; c < rg.num_reads; c++) {
                    ph[k] = rg.hap_begin + a;
                    pr[k] = rg.read_begin + c;
                    pair_region[k] = ri;
                    k++;
                }
        }
    }
    // ---- schedule_pairhmm routing: which pairs its FP32 stage sends to the
    // inter kernel; regions are independent calls, routed in parallel
    std::vector<uint8_t> inter32;
    std::vector<uint32_t> sched_list;
    std::vector<uint64_t> sched_off;
    if (cfg.semantics == RODV_PHMM_SEMANTICS_SCHEDULE) {
        inter32.assign(n_pairs, 0);
        const size_t nr = regions.size();
        std::vector<uint64_t> base(nr + 1, 0);
        for (size_t ri = 0; ri < nr; ri++)
            base[ri + 1] = base[ri] + (uint64_t)regions[ri].num_haps * regions[ri].num_reads;
        std::vector<std::vector<uint32_t>> lists(nr);
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t ri = next++; ri < nr; ri = next++)
                route_region(b, regions[ri], hap_has_n, read_has_n, cfg, inter32.data() + base[ri],
                             lists[ri]);
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nt && t < nr; t++) pool.emplace_back(work);
        work();
        for (auto &t : pool) t.join();
        sched_off.assign(nr + 1, 0);
        for (size_t ri = 0; ri < nr; ri++) {
            sched_off[ri + 1] = sched_off[ri] + lists[ri].size();
            for (uint32_t local : lists[ri]) sched_list.push_back((uint32_t)(base[ri] + local));
        }
        p->any_inter = !sched_list.empty();
        p->n_regions = (uint32_t)nr;
    }
    std::vector<int> bucket_of_key(2 * 4 * 4 * 17, -1);
    std::vector<std::vector<uint32_t>> members;
    uint64_t n_shared = 0;    // pairs in shared-product buckets
    uint32_t max_long_h = 0;  // hap length of pairs whose FP64 rescue may need strips
    uint64_t n_long = 0;      // such pairs
    for (uint64_t k = 0; k < n_pairs; k++) {
        const uint32_t R = b->read_lengths[pr[k]], H = b->hap_lengths[ph[k]];
        p->cells += (uint64_t)R * H;
        if (R > (uint32_t)rescue_long_rows()) {
            max_long_h = std::max(max_long_h, H);
            n_long++;
        }
        const int cls = tune_class(R);
        const TuneEntry *te = nullptr;
        if (ctx->force_class == cls)
            te = &ctx->force;
        else if (cfg.use_tuning && ctx->tuned)
            te = &ctx->tune[cls];
        const bool tab_ok = cfg.prior_table && !hap_has_n[ph[k]];
        int L, MR;
        bool tab;
        shape_for(R, H, cfg, te, tab_ok, &L, &MR, &tab);
        const bool inter = !inter32.empty() && inter32[k];
        const int lidx = L == 16 ? 0 : (L == 32 ? 1 : (L == 64 ? 2 : 3));
        // the shared-product variants exist for the 16-lane table kernels
        const bool share = tab && L == 16 && read_share[pr[k]];
        n_shared += share ? 1 : 0;
        const int key = ((((share ? 1 : 0) * 2 + (inter ? 1 : 0)) * 2 + (tab ? 1 : 0)) * 4 + lidx) * 17 + MR;
        if (bucket_of_key[key] < 0) {
            bucket_of_key[key] = (int)p->buckets.size();
            Bucket bk;
            bk.L = L;
            bk.MR = MR;
            bk.tab = tab;
            bk.share = share;
            bk.inter = inter;
            bk.wpc = te ? te->wpc : 0;
            p->buckets.push_back(bk);
            members.emplace_back();
        }
        members[bucket_of_key[key]].push_back((uint32_t)k);
    }
    // the plain schedule keeps every pair of a bucket in one-pair slots
    std::vector<std::vector<uint32_t>> plain_mem = members;
    // ---- chained items: a bucket of their own next to each table bucket
    if (cfg.haplotype_chains) {
        const size_t nb0 = p->buckets.size();
        // two reads per group where that packs a short read's rows better
        // (16-lane buckets of reads up to 127 rows), then one read per group
        std::vector<Bucket> stacked;
        // modes 1 and 3: chains only where the bucket keeps the device full;
        // 2 and 4: always; 3 and 4 keep one read per group
        const bool always = cfg.haplotype_chains == 2 || cfg.haplotype_chains == 4;
        const bool stack = cfg.haplotype_chains <= 2;
        for (size_t bi = 0; bi < nb0 && stack; bi++) {
            const Bucket &b0 = p->buckets[bi];
            if (!b0.tab || b0.L != 16 || b0.MR > 8) continue;
            Bucket tmpl;
            tmpl.L = b0.L;
            tmpl.MR = b0.MR;
            tmpl.wpc = b0.wpc;
            tmpl.inter = b0.inter;
            tmpl.share = b0.share;
            const int wpc = cfg.waves_per_cu > 0 ? cfg.waves_per_cu
                                                   : (tmpl.wpc > 0 ? tmpl.wpc : kDefaultWavesPerCU);
            form_chains(members[bi], tmpl, ph.data(), pr.data(), pair_region, b->hap_lengths,
                        b->read_lengths, hapc_off, initd_f,
                        always ? 0 : kChainItemsPerWave * (uint64_t)ctx->num_cus * (uint64_t)wpc,
                        2, &stacked);
        }
        for (Bucket &c : stacked) {
            c.chain_off.push_back((uint32_t)c.chain_hap.size());
            c.kern = fwd_chain_kernel_ptr(c.L, c.MR, c.inter, 2, c.share);
            if (!c.kern) return RODV_PHMM_ERR_UNSUPPORTED;
            p->buckets.push_back(std::move(c));
            members.emplace_back();
            plain_mem.emplace_back();
        }
        for (size_t bi = 0; bi < nb0; bi++) {
            if (!p->buckets[bi].tab) continue;
            void *ck = fwd_chain_kernel_ptr(p->buckets[bi].L, p->buckets[bi].MR, p->buckets[bi].inter, 1,
                                            p->buckets[bi].share);
            if (!ck) continue;
            Bucket cb;
            cb.L = p->buckets[bi].L;
            cb.MR = p->buckets[bi].MR;
            cb.wpc = p->buckets[bi].wpc;
            cb.inter = p->buckets[bi].inter;
            cb.share = p->buckets[bi].share;
            cb.tab = true;
            cb.chain = true;
            cb.kern = ck;
            // at least kChainItemsPerWave items per wave of a full grid
            const int wpc = cfg.waves_per_cu > 0 ? cfg.waves_per_cu
                                                   : (cb.wpc > 0 ? cb.wpc : kDefaultWavesPerCU);
            form_chains(members[bi], cb, ph.data(), pr.data(), pair_region, b->hap_lengths,
                        b->read_lengths, hapc_off, initd_f,
                        always ? 0 : kChainItemsPerWave * (uint64_t)ctx->num_cus * (uint64_t)wpc);
            if (cb.chain_off.empty()) continue;
            p->buckets.push_back(std::move(cb));
            members.emplace_back();
            plain_mem.emplace_back();
        }
    }

    // ---- items: length-sorted packing of G pairs per wave
    for (size_t bi = 0; bi < p->buckets.size(); bi++) {
        Bucket &bk = p->buckets[bi];
        const int wpc = cfg.waves_per_cu > 0 ? cfg.waves_per_cu
                                               : (bk.wpc > 0 ? bk.wpc : kDefaultWavesPerCU);
        bk.grid_cap = (int)std::max<uint64_t>(1, (uint64_t)ctx->num_cus * wpc);
        if (bk.chain) {
            bk.n_items = (uint32_t)(bk.chain_off.size() - 1);
            bk.lds_stride = (lds_stride_for(bk.max_h, bk.L) + 255u) & ~255u;
            bk.tab_bytes = fwd_tab_bytes(bk.L, bk.MR);
            bk.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(bk.n_items, (uint64_t)ctx->num_cus * wpc));
            continue;
        }
        bk.kern = bk.tab ? fwd_tab_kernel_ptr(bk.L, bk.MR, bk.inter, bk.share)
                         : fwd_f32_kernel(bk.L, bk.MR, bk.inter);
        if (!bk.kern) {
            return RODV_PHMM_ERR_UNSUPPORTED;
        }
        // LDS for every pair the bucket's kernel may take, plain launches included
        for (uint32_t k : plain_mem[bi]) bk.max_h = std::max(bk.max_h, b->hap_lengths[ph[k]]);
        auto &mem = members[bi];
        std::stable_sort(mem.begin(), mem.end(), [&](uint32_t x, uint32_t y) {
            uint32_t hx = b->hap_lengths[ph[x]], hy = b->hap_lengths[ph[y]];
            if (hx != hy) return hx > hy;
            return b->read_lengths[pr[x]] > b->read_lengths[pr[y]];
        });
        const int G = 64 / bk.L;
        bk.n_items = (uint32_t)((mem.size() + G - 1) / G);
        bk.items.assign((size_t)bk.n_items * G, -1);
        for (size_t k = 0; k < mem.size(); k++) {
            bk.items[k] = (int32_t)mem[k];
            const uint32_t R = b->read_lengths[pr[mem[k]]], H = b->hap_lengths[ph[mem[k]]];
            bk.max_h = std::max(bk.max_h, H);
            if (!bk.tab) bk.max_strips = std::max(bk.max_strips, (uint32_t)strips_for(R, bk.L, bk.MR));
            bk.cells += (uint64_t)R * H;
        }
        if (bk.tab) {
            // one code byte per column, then the prior table.  A multiple of
            // 256 bytes: every group's table then starts on LDS bank 0, which
            // keeps the groups' 16-byte prior reads free of bank conflicts
            bk.lds_stride = (lds_stride_for(bk.max_h, bk.L) + 255u) & ~255u;
            bk.tab_bytes = fwd_tab_bytes(bk.L, bk.MR);
        } else {
            bk.lds_stride = lds_stride_for(bk.max_h, bk.L);
        }
        bk.grid = (int)std::min<uint64_t>(bk.n_items, (uint64_t)ctx->num_cus * wpc);
        if (bk.grid < 1) bk.grid = 1;
        if (bk.max_strips > 1) {
            bk.bnd_elems = 3ull * (bk.max_h + 3ull * bk.L + 2);
            // reads longer than one strip: their strips pipelined over the
            // waves of a workgroup, one workgroup per CU
            void *lk = bk.L == 64 ? fwd_long_kernel_ptr(bk.MR, bk.inter, p->long_tab) : nullptr;
            if (lk) {
                bk.kern = lk;
                // the waves' prior tables (covers the sub-strip launch's too)
                bk.tab_bytes = p->long_tab ? long_tab_bytes(bk.MR, false, fwd_long_waves()) : 0u;
                bk.block = 64 * fwd_long_waves();
                bk.nbnd = (uint32_t)fwd_long_waves() + 1;
                bk.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(bk.n_items, (uint64_t)ctx->num_cus));
                bk.grid_cap = bk.grid;
                // batched: strip 0 of every pair first (units1), then the
                // other strips split into units, so that a long pair's units
                // run on several CUs at once.  units1 covers strip 0's rows
                // as MR / kSubMR sub-strips of kSubMR rows per lane in the
                // same layout, pipelined over the workgroup's waves (a
                // shorter step: the lone pipeline of the longest pair's
                // first strip sets this launch's time)
                constexpr int kSubMR = 4;
                const int sub = bk.MR % kSubMR == 0 ? bk.MR / kSubMR : 1;
                bk.kern1 = sub > 1 ? fwd_long_kernel_ptr(kSubMR, bk.inter, p->long_tab) : lk;
                for (int32_t q : bk.items) {
                    if (q < 0) continue;
                    const uint32_t R = b->read_lengths[pr[q]];
                    const int nstr = strips_for(R, 64, bk.MR);
                    if (nstr < 1 || nstr > 0xffff) return RODV_PHMM_ERR_INVALID_ARG;
                    // sub-strips of strip 0: all of them, or those up to row R
                    const int off = (bk.MR - (int)(R % (uint32_t)bk.MR)) % bk.MR;
                    const int n1 = sub > 1 ? std::min(sub, ((int)R - 1 + off) / (63 * kSubMR) + 1) : 1;
                    LongUnit u1;
                    u1.pair = q;
                    u1.strips = 0u | (uint32_t)n1 << 16;
                    u1.xin = kNoSlot;
                    u1.xout = nstr > 1 ? bk.n_xslots++ : kNoSlot;
                    bk.units1.push_back(u1);
                    if (nstr > 1 && !append_units(bk.units, bk.n_xslots, q, nstr, fwd_long_waves(), 1, u1.xout))
                        return RODV_PHMM_ERR_INVALID_ARG;
                }
                bk.grid = (int)std::max<uint64_t>(
                    1, std::min<uint64_t>(std::max(bk.units.size(), bk.units1.size()), (uint64_t)ctx->num_cus));
            }
        }
    }
    // buckets by descending work, so long launches start first
    std::vector<size_t> order(p->buckets.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return p->buckets[x].cells * p->buckets[x].MR > p->buckets[y].cells * p->buckets[y].MR;
    });
    {
        std::vector<Bucket> sorted;
        std::vector<std::vector<uint32_t>> smem, spm;
        for (size_t k : order) {
            sorted.push_back(std::move(p->buckets[k]));
            smem.push_back(std::move(members[k]));
            spm.push_back(std::move(plain_mem[k]));
        }
        p->buckets = std::move(sorted);
        members = std::move(smem);
        plain_mem = std::move(spm);
    }

    // ---- plain schedule: per region, per bucket, region-ordered items
    if (!no_plain) {
        for (size_t bi = 0; bi < p->buckets.size(); bi++) {
            Bucket &bk = p->buckets[bi];
            const int G = 64 / bk.L;
            auto mem = plain_mem[bi];
            std::stable_sort(mem.begin(), mem.end(), [&](uint32_t x, uint32_t y) {
                if (pair_region[x] != pair_region[y]) return pair_region[x] < pair_region[y];
                uint32_t hx = b->hap_lengths[ph[x]], hy = b->hap_lengths[ph[y]];
                return hx > hy;
            });
            std::vector<int32_t> items;
            size_t k = 0;
            while (k < mem.size()) {
                const uint32_t rg = pair_region[mem[k]];
                const uint32_t begin = (uint32_t)(items.size() / G);
                while (k < mem.size() && pair_region[mem[k]] == rg) {
                    for (int gg = 0; gg < G; gg++) {
                        if (k < mem.size() && pair_region[mem[k]] == rg)
                            items.push_back((int32_t)mem[k++]);
                        else
                            items.push_back(-1);
                    }
                }
                const uint32_t n = (uint32_t)(items.size() / G) - begin;
                const int wpc = cfg.waves_per_cu > 0 ? cfg.waves_per_cu : kDefaultWavesPerCU;
                PlainLaunch pl;
                pl.bucket = (int)bi;
                pl.item_begin = begin;
                pl.n_items = n;
                pl.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(n, (uint64_t)ctx->num_cus * wpc));
                pl.bucket = (int)bi;
                p->plain.push_back(pl);
                (void)rg;
            }
            int32_t *d = nullptr;
            rc = plan_upload(p, &d, items.data(), items.size());
            p->d_plain_items.push_back(d);
            p->plain_items_n.push_back(items.size());
            if (rc) {
                return rc;
            }
        }
        // regions in order: launches sorted by (region of first item)
    }

    // ---- device upload
    rc = plan_upload(p, &p->d_hapc, hapc.data(), hapc.size());
    if (!rc) rc = plan_upload(p, &p->d_hapc_off, hapc_off.data(), hapc_off.size());
    if (!rc) rc = plan_upload(p, &p->d_hap_len, b->hap_lengths, M);
    if (!rc) rc = plan_upload(p, &p->d_initd_f, initd_f.data(), M);
    if (!rc) rc = plan_upload(p, &p->d_initd_d, initd_d.data(), M);
    if (!rc) rc = plan_upload(p, &p->d_rcode, rcode.data(), rcode.size());
    if (!rc) rc = plan_upload(p, &p->d_q, q.data(), q.size());
    if (!rc) rc = plan_upload(p, &p->d_ins, qi.data(), qi.size());
    if (!rc) rc = plan_upload(p, &p->d_del, qd.data(), qd.size());
    if (!rc) rc = plan_upload(p, &p->d_gcp, qc.data(), qc.size());
    if (!rc) rc = plan_upload(p, &p->d_read_off, read_off.data(), N);
    if (!rc) rc = plan_upload(p, &p->d_read_len, b->read_lengths, N);
    if (!rc) rc = plan_upload(p, &p->d_pair_hap, ph.data(), n_pairs);
    if (!rc) rc = plan_upload(p, &p->d_pair_read, pr.data(), n_pairs);
    if (!rc && p->any_inter) {
        rc = plan_upload(p, &p->d_pair_inter, inter32.data(), n_pairs);
        if (!rc) rc = plan_upload(p, &p->d_sched_list, sched_list.data(), sched_list.size());
        if (!rc) rc = plan_upload(p, &p->d_sched_off, sched_off.data(), sched_off.size());
        if (!rc) rc = plan_alloc(p, &p->d_sched_scratch, sched_list.size());
    }
    if (!rc) rc = plan_alloc(p, &p->d_raw32, n_pairs);
    if (!rc) rc = plan_alloc(p, &p->d_raw64, n_pairs);
    if (!rc) rc = plan_alloc(p, &p->d_out, n_pairs);
    if (!rc) rc = plan_alloc(p, &p->d_flags, n_pairs);
    if (!rc) rc = plan_alloc(p, &p->d_rescue, n_pairs);
    {
        // FP64 rescue classes: a pair's capacity is reserved in its class
        uint64_t ncls[kRescueClasses + 1] = {};
        uint32_t max_h_tab = 0;
        for (uint64_t k = 0; k < n_pairs; k++) {
            const uint32_t H = b->hap_lengths[ph[k]];
            const int c = rescue_class(b->read_lengths[pr[k]], H, hap_has_n[ph[k]] != 0);
            ncls[c < 0 ? kRescueClasses : c]++;
            if (c >= 0) max_h_tab = std::max(max_h_tab, H);
        }
        uint64_t acc = 0;
        for (int c = 0; c <= kRescueClasses; c++) {
            p->bin_base[c] = (uint32_t)acc;
            p->bin_cap[c] = (uint32_t)ncls[c];
            acc += ncls[c];
        }
        uint64_t lo = 0, hi = 0;
        for (int c = 0; c < kRescueClasses; c++) (c < 4 ? lo : hi) += ncls[c];
        // two pairs per wave: 2 (classes 4..7, 255 VGPRs) or 3 (0..3) waves per SIMD
        p->rtab_grid[0] = (int)std::min<uint64_t>((hi + 1) / 2, (uint64_t)ctx->num_cus * 8);
        p->rtab_grid[1] = (int)std::min<uint64_t>((lo + 1) / 2, (uint64_t)ctx->num_cus * 12);
        p->rtab_stride = lds_stride_for(max_h_tab, 32);
        if (!rc && (lo || hi)) {
            rc = plan_alloc(p, &p->d_bin, n_pairs);
            if (!rc) rc = plan_upload(p, &p->d_hap_n, hap_has_n.data(), M);
        }
    }
    for (auto &bk : p->buckets) {
        if (!rc) rc = plan_upload(p, &bk.d_items, bk.items.data(), bk.items.size());
        if (!rc && bk.chain) {
            rc = plan_upload(p, &bk.d_chain_off, bk.chain_off.data(), bk.chain_off.size());
            if (!rc) rc = plan_upload(p, &bk.d_chain_hap, bk.chain_hap.data(), bk.chain_hap.size());
        }
        if (!rc && bk.bnd_elems) {
            // the units' slots start as NaN: a row read before its unit
            // wrote it would make the pair's result NaN, not a plausible
            // wrong value (the columns read unwritten lie beyond the
            // haplotype and reach no counted cell)
            const size_t nring = (size_t)bk.grid * bk.nbnd;
            rc = plan_alloc(p, &bk.d_bnd, (nring + bk.n_xslots) * bk.bnd_elems);
            if (!rc) rc = plan_fill(p, bk.d_bnd, 0, nring * bk.bnd_elems * sizeof(float));
            if (!rc)
                rc = plan_fill(p, bk.d_bnd + nring * bk.bnd_elems, 0xff,
                               (size_t)bk.n_xslots * bk.bnd_elems * sizeof(float));
        }
        if (!rc && !bk.units.empty()) rc = plan_upload(p, &bk.d_units, bk.units.data(), bk.units.size());
        if (!rc && !bk.units1.empty()) rc = plan_upload(p, &bk.d_units1, bk.units1.data(), bk.units1.size());
    }
    p->n_counters = 2 + (uint32_t)p->buckets.size() + (uint32_t)p->plain.size() + 2;
    p->long_counter = p->n_counters - 2;
    p->fault_counter = p->n_counters - 1;
    p->spin_limit = spin_limit_hook();
    for (auto &bk : p->buckets) {  // progress counters of the long-read slots
        bk.prog_base = p->n_counters;
        p->n_counters += bk.n_xslots;
        if (!bk.units1.empty()) bk.counter1 = p->n_counters++;
    }
    // FP64 rescue: one pair per wave, two waves per SIMD (its ~240 VGPRs
    // allow no more); boundary rows only if some read > 378 (one strip)
    const int wpc64 = 8;
    p->rescue_grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(n_pairs, (uint64_t)ctx->num_cus * wpc64));
    uint32_t max_h_all = 0;
    for (uint32_t h = 0; h < M; h++) max_h_all = std::max(max_h_all, b->hap_lengths[h]);
    p->rescue_lds = lds_stride_for(max_h_all, 64);
    if (max_long_h) {
        // long reads: workgroups of rescue64_waves waves (4: two per CU; 8:
        // one), 2 waves per SIMD (the kernel's VGPRs allow no more), each
        // with rescue64_waves + 1 boundary-row buffers.
        // Units of every pair that may need them, longest pipeline first:
        // blocks of a strip plus the two-block lag of each further strip
        std::vector<std::pair<uint64_t, uint32_t>> lp;
        lp.reserve(n_long);
        for (uint64_t k = 0; k < n_pairs; k++) {
            const uint32_t R = b->read_lengths[pr[k]], H = b->hap_lengths[ph[k]];
            if (R <= (uint32_t)rescue_long_rows()) continue;
            const uint64_t ns = (uint64_t)strips_for(R, 64, rescue_long_mr());
            lp.push_back({((uint64_t)H + 127) / 64 + 2 * (ns - 1), (uint32_t)k});
        }
        std::stable_sort(lp.begin(), lp.end(), [](const std::pair<uint64_t, uint32_t> &x,
                                                  const std::pair<uint64_t, uint32_t> &y) {
            return x.first > y.first;
        });
        for (const auto &e : lp)
            if (!rc && !append_units(p->long_units, p->n_x64, (int32_t)e.second,
                                     strips_for(b->read_lengths[pr[e.second]], 64, rescue_long_mr()),
                                     p->rescue64_waves))
                rc = RODV_PHMM_ERR_INVALID_ARG;
        p->prog64_base = p->n_counters;
        p->n_counters += p->n_x64;
        p->long_grid = (int)std::max<uint64_t>(
            1, std::min<uint64_t>(p->long_units.size(),
                                  (uint64_t)ctx->num_cus * (uint64_t)(fwd_long_waves() / p->rescue64_waves)));
        const size_t nring = (size_t)p->long_grid * (size_t)(p->rescue64_waves + 1);
        p->bnd64_elems = 3ull * (max_long_h + 3ull * 64 + 2);
        if (!rc) rc = plan_alloc(p, &p->d_bnd64, (nring + p->n_x64) * p->bnd64_elems);
        if (!rc) rc = plan_fill(p, p->d_bnd64, 0, nring * p->bnd64_elems * sizeof(double));
        if (!rc)
            rc = plan_fill(p, p->d_bnd64 + nring * p->bnd64_elems, 0xff,
                           (size_t)p->n_x64 * p->bnd64_elems * sizeof(double));
        if (!rc) rc = plan_upload(p, &p->d_long_units, p->long_units.data(), p->long_units.size());
    }
    p->bin_counter = p->n_counters;
    p->n_counters += 2 * kRescueClasses + 1 + 2;
    if (cfg.extended_range) {
        // one wave per pair whose FP64 sum underflows, with two boundary
        // rows of the longest haplotype
        p->ext_grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(n_pairs, (uint64_t)ctx->num_cus * 8));
        p->ext_counter = p->n_counters++;
        p->bndx_elems = 3ull * (max_h_all + 3ull * 64 + 2);
        if (!rc) rc = plan_alloc(p, &p->d_exp64, n_pairs);
        if (!rc) rc = plan_alloc(p, &p->d_bndx, (size_t)p->ext_grid * 2 * p->bndx_elems);
    }
    if (!rc) rc = plan_alloc(p, &p->d_counters, p->n_counters);
    if (!rc) rc = plan_alloc(p, &p->d_rescued_cells, 1);
    // an arena's uploads and fills reach the device here, in one copy
    if (!rc) rc = plan_flush(p);
    // descriptors of the table kernels' item slots (batched and plain order)
    if (!rc) {
        const FwdArgs a = base_args(p);
        p->d_plain_desc.assign(p->d_plain_items.size(), nullptr);
        for (size_t bi = 0; bi < p->buckets.size() && !rc; bi++) {
            Bucket &bk = p->buckets[bi];
            if (!bk.tab) continue;
            rc = plan_alloc(p, &bk.d_desc, bk.items.size());
            if (!rc && launch_desc(a, bk.d_items, bk.d_desc, bk.items.size(), nullptr) != hipSuccess)
                rc = RODV_PHMM_ERR_DEVICE;
            if (!rc && bi < p->d_plain_items.size()) {
                rc = plan_alloc(p, &p->d_plain_desc[bi], p->plain_items_n[bi]);
                if (!rc && launch_desc(a, p->d_plain_items[bi], p->d_plain_desc[bi], p->plain_items_n[bi],
                                       nullptr) != hipSuccess)
                    rc = RODV_PHMM_ERR_DEVICE;
            }
        }
        if (!rc && hipDeviceSynchronize() != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    }
    if (rc) {
        return rc;
    }
    p->stats.num_pairs = n_pairs;
    p->stats.cells = p->cells;
    p->stats.num_buckets = (uint32_t)p->buckets.size();
    p->stats.shared_pairs = n_shared;
    for (const Bucket &bk : p->buckets)
        if (bk.stack == 2)
            for (size_t k = 0; k + 1 < bk.chain_off.size(); k++)
                p->stats.stacked_pairs += (uint64_t)(bk.chain_off[k + 1] - bk.chain_off[k]) * (bk.items.size() / (bk.chain_off.size() - 1));
    *out = p;
    p = nullptr;  // the caller owns it now
    return RODV_PHMM_OK;
}

int rodv_phmm_plan_create(rodv_phmm_ctx *ctx, const rodv_phmm_batch *b,
                          const rodv_phmm_config *cfg_in, rodv_phmm_plan **out) {
    return guarded([&] { return plan_create_impl(ctx, b, cfg_in, out); });
}

}  // extern "C"

namespace {

FwdArgs base_args(rodv_phmm_plan *p) {
    FwdArgs a;
    std::memset(&a, 0, sizeof(a));
    a.tab = p->ctx->d_tab;
    a.hapc = p->d_hapc;
    a.hapc_off = p->d_hapc_off;
    a.hap_len = p->d_hap_len;
    a.initd_f = p->d_initd_f;
    a.initd_d = p->d_initd_d;
    a.rcode = p->d_rcode;
    a.q = p->d_q;
    a.ins = p->d_ins;
    a.del = p->d_del;
    a.gcp = p->d_gcp;
    a.read_off = p->d_read_off;
    a.read_len = p->d_read_len;
    a.pair_hap = p->d_pair_hap;
    a.pair_read = p->d_pair_read;
    a.pair_inter = p->d_pair_inter;
    a.raw32 = p->d_raw32;
    a.raw64 = p->d_raw64;
    a.flags = p->d_flags;
    a.rescue_list = p->d_rescue;
    a.rescue_count = p->d_counters;
    a.exp64 = p->d_exp64;
    a.bin = p->d_bin;
    a.bin_cnt = p->d_counters + p->bin_counter;
    a.hap_n = p->d_hap_n;
    for (int c = 0; c <= kRescueClasses; c++) {
        a.bin_base[c] = p->bin_base[c];
        a.bin_cap[c] = p->bin_cap[c];
    }
    a.bin_route = p->any_inter ? 1u : 0u;
    return a;
}

// Launch sequence.  mode 0: batched (one persistent launch per bucket);
// mode 1: plain (one launch per region and bucket).  Events, if given,
// bracket the FP32 part, the FP64 rescue and the finalisation.
int enqueue(rodv_phmm_plan *p, hipStream_t st, int mode, hipEvent_t *ev) {
    uint64_t launches = 0;
    HIPCHK(launch_reset(p->d_counters, p->n_counters, st));
    launches++;
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    // schedule_pairhmm's use_double only reroutes pairs from the inter to the
    // intra kernel, which is FP32-first anyway (pairhmm_schedule.cpp:193,
    // 259-262, 307-335): under intra semantics it changes nothing.
    {
        if (mode == 0) {
            // buckets go round the launch stream and the context's side
            // streams, so that one bucket's tail overlaps the next buckets'
            // start (each persistent grid fills the device on its own)
            const int ns = (int)std::min<size_t>(p->buckets.size(), 1 + kSideStreams);
            if (ns > 1) HIPCHK(hipEventRecord(p->ctx->fork_ev, st));
            for (int k = 0; k + 1 < ns; k++) HIPCHK(hipStreamWaitEvent(p->ctx->side[k], p->ctx->fork_ev, 0));
            for (size_t bi = 0; bi < p->buckets.size(); bi++) {
                Bucket &bk = p->buckets[bi];
                if (bk.n_items == 0) continue;  // every pair of the bucket chained
                const int si = (int)(bi % (size_t)ns);
                const hipStream_t bs = si == 0 ? st : p->ctx->side[si - 1];
                FwdArgs a = base_args(p);
                a.items = bk.d_items;
                a.desc = bk.d_desc;
                a.chain_off = bk.d_chain_off;
                a.chain_hap = bk.d_chain_hap;
                a.n_items = bk.n_items;
                a.counter = p->d_counters + 2 + bi;
                a.lds_stride = bk.lds_stride;
                a.bnd = bk.d_bnd;
                a.bnd_elems = bk.bnd_elems;
                a.max_strips = bk.max_strips;
                if (bk.d_units1) {  // fwd_long_kernel over units of strips
                    a.prog = p->d_counters + bk.prog_base;
                    a.fault = p->d_counters + p->fault_counter;
                    a.spin_limit = p->spin_limit;
                    a.xbnd = bk.d_bnd + (size_t)bk.grid * bk.nbnd * bk.bnd_elems;
                    // strip 0 of every pair, finding the pairs whose FP32 sum
                    // is zero; then the other strips of the rest
                    FwdArgs a1 = a;
                    a1.units = bk.d_units1;
                    a1.n_items = (uint32_t)bk.units1.size();
                    a1.counter = p->d_counters + bk.counter1;
                    a1.dead_mode = 1;
                    a1.off_mr = (uint32_t)bk.MR;
                    const size_t lds1 = (size_t)(bk.lds_stride + bk.tab_bytes);
                    HIPCHK(launch_fwd_block(bk.kern1, bk.grid, bk.block, lds1, bs, a1));
                    launches++;
                    if (bk.units.empty()) continue;
                    a.units = bk.d_units;
                    a.n_items = (uint32_t)bk.units.size();
                    a.dead_mode = 2;
                }
                // chained items: G prior tables and two code buffers per wave
                const size_t lds = bk.chain ? (size_t)(64 / bk.L) * bk.tab_bytes + 2u * bk.lds_stride
                                            : (size_t)(64 / bk.L) * (bk.lds_stride + bk.tab_bytes);
                HIPCHK(launch_fwd_block(bk.kern, bk.grid, bk.block, lds, bs, a));
                launches++;
            }
            for (int k = 0; k + 1 < ns; k++) {
                HIPCHK(hipEventRecord(p->ctx->join_ev[k], p->ctx->side[k]));
                HIPCHK(hipStreamWaitEvent(st, p->ctx->join_ev[k], 0));
            }
        } else {
            for (size_t k = 0; k < p->plain.size(); k++) {
                const PlainLaunch &pl = p->plain[k];
                Bucket &bk = p->buckets[pl.bucket];
                const int G = 64 / bk.L;
                FwdArgs a = base_args(p);
                a.items = p->d_plain_items[pl.bucket] + (size_t)pl.item_begin * G;
                if (bk.tab) a.desc = p->d_plain_desc[pl.bucket] + (size_t)pl.item_begin * G;
                a.n_items = pl.n_items;
                a.counter = p->d_counters + 2 + p->buckets.size() + k;
                a.lds_stride = bk.lds_stride;
                a.bnd = bk.d_bnd;
                a.bnd_elems = bk.bnd_elems;
                a.max_strips = bk.max_strips;
                const size_t lds = (size_t)G * (bk.lds_stride + bk.tab_bytes);
                HIPCHK(launch_fwd_block(bk.kern, std::min(pl.grid, bk.grid_cap), bk.block, lds, st, a));
                launches++;
            }
        }
    }
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    if (p->any_inter && p->cfg.simd_width_double > 0) {
        // schedule_pairhmm's FP64 stage groups the rejected inter pairs
        RegroupArgs r;
        r.list = p->d_sched_list;
        r.off = p->d_sched_off;
        r.n_regions = p->n_regions;
        r.wd = p->cfg.simd_width_double;
        r.idle_d = p->cfg.max_idle_ratio_double;
        r.pair_hap = p->d_pair_hap;
        r.pair_read = p->d_pair_read;
        r.hap_len = p->d_hap_len;
        r.read_len = p->d_read_len;
        r.flags = p->d_flags;
        r.scratch = p->d_sched_scratch;
        HIPCHK(launch_regroup(r, st));
        launches++;
    }
    {
        FwdArgs a = base_args(p);
        a.counter = p->d_counters + 1;
        a.n_items_dev = p->d_counters;
        a.lds_stride = p->rescue_lds;
        a.bnd = p->d_bnd64;
        a.bnd_elems = p->bnd64_elems;
        if (p->d_bin) {
            // rescued pairs of the table classes two per wave, the others
            // (bin list kRescueClasses) one per wave below
            HIPCHK(launch_rescue_bin(a, std::max(1, (int)p->ctx->num_cus), st));
            launches++;
            for (int part = 0; part < 2; part++) {
                if (!p->rtab_grid[part]) continue;
                FwdArgs t = a;
                t.counter = a.bin_cnt + 2 * kRescueClasses + 1 + part;
                t.lds_stride = p->rtab_stride;
                const size_t lds = 2u * (p->rtab_stride + rescue_tab_bytes(part == 0 ? 8 : 4));
                HIPCHK(launch_rescue_tab(part, p->share_all, p->rtab_grid[part], lds, st, t));
                launches++;
            }
            a.rescue_list = p->d_bin + p->bin_base[kRescueClasses];
            a.n_items_dev = a.bin_cnt + 2 * kRescueClasses;
        }
        HIPCHK(launch_fwd(fwd_f64_rescue_kernel(p->any_inter), p->rescue_grid, p->rescue_lds, st, a));
        launches++;
        if (p->long_grid) {  // the rescued reads longer than one strip
            a.counter = p->d_counters + p->long_counter;
            a.units = p->d_long_units;
            a.n_items = (uint32_t)p->long_units.size();
            a.prog = p->d_counters + p->prog64_base;
            a.fault = p->d_counters + p->fault_counter;
            a.spin_limit = p->spin_limit;
            a.xbnd = p->d_bnd64 + (size_t)p->long_grid * (p->rescue64_waves + 1) * p->bnd64_elems;
            const size_t llds = (size_t)p->rescue_lds +
                                (p->long_tab ? long_tab_bytes(rescue_long_mr(), true, p->rescue64_waves) : 0u);
            HIPCHK(launch_rescue_long(p->any_inter, p->long_tab, p->rescue64_waves, p->long_grid, llds, st, a));
            launches++;
        }
    }
    if (p->ext_grid) {  // extended range: the pairs whose FP64 sum is zero
        FwdArgs a = base_args(p);
        a.counter = p->d_counters + p->ext_counter;
        a.n_items_dev = p->d_counters;
        a.lds_stride = p->rescue_lds;
        a.bnd = p->d_bndx;
        a.bnd_elems = p->bndx_elems;
        HIPCHK(launch_ext(p->ext_grid, p->rescue_lds, st, a));
        launches++;
    }
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    FinalizeArgs f;
    f.raw32 = p->d_raw32;
    f.raw64 = p->d_raw64;
    f.exp64 = p->d_exp64;
    f.pair_inter = p->d_pair_inter;
    f.out = p->d_out;
    f.flags = p->d_flags;
    f.n = p->n_pairs;
    HIPCHK(launch_finalize(f, st));
    launches++;
    if (ev) HIPCHK(hipEventRecord(ev[3], st));
    p->stats.launches = launches;
    p->last_stream = st;
    return RODV_PHMM_OK;
}

}  // namespace

extern "C" {

static int plan_execute_impl(rodv_phmm_plan *p, void *stream) {
    if (!p) return RODV_PHMM_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(p->ctx->device));
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!p->cfg.use_graph) return enqueue(p, st, 0, nullptr);
    if (!p->graph_exec) {
        // capture on a private stream; the graph is replayed on `st`
        hipStream_t cap;
        HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipStreamDestroy(cap);
            return RODV_PHMM_ERR_DEVICE;
        }
        int rc = enqueue(p, cap, 0, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return RODV_PHMM_ERR_DEVICE;
        p->graph = g;
        if (hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            p->graph_exec = nullptr;
            return RODV_PHMM_ERR_DEVICE;
        }
    }
    HIPCHK(hipGraphLaunch(p->graph_exec, st));
    p->last_stream = st;
    return RODV_PHMM_OK;
}

int rodv_phmm_plan_execute(rodv_phmm_plan *p, void *stream) {
    return guarded([&] { return plan_execute_impl(p, stream); });
}

static int plan_execute_plain_impl(rodv_phmm_plan *p, void *stream) {
    if (!p) return RODV_PHMM_ERR_INVALID_ARG;
    if (p->no_plain) return RODV_PHMM_ERR_UNSUPPORTED;
    HIPCHK(hipSetDevice(p->ctx->device));
    return enqueue(p, reinterpret_cast<hipStream_t>(stream), 1, nullptr);
}

int rodv_phmm_plan_execute_plain(rodv_phmm_plan *p, void *stream) {
    return guarded([&] { return plan_execute_plain_impl(p, stream); });
}

static int plan_profile_impl(rodv_phmm_plan *p, void *stream) {
    if (!p) return RODV_PHMM_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(p->ctx->device));
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t ev[4];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    int rc = enqueue(p, st, 0, ev);
    if (!rc && hipEventSynchronize(ev[3]) != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    if (!rc) {
        float a = 0, b = 0, c = 0;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        (void)hipEventElapsedTime(&c, ev[0], ev[3]);
        p->stats.fp32_kernel_ms = a;
        p->stats.fp64_kernel_ms = b;
        p->stats.total_ms = c;
        uint32_t cnt = 0;
        if (hipMemcpy(&cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost) == hipSuccess)
            p->stats.rescued_pairs = cnt;
        // the FP64 pass's cells: R * H summed over the rescue list
        unsigned long long rc_cells = 0;
        if (launch_reset(reinterpret_cast<uint32_t *>(p->d_rescued_cells), 2, st) != hipSuccess ||
            launch_rescued_cells(base_args(p), p->d_rescued_cells, st) != hipSuccess ||
            hipMemcpyAsync(&rc_cells, p->d_rescued_cells, sizeof(rc_cells), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = RODV_PHMM_ERR_DEVICE;
        p->stats.rescued_cells = rc_cells;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int rodv_phmm_plan_profile(rodv_phmm_plan *p, void *stream) {
    return guarded([&] { return plan_profile_impl(p, stream); });
}

}  // extern "C"

// ---------------------------------------------------------------- autotuner

// (outside the extern "C" block: helpers of C++ linkage, not exported)
namespace {
constexpr uint32_t kMaxTunePairs = 131072;

int lanes_of(uint32_t pairs_per_wave) {
    return pairs_per_wave == 8 ? 8 : pairs_per_wave == 4 ? 16 : pairs_per_wave == 2 ? 32
                                                                 : pairs_per_wave == 1 ? 64 : 0;
}

// FP32 forward of the sub-batch under the context's current (forced) shape:
// best of three profiled executions after a warm-up, in GCUPS.
int time_candidate(rodv_phmm_ctx *ctx, const rodv_phmm_batch *sub, const rodv_phmm_config &cfg,
                   double *gcups) {
    rodv_phmm_plan *p = nullptr;
    int rc = rodv_phmm_plan_create(ctx, sub, &cfg, &p);
    if (rc) return rc;
    double best = 1e30;
    for (int r = 0; r < 4 && !rc; r++) {
        rc = rodv_phmm_plan_profile(p, nullptr);
        if (!rc && r > 0) best = std::min(best, p->stats.fp32_kernel_ms);
    }
    const uint64_t cells = p->cells;
    rodv_phmm_plan_destroy(p);
    if (rc) return rc;
    *gcups = best > 0 ? (double)cells / best / 1e6 : 0.0;
    return RODV_PHMM_OK;
}
}  // namespace

extern "C" {

static int autotune_impl(rodv_phmm_ctx *ctx, const rodv_phmm_batch *sample,
                       const rodv_phmm_config *cfg_in, rodv_phmm_tuning *out) {
    if (!ctx || !sample || !out) return RODV_PHMM_ERR_INVALID_ARG;
    int rc = validate(sample);
    if (rc) return rc;
    rodv_phmm_config cfg;
    rodv_phmm_config_default(&cfg);
    if (cfg_in) {
        if (cfg_in->struct_size != sizeof(rodv_phmm_config)) return RODV_PHMM_ERR_INVALID_ARG;
        cfg = *cfg_in;
    }
    // the shape the tuner forces while timing is cleared on every way out,
    // an exception included
    struct Unforce {
        rodv_phmm_ctx *c;
        ~Unforce() {
            c->force_class = -1;
            c->force = TuneEntry{};
        }
    } unforce{ctx};
    cfg.semantics = RODV_PHMM_SEMANTICS_INTRA;  // shapes do not depend on routing
    cfg.use_graph = 0;
    cfg.use_tuning = 0;
    cfg.waves_per_cu = 0;
    // the sample's pairs by class, every k-th so that each class keeps at
    // most kMaxTunePairs spread over the whole sample
    std::vector<rodv_phmm_region> regions;
    if (sample->num_regions == 0)
        regions.push_back(rodv_phmm_region{0, sample->num_haps, 0, sample->num_reads});
    else
        regions.assign(sample->regions, sample->regions + sample->num_regions);
    std::vector<uint64_t> count(RODV_PHMM_TUNE_CLASSES, 0);
    for (const auto &rg : regions)
        for (uint32_t c = 0; c < rg.num_reads; c++)
            count[tune_class(sample->read_lengths[rg.read_begin + c])] += rg.num_haps;
    std::vector<std::vector<rodv_phmm_region>> picked(RODV_PHMM_TUNE_CLASSES);
    std::vector<uint64_t> seen(RODV_PHMM_TUNE_CLASSES, 0);
    for (const auto &rg : regions)
        for (uint32_t a = 0; a < rg.num_haps; a++)
            for (uint32_t c = 0; c < rg.num_reads; c++) {
                const int k = tune_class(sample->read_lengths[rg.read_begin + c]);
                const uint64_t stride = (count[k] + kMaxTunePairs - 1) / kMaxTunePairs;
                if (seen[k]++ % stride == 0 && picked[k].size() < kMaxTunePairs)
                    picked[k].push_back(rodv_phmm_region{rg.hap_begin + a, 1, rg.read_begin + c, 1});
            }
    rodv_phmm_tuning res{};
    res.struct_size = sizeof(res);
    res.num_classes = RODV_PHMM_TUNE_CLASSES;
    TuneEntry tuned[RODV_PHMM_TUNE_CLASSES];
    for (int k = 0; k < RODV_PHMM_TUNE_CLASSES && !rc; k++) {
        rodv_phmm_tune_entry &e = res.cls[k];
        e.max_read_len = kClassMax[k];
        if (picked[k].size() < 64) continue;  // too few pairs to time: keep the defaults
        e.sample_pairs = (uint32_t)picked[k].size();
        // time every candidate on enough pairs to give each persistent wave
        // several items of the largest (8-pair) kind, repeating the class's
        // sample if needed: on fewer, shapes with fewer pairs per item win
        // by filling the device better, which a full-size batch does not need
        const size_t n_pick = picked[k].size();
        for (size_t i = n_pick; i < tune_pairs_min(ctx->num_cus); i++)
            picked[k].push_back(picked[k][i % n_pick]);
        rodv_phmm_batch sub = *sample;
        sub.num_regions = (uint32_t)picked[k].size();
        sub.regions = picked[k].data();
        // built-in shape
        ctx->force_class = k;
        ctx->force = TuneEntry{};
        rc = time_candidate(ctx, &sub, cfg, &e.default_gcups);
        if (rc) break;
        TuneEntry best{};
        double best_g = e.default_gcups;
        std::vector<TuneEntry> cands;
        if (kClassMax[k] <= 95 && kClassMax[k] >= 47) cands.push_back(TuneEntry{8, 0, 0});
        if (kClassMax[k] <= 255) cands.push_back(TuneEntry{16, 0, 0});
        if (kClassMax[k] <= 319) cands.push_back(TuneEntry{32, 0, 0});
        if (kClassMax[k] <= 319) {
            cands.push_back(TuneEntry{64, 0, 0});
        } else {
            for (int cap : {4, 8, 12, 16}) cands.push_back(TuneEntry{64, cap, 0});
        }
        for (const TuneEntry &c : cands) {
            ctx->force = c;
            double g = 0.0;
            rc = time_candidate(ctx, &sub, cfg, &g);
            if (rc) break;
            if (g > best_g) {
                best_g = g;
                best = c;
            }
        }
        if (rc) break;
        if (best.L == 0)  // the built-in shape won: name it
            best.L = default_tab_lanes(kClassMax[k], cfg.short_lanes);
        for (int w : {4, 8, 16}) {
            TuneEntry c = best;
            c.wpc = w;
            ctx->force = c;
            double g = 0.0;
            rc = time_candidate(ctx, &sub, cfg, &g);
            if (rc) break;
            if (g > best_g * 1.01) {  // a clear gain over the default grid only
                best_g = g;
                best.wpc = w;
            }
        }
        if (rc) break;
        tuned[k] = best;
        e.pairs_per_wave = 64u / (uint32_t)best.L;
        e.max_rows_per_lane = (uint32_t)best.mr_cap;
        e.waves_per_cu = (uint32_t)best.wpc;
        e.gcups = best_g;
    }
    ctx->force_class = -1;
    ctx->force = TuneEntry{};
    if (rc) return rc;
    for (int k = 0; k < RODV_PHMM_TUNE_CLASSES; k++) ctx->tune[k] = tuned[k];
    ctx->tuned = true;
    *out = res;
    return RODV_PHMM_OK;
}

int rodv_phmm_autotune(rodv_phmm_ctx *ctx, const rodv_phmm_batch *sample,
                       const rodv_phmm_config *cfg_in, rodv_phmm_tuning *out) {
    return guarded([&] { return autotune_impl(ctx, sample, cfg_in, out); });
}

static int set_tuning_impl(rodv_phmm_ctx *ctx, const rodv_phmm_tuning *t) {
    if (!ctx) return RODV_PHMM_ERR_INVALID_ARG;
    if (!t) {
        ctx->tuned = false;
        for (auto &e : ctx->tune) e = TuneEntry{};
        return RODV_PHMM_OK;
    }
    if (t->struct_size != sizeof(rodv_phmm_tuning) || t->num_classes != RODV_PHMM_TUNE_CLASSES)
        return RODV_PHMM_ERR_INVALID_ARG;
    TuneEntry tmp[RODV_PHMM_TUNE_CLASSES];
    for (int k = 0; k < RODV_PHMM_TUNE_CLASSES; k++) {
        const rodv_phmm_tune_entry &e = t->cls[k];
        const uint32_t ppw = e.pairs_per_wave, cap = e.max_rows_per_lane;
        if (ppw != 0 && lanes_of(ppw) == 0) return RODV_PHMM_ERR_INVALID_ARG;
        if (cap != 0 && cap != 4 && cap != 8 && cap != 12 && cap != 16) return RODV_PHMM_ERR_INVALID_ARG;
        if (e.waves_per_cu > 32) return RODV_PHMM_ERR_INVALID_ARG;
        tmp[k] = TuneEntry{lanes_of(ppw), (int)cap, (int)e.waves_per_cu};
    }
    for (int k = 0; k < RODV_PHMM_TUNE_CLASSES; k++) ctx->tune[k] = tmp[k];
    ctx->tuned = true;
    return RODV_PHMM_OK;
}

int rodv_phmm_set_tuning(rodv_phmm_ctx *ctx, const rodv_phmm_tuning *t) {
    return guarded([&] { return set_tuning_impl(ctx, t); });
}

static int get_tuning_impl(const rodv_phmm_ctx *ctx, rodv_phmm_tuning *out) {
    if (!ctx || !out) return RODV_PHMM_ERR_INVALID_ARG;
    rodv_phmm_tuning r{};
    r.struct_size = sizeof(r);
    r.num_classes = RODV_PHMM_TUNE_CLASSES;
    for (int k = 0; k < RODV_PHMM_TUNE_CLASSES; k++) {
        r.cls[k].max_read_len = kClassMax[k];
        if (!ctx->tuned) continue;
        const TuneEntry &e = ctx->tune[k];
        r.cls[k].pairs_per_wave = e.L ? 64u / (uint32_t)e.L : 0u;
        r.cls[k].max_rows_per_lane = (uint32_t)e.mr_cap;
        r.cls[k].waves_per_cu = (uint32_t)e.wpc;
    }
    *out = r;
    return RODV_PHMM_OK;
}

int rodv_phmm_get_tuning(const rodv_phmm_ctx *ctx, rodv_phmm_tuning *out) {
    return guarded([&] { return get_tuning_impl(ctx, out); });
}

static int plan_read_impl(rodv_phmm_plan *p, double *out_log10, uint8_t *out_flags,
                        float *out_raw32, double *out_raw64) {
    if (!p) return RODV_PHMM_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->last_stream));
    const size_t n = p->n_pairs;
    if (out_log10) HIPCHK(hipMemcpy(out_log10, p->d_out, n * sizeof(double), hipMemcpyDeviceToHost));
    if (out_flags) HIPCHK(hipMemcpy(out_flags, p->d_flags, n, hipMemcpyDeviceToHost));
    if (out_raw32) HIPCHK(hipMemcpy(out_raw32, p->d_raw32, n * sizeof(float), hipMemcpyDeviceToHost));
    if (out_raw64) HIPCHK(hipMemcpy(out_raw64, p->d_raw64, n * sizeof(double), hipMemcpyDeviceToHost));
    uint32_t cnt = 0, faults = 0;
    HIPCHK(hipMemcpy(&cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&faults, p->d_counters + p->fault_counter, sizeof(faults), hipMemcpyDeviceToHost));
    p->stats.rescued_pairs = cnt;
    // a long read's units waited past their bound for one another: those
    // pairs' results are NaN
    if (faults) return RODV_PHMM_ERR_DEVICE;
    return RODV_PHMM_OK;
}

int rodv_phmm_plan_read(rodv_phmm_plan *p, double *out_log10, uint8_t *out_flags,
                        float *out_raw32, double *out_raw64) {
    return guarded([&] { return plan_read_impl(p, out_log10, out_flags, out_raw32, out_raw64); });
}

int rodv_phmm_plan_stats(const rodv_phmm_plan *p, rodv_phmm_stats *out) {
    if (!p || !out) return RODV_PHMM_ERR_INVALID_ARG;
    *out = p->stats;
    return RODV_PHMM_OK;
}

uint64_t rodv_phmm_plan_num_pairs(const rodv_phmm_plan *p) { return p ? p->n_pairs : 0; }

void rodv_phmm_plan_destroy(rodv_phmm_plan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    if (p->last_stream || true) (void)hipDeviceSynchronize();
    free_plan(p);
}

}  // extern "C"

namespace {

// Largest arena rodv_phmm_forward keeps in its context; a batch whose plan
// needs more allocates (and frees) its own buffers.
constexpr size_t kFwdArenaMax = size_t(256) << 20;

// (Re)allocates the context's forward arena to at least `bytes`.
int grow_arena(Arena &A, size_t bytes) {
    bytes = std::max<size_t>(bytes, size_t(4) << 20);
    bytes = (bytes + ((size_t(1) << 20) - 1)) & ~((size_t(1) << 20) - 1);
    if (A.dev) (void)hipFree(A.dev);
    if (A.host) (void)hipHostFree(A.host);
    A = Arena{};
    if (fail_alloc_hook()) return RODV_PHMM_ERR_OOM;
    if (hipMalloc(reinterpret_cast<void **>(&A.dev), bytes) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&A.host), bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        if (A.dev) (void)hipFree(A.dev);
        A = Arena{};
        return RODV_PHMM_ERR_OOM;
    }
    A.cap = bytes;
    return RODV_PHMM_OK;
}

}  // namespace

namespace rodv {
int ctx_device(const rodv_phmm_ctx *ctx) { return ctx->device; }

int forward_in(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch, const rodv_phmm_config &c,
               Arena *arena, double *out_log10, uint8_t *out_flags) {
    rodv_phmm_plan *p = nullptr;
    int rc = plan_create_impl(ctx, batch, &c, &p, arena, true);
    if (rc) return rc;
    rc = plan_execute_impl(p, nullptr);
    if (!rc) rc = plan_read_impl(p, out_log10, out_flags, nullptr, nullptr);
    rodv_phmm_plan_destroy(p);
    return rc;
}
}  // namespace rodv

extern "C" {

static int forward_impl(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_config *cfg, double *out_log10, uint8_t *out_flags) {
    if (!ctx || !out_log10) return RODV_PHMM_ERR_INVALID_ARG;
    rodv_phmm_config c;
    rodv_phmm_config_default(&c);
    if (cfg) {
        if (cfg->struct_size != sizeof(rodv_phmm_config)) return RODV_PHMM_ERR_INVALID_ARG;
        c = *cfg;
    }
    c.use_graph = 0;  // one-shot: capture would not pay off
    HIPCHK(hipSetDevice(ctx->device));
    // the context's arena, grown until the plan fits (it then stays that
    // size for the next calls); plans larger than kFwdArenaMax allocate
    Arena &A = ctx->fwd_arena;
    for (int attempt = 0; attempt < 8; attempt++) {
        if (!A.dev) {
            const int g = grow_arena(A, 0);
            if (g) return g;
        }
        const int rc = forward_in(ctx, batch, c, &A, out_log10, out_flags);
        if (rc != kArenaFull) return rc;
        const size_t want = std::max(2 * A.cap, 2 * A.need);
        if (want > kFwdArenaMax) break;
        const int g = grow_arena(A, want);
        if (g) return g;
    }
    return forward_in(ctx, batch, c, nullptr, out_log10, out_flags);
}

int rodv_phmm_forward(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_config *cfg, double *out_log10, uint8_t *out_flags) {
    return guarded([&] { return forward_impl(ctx, batch, cfg, out_log10, out_flags); });
}

// Shape of a one-strip read in adj2_kernel: a strip holds (L - 1) * MR rows
// (lane L-1 holds identity rows); the fewest lanes per pair that hold the
// read with at most kAdjRows rows per lane (more pairs per wave, a shorter
// fill and drain of the anti-diagonal pipeline), else 64 lanes with up to
// adj2_max_mr() rows.
constexpr int kAdjRows = 5;
std::pair<int, int> adj_shape(uint32_t R) {
    for (int L : {16, 32}) {
        const int mr = (int)((R + (uint32_t)L - 2) / (uint32_t)(L - 1));
        if (mr <= kAdjRows) return {L, std::max(1, mr)};
    }
    return {64, std::max(1, (int)((R + 62) / 63))};
}

static int adjoint_impl(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_params *params, const double *weights, int ckpt_stride,
                      double *out_log10, double *grad, double *kernel_ms) {
    if (!ctx || !batch || !params || !grad) return RODV_PHMM_ERR_INVALID_ARG;
    if (ckpt_stride == 0) ckpt_stride = 32;
    if (ckpt_stride < 4 || ckpt_stride > 256 || ckpt_stride % 4 != 0) return RODV_PHMM_ERR_INVALID_ARG;
    if (!(params->gap_open > 0.0) || !(params->gap_cont > 0.0)) return RODV_PHMM_ERR_INVALID_ARG;
    for (int b = 0; b < 8; b++)
        if (!(params->scale[b] >= 0.0)) return RODV_PHMM_ERR_INVALID_ARG;
    // upload and validate through the forward planner (same encodings)
    rodv_phmm_config cfg;
    rodv_phmm_config_default(&cfg);
    cfg.use_graph = 0;
    cfg.semantics = RODV_PHMM_SEMANTICS_INTRA;  // no routing: only the uploads are used
    rodv_phmm_plan *p = nullptr;
    int rc = rodv_phmm_plan_create(ctx, batch, &cfg, &p);
    if (rc) return rc;
    const uint64_t n = p->n_pairs;
    // per-pair lengths, bucketed by rows per lane (one pair per wave)
    std::vector<uint32_t> ph(n), pr(n);
    if (hipMemcpy(ph.data(), p->d_pair_hap, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pr.data(), p->d_pair_read, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rodv_phmm_plan_destroy(p);
        return RODV_PHMM_ERR_DEVICE;
    }
    const int maxmr = adj2_max_mr();
    // one-strip reads: buckets by shape (lanes per pair, rows per lane) of
    // adj2_kernel, G = 64 / L pairs per item
    std::vector<std::pair<int, int>> shapes;
    std::vector<std::vector<int32_t>> bucket;
    uint32_t max_h = 0;
    // reads longer than one strip (63 * maxmr rows): the long-read kernel,
    // its strips of 63 * lmr rows in sequence, longest work first
    const int lmr = adj_long_mr();
    std::vector<std::pair<uint64_t, int32_t>> longs;
    uint32_t max_h_long = 0;
    int max_strips = 1;
    for (uint64_t k = 0; k < n; k++) {
        const uint32_t R = batch->read_lengths[pr[k]], H = batch->hap_lengths[ph[k]];
        const int mr = (int)((R + 62) / 63);
        if (mr > maxmr) {
            const int off = (lmr - (int)(R % (uint32_t)lmr)) % lmr;
            const int nstr = ((int)R - 1 + off) / (63 * lmr) + 1;
            if (nstr > adj_long_max_strips()) {
                rodv_phmm_plan_destroy(p);
                return RODV_PHMM_ERR_UNSUPPORTED;  // more strips than the kernel keeps shifts for
            }
            longs.push_back({(uint64_t)R * H, (int32_t)k});
            max_h_long = std::max(max_h_long, H);
            max_strips = std::max(max_strips, nstr);
            continue;
        }
        const std::pair<int, int> sh = adj_shape(R);
        size_t bi = 0;
        while (bi < shapes.size() && shapes[bi] != sh) bi++;
        if (bi == shapes.size()) {
            shapes.push_back(sh);
            bucket.emplace_back();
        }
        bucket[bi].push_back((int32_t)k);
        max_h = std::max(max_h, H);
    }
    std::stable_sort(longs.begin(), longs.end(), [](const std::pair<uint64_t, int32_t> &x,
                                                    const std::pair<uint64_t, int32_t> &y) {
        return x.first > y.first;
    });
    // per-quality tables of the parametrised model (double, then rounded)
    const HostTables &t = tables();
    const double delta = std::pow(10.0, -params->gap_open / 10.0);
    const double gamma = std::pow(10.0, -params->gap_cont / 10.0);
    std::vector<float> xf(128), yf(128);
    std::vector<double> xd(128), yd(128), wq(128);
    std::vector<uint8_t> qbin(128);
    for (int q = 0; q < 128; q++) {
        const int b = std::min(q, 47) / 6;
        const double e = params->scale[b] * t.ph2pr_d[q];
        xd[q] = 1.0 - e;
        yd[q] = e / 3.0;
        xf[q] = (float)xd[q];
        yf[q] = (float)yd[q];
        wq[q] = t.ph2pr_d[q];
        qbin[q] = (uint8_t)b;
    }
    float *d_xf = nullptr, *d_yf = nullptr;
    double *d_xd = nullptr, *d_yd = nullptr, *d_wq = nullptr, *d_out = nullptr, *d_grad = nullptr;
    uint8_t *d_qbin = nullptr, *d_flags = nullptr;
    int32_t *d_items = nullptr, *d_litems = nullptr;
    uint32_t *d_cnt = nullptr, *d_resc = nullptr;
    void *d_ck = nullptr;
    double *d_lck = nullptr, *d_lbnd = nullptr, *d_lab = nullptr;
    // per bucket its items (G slots each, -1 padding the last) and a rescue
    // list of the bucket's pair count
    const size_t nbk = shapes.size();
    std::vector<int32_t> all_items;
    std::vector<size_t> item_begin(nbk + 1, 0), pair_begin(nbk + 1, 0);
    for (size_t bi = 0; bi < nbk; bi++) {
        const size_t G = (size_t)(64 / shapes[bi].first);
        item_begin[bi] = all_items.size();
        pair_begin[bi + 1] = pair_begin[bi] + bucket[bi].size();
        all_items.insert(all_items.end(), bucket[bi].begin(), bucket[bi].end());
        while ((all_items.size() - item_begin[bi]) % G) all_items.push_back(-1);
    }
    item_begin[nbk] = all_items.size();
    // two waves per SIMD: the adjoint kernels use up to 256 VGPRs, so more
    // would not be resident at once
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(n, (uint64_t)ctx->num_cus * 8));
    const uint32_t stride = lds_stride_for(max_h, 64);  // per group of any shape
    const uint64_t d4max = ((uint64_t)max_h + 63 + 3) / 4 * 4;
    const uint64_t nseg_max = d4max / 4 + 2;  // enough for any stride >= 4
    const uint64_t ck_elems = nseg_max * (3ull * maxmr + 3) * 64;
    int maxlds = 65536;
    (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    rc = upload(&d_xf, xf.data(), 128);
    if (!rc) rc = upload(&d_yf, yf.data(), 128);
    if (!rc) rc = upload(&d_xd, xd.data(), 128);
    if (!rc) rc = upload(&d_yd, yd.data(), 128);
    if (!rc) rc = upload(&d_wq, wq.data(), 128);
    if (!rc) rc = upload(&d_qbin, qbin.data(), 128);
    if (!rc) rc = upload(&d_items, all_items.data(), all_items.size());
    if (!rc) rc = dev_alloc(&d_out, n);
    if (!rc) rc = dev_alloc(&d_grad, 10 * n);
    if (!rc) rc = dev_alloc(&d_flags, n);
    if (!rc) rc = dev_alloc(&d_resc, n);
    if (!rc) rc = dev_alloc(&d_cnt, 3 * (size_t)nbk + 3);  // + the long-read kernels' two and their faults
    if (!rc) rc = dev_alloc(reinterpret_cast<double **>(&d_ck), (size_t)grid * ck_elems);
    // long reads, by default a pair's strips over several workgroups (pass A
    // per pair, pass B per strip, in one launch or, for tests or where LDS
    // is short, two; then the totals per pair):
    // per pair every strip's last row but the last strip's, per strip a row
    // of handed-up adjoints, and per pass-B wave the level-1 checkpoints of a
    // strip.  RODV_PHMM_ADJ_LONG_WAVES=1 or 4 (tests) runs each pair on one
    // workgroup of that many waves instead (per block the rows, waves + 1
    // rows of handed-up adjoints, per wave the checkpoints), the same bits.
    const int lhook = adj_long_waves_hook();
    const bool lsplit = lhook != 1 && lhook != adj_long_waves();
    const int lwaves = lhook == 1 ? 1 : adj_long_waves();
    const int lgrid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>(longs.size(), (uint64_t)ctx->num_cus * (lwaves == 1 ? 2 : 1)));
    // columns of a row: up to H + 2L + 4 read by the forward, + a checkpoint
    // segment (<= 256 steps) staged past the last step
    const uint64_t lcols = (uint64_t)max_h_long + 4 * 64 + 8 + 256;
    const uint64_t lrow = 3 * lcols, lab_elems = 2 * lcols;
    const uint64_t lbnd_stride = (uint64_t)std::max(1, max_strips - 1) * lrow;
    const uint64_t lnv = 3ull * lmr + 3;
    int lK1 = ckpt_stride;
    const uint32_t lstride = lds_stride_for(max_h_long, 64);
    auto long_lds = [&](int K1) { return lsplit ? adj_longb_lds(lstride, K1) : adj_long_lds(lstride, K1, lwaves); };
    while (lK1 > 4 && long_lds(lK1) > (size_t)maxlds) lK1 /= 2;
    lK1 = lK1 / 4 * 4;
    const uint64_t ld4max = ((uint64_t)max_h_long + 64 + 3) / 4 * 4;
    const uint64_t lck_elems = (ld4max / (uint64_t)lK1 + 2) * lnv * 64;
    // pass B of the split launches: one-wave workgroups, as many per CU as
    // the LDS holds, at most 4 (each wave holds 512 registers)
    int lds_cu = 65536;
    (void)hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx->device);
    const int bwpc = (int)std::max<size_t>(1, std::min<size_t>(4, (size_t)lds_cu / long_lds(lK1)));
    const uint64_t nlong = longs.size();
    const uint64_t lstrips = (uint64_t)max_strips;
    // the split launches run the long pairs in chunks whose per-pair buffers
    // (rows, handed-up adjoint rows, gradient contributions) fit
    // kLongChunkBytes; every chunk reuses them
    constexpr uint64_t kLongChunkBytes = 8ull << 30;
    const uint64_t pair_bytes = (lbnd_stride + lstrips * lab_elems + lstrips * (uint64_t)lmr * 4 * 64) * 8 +
                                lstrips * 8 + 16;
    const uint64_t chook = adj_long_chunk_hook();
    const uint64_t lchunk =
        std::max<uint64_t>(1, std::min<uint64_t>(nlong, chook ? chook : kLongChunkBytes / pair_bytes));
    // (pair in its chunk << 8 | strip): pairs longest first, strips last
    // first; chunk c's items from bbegin[c]; per pair its strips (lnstr)
    std::vector<uint32_t> bitems, bbegin, lnstr;
    if (lsplit)
        for (uint64_t k = 0; k < nlong; k++) {
            if (k % lchunk == 0) bbegin.push_back((uint32_t)bitems.size());
            const uint32_t R = batch->read_lengths[pr[(size_t)longs[k].second]];
            const int ns = strips_for(R, 64, lmr);
            lnstr.push_back((uint32_t)ns);
            for (int cs = ns - 1; cs >= 0; cs--) bitems.push_back((uint32_t)(k % lchunk) << 8 | (uint32_t)cs);
        }
    bbegin.push_back((uint32_t)bitems.size());
    uint64_t bmax = 1;  // items of the largest chunk
    for (size_t c = 0; c + 1 < bbegin.size(); c++) bmax = std::max<uint64_t>(bmax, bbegin[c + 1] - bbegin[c]);
    const int bgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(bmax, (uint64_t)ctx->num_cus * (uint64_t)bwpc));
    // passes A and B in one launch (adj_longab_kernel) where a workgroup's
    // LDS holds 4 pass-B waves: one workgroup per CU, or fewer for few strips
    const bool lfused = lsplit && adj_long_launches_hook() != 3 && adj_longab_lds(lstride, lK1) <= (size_t)maxlds;
    const int abw = adj_longab_waves();
    const int abgrid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)ctx->num_cus, std::max<uint64_t>(lchunk, (bmax + abw - 1) / abw)));
    // pass A leaves each strip's level-1 checkpoints in HBM for pass B (which
    // then skips its own forward) where a chunk's fit kLongCkBytes: per pair
    // adj_long_ck_elems(H) per strip, at chunk-relative offsets
    constexpr uint64_t kLongCkBytes = 16ull << 30;
    std::vector<uint64_t> ck1off;
    uint64_t ck1max = 0;  // elements of the largest chunk's
    if (lsplit) {
        uint64_t acc = 0;
        for (uint64_t k = 0; k < nlong; k++) {
            if (k % lchunk == 0) acc = 0;
            ck1off.push_back(acc);
            const uint32_t H = batch->hap_lengths[ph[(size_t)longs[k].second]];
            acc += (uint64_t)lnstr[k] * adj_long_ck_elems(H, lK1);
            ck1max = std::max(ck1max, acc);
        }
    }
    const bool lck1 = lsplit && ck1max * 8 <= kLongCkBytes && adj_long_passa_ckpt_hook();
    int *d_lsh = nullptr, *d_lexp = nullptr, *d_lprog = nullptr, *d_lready = nullptr;
    double *d_lsum = nullptr, *d_lpart = nullptr;
    uint32_t *d_bitems = nullptr, *d_lnstr = nullptr, *d_lnext = nullptr;
    double *d_lck1 = nullptr;
    uint64_t *d_lck1_off = nullptr;
    if (!longs.empty()) {
        std::vector<int32_t> li(longs.size());
        for (size_t k = 0; k < longs.size(); k++) li[k] = longs[k].second;
        if (!rc) rc = upload(&d_litems, li.data(), li.size());
        // rows are written before they are read; zeros keep the columns a
        // strip never writes finite
        if (lsplit) {
            const uint64_t nck = lfused ? (uint64_t)abgrid * abw : (uint64_t)bgrid;  // pass-B waves
            if (!rc) rc = dev_alloc(&d_lck, lck1 ? (size_t)1 : (size_t)nck * lck_elems);
            if (!rc && lck1) rc = dev_alloc(&d_lck1, (size_t)ck1max);
            if (!rc && lck1) rc = upload(&d_lck1_off, ck1off.data(), ck1off.size());
            if (!rc) rc = dev_alloc(&d_lbnd, (size_t)lchunk * lbnd_stride);
            if (!rc) rc = dev_alloc(&d_lab, (size_t)lchunk * lstrips * lab_elems);
            if (!rc) rc = dev_alloc(&d_lsh, (size_t)lchunk * lstrips);
            if (!rc) rc = dev_alloc(&d_lexp, (size_t)lchunk);
            if (!rc) rc = dev_alloc(&d_lsum, (size_t)lchunk);
            if (!rc) rc = dev_alloc(&d_lprog, (size_t)lchunk * lstrips);
            if (!rc) rc = dev_alloc(&d_lpart, (size_t)lchunk * lstrips * lmr * 4 * 64);
            if (!rc) rc = upload(&d_bitems, bitems.data(), bitems.size());
            if (!rc) rc = upload(&d_lnstr, lnstr.data(), lnstr.size());
            if (!rc) rc = dev_alloc(&d_lready, (size_t)lchunk);
            if (!rc) rc = dev_alloc(&d_lnext, (size_t)lchunk);
            // progress counts start above every step: nothing published; no
            // pair's pass A done, no strip taken
            if (!rc && (hipMemset(d_lbnd, 0, (size_t)lchunk * lbnd_stride * 8) != hipSuccess ||
                        hipMemset(d_lab, 0, (size_t)lchunk * lstrips * lab_elems * 8) != hipSuccess ||
                        hipMemset(d_lprog, 0x7f, (size_t)lchunk * lstrips * 4) != hipSuccess ||
                        hipMemset(d_lready, 0, (size_t)lchunk * 4) != hipSuccess ||
                        hipMemset(d_lnext, 0, (size_t)lchunk * 4) != hipSuccess))
                rc = RODV_PHMM_ERR_DEVICE;
        } else {
            if (!rc) rc = dev_alloc(&d_lck, (size_t)lgrid * lwaves * lck_elems);
            if (!rc) rc = dev_alloc(&d_lbnd, (size_t)lgrid * lbnd_stride);
            if (!rc) rc = dev_alloc(&d_lab, (size_t)lgrid * (lwaves + 1) * lab_elems);
            if (!rc && (hipMemset(d_lbnd, 0, (size_t)lgrid * lbnd_stride * 8) != hipSuccess ||
                        hipMemset(d_lab, 0, (size_t)lgrid * (lwaves + 1) * lab_elems * 8) != hipSuccess))
                rc = RODV_PHMM_ERR_DEVICE;
        }
    }
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!rc && (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess))
        rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && launch_reset(d_cnt, 3 * (uint32_t)nbk + 3, st) != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && hipEventRecord(ev0, st) != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    for (size_t bi = 0; bi < nbk && !rc; bi++) {
        const int L = shapes[bi].first, mr = shapes[bi].second, G = 64 / L;
        const uint32_t cnt = (uint32_t)((item_begin[bi + 1] - item_begin[bi]) / (size_t)G);
        if (!cnt) continue;
        AdjArgs a;
        std::memset(&a, 0, sizeof(a));
        a.hapc = p->d_hapc;
        a.hapc_off = p->d_hapc_off;
        a.hap_len = p->d_hap_len;
        a.initd_f = p->d_initd_f;
        a.initd_d = p->d_initd_d;
        a.rcode = p->d_rcode;
        a.q = p->d_q;
        a.read_off = p->d_read_off;
        a.read_len = p->d_read_len;
        a.pair_hap = p->d_pair_hap;
        a.pair_read = p->d_pair_read;
        a.items = d_items + item_begin[bi];
        a.n_items = cnt;
        a.xf = d_xf;
        a.yf = d_yf;
        a.xd = d_xd;
        a.yd = d_yd;
        a.wq = d_wq;
        a.qbin = d_qbin;
        a.delta = delta;
        a.gamma = gamma;
        a.mm_d = 1.0 - 2.0 * delta;
        a.g_d = 1.0 - gamma;
        a.mm_f = (float)a.mm_d;
        a.g_f = (float)a.g_d;
        a.dl_f = (float)delta;
        a.gm_f = (float)gamma;
        a.out_log10 = d_out;
        a.out_grad = d_grad;
        a.flags = d_flags;
        a.rescue_list = d_resc + pair_begin[bi];
        a.rescue_count = d_cnt + 3 * bi;
        a.ck = d_ck;
        a.ck_elems = ck_elems;
        a.lds_stride = stride;
        const uint64_t nv = 3ull * mr + 3;
        for (int pass = 0; pass < 2 && !rc; pass++) {
            const size_t esz = pass == 0 ? sizeof(float) : sizeof(double);
            // level-2 windows live in LDS: shrink the checkpoint stride if needed
            int K1 = ckpt_stride;
            // (one state per 8 steps of a segment)
            const size_t codes = (size_t)G * stride;
            while (K1 > 4 && codes + (size_t)((K1 + 7) / 8) * nv * 64 * esz > (size_t)maxlds) K1 /= 2;
            K1 = K1 / 4 * 4;
            a.ckpt = K1;
            a.counter = d_cnt + 3 * bi + 1 + pass;
            a.n_items_dev = pass == 0 ? nullptr : a.rescue_count;
            const size_t lds = codes + (size_t)((K1 + 7) / 8) * nv * 64 * esz;
            // the FP64 instance holds one wave per SIMD (up to 512
            // registers): four workgroups per CU are all that run at once,
            // and the pass over an empty list costs a dispatch of the grid
            const int g = pass == 0 ? grid : std::min(grid, ctx->num_cus * 4);
            if (launch_adj(adj2_kernel_ptr(L, mr, pass == 1), g, lds, st, a) != hipSuccess)
                rc = RODV_PHMM_ERR_DEVICE;
        }
    }
    if (!rc && !longs.empty()) {
        AdjArgs a;
        std::memset(&a, 0, sizeof(a));
        a.hapc = p->d_hapc;
        a.hapc_off = p->d_hapc_off;
        a.hap_len = p->d_hap_len;
        a.initd_f = p->d_initd_f;
        a.initd_d = p->d_initd_d;
        a.rcode = p->d_rcode;
        a.q = p->d_q;
        a.read_off = p->d_read_off;
        a.read_len = p->d_read_len;
        a.pair_hap = p->d_pair_hap;
        a.pair_read = p->d_pair_read;
        a.items = d_litems;
        a.n_items = (uint32_t)longs.size();
        a.counter = d_cnt + 3 * nbk;
        a.xf = d_xf;
        a.yf = d_yf;
        a.xd = d_xd;
        a.yd = d_yd;
        a.wq = d_wq;
        a.qbin = d_qbin;
        a.delta = delta;
        a.gamma = gamma;
        a.mm_d = 1.0 - 2.0 * delta;
        a.g_d = 1.0 - gamma;
        a.out_log10 = d_out;
        a.out_grad = d_grad;
        a.flags = d_flags;
        a.ck = d_lck;
        a.ck_elems = lck_elems;
        a.lds_stride = lstride;
        a.ckpt = lK1;
        a.lbnd = d_lbnd;
        a.lbnd_stride = lbnd_stride;
        a.lrow = lrow;
        a.lab = d_lab;
        a.lab_elems = lab_elems;
        a.spin_limit = spin_limit_hook();
        a.fault = d_cnt + 3 * nbk + 1;
        if (lsplit) {
            a.lsh = d_lsh;
            a.lsum = d_lsum;
            a.lexp = d_lexp;
            a.lprog = d_lprog;
            a.lpart = d_lpart;
            a.bcounter = d_cnt + 3 * nbk + 2;
            a.lstrips = (uint32_t)lstrips;
            for (uint64_t c0 = 0, ci = 0; c0 < nlong && !rc; c0 += lchunk, ci++) {
                const uint64_t cn = std::min<uint64_t>(lchunk, nlong - c0);
                a.items = d_litems + c0;
                a.n_items = (uint32_t)cn;
                a.bitems = d_bitems + bbegin[ci];
                a.n_bitems = bbegin[ci + 1] - bbegin[ci];
                a.lnstr = d_lnstr + c0;
                a.lready = d_lready;
                a.lnext = d_lnext;
                a.lck1 = d_lck1;
                a.lck1_off = d_lck1_off ? d_lck1_off + c0 : nullptr;
                // a later chunk starts as the first: work counters from 0,
                // progress counts above every step, rows zero
                if (ci && (launch_reset(a.counter, 1, st) != hipSuccess ||
                           launch_reset(a.bcounter, 1, st) != hipSuccess ||
                           hipMemsetAsync(d_lbnd, 0, (size_t)cn * lbnd_stride * 8, st) != hipSuccess ||
                           hipMemsetAsync(d_lab, 0, (size_t)cn * lstrips * lab_elems * 8, st) != hipSuccess ||
                           hipMemsetAsync(d_lprog, 0x7f, (size_t)cn * lstrips * 4, st) != hipSuccess ||
                           hipMemsetAsync(d_lready, 0, (size_t)cn * 4, st) != hipSuccess ||
                           hipMemsetAsync(d_lnext, 0, (size_t)cn * 4, st) != hipSuccess)) {
                    rc = RODV_PHMM_ERR_DEVICE;
                    break;
                }
                const int cgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(cn, (uint64_t)ctx->num_cus * 4));
                const hipError_t le =
                    lfused ? launch_adj_long_fused(abgrid, cgrid, adj_longab_lds(lstride, lK1), st, a)
                           : launch_adj_long_split((int)std::min<uint64_t>(cn, (uint64_t)ctx->num_cus), bgrid,
                                                   cgrid, long_lds(lK1), st, a);
                if (le != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
            }
        } else if (launch_adj_long(lwaves, lgrid, long_lds(lK1), st, a) != hipSuccess) {
            rc = RODV_PHMM_ERR_DEVICE;
        }
    }
    std::vector<double> out(n), g(10 * n);
    if (!rc && hipEventRecord(ev1, st) != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && kernel_ms) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        *kernel_ms = ms;
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (!rc && hipMemcpy(out.data(), d_out, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && hipMemcpy(g.data(), d_grad, 10 * n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = RODV_PHMM_ERR_DEVICE;
    uint32_t lfaults = 0;  // hand-over waits of the long-read kernel that ran out
    if (!rc && hipMemcpy(&lfaults, d_cnt + 3 * nbk + 1, 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = RODV_PHMM_ERR_DEVICE;
    if (!rc && lfaults) rc = RODV_PHMM_ERR_DEVICE;
    if (!rc) {
        for (int k = 0; k < 10; k++) grad[k] = 0.0;
        for (uint64_t k = 0; k < n; k++) {
            const double w = weights ? weights[k] : 1.0;
            if (w == 0.0) continue;
            for (int c = 0; c < 10; c++) grad[c] += w * g[10 * k + c];
        }
        if (out_log10) std::memcpy(out_log10, out.data(), n * sizeof(double));
    }
    void *bufs[] = {d_xf, d_yf, d_xd, d_yd, d_wq, d_qbin, d_items, d_out, d_grad, d_flags, d_resc, d_cnt,
                    d_ck, d_litems, d_lck, d_lbnd, d_lab, d_lsh, d_lexp, d_lsum, d_lprog, d_lpart, d_bitems,
                    d_lnstr, d_lready, d_lnext, d_lck1, d_lck1_off};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    rodv_phmm_plan_destroy(p);
    return rc;
}

int rodv_phmm_adjoint(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                      const rodv_phmm_params *params, const double *weights, int ckpt_stride,
                      double *out_log10, double *grad, double *kernel_ms) {
    return guarded([&] { return adjoint_impl(ctx, batch, params, weights, ckpt_stride, out_log10, grad, kernel_ms); });
}

}  // extern "C"
