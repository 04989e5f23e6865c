// phmm_kernels.hip -- gfx950 kernels of the PairHMM forward.
//
// The recurrence is roDV's (pairhmm_impl.cpp:266-289) on the DP interior
// i = 1..R (read rows), j = 1..H (haplotype columns):
//   M[i][j] = prior(i,j) * (mm_i*M[i-1][j-1] + gapm_i*(I[i-1][j-1] + D[i-1][j-1]))
//   I[i][j] = mx_i*M[i-1][j] + xx_i*I[i-1][j]
//   D[i][j] = my_i*M[i][j-1] + yy_i*D[i][j-1]
// with row 0 = (0, 0, K/H) and column 0 = 0; S = sum_j M[R][j] + sum_j I[R][j].
// The FP32 pass reproduces the reference's rounding exactly: the FMA split
// GCC gives the intra kernel (M = fma(D',g, fma(I',g, M'*mm)) * prior,
// I = fma(I^,xx, M^*mx), D = fma(D<,yy, M<*my)), the inter kernel's own M
// order when asked for, the same tables and the same left-to-right sums.
// Compiled with -ffp-contract=off so only the explicit fmas fuse.
//
// Mapping (MI355X, wave64).  A pair owns a group of L lanes (L = 16: four
// pairs per wave; 32: two; 64: one).  Lane l holds MR consecutive read rows
// in registers and sweeps the haplotype one column per step, lagging lane
// l-1 by one column, so a step is an anti-diagonal of lanes.  Rows inside a
// lane pass values in registers; the row above a lane's first row is the
// previous lane's last row, brought over with ds_bpermute (LaneRot; DPP would
// stop two waves from co-issuing).  The rotation makes lane 0 read lane L-1's
// last slot, which is never a read row: it is an identity row that
// reproduces the row-0 boundary (M = I = 0, D = K/H) for free, or -- for
// reads longer than one strip of (L-1)*MR rows -- the feeder that injects the
// previous strip's last row from HBM.  Read rows are shifted down by `off`
// identity rows so that row R always sits in the last register slot of its
// lane.
// Two FP32 variants: fwd_tab_kernel reads each row's emission prior from a
// per-pair LDS table indexed by the haplotype base (N-free haplotypes, one
// strip); fwd_kernel stages one-hot haplotype codes and selects the prior
// per cell with v_bfe_i32 + a bit-select (haplotypes with N, multi-strip
// reads, and the FP64 rescue).  No MFMA: the recurrence is not a contraction.
// fwd_tab_kernel needs no identity lane: all L lanes carry read rows, and
// the row-0 boundary is an identity slot at the top of lane 0 (off >= 1
// when R % MR != 0) or lane L-1 left without read rows (R % MR == 0), so a
// strip holds any R <= L*MR - 1.  An identity slot maps every finite input
// to (0, 0, D unchanged), so what lane L-1 rotates into lane 0 is harmless.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phmm_device.h"
#include "phmm_internal.h"

namespace rodv {
namespace {

using namespace dev;
constexpr int kU = 8;  // steps per unrolled block (even: the register rotation)


// The raw bytes of one read row: qualities (masked as pairhmm_impl.cpp:
// 89-92 masks them) and the base code.
struct RowRaw {
    uint32_t qi, qd, qc, qq, code;
};

__device__ __forceinline__ RowRaw load_raw(const FwdArgs &a, uint64_t idx) {
    RowRaw w;
    w.qi = a.ins[idx] & 127u;
    w.qd = a.del[idx] & 127u;
    w.qc = a.gcp[idx] & 127u;
    w.qq = a.q[idx] & 127u;
    w.code = a.rcode[idx];
    return w;
}

// The row's transition and emission priors from its raw bytes; identity
// rows (real == false) select constants.
template <typename T, int MR, bool INTER>
__device__ __forceinline__ void row_priors(const FwdArgs &a, Priors<T, MR> &P, int s, bool real,
                                           const RowRaw &w) {
    const DevTables *t = a.tab;
    const uint32_t qi = w.qi, qd = w.qd, qc = w.qc, qq = w.qq;
    uint32_t code = w.code;
    uint32_t lo = qi <= qd ? qi : qd, hi = qi <= qd ? qd : qi;
    T mm = Tab<T>::mm(t)[((hi * (hi + 1)) >> 1) + lo];
    T g = Tab<T>::omp(t)[qc];
    T mx = Tab<T>::ph2pr(t)[qi];
    T xx = Tab<T>::ph2pr(t)[qc];
    T my = Tab<T>::ph2pr(t)[qd];
    T X = Tab<T>::omp(t)[qq];
    T Y = INTER ? Tab<T>::pt(t)[qq] : Tab<T>::p3(t)[qq];
    const bool isN = code == kReadN;  // N in the read matches every hap base
    Y = isN ? X : Y;
    code = isN ? 0u : code;
    // identity row: M = 0, I = 0, D carried unchanged
    P.mm[s] = real ? mm : T(0);
    P.g[s] = real ? g : T(0);
    P.mx[s] = real ? mx : T(0);
    P.xx[s] = real ? xx : T(0);
    P.yy[s] = real ? xx : T(1);
    P.my[s] = real ? my : T(0);
    P.X[s] = real ? X : T(0);
    P.Y[s] = real ? Y : T(0);
    P.rpos[s] = real ? code : 0u;
}

// branch-free: the address is always valid (idx points at the read's first
// base for identity rows) and identity rows select constants
template <typename T, int MR, bool INTER>
__device__ __forceinline__ void load_row(const FwdArgs &a, Priors<T, MR> &P, int s, bool real,
                                         uint64_t idx) {
    row_priors<T, MR, INTER>(a, P, s, real, load_raw(a, idx));
}

// Orders one wave's LDS accesses across its lanes without a workgroup
// barrier: the compiler moves no memory access across it, and a wave's LDS
// operations execute in issue order, so waiting for its own ones suffices.
// A __syncthreads() would also wait for every global load and store in
// flight, which the table kernel keeps in flight across items on purpose.
// Valid only for one-wave workgroups.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// min / max of a per-group value over the G groups of the wave (lanes 0,
// L, 2L, ... hold each group's copy)
__device__ __forceinline__ int wave_min4(int v, int G) {
    if (G == 8) {
        int m = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
        for (int k = 1; k < 8; k++) m = min(m, __builtin_amdgcn_readlane(v, 8 * k));
        return m;
    }
    if (G == 1) return __builtin_amdgcn_readfirstlane(v);
    if (G == 2) return min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
    int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}
__device__ __forceinline__ int wave_max4(int v, int G) {
    if (G == 8) {
        int m = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
        for (int k = 1; k < 8; k++) m = max(m, __builtin_amdgcn_readlane(v, 8 * k));
        return m;
    }
    if (G == 1) return __builtin_amdgcn_readfirstlane(v);
    if (G == 2) return max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
    int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}

// ------------------------------------------------------------ the kernel

template <typename T, int L, int MR, bool INTER, bool RESCUE>
__device__ void process_item(const FwdArgs &a, uint8_t *lds, int p) {
    constexpr int G = 64 / L;
    // read rows per strip: lanes 0..L-2; lane L-1 holds identity rows and,
    // in its last slot, the row above the strip (row 0, or the feeder)
    constexpr int C = (L - 1) * MR;
    const int lane = threadIdx.x;
    const int g = lane / L;
    const int l = lane % L;
    const bool valid = p >= 0;

    uint32_t h = 0, r = 0;
    int R = 0, H = 0;
    uint64_t ro = 0;
    if (valid) {
        h = a.pair_hap[p];
        r = a.pair_read[p];
        R = (int)a.read_len[r];
        H = (int)a.hap_len[h];
        ro = a.read_off[r];
    }
    // shift the read down by `off` < MR identity rows (all in lane 0 of the
    // first strip) so that row R lands in the last slot of its lane
    const int off = valid ? (MR - R % MR) % MR : 0;
    const int vR = R - 1 + off;
    const int nstr = valid ? vR / C + 1 : 1;
    const int lR = valid ? (vR % C) / MR : 0;
    const T initd = valid ? Tab<T>::initd(a, h) : T(0);

    // --- stage one-hot hap codes: column j of group g at lds[g*stride + L-1+j]
    __syncthreads();
    for (uint32_t k = lane * 16; k < (uint32_t)G * a.lds_stride; k += 64 * 16)
        *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    if (valid) {
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[h]);
        const int nch = (H + 15) >> 4;
        for (int c = l; c < nch; c += L)
            *reinterpret_cast<uint4 *>(lds + g * a.lds_stride + L + 16 * c) = src[c];
    }
    __syncthreads();
    const uint8_t *cp = lds + g * a.lds_stride + (L - 1) - l;

    const int key = valid ? H + lR : 0x7fffffff;
    const int keymax = valid ? H + lR : -1;
    const int Dmain_last = wave_min4(key, G);
    const int Dend_last = wave_max4(keymax, G);
    const int nstr_w = wave_max4(nstr, G);

    T sm = T(0), si = T(0);
    for (int sidx = 0; sidx < nstr_w; sidx++) {
        const bool last = sidx == nstr_w - 1;
        Priors<T, MR> P;
        State<T, MR> S;
#pragma unroll
        for (int s = 0; s < MR; s++) {
            const int pos = l * MR + s;
            const int i = sidx * C + pos - off + 1;  // 1-based read row
            const bool real = valid && l < L - 1 && i >= 1 && i <= R;
            load_row<T, MR, INTER>(a, P, s, real, ro + (uint64_t)(real ? i - 1 : 0));
            S.M[s] = T(0);
            S.I[s] = T(0);
            S.D[s] = real ? T(0) : initd;
        }
        // step's carry: the row above before the first column and that
        // row's yy and my.  Lane 0's row above is lane L-1's identity slot in
        // the first strip, the previous strip's last read row in the others
        // (column 0 of a read row is zero; its yy and my from the read).
        const LaneRot<L> rot;
        S.Mp = T(0);
        S.Ip = T(0);
        S.Dp = rot(S.D[MR - 1]);
        T yyA = rot(P.yy[MR - 1]), myA = rot(P.my[MR - 1]);
        if (sidx > 0 && l == 0) {
            const uint64_t ip = ro + (uint64_t)(sidx * C - off - 1);  // 0-based row above
            yyA = Tab<T>::ph2pr(a.tab)[a.gcp[ip] & 127u];
            myA = Tab<T>::ph2pr(a.tab)[a.del[ip] & 127u];
            S.Dp = T(0);
        }

        // boundary rows between strips (L = 64 kernels only)
        T *bnd_in = nullptr, *bnd_out = nullptr;
        bool inject = false, store = false;
        const bool feeder = l == L - 1;
        if constexpr (L == 64) {
            if (nstr_w > 1) {
                T *base = reinterpret_cast<T *>(a.bnd) + (uint64_t)blockIdx.x * 2 * a.bnd_elems;
                bnd_in = base + (uint64_t)((sidx + 1) & 1) * a.bnd_elems;
                bnd_out = base + (uint64_t)(sidx & 1) * a.bnd_elems;
                inject = sidx > 0;
                store = !last;
            }
            if (inject) {
                __threadfence();
                // feeder starts at column 1 of the previous strip's last row;
                // its column-0 values are zero
                T bm = ld_agent(bnd_in + 3 * (1 + L)), bi = ld_agent(bnd_in + 3 * (1 + L) + 1),
                  bd = ld_agent(bnd_in + 3 * (1 + L) + 2);
                if (feeder) {
                    S.M[MR - 1] = bm;
                    S.I[MR - 1] = bi;
                    S.D[MR - 1] = bd;
                }
            }
        }
        // a strip that is not the last one runs until lane L-2 has passed
        // column H + L + 1, so the boundary row is written for every column
        // the next strip's feeder reads up to H; its sums are discarded
        const int Dend = last ? Dend_last : H + 2 * L - 1;
        const int Dmain = last ? Dmain_last : Dend;
        sm = T(0);
        si = T(0);

        auto after_step = [&](int d) {
            if constexpr (L == 64) {
                if (store && l == L - 2) {
                    // last read row of this strip (last slot of lane L-2), at
                    // column d-(L-2); column j is stored at index j + L
                    T *o = bnd_out + 3 * (d + 2);
                    st_agent(o, S.M[MR - 1]);
                    st_agent(o + 1, S.I[MR - 1]);
                    st_agent(o + 2, S.D[MR - 1]);
                }
                if (inject) {
                    const T *in = bnd_in + 3 * (d + 1 + L);
                    T bm = ld_agent(in), bi = ld_agent(in + 1), bd = ld_agent(in + 2);
                    if (feeder) {
                        S.M[MR - 1] = bm;
                        S.I[MR - 1] = bi;
                        S.D[MR - 1] = bd;
                    }
                }
            }
        };

        int d = 1;
        // main loop: every lane's sum slot is at a column <= H
        for (; d + kU - 1 <= Dmain; d += kU) {
            uint32_t hc[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) hc[u] = cp[d + u];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                step<T, L, MR, INTER>(S, P, yyA, myA, hc[u], rot);
                sm += S.M[MR - 1];
                si += S.I[MR - 1];
                after_step(d + u);
            }
        }
        // remaining steps: sums only while the lane's column is <= H
        for (; d <= Dend; d++) {
            step<T, L, MR, INTER>(S, P, yyA, myA, cp[d], rot);
            if (last && d - l <= H) {
                sm += S.M[MR - 1];
                si += S.I[MR - 1];
            }
            after_step(d);
        }
    }

    if (valid && l == lR) {
        T Ssum = sm + si;  // pairhmm_impl.cpp:434-437
        if constexpr (!RESCUE) {
            a.raw32[p] = (float)Ssum;
            uint8_t fl = 0;
            if ((float)Ssum < 1e-28f) {  // MIN_ACCEPTED, pairhmm_api.cpp:4,17
                fl = 0x01;
                uint32_t k = atomicAdd(a.rescue_count, 1u);
                a.rescue_list[k] = (uint32_t)p;
            }
            a.flags[p] = fl;
        } else {
            a.raw64[p] = (double)Ssum;
        }
    }
}

// ------------------------------------------------- table-prior variant
// Same lane/row mapping as process_item, FP32, one strip, haplotypes without
// N.  The emission prior of every (hap base, lane, row slot) is written once
// per item into an LDS table; a step reads the lane's MR priors for its
// column with ds_read_b128 instead of selecting them per cell, which takes
// the select (2 VALU per cell) out of the hot loop.  The haplotype is staged
// as u16 byte offsets of its base's table block.
template <int MR>
struct TabLayout {
    static constexpr int MRT = (MR + 3) / 4 * 4;  // slots padded to 16 bytes
};

// One haplotype code (the base's block 0..3 of the prior table): one
// ds_read_u8.  The library is built without unaligned-access-mode, so the
// compiler does not merge the eight byte loads of a block into wider reads
// that would be unaligned for most lanes, and it knows the load zero-extends
// (an atomic load would cost an extra v_and per step).
__device__ __forceinline__ uint32_t ld_code(const uint8_t *p) { return *p; }

// The prior table of a group: [4 hap bases][MRT/4 chunks][L lanes][4 rows]
// of floats, so the 16 lanes of an LDS pass read 16 consecutive 16-byte
// chunks (all 64 banks once) whatever MRT is; a lane-major layout put lanes
// MRT*4 bytes apart, which is 2- and 4-way bank conflicts at MRT 8 and 16.
// The MR priors of one lane for one column: MRT/4 ds_read_b128 from the
// lane's chunk of the column's block.
template <int L, int MR>
__device__ __forceinline__ void load_priors(float (&pr)[MR], const uint8_t *p) {
    constexpr int MRT = TabLayout<MR>::MRT;
    const float4 *tp = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int q = 0; q < MRT / 4; q++) {
        const float4 v = tp[q * L];
        if (4 * q + 0 < MR) pr[4 * q + 0] = v.x;
        if (4 * q + 1 < MR) pr[4 * q + 1] = v.y;
        if (4 * q + 2 < MR) pr[4 * q + 2] = v.z;
        if (4 * q + 3 < MR) pr[4 * q + 3] = v.w;
    }
}

// The row above a lane's first row comes from the previous lane's last slot:
// `up` (this step's column, rotated from the slot's value after the previous
// step) and `dg` (the column before).  dg of step t is up of step t-1 for M
// and I, so each step rotates two values: the new M and I of the last slot
// (up of the next step).  D runs along the row, so the lane recomputes the
// slot's D itself -- D(j) = fma(D(j-1), yy, M(j-1) * my), the slot's own
// instruction on the slot's own inputs (dg of the previous step), hence the
// same bits -- from the slot's yy and my, rotated once per item.
struct Carry {
    float dM, dI, dD;  // dg of this step
    float uM, uI;      // up of this step
    float yyA, myA;    // yy and my of the row above the lane's first row
};

// One step of one lane (all MR rows at the next column), ordered for the LDS
// unit: the next column's priors are read first, then M and D of every slot
// (which need only the previous column), then the I chain down the slots
// (which needs `up`, rotated at the end of the previous step), then the
// rotations for the next step.  LDS results return in issue order, so each
// wait covers only what the next phase needs.
// SH (shared products): the read's deletion prior of each row equals the
// insertion prior of the row below (my_i == mx_{i+1}, as when a read's
// insertion and deletion qualities are the same constant).  Then the
// product M(i,j) * mx_{i+1} that I(i+1,j) takes in this step is also the
// M(i,j) * my_i that D(i,j+1) takes in the next one, the same bits, so it is
// computed once and kept in Pm: 7 FP32 instructions per cell instead of 8.
// Pm[s] holds S.M[s] * my[s] (zero where M is zero: identity slots, and
// every slot after a reset); the lane's last slot computes its own, its row
// below being the next lane's first.
template <int L, int MR, bool INTER, bool SH>
__device__ __forceinline__ void step_tab(State<float, MR> &S, Carry &C, float (&Pm)[MR],
                                         const Priors<float, MR> &P, const float (&pr)[MR],
                                         const LaneRot<L> &rot, float (&pb)[MR], const uint8_t *pnext) {
    // keep the previous step's M and I rotations ahead of this step's LDS
    // reads: the I chain below then waits for them alone
    __builtin_amdgcn_sched_barrier(0);
    load_priors<L, MR>(pb, pnext);
    __builtin_amdgcn_sched_barrier(0);
    float nM[MR], nD[MR];
    float dgM = C.dM, dgI = C.dI, dgD = C.dD;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        float t;
        if (INTER) {  // pairhmm_inter.cpp:279-283 as GCC contracts it
            t = dgI * P.g[s];
            t = fmaT(dgM, P.mm[s], t);
            t = fmaT(dgD, P.g[s], t);
        } else {  // pairhmm_impl.cpp:266-275
            t = dgM * P.mm[s];
            t = fmaT(dgI, P.g[s], t);
            t = fmaT(dgD, P.g[s], t);
        }
        nM[s] = t * pr[s];
        const float v = SH ? Pm[s] : S.M[s] * P.my[s];
        nD[s] = fmaT(S.D[s], P.yy[s], v);
        dgM = S.M[s];
        dgI = S.I[s];
        dgD = S.D[s];
    }
    float upM = C.uM, upI = C.uI;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const float u = upM * P.mx[s];
        if (SH && s > 0) Pm[s - 1] = u;  // nM[s-1] * mx[s] == nM[s-1] * my[s-1]
        const float nI = fmaT(upI, P.xx[s], u);
        upM = nM[s];
        upI = nI;
        S.I[s] = nI;
        S.M[s] = nM[s];
        S.D[s] = nD[s];
    }
    if (SH) Pm[MR - 1] = nM[MR - 1] * P.my[MR - 1];
    C.dD = fmaT(C.dD, C.yyA, C.dM * C.myA);  // the slot above, one column on
    C.dM = C.uM;
    C.dI = C.uI;
    __builtin_amdgcn_sched_barrier(0);
    C.uM = rot(S.M[MR - 1]);
    C.uI = rot(S.I[MR - 1]);
}

// One item of the table kernel: G pairs, one per group of L lanes.  `d` is
// the group's descriptor; the return value is the index of the wave's next
// item, fetched from the work counter once the set-up is done so that the
// atomic's round trip hides behind the sweep.
template <int L, int MR, bool INTER, bool SH>
__device__ uint32_t process_item_tab(const FwdArgs &a, uint8_t *lds, const ItemDesc &d) {
    constexpr int G = 64 / L;
    constexpr int MRT = TabLayout<MR>::MRT;
    constexpr uint32_t kBlk = (uint32_t)L * MRT * 4;  // bytes per hap base
    const int lane = threadIdx.x;
    const int g = lane / L;
    const int l = lane % L;
    const int p = d.p;
    const bool valid = p >= 0;
    const int R = (int)d.R, H = (int)d.H;  // 0 for an empty slot
    const uint64_t ro = d.ro;
    const float initd = d.initd;
    const int off = valid ? (MR - R % MR) % MR : 0;
    const int lR = valid ? (R - 1 + off) / MR : 0;

    // LDS per group: codes (one byte per column, lds_stride bytes) then the
    // prior table
    const uint32_t gstride = a.lds_stride + 4 * kBlk;
    uint8_t *gb = lds + g * gstride;
    uint8_t *codes = gb;
    uint8_t *tab = gb + a.lds_stride;

    // codes outside the haplotype are zero: each group clears its code area
    // (the previous item's LDS reads have all returned: their values were used)
    for (uint32_t k = (uint32_t)l * 16; k < a.lds_stride; k += L * 16)
        *reinterpret_cast<uint4 *>(gb + k) = make_uint4(0, 0, 0, 0);
    wave_lds_sync();

    // the rows' raw bytes first, so that they and the haplotype codes are
    // in flight together
    RowRaw raw[MR];
    bool realv[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = l * MR + s - off + 1;
        realv[s] = valid && i >= 1 && i <= R;  // lane L-1 included
        raw[s] = load_raw(a, ro + (uint64_t)(realv[s] ? i - 1 : 0));
    }
    if (valid) {
        // column j (1-based) at codes[L-1+j]: one-hot 1,2,4,8 -> block 0..3,
        // (b >> 1) - (b >> 3) on four bytes at once; 16 columns per chunk
        // from the 16-byte aligned codes, two chunks per lane and pass, both
        // loads issued before either is used (without a second chunk the
        // first is converted and written twice)
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + d.hoff);
        const int nch = (H + 15) >> 4;
        for (int c = l; c < nch; c += 2 * L) {
            const int c2 = c + L < nch ? c + L : c;
            const uint4 v0 = src[c];
            const uint4 v1 = src[c2];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int h2 = 0; h2 < 2; h2++) {
                const uint4 v = h2 ? v1 : v0;
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
                uint32_t o[4];
#pragma unroll
                for (int k = 0; k < 4; k++)  // 1->0 2->1 4->2 8->3 per byte
                    o[k] = ((w[k] >> 1) & 0x7f7f7f7fu) - ((w[k] >> 3) & 0x1f1f1f1fu);
                *reinterpret_cast<uint4 *>(codes + L + 16 * (h2 ? c2 : c)) = make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
    }

    Priors<float, MR> P;
    float X[MR], Y[MR];
    uint32_t rc[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        row_priors<float, MR, INTER>(a, P, s, realv[s], raw[s]);
        X[s] = P.X[s];
        Y[s] = P.Y[s];
        rc[s] = P.rpos[s];
    }
    // table: block c, chunk s/4, lane l, row s%4 -> prior of row s when the
    // hap base is c
    float *tl = reinterpret_cast<float *>(tab + (uint32_t)l * 16);
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int s = 0; s < MR; s++)
            tl[c * (kBlk / 4) + (s >> 2) * (L * 4) + (s & 3)] = (rc[s] == (uint32_t)c) ? X[s] : Y[s];
    }
    wave_lds_sync();
    uint32_t next = 0;
    if (lane == 0) next = atomicAdd(a.counter, 1u);

    State<float, MR> S;
    float Pm[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        S.M[s] = 0.0f;
        S.I[s] = 0.0f;
        S.D[s] = realv[s] ? 0.0f : initd;  // identity rows carry K/H
        Pm[s] = 0.0f;
    }
    const LaneRot<L> rot;
    Carry C;
    C.dM = 0.0f;
    C.dI = 0.0f;
    C.dD = rot(S.D[MR - 1]);
    C.uM = 0.0f;  // rotations of the all-zero M and I of the first column
    C.uI = 0.0f;
    C.yyA = rot(P.yy[MR - 1]);
    C.myA = rot(P.my[MR - 1]);

    const int key = valid ? H + lR : 0x7fffffff;
    const int keymax = valid ? H + lR : -1;
    const int Dmain = wave_min4(key, G);
    const int Dend = wave_max4(keymax, G);
    const uint8_t *cl = codes + (L - 1) - l;
    const uint8_t *tlane = tab + (uint32_t)l * 16;
    float sm = 0.0f, si = 0.0f;
    // Software pipeline: the priors of column d+1 are read while step d
    // computes, the code of column d+2 one step earlier still.  Codes are read
    // one per step with byte loads: lanes start at different columns, so a
    // wider load would be unaligned for most of them.  Codes past the
    // haplotype are zero (cleared per item), so every prefetch address stays
    // inside the group's table.
    int dd = 1;
    float pa[MR];
    load_priors<L, MR>(pa, tlane + ld_code(cl + dd) * kBlk);
    uint32_t c1 = ld_code(cl + dd + 1);
    for (; dd + kU - 1 <= Dmain; dd += kU) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const uint32_t c2 = ld_code(cl + dd + u + 2);
            float pb[MR];
            step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
#pragma unroll
            for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            c1 = c2;
        }
    }
    for (; dd <= Dend; dd++) {
        const uint32_t c2 = ld_code(cl + dd + 2);
        float pb[MR];
        step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
        if (dd - l <= H) {
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
        }
#pragma unroll
        for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
        c1 = c2;
    }
    if (valid && l == lR) {
        const float Ssum = sm + si;  // pairhmm_impl.cpp:434-437
        a.raw32[p] = Ssum;
        uint8_t fl = 0;
        if (Ssum < 1e-28f) {  // MIN_ACCEPTED, pairhmm_api.cpp:4,17
            fl = 0x01;
            const uint32_t k = atomicAdd(a.rescue_count, 1u);
            a.rescue_list[k] = (uint32_t)p;
        }
        a.flags[p] = fl;
    }
    // every lane has finished its LDS reads (their values were used above)
    // before the next item clears and rewrites the areas
    wave_lds_sync();
    return __builtin_amdgcn_readfirstlane(next);
}

// Persistent: each wave takes items (G pairs) from the work counter, one
// item ahead (process_item_tab fetches the next index), and reads each
// item's descriptors with one load.  One wave per workgroup (the LDS
// ordering in process_item_tab relies on it).
template <int L, int MR, bool INTER, bool SH>
__global__ __launch_bounds__(64) void fwd_tab_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int G = 64 / L;
    const uint32_t n = a.n_items;
    uint32_t it = 0;
    if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    for (uint32_t iter = 0; iter < n && it < n; iter++) {
        const ItemDesc d = a.desc[(uint64_t)it * G + threadIdx.x / L];
        it = process_item_tab<L, MR, INTER, SH>(a, lds, d);
    }
}

// ------------------------------------------------- chained table items
// A chained item is G reads of one region (one per group of L lanes, rows
// laid out as in process_item_tab) scored against the same K haplotypes of
// that region in turn.  The read's priors and LDS table are built once per
// item, and each lane runs from haplotype k's last column straight into
// haplotype k+1's first, where its state is reset to the column-0 boundary
// (the state an item starts from, with the new haplotype's K/H): the L-1
// fill and drain steps and the read's set-up are paid once per item instead
// of once per pair.  Lane l is at global column dd - l at step dd;
// haplotype k occupies global columns (S_k, S_k + H_k].  The groups share
// the haplotype codes, staged in two LDS buffers: haplotype k+1's while k is
// swept.  Two hand-overs never overlap when H_k >= L + 2 for k >= 1, which
// the plan guarantees.

// One haplotype's codes (block 0..3 per column, column j at dst[j - 1]),
// staged by the whole wave: 16 columns per lane and pass.
__device__ __forceinline__ void stage_codes(const FwdArgs &a, uint64_t hoff, int H, uint8_t *dst) {
    const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + hoff);
    const int nch = (H + 15) >> 4;
    for (int c = threadIdx.x; c < nch; c += 64) {
        const uint4 v = src[c];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; k++)  // one-hot 1,2,4,8 -> 0..3 per byte
            o[k] = ((w[k] >> 1) & 0x7f7f7f7fu) - ((w[k] >> 3) & 0x1f1f1f1fu);
        *reinterpret_cast<uint4 *>(dst + 16 * c) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// The FP32 sum of a pair and its rescue flag (pairhmm_api.cpp:17).
__device__ __forceinline__ void put_raw32(const FwdArgs &a, int p, float Ssum) {
    a.raw32[p] = Ssum;
    uint8_t fl = 0;
    if (Ssum < 1e-28f) {
        fl = 0x01;
        const uint32_t k = atomicAdd(a.rescue_count, 1u);
        a.rescue_list[k] = (uint32_t)p;
    }
    a.flags[p] = fl;
}

// Where the NS reads of a group sit in its L * MR slots (slot l * MR + s is
// row slot s of lane l).  NS = 1: the read's row R in the last slot of lane
// lo[0], after off = (MR - R % MR) % MR identity slots at the top (the table
// kernel's layout).  NS = 2: read B ends at the group's last slot, read A at
// the last slot of lane lo[0] = (L * MR - R_B - 1) / MR - 1, so that at least
// one identity slot lies between them; that slot is read B's row 0 (M = I =
// 0, D = K/H at every column, like the slots above a single read).  The
// planner also leaves the group's first slot an identity slot (lane L-1
// holds read B's last row, which the rotation brings into lane 0).  The
// lane holding a read's last row accumulates that pair's sum.
template <int L, int MR, int NS>
struct GroupRows {
    int st[NS], en[NS];  // first and last slot of each read
    int lo[NS];         // the lane of each read's last row
    int p0[NS];         // the read's pair for the chain's first haplotype (-1: none)
    uint64_t ro[NS];
    __device__ __forceinline__ GroupRows(const ItemDesc *d) {
        if constexpr (NS == 1) {
            const bool v = d[0].p >= 0;
            const int R = v ? (int)d[0].R : 0;
            const int off = v ? (MR - R % MR) % MR : 0;
            st[0] = off;
            en[0] = off + R - 1;
            lo[0] = v ? (R - 1 + off) / MR : 0;
            p0[0] = d[0].p;
            ro[0] = d[0].ro;
        } else {
            const bool vb = d[1].p >= 0;
            const int RB = vb ? (int)d[1].R : 0;
            st[1] = L * MR - RB;
            en[1] = L * MR - 1;
            lo[1] = L - 1;
            p0[1] = d[1].p;
            ro[1] = d[1].ro;
            const bool va = d[0].p >= 0;
            const int RA = va ? (int)d[0].R : 0;
            lo[0] = (L * MR - RB - 1) / MR - 1;
            en[0] = (lo[0] + 1) * MR - 1;
            st[0] = en[0] - RA + 1;
            p0[0] = d[0].p;
            ro[0] = d[0].ro;
        }
    }
    // the read of slot i (-1: identity slot)
    __device__ __forceinline__ int read_of(int i) const {
        int r = -1;
#pragma unroll
        for (int j = 0; j < NS; j++)
            if (p0[j] >= 0 && i >= st[j] && i <= en[j]) r = j;
        return r;
    }
};

template <int L, int MR, bool INTER, int NS, bool SH>
__device__ uint32_t process_chain(const FwdArgs &a, uint8_t *lds, uint32_t it) {
    constexpr int G = 64 / L;
    constexpr int MRT = TabLayout<MR>::MRT;
    constexpr uint32_t kBlk = (uint32_t)L * MRT * 4;
    constexpr int P0 = L < 16 ? 16 : L;  // buffer byte of column 1 (16-byte aligned)
    const int lane = threadIdx.x;
    const int g = lane / L;
    const int l = lane % L;
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(a.chain_off[it]);
    const uint32_t K = __builtin_amdgcn_readfirstlane(a.chain_off[it + 1]) - c0;
    ItemDesc dv[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) dv[j] = a.desc[((uint64_t)it * G + g) * NS + j];
    const GroupRows<L, MR, NS> gr(dv);
    const float initd0 = dv[0].initd;  // one haplotype for the group's reads
    // this lane's pair (for the chain's first haplotype) if it holds a read's
    // last row, else -1; the group's first and last such lanes
    int own = -1, lmin = 0x7fffffff, lmax = -1;
#pragma unroll
    for (int j = 0; j < NS; j++)
        if (gr.p0[j] >= 0) {
            if (l == gr.lo[j]) own = gr.p0[j];
            lmin = min(lmin, gr.lo[j]);
            lmax = max(lmax, gr.lo[j]);
        }
    uint8_t *tab = lds + (uint32_t)g * 4u * kBlk;
    uint8_t *buf = lds + (uint32_t)G * 4u * kBlk;  // two code buffers of lds_stride bytes

    RowRaw raw[MR];
    bool realv[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = l * MR + s;
        const int j = gr.read_of(i);
        realv[s] = j >= 0;
        // (selects, not an index: the arrays stay in registers)
        uint64_t ro = gr.ro[0];
        int st = gr.st[0];
#pragma unroll
        for (int q = 1; q < NS; q++)
            if (j == q) {
                ro = gr.ro[q];
                st = gr.st[q];
            }
        raw[s] = load_raw(a, ro + (uint64_t)(j >= 0 ? i - st : 0));
    }
    const int H0 = __builtin_amdgcn_readfirstlane((int)a.chain_hap[c0].H);
    const uint32_t hoff0_lo = __builtin_amdgcn_readfirstlane((uint32_t)a.chain_hap[c0].hoff);
    const uint32_t hoff0_hi = __builtin_amdgcn_readfirstlane((uint32_t)(a.chain_hap[c0].hoff >> 32));
    stage_codes(a, ((uint64_t)hoff0_hi << 32) | hoff0_lo, H0, buf + P0);

    Priors<float, MR> P;
    float X[MR], Y[MR];
    uint32_t rc[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        row_priors<float, MR, INTER>(a, P, s, realv[s], raw[s]);
        X[s] = P.X[s];
        Y[s] = P.Y[s];
        rc[s] = P.rpos[s];
    }
    float *tl = reinterpret_cast<float *>(tab + (uint32_t)l * 16);
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int s = 0; s < MR; s++)
            tl[c * (kBlk / 4) + (s >> 2) * (L * 4) + (s & 3)] = (rc[s] == (uint32_t)c) ? X[s] : Y[s];
    }
    wave_lds_sync();
    uint32_t next = 0;
    if (lane == 0) next = atomicAdd(a.counter, 1u);

    State<float, MR> S;
    float Pm[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        S.M[s] = 0.0f;
        S.I[s] = 0.0f;
        S.D[s] = realv[s] ? 0.0f : initd0;
        Pm[s] = 0.0f;
    }
    const LaneRot<L> rot;
    Carry C;
    C.dM = 0.0f;
    C.dI = 0.0f;
    C.dD = rot(S.D[MR - 1]);
    C.uM = 0.0f;
    C.uI = 0.0f;
    C.yyA = rot(P.yy[MR - 1]);
    C.myA = rot(P.my[MR - 1]);

    const uint8_t *cl = buf + (P0 - 1) - l;  // code of step x: cl[x]
    const uint8_t *tlane = tab + (uint32_t)l * 16;
    int pd = 0;  // pair offset of the current haplotype: own + pd is its pair
    float sm = 0.0f, si = 0.0f;
    int dd = 1;
    int Sk = 0;  // columns of the haplotypes before the current one
    int H = H0;
    float pa[MR];
    load_priors<L, MR>(pa, tlane + ld_code(cl + dd) * kBlk);
    uint32_t c1 = ld_code(cl + dd + 1);
    if constexpr (NS == 2) {
        // Read B's first row lies below an identity slot in a lane l > 0,
        // which computes columns 1 - l .. 0 first: there the identity row's
        // D = K/H makes M nonzero.  So every lane enters the first
        // haplotype as at a hand-over from one that ends at column 0: its
        // state is set to the column-0 boundary in the step that computes
        // its column 1 (the first haplotype has at least L + 2 columns).
        float d0[MR];
#pragma unroll
        for (int s = 0; s < MR; s++) d0[s] = realv[s] ? 0.0f : initd0;
        const float dD0 = rot(d0[MR - 1]);
        for (int u = 0; u < L; u++) {
            const uint32_t c2 = ld_code(cl + dd + 2);
            if (dd - l == 1) {
                sm = 0.0f;
                si = 0.0f;
#pragma unroll
                for (int s = 0; s < MR; s++) {
                    S.M[s] = 0.0f;
                    S.I[s] = 0.0f;
                    S.D[s] = d0[s];
                    Pm[s] = 0.0f;
                }
                C.dM = 0.0f;
                C.dI = 0.0f;
                C.dD = dD0;
            }
            float pb[MR];
            step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
#pragma unroll
            for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            c1 = c2;
            dd++;
        }
    }
    for (uint32_t k = 0; k + 1 < K; k++) {
        const ChainHap *hn = a.chain_hap + c0 + k + 1;
        const int Hn = __builtin_amdgcn_readfirstlane((int)hn->H);
        const float initdn = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(hn->initd)));
        const int pdn = __builtin_amdgcn_readfirstlane(hn->pdelta);
        const uint32_t hlo = __builtin_amdgcn_readfirstlane((uint32_t)hn->hoff);
        const uint32_t hhi = __builtin_amdgcn_readfirstlane((uint32_t)(hn->hoff >> 32));
        const int Sn = Sk + H;
        // the other buffer was last read in the previous hand-over
        uint8_t *nb = buf + ((k + 1) & 1) * a.lds_stride;
        stage_codes(a, ((uint64_t)hhi << 32) | hlo, Hn, nb + P0);
        const uint8_t *cln = nb + (P0 - 1) - l - Sn;
        // every lane, and every code it prefetches, inside haplotype k
        for (; dd + kU - 1 <= Sn - 2; dd += kU) {
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t c2 = ld_code(cl + dd + u + 2);
                float pb[MR];
                step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
                sm += S.M[MR - 1];
                si += S.I[MR - 1];
#pragma unroll
                for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
                c1 = c2;
            }
        }
        for (; dd <= Sn - 2; dd++) {
            const uint32_t c2 = ld_code(cl + dd + 2);
            float pb[MR];
            step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
#pragma unroll
            for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            c1 = c2;
        }
        // hand-over: lane l reads the new codes from step Sn - 1 + l on and
        // enters haplotype k+1 at step Sn + 1 + l
        float dres[MR];
#pragma unroll
        for (int s = 0; s < MR; s++) dres[s] = gr.read_of(l * MR + s) >= 0 ? 0.0f : initdn;
        const float dDres = rot(dres[MR - 1]);
        // steps Sn - 1 .. Sn + L, two at a time with the prior registers
        // swapping roles (no copies)
        auto hand = [&](const float (&pin)[MR], float (&pout)[MR]) {
            if (dd + 1 - l == Sn) cl = cln;
            const uint32_t c2 = ld_code(cl + dd + 2);
            const bool cross = dd - l == Sn + 1;
            if (cross && own >= 0) put_raw32(a, own + pd, sm + si);  // pairhmm_impl.cpp:434-437
            if (cross) {
                sm = 0.0f;
                si = 0.0f;
#pragma unroll
                for (int s = 0; s < MR; s++) {
                    S.M[s] = 0.0f;
                    S.I[s] = 0.0f;
                    S.D[s] = dres[s];
                    Pm[s] = 0.0f;
                }
                C.dM = 0.0f;
                C.dI = 0.0f;
                C.dD = dDres;
            }
            step_tab<L, MR, INTER, SH>(S, C, Pm, P, pin, rot, pout, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
            c1 = c2;
            dd++;
        };
        static_assert((L + 2) % 2 == 0, "hand-over steps in pairs");
        if constexpr (MR < 16) {
            for (int u = 0; u < (L + 2) / 2; u++) {
                float pb[MR];
                hand(pa, pb);
                hand(pb, pa);
            }
        } else {  // in pairs the 16-row shapes would need more than 256 VGPRs
            for (int u = 0; u < L + 2; u++) {
                float pb[MR];
                hand(pa, pb);
#pragma unroll
                for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            }
        }
        Sk = Sn;
        H = Hn;
        pd = pdn;
    }
    // the last haplotype ends like a single pair
    const int key = lmax >= 0 ? Sk + H + lmin : 0x7fffffff;
    const int keymax = lmax >= 0 ? Sk + H + lmax : -1;
    const int Dmain = wave_min4(key, G);
    const int Dend = wave_max4(keymax, G);
    for (; dd + kU - 1 <= Dmain; dd += kU) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const uint32_t c2 = ld_code(cl + dd + u + 2);
            float pb[MR];
            step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
#pragma unroll
            for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            c1 = c2;
        }
    }
    for (; dd <= Dend; dd++) {
        const uint32_t c2 = ld_code(cl + dd + 2);
        float pb[MR];
        step_tab<L, MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
        if (dd - l - Sk <= H) {
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
        }
#pragma unroll
        for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
        c1 = c2;
    }
    if (own >= 0) put_raw32(a, own + pd, sm + si);
    // every lane has finished its LDS reads before the next item rewrites
    // the table and the first code buffer
    wave_lds_sync();
    return __builtin_amdgcn_readfirstlane(next);
}

// Persistent, like fwd_tab_kernel, over chained items.  LDS: the G prior
// tables, then the two code buffers, which start all zero (prefetches past
// a haplotype read codes left there, always a valid block 0..3).
// Up to 12 rows per lane the kernel is held to 168 VGPRs, three waves per
// SIMD where the LDS allows (a few set-up values spill; the step loop keeps
// its registers): 0.7 % faster on the batched workload than at 2 waves, and
// 0.3 % faster than 12 rows at 2 waves without its 100 bytes of spills.
template <int L, int MR, bool INTER, int NS, bool SH>
__global__ __launch_bounds__(64, MR <= 12 ? 3 : 2) void fwd_chain_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int G = 64 / L;
    constexpr uint32_t tabs = (uint32_t)G * 4u * (uint32_t)L * TabLayout<MR>::MRT * 4u;
    for (uint32_t k = threadIdx.x * 16u; k < 2u * a.lds_stride; k += 64u * 16u)
        *reinterpret_cast<uint4 *>(lds + tabs + k) = make_uint4(0, 0, 0, 0);
    wave_lds_sync();
    const uint32_t n = a.n_items;
    uint32_t it = 0;
    if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    for (uint32_t iter = 0; iter < n && it < n; iter++) it = process_chain<L, MR, INTER, NS, SH>(a, lds, it);
}

template <typename T, int L, int MR, bool INTER, bool RESCUE>
__global__ __launch_bounds__(64) void fwd_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int G = 64 / L;
    const uint32_t n = RESCUE ? *a.n_items_dev : a.n_items;
    // an empty list ends every wave before it queues on the counter's atomic
    if (__builtin_amdgcn_readfirstlane(n) == 0) return;
    // a wave takes at most n items: the bound makes every wave exit
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        int p;
        if constexpr (RESCUE)
            p = (int)a.rescue_list[it];
        else
            p = a.items[(uint64_t)it * G + threadIdx.x / L];
        process_item<T, L, MR, INTER, RESCUE>(a, lds, p);
    }
}

// One rescued pair in FP64 with as few rows per lane as its read needs: one
// strip of 63*MR rows for MR = 1..6.  Fewer rows per lane leave fewer
// identity rows to compute; the arithmetic of every read row is the same.
// Longer reads are fwd_rescue_long_kernel's.
constexpr int kRescueStrip = 63 * 6;
template <bool INTER>
__device__ __forceinline__ void rescue_pair(const FwdArgs &a, uint8_t *lds, int p) {
    const int R = __builtin_amdgcn_readfirstlane((int)a.read_len[a.pair_read[p]]);
    if (R <= 63)
        process_item<double, 64, 1, INTER, true>(a, lds, p);
    else if (R <= 126)
        process_item<double, 64, 2, INTER, true>(a, lds, p);
    else if (R <= 189)
        process_item<double, 64, 3, INTER, true>(a, lds, p);
    else if (R <= 252)
        process_item<double, 64, 4, INTER, true>(a, lds, p);
    else if (R <= 315)
        process_item<double, 64, 5, INTER, true>(a, lds, p);
    else if (R <= kRescueStrip)
        process_item<double, 64, 6, INTER, true>(a, lds, p);
}

// ------------------------------------------------- long reads
// Strips [s0, s1) of a pair longer than one strip (64 lanes x MR rows, lane
// 63 the identity lane / feeder, as in process_item), pipelined over the
// kLW waves of a workgroup: wave w takes strips s0 + w, s0 + w + kLW, ...
// The waves advance in blocks of kLB steps with one workgroup barrier per
// block.  Strip s0 + c starts at block (c / kLW) * P + 2 * (c % kLW),
// P = max(blocks per strip, 2 * kLW): a wave's strips do not overlap, and
// every strip starts at least two blocks after the strip above it.  The
// boundary row strip s+1 reads at its step d (column d + 1 of strip s's
// last row) is written by strip s at its step d + L - 1 <= d + kLB - 1, so
// in an earlier block, before the barrier that precedes its use.  Every
// wave runs the same number of blocks, so the barrier count is uniform.
// Boundary rows go through kLW + 1 HBM buffers per workgroup (strip s0 + c
// writes buffer c % (kLW + 1)); a buffer is written again only by strip
// c + kLW + 1, which starts once strip c + 1, its reader, has finished.
//
// A pair split into units of at most kLW strips runs its units on several
// workgroups at once.  The last strip of a unit hands its boundary row to
// the next unit's first strip through a slot: an HBM row and a counter of
// the blocks written, published with release order after each block.  The
// next unit's first strip waits, at the start of its block b, until
// min(b + 2, blocks per strip) blocks are published -- the two-block lag of
// the strips inside a workgroup.  Units are taken from a work counter in
// list order, so the unit a workgroup waits for was taken earlier by a
// running workgroup, and the wait ends.  It is bounded all the same
// (a.spin_limit polls): if the bound runs out, the unit counts it in
// *a.fault (the plan's read then reports a device error), NaN goes into the
// strip and from there into the pair's result, and the wait is not
// repeated.
constexpr int kLW = 8;   // waves per workgroup (2 per SIMD; 4 was 1.5x slower)
// waves per workgroup of the FP64 long-read rescue: units of 4 strips, two
// workgroups per CU (a pair's last unit leaves fewer waves idle)
constexpr int kRescueLW = 4;
// rows per lane of the FP64 long-read rescue (strips of 63 * 7 = 441 rows)
constexpr int kLongRescueMR = 7;
constexpr int kLB = 64;  // steps per block (>= L - 1)

// Waits until *c >= need (the slot's blocks published); false if `limit`
// polls (of ~0.5-1 us each) did not see it.  Wave-uniform.
__device__ __forceinline__ bool wait_published(const uint32_t *c, uint32_t need, uint32_t limit) {
    for (uint32_t k = 0; k < limit; k++) {
        const uint32_t v = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (v >= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the row's stores before the count
            return true;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    return false;
}

// Lane k's value, for every lane (k uniform).
__device__ __forceinline__ float lane_value(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ double lane_value(double v, int k) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), k),
                            __builtin_amdgcn_readlane(__double2loint(v), k));
}

// Long-read prior tables (TAB): per wave, the emission priors of its strip's
// rows for each haplotype base, [4 bases][chunks][64 lanes][16 bytes], a
// lane's MR rows in LT::E-element 16-byte chunks.
template <typename T, int MR>
struct LongTab {
    static constexpr int E = 16 / (int)sizeof(T);               // elements per chunk
    static constexpr int NC = (MR + E - 1) / E;                 // chunks per lane
    static constexpr uint32_t kBlk = 64u * (uint32_t)NC * 16u;  // bytes per base
    static constexpr uint32_t kWave = 4u * kBlk;
};

template <typename T, int MR>
__device__ __forceinline__ void long_priors(T (&pr)[MR], const uint8_t *p) {
    using LT = LongTab<T, MR>;
    const uint4 *tp = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int q = 0; q < LT::NC; q++) {
        const uint4 v = tp[q * 64];
        const T *e = reinterpret_cast<const T *>(&v);
#pragma unroll
        for (int k = 0; k < LT::E; k++)
            if (q * LT::E + k < MR) pr[q * LT::E + k] = e[k];
    }
}

// TAB: the haplotype has no N, and each step reads its slots' priors from
// the wave's table in LDS (one 16-byte read per chunk, the next column's
// issued a step ahead) instead of selecting them per cell from the one-hot
// code (2 VALU instructions per cell in FP32, 4 in FP64).
template <typename T, int MR, bool INTER, bool RESCUE, bool TAB, int W>
__device__ void long_pair(const FwdArgs &a, uint8_t *lds, int p, int s0, int s1, uint32_t xin,
                          uint32_t xout) {
    constexpr int L = 64;
    constexpr int C = (L - 1) * MR;
    using LT = LongTab<T, MR>;
    static_assert(kLB >= L - 1, "a block covers the boundary row's lag");
    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t h = __builtin_amdgcn_readfirstlane(a.pair_hap[p]);
    const uint32_t r = __builtin_amdgcn_readfirstlane(a.pair_read[p]);
    const int R = __builtin_amdgcn_readfirstlane((int)a.read_len[r]);
    const int H = __builtin_amdgcn_readfirstlane((int)a.hap_len[h]);
    const uint64_t ro64 = a.read_off[r];
    const uint64_t ro = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ro64 >> 32)) << 32) |
                        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)ro64);
    // (off_mr: another rows-per-lane layout whose rows these strips cover, a
    // multiple of MR: row R still ends its lane, (R + off) % MR == 0)
    const int omr = a.off_mr ? (int)a.off_mr : MR;
    const int off = (omr - R % omr) % omr;
    const int vR = R - 1 + off;
    const int nstr = vR / C + 1;
    const int lR = (vR % C) / MR;
    const T initd = Tab<T>::initd(a, h);

    // one-hot codes of the haplotype, shared by the waves: column j at
    // lds[L - 1 + j] (TAB: the base's block 0..3 instead)
    for (uint32_t k = threadIdx.x * 16u; k < a.lds_stride; k += 64u * W * 16u)
        *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    __syncthreads();
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[h]);
        const int nch = (H + 15) >> 4;
        for (int c = threadIdx.x; c < nch; c += 64 * W) {
            uint4 v = src[c];
            if constexpr (TAB) {  // one-hot 1, 2, 4, 8 -> 0..3 per byte
                v.x = ((v.x >> 1) & 0x7f7f7f7fu) - ((v.x >> 3) & 0x1f1f1f1fu);
                v.y = ((v.y >> 1) & 0x7f7f7f7fu) - ((v.y >> 3) & 0x1f1f1f1fu);
                v.z = ((v.z >> 1) & 0x7f7f7f7fu) - ((v.z >> 3) & 0x1f1f1f1fu);
                v.w = ((v.w >> 1) & 0x7f7f7f7fu) - ((v.w >> 3) & 0x1f1f1f1fu);
            }
            *reinterpret_cast<uint4 *>(lds + L + 16 * c) = v;
        }
    }
    __syncthreads();
    const uint8_t *cp = lds + (L - 1) - l;
    // this wave's prior table (TAB), after the codes; lane l's chunks
    uint8_t *tabl = lds + a.lds_stride + (uint32_t)w * LT::kWave + (uint32_t)l * 16u;
    T pa[MR];        // TAB: the priors of the current step's column
    uint32_t c1 = 0; // TAB: the code of the next step's column

    if (s1 > nstr) s1 = nstr;
    const int nu = s1 - s0;  // strips of this unit
    const int nb = (H + 2 * L - 1 + kLB - 1) / kLB;  // blocks of a strip (steps 1 .. H + 2L - 1)
    const int P0 = nb > 2 * W ? nb : 2 * W;
    const int nblk = ((nu - 1) / W) * P0 + 2 * ((nu - 1) % W) + nb;
    T *bbase = reinterpret_cast<T *>(a.bnd) + (uint64_t)blockIdx.x * (W + 1) * a.bnd_elems;
    T *xb_in = xin != kNoSlot ? reinterpret_cast<T *>(a.xbnd) + (uint64_t)xin * a.bnd_elems : nullptr;
    T *xb_out = xout != kNoSlot ? reinterpret_cast<T *>(a.xbnd) + (uint64_t)xout * a.bnd_elems : nullptr;
    bool xlive = true;  // the waits for the unit above have not run out

    const LaneRot<L> rot;
    const bool feeder = l == L - 1;
    Priors<T, MR> P;
    State<T, MR> S;
    T yyA = T(0), myA = T(0), sm = T(0), si = T(0);
    T *bnd_in = nullptr, *bnd_out = nullptr;
    bool inject = false, store = false, last = false;
    int Dend = 0, d = 1;
    int cr = w;  // this wave's current strip, s0 + cr
    // the FP32 pass's first launch (dead_mode 1, the first strip of the
    // pair's layout): the bits of the unit's last row's entries at columns
    // 1..H that the storing lane wrote, ORed
    const bool detect = !RESCUE && a.dead_mode == 1;
    uint32_t orbits = 0;
    for (int t = 0; t < nblk; t++) {
        const int st0 = (cr / W) * P0 + 2 * (cr % W);
        bool pub = false;  // this block of the unit's last strip goes to slot xout
        if (cr < nu && t >= st0 && t < st0 + nb) {
            const int cs = s0 + cr;
            if (cr == 0 && xb_in && xlive) {  // the unit above's blocks up to this one's lag
                const int need = t + 2 < nb ? t + 2 : nb;
                xlive = wait_published(a.prog + xin, (uint32_t)need, a.spin_limit);
                if (!xlive && l == 0) atomicAdd(a.fault, 1u);
            }
            if (t == st0) {  // the strip's set-up, as process_item's
#pragma unroll
                for (int s = 0; s < MR; s++) {
                    const int i = cs * C + l * MR + s - off + 1;  // 1-based read row
                    const bool real = l < L - 1 && i >= 1 && i <= R;
                    load_row<T, MR, INTER>(a, P, s, real, ro + (uint64_t)(real ? i - 1 : 0));
                    S.M[s] = T(0);
                    S.I[s] = T(0);
                    S.D[s] = real ? T(0) : initd;
                }
                if constexpr (TAB) {
                    // the wave's table for this strip (its previous strip's
                    // reads of it have all returned: their values were used)
#pragma unroll
                    for (int c = 0; c < 4; c++)
#pragma unroll
                        for (int s = 0; s < MR; s++)
                            reinterpret_cast<T *>(tabl + c * LT::kBlk + (s / LT::E) * (64u * 16u))[s % LT::E] =
                                P.rpos[s] == (uint32_t)c ? P.X[s] : P.Y[s];
                    wave_lds_sync();
                    long_priors<T, MR>(pa, tabl + ld_code(cp + 1) * LT::kBlk);
                    c1 = ld_code(cp + 2);
                }
                S.Mp = T(0);
                S.Ip = T(0);
                S.Dp = rot(S.D[MR - 1]);
                yyA = rot(P.yy[MR - 1]);
                myA = rot(P.my[MR - 1]);
                if (cs > 0 && l == 0) {
                    const uint64_t ip = ro + (uint64_t)(cs * C - off - 1);  // 0-based row above
                    yyA = Tab<T>::ph2pr(a.tab)[a.gcp[ip] & 127u];
                    myA = Tab<T>::ph2pr(a.tab)[a.del[ip] & 127u];
                    S.Dp = T(0);
                }
                last = cs == nstr - 1;
                inject = cs > 0;
                store = !last;
                bnd_out = cr == nu - 1 ? xb_out : bbase + (uint64_t)(cr % (W + 1)) * a.bnd_elems;
                bnd_in = cr == 0 ? xb_in  // strip cs - 1's
                                 : bbase + (uint64_t)((cr + W) % (W + 1)) * a.bnd_elems;
                if (inject) {
                    T bm = ld_agent(bnd_in + 3 * (1 + L)), bi = ld_agent(bnd_in + 3 * (1 + L) + 1),
                      bd = ld_agent(bnd_in + 3 * (1 + L) + 2);
                    if (feeder) {
                        S.M[MR - 1] = bm;
                        S.I[MR - 1] = bi;
                        S.D[MR - 1] = bd;
                    }
                }
                Dend = last ? H + lR : H + 2 * L - 1;
                sm = T(0);
                si = T(0);
                d = 1;
            }
            const int dlim = d + kLB - 1 < Dend ? d + kLB - 1 : Dend;
            // the feeder's inputs of this block's steps, all written before
            // the last barrier: lane k loads column d0 + k + 1 of the strip
            // above (index d0 + k + 1 + L), one coalesced load per block;
            // step d0 + k then takes lane k's values with readlane
            const int d0 = d;
            T fm = T(0), fi = T(0), fd = T(0);
            if (inject && d0 + l <= dlim) {
                const T *in = bnd_in + 3 * (d0 + l + 1 + L);
                fm = ld_agent(in);
                fi = ld_agent(in + 1);
                fd = ld_agent(in + 2);
            }
            if (!xlive && cr == 0) fm = T(__builtin_nan(""));
            // one step; two per iteration, so that the registers of the
            // state and the priors alternate instead of being copied back
            // at the end of every step (the loop holds convergent readlanes,
            // which the compiler does not unroll with a remainder)
            auto one_step = [&](const int d) {
                if constexpr (TAB) {
                    const uint32_t c2 = ld_code(cp + d + 2);
                    T pb[MR];
                    long_priors<T, MR>(pb, tabl + c1 * LT::kBlk);
                    step_with<T, L, MR, INTER>(S, P, yyA, myA, rot, [&](int s) { return pa[s]; });
#pragma unroll
                    for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
                    c1 = c2;
                } else {
                    step<T, L, MR, INTER>(S, P, yyA, myA, cp[d], rot);
                }
                if (last && d - l <= H) {
                    sm += S.M[MR - 1];
                    si += S.I[MR - 1];
                }
                if (store && l == L - 2) {  // column d - (L-2) at index d + 2
                    T *o = bnd_out + 3 * (d + 2);
                    st_agent(o, S.M[MR - 1]);
                    st_agent(o + 1, S.I[MR - 1]);
                    st_agent(o + 2, S.D[MR - 1]);
                    if constexpr (!RESCUE) {
                        const int j = d - (L - 2);
                        if (detect && j >= 1 && j <= H)
                            orbits |= __float_as_uint((float)S.M[MR - 1]) | __float_as_uint((float)S.I[MR - 1]) |
                                      __float_as_uint((float)S.D[MR - 1]);
                    }
                }
                if (inject) {  // column d + 1 of the strip above
                    const int k = d - d0;
                    const T bm = lane_value(fm, k), bi = lane_value(fi, k), bd = lane_value(fd, k);
                    if (feeder) {
                        S.M[MR - 1] = bm;
                        S.I[MR - 1] = bi;
                        S.D[MR - 1] = bd;
                    }
                }
            };
            for (; d + 1 <= dlim; d += 2) {
                one_step(d);
                one_step(d + 1);
            }
            if (d <= dlim) one_step(d++);
            if (t == st0 + nb - 1) {
                if constexpr (!RESCUE) {
                    if (detect && !last && cr == nu - 1) {
                        // the unit's last row all zero at columns 1..H: every
                        // later row, and the FP32 sum, are exactly zero -- the
                        // reference's result, which rescues the pair; the
                        // other strips are not run
                        const uint32_t any = __builtin_amdgcn_readlane(orbits, L - 2);
                        if (l == 0) {
                            if (any == 0u) {
                                a.raw32[p] = 0.0f;
                                a.flags[p] = 0x01 | kFlagDead;
                                const uint32_t k = atomicAdd(a.rescue_count, 1u);
                                a.rescue_list[k] = (uint32_t)p;
                            } else {
                                a.flags[p] = 0;
                            }
                        }
                    }
                }
                if (last && l == lR) {
                    const T Ssum = sm + si;  // pairhmm_impl.cpp:434-437
                    if constexpr (RESCUE) {
                        a.raw64[p] = (double)Ssum;
                    } else {
                        a.raw32[p] = (float)Ssum;
                        uint8_t fl = 0;
                        if ((float)Ssum < 1e-28f) {  // MIN_ACCEPTED, pairhmm_api.cpp:4,17
                            fl = 0x01;
                            const uint32_t k = atomicAdd(a.rescue_count, 1u);
                            a.rescue_list[k] = (uint32_t)p;
                        }
                        a.flags[p] = fl;
                    }
                }
                cr += W;
            }
            pub = bnd_out == xb_out && store && xb_out;
        }
        // this block's boundary stores before the barrier: the rows handed to
        // the next wave of this workgroup need a workgroup-scope release; the
        // row a unit's last strip hands to a unit on another workgroup (pub)
        // an agent-scope one, which writes back this XCD's L2 and is taken by
        // that wave alone, and only where the consumer runs in the same
        // launch (the first launch's units are read by the next launch)
        if (pub && a.dead_mode != 1)
            __threadfence();
        else
            __threadfence_block();
        if (pub && l == L - 2)  // the storing lane: the blocks of the strip written so far
            __hip_atomic_store(a.prog + xout, (uint32_t)(t - st0 + 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
}

// Persistent over the units of the pairs longer than one strip (rescue_pair
// leaves them), one unit per workgroup at a time; a unit whose pair the
// FP32 pass did not reject is skipped, like every other unit of its pair.
// ROUTE: per pair inter or intra FP64 as the regroup decided.
template <bool ROUTE, bool TAB, int W>
__global__ __launch_bounds__(64 * W, kLW / W) void fwd_rescue_long_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t s_it;
    const uint32_t n = a.n_items;
    for (uint32_t iter = 0; iter <= n; iter++) {
        if (threadIdx.x == 0) s_it = atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t it = s_it;
        __syncthreads();  // every wave has read s_it before thread 0 writes it again
        if (it >= n) break;
        const LongUnit u = a.units[it];
        const uint8_t fl = a.flags[u.pair];
        if (!(fl & 0x01u)) continue;  // RESCUED
        const int s0 = (int)(u.strips & 0xffffu), s1 = (int)(u.strips >> 16);
        if (__builtin_amdgcn_readfirstlane((ROUTE && (fl & kFlagGrouped64)) ? 1 : 0))
            long_pair<double, kLongRescueMR, true, true, TAB, W>(a, lds, u.pair, s0, s1, u.xin, u.xout);
        else
            long_pair<double, kLongRescueMR, false, true, TAB, W>(a, lds, u.pair, s0, s1, u.xin, u.xout);
    }
}

// The FP32 pass of a 64-lane select-prior bucket whose reads need more than
// one strip, the pairs' strips pipelined over the workgroup's waves as
// above: over a.units when given, else one whole pair per item.
template <int MR, bool INTER, bool TAB>
__global__ __launch_bounds__(64 * kLW) void fwd_long_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t s_it;
    const uint32_t n = a.n_items;
    for (uint32_t iter = 0; iter <= n; iter++) {
        if (threadIdx.x == 0) s_it = atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t it = s_it;
        __syncthreads();  // every wave has read s_it before thread 0 writes it again
        if (it >= n) break;
        int p, s0 = 0, s1 = 1 << 16;
        uint32_t xin = kNoSlot, xout = kNoSlot;
        if (a.units) {
            const LongUnit u = a.units[it];
            p = u.pair;
            s0 = (int)(u.strips & 0xffffu);
            s1 = (int)(u.strips >> 16);
            xin = u.xin;
            xout = u.xout;
            // the pair's FP32 sum is zero: found by the strip-0 launch, which
            // completed before this one (every unit of the pair skips alike)
            if (a.dead_mode == 2 &&
                __builtin_amdgcn_readfirstlane((a.flags[p] & kFlagDead) ? 1 : 0))
                continue;
        } else {
            p = a.items[it];
        }
        if (p >= 0) long_pair<float, MR, INTER, false, TAB, kLW>(a, lds, p, s0, s1, xin, xout);
    }
}

// FP64 rescue of every pair as intra::computeLikelihoods would
// (pairhmm_api.cpp:20-27): the device-side list of FP32-rejected pairs.
__global__ __launch_bounds__(64) void fwd_rescue_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t n = *a.n_items_dev;
    if (__builtin_amdgcn_readfirstlane(n) == 0) return;  // empty list: no atomic
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        rescue_pair<false>(a, lds, (int)a.rescue_list[it]);
    }
}

// FP64 rescue under schedule_pairhmm routing: pairs the FP64 regroup put in
// an inter group run the inter arithmetic, every other rescued pair the
// intra one (pairhmm_schedule.cpp:273-304, 307-335).  One pair per wave, so
// the choice is uniform.
__global__ __launch_bounds__(64) void fwd_rescue_route_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t n = *a.n_items_dev;
    if (__builtin_amdgcn_readfirstlane(n) == 0) return;  // empty list: no atomic
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        const int p = (int)a.rescue_list[it];
        const uint32_t inter =
            __builtin_amdgcn_readfirstlane((a.flags[p] & kFlagGrouped64) ? 1u : 0u);
        if (inter)
            rescue_pair<true>(a, lds, p);
        else
            rescue_pair<false>(a, lds, p);
    }
}

// ------------------------------------------------- FP64 rescue, prior table
// Rescued pairs of reads of at most 255 rows and N-free haplotypes, two per
// wave: 32 lanes per pair in the table kernel's layout (every lane holds read
// rows; when R % MR != 0 the top slot of lane 0 is an identity slot) with
// MR = ceil((R + 1) / 32) rows per lane.  The emission priors of a pair come
// from its LDS table (two doubles per 16-byte chunk, a lane's chunks for one
// column read with ds_read_b128) instead of a select per cell, and the
// transfers between lanes are ds_bpermute.  The two pairs of a wave share MR
// and the arithmetic (inter or intra), so the rescued pairs are binned first
// (rescue_bin_kernel).  Every read cell is computed with process_item's
// instructions in its order on the same tables: the sums are the same bits.
constexpr int kRtL = 32;

template <int MR>
struct TabLayout64 {
    static constexpr int MRT = (MR + 1) / 2 * 2;  // slots padded to 16 bytes
};

template <int MR>
__device__ __forceinline__ void load_priors64(double (&pr)[MR], const uint8_t *p) {
    constexpr int MRT = TabLayout64<MR>::MRT;
    const double2 *tp = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int q = 0; q < MRT / 2; q++) {
        const double2 v = tp[q * kRtL];
        if (2 * q + 0 < MR) pr[2 * q + 0] = v.x;
        if (2 * q + 1 < MR) pr[2 * q + 1] = v.y;
    }
}

// The carried row above the lane's first row, as Carry (FP32) holds it.
struct Carry64 {
    double dM, dI, dD;
    double uM, uI;
    double yyA, myA;
};

// SH: shared products, as step_tab
template <int MR, bool INTER, bool SH>
__device__ __forceinline__ void step_tab64(State<double, MR> &S, Carry64 &C, double (&Pm)[MR],
                                           const Priors<double, MR> &P, const double (&pr)[MR],
                                           const LaneRot<kRtL> &rot, double (&pb)[MR], const uint8_t *pnext) {
    load_priors64<MR>(pb, pnext);
    double nM[MR], nD[MR];
    double dgM = C.dM, dgI = C.dI, dgD = C.dD;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        double t;
        if (INTER) {  // pairhmm_inter.cpp:279-283
            t = dgI * P.g[s];
            t = fmaT(dgM, P.mm[s], t);
            t = fmaT(dgD, P.g[s], t);
        } else {  // pairhmm_impl.cpp:266-275
            t = dgM * P.mm[s];
            t = fmaT(dgI, P.g[s], t);
            t = fmaT(dgD, P.g[s], t);
        }
        nM[s] = t * pr[s];
        const double v = SH ? Pm[s] : S.M[s] * P.my[s];
        nD[s] = fmaT(S.D[s], P.yy[s], v);
        dgM = S.M[s];
        dgI = S.I[s];
        dgD = S.D[s];
    }
    double upM = C.uM, upI = C.uI;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const double u = upM * P.mx[s];
        if (SH && s > 0) Pm[s - 1] = u;
        const double nI = fmaT(upI, P.xx[s], u);
        upM = nM[s];
        upI = nI;
        S.I[s] = nI;
        S.M[s] = nM[s];
        S.D[s] = nD[s];
    }
    if (SH) Pm[MR - 1] = nM[MR - 1] * P.my[MR - 1];
    C.dD = fmaT(C.dD, C.yyA, C.dM * C.myA);
    C.dM = C.uM;
    C.dI = C.uI;
    C.uM = rot(S.M[MR - 1]);
    C.uI = rot(S.I[MR - 1]);
}

// One item of two pairs (p < 0: an empty slot).  LDS per group: the codes
// (lds_stride bytes) and the table, 4 blocks of 32 * MRT doubles.
template <int MR, bool INTER, bool SH>
__device__ void rescue_tab_item(const FwdArgs &a, uint8_t *lds, int p) {
    constexpr int L = kRtL;
    constexpr int G = 64 / L;
    constexpr int MRT = TabLayout64<MR>::MRT;
    constexpr uint32_t kBlk = (uint32_t)L * MRT * 8;
    constexpr int kU64 = 4;
    const int lane = threadIdx.x;
    const int g = lane / L;
    const int l = lane % L;
    const bool valid = p >= 0;
    uint32_t h = 0, r = 0;
    int R = 0, H = 0;
    uint64_t ro = 0, hoff = 0;
    double initd = 0.0;
    if (valid) {
        h = a.pair_hap[p];
        r = a.pair_read[p];
        R = (int)a.read_len[r];
        H = (int)a.hap_len[h];
        ro = a.read_off[r];
        hoff = a.hapc_off[h];
        initd = a.initd_d[h];
    }
    const int off = valid ? (MR - R % MR) % MR : 0;
    const int lR = valid ? (R - 1 + off) / MR : 0;
    uint8_t *codes = lds + (uint32_t)g * (a.lds_stride + 4u * kBlk);
    uint8_t *tab = codes + a.lds_stride;

    for (uint32_t k = (uint32_t)l * 16; k < a.lds_stride; k += L * 16)
        *reinterpret_cast<uint4 *>(codes + k) = make_uint4(0, 0, 0, 0);
    wave_lds_sync();
    RowRaw raw[MR];
    bool realv[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = l * MR + s - off + 1;
        realv[s] = valid && i >= 1 && i <= R;
        raw[s] = load_raw(a, ro + (uint64_t)(realv[s] ? i - 1 : 0));
    }
    if (valid) {  // column j at codes[L-1+j], block 0..3 per byte
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + hoff);
        const int nch = (H + 15) >> 4;
        for (int c = l; c < nch; c += L) {
            const uint4 v = src[c];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            uint32_t o[4];
#pragma unroll
            for (int k = 0; k < 4; k++) o[k] = ((w[k] >> 1) & 0x7f7f7f7fu) - ((w[k] >> 3) & 0x1f1f1f1fu);
            *reinterpret_cast<uint4 *>(codes + L + 16 * c) = make_uint4(o[0], o[1], o[2], o[3]);
        }
    }
    Priors<double, MR> P;
    double X[MR], Y[MR];
    uint32_t rc[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        row_priors<double, MR, INTER>(a, P, s, realv[s], raw[s]);
        X[s] = P.X[s];
        Y[s] = P.Y[s];
        rc[s] = P.rpos[s];
    }
    double *tl = reinterpret_cast<double *>(tab + (uint32_t)l * 16);
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int s = 0; s < MR; s++)
            tl[c * (kBlk / 8) + (s >> 1) * (L * 2) + (s & 1)] = (rc[s] == (uint32_t)c) ? X[s] : Y[s];
    }
    wave_lds_sync();

    State<double, MR> S;
    double Pm[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) {
        S.M[s] = 0.0;
        S.I[s] = 0.0;
        S.D[s] = realv[s] ? 0.0 : initd;
        Pm[s] = 0.0;
    }
    const LaneRot<L> rot;
    Carry64 C;
    C.dM = 0.0;
    C.dI = 0.0;
    C.dD = rot(S.D[MR - 1]);
    C.uM = 0.0;
    C.uI = 0.0;
    C.yyA = rot(P.yy[MR - 1]);
    C.myA = rot(P.my[MR - 1]);

    const int key = valid ? H + lR : 0x7fffffff;
    const int keymax = valid ? H + lR : -1;
    const int Dmain = wave_min4(key, G);
    const int Dend = wave_max4(keymax, G);
    const uint8_t *cl = codes + (L - 1) - l;
    const uint8_t *tlane = tab + (uint32_t)l * 16;
    double sm = 0.0, si = 0.0;
    int dd = 1;
    double pa[MR];
    load_priors64<MR>(pa, tlane + ld_code(cl + dd) * kBlk);
    uint32_t c1 = ld_code(cl + dd + 1);
    for (; dd + kU64 - 1 <= Dmain; dd += kU64) {
#pragma unroll
        for (int u = 0; u < kU64; u++) {
            const uint32_t c2 = ld_code(cl + dd + u + 2);
            double pb[MR];
            step_tab64<MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
#pragma unroll
            for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
            c1 = c2;
        }
    }
    for (; dd <= Dend; dd++) {
        const uint32_t c2 = ld_code(cl + dd + 2);
        double pb[MR];
        step_tab64<MR, INTER, SH>(S, C, Pm, P, pa, rot, pb, tlane + c1 * kBlk);
        if (dd - l <= H) {
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
        }
#pragma unroll
        for (int s2 = 0; s2 < MR; s2++) pa[s2] = pb[s2];
        c1 = c2;
    }
    if (valid && l == lR) a.raw64[p] = sm + si;  // pairhmm_impl.cpp:434-437
    wave_lds_sync();
}

// v[k] for a wave-uniform k, without indexing the kernel argument at run time
__device__ __forceinline__ uint32_t pick9(const uint32_t (&v)[9], int k) {
    uint32_t r = v[0];
#pragma unroll
    for (int q = 1; q < 9; q++) r = k == q ? v[q] : r;
    return r;
}

// Bins the rescue list (rescue_count entries, appended by the FP32 pass) for
// the FP64 passes.  Pairs of one class and arithmetic in a wave take one
// atomic per wave and class.
__global__ __launch_bounds__(256) void rescue_bin_kernel(FwdArgs a) {
    const uint32_t n = *a.rescue_count;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    // a whole wave walks the loop together (n is the same for all lanes)
    const uint32_t wbase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint32_t e0 = wbase; e0 < n; e0 += stride) {
        const uint32_t e = e0 + lane;
        int key = -1;  // 0..7 inter, 8..15 intra, 16 other
        uint32_t p = 0;
        if (e < n) {
            p = a.rescue_list[e];
            const uint32_t R = a.read_len[a.pair_read[p]];
            const uint32_t hh = a.pair_hap[p];
            const int c = rescue_class(R, a.hap_len[hh], a.hap_n[hh] != 0);
            if (c < 0)
                key = 2 * kRescueClasses;
            else
                key = (a.bin_route && (a.flags[p] & kFlagGrouped64)) ? c : kRescueClasses + c;
        }
        uint64_t todo = __ballot(key >= 0);
        while (todo) {
            const int lead = __builtin_ctzll(todo);
            const int k = __builtin_amdgcn_readlane(key, lead);
            const uint64_t mk = __ballot(key == k);
            uint32_t b = 0;
            if ((int)lane == lead) b = atomicAdd(a.bin_cnt + k, (uint32_t)__popcll(mk));
            b = (uint32_t)__builtin_amdgcn_readlane((int)b, lead);
            if (key == k) {
                const uint32_t slot = b + (uint32_t)__popcll(mk & below);
                const int c = k < kRescueClasses ? k : k < 2 * kRescueClasses ? k - kRescueClasses : kRescueClasses;
                const uint32_t base = pick9(a.bin_base, c);
                a.bin[k >= kRescueClasses && k < 2 * kRescueClasses ? base + pick9(a.bin_cap, c) - 1 - slot
                                                                     : base + slot] = p;
            }
            todo &= ~mk;
        }
    }
}

// Items of the table pass: two pairs of one class and arithmetic, classes
// c in [c_lo, c_hi] with the most rows per lane first, inter before intra.
template <int Q, int C_HI, bool SH>
__device__ __forceinline__ void rescue_tab_switch(const FwdArgs &a, uint8_t *lds, int c, bool inter,
                                                  uint32_t e, uint32_t ck) {
    if (c == Q) {
        int p = -1;
        if (e < ck) p = (int)a.bin[inter ? a.bin_base[Q] + e : a.bin_base[Q] + a.bin_cap[Q] - 1 - e];
        if (inter)
            rescue_tab_item<Q + 1, true, SH>(a, lds, p);
        else
            rescue_tab_item<Q + 1, false, SH>(a, lds, p);
    } else if constexpr (Q < C_HI) {
        rescue_tab_switch<Q + 1, C_HI, SH>(a, lds, c, inter, e, ck);
    }
}

template <int C_LO, int C_HI, bool SH>
__global__ __launch_bounds__(64) void fwd_rescue_tab_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int NK = 2 * (C_HI - C_LO + 1);
    uint32_t cnt[NK], ni[NK];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int c = C_HI - k / 2;
        cnt[k] = __builtin_amdgcn_readfirstlane(a.bin_cnt[(k & 1) ? kRescueClasses + c : c]);
        ni[k] = (cnt[k] + 1) / 2;
        tot += ni[k];
    }
    if (tot == 0) return;  // empty classes: no atomic
    for (uint32_t iter = 0; iter <= tot; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= tot) break;
        int k = 0;
        uint32_t j = it;
#pragma unroll
        for (int q = 0; q + 1 < NK; q++)
            if (k == q && j >= ni[q]) {
                j -= ni[q];
                k = q + 1;
            }
        uint32_t ck = 0;
#pragma unroll
        for (int q = 0; q < NK; q++)
            if (k == q) ck = cnt[q];
        const int c = C_HI - k / 2;
        const bool inter = (k & 1) == 0;
        rescue_tab_switch<C_LO, C_HI, SH>(a, lds, c, inter, 2 * j + threadIdx.x / kRtL, ck);
    }
}

// ------------------------------------------------- extended range
// The FP64 forward of a pair whose FP64 sum underflowed to zero (the
// reference's -inf), again with a finite result: the pair's strips of 63 *
// kExtMR rows run in sequence in one wave, and the row a strip hands to the
// next is scaled by a power of two that takes its largest entry to
// 2^kExtTop.  Scaling by powers of two is exact, so the strip below computes
// the true values times 2^E, E the sum of the shifts, with the same
// roundings as an unbounded exponent range would give (up to entries that
// are subnormal in one of the two).  A strip's entries lie within a factor
// of at most ~H of its input row's largest, so nothing overflows; the bottom
// of a strip is finite while its rows lose less than ~2000 binades to its
// top (about 2 decades per read base at 252 rows), else the pair still
// underflows.  The scaled sum goes to raw64 and E to exp64 (log10 L =
// log10(raw64) - E log10(2) - log10(2^1020)); the pair is flagged
// RODV_PHMM_FLAG_EXTENDED.  Every other pair is left as it was.
constexpr int kExtMR = 4;
constexpr int kExtTop = 960;

template <bool INTER>
__device__ void ext_pair(const FwdArgs &a, uint8_t *lds, int p) {
    using T = double;
    constexpr int L = 64;
    constexpr int MR = kExtMR;
    constexpr int C = (L - 1) * MR;
    const int l = threadIdx.x;
    const uint32_t h = __builtin_amdgcn_readfirstlane(a.pair_hap[p]);
    const uint32_t r = __builtin_amdgcn_readfirstlane(a.pair_read[p]);
    const int R = __builtin_amdgcn_readfirstlane((int)a.read_len[r]);
    const int H = __builtin_amdgcn_readfirstlane((int)a.hap_len[h]);
    const uint64_t ro64 = a.read_off[r];
    const uint64_t ro = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ro64 >> 32)) << 32) |
                        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)ro64);
    const int off = (MR - R % MR) % MR;
    const int vR = R - 1 + off;
    const int nstr = vR / C + 1;
    const int lR = (vR % C) / MR;
    const T initd = Tab<T>::initd(a, h);

    // one-hot codes: column j at lds[L - 1 + j]
    __syncthreads();
    for (uint32_t k = l * 16u; k < a.lds_stride; k += 64u * 16u)
        *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    __syncthreads();
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[h]);
        const int nch = (H + 15) >> 4;
        for (int c = l; c < nch; c += 64) *reinterpret_cast<uint4 *>(lds + L + 16 * c) = src[c];
    }
    __syncthreads();
    const uint8_t *cp = lds + (L - 1) - l;

    // two boundary rows per wave, strip cs writing buffer cs & 1
    T *bbase = reinterpret_cast<T *>(a.bnd) + (uint64_t)blockIdx.x * 2 * a.bnd_elems;
    const LaneRot<L> rot;
    const bool feeder = l == L - 1;
    int E = 0;   // entries are the true values times 2^1020 times 2^E
    int sh = 0;  // shift of the row the current strip reads
    T sm = T(0), si = T(0);
    for (int cs = 0; cs < nstr; cs++) {
        Priors<T, MR> P;
        State<T, MR> S;
#pragma unroll
        for (int s = 0; s < MR; s++) {
            const int i = cs * C + l * MR + s - off + 1;  // 1-based read row
            const bool real = l < L - 1 && i >= 1 && i <= R;
            load_row<T, MR, INTER>(a, P, s, real, ro + (uint64_t)(real ? i - 1 : 0));
            S.M[s] = T(0);
            S.I[s] = T(0);
            S.D[s] = real ? T(0) : initd;
        }
        S.Mp = T(0);
        S.Ip = T(0);
        S.Dp = rot(S.D[MR - 1]);
        T yyA = rot(P.yy[MR - 1]), myA = rot(P.my[MR - 1]);
        if (cs > 0 && l == 0) {
            const uint64_t ip = ro + (uint64_t)(cs * C - off - 1);  // 0-based row above
            yyA = Tab<T>::ph2pr(a.tab)[a.gcp[ip] & 127u];
            myA = Tab<T>::ph2pr(a.tab)[a.del[ip] & 127u];
            S.Dp = T(0);
        }
        const bool last = cs == nstr - 1;
        const bool inject = cs > 0;
        T *bnd_out = bbase + (uint64_t)(cs & 1) * a.bnd_elems;
        const T *bnd_in = bbase + (uint64_t)((cs + 1) & 1) * a.bnd_elems;
        if (inject) {  // column 1 of the row above, scaled
            const T bm = ldexp(ld_agent(bnd_in + 3 * (1 + L)), sh);
            const T bi = ldexp(ld_agent(bnd_in + 3 * (1 + L) + 1), sh);
            const T bd = ldexp(ld_agent(bnd_in + 3 * (1 + L) + 2), sh);
            if (feeder) {
                S.M[MR - 1] = bm;
                S.I[MR - 1] = bi;
                S.D[MR - 1] = bd;
            }
        }
        // a strip that is not the last runs until lane L-2 has stored column
        // H + L + 1, every column the next strip's feeder reads
        const int Dend = last ? H + lR : H + 2 * L - 1;
        T bmax = T(0);  // largest entry the storing lane wrote
        sm = T(0);
        si = T(0);
        for (int d0 = 1; d0 <= Dend; d0 += L) {
            const int dlim = d0 + L - 1 < Dend ? d0 + L - 1 : Dend;
            // the block's feeder inputs, one coalesced load: lane k holds
            // column d0 + k + 1 of the row above (index d0 + k + 1 + L)
            T fm = T(0), fi = T(0), fd = T(0);
            if (inject && d0 + l <= dlim) {
                const T *in = bnd_in + 3 * (d0 + l + 1 + L);
                fm = ldexp(ld_agent(in), sh);
                fi = ldexp(ld_agent(in + 1), sh);
                fd = ldexp(ld_agent(in + 2), sh);
            }
            for (int d = d0; d <= dlim; d++) {
                step<T, L, MR, INTER>(S, P, yyA, myA, cp[d], rot);
                if (last && d - l <= H) {
                    sm += S.M[MR - 1];
                    si += S.I[MR - 1];
                }
                if (!last && l == L - 2) {  // column d - (L-2) at index d + 2
                    T *o = bnd_out + 3 * (d + 2);
                    st_agent(o, S.M[MR - 1]);
                    st_agent(o + 1, S.I[MR - 1]);
                    st_agent(o + 2, S.D[MR - 1]);
                    bmax = fmax(bmax, fmax(S.M[MR - 1], fmax(S.I[MR - 1], S.D[MR - 1])));
                }
                if (inject) {  // column d + 1 of the row above
                    const int k = d - d0;
                    const T bm = lane_value(fm, k), bi = lane_value(fi, k), bd = lane_value(fd, k);
                    if (feeder) {
                        S.M[MR - 1] = bm;
                        S.I[MR - 1] = bi;
                        S.D[MR - 1] = bd;
                    }
                }
            }
        }
        if (!last) {
            __threadfence();  // the row's stores before the next strip's loads
            const T mx = lane_value(bmax, L - 2);
            // an all-zero row leaves every later entry zero: no shift
            sh = mx > T(0) ? kExtTop - ilogb(mx) : 0;
            E += sh;
        }
    }
    if (l == lR) {
        a.raw64[p] = sm + si;  // pairhmm_impl.cpp:434-437
        a.exp64[p] = E;
        a.flags[p] |= kFlagExtended;
    }
}

// Over the rescue list: the pairs whose FP64 sum is exactly zero, in the
// arithmetic (inter or intra) their FP64 pass used.  One pair per wave.
__global__ __launch_bounds__(64) void fwd_ext_kernel(FwdArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t n = *a.n_items_dev;
    if (__builtin_amdgcn_readfirstlane(n) == 0) return;  // empty list: no atomic
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        const int p = (int)__builtin_amdgcn_readfirstlane((int)a.rescue_list[it]);
        if (__builtin_amdgcn_readfirstlane(a.raw64[p] == 0.0 ? 1 : 0) == 0) continue;
        const uint32_t inter = __builtin_amdgcn_readfirstlane(
            (a.pair_inter && (a.flags[p] & kFlagGrouped64)) ? 1u : 0u);
        if (inter)
            ext_pair<true>(a, lds, p);
        else
            ext_pair<false>(a, lds, p);
    }
}

// greedy_grouping (pairhmm_schedule.cpp:46-105) over one region's FP32-rejected
// inter pairs, one wave per region.  The wave first compacts the region's
// rejected pairs, in sorted order, into the scratch list (ballot + prefix
// count).  On that list the reference's scan reduces to: at position i, stop
// if fewer than wd remain; else the window is the next wd entries (every
// accepted window is a contiguous block the scan then skips), accepted when
// idle = 1 - sum(R*H) / (wd * max R * max H) <= idle_d, in the reference's
// integer and double arithmetic.
// The window of W candidates starting at cand[s]: accepted when its idle
// ratio, as pairhmm_schedule.cpp computes it, is at most idle_d.  KW > 0:
// W <= KW, the loop unrolled so that the W candidates' loads are in flight
// together.
template <int KW>
__device__ __forceinline__ bool window_ok(const RegroupArgs &a, const uint32_t *cand, uint32_t s,
                                          uint32_t W) {
    uint32_t mh = 0, mr = 0;
    uint64_t sum = 0;
    auto take = [&](uint32_t k) {
        const uint32_t p = cand[s + k];
        const uint32_t H = a.hap_len[a.pair_hap[p]], R = a.read_len[a.pair_read[p]];
        mh = H > mh ? H : mh;
        mr = R > mr ? R : mr;
        sum += (uint64_t)H * R;
    };
    if constexpr (KW > 0) {
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)KW; k++)
            if (k < W) take(k);
    } else {
        for (uint32_t k = 0; k < W; k++) take(k);
    }
    const uint64_t cap = (uint64_t)mh * mr * W;
    const uint64_t idle = cap > sum ? cap - sum : 0;
    const double ratio = cap == 0 ? 0.0 : (double)idle / (double)cap;
    return ratio <= a.idle_d;
}

// One wave per region.  The region's rejected pairs are compacted in order
// (eight 64-entry chunks per pass, their loads in flight together); then the
// acceptance of the window at every start is decided by the lanes in
// parallel, 64 starts at a time, and the greedy walk over the starts --
// accept at i and go to i + W, else go to i + 1 -- runs on the ballot mask
// in scalar registers.  The walk visits the same starts as the sequential
// loop, so the same windows are accepted.
__global__ __launch_bounds__(64) void regroup_kernel(RegroupArgs a) {
    const uint32_t g = blockIdx.x;
    if (g >= a.n_regions || a.wd == 0) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t *list = a.list + a.off[g];
    uint32_t *cand = a.scratch + a.off[g];
    const uint32_t n = (uint32_t)(a.off[g + 1] - a.off[g]);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    constexpr int kC = 8;
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < n; base += 64 * kC) {
        uint32_t p[kC];
        bool c[kC];
#pragma unroll
        for (int m = 0; m < kC; m++) {
            const uint32_t i = base + 64 * m + lane;
            p[m] = i < n ? list[i] : 0u;
        }
#pragma unroll
        for (int m = 0; m < kC; m++) {
            const uint32_t i = base + 64 * m + lane;
            c[m] = i < n && (a.flags[p[m]] & 0x01) != 0;
        }
#pragma unroll
        for (int m = 0; m < kC; m++) {
            const uint64_t mk = __ballot(c[m]);
            if (c[m]) cand[cnt + (uint32_t)__popcll(mk & below)] = p[m];
            cnt += (uint32_t)__popcll(mk);
        }
    }
    cnt = __builtin_amdgcn_readfirstlane(cnt);
    const uint32_t W = a.wd;
    if (cnt < W) return;
    __syncthreads();  // the compacted list, written by all lanes, is read by all
    const uint32_t nstart = cnt - W + 1;  // window starts 0 .. cnt - W
    uint32_t i = 0;                       // the walk's position (uniform)
    for (uint32_t base = 0; base < nstart && i < nstart; base += 64) {
        const uint32_t s = base + lane;
        bool ok = false;
        if (s < nstart && s >= i) ok = W <= 16 ? window_ok<16>(a, cand, s, W) : window_ok<0>(a, cand, s, W);
        const uint64_t okm = __ballot(ok);
        while (i < nstart && i < base + 64) {
            if ((okm >> (i - base)) & 1ull) {
                if (lane < W) a.flags[cand[i + lane]] |= kFlagGrouped64;
                i += W;
            } else {
                i += 1;
            }
        }
    }
}

// log10 conversion (pairhmm_api.cpp:17-22 and pairhmm_inter_api.h:11-15).
__global__ void finalize_kernel(FinalizeArgs f) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= f.n) return;
    uint8_t fl = f.flags[k];
    const bool inter32 = f.pair_inter && f.pair_inter[k];
    const bool inter64 = (fl & kFlagGrouped64) != 0;
    fl &= (uint8_t)~(kFlagGrouped64 | kFlagDead);
    double out;
    bool inter = false;
    if (fl & 0x01) {
        // FP64 result: intra rescue, inter FP64 group, or the intra FP64
        // fallback of a rejected inter pair -- the same conversion
        double r = f.raw64[k];
        out = log10(r) - 307.050595577260822;
        // extended range: the sum carries 2^exp64 more
        if ((fl & kFlagExtended) && f.exp64) out -= (double)f.exp64[k] * 0.30102999566398119521;
        if (!(r > 0.0)) fl |= 0x02;
        inter = inter64;
    } else {
        float r = f.raw32[k];
        if (inter32)
            out = log10((double)r) - (double)36.1236000061f;
        else
            out = (double)(log10f(r) - 36.1236000061f);
        inter = inter32;
    }
    // within 2^-20 relative of the rescue threshold: the decision is exact
    // here, but would flip under a 1-ulp perturbation of the sum
    const float r32 = f.raw32[k];
    if (fabsf(r32 - 1e-28f) <= 1e-28f * 0x1p-20f) fl |= 0x04;
    if (inter) fl |= 0x08;
    f.out[k] = out;
    f.flags[k] = fl;
}

__global__ void reset_kernel(uint32_t *counters, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) counters[k] = 0;
}

// Sum of R * H over the rescue list (the FP64 pass's cells, for the plan's
// statistics): grid-stride over the list, a wave sum, one atomic per wave.
__global__ void rescued_cells_kernel(const uint32_t *list, const uint32_t *count,
                                     const uint32_t *pair_hap, const uint32_t *pair_read,
                                     const uint32_t *hap_len, const uint32_t *read_len,
                                     unsigned long long *out) {
    const uint32_t n = *count;
    unsigned long long acc = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t p = list[k];
        acc += (unsigned long long)read_len[pair_read[p]] * hap_len[pair_hap[p]];
    }
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

}  // namespace

// ------------------------------------------------------------ launchers

using FwdKernel = void (*)(FwdArgs);

template <typename T, int L, int MR, bool INTER, bool RESCUE>
static void *kernel_ptr() {
    return reinterpret_cast<void *>(&fwd_kernel<T, L, MR, INTER, RESCUE>);
}

// FP32 kernels: L = 16 for MR 1..16, L = 64 for MR in {4, 8, 12, 16}.
void *fwd_f32_kernel(int L, int MR, bool inter) {
#define K16(m) \
    case m: return inter ? kernel_ptr<float, 16, m, true, false>() : kernel_ptr<float, 16, m, false, false>();
#define K64(m) \
    case m: return inter ? kernel_ptr<float, 64, m, true, false>() : kernel_ptr<float, 64, m, false, false>();
    if (L == 16) {
        switch (MR) {
            K16(1) K16(2) K16(3) K16(4) K16(5) K16(6) K16(7) K16(8)
            K16(9) K16(10) K16(11) K16(12) K16(13) K16(14) K16(15) K16(16)
        default: return nullptr;
        }
    }
    if (L == 64) {
        switch (MR) {
            K64(4) K64(8) K64(12) K64(16)
        default: return nullptr;
        }
    }
#undef K16
#undef K64
    return nullptr;
}

// Table-prior FP32 kernels (haplotypes without N, one strip): 8 lanes per
// pair with 6..12 rows per lane, 16 lanes with 1..16, 32 lanes with 2..10,
// 64 lanes with 4, 8, 12, 16.  share: the shared-product variant (16 lanes
// only; nullptr elsewhere).
void *fwd_tab_kernel_ptr(int L, int MR, bool inter, bool share) {
#define TK(l, m) \
    case m: return inter ? reinterpret_cast<void *>(&fwd_tab_kernel<l, m, true, false>) : reinterpret_cast<void *>(&fwd_tab_kernel<l, m, false, false>);
#define TKS(m) \
    case m: return inter ? reinterpret_cast<void *>(&fwd_tab_kernel<16, m, true, true>) : reinterpret_cast<void *>(&fwd_tab_kernel<16, m, false, true>);
    if (share) {
        if (L != 16) return nullptr;
        switch (MR) {
            TKS(1) TKS(2) TKS(3) TKS(4) TKS(5) TKS(6) TKS(7) TKS(8)
            TKS(9) TKS(10) TKS(11) TKS(12) TKS(13) TKS(14) TKS(15) TKS(16)
        default: return nullptr;
        }
    }
    if (L == 8) {
        switch (MR) {
            TK(8, 6) TK(8, 7) TK(8, 8) TK(8, 9) TK(8, 10) TK(8, 11) TK(8, 12)
        default: return nullptr;
        }
    }
    if (L == 16) {
        switch (MR) {
            TK(16, 1) TK(16, 2) TK(16, 3) TK(16, 4) TK(16, 5) TK(16, 6) TK(16, 7) TK(16, 8)
            TK(16, 9) TK(16, 10) TK(16, 11) TK(16, 12) TK(16, 13) TK(16, 14) TK(16, 15) TK(16, 16)
        default: return nullptr;
        }
    }
    if (L == 32) {
        switch (MR) {
            TK(32, 2) TK(32, 3) TK(32, 4) TK(32, 5) TK(32, 6) TK(32, 7) TK(32, 8) TK(32, 9) TK(32, 10)
        default: return nullptr;
        }
    }
    if (L == 64) {
        switch (MR) {
            TK(64, 4) TK(64, 8) TK(64, 12) TK(64, 16)
        default: return nullptr;
        }
    }
#undef TK
#undef TKS
    return nullptr;
}

// Chain kernels exist for every table-kernel shape (the shared-product
// variants for 16 lanes).
void *fwd_chain_kernel_ptr(int L, int MR, bool inter, int S, bool share) {
#define TK2(m, sh) \
    case m: return inter ? reinterpret_cast<void *>(&fwd_chain_kernel<16, m, true, 2, sh>) : reinterpret_cast<void *>(&fwd_chain_kernel<16, m, false, 2, sh>);
    if (S == 2) {  // two reads per group: 16 lanes, 2..16 rows per lane
        if (L != 16) return nullptr;
        if (share) {
            switch (MR) {
                TK2(2, true) TK2(3, true) TK2(4, true) TK2(5, true) TK2(6, true) TK2(7, true) TK2(8, true)
                TK2(9, true) TK2(10, true) TK2(11, true) TK2(12, true) TK2(13, true) TK2(14, true)
                TK2(15, true) TK2(16, true)
            default: return nullptr;
            }
        }
        switch (MR) {
            TK2(2, false) TK2(3, false) TK2(4, false) TK2(5, false) TK2(6, false) TK2(7, false) TK2(8, false)
            TK2(9, false) TK2(10, false) TK2(11, false) TK2(12, false) TK2(13, false) TK2(14, false)
            TK2(15, false) TK2(16, false)
        default: return nullptr;
        }
    }
#undef TK2
    if (S != 1) return nullptr;
#define TK(l, m) \
    case m: return inter ? reinterpret_cast<void *>(&fwd_chain_kernel<l, m, true, 1, false>) : reinterpret_cast<void *>(&fwd_chain_kernel<l, m, false, 1, false>);
#define TKS(m) \
    case m: return inter ? reinterpret_cast<void *>(&fwd_chain_kernel<16, m, true, 1, true>) : reinterpret_cast<void *>(&fwd_chain_kernel<16, m, false, 1, true>);
    if (share) {
        if (L != 16) return nullptr;
        switch (MR) {
            TKS(1) TKS(2) TKS(3) TKS(4) TKS(5) TKS(6) TKS(7) TKS(8)
            TKS(9) TKS(10) TKS(11) TKS(12) TKS(13) TKS(14) TKS(15) TKS(16)
        default: return nullptr;
        }
    }
    if (L == 8) {
        switch (MR) {
            TK(8, 6) TK(8, 7) TK(8, 8) TK(8, 9) TK(8, 10) TK(8, 11) TK(8, 12)
        default: return nullptr;
        }
    }
    if (L == 16) {
        switch (MR) {
            TK(16, 1) TK(16, 2) TK(16, 3) TK(16, 4) TK(16, 5) TK(16, 6) TK(16, 7) TK(16, 8)
            TK(16, 9) TK(16, 10) TK(16, 11) TK(16, 12) TK(16, 13) TK(16, 14) TK(16, 15) TK(16, 16)
        default: return nullptr;
        }
    }
    if (L == 32) {
        switch (MR) {
            TK(32, 2) TK(32, 3) TK(32, 4) TK(32, 5) TK(32, 6) TK(32, 7) TK(32, 8) TK(32, 9) TK(32, 10)
        default: return nullptr;
        }
    }
    if (L == 64) {
        switch (MR) {
            TK(64, 4) TK(64, 8) TK(64, 12) TK(64, 16)
        default: return nullptr;
        }
    }
#undef TK
#undef TKS
    return nullptr;
}
uint32_t fwd_tab_bytes(int L, int MR) { return 4u * (uint32_t)L * (uint32_t)((MR + 3) / 4 * 4) * 4u; }

// FP64 rescue kernel: one pair per wave, 4 rows per lane (255-row strips).
// route: per pair intra or inter arithmetic as the FP64 regroup decided.
void *fwd_f64_rescue_kernel(bool route) {
    return route ? reinterpret_cast<void *>(&fwd_rescue_route_kernel)
                 : reinterpret_cast<void *>(&fwd_rescue_kernel);
}

// Waves per workgroup of the FP32 long-read kernel, and those the FP64
// rescue of the reads longer than one strip may use (kLW or kRescueLW).
int fwd_long_waves() { return kLW; }
int rescue_long_waves() { return kRescueLW; }
// tab: the LDS prior-table variant (every haplotype of the plan N-free)
void *fwd_long_kernel_ptr(int MR, bool inter, bool tab) {
#define LK(m)                                                                                \
    case m:                                                                                  \
        if (tab)                                                                             \
            return inter ? reinterpret_cast<void *>(&fwd_long_kernel<m, true, true>)        \
                         : reinterpret_cast<void *>(&fwd_long_kernel<m, false, true>);      \
        return inter ? reinterpret_cast<void *>(&fwd_long_kernel<m, true, false>)           \
                     : reinterpret_cast<void *>(&fwd_long_kernel<m, false, false>);
    switch (MR) {
        LK(4) LK(8) LK(12) LK(16)
    default: return nullptr;
    }
#undef LK
}
// LDS of the prior tables of a long-read workgroup of `waves` waves (added
// to its code buffer)
uint32_t long_tab_bytes(int MR, bool fp64, int waves) {
    const uint32_t e = fp64 ? 2u : 4u;
    return (uint32_t)waves * 4u * 64u * (((uint32_t)MR + e - 1) / e) * 16u;
}
hipError_t launch_fwd_block(void *kern, int grid, int block, size_t lds_bytes, hipStream_t st, const FwdArgs &a) {
    void *args[] = {const_cast<FwdArgs *>(&a)};
    return hipLaunchKernel(kern, dim3(grid), dim3(block), args, lds_bytes, st);
}
int rescue_long_rows() { return kRescueStrip; }
int rescue_long_mr() { return kLongRescueMR; }
template <int W>
void *rescue_long_kernel(bool route, bool tab) {
    return tab ? (route ? reinterpret_cast<void *>(&fwd_rescue_long_kernel<true, true, W>)
                        : reinterpret_cast<void *>(&fwd_rescue_long_kernel<false, true, W>))
               : (route ? reinterpret_cast<void *>(&fwd_rescue_long_kernel<true, false, W>)
                        : reinterpret_cast<void *>(&fwd_rescue_long_kernel<false, false, W>));
}
// waves: kLW or kRescueLW
hipError_t launch_rescue_long(bool route, bool tab, int waves, int grid, size_t lds_bytes, hipStream_t st,
                              const FwdArgs &a) {
    void *kern = waves == kRescueLW ? rescue_long_kernel<kRescueLW>(route, tab) : rescue_long_kernel<kLW>(route, tab);
    void *args[] = {const_cast<FwdArgs *>(&a)};
    return hipLaunchKernel(kern, dim3(grid), dim3(64 * waves), args, lds_bytes, st);
}

// FP64 rescue in classes: the binning pass, and the table pass over classes
// 4..7 (part 0) or 0..3 (part 1), each a persistent launch of one-wave
// workgroups; LDS per wave: two groups of lds_stride code bytes and a table
// of rescue_tab_bytes(highest MR of the part).
uint32_t rescue_tab_bytes(int MR) { return 4u * kRtL * (uint32_t)((MR + 1) / 2 * 2) * 8u; }
hipError_t launch_rescue_bin(const FwdArgs &a, int grid, hipStream_t st) {
    hipLaunchKernelGGL(rescue_bin_kernel, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_rescue_tab(int part, bool share, int grid, size_t lds_bytes, hipStream_t st,
                             const FwdArgs &a) {
    void *args[] = {const_cast<FwdArgs *>(&a)};
    void *k = share ? (part == 0 ? reinterpret_cast<void *>(&fwd_rescue_tab_kernel<4, 7, true>)
                                 : reinterpret_cast<void *>(&fwd_rescue_tab_kernel<0, 3, true>))
                    : (part == 0 ? reinterpret_cast<void *>(&fwd_rescue_tab_kernel<4, 7, false>)
                                 : reinterpret_cast<void *>(&fwd_rescue_tab_kernel<0, 3, false>));
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, lds_bytes, st);
}

hipError_t launch_regroup(const RegroupArgs &r, hipStream_t st) {
    if (r.n_regions == 0) return hipSuccess;
    hipLaunchKernelGGL(regroup_kernel, dim3(r.n_regions), dim3(64), 0, st, r);
    return hipGetLastError();
}
int fwd_f64_rescue_mr() { return 4; }

hipError_t launch_ext(int grid, size_t lds_bytes, hipStream_t st, const FwdArgs &a) {
    void *args[] = {const_cast<FwdArgs *>(&a)};
    return hipLaunchKernel(reinterpret_cast<void *>(&fwd_ext_kernel), dim3(grid), dim3(64), args, lds_bytes, st);
}

hipError_t launch_fwd(void *kern, int grid, size_t lds_bytes, hipStream_t st, const FwdArgs &a) {
    void *args[] = {const_cast<FwdArgs *>(&a)};
    return hipLaunchKernel(kern, dim3(grid), dim3(64), args, lds_bytes, st);
}

hipError_t launch_finalize(const FinalizeArgs &f, hipStream_t st) {
    if (f.n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t grid = (f.n + block - 1) / block;
    hipLaunchKernelGGL(finalize_kernel, dim3((uint32_t)grid), dim3(block), 0, st, f);
    return hipGetLastError();
}

namespace {
// Item descriptors: one thread per item slot.
__global__ void desc_kernel(FwdArgs a, const int32_t *items, ItemDesc *out, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t p = items[k];
    ItemDesc d;
    d.p = p;
    d.R = 0;
    d.H = 0;
    d.initd = 0.0f;
    d.ro = 0;
    d.hoff = 0;
    if (p >= 0) {
        const uint32_t h = a.pair_hap[p], r = a.pair_read[p];
        d.R = a.read_len[r];
        d.H = a.hap_len[h];
        d.initd = a.initd_f[h];
        d.ro = a.read_off[r];
        d.hoff = a.hapc_off[h];
    }
    out[k] = d;
}
}  // namespace

hipError_t launch_desc(const FwdArgs &a, const int32_t *items, ItemDesc *out, uint64_t n,
                       hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + 255) / 256;
    hipLaunchKernelGGL(desc_kernel, dim3((uint32_t)grid), dim3(256), 0, st, a, items, out, n);
    return hipGetLastError();
}

hipError_t launch_reset(uint32_t *counters, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(reset_kernel, dim3((n + 255) / 256), dim3(256), 0, st, counters, n);
    return hipGetLastError();
}

hipError_t launch_rescued_cells(const FwdArgs &a, unsigned long long *out, hipStream_t st) {
    hipLaunchKernelGGL(rescued_cells_kernel, dim3(256), dim3(256), 0, st, a.rescue_list, a.rescue_count,
                       a.pair_hap, a.pair_read, a.hap_len, a.read_len, out);
    return hipGetLastError();
}

}  // namespace rodv
