// phmm_adjoint.hip -- hand-written adjoint (reverse sweep) of the PairHMM
// forward on gfx950.  The reference has no gradient; this is the model of
// include/rodv_phmm.h (global gap-open delta and gap-continuation gamma,
// per-quality-bin scales of the base error probability).
//
// Reads of one strip: L = 16, 32 or 64 lanes per pair (64 / L pairs per
// wave, adj2_kernel), MR read rows per lane, the select-prior forward's lane
// and row layout (lane L-1 and `off` slots at the top of lane 0 hold
// identity rows, row R in the last slot of its lane).  Three passes:
//  1. forward over all anti-diagonal steps, saving the whole lane state every
//     `ckpt` steps to HBM (level-1 checkpoints);
//  2. for each checkpoint segment, last to first: recompute the segment from
//     its checkpoint, saving the state every 8 steps in LDS (level 2);
//  3. for each 8-step window, last to first, and inside it for each 4-step
//     half, last to first: recompute the 4 states into registers (the second
//     half from a copy advanced 4 steps) and run the adjoint of those steps in
//     reverse, lanes reading the row below their last row from lane l+1.
// Values cross lanes through ds_bpermute (LaneXfer), not DPP, so that two
// waves of a SIMD can co-issue the arithmetic.  Longer reads run in FP64
// strips, each pair's strips over several workgroups (adj_longab_kernel and
// adj_longc_kernel, below).
// Adjoint seeds are sigma = 2^-ilogb(S) so that adjoint x forward products
// stay O(expected counts); the gradient of log10 L follows at the end.
// Per real cell (i, j) with Tb = p * Mb (p the emission prior):
//   Mb = seed + mm_{i+1} Tb(i+1,j+1) + mx_{i+1} Ib(i+1,j) + my_i Db(i,j+1)
//   Ib = seed + gapm_{i+1} Tb(i+1,j+1) + xx_{i+1} Ib(i+1,j)
//   Db = gapm_{i+1} Tb(i+1,j+1) + yy_i Db(i,j+1)
//   dS/ddelta += -2 Tb M(i-1,j-1) + Ib M(i-1,j) + Db M(i,j-1)
//   dS/dgamma += -Tb (I(i-1,j-1) + D(i-1,j-1)) + Ib I(i-1,j) + Db D(i,j-1)
//   dS/de_i   += Mb t(i,j) (match ? -1 : 1/3)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phmm_device.h"
#include "phmm_internal.h"

namespace rodv {
namespace {

using namespace dev;
constexpr int kL = 64;
constexpr int kW = 4;  // steps per register window

template <typename T, int MR>
struct AdjRows {
    // transitions of the row below the lane's last row (0 across the wrap)
    T nb_mm, nb_g, nb_mx, nb_xx;
    // yy and my of the row above the lane's first row (step's D recurrence)
    T ab_yy, ab_my;
};

template <typename T> struct PTab;
template <> struct PTab<float> {
    static __device__ __forceinline__ float X(const AdjArgs &a, uint32_t q) { return a.xf[q]; }
    static __device__ __forceinline__ float Y(const AdjArgs &a, uint32_t q) { return a.yf[q]; }
    static __device__ __forceinline__ float mm(const AdjArgs &a) { return a.mm_f; }
    static __device__ __forceinline__ float g(const AdjArgs &a) { return a.g_f; }
    static __device__ __forceinline__ float dl(const AdjArgs &a) { return a.dl_f; }
    static __device__ __forceinline__ float gm(const AdjArgs &a) { return a.gm_f; }
    static __device__ __forceinline__ float initd(const AdjArgs &a, uint32_t h) { return a.initd_f[h]; }
};
template <> struct PTab<double> {
    static __device__ __forceinline__ double X(const AdjArgs &a, uint32_t q) { return a.xd[q]; }
    static __device__ __forceinline__ double Y(const AdjArgs &a, uint32_t q) { return a.yd[q]; }
    static __device__ __forceinline__ double mm(const AdjArgs &a) { return a.mm_d; }
    static __device__ __forceinline__ double g(const AdjArgs &a) { return a.g_d; }
    static __device__ __forceinline__ double dl(const AdjArgs &a) { return a.delta; }
    static __device__ __forceinline__ double gm(const AdjArgs &a) { return a.gamma; }
    static __device__ __forceinline__ double initd(const AdjArgs &a, uint32_t h) { return a.initd_d[h]; }
};

template <typename T, int MR>
__device__ __forceinline__ void put_state(T *base, int nvec, int k, const State<T, MR> &S) {
    // element v of state k of this lane at base[(k*nvec + v)*64 + lane]
    const int lane = threadIdx.x & 63;
    T *p = base + (size_t)k * nvec * kL + lane;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        p[(3 * s + 0) * kL] = S.M[s];
        p[(3 * s + 1) * kL] = S.I[s];
        p[(3 * s + 2) * kL] = S.D[s];
    }
    p[(3 * MR + 0) * kL] = S.Mp;
    p[(3 * MR + 1) * kL] = S.Ip;
    p[(3 * MR + 2) * kL] = S.Dp;
}

template <typename T, int MR>
__device__ __forceinline__ void get_state(const T *base, int nvec, int k, State<T, MR> &S) {
    const int lane = threadIdx.x & 63;
    const T *p = base + (size_t)k * nvec * kL + lane;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        S.M[s] = p[(3 * s + 0) * kL];
        S.I[s] = p[(3 * s + 1) * kL];
        S.D[s] = p[(3 * s + 2) * kL];
    }
    S.Mp = p[(3 * MR + 0) * kL];
    S.Ip = p[(3 * MR + 1) * kL];
    S.Dp = p[(3 * MR + 2) * kL];
}

// Adjoint of one forward step d for one lane.  Sb = state after step d-1,
// Sa = state after step d (states of step: Mp, Ip, Dp hold the row above
// one column back).  Adjoint registers hold the lane's values at its
// previous reverse step (column j+1); Tb0p holds slot 0's Tb one step older.
template <typename T, int MR>
__device__ __forceinline__ void adj_step(const State<T, MR> &Sb, const State<T, MR> &Sa,
                                         const Priors<T, MR> &P, const AdjRows<T, MR> &A,
                                         uint32_t hc, T seed, T (&Tb)[MR], T (&Ib)[MR],
                                         T (&Db)[MR], T &Tb0p, T (&gd)[MR], T (&gdT)[MR],
                                         T (&gg)[MR], T (&ge)[MR],
                                         const LaneXfer<kL, -1> &from_below) {
    // forward inputs of slot 0 (the row above): at column j-1 they are what
    // step d read (Sb's carry), at column j what step d carried on (Sa's)
    const T dgM0 = Sb.Mp, dgI0 = Sb.Ip, dgD0 = Sb.Dp;
    const T upM0 = Sa.Mp, upI0 = Sa.Ip;
    // adjoint inputs of the last slot from lane l+1
    T bT = from_below(Tb0p);   // Tb(i+1, j+1)
    T bI = from_below(Ib[0]);  // Ib(i+1, j)
    T b_mm = A.nb_mm, b_g = A.nb_g, b_mx = A.nb_mx, b_xx = A.nb_xx;
    T oldT0 = Tb[0];
#pragma unroll
    for (int s = MR - 1; s >= 0; s--) {
        const T sd = (s == MR - 1) ? seed : T(0);
        const T oldT = Tb[s];
        const T Mb = fmaT(P.my[s], Db[s], fmaT(b_mx, bI, fmaT(b_mm, bT, sd)));
        const T gT = b_g * bT;  // fma(b_g, bT, 0) where no seed enters: the same bits
        const T Ibn = fmaT(b_xx, bI, s == MR - 1 ? fmaT(b_g, bT, sd) : gT);
        const T Dbn = fmaT(P.yy[s], Db[s], gT);
        // forward values around cell (i, j)
        const T dgM = s == 0 ? dgM0 : Sb.M[s - 1];
        const T dgI = s == 0 ? dgI0 : Sb.I[s - 1];
        const T dgD = s == 0 ? dgD0 : Sb.D[s - 1];
        const T upM = s == 0 ? upM0 : Sa.M[s - 1];
        const T upI = s == 0 ? upI0 : Sa.I[s - 1];
        const T lfM = Sb.M[s], lfD = Sb.D[s];
        const T t = fmaT(P.g[s], dgD, fmaT(P.g[s], dgI, dgM * P.mm[s]));
        // match: the hap code has the read base's bit (bit 4, always set in
        // the staged codes, for N in the read); the prior and d e / d e_i
        // (-1 on a match, 1/3 otherwise) are both selected by that bit
        T pr, de;
        if constexpr (sizeof(T) == 4) {
            const int m = __builtin_amdgcn_sbfe((int)hc, P.rpos[s], 1);
            pr = __int_as_float((__float_as_int(P.X[s]) & m) | (__float_as_int(P.Y[s]) & ~m));
            de = __int_as_float((__float_as_int(-1.0f) & m) | (__float_as_int(1.0f / 3.0f) & ~m));
        } else {
            const bool match = ((hc >> P.rpos[s]) & 1u) != 0u;
            pr = match ? P.X[s] : P.Y[s];
            de = match ? T(-1) : T(1) / T(3);
        }
        const T Tbn = pr * Mb;
        // dS/ddelta = gd - 2 gdT (the -2 applied once, in the reduction)
        gdT[s] = fmaT(Tbn, dgM, gdT[s]);
        gd[s] = fmaT(Dbn, lfM, fmaT(Ibn, upM, gd[s]));
        gg[s] = fmaT(Dbn, lfD, fmaT(Ibn, upI, fmaT(-Tbn, dgI + dgD, gg[s])));
        ge[s] = fmaT(Mb * t, de, ge[s]);
        Tb[s] = Tbn;
        Ib[s] = Ibn;
        Db[s] = Dbn;
        // the slot above reads this slot as "below"
        bT = oldT;
        bI = Ibn;
        b_mm = P.mm[s];
        b_g = P.g[s];
        b_mx = P.mx[s];
        b_xx = P.xx[s];
    }
    Tb0p = oldT0;
}

template <typename T>
__device__ __forceinline__ T lane_value(T v, int src);  // value of lane `src` (uniform)
template <>
__device__ __forceinline__ float lane_value(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
template <>
__device__ __forceinline__ double lane_value(double v, int src) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                            __builtin_amdgcn_readlane(__double2loint(v), src));
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ------------------------------------------------- short reads, G pairs per wave
// The one-strip adjoint with L = 16, 32 or 64 lanes per pair (G = 64 / L
// pairs per wave, every per-pair scalar per group of lanes) and the model's
// transitions as wave-uniform scalars.  A read row has mm = 1 - 2 delta,
// g = 1 - gamma, mx = my = delta, xx = yy = gamma; an identity row differs
// in yy = 1 alone (its zero emission prior keeps its M at zero, and lane
// 0's row above, lane L-1's last slot, is taken as (M, I) = (0, 0) instead
// of rotated in, so the identity rows' I stay zero above the read).  Below
// row R the identity rows may carry nonzero I, which reaches no read row
// and, in the reverse sweep, gets no adjoint (their adjoints start at zero
// and nothing seeds them).  The identity slots above the read do get
// adjoints: their D = K/H and t enter the gamma and emission accumulators,
// which stay per slot and are masked to the read rows at the end; the
// delta accumulators see only zeros there (M = 0, Tb = 0) and are per lane.
// max over the G groups of a per-group value (each group's lanes hold it)
template <int G>
__device__ __forceinline__ int wave_max_g(int v) {
    int m = __builtin_amdgcn_readfirstlane(v);
#pragma unroll
    for (int k = 1; k < G; k++) m = max(m, __builtin_amdgcn_readlane(v, k * (64 / G)));
    return m;
}
// the value of lane src (per lane)
__device__ __forceinline__ float group_value(float v, int src) { return __shfl(v, src); }
__device__ __forceinline__ double group_value(double v, int src) { return __shfl(v, src); }

template <typename T>
struct UTrans {
    T mm, g, dl, gm;  // 1 - 2 delta, 1 - gamma, delta, gamma
    T third;          // 1/3, d e / d e_i on a mismatch
};

template <typename T, int MR>
struct URows {
    T yy[MR], X[MR], Y[MR];
    uint32_t rpos[MR];
};

// forward step of the uniform-transition model (the intra FMA order).  Every
// row's mx and my are delta, so the product upM * delta that row s's I takes
// is row s-1's M times delta, which row s-1's D takes in the next step: it
// is computed once and kept in Pm (Pm[s] == S.M[s] * delta between steps;
// the lane's last slot computes its own).
template <typename T, int MR>
__device__ __forceinline__ void shared_products(const State<T, MR> &S, T dl, T (&Pm)[MR]) {
#pragma unroll
    for (int s = 0; s < MR; s++) Pm[s] = S.M[s] * dl;
}

template <typename T, int L, int MR>
__device__ __forceinline__ void ustep(State<T, MR> &S, T (&Pm)[MR], const URows<T, MR> &P, const UTrans<T> &u,
                                      T yyA, uint32_t hc, const LaneXfer<L, 1> &rot, T keep) {
    // lane 0's row above is the top boundary: keep = 0 there, 1 elsewhere
    // (the rotated values are finite and >= 0, so the product is the select)
    const T upM0 = rot(S.M[MR - 1]) * keep, upI0 = rot(S.I[MR - 1]) * keep;
    T dgM = S.Mp, dgI = S.Ip, dgD = S.Dp;
    T upM = upM0, upI = upI0;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const T pr = select_prior(hc, P.rpos[s], P.X[s], P.Y[s]);
        T t = dgM * u.mm;
        t = fmaT(dgI, u.g, t);
        t = fmaT(dgD, u.g, t);
        const T nM = t * pr;
        const T up_dl = upM * u.dl;
        const T nI = fmaT(upI, u.gm, up_dl);
        const T nD = fmaT(S.D[s], P.yy[s], Pm[s]);
        if (s > 0) Pm[s - 1] = up_dl;
        dgM = S.M[s];
        dgI = S.I[s];
        dgD = S.D[s];
        upM = nM;
        upI = nI;
        S.M[s] = nM;
        S.I[s] = nI;
        S.D[s] = nD;
    }
    Pm[MR - 1] = S.M[MR - 1] * u.dl;
    S.Dp = fmaT(S.Dp, yyA, S.Mp * u.dl);
    S.Mp = upM0;
    S.Ip = upI0;
}

// Adjoint of one forward step (adj_step with uniform transitions); nb_* are
// the transitions into the row below the lane's last slot (0 for lane L-1,
// whose row below is the wrap).  gd, gdT per lane; gg, ge per slot.
template <typename T, int L, int MR>
__device__ __forceinline__ void adj_step_u(const State<T, MR> &Sb, const State<T, MR> &Sa,
                                           const URows<T, MR> &P, const UTrans<T> &u, T nb_mm, T nb_g,
                                           T nb_mx, T nb_xx, uint32_t hc, T seed, T (&Tb)[MR],
                                           T (&Ib)[MR], T (&Db)[MR], T &Tb0p, T &gd, T &gdT,
                                           T (&gg)[MR], T (&ge)[MR], const LaneXfer<L, -1> &from_below) {
    const T dgM0 = Sb.Mp, dgI0 = Sb.Ip, dgD0 = Sb.Dp;
    const T upM0 = Sa.Mp, upI0 = Sa.Ip;
    T bT = from_below(Tb0p);   // Tb(i+1, j+1)
    T bI = from_below(Ib[0]);  // Ib(i+1, j)
    T b_mm = nb_mm, b_g = nb_g, b_mx = nb_mx, b_xx = nb_xx;
    const T oldT0 = Tb[0];
#pragma unroll
    for (int s = MR - 1; s >= 0; s--) {
        const T sd = (s == MR - 1) ? seed : T(0);
        const T oldT = Tb[s];
        const T Mb = fmaT(u.dl, Db[s], fmaT(b_mx, bI, fmaT(b_mm, bT, sd)));
        const T gT = b_g * bT;  // fma(b_g, bT, 0) where no seed enters: the same bits
        const T Ibn = fmaT(b_xx, bI, s == MR - 1 ? fmaT(b_g, bT, sd) : gT);
        const T Dbn = fmaT(P.yy[s], Db[s], gT);
        const T dgM = s == 0 ? dgM0 : Sb.M[s - 1];
        const T dgI = s == 0 ? dgI0 : Sb.I[s - 1];
        const T dgD = s == 0 ? dgD0 : Sb.D[s - 1];
        const T upM = s == 0 ? upM0 : Sa.M[s - 1];
        const T upI = s == 0 ? upI0 : Sa.I[s - 1];
        const T lfM = Sb.M[s], lfD = Sb.D[s];
        const T sID = dgI + dgD;
        const T t = fmaT(u.g, sID, dgM * u.mm);
        T pr, de;
        if constexpr (sizeof(T) == 4) {
            const int m = __builtin_amdgcn_sbfe((int)hc, P.rpos[s], 1);
            pr = __int_as_float((__float_as_int(P.X[s]) & m) | (__float_as_int(P.Y[s]) & ~m));
            de = __int_as_float((__float_as_int(-1.0f) & m) | (__float_as_int(u.third) & ~m));
        } else {
            const bool match = ((hc >> P.rpos[s]) & 1u) != 0u;
            pr = match ? P.X[s] : P.Y[s];
            de = match ? T(-1) : u.third;
        }
        const T Tbn = pr * Mb;
        gdT = fmaT(Tbn, dgM, gdT);
        gd = fmaT(Dbn, lfM, fmaT(Ibn, upM, gd));
        gg[s] = fmaT(Dbn, lfD, fmaT(Ibn, upI, fmaT(-Tbn, sID, gg[s])));
        ge[s] = fmaT(Mb * t, de, ge[s]);
        Tb[s] = Tbn;
        Ib[s] = Ibn;
        Db[s] = Dbn;
        bT = oldT;
        bI = Ibn;
        b_mm = u.mm;
        b_g = u.g;
        b_mx = u.dl;
        b_xx = u.gm;
    }
    Tb0p = oldT0;
}

// a wave-uniform value held in a VGPR: an instruction with an SGPR operand
// (a value or a select's condition) does not co-issue with another wave's
// VALU instruction on gfx950
__device__ __forceinline__ float in_vgpr(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

template <typename T> struct UT;
template <> struct UT<float> {
    static __device__ __forceinline__ UTrans<float> get(const AdjArgs &a) {
        return UTrans<float>{in_vgpr(a.mm_f), in_vgpr(a.g_f), in_vgpr(a.dl_f), in_vgpr(a.gm_f),
                             in_vgpr(1.0f / 3.0f)};
    }
};
template <> struct UT<double> {
    static __device__ __forceinline__ UTrans<double> get(const AdjArgs &a) {
        return UTrans<double>{a.mm_d, a.g_d, a.delta, a.gamma, 1.0 / 3.0};
    }
};

// G pairs per call (p < 0: an empty group).  Wave-uniform loops run to the
// largest group's bounds; a group past its own is masked (its sums) or
// harmless (its adjoints are zero outside its columns).
template <typename T, int L, int MR, bool RESCUE>
__device__ void adj2_item(const AdjArgs &a, uint8_t *lds, int p) {
    constexpr int G = 64 / L;
    constexpr int NV = 3 * MR + 3;
    constexpr int C = (L - 1) * MR;  // rows of the strip (lane L-1 holds identity rows)
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
    const int off = valid ? (MR - R % MR) % MR : 0;
    const int lR = valid ? (R - 1 + off) / MR : 0;
    const T initd = valid ? PTab<T>::initd(a, h) : T(0);
    const UTrans<T> u = UT<T>::get(a);

    // hap codes of group g: column j at lds[g * lds_stride + L - 1 + j]
    __syncthreads();
    for (uint32_t k = lane * 16; k < (uint32_t)G * a.lds_stride; k += 64 * 16)
        *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    __syncthreads();
    if (valid) {
        const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[h]);
        const int nch = (H + 15) >> 4;
        // one-hot codes with bit 4 set: a read N (row bit 4) matches every column
        for (int c = l; c < nch; c += L) {
            const uint4 v = src[c];
            *reinterpret_cast<uint4 *>(lds + g * a.lds_stride + L + 16 * c) =
                make_uint4(v.x | 0x10101010u, v.y | 0x10101010u, v.z | 0x10101010u, v.w | 0x10101010u);
        }
    }
    __syncthreads();
    const uint8_t *cp = lds + g * a.lds_stride + (L - 1) - l;

    URows<T, MR> P;
    State<T, MR> S0;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = l * MR + s - off + 1;
        const bool real = valid && l < L - 1 && i >= 1 && i <= R;
        const uint64_t idx = ro + (uint64_t)(real ? i - 1 : 0);
        const uint32_t qq = a.q[idx] & 127u;
        const uint32_t code = a.rcode[idx];
        T X = PTab<T>::X(a, qq), Y = PTab<T>::Y(a, qq);
        Y = code == kReadN ? X : Y;
        P.yy[s] = real ? u.gm : T(1);
        P.X[s] = real ? X : T(0);
        P.Y[s] = real ? Y : T(0);
        P.rpos[s] = real ? code : 0u;
        S0.M[s] = T(0);
        S0.I[s] = T(0);
        S0.D[s] = real ? T(0) : initd;
    }
    const LaneXfer<L, 1> up;
    const LaneXfer<L, -1> down;
    const T lane0 = l == 0 ? T(0) : T(1);  // ustep's keep
    S0.Mp = T(0);
    S0.Ip = T(0);
    S0.Dp = up(S0.D[MR - 1]);
    const T yyA = up(P.yy[MR - 1]);
    const bool wrap = l == L - 1;
    const T nb_mm = wrap ? T(0) : u.mm, nb_g = wrap ? T(0) : u.g;
    const T nb_mx = wrap ? T(0) : u.dl, nb_xx = wrap ? T(0) : u.gm;

    // ---- pass 1: forward with level-1 checkpoints
    const int Dend = valid ? H + lR : 0;  // lane lR reaches column H
    const int D4 = (wave_max_g<G>(Dend) + kW - 1) / kW * kW;
    const int K1 = a.ckpt;
    const int nseg = (D4 + K1 - 1) / K1;
    T *ck = reinterpret_cast<T *>(a.ck) + (uint64_t)blockIdx.x * a.ck_elems;
    State<T, MR> S = S0;
    T Pm[MR];
    shared_products(S, u.dl, Pm);
    T sm = T(0), si = T(0);
    put_state(ck, NV, 0, S);
    for (int d = 1; d <= D4; d++) {
        ustep<T, L, MR>(S, Pm, P, u, yyA, cp[d], up, lane0);
        if (d <= Dend && d - l <= H) {
            sm += S.M[MR - 1];
            si += S.I[MR - 1];
        }
        if (d % K1 == 0 && d < D4) put_state(ck, NV, d / K1, S);
    }
    const T Ssum = group_value(sm + si, g * L + lR);
    const bool under = valid && (RESCUE ? !(Ssum > T(0)) : !((float)Ssum >= 1e-28f));
    if (under && l == 0) {
        if (!RESCUE) {  // FP32 underflow: the FP64 instance redoes this pair
            a.flags[p] = 0x01;
            const uint32_t k = atomicAdd(a.rescue_count, 1u);
            a.rescue_list[k] = (uint32_t)p;
        } else {        // FP64 underflow too: no gradient
            a.out_log10[p] = -INFINITY;
            for (int t = 0; t < 10; t++) a.out_grad[10 * (uint64_t)p + t] = 0.0;
            a.flags[p] = 0x03;
        }
    }
    const bool live = valid && !under;
    if (!__builtin_amdgcn_readfirstlane(__ballot(live) != 0ull ? 1 : 0)) return;

    // ---- passes 2 and 3: the reverse sweep (seed 0 in the groups left out)
    const int e = live ? ilogb((double)Ssum) : 0;
    const T sigma = live ? (T)ldexp(1.0, -e) : T(0);
    T *l2 = reinterpret_cast<T *>(lds + G * a.lds_stride);  // level-2 states
    T Tb[MR], Ib[MR], Db[MR], gg[MR], ge[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) Tb[s] = Ib[s] = Db[s] = gg[s] = ge[s] = T(0);
    T Tb0p = T(0), gd = T(0), gdT = T(0);
    // Pw: the shared products of W0 (advanced through the window)
    auto window4 = [&](const State<T, MR> &W0, int dw, T (&Pw)[MR]) {
        State<T, MR> W1, W2, W3, W4;
        W1 = W0;
        ustep<T, L, MR>(W1, Pw, P, u, yyA, cp[dw + 1], up, lane0);
        W2 = W1;
        ustep<T, L, MR>(W2, Pw, P, u, yyA, cp[dw + 2], up, lane0);
        W3 = W2;
        ustep<T, L, MR>(W3, Pw, P, u, yyA, cp[dw + 3], up, lane0);
        W4 = W3;
        ustep<T, L, MR>(W4, Pw, P, u, yyA, cp[dw + 4], up, lane0);
        auto seed_at = [&](int d) {
            const int j = d - l;
            return (l == lR && j >= 1 && j <= H) ? sigma : T(0);
        };
        adj_step_u<T, L, MR>(W3, W4, P, u, nb_mm, nb_g, nb_mx, nb_xx, cp[dw + 4], seed_at(dw + 4), Tb, Ib, Db,
                             Tb0p, gd, gdT, gg, ge, down);
        adj_step_u<T, L, MR>(W2, W3, P, u, nb_mm, nb_g, nb_mx, nb_xx, cp[dw + 3], seed_at(dw + 3), Tb, Ib, Db,
                             Tb0p, gd, gdT, gg, ge, down);
        adj_step_u<T, L, MR>(W1, W2, P, u, nb_mm, nb_g, nb_mx, nb_xx, cp[dw + 2], seed_at(dw + 2), Tb, Ib, Db,
                             Tb0p, gd, gdT, gg, ge, down);
        adj_step_u<T, L, MR>(W0, W1, P, u, nb_mm, nb_g, nb_mx, nb_xx, cp[dw + 1], seed_at(dw + 1), Tb, Ib, Db,
                             Tb0p, gd, gdT, gg, ge, down);
    };
    for (int k = nseg - 1; k >= 0; k--) {
        const int d0 = k * K1;
        const int nsteps = min(K1, D4 - d0);  // a multiple of 4
        const int nwin = (nsteps + 2 * kW - 1) / (2 * kW);
        get_state(ck, NV, k, S);
        shared_products(S, u.dl, Pm);
        __syncthreads();
        // the window starts; the last window's steps are its own pass 3's
        for (int w = 0; w < nwin; w++) {
            put_state(l2, NV, w, S);
            if (w == nwin - 1) break;
            for (int v = 0; v < 2 * kW; v++)
                ustep<T, L, MR>(S, Pm, P, u, yyA, cp[d0 + w * 2 * kW + v + 1], up, lane0);
        }
        __syncthreads();
        for (int w = nwin - 1; w >= 0; w--) {
            const int dw = d0 + w * 2 * kW;
            const int n = min(2 * kW, nsteps - w * 2 * kW);
            // the window's start state is read from LDS for each half, so
            // that it is not held through the second half's window
            if (n == 2 * kW) {
                State<T, MR> X;
                get_state(l2, NV, w, X);
                shared_products(X, u.dl, Pm);
#pragma unroll
                for (int v = 1; v <= kW; v++) ustep<T, L, MR>(X, Pm, P, u, yyA, cp[dw + v], up, lane0);
                window4(X, dw + kW, Pm);
            }
            State<T, MR> W0;
            get_state(l2, NV, w, W0);
            shared_products(W0, u.dl, Pm);
            window4(W0, dw, Pm);
        }
    }

    // ---- reduce over the group's read rows and lanes
    T vg = T(0), vb[8];
#pragma unroll
    for (int b = 0; b < 8; b++) vb[b] = T(0);
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = l * MR + s - off + 1;
        const bool rl = valid && l < L - 1 && i >= 1 && i <= R;
        const uint32_t qq = a.q[ro + (uint64_t)(rl ? i - 1 : 0)] & 127u;
        const T real = rl ? T(1) : T(0);
        const uint32_t bin = a.qbin[qq];
        vg = fmaT(real, gg[s], vg);
        const T c = real * ge[s] * (T)a.wq[qq];
#pragma unroll
        for (int b = 0; b < 8; b++) vb[b] += (bin == (uint32_t)b) ? c : T(0);
    }
    double acc[10];
    acc[0] = (double)fmaT(T(-2), gdT, gd);
    acc[1] = (double)vg;
#pragma unroll
    for (int b = 0; b < 8; b++) acc[2 + b] = (double)vb[b];
#pragma unroll
    for (int t = 0; t < 10; t++) {
        double v = acc[t];
        for (int o = L / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o);
        acc[t] = v;
    }
    if (live && l == 0) {
        // scaled adjoints: acc = sigma * dS/dtheta;  d log10 L = dS / (S ln 10)
        const double sS = (double)sigma * (double)Ssum;
        const double ln10 = 2.302585092994045684;
        const double K_log10 = RESCUE ? 1020.0 * 0.30102999566398119521 : 120.0 * 0.30102999566398119521;
        a.out_log10[p] = log10((double)Ssum) - K_log10;
        double *gr = a.out_grad + 10 * (uint64_t)p;
        gr[0] = -a.delta * acc[0] / (10.0 * sS);
        gr[1] = -a.gamma * acc[1] / (10.0 * sS);
        for (int b = 0; b < 8; b++) gr[2 + b] = acc[2 + b] / (sS * ln10);
        a.flags[p] = RESCUE ? 0x01 : 0;
    }
}

template <typename T, int L, int MR, bool RESCUE>
__global__ __launch_bounds__(64, RESCUE ? 1 : 2) void adj2_kernel(AdjArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int G = 64 / L;
    const uint32_t n = RESCUE ? (*a.n_items_dev + G - 1) / G : a.n_items;
    const uint32_t nlist = RESCUE ? *a.n_items_dev : 0u;
    // an empty FP64 list (the usual case) ends every workgroup before it
    // queues on the shared counter's atomic
    if (__builtin_amdgcn_readfirstlane(n) == 0) return;
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.counter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        const uint32_t k = it * G + threadIdx.x / L;
        int p;
        if constexpr (RESCUE)
            p = k < nlist ? (int)a.rescue_list[k] : -1;
        else
            p = a.items[k];
        adj2_item<T, L, MR, RESCUE>(a, lds, p);
    }
}

// ------------------------------------------------------------ long reads
// Reads longer than one strip of 63 * MR rows, in FP64 (their FP32 sums
// underflow), in two passes over the strips (run by the launches described
// before adj_longa_kernel, or for tests on one workgroup, adj_long_kernel):
//  A. forward over every strip, storing each strip's last row in HBM; the
//     row a strip hands to the next is scaled by a power of two (kept per
//     strip) that takes the largest entry of the row four strips back, in
//     the same scale, to 2^kLongTop, so the scaled entries stay in range
//     however long the read is (as the extended-range forward does);
//  B. strips last to first: recompute the strip's forward from its input
//     row with the two levels of checkpoints of the one-strip kernel, then
//     its reverse sweep.  Below lane 62's last row lies the next strip's
//     first row: before each reverse step lane 63 presents that row's
//     adjoints (Tb one column right, Ib at the column), which the strip
//     below handed up, scaled by the shift between the two strips (the
//     adjoints of scaled entries scale inversely, so adjoint x forward
//     products share one scale across strips); lane 0 hands its first row's
//     adjoints up in turn.  Each strip's per-row gradient accumulators are
//     reduced into per-lane totals (by the strip's rows and quality bins)
//     before the next.
// The feeder's inputs and the adjoints from below are staged in LDS per
// checkpoint segment (one coalesced load each), not loaded per step.
constexpr int kLongMR = 4;
constexpr int kLongTop = 960;
constexpr int kMaxLongStrips = 64;

template <typename T>
__device__ __forceinline__ T ld_ag(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_ag(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Priors, row transitions and initial state of strip cs (parametrised model).
template <int MR>
__device__ __forceinline__ void long_setup(const AdjArgs &a, int cs, int nstr, int R, int off, uint64_t ro,
                                           double initd, Priors<double, MR> &P, AdjRows<double, MR> &A,
                                           State<double, MR> &S0) {
    constexpr int C = (kL - 1) * MR;
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = cs * C + l * MR + s - off + 1;
        const bool real = l < kL - 1 && i >= 1 && i <= R;
        const uint64_t idx = ro + (uint64_t)(real ? i - 1 : 0);
        const uint32_t qq = a.q[idx] & 127u;
        const uint32_t code = a.rcode[idx];
        double X = a.xd[qq], Y = a.yd[qq];
        Y = code == kReadN ? X : Y;  // N in the read: the match prior
        P.mm[s] = real ? a.mm_d : 0.0;
        P.g[s] = real ? a.g_d : 0.0;
        P.mx[s] = real ? a.delta : 0.0;
        P.my[s] = real ? a.delta : 0.0;
        P.xx[s] = real ? a.gamma : 0.0;
        P.yy[s] = real ? a.gamma : 1.0;
        P.X[s] = real ? X : 0.0;
        P.Y[s] = real ? Y : 0.0;
        P.rpos[s] = real ? code : 0u;
        S0.M[s] = 0.0;
        S0.I[s] = 0.0;
        S0.D[s] = real ? 0.0 : initd;
    }
    const LaneXfer<kL, 1> up;
    S0.Mp = 0.0;
    S0.Ip = 0.0;
    S0.Dp = up(S0.D[MR - 1]);
    A.ab_yy = up(P.yy[MR - 1]);
    A.ab_my = up(P.my[MR - 1]);
    if (cs > 0 && l == 0) {  // the row above is the previous strip's last row
        A.ab_yy = a.gamma;
        A.ab_my = a.delta;
        S0.Dp = 0.0;
    }
    const LaneXfer<kL, -1> down;
    const double m = down(P.mm[0]), g = down(P.g[0]), x = down(P.mx[0]), xx = down(P.xx[0]);
    const bool wrap = l == kL - 1;
    // below lane L-2's last row: the next strip's first row
    const bool below = l == kL - 2 && cs < nstr - 1;
    A.nb_mm = wrap ? 0.0 : (below ? a.mm_d : m);
    A.nb_g = wrap ? 0.0 : (below ? a.g_d : g);
    A.nb_mx = wrap ? 0.0 : (below ? a.delta : x);
    A.nb_xx = wrap ? 0.0 : (below ? a.gamma : xx);
}

// Waves of a long-read workgroup: one per SIMD.  Pass A deals the strips to
// the waves in order (wave w takes strips w, w + W, ...), each strip's
// feeder waiting for the strip above to publish the columns of each block.
// Pass B deals the strips to the waves, last first: wave w takes strips
// nstr-1-w, nstr-1-w-W, ...  A strip's reverse reads the adjoints the strip
// below hands up, which that strip writes column by column as its own
// reverse proceeds from the last column.  Before staging the columns of a
// checkpoint segment, a wave waits until the strip below has published (in
// LDS: prog[cs + 1], the lowest step it has finished) every column the
// segment reads; its forward recomputation (pass 1) waits for nothing.
// The strips add their gradient sums into the per-lane totals in the order
// one wave would (the strip below first, waiting for its prog = 0), so the
// results are the same bits for any number of waves.  The handed-up rows go
// through W + 1 HBM rows per workgroup: strip cs writes row cs % (W + 1),
// which strip cs + W + 1 wrote last and strip cs + W, the previous strip of
// the same wave, read last.  Every wait is bounded (a.spin_limit polls);
// one that runs out counts in *a.fault, the strip's handed-up adjoints
// become NaN and its wave waits no more.
constexpr int kLongWaves = 4;
constexpr int kProgNone = 0x7fffffff;
constexpr int kNoMax = -0x40000000;  // a strip's last row is all zero
// Pass A's shifts: strip cs's input row (the last row of strip cs - 1) is
// scaled by 2^shs[cs], chosen from the largest entry of strip
// cs - kShiftLag's last row as the shifts since bring it to strip cs - 1's
// scale: 2^kLongTop if the three strips between change nothing.  Strip
// cs - kShiftLag ran on the same wave (kShiftLag a multiple of the waves),
// so its row is complete when strip cs starts, and 4 strips run at once
// instead of one after another.  A row's total over columns and states grows
// by less than 1 % over a strip (the transitions out of a state sum to at
// most 1, emissions are at most 1) and lies between the row's largest entry
// and 3H times it, so scaled entries stay below 2^(kLongTop + 2 log2(3H) +
// 1) < 2^1023; below, they sit about three strips' decay under
// 2^kLongTop.  Powers of two scale exactly, so the results are those of any
// other choice of shifts except in subnormal entries.
constexpr int kShiftLag = 4;
constexpr int kSafeLog = -500;  // below: pass A again with a shift per strip
constexpr int kPubSegs = 4;     // pass B publishes its progress every 4 checkpoint segments

// LDS of a long-read workgroup: hap codes (lds_stride bytes) | per wave:
// level-2 states, feeder inputs of a segment, adjoints from below of a
// segment | per-lane gradient totals (10 x 64) | scaled sum | the strips'
// shifts | the strips' reverse progress | their forward progress | the
// exponents of their last rows' largest entries | the item taken
__host__ __device__ constexpr uint64_t long_wave_elems(int K1, int MR) {
    return (uint64_t)((K1 + 7) / 8) * (3 * MR + 3) * kL + 3ull * K1 + 2ull * (K1 + 1);
}
__host__ __device__ constexpr uint64_t long_lds_bytes(uint32_t lds_stride, int K1, int MR, int W) {
    return lds_stride + ((uint64_t)W * long_wave_elems(K1, MR) + 10ull * kL + 1) * 8 + 16ull * kMaxLongStrips + 16;
}

// Waits until the published progress *c is <= need, the value seen to
// `seen`; false if `limit` polls (of about half a microsecond each) did not
// see it.  Wave-uniform.
__device__ __forceinline__ bool wait_progress(const int *c, int need, uint32_t limit, int &seen) {
    for (uint32_t k = 0; k < limit; k++) {
        const int v = uni(__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (v <= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // the handed-up rows after the count
            seen = v;
            return true;
        }
        __builtin_amdgcn_s_sleep(16);
    }
    return false;
}
// Waits until the published progress *c is >= need (pass A's forward
// progress); as wait_progress.
__device__ __forceinline__ bool wait_progress_ge(const int *c, int need, uint32_t limit) {
    for (uint32_t k = 0; k < limit; k++) {
        const int v = uni(__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (v >= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // the row's stores after the count
            return true;
        }
        __builtin_amdgcn_s_sleep(16);
    }
    return false;
}
__device__ __forceinline__ void publish_progress(int *c, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the handed-up rows before the count
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// LDS written by other lanes of this wave, before it is read (one wave's
// LDS operations complete in order)
__device__ __forceinline__ void wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// A long pair's per-item scalars (wave-uniform).
struct LongQ {
    int p;
    uint32_t h;
    int R, H, off, nstr, lR;
    uint64_t ro;
    double initd;
};
template <int MR>
__device__ __forceinline__ LongQ long_q(const AdjArgs &a, int p_in) {
    constexpr int C = (kL - 1) * MR;
    LongQ q;
    q.p = uni(p_in);
    q.h = (uint32_t)uni((int)a.pair_hap[q.p]);
    const uint32_t r = (uint32_t)uni((int)a.pair_read[q.p]);
    q.R = uni((int)a.read_len[r]);
    q.H = uni((int)a.hap_len[q.h]);
    const uint64_t ro64 = a.read_off[r];
    q.ro = ((uint64_t)(uint32_t)uni((int)(ro64 >> 32)) << 32) | (uint64_t)(uint32_t)uni((int)(uint32_t)ro64);
    q.off = (MR - q.R % MR) % MR;
    const int vR = q.R - 1 + q.off;
    q.nstr = vR / C + 1;  // <= kMaxLongStrips (checked by the host)
    q.lR = (vR % C) / MR;
    q.initd = a.initd_d[q.h];
    return q;
}

// The haplotype's codes in LDS, column j at lds[L - 1 + j], staged by the
// workgroup's nt threads (one-hot, bit 4 set: a read N matches every
// column); begins and ends with a workgroup barrier.
__device__ __forceinline__ void long_codes(const AdjArgs &a, uint8_t *lds, const LongQ &q, int nt) {
    __syncthreads();  // the previous item's LDS reads are done
    for (uint32_t k = threadIdx.x * 16; k < a.lds_stride; k += (uint32_t)nt * 16)
        *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[q.h]);
    const int nch = (q.H + 15) >> 4;
    __syncthreads();
    for (int c = threadIdx.x; c < nch; c += nt) {
        const uint4 v = src[c];
        *reinterpret_cast<uint4 *>(lds + kL + 16 * c) =
            make_uint4(v.x | 0x10101010u, v.y | 0x10101010u, v.z | 0x10101010u, v.w | 0x10101010u);
    }
    __syncthreads();
}

// Pass B's codes, two per byte: column j at nibble L - 1 + j (low nibble
// first), the one-hot codes (N = 0xf) without the bit that makes a read N
// match, which long_strip_b puts back on every nonzero code (the zeros past
// the haplotype stay zero, as in the byte codes); by one wave on its own
// (one wave's LDS operations complete in order).  Half the LDS of the byte
// codes, so that four pass-B waves fit a CU.
__host__ __device__ constexpr uint32_t packed_stride(uint32_t lds_stride) { return ((lds_stride + 1) / 2 + 15) & ~15u; }
__device__ __forceinline__ uint32_t pack_nibbles(uint32_t x) {  // 4 codes < 16 -> 2 bytes
    return (x & 0xfu) | ((x >> 4) & 0xf0u) | ((x >> 8) & 0xf00u) | ((x >> 12) & 0xf000u);
}
__device__ __forceinline__ void long_codes_packed(const AdjArgs &a, uint8_t *lds, const LongQ &q) {
    const uint32_t l = threadIdx.x & 63;
    wave_sync();  // the previous item's LDS reads are done
    const uint32_t ps = packed_stride(a.lds_stride);
    for (uint32_t k = l * 16; k < ps; k += 64 * 16) *reinterpret_cast<uint4 *>(lds + k) = make_uint4(0, 0, 0, 0);
    const uint4 *src = reinterpret_cast<const uint4 *>(a.hapc + a.hapc_off[q.h]);
    const int nch = (q.H + 15) >> 4;
    wave_sync();
    for (int c = (int)l; c < nch; c += 64) {  // columns 16c + 1 .. 16c + 16 at bytes 32 + 8c ..
        const uint4 v = src[c];
        *reinterpret_cast<uint2 *>(lds + kL / 2 + 8 * c) =
            make_uint2(pack_nibbles(v.x) | pack_nibbles(v.y) << 16, pack_nibbles(v.z) | pack_nibbles(v.w) << 16);
    }
    wave_sync();
}

// Elements of one strip's level-1 checkpoints (every K1 steps, the state at
// step 0 first) for a haplotype of H columns: an upper bound of the steps a
// strip's reverse covers, H + 63 rounded up to 4.
__host__ __device__ constexpr uint64_t long_ck_elems(uint32_t H, int K1, int MR) {
    return ((((uint64_t)H + 64 + 3) / 4 * 4) / (uint64_t)K1 + 2) * (3ull * MR + 3) * kL;
}

// Pass A over the strips from strip `first` in steps of `stride`, the
// shifts taken `lag` strips back; a strip's feeder waits for the strip
// above to publish the columns of each block (LDS progress, bounded waits).
// Each strip's last row to B (row cs at cs * lrow), its shift to shs[cs],
// the exponent of its largest entry to mlog[cs]; the last strip's sum to
// *ssum.  ck1 (may be null): each strip's level-1 checkpoints for pass B,
// strip cs's from ck1 + cs * long_ck_elems(H): the lane state at step 0 and
// every K1 steps below the last step pass B's reverse covers, the states
// pass B's own forward would checkpoint (the same step function, inputs and
// order), so pass B skips that forward.
template <int MR>
__device__ void long_pass_a(const AdjArgs &a, const LongQ &q, const uint8_t *cp, double *B, int *shs, int *progA,
                            int *mlog, double *ssum, const int first, const int stride, const int lag,
                            bool &liveA, double *ck1) {
    using T = double;
    const int l = threadIdx.x & 63;
    const int lane = l;
    const int R = q.R, H = q.H, off = q.off, nstr = q.nstr, lR = q.lR;
    const uint64_t ro = q.ro;
    const T initd = q.initd;
    const LaneXfer<kL, 1> up;
    for (int cs = first; cs < nstr; cs += stride) {
        Priors<T, MR> P;
        AdjRows<T, MR> A;
        State<T, MR> S;
        long_setup<MR>(a, cs, nstr, R, off, ro, initd, P, A, S);
        const bool last = cs == nstr - 1;
        const T *bin = B + (uint64_t)(cs > 0 ? cs - 1 : 0) * a.lrow;
        T *bout = B + (uint64_t)cs * a.lrow;
        int sh = 0;
        if (cs > 0) {
            // the row above up to column 1 (written at the strip above's step 63)
            if (liveA) {
                liveA = wait_progress_ge(progA + cs - 1, 63, a.spin_limit);
                if (!liveA && l == 0) atomicAdd(a.fault, 1u);
            }
            // the shift: strip cs - kShiftLag's last row, brought to the scale
            // of the row above through the shifts since (each published with
            // its strip's first progress), goes to 2^kLongTop
            const int j = cs - lag;
            if (j >= 0 && mlog[j] != kNoMax) {
                int since = 0;
                for (int i = j + 1; i < cs; i++) since += shs[i];
                sh = kLongTop - mlog[j] - since;
            }
            if (lane == 0) shs[cs] = sh;
            // column 1 of the row above (column j at 3 * (j + L))
            const T bm = ldexp(ld_ag(bin + 3 * (1 + kL)), sh), bi = ldexp(ld_ag(bin + 3 * (1 + kL) + 1), sh),
                    bd = ldexp(ld_ag(bin + 3 * (1 + kL) + 2), sh);
            if (l == kL - 1) {
                S.M[MR - 1] = bm;
                S.I[MR - 1] = bi;
                S.D[MR - 1] = bd;
            }
        } else if (lane == 0) {
            shs[0] = 0;
        }
        // pass B's checkpoints of this strip: up to the step before its
        // reverse's last (a multiple of 4)
        constexpr int NV = 3 * MR + 3;
        const int K1 = a.ckpt;
        const int D4b = ((last ? H + lR : H + kL - 2) + kW - 1) / kW * kW;
        T *cks = ck1 ? ck1 + (uint64_t)cs * long_ck_elems((uint32_t)H, K1, MR) : nullptr;
        int nck = K1;  // the next checkpoint's step
        if (cks) put_state(cks, NV, 0, S);
        // a strip that is not the last stores its last row up to column
        // H + 69, past every column the recomputations' feeders read (D4 + 1)
        const int Dend = last ? H + lR : H + 2 * kL + 3;
        T bmax = T(0), sm = T(0), si = T(0);
        for (int d0 = 1; d0 <= Dend; d0 += kL) {
            const int dlim = d0 + kL - 1 < Dend ? d0 + kL - 1 : Dend;
            // the block's feeder columns d0 + 1 .. dlim + 1, written by the
            // strip above up to its step dlim + 63
            if (cs > 0 && liveA) {
                liveA = wait_progress_ge(progA + cs - 1, dlim + 63, a.spin_limit);
                if (!liveA && l == 0) atomicAdd(a.fault, 1u);
            }
            T fm = T(0), fi = T(0), fd = T(0);
            if (cs > 0 && d0 + l <= dlim) {
                const T *in = bin + 3 * (d0 + l + 1 + kL);
                fm = ldexp(ld_ag(in), sh);
                fi = ldexp(ld_ag(in + 1), sh);
                fd = ldexp(ld_ag(in + 2), sh);
            }
            if (!liveA) fm = T(__builtin_nan(""));
            // one step; two per iteration, so that the state's registers
            // alternate instead of being copied back after every step
            auto one_step = [&](const int d) {
                step<T, kL, MR, false>(S, P, A.ab_yy, A.ab_my, cp[d], up);
                if (last && d - l <= H) {
                    sm += S.M[MR - 1];
                    si += S.I[MR - 1];
                }
                if (!last && l == kL - 2) {  // column d - (L-2) at 3 * (d + 2)
                    T *o = bout + 3 * (d + 2);
                    st_ag(o, S.M[MR - 1]);
                    st_ag(o + 1, S.I[MR - 1]);
                    st_ag(o + 2, S.D[MR - 1]);
                    bmax = fmax(bmax, fmax(S.M[MR - 1], fmax(S.I[MR - 1], S.D[MR - 1])));
                }
                if (cs > 0) {  // column d + 1 of the row above
                    const int k = d - d0;
                    const T bm = lane_value(fm, k), bi = lane_value(fi, k), bd = lane_value(fd, k);
                    if (l == kL - 1) {
                        S.M[MR - 1] = bm;
                        S.I[MR - 1] = bi;
                        S.D[MR - 1] = bd;
                    }
                }
                if (cks && d == nck) {
                    if (d < D4b) put_state(cks, NV, d / K1, S);
                    nck += K1;
                }
            };
            int d = d0;
            for (; d + 1 <= dlim; d += 2) {
                one_step(d);
                one_step(d + 1);
            }
            if (d <= dlim) one_step(d);
            if (!last) publish_progress(progA + cs, dlim);  // the row's columns of this block
        }
        if (!last) {
            const T mx = lane_value(bmax, kL - 2);
            if (lane == 0) mlog[cs] = mx > T(0) ? ilogb(mx) : kNoMax;
        } else {
            const T Ssum = lane_value(sm + si, lR);  // pairhmm_impl.cpp:434-437
            if (lane == 0) *ssum = Ssum;
        }
        publish_progress(progA + cs, kProgNone);  // the whole row
    }
}

// Pass A by the W waves of the workgroup, 4 strips at a time, and where a
// strip's row fell more than 2^1460 below 2^kLongTop (the three strips
// without a fresh shift lost more than the range holds: reads unrelated to
// their haplotype) or the sum underflowed, again by wave 0 with each shift
// from the row just above (every strip's row at 2^kLongTop).  Ends with a
// workgroup barrier.
template <int MR, int W>
__device__ void long_forward(const AdjArgs &a, const LongQ &q, const uint8_t *cp, double *B, int *shs, int *progA,
                             int *mlog, double *ssum, double *ck1) {
    static_assert(kShiftLag % W == 0, "strip cs - kShiftLag ran on this wave");
    const int w = uni((int)(threadIdx.x >> 6));
    for (int k = threadIdx.x; k < kMaxLongStrips; k += kL * W) {
        progA[k] = -1;
        mlog[k] = kNoMax;
    }
    __syncthreads();
    bool liveA = true;  // this wave's waits have not run out
    long_pass_a<MR>(a, q, cp, B, shs, progA, mlog, ssum, w, W, kShiftLag, liveA, ck1);
    __syncthreads();
    bool again = !(*ssum > 0.0) || ilogb(*ssum) < kSafeLog;
    for (int k = 0; k + 1 < q.nstr; k++) again = again || mlog[k] < kSafeLog;
    if (again) {
        __syncthreads();  // every wave has read the rows' exponents
        for (int k = threadIdx.x; k < kMaxLongStrips; k += kL * W) {
            progA[k] = -1;
            mlog[k] = kNoMax;
        }
        __syncthreads();
        if (w == 0) long_pass_a<MR>(a, q, cp, B, shs, progA, mlog, ssum, 0, 1, 1, liveA, ck1);
    }
    __syncthreads();
}

// Progress of the strips of one pair spread over several workgroups
// (HBM counters, agent scope): as wait_progress / publish_progress.
__device__ __forceinline__ bool wait_progress_x(const int *c, int need, uint32_t limit, int &seen) {
    for (uint32_t k = 0; k < limit; k++) {
        const int v = uni(__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (v <= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the handed-up rows after the count
            seen = v;
            return true;
        }
        __builtin_amdgcn_s_sleep(16);
    }
    return false;
}
__device__ __forceinline__ void publish_progress_x(int *c, int v) {
    __threadfence();  // the handed-up rows before the count
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pass B of strip cs: its forward recomputed from its input row B[cs - 1]
// (scaled by 2^sh) with level-1 checkpoints in ck, then segment by segment
// from the last the level-2 states in l2 and the reverse sweep, the
// adjoints from below read from ab_in (scaled by 2^sh_below) once the strip
// below has published their steps in *prog_below, this strip's first row's
// adjoints written to ab_out and published in *prog_self.  XWG: the pair's
// strips run on several workgroups (HBM progress), else on the waves of
// one workgroup (LDS progress).  The rows' gradient sums are left in gd,
// gdT, gg and ge.
template <int MR, bool XWG>
__device__ void long_strip_b(const AdjArgs &a, const LongQ &q, const uint8_t *cp, const int cs, const double *B,
                             const int sh, const int sh_below, const double sigma, const double *ab_in,
                             double *ab_out, const int *prog_below, int *prog_self, double *ck, double *l2,
                             double *fst, double *bst, bool &live, double (&gd)[MR], double (&gdT)[MR],
                             double (&gg)[MR], double (&ge)[MR], const double *cka = nullptr) {
    using T = double;
    constexpr int NV = 3 * MR + 3;
    const int l = threadIdx.x & 63;
    const int R = q.R, H = q.H, off = q.off, nstr = q.nstr, lR = q.lR;
    const uint64_t ro = q.ro;
    const T initd = q.initd;
    const int K1 = a.ckpt;
    const LaneXfer<kL, 1> up;
    const LaneXfer<kL, -1> down;
    // the code this lane's rows meet at step d: XWG, cp is the base of the
    // packed codes (long_codes_packed), else the lane's byte codes
    auto code = [&](int d) -> uint32_t {
        if constexpr (XWG) {
            const int n = (kL - 1) - l + d;
            const uint32_t nib = ((uint32_t)cp[n >> 1] >> ((n & 1) << 2)) & 0xfu;
            return nib | ((nib + 15u) & 0x10u);
        } else {
            return cp[d];
        }
    };
    Priors<T, MR> P;
    AdjRows<T, MR> A;
    State<T, MR> S0;
    long_setup<MR>(a, cs, nstr, R, off, ro, initd, P, A, S0);
    const bool last = cs == nstr - 1;
    const T fb = last ? T(0) : ldexp(1.0, sh_below);  // adjoints from below to this strip's scale
    const T *bin = B + (uint64_t)(cs > 0 ? cs - 1 : 0) * a.lrow;
    if (cs > 0) {
        const T bm = ldexp(ld_ag(bin + 3 * (1 + kL)), sh), bi = ldexp(ld_ag(bin + 3 * (1 + kL) + 1), sh),
                bd = ldexp(ld_ag(bin + 3 * (1 + kL) + 2), sh);
        if (l == kL - 1) {
            S0.M[MR - 1] = bm;
            S0.I[MR - 1] = bi;
            S0.D[MR - 1] = bd;
        }
    }
    // the reverse covers every step at which a lane of read rows is at a
    // column 1..H: up to lane lR (last strip) or lane L-2
    const int Dn = last ? H + lR : H + kL - 2;
    const int D4 = (Dn + kW - 1) / kW * kW;
    const int nseg = (D4 + K1 - 1) / K1;
    // feeder inputs of the segment after step d0s: column d0s + 2 + m at fst[3m]
    auto stage_feeder = [&](int d0s) {
        wave_sync();
        if (cs > 0)
            for (int k = l; k < 3 * K1; k += kL) {
                const int c = d0s + 2 + k / 3;
                fst[k] = ldexp(ld_ag(bin + 3 * (c + kL) + k % 3), sh);
            }
        wave_sync();
    };
    // a forward step, then the feeder's injection of column d + 1
    auto fstep = [&](State<T, MR> &S, int d, int d0s) {
        step<T, kL, MR, false>(S, P, A.ab_yy, A.ab_my, code(d), up);
        if (cs > 0) {
            const T *f = fst + 3 * (d - d0s - 1);
            const T bm = f[0], bi = f[1], bd = f[2];
            if (l == kL - 1) {
                S.M[MR - 1] = bm;
                S.I[MR - 1] = bi;
                S.D[MR - 1] = bd;
            }
        }
    };
    // pass 1: the strip's forward with level-1 checkpoints, unless pass A
    // left them in cka
    if (!cka) {
        State<T, MR> S = S0;
        put_state(ck, NV, 0, S);
        for (int k = 0; k < nseg; k++) {
            const int d0s = k * K1;
            stage_feeder(d0s);
            const int n = min(K1, D4 - d0s);
            for (int u = 1; u <= n; u += 2) {  // n is a multiple of 4
                fstep(S, d0s + u, d0s);
                fstep(S, d0s + u + 1, d0s);
            }
            if (n == K1 && d0s + K1 < D4) put_state(ck, NV, k + 1, S);
        }
    }
    T Tb[MR], Ib[MR], Db[MR];
#pragma unroll
    for (int s = 0; s < MR; s++) Tb[s] = Ib[s] = Db[s] = gd[s] = gdT[s] = gg[s] = ge[s] = T(0);
    T Tb0p = T(0);
    // one reverse step: lane L-1 first presents the adjoints of the row
    // below lane L-2's last row (bst[2m] = Tb, bst[2m + 1] = Ib of column
    // d0s - 61 + m); lane 0 then hands its first row's adjoints up
    auto rstep = [&](const State<T, MR> &Sb, const State<T, MR> &Sa, int d, int d0s) {
        if (!last) {
            const T tT = bst[2 * (d - d0s)];          // Tb at column d - 61 = j + 1
            const T iI = bst[2 * (d - d0s) - 1];      // Ib at column d - 62 = j
            if (l == kL - 1) {
                Tb0p = tT;
                Ib[0] = iI;
            }
        }
        const int j = d - l;
        const T seed = (last && l == lR && j >= 1 && j <= H) ? sigma : T(0);
        adj_step<T, MR>(Sb, Sa, P, A, code(d), seed, Tb, Ib, Db, Tb0p, gd, gdT, gg, ge, down);
        if (cs > 0 && l == 0) {
            st_ag(ab_out + 2 * d, Tb[0]);
            st_ag(ab_out + 2 * d + 1, Ib[0]);
        }
    };
    auto window4 = [&](const State<T, MR> &W0, int dw, int d0s) {
        State<T, MR> W1 = W0;
        fstep(W1, dw + 1, d0s);
        State<T, MR> W2 = W1;
        fstep(W2, dw + 2, d0s);
        State<T, MR> W3 = W2;
        fstep(W3, dw + 3, d0s);
        State<T, MR> W4 = W3;
        fstep(W4, dw + 4, d0s);
        rstep(W3, W4, dw + 4, d0s);
        rstep(W2, W3, dw + 3, d0s);
        rstep(W1, W2, dw + 2, d0s);
        rstep(W0, W1, dw + 1, d0s);
    };
    // passes 2 and 3, segment by segment from the last
    int seen = kProgNone;  // the strip below's progress as last seen
    for (int k = nseg - 1; k >= 0; k--) {
        const int d0s = k * K1;
        const int nsteps = min(K1, D4 - d0s);
        const int nwin = (nsteps + 2 * kW - 1) / (2 * kW);
        stage_feeder(d0s);
        if (!last) {
            // columns d0s - 61 .. d0s - 61 + K1 of the rows handed up,
            // once the strip below has finished their steps
            const int lo = max(1, d0s - 61), hi = min(H, d0s - 61 + K1);
            if (live && lo <= hi && seen > lo) {  // (a count seen earlier may cover them)
                live = XWG ? wait_progress_x(prog_below, lo, a.spin_limit, seen)
                           : wait_progress(prog_below, lo, a.spin_limit, seen);
                if (!live && l == 0) atomicAdd(a.fault, 1u);
            }
            const T poison = live ? T(1) : T(__builtin_nan(""));
            for (int m = l; m <= K1; m += kL) {
                const int c = d0s - 61 + m;
                const bool in = c >= 1 && c <= H;  // beyond: no adjoint
                bst[2 * m] = in ? ld_ag(ab_in + 2 * c) * fb * poison : T(0);
                bst[2 * m + 1] = in ? ld_ag(ab_in + 2 * c + 1) * fb * poison : T(0);
            }
        }
        wave_sync();
        State<T, MR> S;
        get_state(cka ? cka : ck, NV, k, S);
        // the window starts; the last window's steps are its own pass 3's
        for (int wi = 0; wi < nwin; wi++) {
            put_state(l2, NV, wi, S);
            if (wi == nwin - 1) break;
            for (int u = 0; u < 2 * kW; u++) fstep(S, d0s + wi * 2 * kW + u + 1, d0s);
        }
        wave_sync();
        for (int wi = nwin - 1; wi >= 0; wi--) {
            const int dw = d0s + wi * 2 * kW;
            const int n = min(2 * kW, nsteps - wi * 2 * kW);
            State<T, MR> W0;
            get_state(l2, NV, wi, W0);
            if (n == 2 * kW) {
                State<T, MR> X = W0;
#pragma unroll
                for (int u = 1; u <= kW; u++) fstep(X, dw + u, d0s);
                window4(X, dw + kW, d0s);
            }
            window4(W0, dw, d0s);
        }
        // steps d0s + 1 .. of this strip done: their handed-up columns,
        // published every kPubSegs segments (each publication's release
        // costs a write-back of the L2 in the XWG case) and at the end
        if (cs > 0 && k % kPubSegs == 0) {
            if constexpr (XWG)
                publish_progress_x(prog_self, d0s + 1);
            else
                publish_progress(prog_self, d0s + 1);
        }
    }
}

// A pair's outputs (one wave): the per-lane totals v summed over the lanes,
// log10 L and d log10 L / d theta written by lane 0.
__device__ __forceinline__ void long_outputs(const AdjArgs &a, int p, double (&v)[10], double sigma, double Ssum,
                                             int E) {
    double acc[10];
#pragma unroll
    for (int t = 0; t < 10; t++) {
        double x = v[t];
        for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
        acc[t] = x;
    }
    if ((threadIdx.x & 63) == 0) {
        const double sS = sigma * Ssum;
        const double ln10 = 2.302585092994045684;
        a.out_log10[p] = log10(Ssum) - (1020.0 + (double)E) * 0.30102999566398119521;
        double *g = a.out_grad + 10 * (uint64_t)p;
        g[0] = -a.delta * acc[0] / (10.0 * sS);
        g[1] = -a.gamma * acc[1] / (10.0 * sS);
        for (int b = 0; b < 8; b++) g[2 + b] = acc[2 + b] / (sS * ln10);
        a.flags[p] = 0x01;
    }
}

template <int MR, int W>
__device__ void adj_long_item(const AdjArgs &a, uint8_t *lds, int p_in) {
    using T = double;
    const int l = threadIdx.x & 63;
    const int lane = l;
    const int w = uni((int)(threadIdx.x >> 6));
    constexpr int C = (kL - 1) * MR;
    const LongQ q = long_q<MR>(a, p_in);
    const int p = q.p, R = q.R, off = q.off, nstr = q.nstr;
    const uint64_t ro = q.ro;
    const int K1 = a.ckpt;

    T *wbase = reinterpret_cast<T *>(lds + a.lds_stride);
    T *l2 = wbase + (size_t)w * long_wave_elems(K1, MR);
    T *fst = l2 + (size_t)((K1 + 7) / 8) * (3 * MR + 3) * kL;
    T *bst = fst + 3 * K1;
    T *tot = wbase + (size_t)W * long_wave_elems(K1, MR);  // per-lane totals, tot[t * 64 + lane]
    T *ssum = tot + 10 * kL;
    int *shs = reinterpret_cast<int *>(ssum + 1);
    int *prog = shs + kMaxLongStrips;
    int *progA = prog + kMaxLongStrips;
    int *mlog = progA + kMaxLongStrips;
    long_codes(a, lds, q, 64 * W);
    const uint8_t *cp = lds + (kL - 1) - l;
    T *B = reinterpret_cast<T *>(a.lbnd) + (uint64_t)blockIdx.x * a.lbnd_stride;  // row of strip cs at cs * lrow
    T *AB = reinterpret_cast<T *>(a.lab) + (uint64_t)blockIdx.x * (W + 1) * a.lab_elems;
    T *ck = reinterpret_cast<T *>(a.ck) + ((uint64_t)blockIdx.x * W + w) * a.ck_elems;

    // ---- A
    long_forward<MR, W>(a, q, cp, B, shs, progA, mlog, ssum, nullptr);
    const T Ssum = *ssum;
    int E = 0;  // entries of the last strip: true values x 2^1020 x 2^E
    for (int k = 1; k < nstr; k++) E += shs[k];
    if (!(Ssum > T(0))) {  // underflow even renormalised: no gradient
        if (threadIdx.x == 0) {
            a.out_log10[p] = -INFINITY;
            for (int t = 0; t < 10; t++) a.out_grad[10 * (uint64_t)p + t] = 0.0;
            a.flags[p] = 0x03;
        }
        return;
    }
    const T sigma = ldexp(1.0, -ilogb(Ssum));
    for (int k = threadIdx.x; k < 10 * kL; k += kL * W) tot[k] = T(0);
    for (int k = threadIdx.x; k < kMaxLongStrips; k += kL * W) prog[k] = kProgNone;
    __syncthreads();

    // ---- B: strips last to first, dealt to the waves
    bool live = true;  // this wave's waits have not run out
    for (int cs = nstr - 1 - w; cs >= 0; cs -= W) {
        const bool last = cs == nstr - 1;
        T gd[MR], gdT[MR], gg[MR], ge[MR];
        long_strip_b<MR, false>(a, q, cp, cs, B, shs[cs], last ? 0 : shs[cs + 1], sigma,
                                AB + (uint64_t)((cs + 1) % (W + 1)) * a.lab_elems,
                                AB + (uint64_t)(cs % (W + 1)) * a.lab_elems, prog + cs + 1, prog + cs, ck, l2, fst,
                                bst, live, gd, gdT, gg, ge);
        // the strip's rows into the per-lane totals, after the strip below's
        if (!last && live) {
            int seen = kProgNone;
            live = wait_progress(prog + cs + 1, 0, a.spin_limit, seen);
            if (!live && l == 0) atomicAdd(a.fault, 1u);
        }
#pragma unroll
        for (int s = 0; s < MR; s++) {
            const int i = cs * C + l * MR + s - off + 1;
            const bool rl = l < kL - 1 && i >= 1 && i <= R;
            const uint32_t qq = a.q[ro + (uint64_t)(rl ? i - 1 : 0)] & 127u;
            const uint32_t bin_q = a.qbin[qq];
            tot[0 * kL + lane] += rl ? fma(T(-2), gdT[s], gd[s]) : T(0);
            tot[1 * kL + lane] += rl ? gg[s] : T(0);
            const T c = rl ? ge[s] * a.wq[qq] : T(0);
            tot[(2 + bin_q) * kL + lane] += c;
        }
        publish_progress(prog + cs, 0);
    }
    __syncthreads();  // every strip's totals
    if (w != 0) return;
    double v[10];
#pragma unroll
    for (int t = 0; t < 10; t++) v[t] = tot[t * kL + lane];
    long_outputs(a, p, v, sigma, Ssum, E);
}


template <int W>
__global__ __launch_bounds__(64 * W, 1) void adj_long_kernel(AdjArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t n = a.n_items;
    int *slot = reinterpret_cast<int *>(lds + long_lds_bytes(a.lds_stride, a.ckpt, kLongMR, W) - 16);
    // a workgroup takes at most n items: the bound makes every wave exit
    for (uint32_t iter = 0; iter <= n; iter++) {
        __syncthreads();  // every wave has read the slot
        if (threadIdx.x == 0) *slot = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t it = (uint32_t)uni(*slot);
        if (it >= n) break;
        adj_long_item<kLongMR, W>(a, lds, a.items[it]);
    }
}

// ---- a pair's strips over several workgroups (the default), in three
// launches:
//  A. pass A of each pair (a.items, longest first) on a workgroup of
//     kLongWaves waves: its rows to the pair's B rows (a.lbnd + li *
//     lbnd_stride), its shifts, scaled sum and exponent to a.lsh, a.lsum,
//     a.lexp;
//  B. one wave per item (pair, strip) of a.bitems -- pairs longest first,
//     each pair's strips last to first -- running pass B of the strip: the
//     adjoints it hands up go through one HBM row per strip, with the steps
//     done published in an HBM count per strip.  A wave waits (bounded) only
//     for the strip below, an earlier item of the list, which a running
//     wave took, so every wait ends.  The rows' gradient contributions go to
//     a.lpart per strip, row and lane;
//  C. per pair, the contributions added strip by strip, last first, in the
//     order of the one-workgroup kernel (the same bits), and the outputs.
// The longest pair's strips then run on as many CUs as it has strips
// instead of on one.
// LDS of pass A's workgroup: codes | scaled sum | shifts | forward progress
// | row exponents | item
struct LongALds {
    double *ssum;
    int *shs, *progA, *mlog, *slot;
    __device__ LongALds(uint8_t *lds, uint32_t lds_stride) {
        ssum = reinterpret_cast<double *>(lds + lds_stride);
        shs = reinterpret_cast<int *>(ssum + 1);
        progA = shs + kMaxLongStrips;
        mlog = progA + kMaxLongStrips;
        slot = mlog + kMaxLongStrips;
    }
};

// Pass A of long pair li by the workgroup's kLongWaves waves: its rows to
// the pair's B rows (a.lbnd + li * lbnd_stride), its shifts, scaled sum and
// exponent to a.lsh, a.lsum, a.lexp.  Ends with a workgroup barrier.
__device__ void long_a_item(const AdjArgs &a, uint8_t *lds, const LongALds &m, uint32_t li) {
    using T = double;
    const LongQ q = long_q<kLongMR>(a, a.items[li]);
    long_codes(a, lds, q, 64 * kLongWaves);
    const uint8_t *cp = lds + (kL - 1) - (int)(threadIdx.x & 63);
    T *B = reinterpret_cast<T *>(a.lbnd) + (uint64_t)li * a.lbnd_stride;
    double *ck1 = a.lck1 ? a.lck1 + a.lck1_off[li] : nullptr;
    long_forward<kLongMR, kLongWaves>(a, q, cp, B, m.shs, m.progA, m.mlog, m.ssum, ck1);
    const T Ssum = *m.ssum;
    for (int k = threadIdx.x; k < q.nstr; k += 64 * kLongWaves) a.lsh[(uint64_t)li * a.lstrips + k] = m.shs[k];
    if (threadIdx.x == 0) {
        int E = 0;  // entries of the last strip: true values x 2^1020 x 2^E
        for (int k = 1; k < q.nstr; k++) E += m.shs[k];
        a.lsum[li] = Ssum;
        a.lexp[li] = E;
        if (!(Ssum > T(0))) {  // underflow even renormalised: no gradient
            a.out_log10[q.p] = -INFINITY;
            for (int t = 0; t < 10; t++) a.out_grad[10 * (uint64_t)q.p + t] = 0.0;
            a.flags[q.p] = 0x03;
        }
    }
    __syncthreads();
}

// Pass B of strip cs of long pair li by one wave (LDS from `lds`: packed
// codes | level-2 states | feeder inputs of a segment | adjoints from below
// of a segment; level-1 checkpoints in ck).  The adjoints it hands up go
// through one HBM row per strip, with the steps done published in an HBM
// count per strip; it waits (bounded) only for the strip below.  The rows'
// gradient contributions go to a.lpart per strip, row and lane.
__device__ void long_b_item(const AdjArgs &a, uint8_t *lds, uint32_t li, int cs, double *ck) {
    using T = double;
    constexpr int MR = kLongMR;
    constexpr int C = (kL - 1) * MR;
    const int l = threadIdx.x & 63;
    const int K1 = a.ckpt;
    T *l2 = reinterpret_cast<T *>(lds + packed_stride(a.lds_stride));
    T *fst = l2 + (size_t)((K1 + 7) / 8) * (3 * MR + 3) * kL;
    T *bst = fst + 3 * K1;
    const T Ssum = a.lsum[li];
    if (!__builtin_amdgcn_readfirstlane(Ssum > T(0) ? 1 : 0)) return;  // no gradient: no strip runs
    const LongQ q = long_q<MR>(a, a.items[li]);
    long_codes_packed(a, lds, q);
    const uint8_t *cp = lds;
    const T *B = reinterpret_cast<const T *>(a.lbnd) + (uint64_t)li * a.lbnd_stride;
    const int *sh = a.lsh + (uint64_t)li * a.lstrips;
    const bool last = cs == q.nstr - 1;
    T *ab = reinterpret_cast<T *>(a.lab) + (uint64_t)li * a.lstrips * a.lab_elems;
    int *pr = a.lprog + (uint64_t)li * a.lstrips;
    bool live = true;
    T gd[MR], gdT[MR], gg[MR], ge[MR];
    const double *cka =
        a.lck1 ? a.lck1 + a.lck1_off[li] + (uint64_t)cs * long_ck_elems((uint32_t)q.H, K1, MR) : nullptr;
    long_strip_b<MR, true>(a, q, cp, cs, B, sh[cs], last ? 0 : sh[cs + 1], ldexp(1.0, -ilogb(Ssum)),
                           ab + (uint64_t)(last ? cs : cs + 1) * a.lab_elems, ab + (uint64_t)cs * a.lab_elems,
                           pr + (last ? cs : cs + 1), pr + cs, ck, l2, fst, bst, live, gd, gdT, gg, ge, cka);
    // the rows' contributions to the totals, as the one-workgroup kernel
    // adds them: per row the delta and gamma sums, the error-prior sum and
    // its quality bin
    T *part = a.lpart + ((uint64_t)li * a.lstrips + cs) * MR * 4 * kL;
#pragma unroll
    for (int s = 0; s < MR; s++) {
        const int i = cs * C + l * MR + s - q.off + 1;
        const bool rl = l < kL - 1 && i >= 1 && i <= q.R;
        const uint32_t qq = a.q[q.ro + (uint64_t)(rl ? i - 1 : 0)] & 127u;
        part[(s * 4 + 0) * kL + l] = rl ? fma(T(-2), gdT[s], gd[s]) : T(0);
        part[(s * 4 + 1) * kL + l] = rl ? gg[s] : T(0);
        part[(s * 4 + 2) * kL + l] = rl ? ge[s] * a.wq[qq] : T(0);
        part[(s * 4 + 3) * kL + l] = (T)a.qbin[qq];
    }
}

// The long pairs' strips over several workgroups, by default in two
// launches (adj_longab_kernel, then adj_longc_kernel), for tests in three
// (adj_longa_kernel, adj_longb_kernel, adj_longc_kernel):
//  A. pass A of each pair (a.items, longest first) on a workgroup of
//     kLongWaves waves (long_a_item);
//  B. one wave per item (pair, strip) -- pairs longest first, each pair's
//     strips last to first -- running pass B of the strip (long_b_item).  A
//     wave waits (bounded) only for the strip below, an earlier item that a
//     running wave took, so every wait ends;
//  C. per pair, the contributions added strip by strip, last first, in the
//     order of the one-workgroup kernel (the same bits), and the outputs.
// The longest pair's strips then run on as many CUs as it has strips
// instead of on one.
__global__ __launch_bounds__(64 * kLongWaves, 1) void adj_longa_kernel(AdjArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const LongALds m(lds, a.lds_stride);
    const uint32_t n = a.n_items;
    for (uint32_t iter = 0; iter <= n; iter++) {
        __syncthreads();  // every wave has read the slot
        if (threadIdx.x == 0) *m.slot = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t li = (uint32_t)uni(*m.slot);
        if (li >= n) break;
        long_a_item(a, lds, m, li);
    }
}

// B over the list a.bitems of (pair << 8 | strip), one-wave workgroups.
__global__ __launch_bounds__(64, 1) void adj_longb_kernel(AdjArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const int l = threadIdx.x & 63;
    double *ck = reinterpret_cast<double *>(a.ck) + (uint64_t)blockIdx.x * a.ck_elems;
    const uint32_t n = a.n_bitems;
    // a wave takes at most n items: the bound makes every wave exit
    for (uint32_t iter = 0; iter <= n; iter++) {
        uint32_t it = 0;
        if (l == 0) it = atomicAdd(a.bcounter, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n) break;
        const uint32_t e = a.bitems[it];
        long_b_item(a, lds, (uint32_t)uni((int)(e >> 8)), uni((int)(e & 0xffu)), ck);
    }
}

// LDS of one pass-B wave (rounded to 16 bytes) and of a workgroup of
// adj_longab_kernel: pass A's, or kLongWaves waves' of pass B.
__host__ __device__ constexpr uint64_t longb_wave_lds(uint32_t lds_stride, int K1) {
    return ((uint64_t)packed_stride(lds_stride) + long_wave_elems(K1, kLongMR) * 8 + 15) & ~15ull;
}
__host__ __device__ constexpr uint64_t longa_lds(uint32_t lds_stride) {
    return (uint64_t)lds_stride + 8 + 12ull * kMaxLongStrips + 16;
}

// A and B in one launch: each workgroup runs pass A of pairs from a.counter
// (longest first) until none is left, then each of its waves on its own
// runs pass B of strips: of the first pair (longest first) whose pass A is
// done (a.lready) with strips left (a.lnext counts those taken, last strip
// first; a.lnstr the pair's strips).  So pass B of the shorter pairs fills
// the CUs whose pass A ended while the longest pairs' pass A runs.  A wave
// finds no strip only while some pair's pass A runs on a workgroup that
// took it, so its polls end (bounded all the same: a.spin_limit polls, then
// a fault count).  Pass A publishes a pair with every wave's agent-scope
// release (its rows, shifts and sum) before the flag.
__global__ __launch_bounds__(64 * kLongWaves, 1) void adj_longab_kernel(AdjArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const LongALds m(lds, a.lds_stride);
    const uint32_t n = a.n_items;
    for (uint32_t iter = 0; iter <= n; iter++) {
        __syncthreads();  // every wave has read the slot
        if (threadIdx.x == 0) *m.slot = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t li = (uint32_t)uni(*m.slot);
        if (li >= n) break;
        long_a_item(a, lds, m, li);
        __threadfence();  // this wave's rows and shifts, at agent scope
        __syncthreads();
        if (threadIdx.x == 0) {
            if (!(a.lsum[li] > 0.0))  // no strip runs: every one taken
                __hip_atomic_store(a.lnext + li, 0x40000000u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            __hip_atomic_store(a.lready + li, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();  // pass A's LDS is free
    const int w = uni((int)(threadIdx.x >> 6));
    const int l = threadIdx.x & 63;
    uint8_t *lw = lds + (uint64_t)w * longb_wave_lds(a.lds_stride, a.ckpt);
    double *ck = reinterpret_cast<double *>(a.ck) + ((uint64_t)blockIdx.x * kLongWaves + w) * a.ck_elems;
    uint32_t k0 = 0;     // the pairs before k0 have no strip left
    uint32_t polls = 0;  // polls since this wave last found a strip
    for (;;) {
        int li = -1, cs = 0;
        bool pending = false;  // a pair with strips left whose pass A runs
        for (uint32_t k = k0; k < n; k++) {
            const uint32_t ns = a.lnstr[k];
            const uint32_t used = uni((int)__hip_atomic_load(a.lnext + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (used >= ns) {
                if (k == k0) k0++;
                continue;
            }
            if (!uni(__hip_atomic_load(a.lready + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                pending = true;
                continue;
            }
            uint32_t t = 0;
            if (l == 0) t = atomicAdd(a.lnext + k, 1u);
            t = __builtin_amdgcn_readfirstlane(t);
            if (t < ns) {
                li = (int)k;
                cs = (int)(ns - 1 - t);
                break;
            }
            if (k == k0) k0++;
        }
        if (li >= 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // pass A's rows, shifts and sum
            long_b_item(a, lw, (uint32_t)li, cs, ck);
            polls = 0;
            continue;
        }
        if (!pending) break;  // every strip taken
        if (++polls > a.spin_limit) {
            if (l == 0) atomicAdd(a.fault, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(32);
    }
}

__global__ __launch_bounds__(64) void adj_longc_kernel(AdjArgs a) {
    using T = double;
    constexpr int MR = kLongMR;
    const int l = threadIdx.x & 63;
    for (uint32_t li = blockIdx.x; li < a.n_items; li += gridDim.x) {
        const T Ssum = a.lsum[li];
        if (!(Ssum > T(0))) continue;
        const LongQ q = long_q<MR>(a, a.items[li]);
        T v[10];
#pragma unroll
        for (int t = 0; t < 10; t++) v[t] = T(0);
        for (int cs = q.nstr - 1; cs >= 0; cs--) {
            const T *part = a.lpart + ((uint64_t)li * a.lstrips + cs) * MR * 4 * kL;
#pragma unroll
            for (int s = 0; s < MR; s++) {
                const T c0 = part[(s * 4 + 0) * kL + l], c1 = part[(s * 4 + 1) * kL + l];
                const T c2 = part[(s * 4 + 2) * kL + l];
                const int b = (int)part[(s * 4 + 3) * kL + l];
                v[0] += c0;
                v[1] += c1;
                // one bin takes c2; the others add +0, which leaves a total
                // unchanged (no total is -0: it starts at +0)
#pragma unroll
                for (int t = 0; t < 8; t++) v[2 + t] += t == b ? c2 : T(0);
            }
        }
        long_outputs(a, q.p, v, ldexp(1.0, -ilogb(Ssum)), Ssum, a.lexp[li]);
    }
}

}  // namespace

// rows per lane of the G-pairs-per-wave kernel: 1 .. adj2_max_mr() at 16,
// 32 and 64 lanes per pair
int adj2_max_mr() { return 6; }
void *adj2_kernel_ptr(int L, int MR, bool rescue) {
#define KB(l, m)                                                                       \
    case m:                                                                            \
        return rescue ? reinterpret_cast<void *>(&adj2_kernel<double, l, m, true>)     \
                      : reinterpret_cast<void *>(&adj2_kernel<float, l, m, false>);
    if (L == 16) switch (MR) { KB(16, 1) KB(16, 2) KB(16, 3) KB(16, 4) KB(16, 5) KB(16, 6) default: return nullptr; }
    if (L == 32) switch (MR) { KB(32, 1) KB(32, 2) KB(32, 3) KB(32, 4) KB(32, 5) KB(32, 6) default: return nullptr; }
    if (L == 64) switch (MR) { KB(64, 1) KB(64, 2) KB(64, 3) KB(64, 4) KB(64, 5) KB(64, 6) default: return nullptr; }
#undef KB
    return nullptr;
}

int adj_long_mr() { return kLongMR; }
int adj_long_max_strips() { return kMaxLongStrips; }
// waves per workgroup of the long-read kernel: 1 or adj_long_waves()
int adj_long_waves() { return kLongWaves; }
size_t adj_long_lds(uint32_t lds_stride, int K1, int waves) {
    return (size_t)long_lds_bytes(lds_stride, K1, kLongMR, waves);
}

hipError_t launch_adj_long(int waves, int grid, size_t lds_bytes, hipStream_t st, const AdjArgs &a) {
    void *args[] = {const_cast<AdjArgs *>(&a)};
    void *k = waves == 1 ? reinterpret_cast<void *>(&adj_long_kernel<1>)
                         : reinterpret_cast<void *>(&adj_long_kernel<kLongWaves>);
    return hipLaunchKernel(k, dim3(grid), dim3(64 * (waves == 1 ? 1 : kLongWaves)), args, lds_bytes, st);
}

// The three launches of the long pairs' strips over several workgroups:
// LDS of pass A's and pass B's workgroups, and the launches in order.
size_t adj_longa_lds(uint32_t lds_stride) { return (size_t)longa_lds(lds_stride); }
size_t adj_longb_lds(uint32_t lds_stride, int K1) { return (size_t)longb_wave_lds(lds_stride, K1); }
size_t adj_longab_lds(uint32_t lds_stride, int K1) {
    const uint64_t b = (uint64_t)kLongWaves * longb_wave_lds(lds_stride, K1);
    return (size_t)(b > longa_lds(lds_stride) ? b : longa_lds(lds_stride));
}
// pass-B waves of one adj_longab_kernel workgroup
int adj_longab_waves() { return kLongWaves; }
// elements of one strip's pass-A checkpoints for pass B
uint64_t adj_long_ck_elems(uint32_t H, int K1) { return long_ck_elems(H, K1, kLongMR); }
hipError_t launch_adj_long_split(int grid_a, int grid_b, int grid_c, size_t lds_b, hipStream_t st,
                                 const AdjArgs &a) {
    void *args[] = {const_cast<AdjArgs *>(&a)};
    hipError_t e = hipLaunchKernel(reinterpret_cast<void *>(&adj_longa_kernel), dim3(grid_a), dim3(64 * kLongWaves),
                                   args, adj_longa_lds(a.lds_stride), st);
    if (e == hipSuccess)
        e = hipLaunchKernel(reinterpret_cast<void *>(&adj_longb_kernel), dim3(grid_b), dim3(64), args, lds_b, st);
    if (e == hipSuccess)
        e = hipLaunchKernel(reinterpret_cast<void *>(&adj_longc_kernel), dim3(grid_c), dim3(64), args, 0, st);
    return e;
}

hipError_t launch_adj_long_fused(int grid_ab, int grid_c, size_t lds_ab, hipStream_t st, const AdjArgs &a) {
    void *args[] = {const_cast<AdjArgs *>(&a)};
    hipError_t e = hipLaunchKernel(reinterpret_cast<void *>(&adj_longab_kernel), dim3(grid_ab),
                                   dim3(64 * kLongWaves), args, lds_ab, st);
    if (e == hipSuccess)
        e = hipLaunchKernel(reinterpret_cast<void *>(&adj_longc_kernel), dim3(grid_c), dim3(64), args, 0, st);
    return e;
}

hipError_t launch_adj(void *kern, int grid, size_t lds_bytes, hipStream_t st, const AdjArgs &a) {
    void *args[] = {const_cast<AdjArgs *>(&a)};
    return hipLaunchKernel(kern, dim3(grid), dim3(64), args, lds_bytes, st);
}

}  // namespace rodv
