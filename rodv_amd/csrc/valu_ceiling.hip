// valu_ceiling.hip -- measures the FP32 VALU issue ceiling of the device the
// forward runs on, for the benchmark's roofline.
//
// The bit-exact forward needs 8 non-packed FP32 instructions per cell (4 for
// M, 2 for I, 2 for D), 7 when the I and D recurrences share the product of
// M and the indel prior (reads whose deletion and insertion priors of
// adjacent rows are equal).  Its attainable ceiling is therefore the rate at
// which the device issues wave64 v_fma_f32 / v_mul_f32, which this library
// measures with four kernels, each 8 independent chains per lane:
//   kind 0: v_fma_f32 only
//   kind 1: v_mul_f32 only
//   kind 2: the forward cell's own instruction mix and dependencies
//   kind 3: the same with the shared product (7 instructions per cell)
// Built with the same flags as the engine (packed FP32 disabled), so no
// v_pk_* instruction is generated.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kChains = 8;
constexpr int kInner = 16;

__global__ __launch_bounds__(256) void fma_chains(float *out, float a, float b, int iters) {
    float acc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) acc[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < kInner; r++) {
#pragma unroll
            for (int k = 0; k < kChains; k++) acc[k] = __builtin_fmaf(acc[k], a, b);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void mul_chains(float *out, float a, float b, int iters) {
    float acc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) acc[k] = (float)(threadIdx.x + k) * 1e-3f + b;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < kInner; r++) {
#pragma unroll
            for (int k = 0; k < kChains; k++) acc[k] = acc[k] * a;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 4 independent cells per step, each the forward's 8 operations:
//   t = M*mm; t = fma(I, g, t); t = fma(D, g, t); M' = t*pr
//   I' = fma(I, xx, M*mx);  D' = fma(D, yy, M*my)
__global__ __launch_bounds__(256) void cell_mix(float *out, float a, float b, int iters) {
    float M[4], I[4], D[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        M[k] = (float)(threadIdx.x + k) * 1e-3f;
        I[k] = M[k] * 0.5f;
        D[k] = M[k] * 0.25f;
    }
    const float mm = a, g = 1.0f - a, pr = 0.9f + b, mx = 0.01f, xx = 0.1f, my = 0.02f, yy = 0.2f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < kInner / 2; r++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float t = M[k] * mm;
                t = __builtin_fmaf(I[k], g, t);
                t = __builtin_fmaf(D[k], g, t);
                const float In = __builtin_fmaf(I[k], xx, M[k] * mx);
                const float Dn = __builtin_fmaf(D[k], yy, M[k] * my);
                M[k] = t * pr;
                I[k] = In;
                D[k] = Dn;
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) s += M[k] + I[k] + D[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// kind 3: as cell_mix, the I and D updates sharing one product
__global__ __launch_bounds__(256) void cell_mix7(float *out, float a, float b, int iters) {
    float M[4], I[4], D[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        M[k] = (float)(threadIdx.x + k) * 1e-3f;
        I[k] = M[k] * 0.5f;
        D[k] = M[k] * 0.25f;
    }
    const float mm = a, g = 1.0f - a, pr = 0.9f + b, mx = 0.01f, xx = 0.1f, yy = 0.2f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < kInner / 2; r++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float t = M[k] * mm;
                t = __builtin_fmaf(I[k], g, t);
                t = __builtin_fmaf(D[k], g, t);
                const float u = M[k] * mx;
                const float In = __builtin_fmaf(I[k], xx, u);
                const float Dn = __builtin_fmaf(D[k], yy, u);
                M[k] = t * pr;
                I[k] = In;
                D[k] = Dn;
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) s += M[k] + I[k] + D[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// Lane-operations per second of `kind` (0 fma, 1 mul, 2 forward-cell mix:
// 8 FP32 instructions per cell, 3 the shared-product mix: 7) with
// `waves_per_simd` waves on every SIMD of
// `device`; best of `reps` timed launches.  Returns 0 on success.
int valu_ceiling(int device, int kind, int waves_per_simd, int reps, double *lane_ops_per_s) {
    if (!lane_ops_per_s || kind < 0 || kind > 3 || waves_per_simd < 1 || waves_per_simd > 8 ||
        reps < 1)
        return 1;
    if (hipSetDevice(device) != hipSuccess) return 3;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 3;
    const int threads = 256;  // 4 waves: one per SIMD
    const int blocks = cus * waves_per_simd;
    const int iters = 2048;
    float *d_out = nullptr;
    if (hipMalloc(&d_out, sizeof(float) * (size_t)blocks * threads) != hipSuccess) return 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 3;
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(fma_chains, dim3(blocks), dim3(threads), 0, 0, d_out, 0.999f,
                               1e-3f, iters);
        else if (kind == 1)
            hipLaunchKernelGGL(mul_chains, dim3(blocks), dim3(threads), 0, 0, d_out, 0.999999f,
                               1.0f, iters);
        else if (kind == 2)
            hipLaunchKernelGGL(cell_mix, dim3(blocks), dim3(threads), 0, 0, d_out, 0.9f, 0.0f,
                               iters);
        else
            hipLaunchKernelGGL(cell_mix7, dim3(blocks), dim3(threads), 0, 0, d_out, 0.9f, 0.0f,
                               iters);
    };
    double best = 0.0;
    if (!rc) {
        launch();  // warm-up
        for (int r = 0; r < reps && !rc; r++) {
            float ms = 0.f;
            if (hipEventRecord(e0, 0) != hipSuccess) rc = 3;
            launch();
            if (!rc && hipEventRecord(e1, 0) != hipSuccess) rc = 3;
            if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = 3;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = 3;
            if (!rc && ms > 0.f) {
                const double ops_per_iter = kind == 2   ? (double)(kInner / 2) * 4 * 8
                                            : kind == 3 ? (double)(kInner / 2) * 4 * 7
                                                        : (double)kInner * kChains;
                const double ops = (double)blocks * threads * iters * ops_per_iter;
                const double r_ = ops / (ms * 1e-3);
                if (r_ > best) best = r_;
            }
        }
    }
    if (!rc && hipGetLastError() != hipSuccess) rc = 3;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_out);
    *lane_ops_per_s = best;
    return rc;
}

}  // extern "C"
