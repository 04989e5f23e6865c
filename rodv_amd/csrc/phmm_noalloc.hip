// phmm_noalloc.hip -- rodv_phmm_forward without device allocation.
//
// rodv_phmm_forward_noalloc runs exactly what rodv_phmm_forward runs (the
// same plan, kernels and routing, hence the same results), but its plan takes
// every device buffer -- uploaded inputs, results, rescue list, counters,
// boundary rows -- from `scratch`, a fixed array in device memory, so the
// call makes no hipMalloc/hipFree.  For callers that cannot allocate device
// memory: serving loops and harnesses that forbid it.  It is synchronous
// (host inputs and outputs, like rodv_phmm_forward), so it is not for use
// under graph capture.
//
// In host code `scratch` names a host-side stand-in: the array's device
// address on each device comes from hipGetSymbolAddress, taken once per
// device (with that device current) at the entry point's first call there,
// and only that address is given to the plan.  Every call on a device uses
// the same array, which keeps the last call's contents; what must start at
// a known value is set on every call (the plan's counters by its reset
// kernel, its boundary rows from the host mirror below).  Calls take turns
// (one mutex for the process), so calls that would overlap on one GPU do
// not race: the second waits.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>

#pragma GCC visibility push(default)
#include "../../include/rodv_phmm.h"
#pragma GCC visibility pop
#include "phmm_arena.h"

constexpr size_t kScratchBytes = size_t(32) << 20;  // RODV_PHMM_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == RODV_PHMM_NOALLOC_SCRATCH_BYTES, "header and library agree");

// Device scratch of the entry point; every device the library runs on holds
// one copy until the process ends.  Kernels reach it only through the
// address hipGetSymbolAddress gives.  Host code names it (below), so the
// compiler gives its device symbol global binding, which is how the runtime
// finds the variable it registers by that name; an attribute that pins the
// symbol as it is written, such as `used`, would keep it local and out of
// the code object's dynamic symbols.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// Host mirror of the scratch's used prefix (the plan's uploads and fills go
// here and reach the device in one copy); static, so it is not allocated.
alignas(256) uint8_t g_mirror[kScratchBytes];

std::mutex g_mu;                       // calls take turns
void *g_scratch_dev[kMaxDevices] = {};  // scratch's address on each device

// scratch's address on `device` (current), once per device
int scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return RODV_PHMM_ERR_UNSUPPORTED;
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return RODV_PHMM_ERR_DEVICE;
        }
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return RODV_PHMM_OK;
}

int forward_noalloc_impl(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch, const rodv_phmm_config *cfg,
                         double *out_log10, uint8_t *out_flags) {
    if (!ctx || !out_log10) return RODV_PHMM_ERR_INVALID_ARG;
    rodv_phmm_config c;
    rodv_phmm_config_default(&c);
    if (cfg) {
        if (cfg->struct_size != sizeof(rodv_phmm_config)) return RODV_PHMM_ERR_INVALID_ARG;
        c = *cfg;
    }
    c.use_graph = 0;
    const int device = rodv::ctx_device(ctx);
    std::lock_guard<std::mutex> lk(g_mu);
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return RODV_PHMM_ERR_DEVICE;
    }
    void *dev = nullptr;
    int rc = scratch_address(device, &dev);
    if (rc) return rc;
    rodv::Arena A;
    A.dev = static_cast<uint8_t *>(dev);
    A.host = g_mirror;
    A.cap = kScratchBytes;
    rc = rodv::forward_in(ctx, batch, c, &A, out_log10, out_flags);
    return rc == rodv::kArenaFull ? RODV_PHMM_ERR_TOO_LARGE : rc;
}

}  // namespace

extern "C" int rodv_phmm_forward_noalloc(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch,
                                         const rodv_phmm_config *cfg, double *out_log10,
                                         uint8_t *out_flags) {
    try {  // nothing throws across the C ABI
        return forward_noalloc_impl(ctx, batch, cfg, out_log10, out_flags);
    } catch (const std::bad_alloc &) {
        return RODV_PHMM_ERR_OOM;
    } catch (...) {
        return RODV_PHMM_ERR_DEVICE;
    }
}
