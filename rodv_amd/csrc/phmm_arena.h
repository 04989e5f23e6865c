// phmm_arena.h -- device memory a plan carves its buffers from (host side).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/rodv_phmm.h"

namespace rodv {

// A block of `cap` bytes at `dev` with a host mirror `host` of the same size.
// A plan built in an arena takes its buffers from it (256-byte aligned) in
// place of allocating each one; its uploads and fills go to the mirror, and
// the mirror's used prefix reaches the device in one copy before the plan's
// first kernel.  Such a plan frees nothing: the arena's owner reuses the
// block for the next plan.
struct Arena {
    uint8_t *dev = nullptr;
    uint8_t *host = nullptr;
    size_t cap = 0;
    size_t used = 0;
    size_t need = 0;  // bytes requested so far by the plan being built
};

// Status of a plan build whose arena is too small (never returned by the C
// ABI: the callers grow the arena or report RODV_PHMM_ERR_TOO_LARGE).
constexpr int kArenaFull = 100;

// The device of a context.
int ctx_device(const rodv_phmm_ctx *ctx);

// One plan of `batch` in `arena` (nullptr: buffers of its own), executed
// once without a graph on the null stream, read back and destroyed.
int forward_in(rodv_phmm_ctx *ctx, const rodv_phmm_batch *batch, const rodv_phmm_config &c,
               Arena *arena, double *out_log10, uint8_t *out_flags);

}  // namespace rodv
