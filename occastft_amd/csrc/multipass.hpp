// multipass.hpp -- large-window path (complex transform sizes 2^14 .. 2^20).
#pragma once

#include "engine.hpp"

struct ostft_ctx;

namespace ostft {

// Two ping-pong complex buffers owned by a context, grown on demand.
struct Scratch {
    float2* a = nullptr;
    float2* b = nullptr;
    uint64_t cap = 0;  // complex elements in each buffer
    hipError_t ensure(uint64_t n);
    void release();
};

// accessors into the context (defined in capi.hip)
float2* ctx_twiddles(ostft_ctx* ctx, int64_t m, int64_t count, int* rc);
Scratch& ctx_scratch(ostft_ctx* ctx);

// Runs the P column passes and the split epilogue.  Returns an OSTFT_* status.
int run_multipass(ostft_ctx* ctx, int log_m, const FusedArgs& a, hipStream_t s);

// Radix split of 2^log_m into passes of at most 2^8 (larger radix first).
int multipass_plan(int log_m, int* log_radix /* [4] */);

}  // namespace ostft
