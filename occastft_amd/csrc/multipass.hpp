// multipass.hpp -- large-window path (complex transform sizes 2^13 .. 2^19).
#pragma once

#include <mutex>

#include "engine.hpp"

struct ostft_ctx;

namespace ostft {

// The complex buffer between the two passes, owned by a context, grown on
// demand up to one chunk of frames.  Calls on different streams take turns:
// each call's launches wait for the previous call's (an event recorded on
// its stream) when the stream changes.
struct Scratch {
    float2* a = nullptr;
    uint64_t cap = 0;                      // complex elements
    uint64_t budget = (uint64_t)1 << 30;  // bytes per chunk of frames
    hipEvent_t done = nullptr;             // recorded after the last call's launches
    hipStream_t last = nullptr;            // the stream of that call
    bool used = false;
    hipError_t ensure(uint64_t n);
    hipError_t acquire(hipStream_t s);  // order stream s after the previous user
    hipError_t handoff(hipStream_t s);  // record the end of this call's launches
    void release();                     // frees the buffer (after the last user's launches)
    void destroy();                     // release + the event
};

// accessors into the context (defined in capi.hip)
// W_m^j, j < count; with `noalloc` only a table already resident (else null
// and OSTFT_ERR_NOT_PREPARED in *rc)
float2* ctx_twiddles(ostft_ctx* ctx, int64_t m, int64_t count, int* rc, bool noalloc = false);
Scratch& ctx_scratch(ostft_ctx* ctx);
std::mutex& ctx_scratch_mutex(ostft_ctx* ctx);  // held while a call queues its scratch launches

// Runs the two passes (the split epilogue inside the last).  Returns an OSTFT_* status.
int run_multipass(ostft_ctx* ctx, int log_m, const FusedArgs& a, hipStream_t s);
// The same through the fixed per-device scratch array (no allocation, no
// event; the tables must be resident).  Calls that overlap on one GPU race.
int run_multipass_static(ostft_ctx* ctx, int device, int log_m, const FusedArgs& a, hipStream_t s);
// Looks up the fixed scratch of the current device, `device` (once per device).
hipError_t multipass_static_scratch_init(int device);
uint64_t multipass_static_scratch_bytes();

// Frames per chunk (both passes run a chunk before the next one starts):
// as many as fit `budget` bytes of scratch, at least one.
uint64_t multipass_chunk_frames(uint64_t budget, int log_m);

// Radix split of 2^log_m into the two passes (returns 2; 0 outside 13..19).
int multipass_plan(int log_m, int* log_radix /* [4] */);

}  // namespace ostft
