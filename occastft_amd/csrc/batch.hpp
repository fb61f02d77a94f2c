// batch.hpp -- many requests in one launch per window-size class (batch.hip).
#pragma once

#include <string>

#include "../../include/occastft_hip.h"
#include "engine.hpp"

struct ostft_ctx;

namespace ostft {

struct BatchState;

// accessors into the context (defined in capi.hip)
BatchState*& ctx_batch(ostft_ctx* ctx);
int ctx_device(const ostft_ctx* ctx);
int ctx_fail(ostft_ctx* ctx, int code, const std::string& msg);
int ctx_hip_fail(ostft_ctx* ctx, hipError_t e, const char* where);
int ctx_radix_plan(const ostft_ctx* ctx, int log_m);
void ctx_set_radix_plan(ostft_ctx* ctx, int log_m, int plan);
int ctx_prepare(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size, float* d_out,
                uint64_t out_cap, FusedArgs* a, int* log_m, uint64_t hop = 0, uint64_t frames = 0,
                bool noalloc = false);

// The batch entry points behind the C ABI (the context's device is current).
int run_batch(ostft_ctx* ctx, const ostft_job* jobs, uint32_t njobs, hipStream_t s);
int batch_config(ostft_ctx* ctx, int exp, int blocks_per_cu, int use_graphs, int* current);
int engine_variant(ostft_ctx* ctx, int exp, int shape, int plan, int* cur_shape, int* cur_plan);
int autotune_batch(ostft_ctx* ctx, int exp, uint64_t samples, const ostft_params* like, int* best_blocks_per_cu,
                   int* best_shape, int* best_plan, double* best_ms);
void batch_release(ostft_ctx* ctx);

}  // namespace ostft
