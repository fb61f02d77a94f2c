// capi.hip -- the C ABI declared in include/occastft_hip.h.
//
// Host side of the engine: option parsing with Default_Pipeline's semantics,
// the reference's float32 frame geometry, resident FP64-generated tables, and
// the choice of kernel path per request.
#include "../../include/occastft_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "batch.hpp"
#include "multipass.hpp"

namespace {
// Staging of one host-buffer call (ostft_activate*): its own stream and
// device buffers, grown on demand and kept.
struct Lane {
    hipStream_t stream = nullptr;
    float* d_in = nullptr;
    uint64_t in_cap = 0;
    float* d_out = nullptr;
    uint64_t out_cap = 0;
    bool busy = false;
};
constexpr size_t MAX_LANES = 4;
}  // namespace

// Concurrency: any number of host threads may call into one context at once
// (the reference runner serves each WebSocket connection on its own thread,
// StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:77-120).  Table caches are
// built under `mu`; host-buffer calls each take a lane; the two-pass scratch
// and the batch descriptor state are shared, and the launches that use them
// are queued under `scratch_mu` / `batch_mu` (the GPU work itself is ordered
// by events, not by these locks).
struct ostft_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // internal: table builds
    std::mutex mu;                 // guards the table caches and `stream`
    std::map<std::pair<int64_t, int64_t>, float2*> twiddles;  // (m, count) -> W_m^j, j < count
    std::map<std::tuple<int, int64_t, uint32_t, int>, float*> windows;  // (kind, N, sigma bits, exact)
    std::map<std::tuple<const float*, int, int>, void*> mfma_tables;  // (window table, log_m, format)
    std::mutex lanes_mu;
    std::condition_variable lanes_cv;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::mutex scratch_mu;
    ostft::Scratch scratch;
    std::mutex batch_mu;
    ostft::BatchState* batch = nullptr;  // batch descriptors, graphs, plans (batch.hip)
    std::mutex err_mu;
    std::unordered_map<std::thread::id, std::string> errors;  // last error per calling thread
    // radix plan of the single-kernel engine per complex size 2^log_m, read
    // once per request: every path (one request, strided, batch, no-alloc)
    // of one context computes a window size with the same plan
    std::atomic<int> radix_plan[16] = {};
};

namespace {

// Restores the caller's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

int fail(ostft_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->errors[std::this_thread::get_id()] = msg;
    }
    return code;
}

int hip_fail(ostft_ctx* ctx, hipError_t e, const char* where)
{
    return fail(ctx, OSTFT_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

bool contains(const char* s, const char* key) { return std::strstr(s, key) != nullptr; }

int transform_of(const ostft_params* p)
{
    if (p->flags & OSTFT_FLAG_MFMA_FP16)
        return ostft::XF_MFMA_FP16;
    if (p->flags & OSTFT_FLAG_MFMA_BF16)
        return ostft::XF_MFMA_BF16;
    return ostft::XF_FP32;
}

// toQuot, StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:26-35: float32 throughout.
#pragma clang fp contract(off)
double to_quot_f32(float fullSize, float overlapRatio, float windowSize)
{
    float q;
    if (overlapRatio == 0.0f)
        q = fullSize / windowSize + 1.0f;
    else
        q = (fullSize) / (windowSize * (1.0f - overlapRatio)) + 1.0f;
    return std::trunc((double)q);
}

uint64_t omove_f32(float windowSize, float overlapRatio)
{
    const float mv = windowSize * (1.0f - overlapRatio);
    return (uint64_t)mv;
}

struct Plan {
    ostft_geometry g;
    int log_m;    // complex transform size M = N/2
    int combine;  // 0: stop before the last radix-2 stage (odd-exponent result)
};

// hop/frames != 0 replace the float32 geometry (the runnerFunction stage
// functions receive OMove and OFullSize already computed by the caller).
int make_plan(const ostft_params* p, uint64_t L, Plan* plan, ostft_ctx* ctx, uint64_t hop_in = 0,
              uint64_t frames_in = 0)
{
    if (!p)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "null params");
    if (p->exp < 6 || p->exp > 20)
        return fail(ctx, OSTFT_ERR_UNSUPPORTED, "Operation not supported: window exponent must be 6..20");
    if (!(p->overlap >= 0.0f && p->overlap < 1.0f))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "overlap ratio must be in [0, 1)");
    if (p->window < OSTFT_WIN_NONE || p->window > OSTFT_WIN_GAUSSIAN)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "unknown window kind");
    if (p->output < OSTFT_OUT_COMPAT_POWER || p->output > OSTFT_OUT_HALF_SPECTRUM)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "unknown output layout");
    if ((p->flags & OSTFT_FLAG_MFMA_FP16) && (p->flags & OSTFT_FLAG_MFMA_BF16))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "choose one of the FP16 and BF16 MFMA forms");
    if ((p->flags & (OSTFT_FLAG_MFMA_FP16 | OSTFT_FLAG_MFMA_BF16)) && !ostft::mfma_supported(p->exp - 1))
        return fail(ctx, OSTFT_ERR_UNSUPPORTED, "the MFMA split-precision form covers windows 64, 256 and 1024");
    ostft_geometry& g = plan->g;
    const uint64_t N = 1ull << p->exp;
    const double q = to_quot_f32((float)L, p->overlap, (float)N);
    if (!(q >= 1.0) || q > 2147483647.0)
        return fail(ctx, OSTFT_ERR_TOO_LARGE, "frame count outside the int32 range of toQuot");
    g.full_size = L;
    g.window_size = N;
    g.frames = (uint64_t)q;
    g.hop = omove_f32((float)N, p->overlap);
    if (hop_in || frames_in) {
        if (!hop_in || !frames_in || frames_in > 2147483647ull)
            return fail(ctx, OSTFT_ERR_INVALID_ARG, "explicit geometry needs hop >= 1 and 1 <= frames < 2^31");
        g.hop = hop_in;
        g.frames = frames_in;
    }
    g.out_full = g.frames * N;
    g.out_half = g.out_full / 2;
    switch (p->output) {
    case OSTFT_OUT_POWER: g.out_floats = g.frames * (N / 2 + 1); break;
    case OSTFT_OUT_HALF_SPECTRUM: g.out_floats = g.frames * (N + 2); break;
    default: g.out_floats = g.out_full; break;
    }
    g.compute_frames = p->output == OSTFT_OUT_HALF_COMPLEX ? (g.out_half + N - 1) / N : g.frames;
    const bool quirk = !(p->flags & OSTFT_FLAG_EXACT_ODD) && p->exp >= 13 && (p->exp & 1)
                       && (p->output == OSTFT_OUT_COMPAT_POWER || p->output == OSTFT_OUT_HALF_COMPLEX
                           || p->output == OSTFT_OUT_MAGNITUDE);
    plan->log_m = p->exp - 1;
    plan->combine = quirk ? 0 : 1;
    return OSTFT_OK;
}

int not_prepared(ostft_ctx* ctx, const char* what)
{
    return fail(ctx, OSTFT_ERR_NOT_PREPARED,
                std::string(what) + " of this window kind and size is not resident yet: run the request once "
                                    "through ostft_run_device (or ostft_activate) first");
}

// W_m^j for j < count, generated once and kept (with `noalloc`: looked up only).
float2* get_twiddles(ostft_ctx* ctx, int64_t m, int64_t count, int* rc, bool noalloc = false)
{
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto key = std::make_pair(m, count);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end())
        return it->second;
    if (noalloc) {
        *rc = not_prepared(ctx, "a twiddle table");
        return nullptr;
    }
    float2* t = nullptr;
    hipError_t e = hipMalloc(&t, sizeof(float2) * (size_t)count);
    if (e == hipSuccess)
        e = ostft::launch_twiddle_table(t, m, count, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (t)
            (void)hipFree(t);
        *rc = hip_fail(ctx, e, "twiddle table");
        return nullptr;
    }
    ctx->twiddles[key] = t;
    return t;
}

// Window coefficients exactly as the reference computes them on the CPU
// (OKL/kernel.okl:33-127 as translated in StandAlone/cross_gpgpu/Serial/kernel/
// compiled.hpp:20-88): the angle is formed in double and stored as float,
// std::cos/std::exp take that float and return float (the C library's cosf /
// expf), the weighted sum is double, the result is float.  Evaluated once per
// (kind, N, sigma) so the table equals the Serial runner's coefficients bit for
// bit; the GPU then reads it from HBM.
#pragma clang fp contract(off)
float compat_window(int kind, int index, int windowSize, float sigma)
{
    constexpr double PI = 3.14159265358979323846;
    switch (kind) {
    case OSTFT_WIN_HANNING: {
        const float angle = 2.0 * PI * index / ((float)(windowSize - 1));
        return 0.5 * (1.0 - std::cos(angle));
    }
    case OSTFT_WIN_HAMMING: {
        const float angle = 2.0 * PI * index / ((float)(windowSize - 1));
        return 0.54 - (0.46 * std::cos(angle));
    }
    case OSTFT_WIN_GAUSSIAN: {
        const int hw = windowSize >> 1;
        float angle = ((float)(index - hw)) / ((float)(hw * sigma));
        angle *= angle;
        angle *= -0.5;
        return std::exp(angle);
    }
    default:
        break;
    }
    const float f = 2.0 * PI * index / ((float)(windowSize));
    const float s = f * 2.0;
    const float t = f * 3.0;
    const float q = f * 4.0;
    switch (kind) {
    case OSTFT_WIN_BLACKMAN: return 0.42 - 0.5 * std::cos(f) + 0.08 * std::cos(s);
    case OSTFT_WIN_NUTTALL:
        return 0.355768 - 0.487396 * std::cos(f) + 0.144232 * std::cos(s) - 0.012604 * std::cos(t);
    case OSTFT_WIN_BLACKMAN_NUTTALL:
        return 0.3635819 - 0.4891775 * std::cos(f) + 0.1365995 * std::cos(s) - 0.0106411 * std::cos(t);
    case OSTFT_WIN_BLACKMAN_HARRIS:
        return 0.35875 - 0.48829 * std::cos(f) + 0.14128 * std::cos(s) - 0.01168 * std::cos(t);
    case OSTFT_WIN_FLATTOP:
        return 0.21557895 - 0.41663158 * std::cos(f) + 0.277263158 * std::cos(s) - 0.083578947 * std::cos(t)
               + 0.006947368 * std::cos(q);
    default: return 1.0f;
    }
}

// Window table for (kind, n, sigma): the reference's coefficients, or with
// `exact` the FP64 formulas generated on the GPU (with `noalloc`: looked up only).
float* get_window(ostft_ctx* ctx, int kind, int64_t n, float sigma, int exact, int* rc, bool noalloc = false)
{
    uint32_t sb = 0;
    std::memcpy(&sb, &sigma, sizeof(sb));
    if (kind != OSTFT_WIN_GAUSSIAN)
        sb = 0;
    const auto key = std::make_tuple(kind, n, sb, exact);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->windows.find(key);
    if (it != ctx->windows.end())
        return it->second;
    if (noalloc) {
        *rc = not_prepared(ctx, "the window table");
        return nullptr;
    }
    float* w = nullptr;
    hipError_t e = hipMalloc(&w, sizeof(float) * (size_t)n);
    if (e == hipSuccess) {
        if (exact) {
            e = ostft::launch_window_table(w, n, kind, sigma, ctx->stream);
        } else {
            std::vector<float> host((size_t)n);
            for (int64_t i = 0; i < n; ++i)
                host[(size_t)i] = compat_window(kind, (int)i, (int)n, sigma);
            e = hipMemcpyAsync(w, host.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(ctx->stream);
        }
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (w)
            (void)hipFree(w);
        *rc = hip_fail(ctx, e, "window table");
        return nullptr;
    }
    ctx->windows[key] = w;
    return w;
}

// The matrix-core form's DFT-matrix fragments for (window table, size,
// format), built on the GPU once and kept (with `noalloc`: looked up only).
void* get_mfma_table(ostft_ctx* ctx, int log_m, int xf, const float* win, const float2* tw_n, int* rc,
                     bool noalloc)
{
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto key = std::make_tuple(win, log_m, xf);
    auto it = ctx->mfma_tables.find(key);
    if (it != ctx->mfma_tables.end())
        return it->second;
    if (noalloc) {
        *rc = not_prepared(ctx, "the matrix-core DFT table");
        return nullptr;
    }
    void* t = nullptr;
    hipError_t e = hipMalloc(&t, ostft::mfma_table_bytes(log_m));
    if (e == hipSuccess)
        e = ostft::launch_mfma_table(log_m, xf, win, tw_n, t, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (t)
            (void)hipFree(t);
        *rc = hip_fail(ctx, e, "matrix-core DFT table");
        return nullptr;
    }
    ctx->mfma_tables[key] = t;
    return t;
}

}  // namespace

namespace ostft {
// used by multipass.hip and batch.hip
float2* ctx_twiddles(ostft_ctx* ctx, int64_t m, int64_t count, int* rc, bool noalloc)
{
    return get_twiddles(ctx, m, count, rc, noalloc);
}
Scratch& ctx_scratch(ostft_ctx* ctx) { return ctx->scratch; }
std::mutex& ctx_scratch_mutex(ostft_ctx* ctx) { return ctx->scratch_mu; }
BatchState*& ctx_batch(ostft_ctx* ctx) { return ctx->batch; }
int ctx_device(const ostft_ctx* ctx) { return ctx->device; }
int ctx_fail(ostft_ctx* ctx, int code, const std::string& msg) { return fail(ctx, code, msg); }
int ctx_hip_fail(ostft_ctx* ctx, hipError_t e, const char* where) { return hip_fail(ctx, e, where); }
int ctx_radix_plan(const ostft_ctx* ctx, int log_m)
{
    return log_m >= 0 && log_m < 16 ? ctx->radix_plan[log_m].load(std::memory_order_relaxed) : 0;
}
void ctx_set_radix_plan(ostft_ctx* ctx, int log_m, int plan)
{
    if (log_m >= 0 && log_m < 16)
        ctx->radix_plan[log_m].store(plan, std::memory_order_relaxed);
}

// Validates one device-resident request and resolves its tables (window,
// twiddles): everything a launch needs, in `a`.  The caller's device is current.
int ctx_prepare(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size, float* d_out,
                uint64_t out_cap, FusedArgs* a, int* log_m, uint64_t hop, uint64_t frames, bool noalloc)
{
    if (!params || !d_out || (!d_in && full_size > 0))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "null argument");
    Plan plan;
    int rc = make_plan(params, full_size, &plan, ctx, hop, frames);
    if (rc != OSTFT_OK)
        return rc;
    if (out_cap < plan.g.out_floats)
        return fail(ctx, OSTFT_ERR_CAPACITY, "output buffer holds " + std::to_string(out_cap) + " floats, need "
                                                 + std::to_string(plan.g.out_floats));
    if ((reinterpret_cast<uintptr_t>(d_out) & 15u) != 0)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "device output must be 16-byte aligned");
    const uint64_t N = plan.g.window_size;
    int win_kind = params->window;
    if (win_kind == OSTFT_WIN_GAUSSIAN && !(params->sigma > 0.0f))
        win_kind = OSTFT_WIN_NONE;  // the reference skips the window for sigma <= 0
    const float* w = nullptr;
    if (win_kind != OSTFT_WIN_NONE) {
        w = get_window(ctx, win_kind, (int64_t)N, params->sigma, (params->flags & OSTFT_FLAG_EXACT_WINDOW) ? 1 : 0,
                       &rc, noalloc);
        if (!w)
            return rc;
    }
    const int64_t M = 1ll << plan.log_m;
    a->in = d_in;
    a->L = full_size;
    a->hop = plan.g.hop;
    a->frames = plan.g.compute_frames;
    a->win = w;
    a->out = d_out;
    a->out_limit = plan.g.out_half;
    a->combine = plan.combine;
    a->out_mode = params->output;
    a->dc = (params->flags & OSTFT_FLAG_REMOVE_DC) ? 1 : 0;
    a->in_stride = 0;
    a->out_stride = 0;
    a->tile_frames = 1;
    a->plan = 0;
    a->reserved = 0;
    if (fused_supported(plan.log_m)) {
        // a tile of G frames stages G N samples; a hop beyond the window (only
        // reachable with an explicit geometry) takes fewer frames per tile, so
        // that (frames - 1) hop + N <= G N
        a->plan = ctx_radix_plan(ctx, plan.log_m);
        if (!fused_plan_ok(plan.log_m, a->plan))
            a->plan = 0;
        a->tile_frames = (int32_t)fused_tile_frames((uint64_t)fused_variant_frames(plan.log_m, a->plan),
                                                    (uint64_t)N, plan.g.hop);
    }
    a->tw = nullptr;
    a->tw_n = get_twiddles(ctx, 2 * M, M, &rc, noalloc);  // W_N^k, k < M
    if (!a->tw_n)
        return rc;
    if (fused_supported(plan.log_m)) {
        a->tw = get_twiddles(ctx, M, M, &rc, noalloc);
        if (!a->tw)
            return rc;
    }
    a->mfma_b = nullptr;
    const int xf = transform_of(params);
    if (xf != XF_FP32 && mfma_gemm_supported(plan.log_m)) {
        a->mfma_b = get_mfma_table(ctx, plan.log_m, xf, w, a->tw_n, &rc, noalloc);
        if (!a->mfma_b)
            return rc;
    }
    *log_m = plan.log_m;
    return OSTFT_OK;
}
}  // namespace ostft

extern "C" {

int ostft_abi_version(void) { return OSTFT_ABI_VERSION; }

const char* ostft_strerror(int status)
{
    switch (status) {
    case OSTFT_OK: return "OK";
    case OSTFT_ERR_INVALID_ARG: return "invalid argument";
    case OSTFT_ERR_UNSUPPORTED: return "Operation not supported";
    case OSTFT_ERR_OPTION: return "--gaussian_window";
    case OSTFT_ERR_CAPACITY: return "output buffer too small";
    case OSTFT_ERR_HIP: return "HIP runtime error";
    case OSTFT_ERR_NO_DEVICE: return "no such HIP device";
    case OSTFT_ERR_TOO_LARGE: return "request too large";
    case OSTFT_ERR_TIMEOUT: return "request still running";
    case OSTFT_ERR_SHUTDOWN: return "server is shutting down";
    case OSTFT_ERR_NOT_PREPARED: return "tables not resident (allocation-free call)";
    default: return "unknown status";
    }
}

const char* ostft_last_error(const ostft_ctx* cctx)
{
    if (!cctx)
        return "";
    ostft_ctx* ctx = const_cast<ostft_ctx*>(cctx);
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    auto it = ctx->errors.find(std::this_thread::get_id());
    // the string lives in a node of the map: its text stays put until this
    // thread records its next error on the context
    return it == ctx->errors.end() ? "" : it->second.c_str();
}

int ostft_window_coefficients(int32_t kind, int32_t n, float sigma, float* out)
{
    if (!out || n < 2 || kind < OSTFT_WIN_NONE || kind > OSTFT_WIN_GAUSSIAN)
        return OSTFT_ERR_INVALID_ARG;
    for (int32_t i = 0; i < n; ++i)
        out[i] = compat_window(kind, i, n, sigma);
    return OSTFT_OK;
}

// Default_Pipeline's option dispatch, RunnerCommon.cpp:216-248 and :309-314.
int ostft_parse_options(const char* options, int32_t exp, float overlap, ostft_params* out)
{
    if (!out)
        return OSTFT_ERR_INVALID_ARG;
    const char* s = options ? options : "";
    ostft_params p;
    std::memset(&p, 0, sizeof(p));
    p.exp = exp;
    p.overlap = overlap;
    p.window = OSTFT_WIN_NONE;
    p.output = OSTFT_OUT_COMPAT_POWER;
    if (contains(s, "--exact_odd_exp"))
        p.flags |= OSTFT_FLAG_EXACT_ODD;
    if (contains(s, "--exact_window"))
        p.flags |= OSTFT_FLAG_EXACT_WINDOW;
    if (contains(s, "--exact_sigma_parse"))
        p.flags |= OSTFT_FLAG_SIGMA_EXACT;
    if (contains(s, "--mfma_fp16"))
        p.flags |= OSTFT_FLAG_MFMA_FP16;
    else if (contains(s, "--mfma_bf16"))
        p.flags |= OSTFT_FLAG_MFMA_BF16;
    static const char* const keys[7] = {"--hanning_window",        "--hamming_window",
                                        "--blackman_window",       "--nuttall_window",
                                        "--blackman_nuttall_window", "--blackman_harris_window",
                                        "--flattop_window"};
    for (int k = 0; k < 7; ++k) {
        if (contains(s, keys[k])) {
            p.window = k + 1;
            break;
        }
    }
    if (p.window == OSTFT_WIN_NONE) {
        const char* gk = std::strstr(s, "--gaussian_window=");
        const char* sk = std::strstr(s, "<<sigma");
        if (gk && sk) {
            // std::string::substr(P1, P2 - P1) with P1 = pos + 19, one past the
            // first sigma character (or pos + 18 with --exact_sigma_parse)
            const size_t len = std::strlen(s);
            const size_t p1 = (size_t)(gk - s) + ((p.flags & OSTFT_FLAG_SIGMA_EXACT) ? 18u : 19u);
            const size_t p2 = (size_t)(sk - s);
            if (p1 > len)
                return OSTFT_ERR_OPTION;
            size_t cnt = p2 - p1;  // wraps like size_t in the reference
            if (cnt > len - p1)
                cnt = len - p1;
            std::string text(s + p1, cnt);
            char* end = nullptr;
            errno = 0;
            const float sigma = std::strtof(text.c_str(), &end);  // std::stof
            if (end == text.c_str() || errno == ERANGE)
                return OSTFT_ERR_OPTION;
            if (sigma > 0.0f) {
                p.window = OSTFT_WIN_GAUSSIAN;
                p.sigma = sigma;
            }
        }
    }
    if (contains(s, "--remove_dc"))
        p.flags |= OSTFT_FLAG_REMOVE_DC;
    if (contains(s, "--half_complex_return"))
        p.output = OSTFT_OUT_HALF_COMPLEX;
    else if (contains(s, "--half_spectrum_return"))
        p.output = OSTFT_OUT_HALF_SPECTRUM;
    else if (contains(s, "--power_return"))
        p.output = OSTFT_OUT_POWER;
    else if (contains(s, "--magnitude_return"))
        p.output = OSTFT_OUT_MAGNITUDE;
    *out = p;
    return OSTFT_OK;
}

int ostft_get_geometry(const ostft_params* params, uint64_t full_size, ostft_geometry* out)
{
    if (!out)
        return OSTFT_ERR_INVALID_ARG;
    Plan plan;
    const int rc = make_plan(params, full_size, &plan, nullptr);
    if (rc == OSTFT_OK)
        *out = plan.g;
    return rc;
}

int ostft_create(int32_t device, ostft_ctx** out)
{
    if (!out)
        return OSTFT_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return OSTFT_ERR_NO_DEVICE;
    DeviceGuard dg(device);
    if (!dg.ok)
        return OSTFT_ERR_NO_DEVICE;
    ostft_ctx* ctx = new ostft_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    // the fixed scratch of the allocation-free entry point: its address on
    // this device, looked up once per device with the device current
    if (e == hipSuccess)
        e = ostft::multipass_static_scratch_init(device);
    if (e != hipSuccess) {
        if (ctx->stream)
            (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return OSTFT_ERR_HIP;
    }
    *out = ctx;
    return OSTFT_OK;
}

void ostft_destroy(ostft_ctx* ctx)
{
    if (!ctx)
        return;
    {
        DeviceGuard dg(ctx->device);
        (void)hipDeviceSynchronize();
        for (auto& kv : ctx->twiddles)
            (void)hipFree(kv.second);
        for (auto& kv : ctx->windows)
            (void)hipFree(kv.second);
        for (auto& kv : ctx->mfma_tables)
            (void)hipFree(kv.second);
        for (auto& ln : ctx->lanes) {
            if (ln->d_in)
                (void)hipFree(ln->d_in);
            if (ln->d_out)
                (void)hipFree(ln->d_out);
            if (ln->stream)
                (void)hipStreamDestroy(ln->stream);
        }
        ctx->scratch.destroy();
        ostft::batch_release(ctx);
        if (ctx->stream)
            (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

}  // extern "C"

namespace {
int run_device_geometry(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size,
                        float* d_out, uint64_t out_cap, void* hip_stream, uint64_t hop, uint64_t frames)
{
    ostft::FusedArgs a;
    int log_m = 0;
    const int rc = ostft::ctx_prepare(ctx, params, d_in, full_size, d_out, out_cap, &a, &log_m, hop, frames);
    if (rc != OSTFT_OK)
        return rc;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    if (a.frames == 0)
        return OSTFT_OK;
    if (ostft::fused_supported(log_m)) {
        hipError_t e = ostft::launch_fused(log_m, transform_of(params), a, 1, s);
        if (e != hipSuccess)
            return hip_fail(ctx, e, "fused STFT launch");
        return OSTFT_OK;
    }
    return ostft::run_multipass(ctx, log_m, a, s);
}
}  // namespace

extern "C" {

int ostft_run_device(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size,
                     float* d_out, uint64_t out_cap, void* hip_stream)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    return run_device_geometry(ctx, params, d_in, full_size, d_out, out_cap, hip_stream, 0, 0);
}

int ostft_run_device_noalloc(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size,
                             float* d_out, uint64_t out_cap, void* hip_stream)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    ostft::FusedArgs a;
    int log_m = 0;
    const int rc = ostft::ctx_prepare(ctx, params, d_in, full_size, d_out, out_cap, &a, &log_m, 0, 0, true);
    if (rc != OSTFT_OK)
        return rc;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    if (a.frames == 0)
        return OSTFT_OK;
    if (ostft::fused_supported(log_m)) {
        // the single kernel needs no scratch
        hipError_t e = ostft::launch_fused(log_m, transform_of(params), a, 1, s);
        return e == hipSuccess ? OSTFT_OK : hip_fail(ctx, e, "fused STFT launch");
    }
    return ostft::run_multipass_static(ctx, ctx->device, log_m, a, s);
}

int ostft_run_strided_device(ostft_ctx* ctx, const ostft_params* params, const float* d_in, uint64_t full_size,
                             uint64_t in_stride, uint32_t count, float* d_out, uint64_t out_stride,
                             void* hip_stream)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    if (count == 0)
        return OSTFT_OK;
    ostft::FusedArgs a;
    int log_m = 0;
    const int rc = ostft::ctx_prepare(ctx, params, d_in, full_size, d_out, out_stride, &a, &log_m);
    if (rc != OSTFT_OK)
        return rc;
    if (out_stride % 4 != 0)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "out_stride must be a multiple of 4 floats");
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    if (a.frames == 0)
        return OSTFT_OK;
    a.in_stride = in_stride;
    a.out_stride = out_stride;
    if (ostft::fused_supported(log_m)) {
        for (uint32_t y0 = 0; y0 < count; y0 += 65535u) {  // grid.y limit
            ostft::FusedArgs ay = a;
            ay.in = d_in + (uint64_t)y0 * in_stride;
            ay.out = d_out + (uint64_t)y0 * out_stride;
            const hipError_t e = ostft::launch_fused(log_m, transform_of(params), ay, std::min(count - y0, 65535u), s);
            if (e != hipSuccess)
                return hip_fail(ctx, e, "strided STFT launch");
        }
        return OSTFT_OK;
    }
    for (uint32_t y = 0; y < count; ++y) {
        ostft::FusedArgs ay = a;
        ay.in = d_in + (uint64_t)y * in_stride;
        ay.out = d_out + (uint64_t)y * out_stride;
        ay.in_stride = ay.out_stride = 0;
        const int r = ostft::run_multipass(ctx, log_m, ay, s);
        if (r != OSTFT_OK)
            return r;
    }
    return OSTFT_OK;
}

int ostft_run_batch_device(ostft_ctx* ctx, const ostft_job* jobs, uint32_t njobs, void* hip_stream)
{
    if (!ctx || (!jobs && njobs > 0))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "null argument");
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    return ostft::run_batch(ctx, jobs, njobs, static_cast<hipStream_t>(hip_stream));
}

int ostft_batch_config(ostft_ctx* ctx, int32_t exp, int32_t blocks_per_cu, int32_t use_graphs, int32_t* current)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    return ostft::batch_config(ctx, exp, blocks_per_cu, use_graphs, current);
}

int ostft_engine_variant(ostft_ctx* ctx, int32_t exp, int32_t shape, int32_t plan, int32_t* cur_shape,
                         int32_t* cur_plan)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    return ostft::engine_variant(ctx, exp, shape, plan, cur_shape, cur_plan);
}

int ostft_autotune_batch(ostft_ctx* ctx, int32_t exp, uint64_t samples, const ostft_params* like,
                         int32_t* best_blocks_per_cu, int32_t* best_shape, int32_t* best_plan, double* best_ms)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    return ostft::autotune_batch(ctx, exp, samples, like, best_blocks_per_cu, best_shape, best_plan, best_ms);
}

int ostft_scratch_budget(ostft_ctx* ctx, uint64_t bytes, uint64_t* current)
{
    if (!ctx)
        return OSTFT_ERR_INVALID_ARG;
    if (bytes) {
        DeviceGuard dg(ctx->device);
        if (!dg.ok)
            return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
        std::lock_guard<std::mutex> lk(ctx->scratch_mu);
        ctx->scratch.budget = bytes;
        if (ctx->scratch.cap * sizeof(float2) > bytes)
            ctx->scratch.release();
    }
    if (current) {
        std::lock_guard<std::mutex> lk(ctx->scratch_mu);
        *current = ctx->scratch.budget;
    }
    return OSTFT_OK;
}

namespace {
int activate_host(ostft_ctx* ctx, const ostft_params& p, const float* in, uint64_t full_size, float* out,
                  uint64_t out_cap, uint64_t* out_len, uint64_t hop, uint64_t frames);
}

int ostft_activate(ostft_ctx* ctx, const float* in, uint64_t full_size, int32_t exp, float overlap,
                   const char* options, float* out, uint64_t out_cap, uint64_t* out_len)
{
    if (!ctx || !out || (!in && full_size > 0))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "null argument");
    ostft_params p;
    int rc = ostft_parse_options(options, exp, overlap, &p);
    if (rc != OSTFT_OK)
        return fail(ctx, rc, "--gaussian_window");
    return activate_host(ctx, p, in, full_size, out, out_cap, out_len, 0, 0);
}

int ostft_activate_geometry(ostft_ctx* ctx, const ostft_params* params, const float* in, uint64_t full_size,
                            uint64_t hop, uint64_t frames, float* out, uint64_t out_cap, uint64_t* out_len)
{
    if (!ctx || !params || !out || (!in && full_size > 0))
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "null argument");
    if (!hop || !frames)
        return fail(ctx, OSTFT_ERR_INVALID_ARG, "explicit geometry needs hop >= 1 and frames >= 1");
    return activate_host(ctx, *params, in, full_size, out, out_cap, out_len, hop, frames);
}

}  // extern "C"

namespace {
// Takes a free lane (creating one, up to MAX_LANES, or waiting for one).
Lane* take_lane(ostft_ctx* ctx, hipError_t* err)
{
    std::unique_lock<std::mutex> lk(ctx->lanes_mu);
    for (;;) {
        for (auto& ln : ctx->lanes)
            if (!ln->busy) {
                ln->busy = true;
                return ln.get();
            }
        if (ctx->lanes.size() < MAX_LANES) {
            std::unique_ptr<Lane> ln(new Lane());
            *err = hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking);
            if (*err != hipSuccess)
                return nullptr;
            ln->busy = true;
            ctx->lanes.push_back(std::move(ln));
            return ctx->lanes.back().get();
        }
        ctx->lanes_cv.wait(lk);
    }
}

void give_lane(ostft_ctx* ctx, Lane* ln)
{
    {
        std::lock_guard<std::mutex> lk(ctx->lanes_mu);
        ln->busy = false;
    }
    ctx->lanes_cv.notify_one();
}

hipError_t grow(float** buf, uint64_t* cap, uint64_t need)
{
    if (*cap >= need)
        return hipSuccess;
    if (*buf)
        (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const hipError_t e = hipMalloc(buf, sizeof(float) * need);
    if (e == hipSuccess)
        *cap = need;
    else
        *buf = nullptr;
    return e;
}

int activate_host(ostft_ctx* ctx, const ostft_params& p, const float* in, uint64_t full_size, float* out,
                  uint64_t out_cap, uint64_t* out_len, uint64_t hop, uint64_t frames)
{
    int rc;
    Plan plan;
    rc = make_plan(&p, full_size, &plan, ctx, hop, frames);
    if (rc != OSTFT_OK)
        return rc;
    if (out_cap < plan.g.out_floats)
        return fail(ctx, OSTFT_ERR_CAPACITY, "output buffer too small");
    DeviceGuard dg(ctx->device);
    if (!dg.ok)
        return fail(ctx, OSTFT_ERR_NO_DEVICE, "hipSetDevice failed");
    hipError_t e = hipSuccess;
    Lane* ln = take_lane(ctx, &e);
    if (!ln)
        return hip_fail(ctx, e, "host staging stream");
    // the lane's buffers and stream are this call's alone until give_lane
    struct Give {
        ostft_ctx* c;
        Lane* l;
        ~Give() { give_lane(c, l); }
    } give{ctx, ln};
    e = grow(&ln->d_in, &ln->in_cap, full_size > 0 ? full_size : 1);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "input staging");
    e = grow(&ln->d_out, &ln->out_cap, plan.g.out_floats > 0 ? plan.g.out_floats : 1);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "output staging");
    if (full_size > 0) {
        e = hipMemcpyAsync(ln->d_in, in, sizeof(float) * full_size, hipMemcpyHostToDevice, ln->stream);
        if (e != hipSuccess)
            return hip_fail(ctx, e, "H2D copy");
    }
    rc = run_device_geometry(ctx, &p, ln->d_in, full_size, ln->d_out, ln->out_cap, ln->stream, hop, frames);
    if (rc != OSTFT_OK) {
        (void)hipStreamSynchronize(ln->stream);  // the lane's buffers may be reused at once
        return rc;
    }
    e = hipMemcpyAsync(out, ln->d_out, sizeof(float) * plan.g.out_floats, hipMemcpyDeviceToHost, ln->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ln->stream);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "D2H copy");
    if (out_len)
        *out_len = plan.g.out_floats;
    return OSTFT_OK;
}
}  // namespace
