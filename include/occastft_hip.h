/*
 * occastft_hip.h -- C ABI of the MI355X (gfx950) STFT engine.
 *
 * This is the boundary a runner process (or any FFI) binds.  It replaces the
 * device half of the reference's runner plugin API:
 *
 *   Runner::ActivateSTFT(VECF&, const int& windowSizeEXP,
 *                        const float& overlapRatio, const std::string& options)
 *       StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:122-127
 *       (Serial: StandAlone/cross_gpgpu/Serial/src/SerialImpl.cpp:30-75,
 *        CUDA:   StandAlone/cross_gpgpu/CUDA/src/CudaImpl.cpp:409-510)
 *   runnerFunction::Default_Pipeline(...)       RunnerInterface.hpp:85-102,
 *                                               RunnerCommon.cpp:184-316
 *   toQuot(...)                                 RunnerInterface.hpp:26-35
 *
 * Conventions: plain C types only; 0 = OSTFT_OK, other values are the
 * OSTFT_ERR_* codes below; no exception crosses this boundary.  The caller
 * owns input and output memory; the context owns device tables and the
 * staging/scratch buffers it grows on demand.
 *
 * Threads: every call except ostft_destroy may be made on one context from
 * several host threads at once, as the reference runner's per-connection
 * threads call ActivateSTFT (StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:
 * 77-120).  Host-buffer calls (ostft_activate*) each take one of up to four
 * staging lanes (own stream and device buffers; a fifth concurrent call waits
 * for a lane); device calls queue on the caller's stream; the two-pass scratch
 * and the batch descriptor blocks are shared, and calls queue their launches
 * on them one at a time, ordered on the GPU by events.  ostft_last_error
 * returns the calling thread's last error on the context.  ostft_destroy must
 * not overlap any other call on the context.
 */
#ifndef OCCASTFT_HIP_H
#define OCCASTFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSTFT_ABI_VERSION 2

/* status codes */
#define OSTFT_OK 0
#define OSTFT_ERR_INVALID_ARG 1    /* null pointer, bad enum, overlap outside [0,1) */
#define OSTFT_ERR_UNSUPPORTED 2    /* window exponent outside 6..20: the reference's
                                      "Operation not supported" (RunnerCommon.cpp:250-253) */
#define OSTFT_ERR_OPTION 3         /* an option string that Default_Pipeline rejects
                                      ("--gaussian_window" sigma parse, RunnerCommon.cpp:224-246) */
#define OSTFT_ERR_CAPACITY 4       /* output buffer smaller than the result */
#define OSTFT_ERR_HIP 5            /* a HIP runtime call failed; see ostft_last_error */
#define OSTFT_ERR_NO_DEVICE 6      /* device index out of range / no GPU */
#define OSTFT_ERR_TOO_LARGE 7      /* sizes beyond what the engine addresses */
#define OSTFT_ERR_TIMEOUT 8        /* ostft_server_wait: the request is still running */
#define OSTFT_ERR_SHUTDOWN 9       /* ostft_server_submit after shutdown began */
#define OSTFT_ERR_NOT_PREPARED 10  /* ostft_run_device_noalloc: a window or twiddle table of
                                      the request is not resident in the context yet */

/* window kinds, in Default_Pipeline's precedence order (RunnerCommon.cpp:216-247) */
enum ostft_window {
    OSTFT_WIN_NONE = 0,
    OSTFT_WIN_HANNING = 1,
    OSTFT_WIN_HAMMING = 2,
    OSTFT_WIN_BLACKMAN = 3,
    OSTFT_WIN_NUTTALL = 4,
    OSTFT_WIN_BLACKMAN_NUTTALL = 5,
    OSTFT_WIN_BLACKMAN_HARRIS = 6,
    OSTFT_WIN_FLATTOP = 7,
    OSTFT_WIN_GAUSSIAN = 8
};

/* output layouts */
enum ostft_output {
    /* Serial/OpenMP runner ToPower: sqrt(Re^2 + Re^2) = sqrt(2)|Re X[k]|, N floats
       per frame (Serial/src/functionImpl.cpp:361-378 passes Real twice) */
    OSTFT_OUT_COMPAT_POWER = 0,
    /* toHalfComplexFormat: the first OHalfSize complex values of the frame-major
       spectrum buffer, interleaved re/im, OFullSize floats (OKL/kernel.okl:196-218) */
    OSTFT_OUT_HALF_COMPLEX = 1,
    /* OKL toPower / CUDA runner: |X[k]|, N floats per frame (OKL/kernel.okl:178-194) */
    OSTFT_OUT_MAGNITUDE = 2,
    /* |X[k]|^2 for k = 0..N/2, N/2+1 floats per frame (new) */
    OSTFT_OUT_POWER = 3,
    /* (Re, Im) X[k] for k = 0..N/2, N+2 floats per frame (new; README.md:231's intent) */
    OSTFT_OUT_HALF_SPECTRUM = 4
};

/* flags */
#define OSTFT_FLAG_REMOVE_DC 0x1u      /* "--remove_dc" */
#define OSTFT_FLAG_EXACT_ODD 0x2u      /* odd exponents >= 13 return the final FFT stage;
                                          without it they return the penultimate radix-2
                                          stage, as every reference runner does
                                          (RunnerCommon.cpp:254-305 + Serial/src/functionImpl.cpp:281-336) */
#define OSTFT_FLAG_EXACT_WINDOW 0x4u   /* window = FP64 formula generated on the GPU, rounded
                                          once; without it the coefficients are the reference's
                                          own float/double evaluation (ostft_window_coefficients) */
#define OSTFT_FLAG_SIGMA_EXACT 0x8u    /* parse "--gaussian_window=<s><<sigma" from the first
                                          character of <s>; without it the first character is
                                          skipped as in RunnerCommon.cpp:228 */
#define OSTFT_FLAG_MFMA_FP16 0x10u     /* "--mfma_fp16": the transform as split-precision
                                          FP16 GEMMs on the matrix cores (hi/lo operands,
                                          3 products, FP32 accumulation); windows 64, 256,
                                          1024 only */
#define OSTFT_FLAG_MFMA_BF16 0x20u     /* "--mfma_bf16": the same with BF16 operands (16
                                          significant bits: about 1e-5 relative error) */

typedef struct ostft_params {
    int32_t exp;      /* window size N = 2^exp, 6 <= exp <= 20 */
    float overlap;    /* overlap ratio in [0, 1) */
    int32_t window;   /* enum ostft_window */
    float sigma;      /* Gaussian sigma (> 0) */
    int32_t output;   /* enum ostft_output */
    uint32_t flags;   /* OSTFT_FLAG_* */
} ostft_params;

typedef struct ostft_geometry {
    uint64_t full_size;      /* L, samples in the signal */
    uint64_t window_size;    /* N */
    uint64_t frames;         /* qtConst = toQuot(L, overlap, N), float32 arithmetic */
    uint64_t hop;            /* OMove = (unsigned)(N * (1 - overlap)) */
    uint64_t out_full;       /* OFullSize = frames * N */
    uint64_t out_half;       /* OHalfSize = OFullSize / 2 */
    uint64_t out_floats;     /* floats the selected output layout writes */
    uint64_t compute_frames; /* frames the engine transforms for that layout */
} ostft_geometry;

typedef struct ostft_ctx ostft_ctx;

/* Parse a reference option string (substring matches, Default_Pipeline's
 * precedence).  exp/overlap are copied into the result.  New options, ignored
 * by the reference runners: "--magnitude_return", "--power_return",
 * "--half_spectrum_return", "--exact_odd_exp", "--exact_window",
 * "--exact_sigma_parse", "--mfma_fp16", "--mfma_bf16".  Returns OSTFT_OK or OSTFT_ERR_OPTION. */
int ostft_parse_options(const char* options, int32_t exp, float overlap, ostft_params* out);

/* Frame geometry of SerialImpl.cpp:36-43 / RunnerInterface.hpp:26-35 in 64 bits. */
int ostft_get_geometry(const ostft_params* params, uint64_t full_size, ostft_geometry* out);

/* Create / destroy a context bound to HIP device `device`. */
int ostft_create(int32_t device, ostft_ctx** out);
void ostft_destroy(ostft_ctx* ctx);

/* Device-resident STFT: `d_in` (L floats) and `d_out` (out_cap floats, 16-byte
 * aligned) are device pointers on the context's device; work is queued on `hip_stream`
 * (a hipStream_t, or NULL for the null stream) and this call does not wait
 * for it.  The first call for a new (window, exp) builds the window and
 * FP64-generated twiddle tables synchronously.  Windows 2^16..2^20 pass
 * through the context's scratch buffer: a call on another stream than the
 * previous such call first makes its stream wait for that call's launches.
 * Replaces the device part of Runner::ActivateSTFT. */
int ostft_run_device(ostft_ctx* ctx, const ostft_params* params, const float* d_in,
                     uint64_t full_size, float* d_out, uint64_t out_cap, void* hip_stream);

/* ostft_run_device that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  Same arguments,
 * kernels and results (bitwise) as ostft_run_device, with two differences:
 *   - the window and twiddle tables of the request must already be resident
 *     in the context (one ostft_run_device or ostft_activate call with the
 *     same window kind, size and sigma builds them); otherwise the call
 *     returns OSTFT_ERR_NOT_PREPARED and launches nothing;
 *   - windows 2^16..2^20 pass their intermediate spectra through a fixed
 *     64 MiB device array per GPU (frames in chunks of 64 MiB / (4 N) bytes:
 *     16 frames at 2^20) instead of the context's growable scratch, with no
 *     event.  Every call on a GPU shares that array and it keeps the last
 *     call's contents (each call overwrites what it reads first): calls that
 *     overlap in time on one GPU race; keep them on one stream or ordered.
 * Windows up to 2^15 use no scratch. */
int ostft_run_device_noalloc(ostft_ctx* ctx, const ostft_params* params, const float* d_in,
                             uint64_t full_size, float* d_out, uint64_t out_cap, void* hip_stream);

/* Strided batch of `count` equal requests: signal y is d_in + y * in_stride
 * (full_size floats), its result goes to d_out + y * out_stride (out_stride a
 * multiple of 4 floats, at least the layout's out_floats).  One launch (for
 * windows up to 2^15); results are bitwise those of ostft_run_device on each
 * signal.  No counterpart in the reference (one signal per ActivateSTFT). */
int ostft_run_strided_device(ostft_ctx* ctx, const ostft_params* params, const float* d_in,
                             uint64_t full_size, uint64_t in_stride, uint32_t count, float* d_out,
                             uint64_t out_stride, void* hip_stream);

/* Host-buffer STFT with the reference's ActivateSTFT signature folded into C:
 * parses `options`, copies `in` to the device, runs, copies the result back
 * and waits.  *out_len receives the floats written (OFullSize for the
 * reference layouts).  Replaces Runner::ActivateSTFT (RunnerInterface.hpp:122-127). */
int ostft_activate(ostft_ctx* ctx, const float* in, uint64_t full_size, int32_t exp,
                   float overlap, const char* options, float* out, uint64_t out_cap,
                   uint64_t* out_len);

/* ---- batches: many requests in one launch sequence ----
 *
 * No counterpart in the reference, which runs one request per
 * Runner::ActivateSTFT (StandAlone/cross_gpgpu/Serial/src/SerialImpl.cpp:30-75);
 * each request here has exactly the semantics of ostft_run_device. */
typedef struct ostft_job {
    const float* in;       /* device pointer, full_size floats */
    uint64_t full_size;
    float* out;            /* device pointer, 16-byte aligned, out_cap floats */
    uint64_t out_cap;
    ostft_params params;
} ostft_job;

/* Runs `njobs` device-resident requests of any window size, overlap, window
 * and output layout: one launch per window size present (windows 2^6..2^15)
 * over a device-side queue of tiles of frames, either one workgroup per tile
 * or a persistent grid whose workgroups pull tiles from a counter; larger
 * windows and requests of the matrix-core form (OSTFT_FLAG_MFMA_*) run after
 * them, request by request.  Every
 * request's result is bitwise identical to running it alone with
 * ostft_run_device.  All requests are validated before anything is launched;
 * a bad one fails the call (its index is in ostft_last_error).  Queued on
 * `hip_stream`, no wait.  Outputs must not overlap. */
int ostft_run_batch_device(ostft_ctx* ctx, const ostft_job* jobs, uint32_t njobs, void* hip_stream);

/* Batch launch form of window exponent `exp` (6..15): blocks_per_cu > 0
 * selects the persistent form with that many workgroups per CU (clamped to
 * the occupancy limit; it also sets workgroup shape 0, the only one the
 * persistent kernel is built for), < 0 the one-workgroup-per-tile form (the
 * default), 0 leaves it; use_graphs 0/1 turns replay of captured launch
 * sequences off/on (-1 leaves it; exp 0 sets only that); *current receives
 * the setting in force (0 for one workgroup per tile). */
int ostft_batch_config(ostft_ctx* ctx, int32_t exp, int32_t blocks_per_cu, int32_t use_graphs,
                       int32_t* current);

/* Engine variant of window exponent `exp` (6..15) on this context.
 *   plan   radix plan of the FP32 single-kernel transform: 0 radix-16 passes
 *          (16 values per thread) with the remainder radix (2, 4 or 8) last,
 *          1 the remainder first; 2 and 3 the same with radix-8 passes (8
 *          values per thread, twice the threads per transform; exponents
 *          8..14).  Plans 1 and 3 need a remainder: plan 1 is not built for
 *          exponents 9 and 13, plan 3 not for 10 and 13.  Every FP32 request
 *          of that window on the context uses the plan (ostft_run_device,
 *          _noalloc, _strided, batches), so batch results stay bitwise equal
 *          to single requests; plans agree to rounding, not bitwise.
 *   shape  frames per workgroup of the batch kernels: 0 the default (256
 *          threads, or one transform when it needs more), 1 512 threads
 *          (exponents up to 13), 2 128 threads (up to 12); plans 2 and 3 have
 *          shape 0 only.  Shapes 1 and 2 run in the one-workgroup-per-tile
 *          form.
 * -1 leaves a setting; *cur_shape / *cur_plan (may be NULL) receive what is
 * in force.  A plan or shape not built for `exp` is OSTFT_ERR_UNSUPPORTED.
 * No counterpart in the reference, whose radix plan is fixed per size. */
int ostft_engine_variant(ostft_ctx* ctx, int32_t exp, int32_t shape, int32_t plan, int32_t* cur_shape,
                         int32_t* cur_plan);

/* Times window exponent `exp` on a synthetic ragged batch of `samples`
 * samples under every built engine variant (radix plan x workgroup shape)
 * and launch form (one workgroup per tile; persistent at every
 * workgroups-per-CU setting, shape 0), keeps the fastest for this context
 * and reports it (*best_blocks_per_cu = 0 for one workgroup per tile; any
 * output pointer may be NULL).  The synthetic requests take the overlap,
 * window, output layout and flags of `like` (its exp is ignored; a
 * matrix-core flag is OSTFT_ERR_INVALID_ARG), or with `like` NULL overlap
 * 0.5, Hann, DC removal and the reference power layout: variants rank
 * differently under different layouts, so tune for the requests the
 * context will serve.  The default variant stays unless a candidate is more
 * than 2% faster, in the sweep and again when the two are timed
 * alternately.  The plan it keeps applies to every FP32 request of that
 * window on the context (see ostft_engine_variant); while it runs, other
 * calls on the context keep the previous plan. */
int ostft_autotune_batch(ostft_ctx* ctx, int32_t exp, uint64_t samples, const ostft_params* like,
                         int32_t* best_blocks_per_cu, int32_t* best_shape, int32_t* best_plan, double* best_ms);

/* Scratch budget of the two-pass path (windows 2^16..2^20), in bytes: its
 * frames go through both passes in chunks whose intermediate spectra fit the
 * budget (at least one frame per chunk).  bytes = 0 leaves it; a budget below
 * the scratch already held frees that scratch.  *current (may be NULL)
 * receives the budget in force.  Default 1 GiB.  No counterpart in the
 * reference, whose runners hold whole-signal intermediate buffers. */
int ostft_scratch_budget(ostft_ctx* ctx, uint64_t bytes, uint64_t* current);

/* ---- serving: continuous batching of host-buffer requests ----
 *
 * No counterpart in the reference, whose runner serves each WebSocket request
 * synchronously on its connection thread (StandAlone/cross_gpgpu/GLOBAL/
 * RunnerCommon.cpp:77-120).  A server owns a context, a device arena and a
 * pinned host mirror of it.  Submitted requests wait in FIFO order; a
 * scheduler thread admits as many as fit in the arena (by their input +
 * output footprint) into the next batch while earlier batches are still on
 * the GPU, copies them in, runs them with ostft_run_batch_device and copies
 * the results out; a completion thread hands results to the callers.  Up to
 * max_inflight batches overlap (copies in, kernels, copies out on separate
 * streams).  All functions are thread-safe. */
typedef struct ostft_server ostft_server;

typedef struct ostft_server_config {
    uint64_t arena_bytes;   /* device arena for admitted requests; 0: 1 GiB */
    uint32_t max_batch;     /* requests per batch; 0: 512 */
    uint32_t max_inflight;  /* batches on the GPU at once; 0: 3 */
    int32_t use_graphs;     /* 1 (default when the config is NULL): replay captured launches */
    int32_t autotune;       /* 1: autotune every batch window size (ostft_autotune_batch) at creation */
    const ostft_params* autotune_like; /* the requests to tune for (NULL: ostft_autotune_batch's default);
                                          read at creation only */
    uint64_t autotune_samples;  /* samples of each tuning batch (32 requests); 0: 2^26 */
} ostft_server_config;

typedef struct ostft_server_stats {
    uint64_t submitted, completed, failed, batches;
    uint64_t frames;          /* frames transformed by completed requests */
    uint64_t bytes_in, bytes_out;
    double latency_p50_ms;    /* submit -> result in the caller's buffer, completed requests */
    double latency_p99_ms;
    double latency_max_ms;
    double latency_mean_ms;
    double mean_batch;        /* requests per batch */
    double span_s;            /* first submit -> last completion */
} ostft_server_stats;

int ostft_server_create(int32_t device, const ostft_server_config* config, ostft_server** out);
/* Finishes every submitted request, then releases everything. */
void ostft_server_destroy(ostft_server* srv);
/* Queues one request.  `in` and `out` are host memory that must stay valid
 * until ostft_server_wait returns for the ticket.  Parameters, capacity and
 * footprint are checked here (the request's own error code is returned and no
 * ticket is issued). */
int ostft_server_submit(ostft_server* srv, const float* in, uint64_t full_size, const ostft_params* params,
                        float* out, uint64_t out_cap, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request; returns its status
 * (OSTFT_ERR_TIMEOUT if it is not done; the ticket stays valid then) and the
 * floats written.  A ticket is released once its status has been returned. */
int ostft_server_wait(ostft_server* srv, uint64_t ticket, int32_t timeout_ms, uint64_t* out_len);
/* Counters and latency percentiles since creation (or the last reset). */
int ostft_server_stats_get(ostft_server* srv, ostft_server_stats* out, int32_t reset);
const char* ostft_server_last_error(const ostft_server* srv);

/* ostft_activate with the frame geometry given instead of derived from the
 * overlap ratio: `hop` (OMove) and `frames` (qtConst = OFullSize / N), as the
 * reference's per-stage functions receive them (runnerFunction::Overlap and
 * ToPower / HalfComplex, StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:38-83).
 * params->overlap is not used.  Used by the record-and-fuse implementation of
 * those functions (occastft_runner_functions.hpp). */
int ostft_activate_geometry(ostft_ctx* ctx, const ostft_params* params, const float* in,
                            uint64_t full_size, uint64_t hop, uint64_t frames, float* out,
                            uint64_t out_cap, uint64_t* out_len);

/* Host-side window coefficients the engine uses by default for `kind`, size n:
 * the reference's own float/double evaluation (OKL/kernel.okl:33-127 as
 * translated in StandAlone/cross_gpgpu/Serial/kernel/compiled.hpp:20-88).
 * Needs no GPU.  Returns OSTFT_OK or OSTFT_ERR_INVALID_ARG. */
int ostft_window_coefficients(int32_t kind, int32_t n, float sigma, float* out);

/* Text for a status code; and the last error detail recorded on a context. */
const char* ostft_strerror(int status);
const char* ostft_last_error(const ostft_ctx* ctx);

/* ABI version (OSTFT_ABI_VERSION). */
int ostft_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* OCCASTFT_HIP_H */
