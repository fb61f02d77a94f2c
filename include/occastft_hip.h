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
 * staging/scratch buffers it grows on demand.  A context is not reentrant:
 * calls on one context must not overlap in time (use one context per host
 * thread).  Calls on different contexts may run concurrently.
 */
#ifndef OCCASTFT_HIP_H
#define OCCASTFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSTFT_ABI_VERSION 1

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
 * "--exact_sigma_parse".  Returns OSTFT_OK or OSTFT_ERR_OPTION. */
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
 * FP64-generated twiddle tables synchronously.  Replaces the device part of Runner::ActivateSTFT. */
int ostft_run_device(ostft_ctx* ctx, const ostft_params* params, const float* d_in,
                     uint64_t full_size, float* d_out, uint64_t out_cap, void* hip_stream);

/* Host-buffer STFT with the reference's ActivateSTFT signature folded into C:
 * parses `options`, copies `in` to the device, runs, copies the result back
 * and waits.  *out_len receives the floats written (OFullSize for the
 * reference layouts).  Replaces Runner::ActivateSTFT (RunnerInterface.hpp:122-127). */
int ostft_activate(ostft_ctx* ctx, const float* in, uint64_t full_size, int32_t exp,
                   float overlap, const char* options, float* out, uint64_t out_cap,
                   uint64_t* out_len);

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
