/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference Serial runner's STFT pipeline (float32,
 * same float/double promotions and the same stage order) plus an FP64 truth
 * pipeline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as a checker or a timed CPU baseline.
 * The product (occastft_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef OCCASTFT_ORACLE_H
#define OCCASTFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* window kinds, in the order Default_Pipeline tests them
 * (StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:216-247) */
enum {
    ORC_WIN_NONE = 0,
    ORC_WIN_HANNING = 1,
    ORC_WIN_HAMMING = 2,
    ORC_WIN_BLACKMAN = 3,
    ORC_WIN_NUTTALL = 4,
    ORC_WIN_BLACKMAN_NUTTALL = 5,
    ORC_WIN_BLACKMAN_HARRIS = 6,
    ORC_WIN_FLATTOP = 7,
    ORC_WIN_GAUSSIAN = 8
};

/* toQuot, RunnerInterface.hpp:26-35 (float32 arithmetic) */
int32_t oracle_toquot(float fullSize, float overlapRatio, float windowSize);

/* geometry block of SerialImpl.cpp:36-43 */
void oracle_geometry(uint32_t fullSize, int32_t exp, float overlap,
                     int32_t* qtConst, uint32_t* oFullSize,
                     uint32_t* oHalfSize, uint32_t* oMove);

/* single window coefficient, OKL/kernel.okl:33-127 (Serial translation
 * StandAlone/cross_gpgpu/Serial/kernel/compiled.hpp:20-88) */
float oracle_window_coef(int32_t kind, int32_t index, int32_t windowSize, float sigma);

/* stage kernels */
void oracle_overlap(const float* in, uint32_t oFullSize, uint32_t fullSize,
                    uint32_t exp, uint32_t oMove, float* outReal);
void oracle_apply_window(float* data, uint32_t oFullSize, uint32_t windowSize,
                         int32_t kind, float sigma);
void oracle_remove_dc(float* data, uint32_t oFullSize, uint32_t windowSize);
void oracle_stockham_stage(const float* inRe, const float* inIm,
                           float* outRe, float* outIm,
                           uint32_t hWindowSize, uint32_t stage,
                           uint32_t oHalfSize, uint32_t exp);
void oracle_to_power(float* out, const float* re, const float* im, uint32_t oFullSize);
void oracle_half_complex(float* out, const float* re, const float* im,
                         uint32_t oHalfSize, int32_t exp);

/* option dispatch of Default_Pipeline (RunnerCommon.cpp:216-248, 309-314).
 * Returns 0, or 1 when the Gaussian sigma text does not parse
 * (the pipeline's "--gaussian_window" error string). */
typedef struct oracle_opts {
    int32_t window;     /* ORC_WIN_* actually applied (NONE if sigma <= 0) */
    float sigma;
    int32_t remove_dc;
    int32_t half_complex;
} oracle_opts;
int32_t oracle_parse_options(const char* options, oracle_opts* out);

/* The whole Serial ActivateSTFT (SerialImpl.cpp:30-75) over Default_Pipeline
 * (RunnerCommon.cpp:184-316) with the Serial runner's EXPC / ToPower call
 * conventions (Serial/src/functionImpl.cpp:207-304).  `out` must hold
 * OFullSize floats.  Returns 0 for "OK", 1 for an option error, 2 for
 * "Operation not supported" (exp < 6), 3 for sizes that overflow 32 bits. */
int32_t oracle_pipeline(const float* in, uint32_t fullSize, int32_t exp,
                        float overlap, const char* options, float* out);

/* ---- FP64 truth ------------------------------------------------------- */

/* naive O(n^2) forward DFT, X[k] = sum x[j] exp(-2 pi i j k / n) */
void oracle_dft64(const double* re, const double* im, int64_t n,
                  double* outRe, double* outIm);
/* iterative radix-2 forward FFT in place (n a power of two) */
void oracle_fft64(double* re, double* im, int64_t n);

/* FP64 pipeline: overlap -> window (FP64 formula) -> optional DC removal ->
 * exact DFT, then the requested output:
 *   out_mode 0: compat power  sqrt(2)*|Re X[k]|, N per frame   (Serial ToPower)
 *   out_mode 1: compat half-complex buffer half                 (toHalfComplexFormat)
 *   out_mode 2: magnitude |X[k]|, N per frame                    (OKL toPower)
 *   out_mode 3: power |X[k]|^2, k = 0..N/2
 *   out_mode 4: half spectrum (Re, Im), k = 0..N/2
 * quirk_odd != 0 and exp odd >= 13: modes 0,1,2 read the penultimate radix-2
 * stage, i.e. [DFT_{N/2}(even samples), DFT_{N/2}(odd samples)].
 * Frames are `frames` (use the float32 toQuot count), hop `hop`.
 * Returns the number of doubles written. */
int64_t oracle_stft64(const float* in, uint64_t fullSize, int32_t exp,
                      uint64_t hop, uint64_t frames, int32_t window, double sigma,
                      int32_t remove_dc, int32_t out_mode, int32_t quirk_odd,
                      double* out);

/* FP64 window coefficient (same formulas, all in double) */
double oracle_window_coef64(int32_t kind, int64_t index, int64_t windowSize, double sigma);

#ifdef __cplusplus
}
#endif
#endif
