// engine.hpp -- internal interface between the C ABI (capi.hip) and the
// kernels (fused_small.hip, tables.hip, multipass.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ostft {

// How a complex FFT of size M is fed from real frames.
enum PackMode : int32_t {
    // z = frame(2P) + i * frame(2P+1): two real frames of N = M samples share
    // one complex transform; the spectra are split in the epilogue.
    PACK_PAIR = 0,
    // z[n] = x[2n] + i * x[2n+1] of one frame of N = 2M samples: the split gives
    // [DFT_M(even samples), DFT_M(odd samples)], which is exactly the radix-2
    // Stockham buffer one stage before the end (the reference's odd-window result).
    PACK_EVEN_ODD = 1,
};

// output layouts (values of enum ostft_output in occastft_hip.h)
enum OutMode : int32_t {
    OUT_COMPAT_POWER = 0,
    OUT_HALF_COMPLEX = 1,
    OUT_MAGNITUDE = 2,
    OUT_POWER = 3,
    OUT_HALF_SPECTRUM = 4,
};

struct FusedArgs {
    const float* in;      // signal, L floats
    uint64_t L;
    uint64_t hop;         // OMove
    uint64_t frames;      // frames to transform
    uint64_t groups;      // complex transforms: ceil(frames/2) (pair) or frames (even/odd)
    const float* win;     // N window coefficients, or nullptr for none
    const float2* tw;     // W_M^j, j in [0, M)
    float* out;
    uint64_t out_limit;   // HALF_COMPLEX: complex values to write (OHalfSize)
    int32_t pack;         // PackMode
    int32_t out_mode;     // OutMode
    int32_t dc;           // subtract the frame mean after windowing
};

// Single-kernel path: complex size M = 2^log_m, 6 <= log_m <= 13.
bool fused_supported(int log_m);
hipError_t launch_fused(int log_m, const FusedArgs& a, hipStream_t s);
// The largest group count one launch handles (grid-size bound).
uint64_t fused_max_groups(int log_m);

// FP64 table generators (run once per table).
// tw[j] = exp(-2 pi i j / m) for j < count (count <= m)
hipError_t launch_twiddle_table(float2* tw, int64_t m, int64_t count, hipStream_t s);
// window formulas of OKL/kernel.okl:33-127 evaluated in FP64, rounded once
hipError_t launch_window_table(float* w, int64_t n, int kind, float sigma, hipStream_t s);

}  // namespace ostft
