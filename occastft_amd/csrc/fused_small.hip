// fused_small.hip -- one kernel per window-size class for windows N = 2^6 ..
// 2^14: overlap gather + window + DC removal + FFT + output epilogue, each
// frame read from HBM once and its spectrum written once.
//
// What it replaces (all in one launch): Overlap_Common, Window_*,
// DCRemove_Common (OKL/EXPCommon.okl:7-192), Stockhoptimized6..11
// (OKL/EXP6.okl..EXP11.okl) or the EXP x StockHamCommon passes
// (OKL/EXPCommon.okl:194-245), and toPower / toHalfComplexFormat
// (OKL/kernel.okl:178-218).
//
// Layout of the work:
//   * A frame of N real samples is one complex transform of M = N/2 points,
//     z[n] = x[2n] + i x[2n+1] (loaded as float2, coalesced).  A group of
//     T = M/16 threads owns it, 16 complex values per thread in VGPRs.
//   * Stockham passes of radix 16 (plus one radix 2/4/8 pass) in registers;
//     between passes the group exchanges through LDS (padded by one complex
//     per 16).  The last pass leaves Z in LDS in natural order; the epilogue
//     reads Z[k] and Z[M-k], forms the even/odd half spectra, applies the last
//     radix-2 stage (or stops before it for the reference's odd-exponent
//     result) and writes positions k and M+k of the frame, coalesced.
//   * Frames never share a transform, so a frame's result is bitwise the same
//     whatever else is in the launch.
//   * Workgroups of max(256, T) threads; blocks are renumbered so each XCD
//     walks a contiguous range of frames (its own L2 sees the overlap halo).
#include "engine.hpp"
#include "epilogue.hpp"
#include "stockham.hpp"

namespace ostft {

template <int LOG_M>
__global__ void __launch_bounds__(Cfg<LOG_M>::WG) stft_fused_kernel(const FusedArgs a)
{
    using C = Cfg<LOG_M>;
    constexpr int M = C::M, T = C::T, G = C::G, PADM = C::PADM;
    __shared__ float2 lds[G * PADM];
    __shared__ float2 red[C::WG / 64];

    const uint32_t tid = threadIdx.x;
    const uint32_t grp = tid / T, t = tid % T;
    const uint64_t f = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * G + grp;  // frame
    const bool active = f < a.frames;
    float2* z = lds + grp * PADM;
    const float* __restrict__ in = a.in;
    const uint64_t L = a.L;

    // ---- gather + window (Overlap_Common + Window_*) ----
    float2 v[16];
    const uint64_t s0 = f * a.hop;
    const bool aligned = ((s0 | (reinterpret_cast<uintptr_t>(in) >> 2)) & 1) == 0;
    const bool inside = active && s0 + 2 * M <= L;
    const float2* __restrict__ w2 = reinterpret_cast<const float2*>(a.win);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const uint32_t n = t + m * T;
        const uint64_t s = s0 + 2u * n;
        float2 x;
        if (inside && aligned) {
            x = *reinterpret_cast<const float2*>(in + s);
        } else {
            x.x = (active && s < L) ? in[s] : 0.0f;
            x.y = (active && s + 1 < L) ? in[s + 1] : 0.0f;
        }
        if (w2) {
            const float2 w = w2[n];
            x = make_float2(x.x * w.x, x.y * w.y);
        }
        v[m] = x;
    }

    // ---- DC removal of the windowed frame (DCRemove_Common) ----
    if (a.dc) {
        float2 s = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            s = cadd(s, v[m]);
        s = group_sum<T, C::WG>(s, red, tid);
        const float mean = (s.x + s.y) / (float)(2 * M);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            v[m] = make_float2(v[m].x - mean, v[m].y - mean);
    }

    // ---- FFT of the packed frame ----
    stockham<LOG_M, 0>(v, z, a.tw, t);

    // ---- last radix-2 stage + output (toPower / toHalfComplexFormat) ----
    if (!active)
        return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t k = t + i * T, km = (M - k) & (M - 1);
        const float2 w = a.combine ? a.tw_n[k] : make_float2(1.0f, 0.0f);
        const BinPair b = finish_bin(z[pad(k)], z[pad(km)], w, a.combine, false);
        write_bin(a.out, a.out_mode, f, M, k, b, a.out_limit);
    }
}

template <int LOG_M>
static hipError_t launch_t(const FusedArgs& a, hipStream_t s)
{
    using C = Cfg<LOG_M>;
    const uint64_t blocks = (a.frames + C::G - 1) / C::G;
    if (blocks == 0)
        return hipSuccess;
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(stft_fused_kernel<LOG_M>, dim3((uint32_t)blocks), dim3(C::WG), 0, s, a);
    return hipGetLastError();
}

bool fused_supported(int log_m) { return log_m >= 5 && log_m <= 13; }

hipError_t launch_fused(int log_m, const FusedArgs& a, hipStream_t s)
{
    switch (log_m) {
    case 5: return launch_t<5>(a, s);
    case 6: return launch_t<6>(a, s);
    case 7: return launch_t<7>(a, s);
    case 8: return launch_t<8>(a, s);
    case 9: return launch_t<9>(a, s);
    case 10: return launch_t<10>(a, s);
    case 11: return launch_t<11>(a, s);
    case 12: return launch_t<12>(a, s);
    case 13: return launch_t<13>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ostft
