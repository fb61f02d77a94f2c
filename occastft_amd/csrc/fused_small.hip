// fused_small.hip -- one kernel per window-size class for complex transform
// sizes M = 2^6 .. 2^13: overlap gather + window + DC removal + Stockham FFT
// + output epilogue, with each frame read from HBM once and its spectrum
// written once.
//
// What it replaces (all in one launch): Overlap_Common, Window_*,
// DCRemove_Common (OKL/EXPCommon.okl:7-192), Stockhoptimized6..11
// (OKL/EXP6.okl..EXP11.okl) or the EXP x StockHamCommon passes
// (OKL/EXPCommon.okl:194-245), and toPower / toHalfComplexFormat
// (OKL/kernel.okl:178-218).
//
// Layout of the work:
//   * A "group" of T = M/16 threads owns one complex transform of size M.
//     Each thread holds 16 complex values in VGPRs.
//   * Real input is packed two-for-one (PACK_PAIR: two frames, or
//     PACK_EVEN_ODD: even/odd samples of one frame), so a real frame costs
//     half a complex transform.
//   * Stockham autosort passes of radix 16 (plus one radix 2/4/8 pass) run in
//     registers; between passes the group exchanges through LDS (padded by
//     one complex per 16 to spread banks).  The last pass leaves the spectrum
//     in LDS in natural order, where the epilogue reads bins k and M-k to
//     split the packed spectra, and writes the output coalesced.
//   * Workgroups of max(256, T) threads; blocks are renumbered so each XCD
//     walks a contiguous range of frames (its own L2 sees the overlap halo).
#include "engine.hpp"
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
    const uint64_t P = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * G + grp;
    const bool active = P < a.groups;
    float2* z = lds + grp * PADM;
    const float* __restrict__ in = a.in;
    const uint64_t L = a.L;

    // ---- gather + window (Overlap_Common + Window_*) ----
    float2 v[16];
    bool second = false;  // does the second packed real sequence exist
    if (a.pack == PACK_PAIR) {
        const uint64_t fa = 2 * P, fb = fa + 1;
        second = active && fb < a.frames;
        const uint64_t sa = fa * a.hop, sb = fb * a.hop;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const uint32_t n = t + m * T;
            const float w = a.win ? a.win[n] : 1.0f;
            const float xa = (active && sa + n < L) ? in[sa + n] : 0.0f;
            const float xb = (second && sb + n < L) ? in[sb + n] : 0.0f;
            v[m] = make_float2(xa * w, xb * w);
        }
    } else {
        second = active;
        const uint64_t s0 = P * a.hop;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const uint32_t n = t + m * T;
            const float2 w = a.win ? reinterpret_cast<const float2*>(a.win)[n] : make_float2(1.0f, 1.0f);
            const uint64_t s = s0 + 2u * n;
            const float x0 = (active && s < L) ? in[s] : 0.0f;
            const float x1 = (active && s + 1 < L) ? in[s + 1] : 0.0f;
            v[m] = make_float2(x0 * w.x, x1 * w.y);
        }
    }

    // ---- DC removal of the windowed frame (DCRemove_Common) ----
    if (a.dc) {
        float2 s = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            s = cadd(s, v[m]);
        s = group_sum<T, C::WG>(s, red, tid);
        float mx, my;
        if (a.pack == PACK_PAIR) {
            mx = s.x / (float)M;
            my = s.y / (float)M;
        } else {
            mx = my = (s.x + s.y) / (float)(2 * M);
        }
#pragma unroll
        for (int m = 0; m < 16; ++m)
            v[m] = make_float2(v[m].x - mx, v[m].y - my);
    }

    // ---- FFT ----
    stockham<LOG_M, 0>(v, z, a.tw, t);

    // ---- split the packed spectra and write (toPower / toHalfComplexFormat) ----
    if (!active)
        return;
    float* __restrict__ out = a.out;
    const uint64_t base = P * (uint64_t)(2 * M);  // first output element of this group (N-per-frame layouts)
    switch (a.out_mode) {
    case OUT_COMPAT_POWER:
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = t + i * T, km = (M - k) & (M - 1);
            const float2 zk = z[pad(k)], zm = z[pad(km)];
            const float ar = 0.5f * (zk.x + zm.x);
            const float br = 0.5f * (zk.y + zm.y);
            nt_store(sqrtf(ar * ar + ar * ar), out + base + k);
            if (second)
                nt_store(sqrtf(br * br + br * br), out + base + M + k);
        }
        break;
    case OUT_MAGNITUDE:
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = t + i * T, km = (M - k) & (M - 1);
            const float2 zk = z[pad(k)], zm = z[pad(km)];
            const float ar = 0.5f * (zk.x + zm.x), ai = 0.5f * (zk.y - zm.y);
            const float br = 0.5f * (zk.y + zm.y), bi = 0.5f * (zm.x - zk.x);
            nt_store(sqrtf(ar * ar + ai * ai), out + base + k);
            if (second)
                nt_store(sqrtf(br * br + bi * bi), out + base + M + k);
        }
        break;
    case OUT_HALF_COMPLEX: {
        float2* __restrict__ o2 = reinterpret_cast<float2*>(out);
        const uint64_t lim = a.out_limit;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = t + i * T, km = (M - k) & (M - 1);
            const float2 zk = z[pad(k)], zm = z[pad(km)];
            const float2 A = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            const float2 B = make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x));
            if (base + k < lim)
                nt_store(A, o2 + base + k);
            if (second && base + M + k < lim)
                nt_store(B, o2 + base + M + k);
        }
        break;
    }
    case OUT_POWER: {
        // pair packing only: frames 2P and 2P+1, N/2+1 bins each
        const uint64_t fa = 2 * P;
        float* oa = out + fa * (uint64_t)(M / 2 + 1);
        float* ob = oa + (M / 2 + 1);
#pragma unroll
        for (int i = 0; i <= 8; ++i) {
            const uint32_t k = t + i * T;
            if (k <= (uint32_t)(M / 2)) {
                const uint32_t km = (M - k) & (M - 1);
                const float2 zk = z[pad(k)], zm = z[pad(km)];
                const float ar = 0.5f * (zk.x + zm.x), ai = 0.5f * (zk.y - zm.y);
                const float br = 0.5f * (zk.y + zm.y), bi = 0.5f * (zm.x - zk.x);
                nt_store(ar * ar + ai * ai, oa + k);
                if (second)
                    nt_store(br * br + bi * bi, ob + k);
            }
        }
        break;
    }
    case OUT_HALF_SPECTRUM: {
        const uint64_t fa = 2 * P;
        float2* oa = reinterpret_cast<float2*>(out) + fa * (uint64_t)(M / 2 + 1);
        float2* ob = oa + (M / 2 + 1);
#pragma unroll
        for (int i = 0; i <= 8; ++i) {
            const uint32_t k = t + i * T;
            if (k <= (uint32_t)(M / 2)) {
                const uint32_t km = (M - k) & (M - 1);
                const float2 zk = z[pad(k)], zm = z[pad(km)];
                nt_store(make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)), oa + k);
                if (second)
                    nt_store(make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)), ob + k);
            }
        }
        break;
    }
    default:
        break;
    }
}

template <int LOG_M>
static hipError_t launch_t(const FusedArgs& a, hipStream_t s)
{
    using C = Cfg<LOG_M>;
    const uint64_t blocks = (a.groups + C::G - 1) / C::G;
    if (blocks == 0)
        return hipSuccess;
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(stft_fused_kernel<LOG_M>, dim3((uint32_t)blocks), dim3(C::WG), 0, s, a);
    return hipGetLastError();
}

bool fused_supported(int log_m) { return log_m >= 6 && log_m <= 13; }

uint64_t fused_max_groups(int log_m)
{
    const uint64_t wg = (1u << log_m) / 16 < 256 ? 256 : (1u << log_m) / 16;
    const uint64_t g = wg / ((1u << log_m) / 16);
    return 0x7fffffffull * g;
}

hipError_t launch_fused(int log_m, const FusedArgs& a, hipStream_t s)
{
    switch (log_m) {
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
