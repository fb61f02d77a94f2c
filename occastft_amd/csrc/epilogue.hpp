// epilogue.hpp -- from the packed half-length spectrum to the output layouts.
//
// A real frame x[0..N) is transformed as one complex sequence of M = N/2
// points, z[n] = x[2n] + i x[2n+1].  With Z = DFT_M(z):
//     E[k] = (Z[k] + conj Z[M-k]) / 2          = DFT_M(x[0::2])[k]
//     O[k] = (Z[k] - conj Z[M-k]) / (2i)       = DFT_M(x[1::2])[k]
// The last radix-2 stage of the N-point transform is
//     X[k] = E[k] + W_N^k O[k],   X[k+M] = E[k] - W_N^k O[k],
// and the buffer one stage earlier -- what the reference returns for odd
// window exponents >= 13 -- is simply [E, O].  Either way bin k of the packed
// spectrum yields the outputs at positions k and M+k of the frame.
#pragma once

#include "engine.hpp"
#include "fft_device.hpp"

namespace ostft {

// Twice the outputs at positions k and M + k: the halvings of E and O are
// left to the output conversion (one exact scaling by a power of two there,
// folded into a constant for the real layouts), which saves vector ALU work.
struct BinPair {
    float2 lo;  // 2 X[k]
    float2 hi;  // 2 X[M + k]
};

// combine != 0: final radix-2 stage with w = W_N^k; otherwise [E, O].
// dc0 != 0 (only for k == 0): remove the frame mean in the frequency domain,
// which moves (E0, O0) to ((E0-O0)/2, (O0-E0)/2).
__device__ __forceinline__ BinPair finish_bin(float2 zk, float2 zm, float2 w, int combine, bool dc0)
{
    float2 E = make_float2(zk.x + zm.x, zk.y - zm.y);  // 2 E
    float2 O = make_float2(zk.y + zm.y, zm.x - zk.x);  // 2 O
    if (dc0) {
        const float d = 0.5f * (E.x - O.x);
        E = make_float2(d, 0.0f);
        O = make_float2(-d, 0.0f);
    }
    BinPair r;
    if (combine) {
        const float2 t = cmul(w, O);
        r.lo = cadd(E, t);
        r.hi = csub(E, t);
    } else {
        r.lo = E;
        r.hi = O;
    }
    return r;
}

// finish_bin with the stage choice fixed at compile time (no dc0 case)
template <bool COMBINE>
__device__ __forceinline__ BinPair finish_pair(float2 zk, float2 zm, float2 w)
{
    const float2 E = make_float2(zk.x + zm.x, zk.y - zm.y);  // 2 E
    const float2 O = make_float2(zk.y + zm.y, zm.x - zk.x);  // 2 O
    BinPair r;
    if constexpr (COMBINE) {
        const float2 t = cmul(w, O);
        r.lo = cadd(E, t);
        r.hi = csub(E, t);
    } else {
        r.lo = E;
        r.hi = O;
    }
    return r;
}

// the output conversions of a doubled value x2 = 2 X:
// sqrt(Re^2 + Re^2) of the Serial runner's ToPower, as sqrt(2)|Re|
__device__ __forceinline__ float compat_power(float2 x2)
{
    constexpr float HALF_SQRT2 = 1.41421356237309504880f * 0.5f;  // exactly half the float sqrt(2)
    return HALF_SQRT2 * fabsf(x2.x);
}
// |X| with the hardware square root (1 ulp)
// (fused multiply-adds written out, as in cmul)
__device__ __forceinline__ float magnitude(float2 x2)
{
    return 0.5f * __builtin_amdgcn_sqrtf(fmaf(x2.x, x2.x, x2.y * x2.y));
}
__device__ __forceinline__ float power_of(float2 x2) { return 0.25f * fmaf(x2.x, x2.x, x2.y * x2.y); }
__device__ __forceinline__ float2 value_of(float2 x2) { return make_float2(0.5f * x2.x, 0.5f * x2.y); }
__device__ __forceinline__ float2 conj_value_of(float2 x2) { return make_float2(0.5f * x2.x, -0.5f * x2.y); }

// Stores of the multi-pass epilogue: plain by default (L2 merges the short
// runs of neighbouring workgroups into whole lines before they reach HBM),
// streaming with OSTFT_MP_NT=1.
#ifndef OSTFT_MP_NT
#define OSTFT_MP_NT 0
#endif
template <typename T>
__device__ __forceinline__ void mp_store(T v, T* p)
{
#if OSTFT_MP_NT
    nt_store(v, p);
#else
    *p = v;
#endif
}

// Write the outputs of bin k of frame f (frame length N = 2M) from the
// doubled pair b.
//   COMPAT_POWER  sqrt(Re^2 + Re^2) at f*N + k and f*N + M + k
//   MAGNITUDE     |.| at the same places
//   HALF_COMPLEX  (Re, Im) at complex index f*N + k, f*N + M + k, below `limit`
//   POWER         |X|^2 at f*(M+1) + k, and X[M] from k == 0
//   HALF_SPECTRUM (Re, Im) at complex index f*(M+1) + k, and X[M] from k == 0
__device__ __forceinline__ void write_bin(float* __restrict__ out, int mode, uint64_t f, uint32_t M, uint32_t k,
                                          const BinPair& b, uint64_t limit)
{
    const uint64_t base = f * (uint64_t)(2 * M);
    switch (mode) {
    case OUT_COMPAT_POWER:
        mp_store(compat_power(b.lo), out + base + k);
        mp_store(compat_power(b.hi), out + base + M + k);
        break;
    case OUT_MAGNITUDE:
        mp_store(magnitude(b.lo), out + base + k);
        mp_store(magnitude(b.hi), out + base + M + k);
        break;
    case OUT_HALF_COMPLEX: {
        float2* o2 = reinterpret_cast<float2*>(out);
        if (base + k < limit)
            mp_store(value_of(b.lo), o2 + base + k);
        if (base + M + k < limit)
            mp_store(value_of(b.hi), o2 + base + M + k);
        break;
    }
    case OUT_POWER: {
        float* o = out + f * (uint64_t)(M + 1);
        mp_store(power_of(b.lo), o + k);
        if (k == 0)
            mp_store(power_of(b.hi), o + M);
        break;
    }
    case OUT_HALF_SPECTRUM: {
        float2* o = reinterpret_cast<float2*>(out) + f * (uint64_t)(M + 1);
        mp_store(value_of(b.lo), o + k);
        if (k == 0)
            mp_store(value_of(b.hi), o + M);
        break;
    }
    default:
        break;
    }
}

}  // namespace ostft
