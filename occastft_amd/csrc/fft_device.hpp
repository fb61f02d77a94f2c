// fft_device.hpp -- register-level FFT building blocks for gfx950 (wave64).
//
// Complex values are float2 (x = real, y = imaginary).  Every size and index
// below is a compile-time constant after inlining, so the arrays live in VGPRs
// and the fixed twiddles of the small DFTs fold into the instruction stream.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace ostft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// The fused multiply-adds are written out, so that every kernel built from
// this code rounds a product the same way: left to the compiler, which of
// the two products goes into the FMA may differ from kernel to kernel (and
// with it the last bit), and frames must be bitwise the same whichever kernel
// ran them.
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

// Streaming stores: outputs are written once and never re-read by the kernel.
#ifndef OSTFT_NT_STORE
#define OSTFT_NT_STORE 1
#endif
__device__ __forceinline__ void nt_store(float v, float* p)
{
#if OSTFT_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
__device__ __forceinline__ void nt_store(float2 v, float2* p)
{
    f32x2 t = {v.x, v.y};
#if OSTFT_NT_STORE
    __builtin_nontemporal_store(t, reinterpret_cast<f32x2*>(p));
#else
    *reinterpret_cast<f32x2*>(p) = t;
#endif
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void nt_store4(float4 v, float4* p)
{
    f32x4 t = {v.x, v.y, v.z, v.w};
#if OSTFT_NT_STORE
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4*>(p));
#else
    *reinterpret_cast<f32x4*>(p) = t;
#endif
}

// compile-time loop: f(std::integral_constant<int, i>) for i in [B, E) step S
template <int B, int E, int S = 1, typename F>
__device__ __forceinline__ void cfor(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        cfor<B + S, E, S>(f);
    }
}

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

template <int R>
constexpr int bitrev(int q)
{
    int r = 0;
    for (int b = 0; b < ilog2(R); ++b)
        if (q & (1 << b))
            r |= 1 << (ilog2(R) - 1 - b);
    return r;
}

// v * W16^K with W16 = exp(-2 pi i / 16), K in [0, 8)
template <int K>
__device__ __forceinline__ float2 mul_w16(float2 v)
{
    constexpr float C1 = 0.92387953251128673848f;  // cos(pi/8)
    constexpr float S1 = 0.38268343236508978178f;  // sin(pi/8)
    constexpr float H = 0.70710678118654752440f;   // sqrt(1/2)
    static_assert(K >= 0 && K < 8, "twiddle exponent");
    if constexpr (K == 0) {
        return v;
    } else if constexpr (K == 4) {  // -i
        return make_float2(v.y, -v.x);
    } else if constexpr (K == 2) {  // (1 - i)/sqrt2
        return make_float2(H * (v.x + v.y), H * (v.y - v.x));
    } else if constexpr (K == 6) {  // (-1 - i)/sqrt2
        return make_float2(H * (v.y - v.x), -H * (v.x + v.y));
    } else {
        constexpr float wr = K == 1 ? C1 : K == 3 ? S1 : K == 5 ? -S1 : -C1;
        constexpr float wi = K == 1 ? -S1 : K == 3 ? -C1 : K == 5 ? -C1 : -S1;
        return cmul(v, make_float2(wr, wi));
    }
}

// One decimation-in-frequency level of a size-R DFT held in v[OFF .. OFF+R).
template <int R, int OFF, int LEN, int NV>
__device__ __forceinline__ void dif_level(float2 (&v)[NV])
{
    if constexpr (LEN >= 2) {
        constexpr int HALF = LEN / 2;
        cfor<0, R, LEN>([&](auto b) {
            constexpr int B = decltype(b)::value;
            cfor<0, HALF>([&](auto j) {
                constexpr int J = decltype(j)::value;
                const float2 a = v[OFF + B + J];
                const float2 c = v[OFF + B + J + HALF];
                v[OFF + B + J] = cadd(a, c);
                v[OFF + B + J + HALF] = mul_w16<J * (16 / LEN)>(csub(a, c));
            });
        });
        dif_level<R, OFF, HALF, NV>(v);
    }
}

// Forward DFT of size R (2, 4, 8 or 16) on v[OFF .. OFF+R), natural order in
// and out: v'[q] = sum_m v[m] exp(-2 pi i q m / R).
template <int R, int OFF, int NV>
__device__ __forceinline__ void dft(float2 (&v)[NV])
{
    static_assert(R == 2 || R == 4 || R == 8 || R == 16, "radix");
    dif_level<R, OFF, R, NV>(v);
    float2 t[R];
    cfor<0, R>([&](auto q) {
        constexpr int Q = decltype(q)::value;
        t[Q] = v[OFF + bitrev<R>(Q)];
    });
    cfor<0, R>([&](auto q) {
        constexpr int Q = decltype(q)::value;
        v[OFF + Q] = t[Q];
    });
}

}  // namespace ostft
