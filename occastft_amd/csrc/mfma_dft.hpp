// mfma_dft.hpp -- the complex transform of the fused kernel recast as small
// dense GEMMs on the matrix cores, with split-precision operands.
//
// For windows N = 64, 256, 1024 (complex size M = N/2 = 32, 128, 512) the
// transform is a Stockham sequence of radix-16 / radix-8 stages on MFMA and
// one final radix-2 stage on the vector ALUs:
//     M = 32 : 16 x 2        M = 128 : 16 x 8        M = 512 : 16 x 16 x 2
// A radix-R stage (stride S, the product of the earlier radices) does, for
// every butterfly b < M/R of every frame (one GEMM column each):
//     V[m] = X[b + m M/R] * W_{SR}^{(b mod S) m}                  (FP32)
//     [Yr; Yi] = [[Fr, -Fi], [Fi, Fr]] [Vr; Vi],  F[q][m] = W_R^{qm}  (MFMA)
//     Y[(b/S) S R + (b mod S) + q S] = Y[q]
// which is a 2R x 2R real matrix times 2R x (columns) real data, tiled into
// 16 x 16 x 16 MFMA blocks (v_mfma_f32_16x16x16_f16, or ..._bf16_1k).
//
// Split precision: each FP32 operand x is written as hi + lo with
// hi = round16(x), lo = round16(x - hi), and the product as
// A_hi B_hi + A_hi B_lo + A_lo B_hi, accumulated in FP32 by the MFMA (the
// A_lo B_lo term is below the FP32 rounding of the sum).  With FP16 the pair
// carries 22 significant bits; with BF16, 16.  Every GEMM column (one
// butterfly of one frame) is scaled by a power of two chosen from its own
// largest component before the split, so its largest value sits at 2^13..2^14
// (FP16 keeps hi and lo normal over 17 octaves below it); the result is scaled
// back exactly.
//
// Lane layouts of the 16x16x16 instructions (lane l, i = l % 16, g = l / 16):
//     A[i][4g + j], B[4g + j][i] (j = 0..3),  D[4g + r][i] (r = 0..3)
#pragma once

#include "fft_device.hpp"
#include "stockham.hpp"

namespace ostft {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

enum Precision : int { PREC_FP16 = 1, PREC_BF16 = 2 };

template <int PREC>
struct Split;

template <>
struct Split<PREC_FP16> {
    using frag = half4_t;
    static __device__ __forceinline__ void split(const float x[4], frag& hi, frag& lo)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const _Float16 h = (_Float16)x[j];
            hi[j] = h;
            lo[j] = (_Float16)(x[j] - (float)h);
        }
    }
    static __device__ __forceinline__ float4_t mma(frag a, frag b, float4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
};

// round to nearest even, as bits (finite inputs)
__device__ __forceinline__ short bf16_bits(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(short b) { return __uint_as_float((uint32_t)(uint16_t)b << 16); }

template <>
struct Split<PREC_BF16> {
    using frag = bf16x4_t;
    static __device__ __forceinline__ void split(const float x[4], frag& hi, frag& lo)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const short h = bf16_bits(x[j]);
            hi[j] = h;
            lo[j] = bf16_bits(x[j] - bf16_value(h));
        }
    }
    static __device__ __forceinline__ float4_t mma(frag a, frag b, float4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};

// 2^(13 - e) for the binary exponent e of m (m > 0 finite), and its inverse:
// the scaled maximum lies in [2^13, 2^14).
__device__ __forceinline__ void pow2_scale(float m, float& s, float& inv)
{
    const int e = (int)((__float_as_uint(m) >> 23) & 0xffu) - 127;
    int k = 13 - e;
    k = k < -120 ? -120 : (k > 120 ? 120 : k);
    s = __uint_as_float((uint32_t)(127 + k) << 23);
    inv = __uint_as_float((uint32_t)(127 - k) << 23);
}

// One radix-R (16 or 8) Stockham stage of M-point transforms on G frames held
// in LDS (frame f at lds + f * PADM, element x at pad(x)), in place.
template <int M, int G, int PADM, int WG, int R, int S, int PREC>
__device__ __forceinline__ void mfma_stage(float2* lds, const float2* __restrict__ tw, const uint32_t tid)
{
    static_assert(R == 16 || R == 8, "MFMA stages are radix 16 or 8");
    using SP = Split<PREC>;
    using frag = typename SP::frag;
    constexpr int NB = M / R;            // butterflies (GEMM columns) per frame
    constexpr int GROUPS = G * NB / 16;  // 16-column blocks
    constexpr int WAVES = WG / 64;
    static_assert(GROUPS % WAVES == 0, "column blocks must divide among the waves");
    constexpr int GPW = GROUPS / WAVES;
    constexpr int NT = 2 * R / 16;       // 16 x 16 tiles per dimension of the 2R x 2R matrix
    const uint32_t lane = tid & 63u, wave = tid >> 6, i = lane & 15u, g = lane >> 4;

    // DFT matrix fragments: lane (i, g) holds A[16 rt + i][16 kt + 4 g + j];
    // entry (q, k) of quadrant (q >= R, k >= R) is Re W, -Im W, Im W, Re W
    // of W = W_R^{(q mod R)(k mod R)} (W_R^x = tw[x M / R])
    frag a_hi[NT][NT], a_lo[NT][NT];
    {
        float2 w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t q = i % R, k = (4u * g + j) % R;
            w[j] = tw[((q * k) % R) * (M / R)];
        }
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
#pragma unroll
            for (int kt = 0; kt < NT; ++kt) {
                float av[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const bool qi = (16 * rt + (int)i) >= R, ki = (16 * kt + 4 * (int)g + j) >= R;
                    av[j] = (!qi && !ki) || (qi && ki) ? w[j].x : (!qi ? -w[j].y : w[j].y);
                }
                SP::split(av, a_hi[rt][kt], a_lo[rt][kt]);
            }
    }

    // gather V for every column block of this wave (all reads before the
    // in-place writes)
    // compile-time loops over the column blocks: u[] must stay in registers
    float2 u[GPW][4];
    cfor<0, GPW>([&](auto gic) {
        constexpr int gi = decltype(gic)::value;
        const uint32_t col = (wave + gi * WAVES) * 16u + i;
        const uint32_t f = col / NB, b = col % NB;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t m = R == 16 ? 4u * g + j : 4u * (g & 1u) + j;
            float2 x = lds[f * PADM + pad(b + m * NB)];
            if constexpr (S > 1)
                x = cmul(x, tw[((b % S) * m) * (M / (S * R))]);
            u[gi][j] = x;
        }
    });
    __syncthreads();

    cfor<0, GPW>([&](auto gic) {
        constexpr int gi = decltype(gic)::value;
        const uint32_t col = (wave + gi * WAVES) * 16u + i;
        const uint32_t f = col / NB, b = col % NB;
        // the column's scale: its largest component over the 4 lanes holding it
        float mx = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            mx = fmaxf(mx, fmaxf(fabsf(u[gi][j].x), fabsf(u[gi][j].y)));
        mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        float sc = 1.0f, inv = 1.0f;
        if (mx > 0.0f)
            pow2_scale(mx, sc, inv);
        // B fragments: B[16 kt + 4 g + j][col] = U[16 kt + 4 g + j],
        // U = [Vr; Vi]
        frag b_hi[NT], b_lo[NT];
#pragma unroll
        for (int kt = 0; kt < NT; ++kt) {
            float bv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool imag = R == 16 ? kt == 1 : g >= 2u;
                bv[j] = (imag ? u[gi][j].y : u[gi][j].x) * sc;
            }
            SP::split(bv, b_hi[kt], b_lo[kt]);
        }
        float4_t acc[NT];
#pragma unroll
        for (int rt = 0; rt < NT; ++rt) {
            acc[rt] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < NT; ++kt) {
                acc[rt] = SP::mma(a_hi[rt][kt], b_hi[kt], acc[rt]);
                acc[rt] = SP::mma(a_hi[rt][kt], b_lo[kt], acc[rt]);
                acc[rt] = SP::mma(a_lo[rt][kt], b_hi[kt], acc[rt]);
            }
        }
        // D[16 rt + 4 g + r][col] = Z[16 rt + 4 g + r], Z = [Yr; Yi]
        float2* fr = lds + f * PADM;
        const uint32_t base = (b / S) * S * R + (b % S);
        if constexpr (R == 16) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t q = 4u * g + r;
                fr[pad(base + q * S)] = make_float2(acc[0][r] * inv, acc[1][r] * inv);
            }
        } else {
            // rows 0..7 are Yr (lanes g < 2), rows 8..15 Yi (lanes g >= 2)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float other = __shfl_xor(acc[0][r], 32, 64);
                if (g < 2u) {
                    const uint32_t q = 4u * g + r;
                    fr[pad(base + q * S)] = make_float2(acc[0][r] * inv, other * inv);
                }
            }
        }
    });
    __syncthreads();
}

// Final radix-2 stage (stride M/2) on the vector ALUs, FP32, in place:
// Y[b] = X[b] + W_M^b X[b + M/2],  Y[b + M/2] = X[b] - W_M^b X[b + M/2].
template <int M, int G, int PADM, int WG>
__device__ __forceinline__ void radix2_last(float2* lds, const float2* __restrict__ tw, const uint32_t tid)
{
    constexpr int H = M / 2;
#pragma unroll
    for (uint32_t e = tid; e < (uint32_t)(G * H); e += WG) {
        const uint32_t f = e / H, b = e % H;
        float2* fr = lds + f * PADM;
        const float2 x0 = fr[pad(b)];
        const float2 x1 = cmul(fr[pad(b + H)], tw[b]);
        fr[pad(b)] = cadd(x0, x1);
        fr[pad(b + H)] = csub(x0, x1);
    }
    __syncthreads();
}

// The whole M-point transform of the G frames in LDS, natural order in and out.
template <int LOG_M, int M, int G, int PADM, int WG, int PREC>
__device__ __forceinline__ void mfma_fft(float2* lds, const float2* __restrict__ tw, const uint32_t tid)
{
    if constexpr (LOG_M == 5) {  // 16 x 2
        mfma_stage<M, G, PADM, WG, 16, 1, PREC>(lds, tw, tid);
        radix2_last<M, G, PADM, WG>(lds, tw, tid);
    } else if constexpr (LOG_M == 7) {  // 16 x 8
        mfma_stage<M, G, PADM, WG, 16, 1, PREC>(lds, tw, tid);
        mfma_stage<M, G, PADM, WG, 8, 16, PREC>(lds, tw, tid);
    } else if constexpr (LOG_M == 9) {  // 16 x 16 x 2
        mfma_stage<M, G, PADM, WG, 16, 1, PREC>(lds, tw, tid);
        mfma_stage<M, G, PADM, WG, 16, 16, PREC>(lds, tw, tid);
        radix2_last<M, G, PADM, WG>(lds, tw, tid);
    } else {
        static_assert(LOG_M == 5 || LOG_M == 7 || LOG_M == 9, "MFMA plans exist for M = 32, 128, 512");
    }
}

}  // namespace ostft
