// stockham.hpp -- the LDS-exchanged Stockham engine shared by the fused
// single-kernel path and the multi-pass column kernels.
//
// A "group" of T = M/VPT threads owns one complex transform of size
// M = 2^LOG_M with VPT = 16 (or 8) values per thread in VGPRs.  Passes are
// radix VPT plus one radix 2/4/8 pass for the remainder; pass PASS with radix
// r and stride Ns (product of the earlier radices) does, for every butterfly
// g < M/r,
//     v[m] = X[g + m*M/r] * W_{Ns*r}^{(g mod Ns)*m},   v = DFT_r(v),
//     Y[(g/Ns)*Ns*r + (g mod Ns) + q*Ns] = v[q],
// which leaves the spectrum in natural order (Stockham autosort).
#pragma once

#include "fft_device.hpp"

// smallest workgroup of the fused kernel (threads), and from complex size
// 2^OSTFT_BIG_LOG_M on
#ifndef OSTFT_WG_MIN
#define OSTFT_WG_MIN 256
#endif
#ifndef OSTFT_WG_BIG
#define OSTFT_WG_BIG 256
#endif
#ifndef OSTFT_BIG_LOG_M
#define OSTFT_BIG_LOG_M 99
#endif
#ifndef OSTFT_DPP_SUM
#define OSTFT_DPP_SUM 1
#endif

namespace ostft {

// Shape and radix plan of the engine for complex size M = 2^LOG_M.  Variant
// V = 4 SHAPE + PLAN.  PLAN bit 1: radix-8 passes with 8 values per thread
// (twice the threads per transform, about half the registers per thread),
// else radix-16 passes with 16; bit 0: the remainder radix first, else last.
// SHAPE: the workgroup (0: at least OSTFT_WG_MIN threads, 1: at least 512,
// 2: at least 128).  A transform's arithmetic depends on the plan only; the
// workgroup decides how many transforms share one.
template <int LOG_M, int V = 0>
struct Cfg {
    static constexpr int M = 1 << LOG_M;
    static constexpr int PLAN = V & 3, SHAPE = V >> 2;
    static constexpr int VPT = (PLAN & 2) ? 8 : 16;  // values per thread
    static constexpr int LV = VPT == 16 ? 4 : 3;
    static constexpr int T = M / VPT;              // threads per transform
    static constexpr int WG_MIN = SHAPE == 1   ? 512
                                  : SHAPE == 2 ? 128
                                               : (LOG_M >= OSTFT_BIG_LOG_M ? OSTFT_WG_BIG : OSTFT_WG_MIN);
    static constexpr int WG = T < WG_MIN ? WG_MIN : T;  // threads per workgroup
    static constexpr int G = WG / T;               // transforms per workgroup
    static constexpr int PADM = M + M / 16;        // padded LDS length per transform
    static constexpr int NPV = LOG_M / LV;         // radix-VPT passes
    static constexpr int REM = LOG_M % LV;
    static constexpr int NPASS = NPV + (REM ? 1 : 0);
    static constexpr int radix(int i)
    {
        return ((PLAN & 1) && REM) ? (i == 0 ? (1 << REM) : VPT) : (i < NPV ? VPT : (1 << REM));
    }
    static constexpr int ns(int i)
    {
        int n = 1;
        for (int k = 0; k < i; ++k)
            n *= radix(k);
        return n;
    }
};

__device__ __forceinline__ uint32_t pad(uint32_t i) { return i + (i >> 4); }

// Workgroup-order remap: hardware hands consecutive workgroups to the 8 XCDs
// round-robin; this gives XCD x the contiguous logical range it would own.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nb)
{
    const uint32_t xcd = bid & 7u, idx = bid >> 3, q = nb >> 3, r = nb & 7u;
    return xcd < r ? xcd * (q + 1u) + idx : r * (q + 1u) + (xcd - r) * q + idx;
}

// x plus its value in the partner lane of a butterfly step over aligned
// blocks of 2*OFF lanes.  Every lane of each aligned OFF-block holds the same
// value when the step runs, so any lane of the other half is a valid partner:
// quad permutations for 1 and 2, the half-row and row mirrors (lane i with
// 7-i, 15-i) for 4 and 8 -- all folded into the add as DPP operands -- a
// swizzle for 16 and a lane permute for 32.
template <int OFF>
__device__ __forceinline__ float lane_pair_add(float x)
{
    const int xi = __builtin_bit_cast(int, x);
    int p;
    if constexpr (OFF == 1)
        p = __builtin_amdgcn_update_dpp(0, xi, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    else if constexpr (OFF == 2)
        p = __builtin_amdgcn_update_dpp(0, xi, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (OFF == 4)
        p = __builtin_amdgcn_update_dpp(0, xi, 0x141, 0xF, 0xF, false);  // row_half_mirror
    else if constexpr (OFF == 8)
        p = __builtin_amdgcn_update_dpp(0, xi, 0x140, 0xF, 0xF, false);  // row_mirror
    else if constexpr (OFF == 16)
        p = __builtin_amdgcn_ds_swizzle(xi, 0x401F);  // bit mode: lane ^ 16 within 32
    else
        p = __builtin_bit_cast(int, __shfl_xor(x, OFF, 64));
    return x + __builtin_bit_cast(float, p);
}

// Sum over the T consecutive lanes of a group (all threads must call it).
template <int T, int WG>
__device__ __forceinline__ float2 group_sum(float2 s, float2* red, uint32_t tid)
{
    constexpr int W = T < 64 ? T : 64;
#if OSTFT_DPP_SUM
    cfor<0, ilog2(W)>([&](auto b) {
        constexpr int OFF = 1 << decltype(b)::value;
        s.x = lane_pair_add<OFF>(s.x);
        s.y = lane_pair_add<OFF>(s.y);
    });
#else
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) {
        s.x += __shfl_xor(s.x, off, 64);
        s.y += __shfl_xor(s.y, off, 64);
    }
#endif
    if constexpr (T > 64) {
        constexpr int WPG = T / 64;
        const uint32_t wave = tid >> 6;
        if ((tid & 63u) == 0)
            red[wave] = s;
        __syncthreads();
        const uint32_t w0 = (wave / WPG) * WPG;
        float2 acc = red[w0];
#pragma unroll
        for (int i = 1; i < WPG; ++i)
            acc = cadd(acc, red[w0 + i]);
        s = acc;
    }
    return s;
}

// Twiddles of every pass for one thread.  Pass p (radix r, stride NS > 1)
// multiplies input m of butterfly g = t + u*T by w^m with
// w = W_M^{(g mod NS) * M/(NS r)}; the thread keeps w (and w^4, w^8 for
// radix 8/16) and forms the other powers by products.  The values depend only
// on the thread's index, so they are loaded once, up front, with the tile's
// other global loads.  (Every power from the table instead -- more loads,
// fewer vector instructions -- measured slower from window 1024 on: 0.32
// against 0.24 ms at 4096 with 50% overlap.)
template <int LOG_M, int V = 0>
struct PassTwiddles {
    float2 b[Cfg<LOG_M, V>::NPASS][Cfg<LOG_M, V>::VPT][3];  // [pass][butterfly u][w, w^4, w^8]; unused slots vanish
};

template <int LOG_M, int V = 0>
__device__ __forceinline__ void load_pass_twiddles(PassTwiddles<LOG_M, V>& P, const float2* __restrict__ tw,
                                                   uint32_t t)
{
    using C = Cfg<LOG_M, V>;
    constexpr int M = C::M, T = C::T;
    cfor<1, C::NPASS>([&](auto p) {
        constexpr int PP = decltype(p)::value;
        constexpr int r = C::radix(PP), NS = C::ns(PP), U = C::VPT / r, STEP = M / (NS * r);
        cfor<0, U>([&](auto u) {
            constexpr int UU = decltype(u)::value;
            const uint32_t e1 = ((t + UU * T) & (NS - 1)) * STEP;
            P.b[PP][UU][0] = tw[e1];
            if constexpr (r >= 8)
                P.b[PP][UU][1] = tw[4 * e1];
            if constexpr (r == 16)
                P.b[PP][UU][2] = tw[8 * e1];
        });
    });
}

// The first pass expects, for butterfly u of the thread, v[u r + m] =
// X[t + u T + m M/r]; the gather delivers v[j] = X[t + j T].  For a first
// radix below VPT (the remainder first) that is a fixed renaming of the
// thread's registers.
template <int LOG_M, int V = 0>
__device__ __forceinline__ void first_pass_order(float2 (&v)[16])
{
    using C = Cfg<LOG_M, V>;
    constexpr int r = C::radix(0), U = C::VPT / r;
    if constexpr (r != C::VPT) {
        float2 w[16];
        cfor<0, U>([&](auto u) {
            constexpr int UU = decltype(u)::value;
            cfor<0, r>([&](auto m) {
                constexpr int MM = decltype(m)::value;
                w[UU * r + MM] = v[UU + MM * U];  // X[t + (u + m U) T], m U T = m M/r
            });
        });
        cfor<0, C::VPT>([&](auto j) { v[decltype(j)::value] = w[decltype(j)::value]; });
    }
}

// Stockham pass PASS: twiddle, radix-r DFTs, scatter to LDS; then gather the
// inputs of the next pass, or stop after pass NRUN - 1 with that pass's
// output in LDS (the spectrum when NRUN is every pass).  No global memory is
// touched here.
template <int LOG_M, int PASS, int V = 0, int NRUN = Cfg<LOG_M, V>::NPASS>
__device__ __forceinline__ void stockham(float2 (&v)[16], float2* z, const PassTwiddles<LOG_M, V>& P, uint32_t t)
{
    using C = Cfg<LOG_M, V>;
    constexpr int M = C::M, T = C::T;
    constexpr int r = C::radix(PASS), NS = C::ns(PASS), U = C::VPT / r;
    if constexpr (NS > 1) {
        cfor<0, U>([&](auto u) {
            constexpr int UU = decltype(u)::value;
            float2 w[16];
            w[1] = P.b[PASS][UU][0];
            if constexpr (r >= 4) {
                w[2] = cmul(w[1], w[1]);
                w[3] = cmul(w[2], w[1]);
            }
            if constexpr (r >= 8) {
                w[4] = P.b[PASS][UU][1];
                w[5] = cmul(w[4], w[1]);
                w[6] = cmul(w[4], w[2]);
                w[7] = cmul(w[4], w[3]);
            }
            if constexpr (r == 16) {
                w[8] = P.b[PASS][UU][2];
                cfor<9, 16>([&](auto m) {
                    constexpr int MM = decltype(m)::value;
                    w[MM] = cmul(w[8], w[MM - 8]);
                });
            }
            cfor<1, r>([&](auto m) {
                constexpr int MM = decltype(m)::value;
                v[UU * r + MM] = cmul(v[UU * r + MM], w[MM]);
            });
        });
    }
    cfor<0, U>([&](auto u) {
        constexpr int UU = decltype(u)::value;
        dft<r, UU * r>(v);
    });
    // pad(base + q NS) = pad(base) + pad(q NS): base = B + s with B a
    // multiple of 16 (when NS r >= 16) and s < NS, and s + (q NS mod 16) < 16
    // when NS < 16; base = g r and q < r when NS r < 16.  Likewise
    // pad(g + m S) = pad(g) + pad(m S) for S = M / r2 (a multiple of 16, or
    // g < S with S a divisor of 16).  So each butterfly needs one padded base
    // and the rest are constant offsets.
    cfor<0, U>([&](auto u) {
        constexpr int UU = decltype(u)::value;
        const uint32_t g = t + UU * T;
        const uint32_t base = (g / NS) * (NS * r) + (g & (NS - 1));
        float2* const zb = z + pad(base);
        cfor<0, r>([&](auto q) {
            constexpr int QQ = decltype(q)::value;
            zb[pad(QQ * NS)] = v[UU * r + QQ];
        });
    });
    __syncthreads();
    if constexpr (PASS + 1 < NRUN) {
        constexpr int r2 = C::radix(PASS + 1), U2 = C::VPT / r2;
        cfor<0, U2>([&](auto u) {
            constexpr int UU = decltype(u)::value;
            const float2* const zg = z + pad(t + UU * T);
            cfor<0, r2>([&](auto m) {
                constexpr int MM = decltype(m)::value;
                v[UU * r2 + MM] = zg[pad(MM * (M / r2))];
            });
        });
        __syncthreads();
        stockham<LOG_M, PASS + 1, V, NRUN>(v, z, P, t);
    }
}

}  // namespace ostft
