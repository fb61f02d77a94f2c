// fused_small.hip -- one kernel per window-size class for windows N = 2^6 ..
// 2^15: overlap gather + window + DC removal + FFT + output epilogue, each
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
//     z[n] = x[2n] + i x[2n+1].  A group of T = M/16 threads owns it, 16
//     complex values per thread in VGPRs.
//   * Stockham passes of radix 16 (plus one radix 2/4/8 pass) in registers;
//     between passes the group exchanges through LDS (padded by one complex
//     per 16).  The last pass leaves Z in LDS in natural order; the epilogue
//     reads Z[k] and Z[M-k], forms the even/odd half spectra, applies the last
//     radix-2 stage (or stops before it for the reference's odd-exponent
//     result) and writes positions k and M+k of the frame -- and, by the
//     conjugate symmetry of a real frame's spectrum, positions M-k and N-k.
//   * Frames never share a transform, so a frame's result is bitwise the same
//     whatever else is in the launch.
//   * A workgroup of max(256, T) threads owns G = WG/T consecutive frames.
//     It stages their whole input span (overlap halo read once) in LDS with
//     16-byte coalesced loads, and writes their whole output block with
//     16-byte coalesced streaming stores; the FFT reuses the staging LDS.
//   * Blocks are renumbered so each XCD walks a contiguous range of frames.
#include "engine.hpp"
#include "epilogue.hpp"
#include "stockham.hpp"

namespace ostft {

#ifndef OSTFT_SETPRIO
#define OSTFT_SETPRIO 0
#endif
#ifndef OSTFT_PAIR_LDS
#define OSTFT_PAIR_LDS 1
#endif
// complex sizes from 2^OSTFT_DIRECT_LOG_M on load each frame's samples straight
// into registers (8-byte loads per thread, the overlap halo from L2) instead of
// staging the workgroup's span in LDS
#ifndef OSTFT_DIRECT_LOG_M
#define OSTFT_DIRECT_LOG_M 99
#endif
// complex sizes from 2^OSTFT_SYM_LOG_M on take the real N-value output
// layouts from half the bins, by the conjugate symmetry of a real frame's
// spectrum (the complex layout always does)
#ifndef OSTFT_SYM_LOG_M
#define OSTFT_SYM_LOG_M 6
#endif
// Minimum waves per SIMD the persistent batch kernel is compiled for: 4 (a
// 128-register budget, which LDS allows anyway) where that compiles without
// spills, else 1 (the compiler's choice).
template <int LOG_M>
struct BatchWaves {
    static constexpr int value = (LOG_M == 9 || LOG_M == 11 || LOG_M >= 13) ? 1 : 4;
};

// The N/2+1 layouts of windows 64 and 128 in per-frame slots (fused_frames):
// one pair of bins per slot and no division per value, where the flattened
// form evaluated a pair per value.
#ifndef OSTFT_SLOT_SMALL
#define OSTFT_SLOT_SMALL 1
#endif

// Fold a last radix-2 Stockham pass into the epilogue (see fused_frames).
#ifndef OSTFT_FOLD2
#define OSTFT_FOLD2 1
#endif

// The per-frame slot epilogue of the N/2+1 layouts is used from
// OSTFT_SLOT_MIN_T threads per transform (windows of 32 x that with 16 values
// per thread) on.
#ifndef OSTFT_SLOT_MIN_T
#define OSTFT_SLOT_MIN_T 8
#endif
// Stores of the N/2+1 layouts (POWER, HALF_SPECTRUM): plain by default.  A
// frame's M+1 values start at any 4- or 8-byte boundary, so a wave's store
// covers partial cache lines at its ends; plain stores let L2 merge them
// with the neighbouring waves' parts, where streaming stores (OSTFT_HALF_NT=1)
// measured up to 1.9x slower (window 512, 2^27 samples).  The N-value
// layouts store whole 16-byte blocks per lane and keep streaming stores.
#ifndef OSTFT_HALF_NT
#define OSTFT_HALF_NT 0
#endif
template <typename V>
__device__ __forceinline__ void half_store(V v, V* p)
{
#if OSTFT_HALF_NT
    nt_store(v, p);
#else
    *p = v;
#endif
}
__device__ __forceinline__ void half_store4(float4 v, float4* p)
{
#if OSTFT_HALF_NT
    nt_store4(v, p);
#else
    *p = v;
#endif
}

// FusedArgs as the batch kernel reads it from the descriptor block: the same
// layout, with the pointers typed as global memory.  A generic pointer read
// from memory would compile to flat accesses; one cast from a global pointer
// lets the compiler use global accesses, as it does for kernel arguments.
#define OSTFT_GLOBAL __attribute__((address_space(1)))
struct FusedArgsGlobal {
    const float OSTFT_GLOBAL* in;
    uint64_t L, hop, frames;
    const float OSTFT_GLOBAL* win;
    const float2 OSTFT_GLOBAL* tw;
    const float2 OSTFT_GLOBAL* tw_n;
    float OSTFT_GLOBAL* out;
    uint64_t out_limit;
    int32_t combine, out_mode, dc, tile_frames;
    uint64_t in_stride, out_stride;
    const void OSTFT_GLOBAL* mfma_b;
    int32_t plan, reserved;
};
static_assert(sizeof(FusedArgsGlobal) == sizeof(FusedArgs), "FusedArgsGlobal mirrors FusedArgs");
static_assert(offsetof(FusedArgsGlobal, out) == offsetof(FusedArgs, out), "FusedArgsGlobal mirrors FusedArgs");
static_assert(offsetof(FusedArgsGlobal, dc) == offsetof(FusedArgs, dc), "FusedArgsGlobal mirrors FusedArgs");
static_assert(offsetof(FusedArgsGlobal, tile_frames) == offsetof(FusedArgs, tile_frames),
              "FusedArgsGlobal mirrors FusedArgs");

__device__ __forceinline__ FusedArgs from_global(const FusedArgsGlobal& g)
{
    FusedArgs a;
    a.in = (const float*)g.in;
    a.L = g.L;
    a.hop = g.hop;
    a.frames = g.frames;
    a.win = (const float*)g.win;
    a.tw = (const float2*)g.tw;
    a.tw_n = (const float2*)g.tw_n;
    a.out = (float*)g.out;
    a.out_limit = g.out_limit;
    a.combine = g.combine;
    a.out_mode = g.out_mode;
    a.dc = g.dc;
    a.tile_frames = g.tile_frames;
    a.in_stride = g.in_stride;
    a.out_stride = g.out_stride;
    a.mfma_b = (const void*)g.mfma_b;
    a.plan = g.plan;
    a.reserved = 0;
    return a;
}

// complex transform sizes from which the frame mean is removed in the
// frequency domain (below: in the time domain)
#ifndef OSTFT_FREQ_DC_LOG_M
#define OSTFT_FREQ_DC_LOG_M 11
#endif

// LDS float index of staged sample q: 4 pad floats after every 128, which
// keeps 16-byte rows aligned for the vector writes.
__device__ __forceinline__ uint32_t spad(uint32_t q) { return q + ((q >> 7) << 2); }

// The work of one workgroup: frames f0 .. f0+a.tile_frames-1 of request `a`
// (fewer at the end of the request; tile_frames <= G, below G only when the
// hop exceeds the window, so that the tile's span fits the staging buffer).  Shared by the one-request kernel and the persistent
// batch kernel, so a frame's result does not depend on which of them ran it.
template <int LOG_M, int V = 0>
__device__ __forceinline__ void fused_frames(const FusedArgs& a, const uint64_t f0, float2* lds, float2* red,
                                             const uint32_t tid)
{
    using C = Cfg<LOG_M, V>;
    constexpr int M = C::M, N = 2 * M, T = C::T, G = C::G, WG = C::WG, PADM = C::PADM, VPT = C::VPT;
    float* sf = reinterpret_cast<float*>(lds);

    const uint32_t grp = tid / T, t = tid % T;
    const uint32_t nf = (uint32_t)min((uint64_t)a.tile_frames, a.frames - f0);  // frames in this tile
    const bool active = grp < nf;
    const uint64_t L = a.L, hop = a.hop;
    const float* __restrict__ in = a.in;

    // ---- 1. issue every global load of the tile before the first wait ----
#if OSTFT_SETPRIO
    __builtin_amdgcn_s_setprio(3);
#endif
    // (a) the block's input span (Overlap_Common): frames f0 .. f0+nf-1 read
    //     samples [f0*hop, (f0+nf-1)*hop + N), loaded once with 16-byte loads
    //     from a 4-sample-aligned origin, zero past L;
    // (b) the window coefficients of this thread's 16 packed samples;
    // (c) the base twiddles of every Stockham pass;
    // (d) W_N^k for the epilogue bins of this thread.
    constexpr bool DIRECT = LOG_M >= OSTFT_DIRECT_LOG_M;
    const uint64_t s_begin = f0 * hop;
    const uint64_t s_al = s_begin & ~3ull;
    const uint64_t s_end = (f0 + nf - 1) * hop + N;
    const uint32_t nvec = (uint32_t)((s_end - s_al + 3) >> 2);
    const bool vec_ok = (reinterpret_cast<uintptr_t>(in) & 15u) == 0;
    constexpr int MAXVEC = G * N / 4 + 2;
    constexpr int ITER = DIRECT ? 1 : (MAXVEC + WG - 1) / WG;
    float4 x[ITER];
    float2 xd[DIRECT ? 16 : 1];
    if constexpr (DIRECT) {
        // the thread's packed samples (x[2n], x[2n+1]), n = t + m T
        const uint64_t s0 = (f0 + grp) * hop;
        const bool pair_ok = ((reinterpret_cast<uintptr_t>(in + s0)) & 7u) == 0;
#pragma unroll
        for (int m = 0; m < VPT; ++m) {
            const uint64_t s = s0 + 2ull * (t + m * T);
            xd[m] = make_float2(0.0f, 0.0f);
            if (active) {
                if (pair_ok && s + 2 <= L) {
                    xd[m] = *reinterpret_cast<const float2*>(in + s);
                } else {
                    xd[m].x = s < L ? in[s] : 0.0f;
                    xd[m].y = s + 1 < L ? in[s + 1] : 0.0f;
                }
            }
        }
    } else if (vec_ok && s_al + 4ull * nvec <= L) {
        // the whole span is inside the signal: 16-byte loads, no per-load
        // checks (values past nvec are never staged)
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            const uint32_t i = tid + it * WG;
            if (i < nvec)
                x[it] = *reinterpret_cast<const float4*>(in + s_al + 4ull * i);
        }
    } else {
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            const uint32_t i = tid + it * WG;
            const uint64_t s = s_al + 4ull * i;
            x[it] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i < nvec) {
                if (vec_ok && s + 4 <= L) {
                    x[it] = *reinterpret_cast<const float4*>(in + s);
                } else {
                    x[it].x = s < L ? in[s] : 0.0f;
                    x[it].y = s + 1 < L ? in[s + 1] : 0.0f;
                    x[it].z = s + 2 < L ? in[s + 2] : 0.0f;
                    x[it].w = s + 3 < L ? in[s + 3] : 0.0f;
                }
            }
        }
    }
    const float2* __restrict__ w2 = reinterpret_cast<const float2*>(a.win);
    float2 wv[16];
    if (w2) {
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            wv[m] = w2[t + m * T];
    }
    PassTwiddles<LOG_M, V> ptw;
    load_pass_twiddles<LOG_M, V>(ptw, a.tw, t);
    // epilogue of the N-value layouts (COMPAT_POWER, MAGNITUDE, HALF_COMPLEX):
    // a real frame's spectrum has X[N-k] = conj X[k], so bins k0 .. k0+B of
    // the packed spectrum give positions k0 .. k0+B-1 and M+k0 .. M+k0+B-1
    // and, conjugated, the B positions below M-k0 and below N-k0.  A thread
    // takes B = 2 bins per iteration for the complex layout (one float4 of two
    // complex values per block) and B = 4 for the real layouts (one float4 per
    // block), so every store instruction of a wave covers contiguous memory.
    // The real layouts of the smallest window, whose blocks would be half a
    // cache line, take 4 bins per iteration for positions k0.. and M+k0..
    // only.
    const int mode = a.out_mode;
    constexpr bool SYM_REAL = LOG_M >= OSTFT_SYM_LOG_M;
    // A last radix-2 pass is folded into the epilogue (windows whose DC
    // removal precedes the FFT): the epilogue reads Z[k] = X[k mod M/2] +
    // W_M^k X[k mod M/2 + M/2] from the buffer the other passes leave, W_M^k
    // being the square of the W_N^k it holds for the bin anyway.  That saves
    // the pass's LDS round trip and barrier.
    constexpr bool FOLD2 = OSTFT_FOLD2 && !DIRECT && LOG_M < OSTFT_FREQ_DC_LOG_M && C::NPASS >= 2 &&
                           C::radix(C::NPASS - 1) == 2;
    constexpr int QPF = M / 8, QITER = (G * QPF) / WG;  // real layouts: quads
    constexpr int PPF = M / 4, PITER = (G * PPF) / WG;  // complex layout: pairs
    static_assert(QITER * WG == G * QPF && PITER * WG == G * PPF, "blocks divide evenly over the workgroup");
    constexpr int OITER4 = (G * M) / (WG * 4);          // real layouts without the symmetry
    // (the N/2+1-bin layouts: slots k <= M/2 of each frame, one per thread
    // and iteration; at the smallest window two consecutive values of the
    // flattened [frame][bin] index)
    constexpr int OITERS = (G * (M / 2 + 1) + WG - 1) / WG;
    // From T >= 8 (windows from 256) the N/2+1 layouts are written frame by
    // frame: slots k < M/2 of a frame as k = t + j T (j < VPT/2) by its own group
    // when a group spans whole waves (T >= 64), else as k = lane + 64 jj
    // (jj < VPT T/128) by a whole wave, frame after frame; the middle slot k = M/2
    // apart.  No per-slot index division, every address a per-thread base
    // plus a constant, and each store instruction one contiguous run.
    constexpr bool SLOT_FRAME = SYM_REAL && T >= OSTFT_SLOT_MIN_T;
    constexpr bool SLOT_WAVE = SLOT_FRAME && T < 64;
    static_assert(!SLOT_WAVE || VPT * T >= 128, "a wave's 64 lanes take slots k < M/2 in whole runs");
    // Below that (windows 64 and 128) the same slots, with 64 / (M/2) frames
    // side by side in a wave (lane = frame x M/2 + slot)
    constexpr bool SLOT_SMALL = OSTFT_SLOT_SMALL && !SLOT_FRAME && !DIRECT && M / 2 <= 64 && 64 % (M / 2) == 0 &&
                                (64 * VPT / M) % (128 / M) == 0;
    // at the largest transforms (one frame per workgroup, four passes) the
    // slot twiddles are loaded after the FFT, which keeps their registers free
    // through it
    constexpr bool SLOT_LATE_TW = SLOT_FRAME && T >= 512;
    static_assert(!SLOT_FRAME || OITERS >= VPT / 2 + 1, "slot twiddles");
    constexpr int OITERP = (G * (M + 1) + 2 * WG - 1) / (2 * WG);
    constexpr int OITERM = OITERS > OITERP ? OITERS : OITERP;
    constexpr int NEB = OITERM > PITER ? (OITERM > OITER4 ? OITERM : OITER4) : (PITER > OITER4 ? PITER : OITER4);
    const float2* __restrict__ twn = a.tw_n;
    const bool vec_out = mode == OUT_COMPAT_POWER || mode == OUT_MAGNITUDE || mode == OUT_HALF_COMPLEX;
    float2 eb[NEB];  // W_N^k0 for each epilogue iteration
    float2 ec1 = make_float2(1.0f, 0.0f), ec2 = ec1, ec3 = ec1, ec4 = ec1;
#pragma unroll
    for (int it = 0; it < NEB; ++it)
        eb[it] = ec1;
    const bool need_w = a.combine || FOLD2;  // the fold needs W_N^k at every bin
    if (mode == OUT_HALF_COMPLEX && need_w) {
#pragma unroll
        for (int it = 0; it < PITER; ++it)
            eb[it] = twn[2 * ((tid + it * WG) & (PPF - 1))];
        ec1 = twn[1];
        ec2 = twn[2];
    } else if (vec_out && need_w) {
        if constexpr (SYM_REAL) {
#pragma unroll
            for (int it = 0; it < QITER; ++it)
                eb[it] = twn[4 * ((tid + it * WG) & (QPF - 1))];
        } else {
#pragma unroll
            for (int it = 0; it < OITER4; ++it)
                eb[it] = twn[(4 * (tid + it * WG)) & (M - 1)];
        }
        ec1 = twn[1];
        ec2 = twn[2];
        ec3 = twn[3];
        ec4 = twn[4];
    } else if (need_w) {
        if constexpr (SLOT_WAVE) {
            // the slot of lane l and step jj is the same in every frame
#pragma unroll
            for (int jj = 0; jj < VPT * T / 128; ++jj)
                eb[jj] = twn[(tid & 63u) + 64 * jj];
            eb[VPT * T / 128] = twn[M / 2];
        } else if constexpr (SLOT_FRAME) {
            if constexpr (!SLOT_LATE_TW) {
#pragma unroll
                for (int j = 0; j < VPT / 2; ++j)
                    eb[j] = twn[t + j * T];
                eb[VPT / 2] = twn[M / 2];
            }
        } else if constexpr (SLOT_SMALL) {
            // the slot of a lane is the same in every frame it takes
            eb[0] = twn[(tid & 63u) & (uint32_t)(M / 2 - 1)];
            eb[1] = twn[M / 2];
        } else if constexpr (SYM_REAL) {
#pragma unroll
            for (int it = 0; it < OITERS; ++it) {
                const uint32_t e = tid + it * WG;
                if (e < (uint32_t)(G * (M / 2 + 1)))
                    eb[it] = twn[e % (uint32_t)(M / 2 + 1)];
            }
        } else {
#pragma unroll
            for (int it = 0; it < OITERP; ++it) {
                const uint32_t e1 = 2u * (tid + it * WG);
                if (e1 < (uint32_t)(G * (M + 1))) {
                    const uint32_t k1 = e1 % (uint32_t)(M + 1);
                    eb[it] = twn[k1 == (uint32_t)M ? 0u : k1];
                }
            }
            ec1 = twn[1];
        }
    }

#if OSTFT_SETPRIO
    __builtin_amdgcn_s_setprio(0);
#endif
    // ---- 2. each group gathers its frame, packed z[n] = x[2n] + i x[2n+1],
    //         and applies the window (Window_*) ----
    float2 v[16];
    if constexpr (DIRECT) {
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            v[m] = w2 ? make_float2(xd[m].x * wv[m].x, xd[m].y * wv[m].y) : xd[m];
    } else {
        // stage the span in LDS; 4*WG is a multiple of 128, so vector i and
        // vector i + WG are a fixed distance apart in the padded layout
        constexpr uint32_t SROW = 4 * WG + WG / 8;
        float* const dst = sf + spad(4 * tid);
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            if (tid + it * WG < nvec)
                *reinterpret_cast<float4*>(dst + it * SROW) = x[it];
        }
        __syncthreads();

        const uint32_t q0 = (uint32_t)(grp * hop + (s_begin - s_al));
        // an even q0 keeps every (x[2n], x[2n+1]) pair inside one padded row
        const bool pair_ok = (q0 & 1u) == 0;
        // sample q0 + 2(t + mT): with m = KB a + b, 2T m = 2T b + 128 a (T < 64,
        // KB = 64/T) or m (2T), a multiple of 128 (T >= 64, KB = 1), so it is
        // staged at spad(q0 + 2t + 2T b) + a ASTEP -- KB base addresses and
        // immediate offsets instead of one padded index per value
        constexpr int KB = T >= 64 ? 1 : (64 / T < VPT ? 64 / T : VPT);
        constexpr uint32_t ASTEP = T >= 64 ? 2 * T + T / 16 : 132;
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            v[m] = make_float2(0.0f, 0.0f);
        if (active) {
            const uint32_t qt = q0 + 2 * t;
            if (OSTFT_PAIR_LDS && pair_ok) {
                cfor<0, KB>([&](auto bb) {
                    constexpr int B = decltype(bb)::value;
                    const float* const base = sf + spad(qt + 2 * T * B);
                    cfor<0, VPT / KB>([&](auto aa) {
                        constexpr int A = decltype(aa)::value;
                        v[KB * A + B] = *reinterpret_cast<const float2*>(base + A * ASTEP);
                    });
                });
            } else {
                cfor<0, KB>([&](auto bb) {
                    constexpr int B = decltype(bb)::value;
                    const float* const b0 = sf + spad(qt + 2 * T * B);
                    const float* const b1 = sf + spad(qt + 2 * T * B + 1);
                    cfor<0, VPT / KB>([&](auto aa) {
                        constexpr int A = decltype(aa)::value;
                        v[KB * A + B] = make_float2(b0[A * ASTEP], b1[A * ASTEP]);
                    });
                });
            }
        }
        if (w2) {
#pragma unroll
            for (int m = 0; m < VPT; ++m)
                v[m] = make_float2(v[m].x * wv[m].x, v[m].y * wv[m].y);
        }
        __syncthreads();  // the FFT reuses the staging buffer
    }

    // ---- 3. DC removal of the windowed frame (DCRemove_Common) ----
    // Windows below 2^OSTFT_FREQ_DC_LOG_M+1
    // subtract the frame mean here (sum by in-row DPP butterflies); larger
    // ones remove it after the FFT, where subtracting the mean from every
    // sample only changes bin 0.
    constexpr bool FREQ_DC = LOG_M >= OSTFT_FREQ_DC_LOG_M;
    if (!FREQ_DC && a.dc) {
        float2 s = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            s = cadd(s, v[m]);
        s = group_sum<T, WG>(s, red, tid);
        const float mean = (s.x + s.y) / (float)N;
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            v[m] = make_float2(v[m].x - mean, v[m].y - mean);
    }

    // ---- 4. FFT of the packed frame; Z stays in LDS, natural order ----
    {
        first_pass_order<LOG_M, V>(v);
        stockham<LOG_M, 0, V, FOLD2 ? C::NPASS - 1 : C::NPASS>(v, lds + grp * PADM, ptw, t);
        if constexpr (FREQ_DC) {
            // z[n] - mu (1 + i) for all n, mu = (Re Z0 + Im Z0) / N, moves
            // Z0 to ((Re Z0 - Im Z0) / 2, (Im Z0 - Re Z0) / 2)
            if (a.dc) {
                float2* z = lds + grp * PADM;
                if (t == 0 && active) {
                    const float2 z0 = z[0];
                    z[0] = make_float2(0.5f * (z0.x - z0.y), 0.5f * (z0.y - z0.x));
                }
                __syncthreads();
            }
        }
    }

    // ---- 5. last radix-2 stage + output, the whole block's output at once ----
    // The layout and the stage choice are uniform; each combination gets its
    // own straight-line code (no per-value selects between layouts).  Only the
    // reference's odd-exponent result skips the last stage, and the single
    // kernel meets it at exponents 13 and 15 alone.
    float* __restrict__ out = a.out;
    constexpr bool VAR_COMBINE = LOG_M == 12 || LOG_M == 14;
    // Z[k] of a frame's buffer z, k < M (k = M reads as 0); wm = W_M^k, used
    // only when folding
    auto zat = [&](const float2* z, uint32_t k, float2 wm) -> float2 {
        if constexpr (FOLD2) {
            const uint32_t k1 = k & (uint32_t)(M / 2 - 1);
            return cadd(z[pad(k1)], cmul(wm, z[pad(k1 + M / 2)]));
        } else {
            (void)wm;
            return z[pad(k & (uint32_t)(M - 1))];
        }
    };
    auto epilogue = [&](auto mode_c, auto comb_c) {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr bool COMB = decltype(comb_c)::value;
        // mirrored blocks: X[M-k] = conj X[M+k] and X[N-k] = conj X[k] after
        // the last radix-2 stage; for [E, O] (no last stage) E[M-k] = conj E[k]
        // and O[M-k] = conj O[k] go to positions M-k and N-k instead
        constexpr uint32_t hm = COMB ? (uint32_t)M : (uint32_t)N;  // mirror of the M+k block
        constexpr uint32_t lm = COMB ? (uint32_t)N : (uint32_t)M;  // mirror of the k block
        if constexpr (MODE == OUT_HALF_COMPLEX) {
            const uint64_t lim = a.out_limit;  // complex values
            float4* o = reinterpret_cast<float4*>(out);
#pragma unroll
            for (int it = 0; it < PITER; ++it) {
                const uint32_t i = tid + it * WG;
                const uint32_t g = i / PPF, k0 = 2 * (i & (PPF - 1));
                if (g >= nf)
                    break;
                // Z[k0], Z[k0+1] and Z[M-k0-2], Z[M-k0-1] each share a row
                // of the padded layout
                const float2* z = lds + g * PADM;
                const float2* zf = z + pad(k0);
                const float2* zm = z + pad((uint32_t)M - k0 - 2);
                const float2 w1 = cmul(eb[it], ec1), w2 = cmul(eb[it], ec2);
                float2 zk[3], zr[3];  // Z[k0 + j], Z[M - k0 - j]
                if constexpr (FOLD2) {
                    const float2 ws[3] = {eb[it], w1, w2};
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const float2 wm = cmul(ws[j], ws[j]);
                        zk[j] = zat(z, k0 + j, wm);
                        zr[j] = zat(z, (uint32_t)M - k0 - j, make_float2(wm.x, -wm.y));
                    }
                } else {
                    zk[0] = zf[0];
                    zk[1] = zf[1];
                    zk[2] = z[pad(k0 + 2)];
                    zr[0] = z[pad(((uint32_t)M - k0) & (M - 1))];
                    zr[1] = zm[1];
                    zr[2] = zm[0];
                }
                const BinPair b0 = finish_pair<COMB>(zk[0], zr[0], eb[it]);
                const BinPair b1 = finish_pair<COMB>(zk[1], zr[1], w1);
                const BinPair b2 = finish_pair<COMB>(zk[2], zr[2], w2);
                const uint64_t base = (f0 + g) * (uint64_t)N;
                const uint64_t c0 = base + k0, c1 = base + M + k0, c2 = base + hm - k0 - 2, c3 = base + lm - k0 - 2;
                const float2 v0 = value_of(b0.lo), v1 = value_of(b1.lo);
                const float2 h0 = value_of(b0.hi), h1 = value_of(b1.hi);
                const float2 mh2 = conj_value_of(b2.hi), mh1 = conj_value_of(b1.hi);
                const float2 ml2 = conj_value_of(b2.lo), ml1 = conj_value_of(b1.lo);
                if (c0 < lim)
                    nt_store4(make_float4(v0.x, v0.y, v1.x, v1.y), o + c0 / 2);
                if (c1 < lim)
                    nt_store4(make_float4(h0.x, h0.y, h1.x, h1.y), o + c1 / 2);
                if (c2 < lim)
                    nt_store4(make_float4(mh2.x, mh2.y, mh1.x, mh1.y), o + c2 / 2);
                if (c3 < lim)
                    nt_store4(make_float4(ml2.x, ml2.y, ml1.x, ml1.y), o + c3 / 2);
            }
        } else if constexpr (MODE == OUT_COMPAT_POWER || MODE == OUT_MAGNITUDE) {
            auto conv = [](float2 x2) {
                if constexpr (MODE == OUT_COMPAT_POWER)
                    return compat_power(x2);
                else
                    return magnitude(x2);
            };
            if constexpr (SYM_REAL) {
#pragma unroll
                for (int it = 0; it < QITER; ++it) {
                    const uint32_t i = tid + it * WG;
                    const uint32_t g = i / QPF, k0 = 4 * (i & (QPF - 1));
                    if (g >= nf)
                        break;
                    // Z[k0 .. k0+3] and Z[M-k0-4 .. M-k0-1] each share a row of
                    // the padded layout
                    const float2* z = lds + g * PADM;
                    const float2* zf = z + pad(k0);
                    const float2* zm = z + pad((uint32_t)M - k0 - 4);
                    const float2 wk[5] = {eb[it], cmul(eb[it], ec1), cmul(eb[it], ec2), cmul(eb[it], ec3),
                                          cmul(eb[it], ec4)};
                    BinPair b[5];
                    if constexpr (FOLD2) {
#pragma unroll
                        for (int j = 0; j < 5; ++j) {
                            const float2 wm = cmul(wk[j], wk[j]);
                            b[j] = finish_pair<COMB>(zat(z, k0 + j, wm),
                                                     zat(z, (uint32_t)M - k0 - j, make_float2(wm.x, -wm.y)), wk[j]);
                        }
                    } else {
                        b[0] = finish_pair<COMB>(zf[0], z[pad(((uint32_t)M - k0) & (M - 1))], wk[0]);
#pragma unroll
                        for (int j = 1; j < 4; ++j)
                            b[j] = finish_pair<COMB>(zf[j], zm[4 - j], wk[j]);
                        b[4] = finish_pair<COMB>(z[pad(k0 + 4)], zm[0], wk[4]);
                    }
                    // the value of a conjugate is the value itself
                    float lo[5], hi[5];
#pragma unroll
                    for (int j = 0; j < 5; ++j) {
                        lo[j] = conv(b[j].lo);
                        hi[j] = conv(b[j].hi);
                    }
                    float4* o = reinterpret_cast<float4*>(out + (f0 + g) * (uint64_t)N);
                    nt_store4(make_float4(lo[0], lo[1], lo[2], lo[3]), o + k0 / 4);
                    nt_store4(make_float4(hi[0], hi[1], hi[2], hi[3]), o + (M + k0) / 4);
                    nt_store4(make_float4(hi[4], hi[3], hi[2], hi[1]), o + (hm - k0 - 4) / 4);
                    nt_store4(make_float4(lo[4], lo[3], lo[2], lo[1]), o + (lm - k0 - 4) / 4);
                }
            } else {
                const uint32_t nbins = nf * (uint32_t)M;
#pragma unroll
                for (int it = 0; it < OITER4; ++it) {
                    const uint32_t e = (tid + it * WG) * 4;
                    if (e >= nbins)
                        break;
                    const uint32_t g = e >> LOG_M, k0 = e & (M - 1);
                    const float2* z = lds + g * PADM;
                    const float2 wk[4] = {eb[it], cmul(eb[it], ec1), cmul(eb[it], ec2), cmul(eb[it], ec3)};
                    float lo[4], hi[4];
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) {
                        const uint32_t k = k0 + j, km = (M - k) & (M - 1);
                        const float2 wm = cmul(wk[j], wk[j]);  // W_M^k (the fold only)
                        const BinPair b = finish_pair<COMB>(zat(z, k, wm), zat(z, km, make_float2(wm.x, -wm.y)),
                                                            wk[j]);
                        lo[j] = conv(b.lo);
                        hi[j] = conv(b.hi);
                    }
                    float4* o = reinterpret_cast<float4*>(out + (f0 + g) * (uint64_t)N);
                    nt_store4(make_float4(lo[0], lo[1], lo[2], lo[3]), o + k0 / 4);
                    nt_store4(make_float4(hi[0], hi[1], hi[2], hi[3]), o + (M + k0) / 4);
                }
            }
        } else if constexpr (SLOT_WAVE) {
            // POWER / HALF_SPECTRUM, M+1 values per frame, 64 / T frames per
            // wave, one after the other (slots as in the SYM_REAL branch
            // below).  pad(l + 64 jj) = pad(l) + 68 jj and pad(M - 64 jj - l)
            // = PADM - 68 jj - (l + ceil(l/16)).
            constexpr uint32_t FW = 64 / T, JW = VPT * T / 128;
            const uint32_t lane = tid & 63u, wv = tid >> 6;
            const uint32_t lu = lane + ((lane + 15u) >> 4);
#pragma unroll
            for (uint32_t fi = 0; fi < FW; ++fi) {
                const uint32_t g = wv * FW + fi;
                if (g < nf) {
                    const float2* z = lds + g * PADM;
                    const float2* zl = z + pad(lane);
                    const float2* zh = z + (PADM - 68 * (JW - 1)) - lu;  // + 68 (JW - 1 - jj)
                    const uint64_t o = (f0 + g) * (uint64_t)(M + 1);
#pragma unroll
                    for (uint32_t jj = 0; jj < JW; ++jj) {
                        const uint32_t k = lane + 64 * jj;
                        float2 zk, zm;
                        if constexpr (FOLD2) {
                            const float2 wm = cmul(eb[jj], eb[jj]);
                            zk = zat(z, k, wm);
                            zm = zat(z, (uint32_t)M - k, make_float2(wm.x, -wm.y));
                        } else {
                            zk = zl[68 * jj];
                            zm = (jj == 0 && lane == 0) ? z[0] : zh[68 * (JW - 1 - jj)];
                        }
                        const BinPair b = finish_pair<COMB>(zk, zm, eb[jj]);
                        const float2 x2 = (COMB || k == 0) ? b.hi : b.lo;  // position M-k, before conj
                        if constexpr (MODE == OUT_POWER) {
                            half_store(power_of(b.lo), out + o + lane + 64 * jj);
                            half_store(power_of(x2), out + o + (M - lane) - 64 * jj);
                        } else {
                            float2* o2 = reinterpret_cast<float2*>(out) + o;
                            half_store(value_of(b.lo), o2 + lane + 64 * jj);
                            half_store(k == 0 ? value_of(x2) : conj_value_of(x2), o2 + (M - lane) - 64 * jj);
                        }
                    }
                }
            }
            // the middle slots of the wave's frames, one per lane
            const uint32_t gm = wv * FW + lane;
            if (lane < FW && gm < nf) {
                const float2* z = lds + gm * PADM;
                const float2 zh2 = zat(z, M / 2, cmul(eb[JW], eb[JW]));  // W_M^(M/2) = -1 when folding
                const BinPair b = finish_pair<COMB>(zh2, zh2, eb[JW]);
                const uint64_t o = (f0 + gm) * (uint64_t)(M + 1) + M / 2;
                if constexpr (MODE == OUT_POWER)
                    half_store(power_of(b.lo), out + o);
                else
                    half_store(value_of(b.lo), reinterpret_cast<float2*>(out) + o);
            }
        } else if constexpr (SLOT_FRAME) {
            // POWER / HALF_SPECTRUM, M+1 values per frame, one frame per group
            // (slots as in the SYM_REAL branch below).  With T a multiple of 16,
            // pad(t + jT) = pad(t) + 17 jT/16 and pad(M - jT - t) =
            // 17 (M - jT)/16 - (t + ceil(t/16)): every LDS and output address
            // is a per-thread base plus a constant.  (Slot 0's partner M - 0
            // wraps to 0.)
            if constexpr (SLOT_LATE_TW) {
                if (COMB) {
#pragma unroll
                    for (int j = 0; j < VPT / 2; ++j)
                        eb[j] = twn[t + j * T];
                    eb[VPT / 2] = twn[M / 2];
                }
            }
            if (active) {
                constexpr uint32_t PT = T + T / 16;  // padded stride of j
                const float2* z = lds + grp * PADM;
                const float2* zl = z + pad(t);
                const float2* zh = z + (PADM - (VPT / 2 - 1) * PT) - (t + ((t + 15u) >> 4));  // + (VPT/2 - 1 - j) PT
                const uint64_t o = (f0 + grp) * (uint64_t)(M + 1);
#pragma unroll
                for (int j = 0; j < VPT / 2; ++j) {
                    const uint32_t k = t + j * T;
                    float2 zk, zm;
                    if constexpr (FOLD2) {
                        const float2 wm = cmul(eb[j], eb[j]);
                        zk = zat(z, k, wm);
                        zm = zat(z, (uint32_t)M - k, make_float2(wm.x, -wm.y));
                    } else {
                        zk = zl[j * PT];
                        zm = (j == 0 && t == 0) ? z[0] : zh[(VPT / 2 - 1 - j) * PT];
                    }
                    const BinPair b = finish_pair<COMB>(zk, zm, eb[j]);
                    const float2 x2 = (COMB || k == 0) ? b.hi : b.lo;  // position M-k, before conj
                    if constexpr (MODE == OUT_POWER) {
                        float* ol = out + o + t;
                        float* oh = out + o + (M - t);
                        half_store(power_of(b.lo), ol + j * T);
                        half_store(power_of(x2), oh - j * T);
                    } else {
                        float2* ol = reinterpret_cast<float2*>(out) + o + t;
                        float2* oh = reinterpret_cast<float2*>(out) + o + (M - t);
                        half_store(value_of(b.lo), ol + j * T);
                        half_store(k == 0 ? value_of(x2) : conj_value_of(x2), oh - j * T);
                    }
                }
                if (t == 0) {  // slot M/2: its own mirror
                    const float2 zh2 = zat(z, M / 2, cmul(eb[VPT / 2], eb[VPT / 2]));
                    const BinPair b = finish_pair<COMB>(zh2, zh2, eb[VPT / 2]);
                    if constexpr (MODE == OUT_POWER)
                        half_store(power_of(b.lo), out + o + M / 2);
                    else
                        half_store(value_of(b.lo), reinterpret_cast<float2*>(out) + o + M / 2);
                }
            }
        } else if constexpr (SLOT_SMALL) {
            // POWER / HALF_SPECTRUM at windows 64 and 128, M+1 values per
            // frame: a wave's FW frames, FS at a time, lane = frame within the
            // step x M/2 + slot k (slots as in the SLOT_WAVE branch); the
            // middle slots apart, one per lane
            constexpr uint32_t S = M / 2, FS = 64 / S, FW = 64 * VPT / M;
            const uint32_t lane = tid & 63u, wv = tid >> 6;
            const uint32_t k = lane & (S - 1u), fs = lane / S;
            const float2 wm = cmul(eb[0], eb[0]);  // W_M^k (the fold only)
#pragma unroll
            for (uint32_t st = 0; st < FW / FS; ++st) {
                const uint32_t g = wv * FW + st * FS + fs;
                if (g < nf) {
                    const float2* z = lds + g * PADM;
                    const float2 zk = zat(z, k, wm);
                    const float2 zm = FOLD2 ? zat(z, (uint32_t)M - k, make_float2(wm.x, -wm.y))
                                            : z[pad(((uint32_t)M - k) & (M - 1))];
                    const BinPair b = finish_pair<COMB>(zk, zm, eb[0]);
                    const float2 x2 = (COMB || k == 0) ? b.hi : b.lo;  // position M-k, before conj
                    const uint64_t o = (f0 + g) * (uint64_t)(M + 1);
                    if constexpr (MODE == OUT_POWER) {
                        half_store(power_of(b.lo), out + o + k);
                        half_store(power_of(x2), out + o + (M - k));
                    } else {
                        float2* o2 = reinterpret_cast<float2*>(out) + o;
                        half_store(value_of(b.lo), o2 + k);
                        half_store(k == 0 ? value_of(x2) : conj_value_of(x2), o2 + (M - k));
                    }
                }
            }
            const uint32_t gm = wv * FW + lane;
            if (lane < FW && gm < nf) {
                const float2* z = lds + gm * PADM;
                const float2 zh2 = zat(z, M / 2, cmul(eb[1], eb[1]));  // W_M^(M/2) = -1 when folding
                const BinPair b = finish_pair<COMB>(zh2, zh2, eb[1]);
                const uint64_t o = (f0 + gm) * (uint64_t)(M + 1) + M / 2;
                if constexpr (MODE == OUT_POWER)
                    half_store(power_of(b.lo), out + o);
                else
                    half_store(value_of(b.lo), reinterpret_cast<float2*>(out) + o);
            }
        } else if constexpr (SYM_REAL) {
            // (window 128 built with OSTFT_SLOT_SMALL=0, the measured
            // alternative to the slot form above)
            // POWER / HALF_SPECTRUM: M+1 values per frame.  Slot k = 0 .. M/2 of
            // a frame gives positions k and M-k from one pair (Z[k], Z[M-k]):
            // X[k] and X[M-k] = conj X[M+k] after the last radix-2 stage, E[k]
            // and E[M-k] = conj E[k] for [E, O]; position M (from slot 0) is
            // X[M] or O[0].  Consecutive slots sit on consecutive lanes, so each
            // store instruction covers contiguous memory (ascending for k,
            // descending for M-k).
            static_assert(!FOLD2, "no built variant folds here");
            constexpr uint32_t SPF = M / 2 + 1;  // slots per frame
            const uint32_t nslot = nf * SPF;
            const uint64_t obase = f0 * (uint64_t)(M + 1);
#pragma unroll
            for (int it = 0; it < OITERS; ++it) {
                const uint32_t e = tid + it * WG;
                if (e >= nslot)
                    break;
                const uint32_t g = e / SPF, k = e - g * SPF;
                const float2* z = lds + g * PADM;
                const BinPair b = finish_pair<COMB>(z[pad(k)], z[pad(((uint32_t)M - k) & (M - 1))], eb[it]);
                const float2 x1 = b.lo;                            // position k (doubled)
                const float2 x2 = (k == 0 || COMB) ? b.hi : b.lo;  // position M-k, before conj
                const bool second = k != (uint32_t)(M / 2);
                const uint64_t o = obase + (uint64_t)g * (M + 1);
                if constexpr (MODE == OUT_POWER) {
                    half_store(power_of(x1), out + o + k);
                    if (second)
                        half_store(power_of(x2), out + o + (M - k));
                } else {
                    float2* o2 = reinterpret_cast<float2*>(out) + o;
                    half_store(value_of(x1), o2 + k);
                    if (second)
                        half_store(k == 0 ? value_of(x2) : conj_value_of(x2), o2 + (M - k));
                }
            }
        } else {
            // (window 64 built with OSTFT_SLOT_SMALL=0, the measured
            // alternative to the slot form above)
            // POWER / HALF_SPECTRUM at the smallest window: two consecutive
            // values of the flattened [frame][bin] index per thread and
            // iteration, one 8-byte (POWER) or 16-byte (HALF_SPECTRUM) store
            // where aligned (the slot form's runs would be shorter than a cache
            // line)
            const uint32_t nval = nf * (uint32_t)(M + 1);
            const uint64_t obase = f0 * (uint64_t)(M + 1);
            const float2 one = make_float2(1.0f, 0.0f);
#pragma unroll
            for (int it = 0; it < OITERP; ++it) {
                const uint32_t e1 = 2u * (tid + it * WG);
                if (e1 >= nval)
                    break;
                const bool has2 = e1 + 1 < nval;
                const uint32_t g1 = e1 / (uint32_t)(M + 1), k1 = e1 - g1 * (uint32_t)(M + 1);
                const uint32_t kb1 = k1 == (uint32_t)M ? 0u : k1;
                const float2 w1 = eb[it];
                const float2* z1 = lds + g1 * PADM;
                const float2 wm1 = cmul(w1, w1);  // W_M^kb1 (the fold only)
                const BinPair b1 = finish_pair<COMB>(zat(z1, kb1, wm1),
                                                     zat(z1, (M - kb1) & (M - 1), make_float2(wm1.x, -wm1.y)), w1);
                const float2 x1 = k1 == (uint32_t)M ? b1.hi : b1.lo;  // doubled
                // the next value: bin k1 + 1 of the same frame, or bin 0 of the next
                const uint32_t g2 = k1 == (uint32_t)M ? g1 + 1 : g1, k2 = k1 == (uint32_t)M ? 0u : k1 + 1;
                const uint32_t kb2 = k2 == (uint32_t)M ? 0u : k2;
                float2 x2 = make_float2(0.0f, 0.0f);
                if (has2) {
                    const float2 w2 = kb2 != 0u ? cmul(w1, ec1) : one;
                    const float2* z2 = lds + g2 * PADM;
                    const float2 wm2 = cmul(w2, w2);
                    const BinPair b2 = finish_pair<COMB>(zat(z2, kb2, wm2),
                                                         zat(z2, (M - kb2) & (M - 1), make_float2(wm2.x, -wm2.y)), w2);
                    x2 = k2 == (uint32_t)M ? b2.hi : b2.lo;
                }
                const uint64_t o = obase + e1;
                const bool pair = has2 && (o & 1u) == 0;
                if constexpr (MODE == OUT_POWER) {
                    const float p1 = power_of(x1), p2 = power_of(x2);
                    if (pair) {
                        half_store(make_float2(p1, p2), reinterpret_cast<float2*>(out + o));
                    } else {
                        half_store(p1, out + o);
                        if (has2)
                            half_store(p2, out + o + 1);
                    }
                } else {
                    float2* o2 = reinterpret_cast<float2*>(out) + o;
                    const float2 v1 = value_of(x1), v2 = value_of(x2);
                    if (pair) {
                        half_store4(make_float4(v1.x, v1.y, v2.x, v2.y), reinterpret_cast<float4*>(o2));
                    } else {
                        half_store(v1, o2);
                        if (has2)
                            half_store(v2, o2 + 1);
                    }
                }
            }
        }
    };
    auto by_stage = [&](auto mode_c) {
        if constexpr (VAR_COMBINE) {
            if (a.combine)
                epilogue(mode_c, std::true_type{});
            else
                epilogue(mode_c, std::false_type{});
        } else {
            epilogue(mode_c, std::true_type{});
        }
    };
    switch (mode) {
    case OUT_COMPAT_POWER: by_stage(std::integral_constant<int, OUT_COMPAT_POWER>{}); break;
    case OUT_MAGNITUDE: by_stage(std::integral_constant<int, OUT_MAGNITUDE>{}); break;
    case OUT_HALF_COMPLEX: by_stage(std::integral_constant<int, OUT_HALF_COMPLEX>{}); break;
    case OUT_POWER: by_stage(std::integral_constant<int, OUT_POWER>{}); break;
    case OUT_HALF_SPECTRUM: by_stage(std::integral_constant<int, OUT_HALF_SPECTRUM>{}); break;
    default: break;
    }
}

// minimum waves per SIMD the one-request kernel is compiled for
#ifndef OSTFT_FUSED_WAVES
#define OSTFT_FUSED_WAVES 4
#endif

// the one-request kernel, radix plan PLAN (workgroup shape 0)
template <int LOG_M, int PLAN>
__global__ void __launch_bounds__((Cfg<LOG_M, PLAN>::WG)) __attribute__((amdgpu_waves_per_eu(OSTFT_FUSED_WAVES)))
    stft_fused_kernel(const FusedArgs a)
{
    using C = Cfg<LOG_M, PLAN>;
    __shared__ float2 lds[C::G * C::PADM];
    __shared__ float2 red[C::WG / 64];
    FusedArgs ay = a;
    ay.in += (uint64_t)blockIdx.y * a.in_stride;
    ay.out += (uint64_t)blockIdx.y * a.out_stride;
    fused_frames<LOG_M, PLAN>(ay, (uint64_t)xcd_block(blockIdx.x, gridDim.x) * (uint32_t)a.tile_frames, lds, red,
                              threadIdx.x);
}

// Batch of requests of one window size, two forms over the same tile queue
// (tile t = G frames of request tile_job[t], whose first tile is
// jobs[j].tile_begin):
//   * dispatch: one workgroup per tile, the hardware dispatcher walks the
//     queue (XCD-aware order);
//   * persistent: a fixed grid whose workgroups claim tiles from the class's
//     counter until the tiles run out.
// On gfx9 one wave's vector loads and stores retire in order, so a persistent
// workgroup's wait for its next tile's loads also waits for the acknowledgement
// of its previous tile's stores; a dispatched workgroup simply ends and leaves
// its stores to drain.  Which form is faster for a window size is measured
// (ostft_autotune_batch).
template <int LOG_M, int V>
__global__ void __launch_bounds__((Cfg<LOG_M, V>::WG)) __attribute__((amdgpu_waves_per_eu(4)))
    stft_batch_dispatch_kernel(const char* __restrict__ blk, const uint32_t header_off)
{
    using C = Cfg<LOG_M, V>;
    __shared__ float2 lds[C::G * C::PADM];
    __shared__ float2 red[C::WG / 64];
    const BatchClassHeader* h = reinterpret_cast<const BatchClassHeader*>(blk + header_off);
    const uint32_t ntiles = h->ntiles;
    // the grid is ntiles (or 1 for an empty class): a permutation of the tiles
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    if (tile >= ntiles)
        return;
    const BatchJob* jobs = reinterpret_cast<const BatchJob*>(blk + h->jobs_off);
    const uint32_t j = reinterpret_cast<const uint32_t*>(blk + h->tiles_off)[tile];
    const FusedArgs a = from_global(*reinterpret_cast<const FusedArgsGlobal*>(&jobs[j].a));
    fused_frames<LOG_M, V>(a, (uint64_t)(tile - jobs[j].tile_begin) * (uint32_t)a.tile_frames, lds, red,
                           threadIdx.x);
}

template <int LOG_M, int PLAN>
__global__ void __launch_bounds__((Cfg<LOG_M, PLAN>::WG))
    __attribute__((amdgpu_waves_per_eu(BatchWaves<LOG_M>::value)))
    stft_batch_kernel(const char* __restrict__ blk, uint32_t* __restrict__ counter, const uint32_t header_off)
{
    // `blk` is read-only here (the counter is reached only through `counter`),
    // which lets the compiler read the descriptors with scalar loads
    using C = Cfg<LOG_M, PLAN>;
    __shared__ float2 lds[C::G * C::PADM];
    __shared__ float2 red[C::WG / 64];
    __shared__ uint32_t claim;
    const BatchClassHeader* h = reinterpret_cast<const BatchClassHeader*>(blk + header_off);
    const uint32_t ntiles = h->ntiles;
    const BatchJob* jobs = reinterpret_cast<const BatchJob*>(blk + h->jobs_off);
    const uint32_t* tile_job = reinterpret_cast<const uint32_t*>(blk + h->tiles_off);
    if (threadIdx.x == 0)
        claim = atomicAdd(counter, 1u);
    __syncthreads();
    for (;;) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(claim);
        if (tile >= ntiles)
            break;  // uniform: every thread read the same claim
        // claim the next tile now: the atomic's round trip overlaps this
        // tile's loads instead of stalling the workgroup between tiles
        uint32_t next = 0;
        if (threadIdx.x == 0)
            next = atomicAdd(counter, 1u);
        const uint32_t j = tile_job[tile];
        const FusedArgs a = from_global(*reinterpret_cast<const FusedArgsGlobal*>(&jobs[j].a));
        // the thread index, re-stated opaquely every tile: otherwise the
        // compiler hoists every thread-dependent address out of the loop and
        // keeps them all live in registers (about twice the one-request
        // kernel's budget)
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        fused_frames<LOG_M, PLAN>(a, (uint64_t)(tile - jobs[j].tile_begin) * (uint32_t)a.tile_frames, lds, red, tid);
        // every thread read `claim` before the body's first barrier, so it
        // can be overwritten now; the barrier publishes it and frees the LDS
        if (threadIdx.x == 0)
            claim = next;
        __syncthreads();
    }
}

// host-side compile-time loop: f(std::integral_constant<int, i>) for i in [B, E)
template <int B, int E, typename F>
static void host_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        host_for<B + 1, E>(f);
    }
}

// variant validity (see FUSED_VARIANTS in engine.hpp): the remainder-first
// plans only where the size has a remainder radix; the radix-8 plans for
// complex sizes 2^7..2^13 in the default workgroup shape; the other shapes
// only where they change the workgroup (512 threads while a transform has
// fewer, 128 while it has fewer than 256)
template <int LOG_M, int V>
constexpr bool variant_ok()
{
    if constexpr (V < 0 || V >= FUSED_VARIANTS) {
        return false;
    } else {
        using C = Cfg<LOG_M, V>;
        constexpr int PLAN = V & 3, SHAPE = V >> 2;
        return SHAPE <= 2 && !((PLAN & 2) && (LOG_M < 7 || LOG_M > 13 || SHAPE != 0)) &&
               !((PLAN & 1) && C::REM == 0) && !(SHAPE == 1 && C::T >= 512) && !(SHAPE == 2 && C::T >= 256);
    }
}

// f(integral_constant<int, V>) for the runtime variant v, if it is built
template <int LOG_M, int NV = FUSED_VARIANTS, typename F>
static hipError_t with_variant(int v, F&& f)
{
    hipError_t e = hipErrorInvalidValue;
    host_for<0, NV>([&](auto vc) {
        constexpr int V = decltype(vc)::value;
        if constexpr (variant_ok<LOG_M, V>()) {
            if (v == V)
                e = f(vc);
        }
    });
    return e;
}

// the one-request kernel runs the request's radix plan in the default shape
template <int LOG_M>
static hipError_t launch_t(const FusedArgs& a, uint32_t count, hipStream_t s)
{
    if (a.plan < 0 || a.plan > 3)
        return hipErrorInvalidValue;
    return with_variant<LOG_M, 4>(a.plan, [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        using C = Cfg<LOG_M, P>;
        if (a.tile_frames < 1 || a.tile_frames > C::G)
            return hipErrorInvalidValue;
        const uint64_t blocks = (a.frames + a.tile_frames - 1) / a.tile_frames;
        if (blocks == 0 || count == 0)
            return hipSuccess;
        if (blocks > 0x7fffffffull || count > 65535u)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((stft_fused_kernel<LOG_M, P>), dim3((uint32_t)blocks, count), dim3(C::WG), 0, s, a);
        return hipGetLastError();
    });
}

template <int LOG_M>
static hipError_t launch_xf(int xf, const FusedArgs& a, uint32_t count, hipStream_t s)
{
    // the matrix-core forms run the GEMM real DFT of mfma_real_dft.hip
    if (xf != XF_FP32 && mfma_gemm_supported(LOG_M))
        return launch_mfma_gemm(LOG_M, xf, a, count, s);
    if (xf != XF_FP32)
        return hipErrorInvalidValue;
    return launch_t<LOG_M>(a, count, s);
}

bool mfma_supported(int log_m) { return log_m == 5 || log_m == 7 || log_m == 9; }

template <int LOG_M>
static hipError_t launch_batch_t(int v, const BatchClassArgs& b, uint32_t grid, bool persistent, hipStream_t s)
{
    if (grid == 0)
        return hipErrorInvalidValue;
    char* blk = static_cast<char*>(b.base);
    return with_variant<LOG_M>(v, [&](auto vc) -> hipError_t {
        constexpr int V = decltype(vc)::value;
        using C = Cfg<LOG_M, V>;
        if (persistent) {
            if constexpr (C::SHAPE == 0) {
                uint32_t* counter =
                    reinterpret_cast<uint32_t*>(blk + b.header_off + offsetof(BatchClassHeader, counter));
                hipLaunchKernelGGL((stft_batch_kernel<LOG_M, C::PLAN>), dim3(grid), dim3(C::WG), 0, s,
                                   static_cast<const char*>(blk), counter, (uint32_t)b.header_off);
            } else {
                return hipErrorInvalidValue;  // the persistent form has workgroup shape 0 only
            }
        } else {
            hipLaunchKernelGGL((stft_batch_dispatch_kernel<LOG_M, V>), dim3(grid), dim3(C::WG), 0, s,
                               static_cast<const char*>(blk), (uint32_t)b.header_off);
        }
        return hipGetLastError();
    });
}

template <int LOG_M>
static hipError_t batch_info_t(int v, BatchClassInfo* info)
{
    return with_variant<LOG_M>(v, [&](auto vc) -> hipError_t {
        constexpr int V = decltype(vc)::value;
        using C = Cfg<LOG_M, V>;
        info->frames_per_tile = C::G;
        info->threads = C::WG;
        int per_cu = 0;
        hipError_t e;
        if constexpr (C::SHAPE == 0)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stft_batch_kernel<LOG_M, C::PLAN>, C::WG, 0);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stft_batch_dispatch_kernel<LOG_M, V>, C::WG,
                                                             0);
        info->max_blocks_per_cu = per_cu;
        return e;
    });
}

template <int LOG_M>
static const void* batch_fn_t(int v, bool persistent)
{
    const void* fn = nullptr;
    (void)with_variant<LOG_M>(v, [&](auto vc) {
        constexpr int V = decltype(vc)::value;
        using C = Cfg<LOG_M, V>;
        if constexpr (C::SHAPE == 0) {
            fn = persistent ? reinterpret_cast<const void*>(&stft_batch_kernel<LOG_M, C::PLAN>)
                            : reinterpret_cast<const void*>(&stft_batch_dispatch_kernel<LOG_M, V>);
        } else {
            fn = persistent ? nullptr : reinterpret_cast<const void*>(&stft_batch_dispatch_kernel<LOG_M, V>);
        }
        return hipSuccess;
    });
    return fn;
}

bool fused_supported(int log_m) { return log_m >= 5 && log_m <= 14; }

// f(integral_constant<int, LOG_M>) for a runtime log_m in 5 .. 14
template <typename R, typename F>
static R with_log_m(int log_m, R fail, F&& f)
{
    R r = fail;
    host_for<5, 15>([&](auto lc) {
        if (log_m == decltype(lc)::value)
            r = f(lc);
    });
    return r;
}

uint64_t fused_tile_frames(uint64_t G, uint64_t N, uint64_t hop)
{
    return hop > N ? std::min<uint64_t>(G, (G - 1) * N / hop + 1) : G;
}

bool fused_variant_ok(int log_m, int variant)
{
    return with_log_m(log_m, false, [&](auto lc) {
        bool ok = false;
        host_for<0, FUSED_VARIANTS>([&](auto vc) {
            if (variant == decltype(vc)::value)
                ok = variant_ok<decltype(lc)::value, decltype(vc)::value>();
        });
        return ok;
    });
}

bool fused_plan_ok(int log_m, int plan) { return plan >= 0 && plan < 4 && fused_variant_ok(log_m, plan); }

int fused_variant_frames(int log_m, int variant)
{
    return with_log_m(log_m, 0, [&](auto lc) {
        int g = 0;
        host_for<0, FUSED_VARIANTS>([&](auto vc) {
            constexpr int L = decltype(lc)::value, V = decltype(vc)::value;
            if constexpr (variant_ok<L, V>()) {
                if (variant == V)
                    g = Cfg<L, V>::G;
            }
        });
        return g;
    });
}

hipError_t launch_batch(int log_m, int variant, const BatchClassArgs& b, uint32_t grid, bool persistent,
                        hipStream_t s)
{
    return with_log_m(log_m, hipErrorInvalidValue, [&](auto lc) {
        return launch_batch_t<decltype(lc)::value>(variant, b, grid, persistent, s);
    });
}

const void* batch_kernel_fn(int log_m, int variant, bool persistent)
{
    return with_log_m(log_m, (const void*)nullptr,
                      [&](auto lc) { return batch_fn_t<decltype(lc)::value>(variant, persistent); });
}

hipError_t batch_class_info(int log_m, int variant, BatchClassInfo* info)
{
    return with_log_m(log_m, hipErrorInvalidValue,
                      [&](auto lc) { return batch_info_t<decltype(lc)::value>(variant, info); });
}

hipError_t launch_fused(int log_m, int xf, const FusedArgs& a, uint32_t count, hipStream_t s)
{
    return with_log_m(log_m, hipErrorInvalidValue,
                      [&](auto lc) { return launch_xf<decltype(lc)::value>(xf, a, count, s); });
}

}  // namespace ostft
