// mfma_real_dft.hip -- the split-precision matrix-core form of the STFT
// (--mfma_fp16 / --mfma_bf16) for window 64: each frame's real DFT as one
// dense GEMM on the matrix cores, with the window folded into the DFT matrix.
//
// What it replaces, in one launch: Overlap_Common, Window_*, DCRemove_Common
// (OKL/EXPCommon.okl:7-192), Stockhoptimized6 (OKL/EXP6.okl:7-46) and
// toPower / toHalfComplexFormat (OKL/kernel.okl:178-218).
//
// The GEMM (one wave, 16 frames): C = U B with
//     U[f][n] = x[(f0 + f) hop + n] s_f          (frames x 64 samples)
//     B[n][c] = 2^10 w[n] Re W^(c n)       c = 0..31   (Re X[c])
//             = 2^10 w[n] (-1)^n           c = 32      (Re X[32])
//             = 2^10 w[n] Im W^((c-32) n)  c = 33..63  (Im X[c-32])
// (W = exp(-2 pi i / 64)), on v_mfma_f32_16x16x32_{f16,bf16}: 2 K-steps x 4
// column tiles.  Split precision: every operand is hi + lo with hi =
// round16(v), lo = round16(v - hi), and each product is A_hi B_hi + A_hi B_lo
// + A_lo B_hi accumulated in FP32 (the A_lo B_lo term lies below the FP32
// rounding of the sum).  FP16 pairs carry 22 significant bits, BF16 pairs 16.
// s_f = 2^(13 - e) from the exponent e of the frame's largest |x| keeps U in
// FP16's normal range (exactly undone at the end, with the 2^10 of B).
//
// DC removal costs nothing here: subtracting the frame mean mu from the
// windowed frame changes only bin 0, X[0] = sum(w x) - N mu = 0.
//
// Lane layouts of v_mfma_f32_16x16x32_* (lane l, g = l / 16, checked on the
// GPU with an in-bounds probe): A[l % 16][8g + t], B[8g + t][l % 16]
// (t = 0..7), D[4g + r][l % 16] (r = 0..3).  So a lane loads 8 consecutive
// samples of one frame per K-step, and holds bins j and 16 + j (j = l % 16;
// and Re X[32] at j = 0) of 4 frames.
//
// B is the same for every frame of a request: its split fragments, in lane
// order, are built once per (window table, format) on the GPU
// (mfma_table_kernel, 16 KB) and each workgroup copies them to LDS; every
// wave then reads its fragments from there (ds_read_b128).  Workgroups of 4
// waves are persistent: each wave runs groups of two 16-frame tiles in a
// loop, the next group's samples in flight while the current one's bins are
// stored.
#include "engine.hpp"
#include "fft_device.hpp"
#include "stockham.hpp"

#include <atomic>

namespace ostft {

namespace {

typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 b16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef float f32x16v __attribute__((ext_vector_type(16)));

template <int XF>
struct Prec;

template <>
struct Prec<XF_MFMA_FP16> {
    using frag = h16x8;
    using elem = _Float16;
    static __device__ __forceinline__ f32x4v mma(frag a, frag b, f32x4v c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ f32x16v mma32(frag a, frag b, f32x16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Prec<XF_MFMA_BF16> {
    using frag = b16x8;
    using elem = __bf16;
    static __device__ __forceinline__ f32x4v mma(frag a, frag b, f32x4v c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ f32x16v mma32(frag a, frag b, f32x16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// v = hi + lo, hi = round(v), lo = round(v - hi) in the 16-bit format
template <int XF>
__device__ __forceinline__ void split8(const float (&v)[8], typename Prec<XF>::frag& hi,
                                       typename Prec<XF>::frag& lo)
{
    using E = typename Prec<XF>::elem;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const E e = (E)v[t];
        hi[t] = e;
        lo[t] = (E)(v[t] - (float)e);
    }
}

constexpr int N64 = 64;
constexpr int TILE = 16;            // frames per wave tile (GEMM rows)
constexpr int WAVES = 4;            // waves per workgroup
constexpr int KS = 2, CT = 4;       // K-steps of 32, column tiles of 16
constexpr int NFRAG = KS * CT * 2;  // B fragments per lane (hi and lo)
constexpr float B_SCALE = 1024.0f;  // 2^10: keeps the window tapers of B in FP16's normal range
constexpr size_t TABLE_BYTES = (size_t)NFRAG * 64 * 16;

// fragment (ks, ct, hl) of lane l at index (((ks * CT + ct) * 2 + hl) * 64 + l)
__host__ __device__ constexpr int frag_index(int ks, int ct, int hl, int lane)
{
    return ((ks * CT + ct) * 2 + hl) * 64 + lane;
}

// One wave: every lane builds its 16 fragments of B from the window and the
// FP64-generated twiddle table (W^e = tw_n[e] for e < 32, -tw_n[e - 32] above).
template <int XF>
__global__ void __launch_bounds__(64) mfma_table_kernel(const float* __restrict__ win,
                                                        const float2* __restrict__ twn,
                                                        typename Prec<XF>::frag* __restrict__ table)
{
    const uint32_t lane = threadIdx.x, j = lane & 15u, g = lane >> 4;
    for (int ks = 0; ks < KS; ++ks) {
        for (int ct = 0; ct < CT; ++ct) {
            const uint32_t c = 16u * ct + j;  // GEMM column
            float v[8];
            for (int t = 0; t < 8; ++t) {
                const uint32_t n = 32u * ks + 8u * g + t;
                const float w = (win ? win[n] : 1.0f) * B_SCALE;
                if (c == 32u) {
                    v[t] = (n & 1u) ? -w : w;
                } else {
                    const uint32_t bin = c < 32u ? c : c - 32u;
                    const uint32_t e = (bin * n) & 63u;
                    const float2 tw = twn[e & 31u];
                    const float sg = e >= 32u ? -1.0f : 1.0f;
                    v[t] = w * (sg * (c < 32u ? tw.x : tw.y));
                }
            }
            typename Prec<XF>::frag hi, lo;
            split8<XF>(v, hi, lo);
            table[frag_index(ks, ct, 0, lane)] = hi;
            table[frag_index(ks, ct, 1, lane)] = lo;
        }
    }
}

// 2^(13 - e) for the exponent e of m > 0 (finite), clamped; its inverse
// times 2^-10 (B's scale) undoes the whole scaling exactly.
__device__ __forceinline__ void frame_scale(float m, float& s, float& inv)
{
    const int e = (int)((__float_as_uint(m) >> 23) & 0xffu) - 127;
    int k = 13 - e;
    k = k < -100 ? -100 : (k > 100 ? 100 : k);
    s = __uint_as_float((uint32_t)(127 + k) << 23);
    inv = __uint_as_float((uint32_t)(127 - k - 10) << 23);
}

// x[s .. s+8) of signal `in` (zero past L), 16-byte loads where possible
__device__ __forceinline__ void load8(const float* __restrict__ in, uint64_t L, uint64_t s, bool row_ok,
                                      float (&v)[8])
{
    if (row_ok && s + 8 <= L && ((reinterpret_cast<uintptr_t>(in + s) & 15u) == 0)) {
        const float4 p = *reinterpret_cast<const float4*>(in + s);
        const float4 q = *reinterpret_cast<const float4*>(in + s + 4);
        v[0] = p.x, v[1] = p.y, v[2] = p.z, v[3] = p.w, v[4] = q.x, v[5] = q.y, v[6] = q.z, v[7] = q.w;
    } else {
#pragma unroll
        for (int t = 0; t < 8; ++t)
            v[t] = (row_ok && s + t < L) ? in[s + t] : 0.0f;
    }
}

// Output stores of the matrix-core kernels: plain stores by default (they
// measure faster here than streaming ones), streaming with OSTFT_MFMA_NT=1.
#ifndef OSTFT_MFMA_NT
#define OSTFT_MFMA_NT 0
#endif
template <typename T>
__device__ __forceinline__ void out_store(T v, T* p)
{
#if OSTFT_MFMA_NT
    nt_store(v, p);
#else
    *p = v;
#endif
}
__device__ __forceinline__ void out_store4(float4 v, float4* p)
{
#if OSTFT_MFMA_NT
    nt_store4(v, p);
#else
    *p = v;
#endif
}

// Bin k (0 <= k <= N/2) of frame f in the output layout MODE, and for the
// N-value layouts its mirror N - k (the spectrum of a real frame is
// Hermitian: X[N - k] = conj X[k]).
template <int MODE, int N>
__device__ __forceinline__ void put_bin(float* __restrict__ out, uint64_t limit, uint64_t f, uint32_t k, float re,
                                        float im)
{
    if constexpr (MODE == OUT_POWER) {
        out_store(fmaf(re, re, im * im), out + f * (uint64_t)(N / 2 + 1) + k);
    } else if constexpr (MODE == OUT_HALF_SPECTRUM) {
        out_store(make_float2(re, im), reinterpret_cast<float2*>(out) + f * (uint64_t)(N / 2 + 1) + k);
    } else if constexpr (MODE == OUT_COMPAT_POWER || MODE == OUT_MAGNITUDE) {
        // sqrt(2)|Re X[k]| (the Serial runner's ToPower) or |X[k]|
        const float v = MODE == OUT_COMPAT_POWER ? 1.41421356237309504880f * fabsf(re)
                                                 : __builtin_amdgcn_sqrtf(fmaf(re, re, im * im));
        float* o = out + f * (uint64_t)N;
        out_store(v, o + k);
        if (k != 0 && k != N / 2)
            out_store(v, o + (N - k));
    } else {
        // the first OHalfSize complex values of the frame-major spectrum
        float2* o = reinterpret_cast<float2*>(out);
        const uint64_t c = f * (uint64_t)N;
        if (c + k < limit)
            out_store(make_float2(re, im), o + c + k);
        if (k != 0 && k != N / 2 && c + N - k < limit)
            out_store(make_float2(re, -im), o + c + (N - k));
    }
}

// floats per frame of each output layout
template <int MODE, int N>
__host__ __device__ constexpr int layout_floats()
{
    return MODE == OUT_POWER ? N / 2 + 1 : MODE == OUT_HALF_SPECTRUM ? N + 2 : MODE == OUT_HALF_COMPLEX ? 2 * N : N;
}

// Bin k (0 <= k <= N/2) into a frame's output block `fr` of layout MODE,
// with its mirror N - k for the N-value layouts (X[N - k] = conj X[k]).
template <int MODE, int N>
__device__ __forceinline__ void stage_bin(float* fr, uint32_t k, float re, float im)
{
    if constexpr (MODE == OUT_POWER) {
        fr[k] = fmaf(re, re, im * im);
    } else if constexpr (MODE == OUT_HALF_SPECTRUM) {
        *reinterpret_cast<float2*>(fr + 2 * k) = make_float2(re, im);
    } else if constexpr (MODE == OUT_COMPAT_POWER || MODE == OUT_MAGNITUDE) {
        // sqrt(2)|Re X[k]| (the Serial runner's ToPower) or |X[k]|
        const float v = MODE == OUT_COMPAT_POWER ? 1.41421356237309504880f * fabsf(re)
                                                 : __builtin_amdgcn_sqrtf(fmaf(re, re, im * im));
        fr[k] = v;
        if (k != 0 && k != N / 2)
            fr[N - k] = v;
    } else {
        *reinterpret_cast<float2*>(fr + 2 * k) = make_float2(re, im);
        if (k != 0 && k != N / 2)
            *reinterpret_cast<float2*>(fr + 2 * (N - k)) = make_float2(re, -im);
    }
}

// n <= NMAX floats from LDS to global memory (dst 16-byte aligned), every
// LDS read of a lane issued before its first store
template <int NMAX>
__device__ __forceinline__ void write_block_n(float* __restrict__ dst, const float* src, uint32_t n, uint32_t lane)
{
    constexpr int IT = (NMAX / 4 + 63) / 64;
    const uint32_t n4 = n / 4;
    float4 v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i)
        if (lane + 64u * i < n4)
            v[i] = reinterpret_cast<const float4*>(src)[lane + 64u * i];
#pragma unroll
    for (int i = 0; i < IT; ++i)
        if (lane + 64u * i < n4)
            out_store4(v[i], reinterpret_cast<float4*>(dst) + lane + 64u * i);
    if (4 * n4 + lane < n)
        out_store(src[4 * n4 + lane], dst + 4 * n4 + lane);
}

constexpr int STAGE_FLOATS = TILE * (N64 + 2);  // one tile's largest single-round output block

// tiles per wave and loop iteration of the window-64 kernel: both tiles'
// samples (8 KB a wave) are in flight together
#ifndef OSTFT_MFMA_TPW
#define OSTFT_MFMA_TPW 2
#endif
constexpr int TPW = OSTFT_MFMA_TPW;
// minimum waves per SIMD it is compiled for (its LDS allows 4 workgroups per
// CU; 4 waves would leave too few registers for two tiles)
#ifndef OSTFT_MFMA_MINWAVES
#define OSTFT_MFMA_MINWAVES 3
#endif

// this lane's samples of the tiles of group `grp` (tile grp TPW + u: frame j,
// n = 32 ks + 8 g + t)
__device__ __forceinline__ void load64_group(const float* __restrict__ in, uint64_t L, uint64_t hop,
                                             uint64_t frames, uint64_t tiles, uint64_t grp, uint32_t j, uint32_t g,
                                             float (&x)[TPW][KS][8])
{
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const uint64_t tile = grp * TPW + u;
        const uint64_t f = tile * TILE + j;
        const bool row_ok = f < frames && tile < tiles;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            load8(in, L, f * hop + 32u * ks + 8u * g, row_ok, x[u][ks]);
    }
}

// Persistent: wave w of workgroup b takes tile groups b WAVES + w, then
// every gridDim.x WAVES-th one, so B is copied to LDS once per workgroup; the
// next group's samples load as soon as the current ones are in the matrix cores.
template <int XF>
__global__ void __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(OSTFT_MFMA_MINWAVES)))
stft_mfma64_kernel(const FusedArgs a, const uint32_t tiles_per_signal)
{
    using P = Prec<XF>;
    using frag = typename P::frag;
    __shared__ frag bfr[NFRAG * 64];                 // B fragments, lane order (16 KB)
    __shared__ float inv_row[WAVES][TPW][TILE];      // the frames' inverse scales
    __shared__ __align__(16) float out_lds[WAVES][STAGE_FLOATS];  // each wave's output block

    const uint32_t tid = threadIdx.x, lane0 = tid & 63u, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t y = blockIdx.y;  // signal of a strided batch
    const float* __restrict__ in = a.in + (uint64_t)y * a.in_stride;
    float* __restrict__ out = a.out + (uint64_t)y * a.out_stride;
    const uint64_t hop = a.hop, L = a.L, frames = a.frames;
    const uint64_t groups = ((uint64_t)tiles_per_signal + TPW - 1) / TPW;

    // ---- the first group's samples and the B table ----
    const uint64_t stride = (uint64_t)gridDim.x * WAVES;
    uint64_t grp = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * WAVES + wave;
    float x[TPW][KS][8];
    load64_group(in, L, hop, frames, tiles_per_signal, grp, lane0 & 15u, lane0 >> 4, x);
    {
        const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.mfma_b);
        for (uint32_t e = tid; e < NFRAG * 64; e += WAVES * 64)
            reinterpret_cast<uint4*>(bfr)[e] = src[e];
    }
    __syncthreads();

    const int mode = a.out_mode;
    const bool dc = a.dc != 0;
    for (; grp < groups; grp += stride) {  // uniform over the wave
        // B is read from LDS, and the lane's output offsets formed, where they
        // are used: kept in registers over the loop they would not fit
        asm volatile("" ::: "memory");
        uint32_t lane = lane0;
        asm volatile("" : "+v"(lane));
        const uint32_t j = lane & 15u, g = lane >> 4;
        f32x4v acc[TPW][CT];
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            // per-frame scale from the largest |x| of the frame (lanes j + 16 g')
            float mx = 0.0f;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    mx = fmaxf(mx, fabsf(x[u][ks][t]));
            mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
            mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
            float s = 1.0f, inv = 1.0f / B_SCALE;
            if (mx > 0.0f)
                frame_scale(mx, s, inv);
            if (g == 0)
                inv_row[wave][u][j] = inv;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc[u][ct] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                float v[8];
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    v[t] = x[u][ks][t] * s;
                frag ah, al;
                split8<XF>(v, ah, al);
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    const frag bh = bfr[frag_index(ks, ct, 0, lane)];
                    const frag bl = bfr[frag_index(ks, ct, 1, lane)];
                    acc[u][ct] = P::mma(ah, bh, acc[u][ct]);
                    acc[u][ct] = P::mma(ah, bl, acc[u][ct]);
                    acc[u][ct] = P::mma(al, bh, acc[u][ct]);
                }
            }
        }
        load64_group(in, L, hop, frames, tiles_per_signal, grp + stride, j, g, x);
        // the rows' inverse scales (written by lanes 0..15 of this wave)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        // lane (j, g) holds, for frames 4g + r: Re X[j] (tile 0), Re X[16+j]
        // (tile 1), Im X[j] or Re X[32] at j = 0 (tile 2), Im X[16+j] (tile 3).
        // A tile's output block (frames f0 .. f0+15, contiguous in every
        // layout) is assembled in this wave's LDS and written with 16-byte
        // stores over consecutive lanes.
        float* stage = out_lds[wave];
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const uint64_t tile = grp * TPW + u;
            if (tile >= tiles_per_signal)
                break;  // uniform over the wave
            const uint64_t f0 = tile * TILE;
            const uint32_t nf = (uint32_t)(frames - f0 < (uint64_t)TILE ? frames - f0 : (uint64_t)TILE);
            const float4 iv = *reinterpret_cast<const float4*>(&inv_row[wave][u][4 * g]);
            auto emit = [&](auto mode_c) {
                constexpr int MODE = decltype(mode_c)::value;
                constexpr int P = layout_floats<MODE, N64>();  // floats per frame
                constexpr int FR = STAGE_FLOATS / P < TILE ? STAGE_FLOATS / P : TILE;  // frames per round
#pragma unroll
                for (int r0 = 0; r0 < TILE; r0 += FR) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t fl = 4u * g + r;
                        if (fl < (uint32_t)r0 || fl >= (uint32_t)(r0 + FR))
                            continue;
                        const float sc = r == 0 ? iv.x : r == 1 ? iv.y : r == 2 ? iv.z : iv.w;
                        float re0 = acc[u][0][r] * sc, im0 = acc[u][2][r] * sc;
                        const float re1 = acc[u][1][r] * sc, im1 = acc[u][3][r] * sc;
                        float* fr = stage + (fl - r0) * P;
                        if (j == 0) {
                            stage_bin<MODE, N64>(fr, N64 / 2, im0, 0.0f);  // Re X[32]
                            im0 = 0.0f;
                            if (dc)
                                re0 = 0.0f;
                        }
                        stage_bin<MODE, N64>(fr, j, re0, im0);
                        stage_bin<MODE, N64>(fr, 16u + j, re1, im1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t fr0 = (uint32_t)r0 < nf ? (uint32_t)r0 : nf;
                    const uint32_t fr1 = (uint32_t)(r0 + FR) < nf ? (uint32_t)(r0 + FR) : nf;
                    uint64_t begin = (f0 + fr0) * (uint64_t)P, end = (f0 + fr1) * (uint64_t)P;
                    if constexpr (MODE == OUT_HALF_COMPLEX) {
                        // the first OHalfSize complex values of the frame-major spectrum
                        const uint64_t lim = 2 * a.out_limit;
                        end = end < lim ? end : lim;
                        begin = begin < end ? begin : end;
                    }
                    write_block_n<FR * P>(out + begin, stage, (uint32_t)(end - begin), lane);
                    __builtin_amdgcn_wave_barrier();  // the next round rewrites the stage
                }
            };
            switch (mode) {
            case OUT_POWER: emit(std::integral_constant<int, OUT_POWER>{}); break;
            case OUT_HALF_SPECTRUM: emit(std::integral_constant<int, OUT_HALF_SPECTRUM>{}); break;
            case OUT_COMPAT_POWER: emit(std::integral_constant<int, OUT_COMPAT_POWER>{}); break;
            case OUT_MAGNITUDE: emit(std::integral_constant<int, OUT_MAGNITUDE>{}); break;
            case OUT_HALF_COMPLEX: emit(std::integral_constant<int, OUT_HALF_COMPLEX>{}); break;
            default: break;
            }
        }
    }
}

// compute units of the current device (cached per device)
uint32_t device_cus()
{
    static std::atomic<int> cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return 256;
    int c = cached[dev].load(std::memory_order_relaxed);
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
            c = 256;
        cached[dev].store(c, std::memory_order_relaxed);
    }
    return (uint32_t)c;
}

// waves per compute unit a persistent launch asks for, over all signals of
// the batch (about twice what one CU holds at once)
#ifndef OSTFT_MFMA_GRID_WAVES_PER_CU
#define OSTFT_MFMA_GRID_WAVES_PER_CU 32
#endif

template <int XF>
hipError_t launch64(const FusedArgs& a, uint32_t count, hipStream_t s)
{
    const uint64_t tps = (a.frames + TILE - 1) / TILE;
    if (tps == 0 || count == 0)
        return hipSuccess;
    if (!a.mfma_b)
        return hipErrorInvalidValue;
    if (tps > 0xffffffffull || count > 65535u)
        return hipErrorInvalidValue;
    const uint64_t need = (tps + WAVES * TPW - 1) / (WAVES * TPW);  // one tile group per wave
    const uint64_t total = (uint64_t)device_cus() * OSTFT_MFMA_GRID_WAVES_PER_CU / WAVES;
    const uint64_t cap = (total + count - 1) / count;
    const uint64_t blocks = need < cap ? need : cap;
    hipLaunchKernelGGL(stft_mfma64_kernel<XF>, dim3((uint32_t)blocks, count), dim3(WAVES * 64), 0, s, a,
                       (uint32_t)tps);
    return hipGetLastError();
}

// ---- windows 256 and 1024: two stages -------------------------------------
//
// With R = N / 64, n = n1 + R n2 (n1 < R, n2 < 64) and k = k2 + 64 k1:
//     Y[n1][k2]      = sum_n2 u[n1 + R n2] W_64^(n2 k2)                  (matrix cores)
//     X[k2 + 64 k1]  = sum_n1 (Y[n1][k2] W_N^(n1 k2)) W_R^(n1 k1)         (registers, FP32)
// The first stage is the window-64 GEMM above with GEMM rows (frame, n1),
// K = n2 and the window applied to U (it depends on n1 and n2), on
// v_mfma_f32_32x32x16_*: lane (j, h) of the accumulators holds column j
// (Re Y[n1][j] in tile 0; Im Y[n1][j], or Re Y[n1][32] at j = 0, in tile 1)
// of rows 8q + 4h + r (q, r = 0..3).  With rows (frame f, n1) = R f + n1:
//   R = 4:  row 8q + 4h + r is frame 2q + h, n1 = r: each lane holds every
//           n1 of 4 frames;
//   R = 16: row 8q + 4h + r is frame q/2, n1 = 8(q%2) + 4h + r: lanes j and
//           j + 32 swap halves, after which lane (j, h) holds every n1 of frame h.
// Y is Hermitian in k2 (u is real), so lane j >= 1 with T[n1] = Y[n1][j]
// W_N^(n1 j) and Z = DFT_R(T) gives R bins:
//     X[j + 64 k1] = Z[k1],   X[64 - j + 64 k1] = conj Z[R - 1 - k1]   (k1 < R/2)
// (T'[n1] = conj(Y[n1][j]) W_N^(n1 (64 - j)) = conj(T[n1]) W_R^n1 for
// k2 = 64 - j).  Lane 0 holds the two real columns Y0 = Y[.][0] and
// Y32 = Y[.][32]; the bins they give,
//     X[32 m] = sum_n1 W_2R^(n1 m) Y_(m mod 2)[n1]          (m = 0..R),
// are the real DFT V of length 2R of v[n1] = Y0[n1] + Y32[n1],
// v[R + n1] = Y0[n1] - Y32[n1], over two: (v[n1] + (-1)^m v[R + n1]) / 2 =
// Y_(m mod 2)[n1].  Lane 0 runs the same R-point complex DFT as the other
// lanes, on z[p] = v[2p] + i v[2p + 1], and untangles the even and odd halves:
//     P = Z[m] + conj Z[R - m],  Q = -i (Z[m] - conj Z[R - m]),
//     4 V[m] = P + W_2R^m Q,     4 V[R - m] = conj(P - W_2R^m Q)    (m <= R/2)
// (the 4: the halving of v and that of the even/odd split), so every lane
// runs one DFT and the N/2 + 1 bins come out in R + 1 rounds.

constexpr int RKS = 4, RCT = 2;        // K-steps of 16, column tiles of 32 (32x32x16)
constexpr int RNFRAG = RKS * RCT * 2;  // B fragments per lane
constexpr size_t RB_BYTES = (size_t)RNFRAG * 64 * 16;  // B fragments
constexpr size_t RW_BYTES = (size_t)RKS * 64 * 8 * 4;  // window fragments
// twiddles W_N^(n1 kk) (n1 < R; kk = j, and 32 at j = 0, where they are
// W_2R^n1), n1-major
__host__ __device__ constexpr size_t rt_bytes(int r) { return (size_t)r * 32 * 8; }

#ifndef OSTFT_MFMA_R4_WAVES
#define OSTFT_MFMA_R4_WAVES 8
#endif
#ifndef OSTFT_MFMA_R4_STAGE
#define OSTFT_MFMA_R4_STAGE 1032
#endif
#ifndef OSTFT_MFMA_R4_DIRECT
#define OSTFT_MFMA_R4_DIRECT 0
#endif
#ifndef OSTFT_MFMA_R16_DIRECT
#define OSTFT_MFMA_R16_DIRECT 0
#endif
#ifndef OSTFT_MFMA_R16_WAVES
#define OSTFT_MFMA_R16_WAVES 4
#endif
// minimum waves per SIMD the two-stage kernels are compiled for (their LDS
// allows two workgroups per CU)
#ifndef OSTFT_MFMA_RMINWAVES
#define OSTFT_MFMA_RMINWAVES 4
#endif

// launch shape of the two-stage kernels: waves per workgroup (they share the
// table's LDS copy), tiles per wave, and the output stage (floats per wave)
template <int LOG_R>
struct RCfg;
// window 256 (8 frames a tile): output through the stage, 4 frames of the
// half spectrum (258 floats each) a round
template <>
struct RCfg<2> {
    static constexpr int WAVES = OSTFT_MFMA_R4_WAVES, STAGE = OSTFT_MFMA_R4_STAGE;
    static constexpr bool DIRECT = OSTFT_MFMA_R4_DIRECT;
};
// window 1024 (2 frames a tile): a one-frame stage, or direct stores (each
// store instruction writes 32 consecutive bins of both frames)
template <>
struct RCfg<4> {
    static constexpr int WAVES = OSTFT_MFMA_R16_WAVES;
    static constexpr bool DIRECT = OSTFT_MFMA_R16_DIRECT;
    static constexpr int STAGE = DIRECT ? 4 : 1032;
};

// frames per output round: the largest divisor of the tile's frames whose
// blocks fit the stage (0: the layout does not fit)
__host__ __device__ constexpr int stage_rounds_frames(int fpt, int pf, int stage)
{
    int fr = fpt;
    while (fr > 0 && (fr * pf > stage || fpt % fr != 0))
        --fr;
    return fr;
}

__host__ __device__ constexpr int rfrag_index(int ks, int ct, int hl, int lane)
{
    return ((ks * RCT + ct) * 2 + hl) * 64 + lane;
}

// One wave builds the table of window N = 64 R: every lane's 16 B fragments
// (W_64^e = W_N^(eR) = tw_n[eR] for e < 32, -tw_n[(e - 32) R] above), its
// 32 window values w[n1 + R (16 ks + 8 h + t)] (n1 = (lane % 32) % R) and
// the twiddle rows.
template <int LOG_R, int XF>
__global__ void __launch_bounds__(64) mfma_rtable_kernel(const float* __restrict__ win,
                                                         const float2* __restrict__ twn, char* __restrict__ table)
{
    constexpr int R = 1 << LOG_R;
    using frag = typename Prec<XF>::frag;
    frag* bt = reinterpret_cast<frag*>(table);
    float* wt = reinterpret_cast<float*>(table + RB_BYTES);
    float2* tt = reinterpret_cast<float2*>(table + RB_BYTES + RW_BYTES);
    const uint32_t lane = threadIdx.x, j = lane & 31u, h = lane >> 5;
    for (int ks = 0; ks < RKS; ++ks) {
        for (int ct = 0; ct < RCT; ++ct) {
            float v[8];
            for (int t = 0; t < 8; ++t) {
                const uint32_t n2 = 16u * ks + 8u * h + t;
                if (ct == 1 && j == 0) {
                    v[t] = (n2 & 1u) ? -1.0f : 1.0f;
                } else {
                    const uint32_t e = (j * n2) & 63u;
                    const float2 tw = twn[(e & 31u) * R];
                    const float sg = e >= 32u ? -1.0f : 1.0f;
                    v[t] = sg * (ct == 0 ? tw.x : tw.y);
                }
            }
            frag hi, lo;
            split8<XF>(v, hi, lo);
            bt[rfrag_index(ks, ct, 0, lane)] = hi;
            bt[rfrag_index(ks, ct, 1, lane)] = lo;
        }
        for (int t = 0; t < 8; ++t) {
            const uint32_t n = (j % R) + R * (16u * ks + 8u * h + t);
            wt[(ks * 64 + lane) * 8 + t] = win ? win[n] : 1.0f;
        }
    }
    // n1 kk <= 32 (R - 1) < N/2: inside the table
    if (h == 0)
        for (int n1 = 0; n1 < R; ++n1)
            tt[n1 * 32 + j] = twn[n1 * (j == 0 ? 32u : j)];
}

// Lane q of each quad holds row q of a 4 x 4 matrix in a[0..3]; afterwards
// it holds column q.  The off-diagonal 2 x 2 blocks are exchanged with the
// lane two apart, then the off-diagonal elements of each block with the lane
// one apart (quad permutations as DPP moves; each lane sends the values its
// partner keeps).
__device__ __forceinline__ float quad_xor1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // [1,0,3,2]
}
__device__ __forceinline__ float quad_xor2(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // [2,3,0,1]
}
__device__ __forceinline__ void quad_transpose(float (&a)[4], uint32_t q)
{
    const bool hi = (q & 2u) != 0, odd = (q & 1u) != 0;
    const float s0 = quad_xor2(hi ? a[0] : a[2]), s1 = quad_xor2(hi ? a[1] : a[3]);
    if (hi) {
        a[0] = s0;
        a[1] = s1;
    } else {
        a[2] = s0;
        a[3] = s1;
    }
    const float r0 = quad_xor1(odd ? a[0] : a[1]), r1 = quad_xor1(odd ? a[2] : a[3]);
    if (odd) {
        a[0] = r0;
        a[2] = r1;
    } else {
        a[1] = r0;
        a[3] = r1;
    }
}

// this lane's samples of tile `tile`: x[ks][t] = x[f hop + n1 + R (16 ks + 8 h + t)]
// (row (f, n1) of the GEMM), zeros past the signal, its frames or its tiles
// VEC (window 256 only; the caller checks that the input, the hop and the
// signal stride are multiples of 4 floats): a tile inside the signal loads
// 16-byte blocks, left for quad_transpose_tile where they are used (the
// transposes' temporaries would not fit beside the tile in flight); the
// return value says so
template <int R, bool VEC = false>
__device__ __forceinline__ bool load_rtile(const float* __restrict__ in, uint64_t L, uint64_t hop, uint64_t frames,
                                           uint64_t tiles, uint64_t tile, uint32_t n1r, uint32_t fr, uint32_t h,
                                           float (&x)[RKS][8])
{
    constexpr int N = 64 * R, FPT = 32 / R;
    const uint64_t f = tile * FPT + fr;
    // a tile wholly inside the signal (uniform over the wave) loads unguarded
    if (tile < tiles && (tile + 1) * FPT <= frames && ((tile + 1) * FPT - 1) * hop + N <= L) {
        if constexpr (VEC && R == 4) {
            // the quad's lanes (n1r = 0..3, one frame and h) read the 32
            // consecutive samples of step ks together: lane q loads those of
            // t = q and t = q + 4 as two 16-byte blocks (values n1r = 0..3)
            const float* __restrict__ qsrc = in + f * hop + 32u * h;
#pragma unroll
            for (int ks = 0; ks < RKS; ++ks) {
                const float4 u0 = *reinterpret_cast<const float4*>(qsrc + 4u * (16u * ks + n1r));
                const float4 u1 = *reinterpret_cast<const float4*>(qsrc + 4u * (16u * ks + n1r + 4u));
                x[ks][0] = u0.x;
                x[ks][1] = u0.y;
                x[ks][2] = u0.z;
                x[ks][3] = u0.w;
                x[ks][4] = u1.x;
                x[ks][5] = u1.y;
                x[ks][6] = u1.z;
                x[ks][7] = u1.w;
            }
            return true;
        }
        const float* __restrict__ src = in + f * hop + n1r + (uint64_t)R * 8u * h;
#pragma unroll
        for (int ks = 0; ks < RKS; ++ks)
#pragma unroll
            for (int t = 0; t < 8; ++t)
                x[ks][t] = src[R * (16 * ks + t)];
    } else {
        const bool row_ok = f < frames && tile < tiles;
        const uint64_t base = f * hop + n1r;
#pragma unroll
        for (int ks = 0; ks < RKS; ++ks)
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const uint64_t s = base + (uint64_t)R * (16u * ks + 8u * h + t);
                x[ks][t] = (row_ok && s < L) ? in[s] : 0.0f;
            }
    }
    return false;
}

// Persistent: wave w of workgroup b takes tiles b W + w, then every
// gridDim.x W-th one, so the table is copied to LDS once per workgroup; the
// next tile's samples load while the current one's bins are stored.
template <int LOG_R, int XF, bool VEC>
__global__ void __launch_bounds__(RCfg<LOG_R>::WAVES * 64) __attribute__((amdgpu_waves_per_eu(OSTFT_MFMA_RMINWAVES)))
stft_mfma_r_kernel(const FusedArgs a, const uint32_t tiles_per_signal)
{
    using C = RCfg<LOG_R>;
    constexpr int W = C::WAVES, STAGE = C::STAGE;
    constexpr int R = 1 << LOG_R, N = 64 * R, FPT = 32 / R;  // frames per tile
    using P = Prec<XF>;
    using frag = typename P::frag;
    __shared__ frag bfr[RNFRAG * 64];                    // 16 KB
    __shared__ __align__(16) float wfr[RKS * 64 * 8];    // 8 KB
    __shared__ __align__(16) float2 twl[R * 32];
    __shared__ float inv_row[W][FPT];
    __shared__ __align__(16) float stage_all[W][STAGE];

    const uint32_t tid = threadIdx.x, lane0 = tid & 63u, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t j0 = lane0 & 31u, h0 = lane0 >> 5;
    const uint32_t n1r0 = j0 % R, fr0 = j0 / R;  // GEMM row j0 of this lane's A operand: frame fr0, n1r0
    const uint32_t y = blockIdx.y;
    const float* __restrict__ in = a.in + (uint64_t)y * a.in_stride;
    float* __restrict__ out = a.out + (uint64_t)y * a.out_stride;
    const uint64_t hop = a.hop, L = a.L, frames = a.frames;

    // ---- the first tile's samples and the table ----
    const uint64_t stride = (uint64_t)gridDim.x * W;
    uint64_t tile = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * W + wave;
    float x[RKS][8];
    bool raw = load_rtile<R, VEC>(in, L, hop, frames, tiles_per_signal, tile, n1r0, fr0, h0, x);
    {
        constexpr uint32_t NB = RB_BYTES / 16, NW = RW_BYTES / 16, NU = NB + NW + rt_bytes(R) / 16;
        const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.mfma_b);
        for (uint32_t e = tid; e < NU; e += W * 64) {
            const uint4 v = src[e];
            if (e < NB)
                reinterpret_cast<uint4*>(bfr)[e] = v;
            else if (e < NB + NW)
                reinterpret_cast<uint4*>(wfr)[e - NB] = v;
            else
                reinterpret_cast<uint4*>(twl)[e - NB - NW] = v;
        }
    }
    __syncthreads();

    const int mode = a.out_mode;
    const bool dc = a.dc != 0;
    float* stage = stage_all[wave];
    for (; tile < tiles_per_signal; tile += stride) {  // uniform over the wave
        // the table's fragments are read from LDS, and the lane's output
        // offsets formed, where they are used: kept in registers over the
        // loop they would not fit
        asm volatile("" ::: "memory");
        uint32_t lane = lane0, j = j0, h = h0, n1r = n1r0, fr = fr0;
        asm volatile("" : "+v"(lane), "+v"(j), "+v"(h), "+v"(n1r), "+v"(fr));
        const uint64_t f0 = tile * FPT;
        if constexpr (VEC) {
            if (raw) {  // 16-byte blocks of the quad: one transpose per block
#pragma unroll
                for (int ks = 0; ks < RKS; ++ks)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        float q4[4] = {x[ks][4 * b], x[ks][4 * b + 1], x[ks][4 * b + 2], x[ks][4 * b + 3]};
                        quad_transpose(q4, n1r);
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            x[ks][4 * b + t] = q4[t];
                    }
            }
        }
        // windowed samples and the frame's largest |value| (lanes R fr .. R fr + R - 1, both h)
        float mx = 0.0f;
#pragma unroll
        for (int ks = 0; ks < RKS; ++ks) {
            const float4 w0 = *reinterpret_cast<const float4*>(&wfr[(ks * 64 + lane) * 8]);
            const float4 w1 = *reinterpret_cast<const float4*>(&wfr[(ks * 64 + lane) * 8 + 4]);
            const float wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                x[ks][t] *= wv[t];
                mx = fmaxf(mx, fabsf(x[ks][t]));
            }
        }
#pragma unroll
        for (int o = 1; o < R; o <<= 1)
            mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        float s = 1.0f, inv = 1.0f;
        if (mx > 0.0f) {
            frame_scale(mx, s, inv);
            inv *= B_SCALE;  // B carries no 2^10 here
        }
        if (h == 0 && n1r == 0)
            inv_row[wave][fr] = inv;
        f32x16v acc0 = {}, acc1 = {};
#pragma unroll
        for (int ks = 0; ks < RKS; ++ks) {
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                v[t] = x[ks][t] * s;
            frag ah, al;
            split8<XF>(v, ah, al);
            const frag b0h = bfr[rfrag_index(ks, 0, 0, lane)], b0l = bfr[rfrag_index(ks, 0, 1, lane)];
            const frag b1h = bfr[rfrag_index(ks, 1, 0, lane)], b1l = bfr[rfrag_index(ks, 1, 1, lane)];
            acc0 = P::mma32(ah, b0h, acc0);
            acc1 = P::mma32(ah, b1h, acc1);
            acc0 = P::mma32(ah, b0l, acc0);
            acc1 = P::mma32(ah, b1l, acc1);
            acc0 = P::mma32(al, b0h, acc0);
            acc1 = P::mma32(al, b1h, acc1);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        // ---- the radix-R stage, per frame slot of this lane ----
        // slot q (R = 4: frames 2q + h, q = 0..3; R = 16: frame h, one slot)
        constexpr int SLOTS = R == 4 ? 4 : 1;
        const uint32_t nf = (uint32_t)(frames - f0 < (uint64_t)FPT ? frames - f0 : (uint64_t)FPT);
        // Y[n1] of each slot (complex; lane 0: Re Y[.][0] in .x, Re Y[.][32]
        // in .y), gathered with every lane active (the R = 16 exchange is
        // cross-lane)
        float2 TT[SLOTS][R];
        if constexpr (R == 4) {
#pragma unroll
            for (int q = 0; q < SLOTS; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    TT[q][r] = make_float2(acc0[4 * q + r], acc1[4 * q + r]);
        } else {
            // element m < 8 is frame 0, n1 = 8b + 4h + r (b = m / 4, r = m % 4),
            // element 8 + m frame 1 at the same n1: v_permlane32_swap trades
            // the upper lanes' element m for the lower lanes' element 8 + m, after
            // which lane (j, h) holds n1 = 8b + r in element m and 8b + 4 + r in
            // element 8 + m, of frame h
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const auto re = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc0[m]),
                                                                 __float_as_uint(acc0[8 + m]), false, false);
                const auto im = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc1[m]),
                                                                 __float_as_uint(acc1[8 + m]), false, false);
                const int b = m / 4, r = m % 4;
                TT[0][8 * b + r] = make_float2(__uint_as_float(re[0]), __uint_as_float(im[0]));
                TT[0][8 * b + 4 + r] = make_float2(__uint_as_float(re[1]), __uint_as_float(im[1]));
            }
        }
        // every slot's bins, computed once with every lane active: Z[q][.]
        // and, on lane 0, nyq[q] = X[N/2]; scl[q] undoes the frame's scaling
        float2 Z[SLOTS][R], nyq[SLOTS];
        float scl[SLOTS];
#pragma unroll
        for (int q = 0; q < SLOTS; ++q) {
            float2 (&T)[R] = Z[q];
#pragma unroll
            for (int p = 0; p < R; ++p) {
                // lane 0: z[p] = v[2p] + i v[2p + 1]
                const int c0 = (2 * p) % R, c1 = (2 * p + 1) % R;
                const float2 z = 2 * p < R ? make_float2(TT[q][c0].x + TT[q][c0].y, TT[q][c1].x + TT[q][c1].y)
                                           : make_float2(TT[q][c0].x - TT[q][c0].y, TT[q][c1].x - TT[q][c1].y);
                const float2 t = p == 0 ? TT[q][0] : cmul(TT[q][p], twl[p * 32 + j]);
                T[p] = j == 0 ? z : t;
            }
            dft<R, 0>(T);
            const uint32_t fl = R == 4 ? 2u * q + h : h;  // frame of the slot within the tile
            scl[q] = inv_row[wave][fl < (uint32_t)FPT ? fl : 0];
            nyq[q] = make_float2(0.0f, 0.0f);
            if (j == 0) {
                // V[m] (times 4) from Z, then placed where the other lanes
                // keep the bins of the same output round: V[2 k1] in T[k1],
                // conj V[2 k1 + 1] in T[R - 1 - k1]
                float2 V[R + 1];
#pragma unroll
                for (int m = 0; m <= R / 2; ++m) {
                    const float2 A = T[m], B = T[(R - m) % R];  // Z[m], Z[R - m]
                    const float2 Pv = make_float2(A.x + B.x, A.y - B.y);
                    const float2 Q = make_float2(A.y + B.y, B.x - A.x);  // -i (Z[m] - conj Z[R - m])
                    const float2 WQ = m == 0 ? Q : cmul(Q, twl[m * 32]);  // W_2R^m
                    V[m] = make_float2(Pv.x + WQ.x, Pv.y + WQ.y);
                    if (2 * m != R)
                        V[R - m] = make_float2(Pv.x - WQ.x, WQ.y - Pv.y);
                }
                if (dc)
                    V[0] = make_float2(0.0f, 0.0f);
#pragma unroll
                for (int k1 = 0; k1 < R / 2; ++k1) {
                    T[k1] = V[2 * k1];
                    T[R - 1 - k1] = make_float2(V[2 * k1 + 1].x, -V[2 * k1 + 1].y);
                }
                nyq[q] = V[R];
                scl[q] *= 0.25f;
            }
        }
        // the next tile's samples load while this one's bins are stored
        raw = load_rtile<R, VEC>(in, L, hop, frames, tiles_per_signal, tile + stride, n1r, fr, h, x);
        // slot q's bins to sink(k, re, im):
        //   lane j >= 1: X[j + 64 k1] = Z[k1], X[64 - j + 64 k1] = conj Z[R - 1 - k1];
        //   lane 0:      X[64 k1] = V[2 k1],   X[32 + 64 k1] = V[2 k1 + 1], X[N/2] = V[R]
        const uint32_t kb = j == 0 ? 32u : 64u - j;
        auto put = [&](int q, auto&& sink) {
            const float sc = scl[q];
#pragma unroll
            for (int k1 = 0; k1 < R / 2; ++k1) {
                sink(j + 64u * k1, Z[q][k1].x * sc, Z[q][k1].y * sc);
                const float2 zc = Z[q][R - 1 - k1];
                sink(kb + 64u * k1, zc.x * sc, -zc.y * sc);
            }
            if (j == 0)
                sink(32u * R, nyq[q].x * sc, nyq[q].y * sc);
        };
        auto emit = [&](auto mode_c) {
            constexpr int MODE = decltype(mode_c)::value;
            constexpr int PF = layout_floats<MODE, N>();
            constexpr int FR = C::DIRECT ? 0 : stage_rounds_frames(FPT, PF, STAGE);  // frames per output round
            if constexpr (FR == 0) {
                // direct stores: each store instruction writes 32 consecutive
                // bins of each of two frames
#pragma unroll
                for (int q = 0; q < SLOTS; ++q) {
                    const uint32_t fl = R == 4 ? 2u * q + h : h;
                    if (fl >= nf)
                        continue;
                    const uint64_t f = f0 + fl;
                    put(q, [&](uint32_t k, float re, float im) { put_bin<MODE, N>(out, a.out_limit, f, k, re, im); });
                }
            } else {
#pragma unroll
                for (int r0 = 0; r0 < FPT; r0 += FR) {
#pragma unroll
                    for (int q = 0; q < SLOTS; ++q) {
                        const uint32_t fl = R == 4 ? 2u * q + h : h;  // frame within the tile
                        if (fl < (uint32_t)r0 || fl >= (uint32_t)(r0 + FR) || fl >= nf)
                            continue;
                        float* frb = stage + (fl - r0) * PF;
                        put(q, [&](uint32_t k, float re, float im) { stage_bin<MODE, N>(frb, k, re, im); });
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t fr0 = (uint32_t)r0 < nf ? (uint32_t)r0 : nf;
                    const uint32_t fr1 = (uint32_t)(r0 + FR) < nf ? (uint32_t)(r0 + FR) : nf;
                    uint64_t begin = (f0 + fr0) * (uint64_t)PF, end = (f0 + fr1) * (uint64_t)PF;
                    if constexpr (MODE == OUT_HALF_COMPLEX) {
                        const uint64_t lim = 2 * a.out_limit;
                        end = end < lim ? end : lim;
                        begin = begin < end ? begin : end;
                    }
                    write_block_n<FR * PF>(out + begin, stage, (uint32_t)(end - begin), lane);
                    __builtin_amdgcn_wave_barrier();  // the next round rewrites the stage
                }
            }
        };
        switch (mode) {
        case OUT_POWER: emit(std::integral_constant<int, OUT_POWER>{}); break;
        case OUT_HALF_SPECTRUM: emit(std::integral_constant<int, OUT_HALF_SPECTRUM>{}); break;
        case OUT_COMPAT_POWER: emit(std::integral_constant<int, OUT_COMPAT_POWER>{}); break;
        case OUT_MAGNITUDE: emit(std::integral_constant<int, OUT_MAGNITUDE>{}); break;
        case OUT_HALF_COMPLEX: emit(std::integral_constant<int, OUT_HALF_COMPLEX>{}); break;
        default: break;
        }
    }
}

template <int LOG_R, int XF>
hipError_t launch_r(const FusedArgs& a, uint32_t count, hipStream_t s)
{
    using C = RCfg<LOG_R>;
    constexpr int FPT = 32 >> LOG_R;
    const uint64_t tps = (a.frames + FPT - 1) / FPT;
    if (tps == 0 || count == 0)
        return hipSuccess;
    if (!a.mfma_b)
        return hipErrorInvalidValue;
    if (tps > 0xffffffffull || count > 65535u)
        return hipErrorInvalidValue;
    const uint64_t need = (tps + C::WAVES - 1) / C::WAVES;  // one tile per wave
    const uint64_t total = (uint64_t)device_cus() * OSTFT_MFMA_GRID_WAVES_PER_CU / C::WAVES;
    const uint64_t cap = (total + count - 1) / count;
    const uint64_t blocks = need < cap ? need : cap;
    // 16-byte input loads where every frame starts on a 16-byte boundary
    const bool vec = LOG_R == 2 && (reinterpret_cast<uintptr_t>(a.in) & 15u) == 0 && (a.hop & 3u) == 0 &&
                     (count == 1 || (a.in_stride & 3u) == 0);
    if (vec)
        hipLaunchKernelGGL((stft_mfma_r_kernel<LOG_R, XF, LOG_R == 2>), dim3((uint32_t)blocks, count),
                           dim3(C::WAVES * 64), 0, s, a, (uint32_t)tps);
    else
        hipLaunchKernelGGL((stft_mfma_r_kernel<LOG_R, XF, false>), dim3((uint32_t)blocks, count),
                           dim3(C::WAVES * 64), 0, s, a, (uint32_t)tps);
    return hipGetLastError();
}

template <int XF>
hipError_t launch_table(int log_m, const float* win, const float2* tw_n, void* table, hipStream_t s)
{
    if (log_m == 5)
        hipLaunchKernelGGL(mfma_table_kernel<XF>, dim3(1), dim3(64), 0, s, win, tw_n,
                           static_cast<typename Prec<XF>::frag*>(table));
    else if (log_m == 7)
        hipLaunchKernelGGL((mfma_rtable_kernel<2, XF>), dim3(1), dim3(64), 0, s, win, tw_n,
                           static_cast<char*>(table));
    else if (log_m == 9)
        hipLaunchKernelGGL((mfma_rtable_kernel<4, XF>), dim3(1), dim3(64), 0, s, win, tw_n,
                           static_cast<char*>(table));
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int XF>
hipError_t launch_gemm(int log_m, const FusedArgs& a, uint32_t count, hipStream_t s)
{
    switch (log_m) {
    case 5: return launch64<XF>(a, count, s);
    case 7: return launch_r<2, XF>(a, count, s);
    case 9: return launch_r<4, XF>(a, count, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool mfma_gemm_supported(int log_m) { return log_m == 5 || log_m == 7 || log_m == 9; }

size_t mfma_table_bytes(int log_m)
{
    return log_m == 5 ? TABLE_BYTES : log_m == 7 ? RB_BYTES + RW_BYTES + rt_bytes(4) : log_m == 9 ? RB_BYTES + RW_BYTES + rt_bytes(16) : 0;
}

hipError_t launch_mfma_table(int log_m, int xf, const float* win, const float2* tw_n, void* table, hipStream_t s)
{
    if (xf == XF_MFMA_FP16)
        return launch_table<XF_MFMA_FP16>(log_m, win, tw_n, table, s);
    if (xf == XF_MFMA_BF16)
        return launch_table<XF_MFMA_BF16>(log_m, win, tw_n, table, s);
    return hipErrorInvalidValue;
}

hipError_t launch_mfma_gemm(int log_m, int xf, const FusedArgs& a, uint32_t count, hipStream_t s)
{
    if (xf == XF_MFMA_FP16)
        return launch_gemm<XF_MFMA_FP16>(log_m, a, count, s);
    if (xf == XF_MFMA_BF16)
        return launch_gemm<XF_MFMA_BF16>(log_m, a, count, s);
    return hipErrorInvalidValue;
}

}  // namespace ostft
