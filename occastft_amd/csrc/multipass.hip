// multipass.hip -- STFT for complex transform sizes M = 2^15 .. 2^19
// (windows 2^16 .. 2^20; the plan also covers M = 2^13, 2^14, which the
// single kernel handles).
//
// Same pipeline as the fused kernel (gather + window, FFT, split epilogue),
// but the transform no longer fits one workgroup's LDS, so it is cut into two
// Stockham passes through HBM-resident scratch (this replaces the EXP
// separate radix-2 StockHamCommon launches of OKL/EXPCommon.okl:194-245, one
// per stage, and the separate ToPower / toHalfComplexFormat passes).
//
// Both passes work on 8192 complex values per workgroup (512 threads, 70 KB of
// LDS): CB = 8192/R butterflies of one frame, gathered as R rows of CB
// contiguous values (stride M/R), FFT'd as CB independent R-point transforms
// with the fused kernel's LDS Stockham engine.
//   * first pass (radix R0 <= 2^11): reads the signal itself (overlap gather +
//     window), writes the Stockham outputs Y[g R0 + q] (each butterfly's R0
//     results contiguous);
//   * last pass (radix R1 <= 2^8, stride Ns = R0): applies the inter-pass
//     twiddle W_M^{g m} from a two-level FP64 table, and holds a block of
//     butterflies g together with its mirror block Ns - g (0 and Ns/2 are their
//     own mirrors), so both Z[k] and Z[M-k] of every output index k are in its
//     LDS: the split epilogue (last radix-2 stage, DC correction, output layout)
//     runs there and the spectrum is written once.
// As in the fused kernel, a frame of N real samples is one complex transform
// of M = N/2 points (even samples real, odd imaginary).  DC removal is applied
// in the frequency domain: subtracting the frame mean only changes the k = 0
// terms.
#include "../../include/occastft_hip.h"

#include <mutex>
#include <string>

#include "batch.hpp"
#include "epilogue.hpp"
#include "multipass.hpp"
#include "stockham.hpp"

namespace ostft {

hipError_t Scratch::ensure(uint64_t n)
{
    if (cap >= n)
        return hipSuccess;
    release();
    hipError_t e = hipMalloc(&a, sizeof(float2) * n);
    if (e != hipSuccess) {
        release();
        return e;
    }
    cap = n;
    return hipSuccess;
}

hipError_t Scratch::acquire(hipStream_t s)
{
    if (!done) {
        const hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
        if (e != hipSuccess) {
            done = nullptr;
            return e;
        }
    }
    return used && last != s ? hipStreamWaitEvent(s, done, 0) : hipSuccess;
}

hipError_t Scratch::handoff(hipStream_t s)
{
    const hipError_t e = hipEventRecord(done, s);
    if (e == hipSuccess) {
        last = s;
        used = true;
    }
    return e;
}

void Scratch::release()
{
    if (a) {
        // the last call's launches (on whatever stream) must be done with it;
        // other streams and contexts on the device are not waited for
        if (used && done)
            (void)hipEventSynchronize(done);
        (void)hipFree(a);
    }
    a = nullptr;
    cap = 0;
}

void Scratch::destroy()
{
    release();
    if (done)
        (void)hipEventDestroy(done);
    done = nullptr;
    used = false;
}

// Extra complex values per LDS column: the workgroup's global <-> LDS
// transposition writes one value per column across consecutive lanes, so an
// odd column stride spreads those lanes over all banks (R + R/16 is a
// multiple of 4: up to 16 lanes of a wave on one bank).
// First pass: 16-byte loads of two neighbouring packed samples of a row
// where the signal and the window allow it (half the load instructions).
#ifndef OSTFT_MP_VEC4
#define OSTFT_MP_VEC4 1
#endif
// First pass: its scratch rows written two values per 16-byte store.
#ifndef OSTFT_MP_VEC4_ST
#define OSTFT_MP_VEC4_ST 1
#endif
#ifndef OSTFT_MP_COL_PAD
#define OSTFT_MP_COL_PAD 1
#endif

// first pass: frames per workgroup-order group (see first_pass_kernel)
#ifndef OSTFT_MP_FRAME_GROUP
#define OSTFT_MP_FRAME_GROUP 1
#endif

constexpr int MP_THREADS = 512;
constexpr int MP_ELEMS = 16 * MP_THREADS;  // complex values per workgroup

struct PassArgs {
    const float* sig;  // first pass: the signal
    uint64_t L, hop;
    uint64_t frame0, frames;  // frames [frame0, frames) of this launch; scratch row 0 = frame0
    const float* win;
    float2* dst;
    const float2* tw_r;  // W_R^j, j < R
    uint32_t log_m;
};

template <int LOG_R>
__global__ void __launch_bounds__(MP_THREADS) first_pass_kernel(const PassArgs a)
{
    constexpr int R = 1 << LOG_R, CB = MP_ELEMS / R, T = R / 16, PADR = R + R / 16 + OSTFT_MP_COL_PAD;
    static_assert(CB * T == MP_THREADS, "MP_THREADS threads per block");
    __shared__ float2 lds[CB * PADR];

    const uint32_t tid = threadIdx.x;
    const uint64_t M = 1ull << a.log_m;
    const uint64_t bpg = M / MP_ELEMS;  // blocks per transform = (M / R) / CB
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    // Workgroup order: frames in groups of FG; within a group, consecutive
    // workgroups (one XCD, about the same time) take the GB blocks of a block
    // group for each frame of the group in turn, then the next block group.
    // The GB blocks share their 128-byte lines (a block's rows are CB values
    // long), overlapping frames of a group share rows and the group shares
    // each window segment, so those come from HBM about once per group; FG = 1
    // is plain frame order.
    constexpr uint32_t GB = CB >= 16 ? 1u : 16u / CB;  // bpg is a multiple of GB (M >= 2^13)
    const uint64_t F = a.frames - a.frame0, fg = blk / (OSTFT_MP_FRAME_GROUP * bpg);
    const uint64_t p0 = fg * OSTFT_MP_FRAME_GROUP;
    if (p0 >= F)
        return;  // uniform over the block
    const uint64_t q = F - p0 < OSTFT_MP_FRAME_GROUP ? F - p0 : OSTFT_MP_FRAME_GROUP;  // frames in this group
    const uint64_t i = blk - fg * OSTFT_MP_FRAME_GROUP * bpg, r = i % (q * GB);
    const uint64_t P = a.frame0 + p0 + r / GB, bi = (i / (q * GB)) * GB + r % GB;
    const uint32_t g0 = (uint32_t)bi * CB;
    const uint32_t stride = (uint32_t)(M >> LOG_R);
    PassTwiddles<LOG_R> tws;
    load_pass_twiddles<LOG_R>(tws, a.tw_r, tid % T);

    // rows of CB contiguous packed samples z[n] = x[2n] + i x[2n+1], windowed
    const uint64_t fs = P * a.hop;
    const float2* win2 = reinterpret_cast<const float2*>(a.win);
    const uint64_t n_last = g0 + (CB - 1) + (uint64_t)(R - 1) * stride;  // the block's last packed index
    const bool inside = fs + 2 * n_last + 2 <= a.L;
    if (OSTFT_MP_VEC4 && CB >= 2 && inside && (reinterpret_cast<uintptr_t>(a.sig + fs) & 15u) == 0 &&
        (reinterpret_cast<uintptr_t>(win2) & 15u) == 0) {
        // pairs of neighbouring packed samples (columns c, c + 1 of a row;
        // n even, so each pair is one 16-byte block of signal and of window)
        constexpr int CH = CB >= 2 ? CB / 2 : 1;
        float4 xv[8], wv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            const uint32_t n = g0 + 2u * (e % CH) + (e / CH) * stride;
            xv[j] = *reinterpret_cast<const float4*>(a.sig + fs + 2ull * n);
            wv[j] = win2 ? *reinterpret_cast<const float4*>(win2 + n) : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            const uint32_t c = 2u * (e % CH), m = e / CH;
            lds[c * PADR + pad(m)] = make_float2(xv[j].x * wv[j].x, xv[j].y * wv[j].y);
            lds[(c + 1) * PADR + pad(m)] = make_float2(xv[j].z * wv[j].z, xv[j].w * wv[j].w);
        }
    } else if ((reinterpret_cast<uintptr_t>(a.sig + fs) & 7u) == 0 && inside) {
        // the whole block lies inside the signal and its pairs are 8-byte
        // aligned: every load is issued before the first one is used
        float2 xv[16], wv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            const uint32_t n = g0 + e % CB + (e / CB) * stride;
            xv[j] = *reinterpret_cast<const float2*>(a.sig + fs + 2ull * n);
            wv[j] = win2 ? win2[n] : make_float2(1.0f, 1.0f);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            lds[(e % CB) * PADR + pad(e / CB)] = make_float2(xv[j].x * wv[j].x, xv[j].y * wv[j].y);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            const uint32_t c = e % CB, m = e / CB;
            const uint32_t n = g0 + c + m * stride;
            const uint64_t s0 = fs + 2ull * n;
            const float2 w = win2 ? win2[n] : make_float2(1.0f, 1.0f);
            float2 xv;
            xv.x = s0 < a.L ? a.sig[s0] : 0.0f;
            xv.y = s0 + 1 < a.L ? a.sig[s0 + 1] : 0.0f;
            lds[c * PADR + pad(m)] = make_float2(xv.x * w.x, xv.y * w.y);
        }
    }
    __syncthreads();

    const uint32_t col = tid / T, t = tid % T;
    float2* z = lds + col * PADR;
    float2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        v[j] = z[pad(t + j * T)];
    __syncthreads();
    stockham<LOG_R, 0>(v, z, tws, t);

    // Y[g R + q]: each butterfly's R outputs are contiguous
    const uint64_t obase = (P - a.frame0) * M + ((uint64_t)g0 << LOG_R);
    if constexpr (OSTFT_MP_VEC4_ST) {
        // two neighbouring outputs of a butterfly per 16-byte store (obase
        // and e even)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t e = 2u * (tid + MP_THREADS * j);
            const uint32_t c = e >> LOG_R, q = e & (R - 1);
            const float2 v0 = lds[c * PADR + pad(q)], v1 = lds[c * PADR + pad(q + 1)];
            *reinterpret_cast<float4*>(a.dst + obase + e) = make_float4(v0.x, v0.y, v1.x, v1.y);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t e = tid + MP_THREADS * j;
            const uint32_t c = e >> LOG_R, q = e & (R - 1);
            a.dst[obase + e] = lds[c * PADR + pad(q)];
        }
    }
}

struct LastArgs {
    const float2* src;        // first pass output
    const float2* tw_r;       // W_R^j, j < R
    const float2* tw_fine;    // W_M^j, j < 2^fine_bits
    const float2* tw_coarse;  // W_M^(j << fine_bits)
    const float2* tw_n;       // W_N^k, k < M
    const float2* tw_2r;      // W_2R^q = W_N^(q Ns), q < R
    float* out;
    uint64_t frame0, frames, out_limit;  // frames [frame0, frames); scratch row 0 = frame0
    uint32_t fine_bits, log_m;
    int32_t combine, out_mode, dc;
};

template <int LOG_R>
__global__ void __launch_bounds__(MP_THREADS) last_pass_kernel(const LastArgs a)
{
    constexpr int R = 1 << LOG_R, CB2 = MP_ELEMS / R, CB = CB2 / 2, T = R / 16, PADR = R + R / 16 + OSTFT_MP_COL_PAD;
    static_assert(CB2 * T == MP_THREADS && CB >= 1, "MP_THREADS threads per block");
    __shared__ float2 lds[CB2 * PADR];

    const uint32_t tid = threadIdx.x;
    const uint64_t M = 1ull << a.log_m;
    const uint32_t Ns = (uint32_t)(M >> LOG_R);  // butterflies of the pass (= its stride)
    const uint64_t bpf = M / MP_ELEMS;           // blocks per frame = Ns / CB2
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t P = a.frame0 + blk / bpf;
    if (P >= a.frames)
        return;  // uniform over the block
    const uint32_t b = (uint32_t)(blk % bpf);
    PassTwiddles<LOG_R> tws;
    load_pass_twiddles<LOG_R>(tws, a.tw_r, tid % T);

    // column c < CB: butterfly b CB + c; column CB + h: its mirror Ns - (b CB + h),
    // with Ns/2 in the place of Ns (= butterfly 0, already in column 0)
    auto butterfly = [&](uint32_t c) -> uint32_t {
        if (c < (uint32_t)CB)
            return b * CB + c;
        const uint32_t h = c - CB;
        return (b == 0 && h == 0) ? Ns / 2 : Ns - (b * CB + h);
    };
    // The thread's 16 values share one column: c = tid mod CB2 (butterfly g),
    // rows m = m0 + T j.  Inter-pass twiddle W_M^{g m} (g m < M): from the
    // two-level FP64 table at every fourth row, and by products with
    // W_M^{g T} in between (at most three chained roundings).
    const uint32_t fmask = (1u << a.fine_bits) - 1u;
    auto twiddle = [&](uint32_t ex) {
        return cmul(a.tw_fine[ex & fmask], a.tw_coarse[ex >> a.fine_bits]);
    };
    const float2* src = a.src + (P - a.frame0) * M;
    const uint32_t cc = tid % CB2, m0 = tid / CB2;
    const uint32_t gc = butterfly(cc);
    float2 xs[16];  // the scratch values first: nothing below waits on them until the product
#pragma unroll
    for (int j = 0; j < 16; ++j)
        xs[j] = src[gc + (uint64_t)(m0 + T * j) * Ns];
    // the epilogue's W_N^g and W_2R^q, loaded with the rest
    float2 wg = make_float2(1.0f, 0.0f), w2r[16];
    if (a.combine) {
        wg = a.tw_n[gc];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            w2r[j] = a.tw_2r[m0 + T * j];
    }
    const float2 step = twiddle(gc * T);
    float2 anchor[4];
#pragma unroll
    for (int j = 0; j < 16; j += 4)
        anchor[j / 4] = twiddle(gc * (m0 + T * j));
    float2 tw = anchor[0];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t m = m0 + T * j;
        tw = (j % 4 == 0) ? anchor[j / 4] : cmul(tw, step);
        lds[cc * PADR + pad(m)] = cmul(xs[j], tw);
    }
    __syncthreads();

    const uint32_t col = tid / T, t = tid % T;
    float2* z = lds + col * PADR;
    float2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        v[j] = z[pad(t + j * T)];
    __syncthreads();
    stockham<LOG_R, 0>(v, z, tws, t);
    // column c, row q now holds Z[k], k = butterfly(c) + q Ns

    // split: Z[M - k] is column mirror(c), row R - 1 - q (row (R - q) mod R
    // for butterfly 0).  The thread's bins are k = g + q Ns, q = m0 + T j, of
    // its column c = cc: W_N^k = W_N^g W_2R^q.
    uint32_t mc;
    if (gc == 0 || (gc == Ns / 2 && b == 0))
        mc = cc;
    else
        mc = cc < (uint32_t)CB ? cc + CB : cc - CB;
    // one straight-line body per output layout and stage choice
    auto epilogue = [&](auto mode_c, auto comb_c) {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr bool COMB = decltype(comb_c)::value;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t q = m0 + T * j;
            const uint32_t mq = gc == 0 ? (R - q) & (R - 1) : R - 1 - q;
            const uint32_t k = gc + q * Ns;
            const float2 wk = COMB ? cmul(wg, w2r[j]) : wg;
            const BinPair bp = finish_bin(lds[cc * PADR + pad(q)], lds[mc * PADR + pad(mq)], wk, COMB,
                                          a.dc && k == 0);
            write_bin(a.out, MODE, P, (uint32_t)M, k, bp, a.out_limit);
        }
    };
    auto by_stage = [&](auto mode_c) {
        if (a.combine)
            epilogue(mode_c, std::true_type{});
        else
            epilogue(mode_c, std::false_type{});
    };
    switch (a.out_mode) {
    case OUT_COMPAT_POWER: by_stage(std::integral_constant<int, OUT_COMPAT_POWER>{}); break;
    case OUT_MAGNITUDE: by_stage(std::integral_constant<int, OUT_MAGNITUDE>{}); break;
    case OUT_HALF_COMPLEX: by_stage(std::integral_constant<int, OUT_HALF_COMPLEX>{}); break;
    case OUT_POWER: by_stage(std::integral_constant<int, OUT_POWER>{}); break;
    case OUT_HALF_SPECTRUM: by_stage(std::integral_constant<int, OUT_HALF_SPECTRUM>{}); break;
    default: break;
    }
}

int multipass_plan(int log_m, int* lr)
{
    // two passes: the last radix small, so that a block and its mirror block
    // (CB = 4096 / R1 butterflies each) write long runs of consecutive bins
    // (R1 = 2^6: 256-byte runs of the real layouts; 2^8 gave 64-byte runs and
    // a 4-5% slower large-window step, 2^9 was 1.7x slower), the first radix
    // at most 2^11 (its column segments are CB = 8192 / R0 values)
    if (log_m < 13 || log_m > 19)
        return 0;
    lr[1] = log_m / 2 > 6 ? 6 : log_m / 2;
    if (log_m - lr[1] > 11)
        lr[1] = log_m - 11;
    lr[0] = log_m - lr[1];
    return 2;
}

template <int LOG_R>
static hipError_t launch_first(const PassArgs& p, hipStream_t s)
{
    const uint64_t blocks = (p.frames - p.frame0) * ((1ull << p.log_m) / MP_ELEMS);
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(first_pass_kernel<LOG_R>, dim3((uint32_t)blocks), dim3(MP_THREADS), 0, s, p);
    return hipGetLastError();
}

template <int LOG_R>
static hipError_t launch_last(const LastArgs& p, hipStream_t s)
{
    const uint64_t blocks = (p.frames - p.frame0) * ((1ull << p.log_m) / MP_ELEMS);
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(last_pass_kernel<LOG_R>, dim3((uint32_t)blocks), dim3(MP_THREADS), 0, s, p);
    return hipGetLastError();
}

static hipError_t first_pass(int log_r, const PassArgs& p, hipStream_t s)
{
    switch (log_r) {
    case 7: return launch_first<7>(p, s);
    case 8: return launch_first<8>(p, s);
    case 9: return launch_first<9>(p, s);
    case 10: return launch_first<10>(p, s);
    case 11: return launch_first<11>(p, s);
    default: return hipErrorInvalidValue;
    }
}

static hipError_t last_pass(int log_r, const LastArgs& p, hipStream_t s)
{
    switch (log_r) {
    case 6: return launch_last<6>(p, s);
    case 7: return launch_last<7>(p, s);
    case 8: return launch_last<8>(p, s);
    default: return hipErrorInvalidValue;
    }
}

uint64_t multipass_chunk_frames(uint64_t budget, int log_m)
{
    const uint64_t n = budget / (sizeof(float2) << log_m);
    return n ? n : 1;
}

namespace {

// The twiddle tables of the two passes (resident in the context).  With
// `noalloc` a table that is not resident yet is an error instead of being
// built.
struct PassTables {
    const float2 *fine, *coarse, *r0, *r1, *r2r1;
    uint32_t fine_bits;
};

int pass_tables(ostft_ctx* ctx, int log_m, const int* lr, bool noalloc, PassTables* t)
{
    int rc = OSTFT_OK;
    const int64_t M = (int64_t)1 << log_m;
    t->fine_bits = (uint32_t)(log_m + 1) / 2;
    if (!(t->fine = ctx_twiddles(ctx, M, (int64_t)1 << t->fine_bits, &rc, noalloc)))
        return rc;
    if (!(t->coarse = ctx_twiddles(ctx, M >> t->fine_bits, M >> t->fine_bits, &rc, noalloc)))
        return rc;
    if (!(t->r0 = ctx_twiddles(ctx, (int64_t)1 << lr[0], (int64_t)1 << lr[0], &rc, noalloc)))
        return rc;
    if (!(t->r1 = ctx_twiddles(ctx, (int64_t)1 << lr[1], (int64_t)1 << lr[1], &rc, noalloc)))
        return rc;
    if (!(t->r2r1 = ctx_twiddles(ctx, (int64_t)2 << lr[1], (int64_t)1 << lr[1], &rc, noalloc)))
        return rc;
    return OSTFT_OK;
}

// The fewest chunks of at most `cap` frames, of equal size (a request of
// 2^k + 1 frames gets no one-frame chunk at its end).
uint64_t even_chunk(uint64_t frames, uint64_t cap)
{
    const uint64_t nchunks = (frames + cap - 1) / cap;
    return nchunks ? (frames + nchunks - 1) / nchunks : 1;
}

// Both passes over the request's frames, `chunk` frames at a time through
// `buf` (chunk * M complex values).
int run_chunks(ostft_ctx* ctx, int log_m, const int* lr, const FusedArgs& a, const PassTables& t, float2* buf,
               uint64_t chunk, hipStream_t s)
{
    PassArgs p;
    p.sig = a.in;
    p.L = a.L;
    p.hop = a.hop;
    p.win = a.win;
    p.dst = buf;
    p.tw_r = t.r0;
    p.log_m = (uint32_t)log_m;

    LastArgs la;
    la.src = buf;
    la.tw_r = t.r1;
    la.tw_fine = t.fine;
    la.tw_coarse = t.coarse;
    la.tw_n = a.tw_n;
    la.tw_2r = t.r2r1;
    la.out = a.out;
    la.out_limit = a.out_limit;
    la.fine_bits = t.fine_bits;
    la.log_m = (uint32_t)log_m;
    la.combine = a.combine;
    la.out_mode = a.out_mode;
    la.dc = a.dc;

    for (uint64_t f0 = 0; f0 < a.frames; f0 += chunk) {
        const uint64_t f1 = a.frames - f0 < chunk ? a.frames : f0 + chunk;
        p.frame0 = la.frame0 = f0;
        p.frames = la.frames = f1;
        hipError_t e = first_pass(lr[0], p, s);
        if (e == hipSuccess)
            e = last_pass(lr[1], la, s);
        if (e != hipSuccess)
            return ctx_hip_fail(ctx, e, "multi-pass launch");
    }
    return OSTFT_OK;
}

// Scratch of the allocation-free entry point: a fixed device array, one
// per GPU, shared by every such call on that GPU (calls that overlap in
// time on one GPU race on it).  Host code reaches it only through the
// address hipGetSymbolAddress returns for the device (the array's name is a
// host-side stand-in, not a device address).
constexpr uint64_t STATIC_SCRATCH_BYTES = 64ull << 20;
__device__ __align__(256) float2 scratch[STATIC_SCRATCH_BYTES / sizeof(float2)];

constexpr int MAX_DEVICES = 64;
std::mutex static_mu;
float2* static_ptr[MAX_DEVICES] = {};

}  // namespace

uint64_t multipass_static_scratch_bytes() { return STATIC_SCRATCH_BYTES; }

hipError_t multipass_static_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(static_mu);
    if (static_ptr[device])
        return hipSuccess;
    // the caller made `device` current: this is that device's copy
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    static_ptr[device] = static_cast<float2*>(ptr);
    return hipSuccess;
}

int run_multipass(ostft_ctx* ctx, int log_m, const FusedArgs& a, hipStream_t s)
{
    int lr[4];
    if (multipass_plan(log_m, lr) != 2)
        return OSTFT_ERR_UNSUPPORTED;
    const uint64_t M = 1ull << log_m;
    // one call at a time queues launches on the shared scratch; the event
    // chain (acquire / handoff) orders their GPU work across streams
    std::lock_guard<std::mutex> lk(ctx_scratch_mutex(ctx));
    Scratch& sc = ctx_scratch(ctx);
    // Frames go through both passes a chunk at a time, so the scratch
    // allocation stays bounded (1 GiB by default) whatever the signal
    // length.  Chunks small enough for their scratch to stay in the 256 MB
    // memory-side cache between the passes measured slower (64 MiB: +15%,
    // 128 MiB: +3% on windows 2^16..2^20): the passes are not bound by the
    // round trip's HBM traffic, and every chunk adds two launch tails.
    const uint64_t chunk = even_chunk(a.frames, multipass_chunk_frames(sc.budget, log_m));
    hipError_t e = sc.ensure((a.frames < chunk ? a.frames : chunk) * M);
    if (e != hipSuccess)
        return ctx_hip_fail(ctx, e, "multi-pass scratch allocation");
    e = sc.acquire(s);
    if (e != hipSuccess)
        return ctx_hip_fail(ctx, e, "multi-pass scratch ordering");
    PassTables t;
    int rc = pass_tables(ctx, log_m, lr, false, &t);
    if (rc == OSTFT_OK)
        rc = run_chunks(ctx, log_m, lr, a, t, sc.a, chunk, s);
    if (rc != OSTFT_OK)
        return rc;
    e = sc.handoff(s);
    return e == hipSuccess ? OSTFT_OK : ctx_hip_fail(ctx, e, "multi-pass scratch event");
}

int run_multipass_static(ostft_ctx* ctx, int device, int log_m, const FusedArgs& a, hipStream_t s)
{
    int lr[4];
    if (multipass_plan(log_m, lr) != 2)
        return OSTFT_ERR_UNSUPPORTED;
    const uint64_t frame_bytes = sizeof(float2) << log_m;
    if (frame_bytes > STATIC_SCRATCH_BYTES)
        return ctx_fail(ctx, OSTFT_ERR_TOO_LARGE,
                        "one frame's intermediate spectrum (" + std::to_string(frame_bytes)
                            + " B) exceeds the fixed scratch of the allocation-free entry point ("
                            + std::to_string(STATIC_SCRATCH_BYTES) + " B)");
    float2* buf = nullptr;
    {
        std::lock_guard<std::mutex> lk(static_mu);
        buf = device >= 0 && device < MAX_DEVICES ? static_ptr[device] : nullptr;
    }
    if (!buf)
        return ctx_fail(ctx, OSTFT_ERR_HIP, "fixed scratch of this device was not set up at context creation");
    PassTables t;
    const int rc = pass_tables(ctx, log_m, lr, true, &t);
    if (rc != OSTFT_OK)
        return rc;
    return run_chunks(ctx, log_m, lr, a, t, buf, even_chunk(a.frames, STATIC_SCRATCH_BYTES / frame_bytes), s);
}

}  // namespace ostft
