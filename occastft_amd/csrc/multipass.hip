// multipass.hip -- STFT for complex transform sizes M = 2^13 .. 2^20.
//
// Same pipeline as the fused kernel (gather + window, FFT, split epilogue),
// but the transform no longer fits one workgroup's LDS, so it is cut into
// 2 or 3 Stockham passes of radix R = 2^5 .. 2^8 through HBM-resident
// scratch (this replaces the EXP separate radix-2 StockHamCommon launches of
// OKL/EXPCommon.okl:194-245, one per stage).  Each pass kernel owns CB = 4096/R
// consecutive butterflies of one transform: it gathers CB-wide contiguous rows
// (R of them, stride M/R), applies the inter-pass twiddle W_{Ns*R}^{(g mod Ns) m}
// from a two-level table, runs CB independent R-point FFTs with the same LDS
// Stockham engine as the fused kernel, and scatters the Stockham outputs
// Y[(g/Ns)Ns R + (g mod Ns) + q Ns] in CB-wide contiguous rows.
//
// As in the fused kernel, a frame of N real samples is one complex transform
// of M = N/2 points (even samples real, odd imaginary).  The first pass reads
// the signal itself (overlap gather + window); the last pass leaves the packed
// spectrum Z in natural order, and a split kernel reads Z[k], Z[M-k], applies
// the final radix-2 stage (epilogue.hpp) and writes the requested output.  DC
// removal is applied in the frequency domain there: subtracting the frame mean
// only changes the k = 0 terms.
#include "../../include/occastft_hip.h"
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
    if (e == hipSuccess)
        e = hipMalloc(&b, sizeof(float2) * n);
    if (e != hipSuccess) {
        release();
        return e;
    }
    cap = n;
    return hipSuccess;
}

void Scratch::release()
{
    if (a || b)
        (void)hipDeviceSynchronize();
    if (a)
        (void)hipFree(a);
    if (b)
        (void)hipFree(b);
    a = b = nullptr;
    cap = 0;
}

struct PassArgs {
    const float* sig;  // first pass: the signal
    uint64_t L, hop, frames;
    const float* win;
    int32_t first;
    const float2* src;       // later passes: previous pass output
    float2* dst;
    const float2* tw_r;      // W_R^j, j < R
    const float2* tw_fine;   // W_M^j, j < 2^fine_bits
    const float2* tw_coarse; // W_M^(j << fine_bits)
    uint32_t fine_bits;
    uint32_t log_m;
    uint32_t log_ns;         // log2 of the pass stride Ns
};

template <int LOG_R>
__global__ void __launch_bounds__(256) column_pass_kernel(const PassArgs a)
{
    constexpr int R = 1 << LOG_R, CB = 4096 / R, T = R / 16, PADR = R + R / 16;
    static_assert(CB * T == 256, "256 threads per block");
    __shared__ float2 lds[CB * PADR];

    const uint32_t tid = threadIdx.x;
    const uint64_t M = 1ull << a.log_m;
    const uint64_t bpg = M >> 12;  // blocks per transform = (M / R) / CB
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t P = blk / bpg;
    if (P >= a.frames)
        return;  // uniform over the block
    const uint32_t g0 = (uint32_t)(blk - P * bpg) * CB;
    const uint32_t stride = (uint32_t)(M >> LOG_R);
    const uint32_t ns = 1u << a.log_ns;
    PassTwiddles<LOG_R> tws;
    load_pass_twiddles<LOG_R>(tws, a.tw_r, tid % T);

#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t e = tid + 256u * j;
        const uint32_t c = e % CB, m = e / CB;
        const uint32_t n = g0 + c + m * stride;
        float2 x;
        if (a.first) {
            const uint64_t s0 = P * a.hop + 2ull * n;
            const float2 w = a.win ? reinterpret_cast<const float2*>(a.win)[n] : make_float2(1.0f, 1.0f);
            const float x0 = s0 < a.L ? a.sig[s0] : 0.0f;
            const float x1 = s0 + 1 < a.L ? a.sig[s0 + 1] : 0.0f;
            x = make_float2(x0 * w.x, x1 * w.y);
        } else {
            x = a.src[P * M + n];
            if (ns > 1) {
                // W_{Ns R}^{((g0+c) mod Ns) m} = W_M^ex, ex = ((g0+c) mod Ns) m M/(Ns R)
                const uint64_t ex = ((uint64_t)((g0 + c) & (ns - 1)) * m) << (a.log_m - a.log_ns - LOG_R);
                const float2 wf = a.tw_fine[ex & ((1ull << a.fine_bits) - 1)];
                const float2 wc = a.tw_coarse[ex >> a.fine_bits];
                x = cmul(x, cmul(wf, wc));
            }
        }
        lds[c * PADR + pad(m)] = x;
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

    const uint64_t obase = P * M;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t e = tid + 256u * j;
        uint32_t c, q;
        if (ns == 1) {
            c = e / R;
            q = e % R;
        } else {
            c = e % CB;
            q = e / CB;
        }
        const uint32_t g = g0 + c;
        const uint64_t addr = ((uint64_t)(g >> a.log_ns) << (a.log_ns + LOG_R)) + (g & (ns - 1)) + (uint64_t)q * ns;
        a.dst[obase + addr] = lds[c * PADR + pad(q)];
    }
}

struct SplitArgs {
    const float2* Z;
    const float2* tw_n;  // W_N^k, k < M
    float* out;
    uint64_t frames, out_limit;
    uint32_t log_m;
    int32_t combine, out_mode, dc;
};

__global__ void __launch_bounds__(256) split_kernel(const SplitArgs a)
{
    const uint64_t M = 1ull << a.log_m;
    const uint64_t total = a.frames * M;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256u) {
        const uint64_t f = i >> a.log_m;
        const uint32_t k = (uint32_t)(i & (M - 1));
        const uint32_t km = (uint32_t)((M - k) & (M - 1));
        const float2 w = a.combine ? a.tw_n[k] : make_float2(1.0f, 0.0f);
        const BinPair b = finish_bin(a.Z[f * M + k], a.Z[f * M + km], w, a.combine, a.dc && k == 0);
        write_bin(a.out, a.out_mode, f, (uint32_t)M, k, b, a.out_limit);
    }
}

int multipass_plan(int log_m, int* lr)
{
    const int np = (log_m + 7) / 8;
    int left = log_m;
    for (int i = 0; i < np; ++i) {
        const int parts = np - i;
        lr[i] = (left + parts - 1) / parts;
        left -= lr[i];
    }
    return np;
}

template <int LOG_R>
static hipError_t launch_pass(const PassArgs& p, hipStream_t s)
{
    const uint64_t blocks = p.frames * ((1ull << p.log_m) >> 12);
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(column_pass_kernel<LOG_R>, dim3((uint32_t)blocks), dim3(256), 0, s, p);
    return hipGetLastError();
}

int run_multipass(ostft_ctx* ctx, int log_m, const FusedArgs& a, hipStream_t s)
{
    if (log_m < 13 || log_m > 19)
        return OSTFT_ERR_UNSUPPORTED;
    int lr[4];
    const int np = multipass_plan(log_m, lr);
    const uint64_t M = 1ull << log_m;
    Scratch& sc = ctx_scratch(ctx);
    hipError_t e = sc.ensure(a.frames * M);
    if (e != hipSuccess)
        return OSTFT_ERR_HIP;
    int rc = OSTFT_OK;
    const uint32_t fb = (uint32_t)(log_m + 1) / 2;
    const float2* tw_fine = ctx_twiddles(ctx, (int64_t)M, (int64_t)1 << fb, &rc);
    if (!tw_fine)
        return rc;
    const float2* tw_coarse = ctx_twiddles(ctx, (int64_t)(M >> fb), (int64_t)(M >> fb), &rc);
    if (!tw_coarse)
        return rc;

    PassArgs p;
    p.sig = a.in;
    p.L = a.L;
    p.hop = a.hop;
    p.frames = a.frames;
    p.win = a.win;
    p.tw_fine = tw_fine;
    p.tw_coarse = tw_coarse;
    p.fine_bits = fb;
    p.log_m = (uint32_t)log_m;
    const float2* src = nullptr;
    float2* dst = sc.a;
    uint32_t log_ns = 0;
    for (int i = 0; i < np; ++i) {
        p.first = i == 0;
        p.src = src;
        p.dst = dst;
        p.log_ns = log_ns;
        p.tw_r = ctx_twiddles(ctx, (int64_t)1 << lr[i], (int64_t)1 << lr[i], &rc);
        if (!p.tw_r)
            return rc;
        switch (lr[i]) {
        case 5: e = launch_pass<5>(p, s); break;
        case 6: e = launch_pass<6>(p, s); break;
        case 7: e = launch_pass<7>(p, s); break;
        case 8: e = launch_pass<8>(p, s); break;
        default: return OSTFT_ERR_UNSUPPORTED;
        }
        if (e != hipSuccess)
            return OSTFT_ERR_HIP;
        log_ns += (uint32_t)lr[i];
        src = dst;
        dst = dst == sc.a ? sc.b : sc.a;
    }
    SplitArgs sa;
    sa.Z = src;
    sa.tw_n = a.tw_n;
    sa.out = a.out;
    sa.frames = a.frames;
    sa.out_limit = a.out_limit;
    sa.log_m = (uint32_t)log_m;
    sa.combine = a.combine;
    sa.out_mode = a.out_mode;
    sa.dc = a.dc;
    const uint64_t total = a.frames * M;
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536ull * 16)
        blocks = 65536ull * 16;
    hipLaunchKernelGGL(split_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, sa);
    e = hipGetLastError();
    return e == hipSuccess ? OSTFT_OK : OSTFT_ERR_HIP;
}

}  // namespace ostft
