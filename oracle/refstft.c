/*
 * refstft.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A C restatement of the reference's Serial/OpenMP STFT runner, stage by
 * stage, keeping the float/double promotions of the OCCA Serial translation
 * (StandAlone/cross_gpgpu/Serial/kernel/compiled.hpp) so that, built with
 * the same compiler flags, it reproduces the reference kernels bit for bit
 * for every window size the reference computes without undefined behaviour
 * (2^6 .. 2^16).  The OpenMP pragmas sit on the same outer loops as the
 * reference OpenMP runner (StandAlone/cross_gpgpu/OpenMP/kernel/compiled.hpp),
 * so this file doubles as the "port" CPU baseline.
 *
 * Reading this file: C's cos()/sin()/exp()/sqrt() are double functions, while
 * the reference is C++ where <math.h> overloads them for float.  Every place
 * the reference calls cos(float) is therefore written cosf() here.
 *
 * Not reproduced: the int32 overflow in segmentK for 2^17 .. 2^20 windows
 * (OKL/kernel.okl:141-146) -- that is undefined behaviour; the index is
 * computed in 64 bits here, which is what the reference intends.
 */
#include "oracle.h"

#include <errno.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* geometry                                                                 */
/* ------------------------------------------------------------------------ */

/* RunnerInterface.hpp:26-35 -- every operand is float */
int32_t oracle_toquot(float fullSize, float overlapRatio, float windowSize)
{
    if (overlapRatio == 0.0f) {
        float q = fullSize / windowSize + 1.0f;
        return (int32_t)q;
    } else {
        float q = (fullSize) / (windowSize * (1.0f - overlapRatio)) + 1.0f;
        return (int32_t)q;
    }
}

/* SerialImpl.cpp:36-43 */
void oracle_geometry(uint32_t fullSize, int32_t exp, float overlap,
                     int32_t* qtConst, uint32_t* oFullSize,
                     uint32_t* oHalfSize, uint32_t* oMove)
{
    const int32_t windowSize = 1 << exp;
    const int32_t qt = oracle_toquot((float)fullSize, overlap, (float)windowSize);
    /* int * int in the reference; done in 64 bits and truncated to the
     * reference's unsigned 32-bit result */
    const uint32_t ofull = (uint32_t)((int64_t)qt * (int64_t)windowSize);
    *qtConst = qt;
    *oFullSize = ofull;
    *oHalfSize = ofull / 2u;
    {
        float mv = (float)windowSize * (1.0f - overlap);
        *oMove = (uint32_t)mv;
    }
}

/* ------------------------------------------------------------------------ */
/* windows: OKL/kernel.okl:33-127, Serial compiled.hpp:20-88                */
/* ------------------------------------------------------------------------ */

static float hanning_w(int32_t index, int32_t windowSize)
{
    float angle = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize - 1));
    return (float)(0.5 * (1.0 - (double)cosf(angle)));
}

static float hamming_w(int32_t index, int32_t windowSize)
{
    float angle = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize - 1));
    return (float)(0.54 - (0.46 * (double)cosf(angle)));
}

static float blackman_w(int32_t index, int32_t windowSize)
{
    float fa = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize));
    float sa = (float)((double)fa * 2.0);
    return (float)(0.42 - 0.5 * (double)cosf(fa) + 0.08 * (double)cosf(sa));
}

static float nuttall_w(int32_t index, int32_t windowSize)
{
    float fa = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize));
    float sa = (float)((double)fa * 2.0);
    float ta = (float)((double)fa * 3.0);
    return (float)(0.355768 - 0.487396 * (double)cosf(fa) + 0.144232 * (double)cosf(sa)
                   - 0.012604 * (double)cosf(ta));
}

static float blackman_nuttall_w(int32_t index, int32_t windowSize)
{
    float fa = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize));
    float sa = (float)((double)fa * 2.0);
    float ta = (float)((double)fa * 3.0);
    return (float)(0.3635819 - 0.4891775 * (double)cosf(fa) + 0.1365995 * (double)cosf(sa)
                   - 0.0106411 * (double)cosf(ta));
}

static float blackman_harris_w(int32_t index, int32_t windowSize)
{
    float fa = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize));
    float sa = (float)((double)fa * 2.0);
    float ta = (float)((double)fa * 3.0);
    return (float)(0.35875 - 0.48829 * (double)cosf(fa) + 0.14128 * (double)cosf(sa)
                   - 0.01168 * (double)cosf(ta));
}

static float flattop_w(int32_t index, int32_t windowSize)
{
    float fa = (float)(2.0 * ORC_PI * (double)index / (double)(float)(windowSize));
    float sa = (float)((double)fa * 2.0);
    float ta = (float)((double)fa * 3.0);
    float qa = (float)((double)fa * 4.0);
    return (float)(0.21557895 - 0.41663158 * (double)cosf(fa) + 0.277263158 * (double)cosf(sa)
                   - 0.083578947 * (double)cosf(ta) + 0.006947368 * (double)cosf(qa));
}

static float gaussian_w(int32_t index, int32_t windowSize, float sigma)
{
    const int32_t hw = windowSize >> 1;
    float denom = (float)hw * sigma;           /* int * float -> float */
    float angle = ((float)(index - hw)) / denom;
    angle *= angle;
    angle = (float)((double)angle * -0.5);     /* angle *= -0.5 */
    return expf(angle);
}

float oracle_window_coef(int32_t kind, int32_t index, int32_t windowSize, float sigma)
{
    switch (kind) {
    case ORC_WIN_HANNING: return hanning_w(index, windowSize);
    case ORC_WIN_HAMMING: return hamming_w(index, windowSize);
    case ORC_WIN_BLACKMAN: return blackman_w(index, windowSize);
    case ORC_WIN_NUTTALL: return nuttall_w(index, windowSize);
    case ORC_WIN_BLACKMAN_NUTTALL: return blackman_nuttall_w(index, windowSize);
    case ORC_WIN_BLACKMAN_HARRIS: return blackman_harris_w(index, windowSize);
    case ORC_WIN_FLATTOP: return flattop_w(index, windowSize);
    case ORC_WIN_GAUSSIAN: return gaussian_w(index, windowSize, sigma);
    default: return 1.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* stage kernels                                                            */
/* ------------------------------------------------------------------------ */

/* Overlap_Common, OKL/EXPCommon.okl:7-27 (frame-major gather, zero past L) */
void oracle_overlap(const float* in, uint32_t oFullSize, uint32_t fullSize,
                    uint32_t exp, uint32_t oMove, float* outReal)
{
    const int64_t total = (int64_t)oFullSize;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t o = 0; o < total; ++o) {
        const uint32_t idx = (uint32_t)o;
        const uint32_t win = idx >> exp;
        const uint32_t local = idx & ((1u << exp) - 1u);
        const uint32_t origin = win * oMove + local;
        const uint32_t inside = origin < fullSize;
        /* the reference reads in[0] * 0 when outside; an empty input has no
         * in[0], so read nothing then */
        const float v = (inside || fullSize > 0) ? in[origin * inside] : 0.0f;
        outReal[idx] = v * (float)inside;
    }
}

/* Window_* kernels, OKL/EXPCommon.okl:29-148: data[i] *= w(i mod N) */
void oracle_apply_window(float* data, uint32_t oFullSize, uint32_t windowSize,
                         int32_t kind, float sigma)
{
    const int64_t total = (int64_t)oFullSize;
    if (kind == ORC_WIN_NONE)
        return;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t o = 0; o < total; ++o) {
        const uint32_t g = (uint32_t)o;
        data[g] *= oracle_window_coef(kind, (int32_t)(g & (windowSize - 1u)),
                                      (int32_t)windowSize, sigma);
    }
}

/* DCRemove_Common, OKL/EXPCommon.okl:151-192: 64 lane-strided partial sums,
 * then a pairwise tree 32,16,..,1, then subtract sum/N */
void oracle_remove_dc(float* data, uint32_t oFullSize, uint32_t windowSize)
{
    const int64_t frames = (int64_t)(oFullSize / windowSize);
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t f = 0; f < frames; ++f) {
        float added[128];
        float* x = data + (size_t)f * windowSize;
        unsigned i, w, seg;
        for (i = 0; i < 64; ++i)
            added[i] = 0.0f;
        for (w = 0; w < windowSize; w += 64) {
            for (i = 0; i < 64; ++i)
                added[i + 64] = x[w + i];
            for (i = 0; i < 64; ++i)
                added[i] += added[i + 64];
        }
        for (seg = 32; seg > 0; seg >>= 1) {
            for (i = 0; i < 64; ++i) {
                const unsigned inSeg = i < seg;
                const float l = added[i];
                const float r = added[i + seg];
                added[i] = (l + r) * (float)inSeg;
            }
        }
        for (w = 0; w < windowSize; w += 64)
            for (i = 0; i < 64; ++i)
                x[w + i] -= (added[0] / (float)windowSize);
    }
}

/* twiddle + segmentK, OKL/kernel.okl:141-156 */
static inline void ref_twiddle(int64_t k, int64_t windowSize, float* re, float* im)
{
    float angle = (float)(-2.0 * ORC_PI * (double)((float)k / (float)windowSize));
    *re = cosf(angle);
    *im = sinf(angle);
}

/* One radix-2 Stockham stage over all frames: StockHamCommon,
 * OKL/EXPCommon.okl:194-245.  The shared-memory kernels Stockhoptimized6..11
 * (OKL/EXP6.okl .. OKL/EXP11.okl, macros OKL/kernel.okl:223-352) perform the
 * same per-element operations stage by stage, so one routine serves both. */
void oracle_stockham_stage(const float* inRe, const float* inIm,
                           float* outRe, float* outIm,
                           uint32_t hWindowSize, uint32_t stage,
                           uint32_t oHalfSize, uint32_t exp)
{
    const int64_t total = (int64_t)oHalfSize;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t o = 0; o < total; ++o) {
        const uint32_t oidx0 = (uint32_t)o;
        const uint32_t fws = hWindowSize << 1;
        const uint32_t gItr = oidx0 >> (exp - 1u);
        const uint32_t gIdx = oidx0 & (hWindowSize - 1u);
        const uint32_t oidx = gItr * fws + gIdx;
        const float lr = inRe[oidx];
        const float li = inIm[oidx];
        const float rr = inRe[oidx + hWindowSize];
        const float ri = inIm[oidx + hWindowSize];
        const uint32_t seg = 1u << stage;
        const uint32_t segItr = gIdx >> stage;
        const uint32_t segIdx = gIdx & (seg - 1u);
        const uint32_t ls = segItr * (seg << 1) + segIdx + gItr * fws;
        const uint32_t rs = ls + seg;
        /* segmentK(lsave, seg, H) = ((lsave % seg) * H) / seg, in 64 bits */
        const int64_t k = ((int64_t)(oidx % seg) * (int64_t)hWindowSize) / (int64_t)seg;
        float twr, twi, rt, it;
        ref_twiddle(k, (int64_t)fws, &twr, &twi);
        rt = (rr * twr) - (ri * twi); /* RMult */
        it = (rr * twi) + (ri * twr); /* IMult */
        outRe[ls] = lr + rt;
        outIm[ls] = li + it;
        outRe[rs] = lr - rt;
        outIm[rs] = li - it;
    }
}

/* toPower, OKL/kernel.okl:178-194 */
void oracle_to_power(float* out, const float* re, const float* im, uint32_t oFullSize)
{
    const int64_t total = (int64_t)oFullSize;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t o = 0; o < total; ++o) {
        const float r = re[o];
        const float i = im[o];
        out[o] = sqrtf(r * r + i * i);
    }
}

/* toHalfComplexFormat, OKL/kernel.okl:196-218: the read index equals the
 * write index, so this copies the first OHalfSize complex values */
void oracle_half_complex(float* out, const float* re, const float* im,
                         uint32_t oHalfSize, int32_t exp)
{
    const int64_t total = (int64_t)oHalfSize;
    const uint32_t mradix = (uint32_t)exp - 1u;
    const uint32_t hw = 1u << mradix;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t o = 0; o < total; ++o) {
        const uint32_t gid = (uint32_t)o;
        const uint32_t rd = (gid >> mradix) * hw + gid % hw;
        out[2 * (size_t)gid] = re[rd];
        out[2 * (size_t)gid + 1] = im[rd];
    }
}

/* ------------------------------------------------------------------------ */
/* option dispatch: RunnerCommon.cpp:216-248 and :309-314                   */
/* ------------------------------------------------------------------------ */

int32_t oracle_parse_options(const char* options, oracle_opts* o)
{
    static const char* const keys[7] = {
        "--hanning_window", "--hamming_window", "--blackman_window",
        "--nuttall_window", "--blackman_nuttall_window",
        "--blackman_harris_window", "--flattop_window"};
    int k;
    memset(o, 0, sizeof(*o));
    o->window = ORC_WIN_NONE;
    for (k = 0; k < 7; ++k) {
        if (strstr(options, keys[k])) {
            o->window = k + 1;
            break;
        }
    }
    if (o->window == ORC_WIN_NONE) {
        const char* g = strstr(options, "--gaussian_window=");
        if (g) {
            const char* s = strstr(options, "<<sigma");
            if (s) {
                /* substr(P1, P2 - P1) with P1 = pos + 19 (one past the '=') */
                size_t len = strlen(options);
                size_t p1 = (size_t)(g - options) + 19u;
                size_t p2 = (size_t)(s - options);
                size_t cnt = p2 - p1; /* unsigned wrap as in std::string */
                char buf[256];
                char* end = NULL;
                float sigma;
                if (p1 > len)
                    return 1; /* std::out_of_range from substr */
                if (cnt > len - p1)
                    cnt = len - p1;
                if (cnt >= sizeof(buf))
                    cnt = sizeof(buf) - 1;
                memcpy(buf, options + p1, cnt);
                buf[cnt] = 0;
                errno = 0;
                sigma = strtof(buf, &end); /* std::stof */
                if (end == buf || errno == ERANGE)
                    return 1;
                if (sigma > 0.0f) {
                    o->window = ORC_WIN_GAUSSIAN;
                    o->sigma = sigma;
                }
            }
        }
    }
    o->remove_dc = strstr(options, "--remove_dc") != NULL;
    o->half_complex = strstr(options, "--half_complex_return") != NULL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* the whole Serial runner                                                  */
/* ------------------------------------------------------------------------ */

int32_t oracle_pipeline(const float* in, uint32_t fullSize, int32_t exp,
                        float overlap, const char* options, float* out)
{
    int32_t qt;
    uint32_t ofull, ohalf, omove, N, H;
    oracle_opts o;
    float *re, *im, *sre, *sim, *res;
    int32_t s;

    if (exp < 1 || exp > 30)
        return 2;
    oracle_geometry(fullSize, exp, overlap, &qt, &ofull, &ohalf, &omove);
    if ((int64_t)qt * ((int64_t)1 << exp) >= ((int64_t)1 << 32))
        return 3;
    N = 1u << exp;
    H = N >> 1;
    memset(out, 0, (size_t)ofull * sizeof(float)); /* std::vector outMem(OFullSize) */
    if (oracle_parse_options(options, &o) != 0)
        return 1;
    if (exp < 6)
        return 2; /* RunnerCommon.cpp:250-253 */

    re = (float*)calloc(ofull, sizeof(float));
    im = (float*)calloc(ofull, sizeof(float));
    sre = (float*)calloc(ofull, sizeof(float));
    sim = (float*)calloc(ofull, sizeof(float));
    if (!re || !im || !sre || !sim) {
        free(re); free(im); free(sre); free(sim);
        return 3;
    }

    oracle_overlap(in, ofull, fullSize, (uint32_t)exp, omove, re);
    oracle_apply_window(re, ofull, N, o.window, o.sigma);
    if (o.remove_dc)
        oracle_remove_dc(re, ofull, N);

    /* stage s reads (s even ? re : sre) and writes the other buffer
     * (Serial/src/functionImpl.cpp:207-262).  Default_Pipeline then always
     * reads `re` (RunnerCommon.cpp:254-314): the result for even exp, and
     * the output of stage exp-2 for odd exp >= 12.  The shared-memory kernels
     * for exp <= 11 write their final stage to `re` directly. */
    for (s = 0; s < exp; ++s) {
        if (s % 2 == 0)
            oracle_stockham_stage(re, im, sre, sim, H, (uint32_t)s, ohalf, (uint32_t)exp);
        else
            oracle_stockham_stage(sre, sim, re, im, H, (uint32_t)s, ohalf, (uint32_t)exp);
    }
    res = re;
    if (exp <= 11 && (exp % 2) == 1) {
        /* Stockhoptimized writes its last stage to Rout/Iout */
        memcpy(re, sre, (size_t)ofull * sizeof(float));
        memcpy(im, sim, (size_t)ofull * sizeof(float));
    }
    if (o.half_complex)
        oracle_half_complex(out, res, im, ohalf, exp);
    else
        oracle_to_power(out, res, res, ofull); /* Serial passes Real twice */

    free(re); free(im); free(sre); free(sim);
    return 0;
}
