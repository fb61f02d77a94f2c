/*
 * fp64.c -- TEST INFRASTRUCTURE ONLY (the FP64 truth oracle).
 *
 * An FP64 STFT with the reference pipeline's stage order
 * (StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:206-314: overlap, window,
 * DC removal of the windowed frame, FFT, output) but exact arithmetic:
 * window formulas of OKL/kernel.okl:33-127 evaluated in double, the mean
 * subtracted in double, and a double FFT.  It pins the float32 restatement
 * and the GPU engine from the mathematics rather than from the reference's
 * rounding:
 *   - oracle_dft64 is the DFT definition, O(n^2);
 *   - oracle_fft64 is an iterative radix-2 FFT checked against it;
 *   - the odd-window quirk (penultimate Stockham stage) is stated in closed
 *     form: the radix-2 Stockham buffer after exp-1 of exp stages holds
 *     [DFT_{N/2}(x[0::2]), DFT_{N/2}(x[1::2])] (StandAlone/cross_gpgpu/Serial/
 *     src/functionImpl.cpp:207-262 with the by-value result pointers of
 *     RunnerCommon.cpp:254-305).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

void oracle_dft64(const double* re, const double* im, int64_t n,
                  double* outRe, double* outIm)
{
    int64_t k, j;
    for (k = 0; k < n; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (j = 0; j < n; ++j) {
            /* reduce j*k mod n before the angle: exact for n < 2^31 */
            const int64_t m = (j * k) % n;
            const long double a = -2.0L * (long double)ORC_PI * (long double)m / (long double)n;
            const long double c = cosl(a), s = sinl(a);
            sr += (long double)re[j] * c - (long double)im[j] * s;
            si += (long double)re[j] * s + (long double)im[j] * c;
        }
        outRe[k] = (double)sr;
        outIm[k] = (double)si;
    }
}

void oracle_fft64(double* re, double* im, int64_t n)
{
    int64_t i, j, len;
    /* bit reversal */
    for (i = 1, j = 0; i < n; ++i) {
        int64_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (len = 2; len <= n; len <<= 1) {
        const int64_t half = len >> 1;
        int64_t base, k;
        for (k = 0; k < half; ++k) {
            /* twiddle from the exact angle each time (no recurrence) */
            const double a = -2.0 * ORC_PI * (double)k / (double)len;
            const double wr = cos(a), wi = sin(a);
            for (base = 0; base < n; base += len) {
                const int64_t p = base + k, q = p + half;
                const double tr = re[q] * wr - im[q] * wi;
                const double ti = re[q] * wi + im[q] * wr;
                re[q] = re[p] - tr;
                im[q] = im[p] - ti;
                re[p] += tr;
                im[p] += ti;
            }
        }
    }
}

double oracle_window_coef64(int32_t kind, int64_t index, int64_t windowSize, double sigma)
{
    const double n = (double)index;
    const double N = (double)windowSize;
    const double t = 2.0 * ORC_PI * n / N;
    switch (kind) {
    case ORC_WIN_HANNING: return 0.5 * (1.0 - cos(2.0 * ORC_PI * n / (N - 1.0)));
    case ORC_WIN_HAMMING: return 0.54 - 0.46 * cos(2.0 * ORC_PI * n / (N - 1.0));
    case ORC_WIN_BLACKMAN: return 0.42 - 0.5 * cos(t) + 0.08 * cos(2.0 * t);
    case ORC_WIN_NUTTALL:
        return 0.355768 - 0.487396 * cos(t) + 0.144232 * cos(2.0 * t) - 0.012604 * cos(3.0 * t);
    case ORC_WIN_BLACKMAN_NUTTALL:
        return 0.3635819 - 0.4891775 * cos(t) + 0.1365995 * cos(2.0 * t) - 0.0106411 * cos(3.0 * t);
    case ORC_WIN_BLACKMAN_HARRIS:
        return 0.35875 - 0.48829 * cos(t) + 0.14128 * cos(2.0 * t) - 0.01168 * cos(3.0 * t);
    case ORC_WIN_FLATTOP:
        return 0.21557895 - 0.41663158 * cos(t) + 0.277263158 * cos(2.0 * t)
               - 0.083578947 * cos(3.0 * t) + 0.006947368 * cos(4.0 * t);
    case ORC_WIN_GAUSSIAN: {
        const double h = (double)(windowSize >> 1);
        const double a = (n - h) / (h * sigma);
        return exp(-0.5 * a * a);
    }
    default: return 1.0;
    }
}

int64_t oracle_stft64(const float* in, uint64_t fullSize, int32_t exp,
                      uint64_t hop, uint64_t frames, int32_t window, double sigma,
                      int32_t remove_dc, int32_t out_mode, int32_t quirk_odd,
                      double* out)
{
    const int64_t N = (int64_t)1 << exp;
    const int64_t M = N / 2;
    const int use_quirk = quirk_odd && exp >= 13 && (exp % 2) == 1
                          && (out_mode == 0 || out_mode == 1 || out_mode == 2);
    double* xr = (double*)malloc(sizeof(double) * (size_t)N);
    double* xi = (double*)malloc(sizeof(double) * (size_t)N);
    double* er = (double*)malloc(sizeof(double) * (size_t)N);
    double* ei = (double*)malloc(sizeof(double) * (size_t)N);
    double* w = (double*)malloc(sizeof(double) * (size_t)N);
    const uint64_t ofull = frames * (uint64_t)N;
    const uint64_t ohalf = ofull / 2;
    int64_t written = 0;
    uint64_t f;
    int64_t n, k;

    for (n = 0; n < N; ++n)
        w[n] = oracle_window_coef64(window, n, N, sigma);

    for (f = 0; f < frames; ++f) {
        double mean = 0.0;
        /* half-complex compat only needs frames whose spectra land in the
         * first OHalfSize complex values */
        if (out_mode == 1 && f * (uint64_t)N >= ohalf)
            break;
        for (n = 0; n < N; ++n) {
            const uint64_t src = f * hop + (uint64_t)n;
            const double v = src < fullSize ? (double)in[src] : 0.0;
            xr[n] = v * w[n];
            xi[n] = 0.0;
        }
        if (remove_dc) {
            for (n = 0; n < N; ++n)
                mean += xr[n];
            mean /= (double)N;
            for (n = 0; n < N; ++n)
                xr[n] -= mean;
        }
        if (use_quirk) {
            /* [DFT_{N/2}(even), DFT_{N/2}(odd)] */
            for (n = 0; n < M; ++n) {
                er[n] = xr[2 * n];
                ei[n] = 0.0;
                er[M + n] = xr[2 * n + 1];
                ei[M + n] = 0.0;
            }
            oracle_fft64(er, ei, M);
            oracle_fft64(er + M, ei + M, M);
        } else {
            memcpy(er, xr, sizeof(double) * (size_t)N);
            memcpy(ei, xi, sizeof(double) * (size_t)N);
            oracle_fft64(er, ei, N);
        }
        switch (out_mode) {
        case 0:
            for (k = 0; k < N; ++k)
                out[f * N + k] = sqrt(2.0) * fabs(er[k]);
            written += N;
            break;
        case 1:
            for (k = 0; k < N; ++k) {
                const uint64_t g = f * (uint64_t)N + (uint64_t)k;
                if (g >= ohalf)
                    break;
                out[2 * g] = er[k];
                out[2 * g + 1] = ei[k];
                written += 2;
            }
            break;
        case 2:
            for (k = 0; k < N; ++k)
                out[f * N + k] = sqrt(er[k] * er[k] + ei[k] * ei[k]);
            written += N;
            break;
        case 3:
            for (k = 0; k <= M; ++k)
                out[f * (M + 1) + k] = er[k] * er[k] + ei[k] * ei[k];
            written += M + 1;
            break;
        case 4:
            for (k = 0; k <= M; ++k) {
                out[f * (N + 2) + 2 * k] = er[k];
                out[f * (N + 2) + 2 * k + 1] = ei[k];
            }
            written += N + 2;
            break;
        default:
            break;
        }
    }
    if (out_mode == 1) {
        /* the rest of the OFullSize output buffer stays zero */
        for (; written < (int64_t)ofull; ++written)
            out[written] = 0.0;
    }
    free(xr); free(xi); free(er); free(ei); free(w);
    return written;
}
