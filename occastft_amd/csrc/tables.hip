// tables.hip -- FP64 generators for the twiddle table and the exact window table.
//
// FP64 is used on the GPU only here.  Tables are built once per (size, kind)
// and kept resident by the context.
#include "engine.hpp"

namespace ostft {

// tw[j] = exp(-2 pi i j / M) for j < count: sincospi of the exact rational -2j/M.
__global__ void twiddle_kernel(float2* tw, int64_t m, int64_t count)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count)
        return;
    double s, c;
    sincospi(-2.0 * (double)j / (double)m, &s, &c);
    tw[j] = make_float2((float)c, (float)s);
}

hipError_t launch_twiddle_table(float2* tw, int64_t m, int64_t count, hipStream_t s)
{
    const int64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(twiddle_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, tw, m, count);
    return hipGetLastError();
}

constexpr double kPi = 3.14159265358979323846;

// The same windows evaluated exactly in FP64 and rounded once.
__device__ float window_exact(int kind, int64_t n, int64_t N, float sigma)
{
    const double t = 2.0 * (double)n / (double)N;  // angle / pi
    switch (kind) {
    case 1: return (float)(0.5 * (1.0 - cospi(2.0 * (double)n / (double)(N - 1))));
    case 2: return (float)(0.54 - 0.46 * cospi(2.0 * (double)n / (double)(N - 1)));
    case 3: return (float)(0.42 - 0.5 * cospi(t) + 0.08 * cospi(2.0 * t));
    case 4: return (float)(0.355768 - 0.487396 * cospi(t) + 0.144232 * cospi(2.0 * t) - 0.012604 * cospi(3.0 * t));
    case 5:
        return (float)(0.3635819 - 0.4891775 * cospi(t) + 0.1365995 * cospi(2.0 * t) - 0.0106411 * cospi(3.0 * t));
    case 6: return (float)(0.35875 - 0.48829 * cospi(t) + 0.14128 * cospi(2.0 * t) - 0.01168 * cospi(3.0 * t));
    case 7:
        return (float)(0.21557895 - 0.41663158 * cospi(t) + 0.277263158 * cospi(2.0 * t)
                       - 0.083578947 * cospi(3.0 * t) + 0.006947368 * cospi(4.0 * t));
    case 8: {
        const double h = (double)(N >> 1);
        const double a = ((double)n - h) / (h * (double)sigma);
        return (float)exp(-0.5 * a * a);
    }
    default:
        return 1.0f;
    }
}

__global__ void window_kernel(float* w, int64_t N, int kind, float sigma)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N)
        return;
    w[n] = window_exact(kind, n, N, sigma);
}

hipError_t launch_window_table(float* w, int64_t n, int kind, float sigma, hipStream_t s)
{
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(window_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, w, n, kind, sigma);
    return hipGetLastError();
}

}  // namespace ostft
