// engine.hpp -- internal interface between the C ABI (capi.hip) and the
// kernels (fused_small.hip, tables.hip, multipass.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ostft {

// output layouts (values of enum ostft_output in occastft_hip.h)
enum OutMode : int32_t {
    OUT_COMPAT_POWER = 0,
    OUT_HALF_COMPLEX = 1,
    OUT_MAGNITUDE = 2,
    OUT_POWER = 3,
    OUT_HALF_SPECTRUM = 4,
};

// One request, as the kernels see it.  Each frame of N = 2M real samples is
// one complex transform of M points (even samples real, odd samples imaginary).
struct FusedArgs {
    const float* in;      // signal, L floats
    uint64_t L;
    uint64_t hop;         // OMove
    uint64_t frames;      // frames to transform
    const float* win;     // N window coefficients, or nullptr for none
    const float2* tw;     // W_M^j, j in [0, M)
    const float2* tw_n;   // W_N^k, k in [0, M): the final radix-2 stage
    float* out;
    uint64_t out_limit;   // HALF_COMPLEX: complex values to write (OHalfSize)
    int32_t combine;      // 1: full FFT; 0: stop one radix-2 stage early (odd-exponent result)
    int32_t out_mode;     // OutMode
    int32_t dc;           // subtract the frame mean after windowing
    int32_t tile_frames;  // frames per workgroup tile of the single kernel (<= its G;
                          // fewer when hop > N, so a tile's span fits its staging LDS)
    uint64_t in_stride;   // strided batches: floats between signal y's input and y+1's
    uint64_t out_stride;  //                  and between their outputs (blockIdx.y = y)
    const void* mfma_b;   // matrix-core form: its DFT-matrix fragment table (mfma_real_dft.hip)
    int32_t plan;         // radix plan of the FP32 engine (0 or 1, see Cfg in stockham.hpp)
    int32_t reserved;
};

// transform of the fused kernel: Stockham FFT in FP32, or the split-precision
// MFMA DFT (windows 64, 256, 1024)
enum Transform : int { XF_FP32 = 0, XF_MFMA_FP16 = 1, XF_MFMA_BF16 = 2 };
bool mfma_supported(int log_m);
// complex sizes whose matrix-core form is the GEMM real DFT of
// mfma_real_dft.hip; its fragment table (per window table and format, built
// once with launch_mfma_table, mfma_table_bytes bytes) goes in FusedArgs::mfma_b
bool mfma_gemm_supported(int log_m);
size_t mfma_table_bytes(int log_m);
hipError_t launch_mfma_table(int log_m, int xf, const float* win, const float2* tw_n, void* table, hipStream_t s);
hipError_t launch_mfma_gemm(int log_m, int xf, const FusedArgs& a, uint32_t count, hipStream_t s);

// Single-kernel path: complex size M = 2^log_m, 5 <= log_m <= 14 (N = 64 .. 32768).
bool fused_supported(int log_m);
// frames per tile for a workgroup of G frames, window N and hop: G, fewer when
// the hop exceeds the window, so that the tile's span (tf - 1) hop + N fits
// the G N samples of the staging buffer
uint64_t fused_tile_frames(uint64_t G, uint64_t N, uint64_t hop);
// Variants of the FP32 engine (Cfg<LOG_M, V> in stockham.hpp), V = 4 shape +
// plan.  Plan bit 1: radix-8 passes with 8 values per thread (complex sizes
// 2^7..2^13), else radix 16 with 16; bit 0: the remainder radix first.  Shape:
// the workgroup of the batch kernels (0: 256 threads or one transform, 1: 512,
// 2: 128).  Only variants that differ from the others are built: a
// remainder-first plan where the size has a remainder radix, the radix-8
// plans in shape 0, shapes that change the workgroup.
constexpr int FUSED_VARIANTS = 12;
bool fused_variant_ok(int log_m, int variant);
// frames per workgroup (G) of variant `variant`, 0 if not built
int fused_variant_frames(int log_m, int variant);
bool fused_plan_ok(int log_m, int plan);
// `count` signals (grid.y), signal y at a.in + y in_stride / a.out + y out_stride
hipError_t launch_fused(int log_m, int xf, const FusedArgs& a, uint32_t count, hipStream_t s);

// ---- batches: many requests of one window size in one persistent launch ----
// A batch descriptor block lives in device memory (uploaded by the host before
// the launch); the kernel arguments only name where, so one captured launch
// sequence serves every batch.
struct BatchJob {
    FusedArgs a;
    uint64_t tile_begin;  // first tile of this request in its class
};
struct BatchClassHeader {
    uint64_t jobs_off;    // byte offset of BatchJob[njobs] from the block base
    uint64_t tiles_off;   // byte offset of uint32 tile_job[ntiles]
    uint32_t ntiles;      // tiles of G frames in the class (0: the launch exits at once)
    uint32_t counter;     // next unclaimed tile; uploaded as 0
};
struct BatchClassArgs {
    void* base;           // descriptor block
    uint64_t header_off;  // byte offset of this class's BatchClassHeader
};
struct BatchClassInfo {
    int frames_per_tile;    // G
    int threads;            // workgroup size
    int max_blocks_per_cu;  // occupancy limit of the persistent batch kernel
};
// persistent: a fixed grid claiming tiles from the class counter (workgroup
// shape 0 only); otherwise one workgroup per tile (grid = the class's tile count)
hipError_t launch_batch(int log_m, int variant, const BatchClassArgs& b, uint32_t grid, bool persistent,
                        hipStream_t s);
hipError_t batch_class_info(int log_m, int variant, BatchClassInfo* info);
// the kernel's host handle (graph nodes name it)
const void* batch_kernel_fn(int log_m, int variant, bool persistent);

// FP64 table generators (run once per table).
// tw[j] = exp(-2 pi i j / m) for j < count (count <= m)
hipError_t launch_twiddle_table(float2* tw, int64_t m, int64_t count, hipStream_t s);
// window formulas of OKL/kernel.okl:33-127 evaluated in FP64, rounded once
hipError_t launch_window_table(float* w, int64_t n, int kind, float sigma, hipStream_t s);

}  // namespace ostft
