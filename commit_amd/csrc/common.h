// Shared by the translation units of libcommit_hip.so: error reporting and
// the deterministic block reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace commit_internal {
extern thread_local std::string g_err;
int fail(int code, const char *fmt, ...);
}  // namespace commit_internal

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return commit_internal::fail(COMMIT_HIP_EDEVICE, "%s failed: %s", #expr,            \
                                         hipGetErrorString(e_));                                \
    } while (0)

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
// Deterministic sums: kSumBlocks blocks each reduce a fixed contiguous chunk in
// a fixed order, then one block reduces the partials.
constexpr int kSumBlocks = 256;
// Blocks of the residual combination (commit_hip_residual_combine), each
// summing a fixed range of rows; their partials are reduced by one block.
constexpr int kCombBlocks = 2048;

// Sum over a 256-thread block; the result is valid in thread 0.  Fixed order:
// xor-butterfly inside each wavefront, then the 4 wavefront sums in order.
__device__ __forceinline__ double block_sum_256(double v, double *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / kWave; w++) t += sh[w];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kBlock) void k_sum_stage1(const double *__restrict__ in, size_t n,
                                                       double *__restrict__ part)
{
    __shared__ double sh[kBlock / kWave];
    const size_t chunk = (n + kSumBlocks - 1) / kSumBlocks;
    const size_t b = (size_t)blockIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
    double v = 0.0;
    for (size_t i = b + threadIdx.x; i < e; i += kBlock) v += in[i];
    v = block_sum_256(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_sum_stage2(const double *__restrict__ part, int np,
                                                       double *__restrict__ out)
{
    __shared__ double sh[kBlock / kWave];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += kBlock) v += part[i];
    v = block_sum_256(v, sh);
    if (threadIdx.x == 0) *out = v;
}

}  // namespace
