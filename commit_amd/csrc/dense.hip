// y = A x on the FP64 matrix cores, for dense x: the IC part of a tile of 16
// voxels is one dense product
//     Y_tile[16 voxels][nS] = C[16 voxels][ndirs] * LUT[ndirs][nS]
// with C[v][o] = sum of l * x[f] * eval over the voxel's segments of
// orientation o (setup_operator.py:62-103 regrouped by orientation).  The
// work grows from nnz*nS to nVOX*ndirs*nS multiply-adds (x3.6 at 139 segments
// per voxel and 500 orientations) but the LUT is read once per 16 voxels
// instead of once per segment, which moves the bound from the L2->CU stream of
// LUT rows to the FP64 matrix-core rate.  Products and sums are FP64 (the
// float32 LUT widens exactly), so results agree with the per-segment kernel
// to FP64 rounding.  EC and ISO entries (a few per voxel) are added per
// element after the product.
//
// v_mfma_f64_16x16x4f64: D[16x16] += A[16x4] * B[4x16]; lane l supplies
// A[l%16][l/16] and B[l/16][l%16] and receives D[4r + l/16][l%16] in its
// register r = 0..3 (rows interleaved over the lane groups, unlike the f32
// form; established with one-hot operands).
//   A = C (row = voxel of the tile, k = orientation), read from LDS where C is
//       stored orientation-major (Ct[o][16]): the 64 lanes of one read cover
//       512 contiguous bytes;
//   B = LUT (k = orientation, column = sample), read from the padded LUT rows.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/commit_hip.h"
#include "common.h"
#include "plan.h"

using commit_internal::fail;

namespace {

using commit_internal::Comp;
using commit_internal::DevDict;

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTV = 16;                    // voxels per tile (MFMA M)
constexpr int kMaxDirs = 512;              // Ct[ndirs][16] doubles in LDS: 64 KB
constexpr int kMaxNTW = 5;                 // sample tiles per wavefront (nS <= 320)

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T>
struct DenseOut {
    T *y;
    const T *yobs;        // residual form (null: plain product)
    const T *sqrtw;
    double *rpart;        // [nVOX] per-voxel ||r||^2, or null
};

// Wavefront wb: the coefficients of voxels v0+4wb .. v0+4wb+3 (one
// contiguous entry range) into columns 4wb..4wb+3 of a C buffer: its columns
// are zeroed first, then each entry adds l * x * eval at [o][voxel] with an
// LDS atomic (lanes of one instruction in the LDS's fixed lane order, chunks
// in entry order).
template <typename T>
__device__ __forceinline__ void dense_build(const DevDict &d, const T *__restrict__ xb, int v0, int wb, int nrow,
                                            double *Cb, int lane)
{
    for (int k = lane; k < nrow * 4; k += kWave) Cb[(k >> 2) * kTV + 4 * wb + (k & 3)] = 0.0;
    const int va = min(v0 + 4 * wb, d.nVOX), vz = min(v0 + 4 * wb + 4, d.nVOX);
    if (va >= vz) return;
    const uint32_t b = d.ic.ptr[va], e = d.ic.ptr[vz];
    uint32_t p1 = va + 1 < vz ? d.ic.ptr[va + 1] : e;
    uint32_t p2 = va + 2 < vz ? d.ic.ptr[va + 2] : e;
    uint32_t p3 = va + 3 < vz ? d.ic.ptr[va + 3] : e;
    const int vb = 4 * wb;
    // four chunks of 64 entries per round, all loads issued before any use
    // (entries past e read entry b and are dropped)
    for (uint32_t base = b; base < e; base += 4 * kWave) {
        int o[4];
        uint32_t col[4];
        float len[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t k = base + u * kWave + lane;
            ok[u] = k < e;
            const uint32_t ks = ok[u] ? k : b;
            o[u] = d.ic.o[ks];
            col[u] = d.ic.col[ks];
            len[u] = d.ic.len[ks];
        }
        T xv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = xb[col[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t k = base + u * kWave + lane;
            const double c = ok[u] ? (double)xv[u] * (double)len[u] : 0.0;   // (setup_operator.py:63-96)
            const int vl = vb + (k >= p1) + (k >= p2) + (k >= p3);
            if (c != 0.0) atomicAdd(&Cb[o[u] * kTV + vl], c);
        }
    }
}

// Matrix-core wavefront w (< 4): acc[u] += C * LUT_i for sample tiles
// w + 4u.  Branch-free (exact waits): orientations past ndirs meet zero C
// entries (rows clamped to the LUT, whose last RF has three zero rows after
// it); tiles past the last sample tile read column q of tile 0 and are
// discarded.
__device__ __forceinline__ void dense_mma(const DevDict &d, const float *__restrict__ L, const double *Cb, int w,
                                          int g, int q, int nK, int nNT, d4 (&acc)[kMaxNTW])
{
    int soff[kMaxNTW];
#pragma unroll
    for (int u = 0; u < kMaxNTW; u++) {
        const int j = w + kWavesPerBlock * u;
        soff[u] = (j < nNT ? 16 * j : 0) + q;
    }
    const int rmax = d.ndirs + 2;
    float bq[4][kMaxNTW];
    auto fetch = [&](float(&b)[kMaxNTW], int t) {
        const float *row = L + (size_t)min(4 * t + g, rmax) * d.nSp;
#pragma unroll
        for (int u = 0; u < kMaxNTW; u++) b[u] = row[soff[u]];
    };
    fetch(bq[0], 0);
    fetch(bq[1], 1);
    fetch(bq[2], 2);
    for (int t0 = 0; t0 < nK; t0 += 4) {
#pragma unroll
        for (int tt = 0; tt < 4; tt++) {
            const int t = t0 + tt;
            // keep this step's loads after the previous step's products, so
            // that each row arrives three steps (15 products) after its request
            __builtin_amdgcn_sched_barrier(0);
            fetch(bq[(tt + 3) & 3], min(t + 3, nK - 1));
            const double a = Cb[(4 * t + g) * kTV + q];         // A[q][g]: voxel q, orientation 4t+g
#pragma unroll
            for (int u = 0; u < kMaxNTW; u++) acc[u] = mfma64(a, (double)bq[tt][u], acc[u]);
        }
    }
}

// EC / ISO entries, residual form, stores of one tile; lane (g, q) of sample
// tile j holds voxels 4r+g (register r) at sample 16j + q.  rs: this
// wavefront's per-voxel ||r||^2 partials.
template <typename T>
__device__ __forceinline__ void dense_store(const DevDict &d, const T *__restrict__ x, const DenseOut<T> &out,
                                            int v0, int w, int g, int q, int nNT, const d4 (&acc)[kMaxNTW],
                                            double *rs)
{
    double sq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kMaxNTW; u++) {
        const int j = w + kWavesPerBlock * u;
        if (j >= nNT) break;
        const int s = 16 * j + q;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int v = v0 + 4 * r + g;
            if (v >= d.nVOX || s >= d.nS) continue;
            double yv = acc[u][r];
            for (int i = 0; i < d.nEC; i++) {                 // (setup_operator.py:111-165)
                const float *L = d.wmh + (size_t)i * d.ndirs * d.nSp;
                const T *xe = x + (size_t)d.nIC * d.nF + (size_t)i * d.nE;
                for (uint32_t k = d.ec.ptr[v]; k < d.ec.ptr[v + 1]; k++)
                    yv = fma((double)xe[d.ec.col[k]], (double)L[(size_t)d.ec.o[k] * d.nSp + s], yv);
            }
            for (int i = 0; i < d.nISO; i++) {                // (setup_operator.py:167-218)
                const float *L = d.isol + (size_t)i * d.nSp;
                const T *xs = x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE + (size_t)i * d.nI;
                for (uint32_t k = d.iso.ptr[v]; k < d.iso.ptr[v + 1]; k++)
                    yv = fma((double)xs[d.iso.col[k]], (double)L[s], yv);
            }
            const size_t row = (size_t)v * d.nS + s;
            if (out.yobs) yv -= (double)out.yobs[row];        // (solvers.py:308)
            if (out.sqrtw) yv *= (double)out.sqrtw[row];      // (solvers.py:306)
            if (out.y) __builtin_nontemporal_store((T)yv, out.y + row);
            sq[r] = fma(yv, yv, sq[r]);
        }
    }
    if (out.rpart) {
#pragma unroll
        for (int r = 0; r < 4; r++) {                         // the 16 sample lanes of the group
            double t = sq[r];
            t += __shfl_xor(t, 1);
            t += __shfl_xor(t, 2);
            t += __shfl_xor(t, 4);
            t += __shfl_xor(t, 8);
            if (q == 0) rs[4 * r + g] = t;
        }
    }
}

// One block (4 wavefronts) per tile of 16 voxels, two blocks per CU (64 KB
// of LDS each): while one block builds its C tile the other's products run.
// Per RF: wavefront w builds voxels 4w..4w+3, barrier, the four wavefronts
// multiply, barrier.  Per-voxel ||r||^2 partials are summed over the four
// wavefronts in order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_A_dense(DevDict d, const T *__restrict__ x, const T *__restrict__ xic,
                                                    DenseOut<T> out, int ntiles)
{
    __shared__ __align__(16) double Ct[kMaxDirs * kTV];
    __shared__ double rsum[kWavesPerBlock][kTV];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = lane >> 4, q = lane & 15;
    const int nK = (((d.ndirs + 3) >> 2) + 3) & ~3;       // K-steps of 4 orientations, a multiple of 4
    const int nNT = (d.nS + 15) >> 4;
    const size_t lut_rf = (size_t)d.ndirs * d.nSp;
    const int v0 = (int)blockIdx.x * kTV;
    if ((int)blockIdx.x >= ntiles) return;
    d4 acc[kMaxNTW];
#pragma unroll
    for (int u = 0; u < kMaxNTW; u++) acc[u] = d4{0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < d.nIC; i++) {
        dense_build<T>(d, xic + (size_t)i * d.nF, v0, w, 4 * nK, Ct, lane);
        __syncthreads();
        dense_mma(d, d.wmr + (size_t)i * lut_rf, Ct, w, g, q, nK, nNT, acc);
        __syncthreads();
    }
    dense_store<T>(d, x, out, v0, w, g, q, nNT, acc, rsum[w]);
    if (out.rpart) {
        __syncthreads();
        if (threadIdx.x < kTV) {
            const int v = v0 + (int)threadIdx.x;
            if (v < d.nVOX) {
                double t = 0.0;
                for (int ww = 0; ww < kWavesPerBlock; ww++) t += rsum[ww][threadIdx.x];
                out.rpart[v] = t;
            }
        }
    }
}

}  // namespace

namespace commit_internal {

// Launch of the dense-tile A x; returns COMMIT_HIP_EUNSUPPORTED when the plan
// is outside what the kernel is built for (the caller then uses the
// per-segment kernel).
template <typename T>
int launch_A_dense(commit_hip_plan *p, const T *x, const T *xic, T *y, const T *yobs, const T *sqrtw, double *rpart,
                   hipStream_t st)
{
    const DevDict &d = p->d;
    if (p->nolut || (((d.ndirs + 3) / 4 + 3) & ~3) * 4 > kMaxDirs || (d.nS + 15) / 16 > kMaxNTW * kWavesPerBlock)
        return COMMIT_HIP_EUNSUPPORTED;
    const int ntiles = (d.nVOX + kTV - 1) / kTV;
    if (ntiles == 0) return 0;
    DenseOut<T> o{y, yobs, sqrtw, rpart};
    hipLaunchKernelGGL(k_A_dense<T>, dim3((unsigned)ntiles), dim3(kBlock), 0, st, d, x, xic, o, ntiles);
    HIP_TRY(hipGetLastError());
    return 0;
}

template int launch_A_dense<double>(commit_hip_plan *, const double *, const double *, double *, const double *,
                                    const double *, double *, hipStream_t);
template int launch_A_dense<float>(commit_hip_plan *, const float *, const float *, float *, const float *,
                                   const float *, double *, hipStream_t);

}  // namespace commit_internal
