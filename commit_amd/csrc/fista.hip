// FISTA vector kernels of libcommit_hip.so: the proximal operators and the
// deterministic reductions of one FISTA step (solvers.py:244-391,
// proximals.pyx:13-186).  The operator products are in commit_hip.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/commit_hip.h"
#include "common.h"

using commit_internal::fail;

namespace {

struct ProxArgs {
    int64_t start[3], end[3];
    int reg[3], nn[3], weighted[3];
    double thr[3];       // lambda * scale
    const double *w;
};

__device__ __forceinline__ double soft(double v, double t)
{
    // proximals.pyx:25-39 / 42-56, branch for branch
    if (v > t) return v - t;
    if (v < -t) return v + t;
    return 0.0;
}

// x = xin - step*grad, then soft-thresholding and (for compartments without a
// group penalty) non-negativity, element by element.
// xin may alias x (in-place prox): each element is read and written by one thread.
// T: storage type of the vectors (double, or float for the FP32 mode); the
// arithmetic is FP64 and the result is rounded once on store.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prox_elem(ProxArgs P, int64_t n, const T *xin,
                                                      const T *__restrict__ grad, double step, T *x)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = (double)xin[i];
        if (grad) v = v - step * (double)grad[i];                         // (solvers.py:299)
        for (int c = 0; c < 3; c++) {
            if (i < P.start[c] || i >= P.end[c]) continue;
            const int r = P.reg[c];
            if (r == COMMIT_HIP_REG_LASSO || r == COMMIT_HIP_REG_SPARSE_GROUP_LASSO)
                v = soft(v, P.weighted[c] ? P.w[i] * P.thr[c] : P.thr[c]);
            if (P.nn[c] && r != COMMIT_HIP_REG_GROUP_LASSO && r != COMMIT_HIP_REG_SPARSE_GROUP_LASSO)
                v = (v <= 0.0) ? 0.0 : v;                                  // (proximals.pyx:19-21)
            break;
        }
        x[i] = (T)v;
    }
}

// Group shrinkage (proximals.pyx:103-134), one wavefront per group: for groups
// that do not overlap the order does not matter.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_group_prox_par(int64_t ngroups, const int32_t *__restrict__ gptr,
                                                           const int32_t *__restrict__ gidx,
                                                           const double *__restrict__ gw, double lam,
                                                           T *__restrict__ x)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t g = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)));
    if (g >= ngroups) return;
    const int b = gptr[g], e = gptr[g + 1];
    double s = 0.0;
    for (int k = b + lane; k < e; k += kWave) {
        const double v = (double)x[gidx[k]];
        s = fma(v, v, s);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    const double nrm = sqrt(s);
    const double wl = gw[g] * lam;
    const double f = (nrm <= wl) ? 0.0 : (nrm - wl) / nrm;
    for (int k = b + lane; k < e; k += kWave) {
        const int32_t q = gidx[k];
        x[q] = (f == 0.0) ? (T)0 : (T)((double)x[q] * f);
    }
}

// Overlapping (hierarchical) groups: one block applies the groups in order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_group_prox_seq(int64_t ngroups, const int32_t *__restrict__ gptr,
                                                           const int32_t *__restrict__ gidx,
                                                           const double *__restrict__ gw, double lam,
                                                           T *__restrict__ x)
{
    __shared__ double sh[kBlock / kWave];
    __shared__ double fac;
    for (int64_t g = 0; g < ngroups; g++) {
        const int b = gptr[g], e = gptr[g + 1];
        double s = 0.0;
        for (int k = b + (int)threadIdx.x; k < e; k += kBlock) {
            const double v = (double)x[gidx[k]];
            s = fma(v, v, s);
        }
        s = block_sum_256(s, sh);
        if (threadIdx.x == 0) {
            const double nrm = sqrt(s);
            const double wl = gw[g] * lam;
            fac = (nrm <= wl) ? 0.0 : (nrm - wl) / nrm;
        }
        __syncthreads();
        const double f = fac;
        for (int k = b + (int)threadIdx.x; k < e; k += kBlock) {
            const int32_t q = gidx[k];
            x[q] = (f == 0.0) ? (T)0 : (T)((double)x[q] * f);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_nn_range(T *__restrict__ x, int64_t b, int64_t e)
{
    for (int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (x[i] <= (T)0) ? (T)0 : x[i];
}

template <typename T>
struct SumArgs {
    const T *x, *xhat, *grad, *prev;
    const double *w;
    int64_t l1b[3], l1e[3];
    int l1w[3];
};

// Stage 1 of the seven elementwise sums; part[t * kSumBlocks + block].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sums_stage1(SumArgs<T> a, int64_t n, double *__restrict__ part)
{
    __shared__ double sh[kBlock / kWave];
    const int64_t chunk = (n + kSumBlocks - 1) / kSumBlocks;
    const int64_t b = (int64_t)blockIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = b + threadIdx.x; i < e; i += kBlock) {
        const double xi = (double)a.x[i];
        if (a.xhat) {
            const double d = xi - (double)a.xhat[i];
            s[0] = fma(d, (double)a.grad[i], s[0]);
            s[1] = fma(d, d, s[1]);
        }
        if (a.prev) {
            const double d = xi - (double)a.prev[i];
            s[2] = fma(d, d, s[2]);
        }
        s[3] = fma(xi, xi, s[3]);
#pragma unroll
        for (int t = 0; t < 3; t++)
            if (i >= a.l1b[t] && i < a.l1e[t]) s[4 + t] += fabs(a.l1w[t] ? a.w[i] * xi : xi);
    }
#pragma unroll
    for (int t = 0; t < 7; t++) {
        const double v = block_sum_256(s[t], sh);
        if (threadIdx.x == 0) part[t * kSumBlocks + blockIdx.x] = v;
    }
}

// gn[g] = group_w[g] * ||x_g||  (proximals.pyx:80-99), one wavefront per group
template <typename T>
__global__ __launch_bounds__(kBlock) void k_group_wnorm(int64_t ngroups, const int32_t *__restrict__ gptr,
                                                        const int32_t *__restrict__ gidx,
                                                        const double *__restrict__ gw, const T *__restrict__ x,
                                                        double *__restrict__ gn)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t g = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)));
    if (g >= ngroups) return;
    double s = 0.0;
    for (int k = gptr[g] + lane; k < gptr[g + 1]; k += kWave) {
        const double v = (double)x[gidx[k]];
        s = fma(v, v, s);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) gn[g] = gw[g] * sqrt(s);
}

// out[t] = sum of part[t * kSumBlocks ...], one block per term
__global__ __launch_bounds__(kBlock) void k_sums_stage2(const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double sh[kBlock / kWave];
    const double v = block_sum_256(part[blockIdx.x * kSumBlocks + threadIdx.x], sh);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a, const T *__restrict__ x, double b,
                                                  const T *__restrict__ y, T *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

unsigned elem_grid(int64_t n)
{
    const int64_t g = (n + kBlock - 1) / kBlock;
    return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

unsigned wave_grid(int64_t nw)
{
    const int64_t g = (nw + kWavesPerBlock - 1) / kWavesPerBlock;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

static_assert(kSumBlocks == kBlock, "stage 2 reduces one partial per thread");

namespace {

template <typename T>
int prox_impl(const commit_hip_prox_params *p, int64_t n, const T *xin, const T *grad, double step, double scale,
              T *x, void *stream)
{
    if (!p || !xin || !x || n < 0) return fail(COMMIT_HIP_EINVAL, "NULL argument or negative size");
    if (n == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    ProxArgs a{};
    for (int c = 0; c < 3; c++) {
        a.start[c] = p->start[c];
        a.end[c] = p->start[c] + p->size[c];
        if (a.start[c] < 0 || a.end[c] > n || p->size[c] < 0)
            return fail(COMMIT_HIP_EINVAL, "compartment %d range [%lld, %lld) outside x", c, (long long)a.start[c],
                        (long long)a.end[c]);
        a.reg[c] = p->reg[c];
        if (c > 0 && (a.reg[c] == COMMIT_HIP_REG_GROUP_LASSO || a.reg[c] == COMMIT_HIP_REG_SPARSE_GROUP_LASSO))
            return fail(COMMIT_HIP_EUNSUPPORTED, "group penalties exist only for the IC compartment");
        a.nn[c] = p->nn[c];
        a.weighted[c] = p->weighted[c];
        a.thr[c] = p->lam[c] * scale;
        if (a.weighted[c] && !p->w) return fail(COMMIT_HIP_EINVAL, "weighted lasso without weights");
    }
    a.w = p->w;
    hipLaunchKernelGGL(k_prox_elem<T>, dim3(elem_grid(n)), dim3(kBlock), 0, st, a, n, xin, grad, step, x);
    HIP_TRY(hipGetLastError());
    const int r = p->reg[0];
    if (r == COMMIT_HIP_REG_GROUP_LASSO || r == COMMIT_HIP_REG_SPARSE_GROUP_LASSO) {
        const double lam = p->lam_group * scale;
        if (lam != 0.0 && p->ngroups > 0) {          // proximals.pyx:116-117
            if (!p->group_ptr || !p->group_idx || !p->group_w) return fail(COMMIT_HIP_EINVAL, "group arrays missing");
            if (p->groups_overlap)
                hipLaunchKernelGGL(k_group_prox_seq<T>, dim3(1), dim3(kBlock), 0, st, p->ngroups, p->group_ptr,
                                   p->group_idx, p->group_w, lam, x);
            else
                hipLaunchKernelGGL(k_group_prox_par<T>, dim3(wave_grid(p->ngroups)), dim3(kBlock), 0, st, p->ngroups,
                                   p->group_ptr, p->group_idx, p->group_w, lam, x);
            HIP_TRY(hipGetLastError());
        }
        if (p->nn[0] && p->size[0] > 0) {
            hipLaunchKernelGGL(k_nn_range<T>, dim3(elem_grid(p->size[0])), dim3(kBlock), 0, st, x, a.start[0], a.end[0]);
            HIP_TRY(hipGetLastError());
        }
    }
    return 0;
}

template <typename T>
int sums_impl(int64_t n, const T *x, const T *xhat, const T *grad, const T *prev, const int64_t *l1_range,
              const int *l1_weighted, const double *w, int64_t ngroups, const int32_t *group_ptr,
              const int32_t *group_idx, const double *group_w, double *out, double *work, void *stream)
{
    if (!x || !out || !work || n < 0 || ngroups < 0) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if ((xhat == nullptr) != (grad == nullptr)) return fail(COMMIT_HIP_EINVAL, "xhat and grad go together");
    hipStream_t st = static_cast<hipStream_t>(stream);
    SumArgs<T> a{};
    a.x = x;
    a.xhat = xhat;
    a.grad = grad;
    a.prev = prev;
    a.w = w;
    for (int t = 0; t < 3; t++) {
        a.l1b[t] = l1_range ? l1_range[2 * t] : 0;
        a.l1e[t] = l1_range ? l1_range[2 * t + 1] : 0;
        a.l1w[t] = l1_weighted ? l1_weighted[t] : 0;
        if (a.l1w[t] && !w) return fail(COMMIT_HIP_EINVAL, "weighted l1 term without weights");
    }
    double *part = work;
    double *gn = work + 8 * kSumBlocks;
    hipLaunchKernelGGL(k_sums_stage1<T>, dim3(kSumBlocks), dim3(kBlock), 0, st, a, n, part);
    HIP_TRY(hipGetLastError());
    if (ngroups > 0) {
        if (!group_ptr || !group_idx || !group_w) return fail(COMMIT_HIP_EINVAL, "group arrays missing");
        hipLaunchKernelGGL(k_group_wnorm<T>, dim3(wave_grid(ngroups)), dim3(kBlock), 0, st, ngroups, group_ptr,
                           group_idx, group_w, x, gn);
        hipLaunchKernelGGL(k_sum_stage1, dim3(kSumBlocks), dim3(kBlock), 0, st, gn, (size_t)ngroups,
                           part + 7 * kSumBlocks);
    } else {   // no groups: a zero-length sum writes zero partials (no memset node in captured graphs)
        hipLaunchKernelGGL(k_sum_stage1, dim3(kSumBlocks), dim3(kBlock), 0, st, gn, (size_t)0,
                           part + 7 * kSumBlocks);
    }
    hipLaunchKernelGGL(k_sums_stage2, dim3(8), dim3(kBlock), 0, st, part, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Workspace of commit_hip_fista_sums_noalloc: a file-scope device array that
// stands in for the `work` buffer commit_hip_fista_sums takes from its caller
// (one copy per GPU, held until the process ends).  Its name is a host-side
// stand-in: kernels get the device address that hipGetSymbolAddress returns,
// looked up once per device and kept in g_scratch_ptr.  Every partial it holds
// is written before it is read in each call, so nothing needs clearing.
constexpr size_t kScratchDoubles = COMMIT_HIP_SUMS_WORK((size_t)COMMIT_HIP_NOALLOC_MAX_GROUPS);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (the current device), looked up on the
// first call on that device.
int scratch_ptr(double **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(COMMIT_HIP_EUNSUPPORTED, "device %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(COMMIT_HIP_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

template <typename T>
int axpby_impl(int64_t n, double a, const T *x, double b, const T *y, T *out, void *stream)
{
    if (!x || !y || !out || n < 0) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_axpby<T>, dim3(elem_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, a, x, b,
                       y, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int commit_hip_prox(const commit_hip_prox_params *p, int64_t n, const double *xin, const double *grad,
                    double step, double scale, double *x, void *stream)
{
    return prox_impl<double>(p, n, xin, grad, step, scale, x, stream);
}

int commit_hip_prox_f32(const commit_hip_prox_params *p, int64_t n, const float *xin, const float *grad,
                        double step, double scale, float *x, void *stream)
{
    return prox_impl<float>(p, n, xin, grad, step, scale, x, stream);
}

int commit_hip_fista_sums(int64_t n, const double *x, const double *xhat, const double *grad,
                          const double *prev, const int64_t *l1_range, const int *l1_weighted,
                          const double *w, int64_t ngroups, const int32_t *group_ptr,
                          const int32_t *group_idx, const double *group_w, double *out, double *work,
                          void *stream)
{
    return sums_impl<double>(n, x, xhat, grad, prev, l1_range, l1_weighted, w, ngroups, group_ptr, group_idx,
                             group_w, out, work, stream);
}

int commit_hip_fista_sums_f32(int64_t n, const float *x, const float *xhat, const float *grad,
                              const float *prev, const int64_t *l1_range, const int *l1_weighted,
                              const double *w, int64_t ngroups, const int32_t *group_ptr,
                              const int32_t *group_idx, const double *group_w, double *out, double *work,
                              void *stream)
{
    return sums_impl<float>(n, x, xhat, grad, prev, l1_range, l1_weighted, w, ngroups, group_ptr, group_idx,
                            group_w, out, work, stream);
}

int commit_hip_fista_sums_noalloc(int64_t n, const double *x, const double *xhat, const double *grad,
                                  const double *prev, const int64_t *l1_range, const int *l1_weighted,
                                  const double *w, int64_t ngroups, const int32_t *group_ptr,
                                  const int32_t *group_idx, const double *group_w, double *out, void *stream)
{
    if (ngroups > COMMIT_HIP_NOALLOC_MAX_GROUPS)
        return fail(COMMIT_HIP_EUNSUPPORTED, "%lld groups exceed the %d of the fixed workspace", (long long)ngroups,
                    COMMIT_HIP_NOALLOC_MAX_GROUPS);
    double *work = nullptr;
    const int rc = scratch_ptr(&work);
    if (rc) return rc;
    return sums_impl<double>(n, x, xhat, grad, prev, l1_range, l1_weighted, w, ngroups, group_ptr, group_idx,
                             group_w, out, work, stream);
}

int commit_hip_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out, void *stream)
{
    return axpby_impl<double>(n, a, x, b, y, out, stream);
}

int commit_hip_axpby_f32(int64_t n, double a, const float *x, double b, const float *y, float *out, void *stream)
{
    return axpby_impl<float>(n, a, x, b, y, out, stream);
}

}  // extern "C"
