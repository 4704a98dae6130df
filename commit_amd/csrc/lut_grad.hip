// Gradient of y = A(LUT) x with respect to the response-function tables, and
// replacement of the plan's tables from device memory: the two pieces a
// differentiable operator needs to fit the compartment diffusivities (the
// tables are functions of them) by gradient.  No counterpart in the
// reference, whose tables are fixed inputs (setup_operator.py:93,151,299).
//
// With gy = dL/dy (rows v*nS + s):
//   dL/dwmr[i][o][s] = sum over IC segments (f, v, o, l):  l * x[i*nF + f] * eval * gy[v*nS + s]
//   dL/dwmh[i][o][s] = sum over EC peaks (v, o) of column e: x[EC i, e] * gy[v*nS + s]
//   dL/diso[i][s]    = sum over ISO voxels v of column k:   x[ISO i, k] * gy[v*nS + s]
// (the LUT entry multiplies exactly these terms in COMMIT_A, setup_operator.py:62-218).
//
// An orientation-major walk per compartment (IC, EC, ISO).  The plan keeps
// (built on first use) each compartment's entries listed by orientation, with
// each entry's voxel, cut into chunks of at most kOmChunk entries of one
// orientation.  One wavefront per chunk accumulates the chunk's gradient row
// in registers (lane = sample, NJ per lane, FP64): per entry it gathers the gy
// row of the entry's voxel (rows of the next two entries in flight) and adds
// coefficient x gy.  No LDS tables, no atomics; a second kernel sums each
// orientation's chunk rows in chunk order, so the result is deterministic.
// (ISO has a single row per RF: one "orientation".)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/commit_hip.h"
#include "common.h"
#include "plan.h"
#include "rows.h"

using commit_internal::fail;

namespace {

using commit_internal::Comp;
using commit_internal::DevDict;
using commit_internal::OmList;

constexpr int kOmChunk = 1024;      // entries per wavefront of the orientation-major walk

// Pad [rows][nS] float tables into the plan's [rows][nSp] layout.
__global__ void k_pad_lut(const float *__restrict__ src, float *__restrict__ dst, int64_t rows, int nS, int nSp)
{
    const int64_t total = rows * nSp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / nSp;
        const int s = (int)(t - r * nSp);
        dst[t] = s < nS ? src[r * nS + s] : 0.0f;
    }
}

// NJ loads of the gy row of one voxel (this lane's sample 0; tail: the lane's
// last sample is < nS).
template <int NJ, typename T>
__device__ __forceinline__ void gy_row(T (&g)[NJ], const T *__restrict__ row, bool tail)
{
#pragma unroll
    for (int j = 0; j < NJ - 1; j++) g[j] = row[kWave * j];
    g[NJ - 1] = tail ? row[kWave * (NJ - 1)] : (T)0;
}

template <int NJ, typename T>
__device__ __forceinline__ void gy_fma(double (&acc)[NJ], double c, const T (&g)[NJ])
{
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = fma(c, (double)g[j], acc[j]);
}

// part[chunk][s] = sum over the chunk's entries (one orientation of one RF)
// of coefficient * gy[v*nS + s], entries in list order; coefficient
// l * x * eval for IC (setup_operator.py:62-96), x for EC and ISO (:111-218).
template <int NJ, typename T>
__global__ __launch_bounds__(kWave) void k_lut_grad_om(DevDict d, Comp cp, const uint32_t *__restrict__ ent,
                                                       const uint32_t *__restrict__ vox,
                                                       const uint32_t *__restrict__ chunks, const T *__restrict__ xb,
                                                       const uint32_t *__restrict__ ev, const T *__restrict__ gy,
                                                       double *__restrict__ part)
{
    const int lane = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const uint32_t k0 = chunks[2 * c], k1 = chunks[2 * c + 1];
    const bool tail = lane + kWave * (NJ - 1) < d.nS;
    const T *gl = gy + lane;
    double acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = 0.0;
    for (uint32_t base = k0; base < k1; base += kWave) {
        const uint32_t k = base + lane;
        const bool ok = k < k1;
        const uint32_t ks = ok ? k : k0;
        const uint32_t e = ent[ks], v = vox[ks];
        const uint32_t col = cp.col[e];
        double cf = (double)xb[col];
        if (cp.len) cf *= (double)cp.len[e];
        if (ev && ev[col] == 0u) cf = 0.0;
        if (!ok) cf = 0.0;
        unsigned long long m = __ballot(cf != 0.0);
        // three gy rows rotate through the pipeline; entries keep their order
        T ga[NJ], gb[NJ], gc[NJ];
        int ta = take(m), tb = -1, tc = -1;
        if (ta >= 0) gy_row<NJ>(ga, gl + (size_t)(uint32_t)__builtin_amdgcn_readlane((int)v, ta) * d.nS, tail);
        if ((tb = take(m)) >= 0) gy_row<NJ>(gb, gl + (size_t)(uint32_t)__builtin_amdgcn_readlane((int)v, tb) * d.nS, tail);
        while (true) {
            if (ta < 0) break;
            if ((tc = take(m)) >= 0)
                gy_row<NJ>(gc, gl + (size_t)(uint32_t)__builtin_amdgcn_readlane((int)v, tc) * d.nS, tail);
            gy_fma<NJ>(acc, readlane_f64(cf, ta), ga);
            if (tb < 0) break;
            if ((ta = take(m)) >= 0)
                gy_row<NJ>(ga, gl + (size_t)(uint32_t)__builtin_amdgcn_readlane((int)v, ta) * d.nS, tail);
            gy_fma<NJ>(acc, readlane_f64(cf, tb), gb);
            if (tc < 0) break;
            if ((tb = take(m)) >= 0)
                gy_row<NJ>(gb, gl + (size_t)(uint32_t)__builtin_amdgcn_readlane((int)v, tb) * d.nS, tail);
            gy_fma<NJ>(acc, readlane_f64(cf, tc), gc);
        }
    }
    double *dst = part + (size_t)c * d.nS;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int s = lane + kWave * j;
        if (s < d.nS) dst[s] = acc[j];
    }
}

// out[o][s] = sum of the chunk rows of orientation o, in chunk order.
__global__ void k_lut_grad_om_reduce(const double *__restrict__ part, const uint32_t *__restrict__ cptr, int ndirs,
                                     int nS, double *__restrict__ out)
{
    const int64_t total = (int64_t)ndirs * nS;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t / nS), s = (int)(t - (int64_t)o * nS);
        double acc = 0.0;
        for (uint32_t c = cptr[o]; c < cptr[o + 1]; c++) acc += part[(size_t)c * nS + s];
        out[t] = acc;
    }
}

// Orientation-major list of one compartment's entries, their voxels, the
// chunks and the chunk partial rows (once per plan; the structure never
// changes).  ndirs: orientations (1 for ISO, whose entries have none).
int om_build(commit_hip_plan *p, const Comp &cp, int ndirs, OmList &L)
{
    if (L.cptr) return 0;                                  // built (ndirs+1 pointers, never empty)
    const DevDict &d = p->d;
    const size_t n = cp.m;
    std::vector<uint32_t> ptr((size_t)d.nVOX + 1);
    std::vector<uint16_t> o(n, 0);
    HIP_TRY(hipMemcpy(ptr.data(), cp.ptr, ptr.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (cp.o && n) HIP_TRY(hipMemcpy(o.data(), cp.o, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> optr((size_t)ndirs + 1, 0u), ent(n), vox(n);
    for (size_t k = 0; k < n; k++) optr[o[k] + 1u]++;
    for (int q = 0; q < ndirs; q++) optr[q + 1] += optr[q];
    std::vector<uint32_t> pos(optr.begin(), optr.end() - 1);
    for (int v = 0; v < d.nVOX; v++)                       // stable: voxel order inside an orientation
        for (uint32_t k = ptr[v]; k < ptr[v + 1]; k++) {
            const uint32_t at = pos[o[k]]++;
            ent[at] = k;
            vox[at] = (uint32_t)v;
        }
    std::vector<uint32_t> chunks, cptr((size_t)ndirs + 1, 0u);
    for (int q = 0; q < ndirs; q++) {
        cptr[q] = (uint32_t)(chunks.size() / 2);
        for (uint32_t k = optr[q]; k < optr[q + 1]; k += kOmChunk) {
            chunks.push_back(k);
            chunks.push_back(std::min(optr[q + 1], k + (uint32_t)kOmChunk));
        }
    }
    cptr[ndirs] = (uint32_t)(chunks.size() / 2);
    const int nch = (int)(chunks.size() / 2);
    auto up = [&](const std::vector<uint32_t> &h, uint32_t **dst) -> int {
        *dst = nullptr;
        if (h.empty()) return 0;
        void *mm = nullptr;
        const int r = commit_internal::dev_alloc(p, h.size() * sizeof(uint32_t), &mm);
        if (r) return r;
        HIP_TRY(hipMemcpy(mm, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        *dst = static_cast<uint32_t *>(mm);
        return 0;
    };
    int rc;
    if ((rc = up(ent, &L.ent)) || (rc = up(vox, &L.vox)) || (rc = up(chunks, &L.chunk))) return rc;
    void *m = nullptr;
    if (nch > 0 && (rc = commit_internal::dev_alloc(p, (size_t)nch * d.nS * sizeof(double), &m))) return rc;
    L.part = static_cast<double *>(m);
    L.nchunks = nch;
    return up(cptr, &L.cptr);                              // last: marks the list as built
}

// out[o][s] for orientations o < ndirs of one RF: x block xb, eval ev (IC).
template <typename T>
int launch_om(commit_hip_plan *p, const Comp &cp, const OmList &L, int ndirs, const T *xb, const uint32_t *ev,
              const T *gy, double *out, hipStream_t st)
{
    const DevDict &d = p->d;
    if (L.nchunks > 0) {
        switch (d.nSp / kWave) {
#define CASE(NJ)                                                                                                  \
    case NJ:                                                                                                      \
        hipLaunchKernelGGL((k_lut_grad_om<NJ, T>), dim3((unsigned)L.nchunks), dim3(kWave), 0, st, d, cp, L.ent,   \
                           L.vox, L.chunk, xb, ev, gy, L.part);                                                   \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds 512", d.nS);
        }
        HIP_TRY(hipGetLastError());
    }
    const unsigned rg = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)ndirs * d.nS + 255) / 256, 4096));
    hipLaunchKernelGGL(k_lut_grad_om_reduce, dim3(rg), dim3(256), 0, st, L.part, L.cptr, ndirs, d.nS, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
int lut_grad(commit_hip_plan *plan, const T *x, const T *gy, double *gwmr, double *gwmh, double *giso,
             hipStream_t st)
{
    const DevDict &d = plan->d;
    int rc;
    if (gwmr && d.nIC > 0) {
        if ((rc = om_build(plan, d.ic, d.ndirs, plan->lg_om[0]))) return rc;
        for (int i = 0; i < d.nIC; i++)
            if ((rc = launch_om<T>(plan, d.ic, plan->lg_om[0], d.ndirs, x + (size_t)i * d.nF,
                                   d.eval + (size_t)i * d.nF, gy, gwmr + (size_t)i * d.ndirs * d.nS, st)))
                return rc;
    }
    if (gwmh && d.nEC > 0) {
        if ((rc = om_build(plan, d.ec, d.ndirs, plan->lg_om[1]))) return rc;
        for (int i = 0; i < d.nEC; i++)
            if ((rc = launch_om<T>(plan, d.ec, plan->lg_om[1], d.ndirs, x + (size_t)d.nIC * d.nF + (size_t)i * d.nE,
                                   (const uint32_t *)nullptr, gy, gwmh + (size_t)i * d.ndirs * d.nS, st)))
                return rc;
    }
    if (giso && d.nISO > 0) {   // ISO: one row per RF, a single "orientation"
        if ((rc = om_build(plan, d.iso, 1, plan->lg_om[2]))) return rc;
        for (int i = 0; i < d.nISO; i++)
            if ((rc = launch_om<T>(plan, d.iso, plan->lg_om[2], 1,
                                   x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE + (size_t)i * d.nI,
                                   (const uint32_t *)nullptr, gy, giso + (size_t)i * d.nS, st)))
                return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int commit_hip_plan_set_lut(commit_hip_plan *plan, const float *wmr, const float *wmh, const float *iso, void *stream)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (plan->nolut) return fail(COMMIT_HIP_EUNSUPPORTED, "a LUT-free plan has no tables");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevDict &d = plan->d;
    auto pad = [&](const float *src, const float *dst, int64_t rows) -> int {
        if (!src || rows == 0) return 0;
        const unsigned g = (unsigned)std::min<int64_t>((rows * d.nSp + 255) / 256, 4096);
        hipLaunchKernelGGL(k_pad_lut, dim3(g), dim3(256), 0, st, src, const_cast<float *>(dst), rows, d.nS, d.nSp);
        HIP_TRY(hipGetLastError());
        return 0;
    };
    int rc;
    if ((rc = pad(wmr, d.wmr, (int64_t)d.nIC * d.ndirs))) return rc;
    if ((rc = pad(wmh, d.wmh, (int64_t)d.nEC * d.ndirs))) return rc;
    if ((rc = pad(iso, d.isol, (int64_t)d.nISO))) return rc;
    return 0;
}

int commit_hip_lut_grad(commit_hip_plan *plan, const void *x, const void *gy, double *gwmr, double *gwmh,
                        double *giso, void *stream)
{
    if (!plan || !x || !gy) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (plan->nolut) return fail(COMMIT_HIP_EUNSUPPORTED, "a LUT-free plan has no tables");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (plan->precision == COMMIT_HIP_FP32)
        return lut_grad<float>(plan, static_cast<const float *>(x), static_cast<const float *>(gy), gwmr, gwmh, giso,
                               st);
    return lut_grad<double>(plan, static_cast<const double *>(x), static_cast<const double *>(gy), gwmr, gwmh, giso,
                            st);
}

}  // extern "C"
