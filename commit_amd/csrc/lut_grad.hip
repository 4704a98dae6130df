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
// IC tables (the bulk of the work): an orientation-major walk.  The plan keeps
// (built on first use) the IC entries listed by orientation, with each entry's
// voxel, cut into chunks of at most kOmChunk entries of one orientation.  One
// wavefront per chunk accumulates the chunk's gradient row in registers (lane
// = sample, NJ per lane, FP64): per entry it gathers the gy row of the entry's
// voxel (rows of the next two entries in flight) and adds coefficient x gy.
// No LDS tables, no atomics; a second kernel sums each orientation's chunk
// rows in chunk order, so the result is deterministic.
//
// EC and ISO tables (a few entries per voxel) -- scatter-reduction with privatised LDS tables: a work-group is one wavefront
// that owns an FP64 table of 64 orientations x 64 samples (32 KB of LDS, so
// five work-groups share a CU) for one (RF, sample chunk, orientation part)
// and walks a slab of voxels.  Lane s only ever touches column s, and an entry
// is one LDS add without return (ds_add_f64) issued in entry order by the
// table's only wavefront, so the wavefront never waits on the table and the
// result is deterministic.  The slab tables go to a partial buffer and a
// second kernel adds the slabs in order.
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

constexpr int kOrPart = 64;         // orientations per LDS table
constexpr int kSlabs = 64;          // voxel slabs (work-groups per table)
constexpr int kDepth = 4;           // voxels in flight in the walk of k_lut_grad_part
constexpr int kOmChunk = 1024;      // IC entries per wavefront of the orientation-major walk

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

// part[((slab * nparts + h) * nchunk + c) * kOrPart * 64 + (o - h*kOrPart) * 64 + lane]
// for one RF of one compartment.  xb: x block of the RF (IC: + eval applied
// through ev), lengths for IC.
template <typename T>
__global__ __launch_bounds__(kWave) void k_lut_grad_part(DevDict d, Comp cp, int ndirs, const T *__restrict__ xb,
                                                         const uint32_t *__restrict__ ev, const T *__restrict__ gy,
                                                         int nparts, int nchunk, double *__restrict__ part)
{
    __shared__ double tab[kOrPart * kWave];
    const int lane = threadIdx.x;
    const int bid = blockIdx.x;
    const int slab = bid / (nparts * nchunk), rem = bid - slab * nparts * nchunk;
    const int h = rem / nchunk, c = rem - h * nchunk;
    const int o0 = h * kOrPart;
    const int no = min(kOrPart, ndirs - o0);
    for (int k = lane; k < kOrPart * kWave; k += kWave) tab[k] = 0.0;
    __builtin_amdgcn_wave_barrier();
    const int s = c * kWave + lane;
    const int sc = s < d.nS ? s : 0;
    const int v0 = (int)(((int64_t)d.nVOX * slab) / kSlabs), v1 = (int)(((int64_t)d.nVOX * (slab + 1)) / kSlabs);
    // Software pipeline over the slab's voxels (first 64 entries of each),
    // kDepth voxels deep: while voxel v is applied, the records and gy of
    // voxel v+kDepth, the x gathers of v+kDepth-1 and the entry range of
    // v+kDepth+1 load.  Stage data lives in rings of kDepth+1 slots indexed at
    // compile time (the voxel loop is unrolled by the ring size).  Later
    // chunks of a long voxel load in place; stage data of voxels past v1 is
    // never applied.
    auto range = [&](int v, uint32_t &b, uint32_t &e) {
        const int vc = v < v1 ? v : v1 - 1;
        b = cp.ptr[vc];
        e = v < v1 ? cp.ptr[vc + 1] : b;
    };
    auto recs = [&](uint32_t b, uint32_t e, uint32_t base, int &o, uint32_t &col, float &len) {
        const uint32_t k = base + lane;
        const uint32_t ks = k < e ? k : (b < e ? b : 0u);
        o = cp.o ? (int)cp.o[ks] : 0;
        col = cp.col[ks];
        len = cp.len ? cp.len[ks] : 1.0f;
    };
    auto gyv = [&](int v) { return (double)gy[(size_t)min(v, v1 - 1) * d.nS + sc]; };
    // x gather and eval of an entry's column, loaded in the same stage
    auto xload = [&](uint32_t col, T &xv, uint32_t &evv) {
        xv = xb[col];
        evv = ev ? ev[col] : 1u;
    };
    auto apply = [&](uint32_t e, uint32_t base, int o, float len, T xv, uint32_t evv, double g) {
        const bool mine = base + lane < e && o >= o0 && o < o0 + no;        // this part's orientations
        double cf = (double)xv;
        if (cp.len) cf *= (double)len;
        if (evv == 0u) cf = 0.0;
        if (!mine) cf = 0.0;
        unsigned long long m = __ballot(cf != 0.0);
        while (m) {
            const int t = __builtin_ctzll(m);
            m &= m - 1;
            const int ot = __builtin_amdgcn_readlane(o, t) - o0;
            const long long cb = __double_as_longlong(cf);
            const int lo = __builtin_amdgcn_readlane((int)(cb & 0xffffffffll), t);
            const int hi = __builtin_amdgcn_readlane((int)(cb >> 32), t);
            const double ct = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
            atomicAdd(tab + ot * kWave + lane, ct * g);          // lanes hit distinct addresses
        }
    };
    constexpr int R = kDepth + 1;                                 // ring slots
    if (v0 < v1 && cp.m > 0) {                                  // (record loads clamp to entry 0)
        uint32_t rb[R], re[R], rc[R], rev[R];
        int ro[R];
        float rl[R];
        double rg[R];
        T rx[R];
#pragma unroll
        for (int j = 0; j < R; j++) range(v0 + j, rb[j], re[j]);
#pragma unroll
        for (int j = 0; j < kDepth; j++) {
            recs(rb[j], re[j], rb[j], ro[j], rc[j], rl[j]);
            rg[j] = gyv(v0 + j);
        }
#pragma unroll
        for (int j = 0; j < kDepth - 1; j++) xload(rc[j], rx[j], rev[j]);
        for (int vb = v0; vb < v1; vb += R) {
#pragma unroll
            for (int j = 0; j < R; j++) {
                const int v = vb + j;
                if (v >= v1) break;
                const int sr = (j + kDepth) % R, sx = (j + kDepth - 1) % R;
                recs(rb[sr], re[sr], rb[sr], ro[sr], rc[sr], rl[sr]);   // stage: records and gy of v+kDepth
                rg[sr] = gyv(v + kDepth);
                xload(rc[sx], rx[sx], rev[sx]);                    // stage: x and eval of v+kDepth-1
                if (rb[j] < re[j]) {
                    const double g = s < d.nS ? rg[j] : 0.0;
                    apply(re[j], rb[j], ro[j], rl[j], rx[j], rev[j], g);
                    for (uint32_t base = rb[j] + kWave; base < re[j]; base += kWave) {
                        int o;
                        uint32_t col;
                        float len;
                        recs(rb[j], re[j], base, o, col, len);
                        T xv;
                        uint32_t evv;
                        xload(col, xv, evv);
                        apply(re[j], base, o, len, xv, evv, rg[j]);
                    }
                }
                range(v + kDepth + 1, rb[j], re[j]);              // stage: entry range of v+kDepth+1
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    double *dst = part + ((size_t)bid) * kOrPart * kWave;
    for (int k = lane; k < no * kWave; k += kWave) dst[k] = tab[k];
}

// out[o][s] (+ RF offset) = sum over slabs of the partial tables, in slab order.
__global__ void k_lut_grad_reduce(const double *__restrict__ part, int ndirs, int nS, int nparts, int nchunk,
                                  double *__restrict__ out)
{
    const int64_t total = (int64_t)ndirs * nS;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t / nS), s = (int)(t - (int64_t)o * nS);
        const int h = o / kOrPart, c = s / kWave, lane = s - c * kWave;
        double acc = 0.0;
        for (int p = 0; p < kSlabs; p++) {
            const size_t bid = ((size_t)p * nparts + h) * nchunk + c;
            acc += part[bid * kOrPart * kWave + (size_t)(o - h * kOrPart) * kWave + lane];
        }
        out[t] = acc;
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

// part[chunk][s] = sum over the chunk's entries (orientation o of RF i) of
// l * x * eval * gy[v*nS + s], entries in list order.
template <int NJ, typename T>
__global__ __launch_bounds__(kWave) void k_lut_grad_om(DevDict d, const uint32_t *__restrict__ ent,
                                                       const uint32_t *__restrict__ vox,
                                                       const uint32_t *__restrict__ chunks, const T *__restrict__ xb,
                                                       const uint32_t *__restrict__ ev, const T *__restrict__ gy,
                                                       double *__restrict__ part)
{
    const int lane = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const uint32_t k0 = chunks[3 * c + 1], k1 = chunks[3 * c + 2];
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
        const uint32_t col = d.ic.col[e];
        double cf = (double)xb[col];
        cf *= (double)d.ic.len[e];
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

// Orientation-major list of the IC entries, their voxels, the chunks and the
// chunk partial rows (once per plan; the IC structure never changes).
int om_build(commit_hip_plan *p)
{
    if (p->lg_cptr) return 0;                              // built (ndirs+1 pointers, never empty)
    const DevDict &d = p->d;
    const size_t n = (size_t)d.n;
    std::vector<uint32_t> ptr((size_t)d.nVOX + 1);
    std::vector<uint16_t> o(n);
    HIP_TRY(hipMemcpy(ptr.data(), d.ic.ptr, ptr.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(o.data(), d.ic.o, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> optr((size_t)d.ndirs + 1, 0u), ent(n), vox(n);
    for (size_t k = 0; k < n; k++) optr[o[k] + 1u]++;
    for (int q = 0; q < d.ndirs; q++) optr[q + 1] += optr[q];
    std::vector<uint32_t> pos(optr.begin(), optr.end() - 1);
    for (int v = 0; v < d.nVOX; v++)                       // stable: voxel order inside an orientation
        for (uint32_t k = ptr[v]; k < ptr[v + 1]; k++) {
            const uint32_t at = pos[o[k]]++;
            ent[at] = k;
            vox[at] = (uint32_t)v;
        }
    std::vector<uint32_t> chunks, cptr((size_t)d.ndirs + 1, 0u);
    for (int q = 0; q < d.ndirs; q++) {
        cptr[q] = (uint32_t)(chunks.size() / 3);
        for (uint32_t k = optr[q]; k < optr[q + 1]; k += kOmChunk) {
            chunks.push_back((uint32_t)q);
            chunks.push_back(k);
            chunks.push_back(std::min(optr[q + 1], k + (uint32_t)kOmChunk));
        }
    }
    cptr[d.ndirs] = (uint32_t)(chunks.size() / 3);
    const int nch = (int)(chunks.size() / 3);
    int rc;
    void *m = nullptr;
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
    if ((rc = up(ent, &p->lg_ent)) || (rc = up(vox, &p->lg_vox)) || (rc = up(chunks, &p->lg_chunk)) ||
        (rc = up(cptr, &p->lg_cptr)))
        return rc;
    if (nch > 0 && (rc = commit_internal::dev_alloc(p, (size_t)nch * d.nS * sizeof(double), &m))) return rc;
    p->lg_part = static_cast<double *>(m);
    p->lg_nchunks = nch;
    return 0;
}

template <typename T>
int launch_om(commit_hip_plan *p, const T *xb, const uint32_t *ev, const T *gy, double *out, hipStream_t st)
{
    const DevDict &d = p->d;
    if (p->lg_nchunks > 0) {
        switch (d.nSp / kWave) {
#define CASE(NJ)                                                                                                  \
    case NJ:                                                                                                      \
        hipLaunchKernelGGL((k_lut_grad_om<NJ, T>), dim3((unsigned)p->lg_nchunks), dim3(kWave), 0, st, d, p->lg_ent, \
                           p->lg_vox, p->lg_chunk, xb, ev, gy, p->lg_part);                                       \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds 512", d.nS);
        }
        HIP_TRY(hipGetLastError());
    }
    const unsigned rg = (unsigned)std::min<int64_t>(((int64_t)d.ndirs * d.nS + 255) / 256, 4096);
    hipLaunchKernelGGL(k_lut_grad_om_reduce, dim3(rg), dim3(256), 0, st, p->lg_part, p->lg_cptr, d.ndirs, d.nS, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
int lut_grad(commit_hip_plan *plan, const T *x, const T *gy, double *gwmr, double *gwmh, double *giso,
             hipStream_t st)
{
    const DevDict &d = plan->d;
    const int nparts = (d.ndirs + kOrPart - 1) / kOrPart;
    const int nchunk = (d.nS + kWave - 1) / kWave;
    const size_t blocks = (size_t)kSlabs * nparts * nchunk;
    const size_t need = blocks * kOrPart * kWave;
    if (plan->lgrad_elems < need) {
        void *p = nullptr;
        int rc = commit_internal::dev_alloc(plan, need * sizeof(double), &p);
        if (rc) return rc;
        plan->lgrad = static_cast<double *>(p);
        plan->lgrad_elems = need;
    }
    const unsigned rg = (unsigned)std::min<int64_t>(((int64_t)d.ndirs * d.nS + 255) / 256, 4096);
    if (gwmr && d.nIC > 0) {
        int rc = om_build(plan);
        if (rc) return rc;
        for (int i = 0; i < d.nIC; i++)
            if ((rc = launch_om<T>(plan, x + (size_t)i * d.nF, d.eval + (size_t)i * d.nF, gy,
                                   gwmr + (size_t)i * d.ndirs * d.nS, st)))
                return rc;
    }
    if (gwmh)
        for (int i = 0; i < d.nEC; i++) {
            hipLaunchKernelGGL(k_lut_grad_part<T>, dim3((unsigned)blocks), dim3(kWave), 0, st, d, d.ec, d.ndirs,
                               x + (size_t)d.nIC * d.nF + (size_t)i * d.nE, (const uint32_t *)nullptr, gy, nparts,
                               nchunk, plan->lgrad);
            hipLaunchKernelGGL(k_lut_grad_reduce, dim3(rg), dim3(256), 0, st, plan->lgrad, d.ndirs, d.nS, nparts,
                               nchunk, gwmh + (size_t)i * d.ndirs * d.nS);
        }
    if (giso) {   // ISO: one row per RF (a single "orientation"), same scatter-reduction
        const size_t b1 = (size_t)kSlabs * nchunk;
        const unsigned r1 = (unsigned)std::max<int64_t>(1, ((int64_t)d.nS + 255) / 256);
        for (int i = 0; i < d.nISO; i++) {
            hipLaunchKernelGGL(k_lut_grad_part<T>, dim3((unsigned)b1), dim3(kWave), 0, st, d, d.iso, 1,
                               x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE + (size_t)i * d.nI,
                               (const uint32_t *)nullptr, gy, 1, nchunk, plan->lgrad);
            hipLaunchKernelGGL(k_lut_grad_reduce, dim3(r1), dim3(256), 0, st, plan->lgrad, 1, d.nS, 1, nchunk,
                               giso + (size_t)i * d.nS);
        }
    }
    HIP_TRY(hipGetLastError());
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
