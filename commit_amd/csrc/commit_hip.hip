// libcommit_hip.so — MI355X (gfx950, CDNA4) implementation of the COMMIT
// implicit operator: y = A x and x = A^T y over the IC / EC / ISO compartments.
//
// Reference algorithm: setup_operator.py:48-915 (generated COMMIT_A/COMMIT_At
// and their LUT-free variants), driven from operator.pyx:114-184.  The design
// here is GPU-first and shares none of the reference's structure:
//
//   * the plan builds, once, a voxel-major CSR layout of the segments (the
//     order load_dictionary produces, core.pyx:509-513) and a fibre-major
//     permutation of it, and keeps both resident in HBM;
//   * A x: one wavefront per voxel, the nS samples of the voxel spread over
//     the 64 lanes (NJ = ceil(nS/64) FP64 accumulators per lane); segment
//     records are read cooperatively (one per lane) and broadcast with
//     v_readlane, so the LUT row reads of a segment are fully coalesced;
//   * A^T y: two atomic-free passes.  Pass 1 (voxel-major) stages the voxel's
//     y row in LDS and computes, one segment per lane, the FP64 dot product of
//     its LUT row with it; pass 2 (fibre-major) sums each fibre's products in
//     ascending voxel order.  EC and ISO columns are owned by a single voxel and
//     are written directly by pass 1.
//
// Entry points are declared in include/commit_hip.h.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/commit_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(COMMIT_HIP_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kMaxNJ = 8;            // nS <= 512 samples per voxel
constexpr int kMaxRF = 20;           // the reference's own limit per compartment

// ------------------------------------------------------------------ device --
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct DevDict {
    // IC, voxel-major CSR
    const uint32_t *ic_ptr;   // [nVOX+1]
    const uint32_t *ic_f;     // [n]
    const uint16_t *ic_o;     // [n]
    const float *ic_l;        // [n]
    const uint32_t *eval;     // [nF*nIC]
    // fibre-major view of the same segments
    const uint32_t *fib_ptr;  // [nF+1]
    const uint32_t *fib_perm; // [n] voxel-major position of each fibre-major segment
    // EC, voxel-major CSR
    const uint32_t *ec_ptr;   // [nVOX+1]
    const uint32_t *ec_e;     // [nE] original peak index (x column)
    const uint16_t *ec_o;     // [nE]
    // ISO, voxel-major CSR
    const uint32_t *iso_ptr;  // [nVOX+1]
    const uint32_t *iso_k;    // [nV] original ISO index (x column)
    // LUTs (float32, row-major [rf][dir][s])
    const float *wmr, *wmh, *iso;
    int nS, ndirs, nF, n, nE, nI, nVOX;
    int nIC, nEC, nISO;
};

// Accumulate, for every sample owned by this lane, sum_t c_t * row_t[s] over the
// (wave-uniform) records t < cnt.  c and o live one record per lane.
template <int NJ>
__device__ __forceinline__ void accumulate_rows(double (&acc)[NJ], double c, int o, int cnt,
                                                const float *__restrict__ lut, int nS, int lane,
                                                bool same_row)
{
    for (int t = 0; t < cnt; t++) {
        const double ct = readlane_f64(c, t);
        if (ct != 0.0) {
            const int ot = same_row ? 0 : __builtin_amdgcn_readlane(o, t);
            const float *row = lut + (size_t)ot * nS;
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int s = lane + kWave * j;
                if (s < nS) acc[j] = fma(ct, (double)row[s], acc[j]);
            }
        }
    }
}

// y = A x, FP64.  One wavefront per voxel; lanes own samples lane + 64 j.
template <int NJ>
__global__ __launch_bounds__(kBlock) void k_A_fp64(DevDict d, const double *__restrict__ x,
                                                    double *__restrict__ y)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int v = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (v >= d.nVOX) return;
    double acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = 0.0;
    const size_t lut_stride = (size_t)d.ndirs * d.nS;

    // intra-cellular: c = l * x[f + i nF] * eval[f + i nF]    (setup_operator.py:62-103)
    {
        const uint32_t b = d.ic_ptr[v], e = d.ic_ptr[v + 1];
        for (uint32_t base = b; base < e; base += kWave) {
            const uint32_t k = base + lane;
            const int cnt = (int)min((uint32_t)kWave, e - base);
            uint32_t f = 0;
            int o = 0;
            double l = 0.0;
            if (k < e) {
                f = d.ic_f[k];
                o = d.ic_o[k];
                l = (double)d.ic_l[k];
            }
            for (int i = 0; i < d.nIC; i++) {
                double c = 0.0;
                if (k < e) {
                    const size_t col = f + (size_t)i * d.nF;
                    c = l * (x[col] * (double)d.eval[col]);
                }
                accumulate_rows<NJ>(acc, c, o, cnt, d.wmr + i * lut_stride, d.nS, lane, false);
            }
        }
    }
    // extra-cellular: c = x[nIC nF + i nE + e]                  (setup_operator.py:111-165)
    if (d.nEC > 0) {
        const uint32_t b = d.ec_ptr[v], e = d.ec_ptr[v + 1];
        const double *xe = x + (size_t)d.nIC * d.nF;
        for (uint32_t base = b; base < e; base += kWave) {
            const uint32_t k = base + lane;
            const int cnt = (int)min((uint32_t)kWave, e - base);
            uint32_t eo = 0;
            int o = 0;
            if (k < e) {
                eo = d.ec_e[k];
                o = d.ec_o[k];
            }
            for (int i = 0; i < d.nEC; i++) {
                const double c = (k < e) ? xe[eo + (size_t)i * d.nE] : 0.0;
                accumulate_rows<NJ>(acc, c, o, cnt, d.wmh + i * lut_stride, d.nS, lane, false);
            }
        }
    }
    // isotropic: c = x[nIC nF + nEC nE + i nV + k]               (setup_operator.py:167-218)
    if (d.nISO > 0) {
        const uint32_t b = d.iso_ptr[v], e = d.iso_ptr[v + 1];
        const double *xi = x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE;
        for (uint32_t base = b; base < e; base += kWave) {
            const uint32_t k = base + lane;
            const int cnt = (int)min((uint32_t)kWave, e - base);
            const uint32_t ko = (k < e) ? d.iso_k[k] : 0u;
            for (int i = 0; i < d.nISO; i++) {
                const double c = (k < e) ? xi[ko + (size_t)i * d.nI] : 0.0;
                accumulate_rows<NJ>(acc, c, 0, cnt, d.iso + (size_t)i * d.nS, d.nS, lane, true);
            }
        }
    }
    double *yv = y + (size_t)v * d.nS;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int s = lane + kWave * j;
        if (s < d.nS) yv[s] = acc[j];
    }
}

// Dot product of a float32 LUT row with the voxel's y row staged in LDS.
__device__ __forceinline__ double row_dot(const float *__restrict__ row, const double *yr, int nS,
                                          bool vec4)
{
    double acc = 0.0;
    if (vec4) {
        const float4 *r4 = reinterpret_cast<const float4 *>(row);
        const double2 *y2 = reinterpret_cast<const double2 *>(yr);
        for (int q = 0; q < (nS >> 2); q++) {
            const float4 a = r4[q];
            const double2 b0 = y2[2 * q], b1 = y2[2 * q + 1];
            acc = fma((double)a.x, b0.x, acc);
            acc = fma((double)a.y, b0.y, acc);
            acc = fma((double)a.z, b1.x, acc);
            acc = fma((double)a.w, b1.y, acc);
        }
    } else {
        for (int s = 0; s < nS; s++) acc = fma((double)row[s], yr[s], acc);
    }
    return acc;
}

// A^T y pass 1, FP64.  One wavefront per voxel; the voxel's y row is staged in
// LDS and each lane takes one segment (peak, ISO entry) of the voxel.
__global__ __launch_bounds__(kBlock) void k_At_voxel_fp64(DevDict d, const double *__restrict__ y,
                                                          double *__restrict__ z,
                                                          double *__restrict__ x, int nSp)
{
    extern __shared__ double ylds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x >> 6;
    const int v = blockIdx.x * kWavesPerBlock + w;
    const bool valid = v < d.nVOX;
    double *yr = ylds + (size_t)w * nSp;
    if (valid)
        for (int s = lane; s < d.nS; s += kWave) yr[s] = y[(size_t)v * d.nS + s];
    __syncthreads();
    if (!valid) return;
    const size_t lut_stride = (size_t)d.ndirs * d.nS;
    const bool vec4 = (d.nS & 3) == 0;

    // intra-cellular: z[i n + k] = l * (LUT_IC_i[o] . y[v]) * eval   (setup_operator.py:412-428)
    {
        const uint32_t b = d.ic_ptr[v], e = d.ic_ptr[v + 1];
        for (uint32_t k = b + lane; k < e; k += kWave) {
            const uint32_t f = d.ic_f[k];
            const int o = d.ic_o[k];
            const double l = (double)d.ic_l[k];
            for (int i = 0; i < d.nIC; i++) {
                const double dot = row_dot(d.wmr + i * lut_stride + (size_t)o * d.nS, yr, d.nS, vec4);
                z[(size_t)i * d.n + k] = l * dot * (double)d.eval[f + (size_t)i * d.nF];
            }
        }
    }
    // extra-cellular: x[nIC nF + i nE + e] = LUT_EC_i[o_e] . y[v]       (setup_operator.py:444-504)
    if (d.nEC > 0) {
        const uint32_t b = d.ec_ptr[v], e = d.ec_ptr[v + 1];
        double *xe = x + (size_t)d.nIC * d.nF;
        for (uint32_t k = b + lane; k < e; k += kWave) {
            const uint32_t eo = d.ec_e[k];
            const int o = d.ec_o[k];
            for (int i = 0; i < d.nEC; i++)
                xe[eo + (size_t)i * d.nE] = row_dot(d.wmh + i * lut_stride + (size_t)o * d.nS, yr, d.nS, vec4);
        }
    }
    // isotropic: x[nIC nF + nEC nE + i nV + k] = LUT_ISO_i . y[v]       (setup_operator.py:506-563)
    if (d.nISO > 0) {
        const uint32_t b = d.iso_ptr[v], e = d.iso_ptr[v + 1];
        double *xi = x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE;
        for (uint32_t k = b + lane; k < e; k += kWave) {
            const uint32_t ko = d.iso_k[k];
            for (int i = 0; i < d.nISO; i++)
                xi[ko + (size_t)i * d.nI] = row_dot(d.iso + (size_t)i * d.nS, yr, d.nS, vec4);
        }
    }
}

// A^T y pass 2: x[f + i nF] = sum over the fibre's segments, ascending voxel order.
__global__ __launch_bounds__(kBlock) void k_At_fibre_fp64(DevDict d, const double *__restrict__ z,
                                                          double *__restrict__ x)
{
    const size_t total = (size_t)d.nF * d.nIC;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(g / d.nF);
        const uint32_t f = (uint32_t)(g - (size_t)i * d.nF);
        const double *zi = z + (size_t)i * d.n;
        double acc = 0.0;
        for (uint32_t k = d.fib_ptr[f]; k < d.fib_ptr[f + 1]; k++) acc += zi[d.fib_perm[k]];
        x[g] = acc;
    }
}

__global__ void k_zero_fp64(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---------------------------------------------------------------- host side --

// Stable counting sort of keys in [0, nkeys): perm lists input positions in key
// order (ties keep input order, like numpy's kind='mergesort'); ptr is the CSR.
bool counting_sort(const uint32_t *keys, size_t n, uint32_t nkeys, std::vector<uint32_t> &perm,
                   std::vector<uint32_t> &ptr)
{
    ptr.assign((size_t)nkeys + 1, 0u);
    for (size_t i = 0; i < n; i++) {
        if (keys[i] >= nkeys) return false;
        ptr[keys[i] + 1]++;
    }
    for (size_t k = 0; k < nkeys; k++) ptr[k + 1] += ptr[k];
    std::vector<uint32_t> pos(ptr.begin(), ptr.end() - 1);
    perm.resize(n);
    for (size_t i = 0; i < n; i++) perm[pos[keys[i]]++] = (uint32_t)i;
    return true;
}

}  // namespace

struct commit_hip_plan {
    int device = 0;
    int precision = COMMIT_HIP_FP64;
    int nolut = 0;
    int nS_ref = 0;            // nS as the reference reports it (rows = nVOX * nS_ref)
    DevDict d{};
    std::vector<void *> allocs;
    double *zbuf = nullptr;    // [nIC][n] per-segment products of A^T pass 1
    int64_t bytes = 0;
};

namespace {

template <typename T>
int upload(commit_hip_plan *p, const T *host, size_t count, const T **dst)
{
    *dst = nullptr;
    if (count == 0) return 0;
    void *ptr = nullptr;
    if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess)
        return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
    p->allocs.push_back(ptr);
    p->bytes += (int64_t)(count * sizeof(T));
    HIP_TRY(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<const T *>(ptr);
    return 0;
}

void release(commit_hip_plan *p)
{
    for (void *a : p->allocs) (void)hipFree(a);
    p->allocs.clear();
}

int launch_A(commit_hip_plan *p, const double *x, double *y, hipStream_t st)
{
    const DevDict &d = p->d;
    const int nj = (d.nS + kWave - 1) / kWave;
    const unsigned grid = (unsigned)((d.nVOX + kWavesPerBlock - 1) / kWavesPerBlock);
    if (grid == 0) return 0;
    switch (nj) {
#define CASE(NJ) \
    case NJ: hipLaunchKernelGGL(k_A_fp64<NJ>, dim3(grid), dim3(kBlock), 0, st, d, x, y); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds %d", d.nS, kMaxNJ * kWave);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_At(commit_hip_plan *p, const double *y, double *x, hipStream_t st)
{
    const DevDict &d = p->d;
    const int nSp = (d.nS + 1) & ~1;   // keep each wave's row 16-byte aligned
    const unsigned grid = (unsigned)((d.nVOX + kWavesPerBlock - 1) / kWavesPerBlock);
    if (grid > 0) {
        const size_t shm = (size_t)kWavesPerBlock * nSp * sizeof(double);
        hipLaunchKernelGGL(k_At_voxel_fp64, dim3(grid), dim3(kBlock), shm, st, d, y, p->zbuf, x, nSp);
        HIP_TRY(hipGetLastError());
    }
    const size_t total = (size_t)d.nF * d.nIC;
    if (total > 0) {
        unsigned g2 = (unsigned)std::min<size_t>((total + kBlock - 1) / kBlock, 65536);
        hipLaunchKernelGGL(k_At_fibre_fp64, dim3(g2), dim3(kBlock), 0, st, d, p->zbuf, x);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace

extern "C" {

const char *commit_hip_last_error(void) { return g_err.c_str(); }

const char *commit_hip_version(void) { return "commit_hip 0.1 (gfx950)"; }

int commit_hip_plan_create(commit_hip_plan **plan, int nS, int ndirs, int nF, int n, int nE, int nV,
                           int nVOX, const uint32_t *ICf, const uint32_t *ICeval, const uint32_t *ICv,
                           const uint16_t *ICo, const float *ICl, const uint32_t *ECv,
                           const uint16_t *ECo, const uint32_t *ISOv, const float *wmrSFP,
                           const float *wmhSFP, const float *isoSFP, uint32_t nIC, uint32_t nEC,
                           uint32_t nISO, int nolut, int precision, int device)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "plan pointer is NULL");
    *plan = nullptr;
    if (precision != COMMIT_HIP_FP64)
        return fail(COMMIT_HIP_EUNSUPPORTED, "precision %d is not built in this library", precision);
    if (nS < 1 || ndirs < 1 || nF < 0 || n < 0 || nE < 0 || nV < 0 || nVOX < 0)
        return fail(COMMIT_HIP_EINVAL, "negative or zero size (nS=%d ndirs=%d nF=%d n=%d nE=%d nV=%d nVOX=%d)",
                    nS, ndirs, nF, n, nE, nV, nVOX);
    if (nIC > (uint32_t)kMaxRF || nEC > (uint32_t)kMaxRF || nISO > (uint32_t)kMaxRF)
        return fail(COMMIT_HIP_EUNSUPPORTED, "more than %d response functions in a compartment", kMaxRF);
    if (nolut && (nIC != 1 || nEC != 0 || nISO > 1))
        return fail(COMMIT_HIP_EUNSUPPORTED, "LUT-free operator needs nIC=1, nEC=0, nISO<=1");
    if (!nolut && nS > kMaxNJ * kWave)
        return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds %d", nS, kMaxNJ * kWave);
    if ((n > 0 && nIC == 0) || (n > 0 && (!ICf || !ICv || !ICo || !ICl || !ICeval)))
        return fail(COMMIT_HIP_EINVAL, "IC arrays missing");
    if (nEC > 0 && nE > 0 && (!ECv || !ECo)) return fail(COMMIT_HIP_EINVAL, "EC arrays missing");
    if (nISO > 0 && nV > 0 && !ISOv) return fail(COMMIT_HIP_EINVAL, "ISO arrays missing");
    if (!nolut && ((nIC && !wmrSFP) || (nEC && !wmhSFP) || (nISO && !isoSFP)))
        return fail(COMMIT_HIP_EINVAL, "LUT arrays missing");
    for (int k = 0; k < n; k++) {
        if ((int64_t)ICf[k] >= nF) return fail(COMMIT_HIP_EINVAL, "ICf[%d]=%u >= nF=%d", k, ICf[k], nF);
        if (!nolut && (int)ICo[k] >= ndirs) return fail(COMMIT_HIP_EINVAL, "ICo[%d]=%u >= ndirs=%d", k, ICo[k], ndirs);
    }
    if (nEC > 0)
        for (int e = 0; e < nE; e++)
            if ((int)ECo[e] >= ndirs) return fail(COMMIT_HIP_EINVAL, "ECo[%d]=%u >= ndirs=%d", e, ECo[e], ndirs);

    HIP_TRY(hipSetDevice(device));
    commit_hip_plan *p = new commit_hip_plan();
    p->device = device;
    p->precision = precision;
    p->nolut = nolut;
    p->nS_ref = nS;
    DevDict &d = p->d;
    d.nS = nolut ? 1 : nS;
    d.ndirs = nolut ? 1 : ndirs;
    d.nF = nF;
    d.n = n;
    d.nE = (nEC > 0) ? nE : 0;
    d.nI = (nISO > 0) ? nV : 0;
    d.nVOX = nVOX;
    d.nIC = (int)nIC;
    d.nEC = (int)nEC;
    d.nISO = (int)nISO;
    int rc = 0;
    {
        // IC: voxel-major CSR (stable), then the fibre-major permutation of it
        std::vector<uint32_t> perm, ptr;
        if (!counting_sort(ICv, (size_t)n, (uint32_t)nVOX, perm, ptr)) {
            rc = fail(COMMIT_HIP_EINVAL, "ICv holds a voxel index >= nVOX=%d", nVOX);
            goto done;
        }
        std::vector<uint32_t> f(n);
        std::vector<uint16_t> o(n);
        std::vector<float> l(n);
        for (int k = 0; k < n; k++) {
            f[k] = ICf[perm[k]];
            o[k] = nolut ? 0 : ICo[perm[k]];
            l[k] = ICl[perm[k]];
        }
        if ((rc = upload(p, ptr.data(), ptr.size(), &d.ic_ptr))) goto done;
        if ((rc = upload(p, f.data(), (size_t)n, &d.ic_f))) goto done;
        if ((rc = upload(p, o.data(), (size_t)n, &d.ic_o))) goto done;
        if ((rc = upload(p, l.data(), (size_t)n, &d.ic_l))) goto done;
        std::vector<uint32_t> fperm, fptr;
        counting_sort(f.data(), (size_t)n, (uint32_t)nF, fperm, fptr);
        if ((rc = upload(p, fptr.data(), fptr.size(), &d.fib_ptr))) goto done;
        if ((rc = upload(p, fperm.data(), (size_t)n, &d.fib_perm))) goto done;
        if ((rc = upload(p, ICeval, (size_t)nF * nIC, &d.eval))) goto done;
    }
    {
        // EC: voxel-major CSR; ec_e keeps the original peak index (x column)
        std::vector<uint32_t> perm, ptr;
        const size_t ne = (size_t)d.nE;
        if (!counting_sort(ne ? ECv : nullptr, ne, (uint32_t)nVOX, perm, ptr)) {
            rc = fail(COMMIT_HIP_EINVAL, "ECv holds a voxel index >= nVOX=%d", nVOX);
            goto done;
        }
        std::vector<uint16_t> o(ne);
        for (size_t k = 0; k < ne; k++) o[k] = ECo[perm[k]];
        if ((rc = upload(p, ptr.data(), ptr.size(), &d.ec_ptr))) goto done;
        if ((rc = upload(p, perm.data(), ne, &d.ec_e))) goto done;
        if ((rc = upload(p, o.data(), ne, &d.ec_o))) goto done;
    }
    {
        std::vector<uint32_t> perm, ptr;
        const size_t ni = (size_t)d.nI;
        if (!counting_sort(ni ? ISOv : nullptr, ni, (uint32_t)nVOX, perm, ptr)) {
            rc = fail(COMMIT_HIP_EINVAL, "ISOv holds a voxel index >= nVOX=%d", nVOX);
            goto done;
        }
        if ((rc = upload(p, ptr.data(), ptr.size(), &d.iso_ptr))) goto done;
        if ((rc = upload(p, perm.data(), ni, &d.iso_k))) goto done;
    }
    if (nolut) {
        const float one = 1.0f;
        if ((rc = upload(p, &one, 1, &d.wmr))) goto done;
        if (nISO) { if ((rc = upload(p, &one, 1, &d.iso))) goto done; }
    } else {
        const size_t lut = (size_t)ndirs * nS;
        if (nIC && (rc = upload(p, wmrSFP, lut * nIC, &d.wmr))) goto done;
        if (nEC && (rc = upload(p, wmhSFP, lut * nEC, &d.wmh))) goto done;
        if (nISO && (rc = upload(p, isoSFP, (size_t)nS * nISO, &d.iso))) goto done;
    }
    if (n > 0 && nIC > 0) {
        void *z = nullptr;
        const size_t zb = (size_t)n * nIC * sizeof(double);
        if (hipMalloc(&z, zb) != hipSuccess) {
            rc = fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", zb);
            goto done;
        }
        p->allocs.push_back(z);
        p->bytes += (int64_t)zb;
        p->zbuf = static_cast<double *>(z);
    }
done:
    if (rc) {
        release(p);
        delete p;
        return rc;
    }
    *plan = p;
    return 0;
}

int commit_hip_plan_destroy(commit_hip_plan *plan)
{
    if (!plan) return 0;
    (void)hipSetDevice(plan->device);
    release(plan);
    delete plan;
    return 0;
}

int commit_hip_plan_set_eval(commit_hip_plan *plan, const uint32_t *ICeval)
{
    if (!plan || !ICeval) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    const size_t cnt = (size_t)plan->d.nF * plan->d.nIC;
    if (cnt == 0) return 0;
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipMemcpy(const_cast<uint32_t *>(plan->d.eval), ICeval, cnt * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    return 0;
}

int commit_hip_plan_shape(const commit_hip_plan *plan, int64_t *nrows, int64_t *ncols)
{
    if (!plan || !nrows || !ncols) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    const DevDict &d = plan->d;
    *nrows = (int64_t)d.nVOX * plan->nS_ref;
    *ncols = (int64_t)d.nF * d.nIC + (int64_t)d.nE * d.nEC + (int64_t)d.nI * d.nISO;
    return 0;
}

int64_t commit_hip_plan_device_bytes(const commit_hip_plan *plan) { return plan ? plan->bytes : -1; }

int commit_hip_A(commit_hip_plan *plan, const void *x, void *y, void *stream)
{
    if (!plan || !x || !y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *yd = static_cast<double *>(y);
    int rc = launch_A(plan, static_cast<const double *>(x), yd, st);
    if (rc) return rc;
    if (plan->nolut && plan->nS_ref > 1) {
        // LUT-free rows are indexed by voxel only (setup_operator.py:726-729);
        // the remaining rows of the reference-sized output stay zero.
        const size_t tail = (size_t)plan->d.nVOX * (plan->nS_ref - 1);
        hipLaunchKernelGGL(k_zero_fp64, dim3(1024), dim3(256), 0, st, yd + plan->d.nVOX, tail);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int commit_hip_At(commit_hip_plan *plan, const void *y, void *x, void *stream)
{
    if (!plan || !x || !y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    HIP_TRY(hipSetDevice(plan->device));
    return launch_At(plan, static_cast<const double *>(y), static_cast<double *>(x),
                     static_cast<hipStream_t>(stream));
}

}  // extern "C"
