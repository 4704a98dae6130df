// libcommit_hip.so — MI355X (gfx950, CDNA4) implementation of the COMMIT
// implicit operator: y = A x and x = A^T y over the IC / EC / ISO compartments.
//
// Reference algorithm: setup_operator.py:48-915 (generated COMMIT_A/COMMIT_At
// and their LUT-free variants), driven from operator.pyx:114-184.  The design
// here is GPU-first and shares none of the reference's structure.
//
// Device layout (built once per plan, resident in HBM)
//   * every compartment (IC segments, EC peaks, ISO voxels) as a voxel-major
//     CSR: entry ranges per voxel, entries in the order load_dictionary leaves
//     them (stable by voxel, core.pyx:509-513); entry = orientation (u16),
//     x column (u32), length (f32, IC only) — 10 bytes per IC segment;
//   * a fibre-major permutation of the IC entries (ascending voxel inside each
//     fibre) for the adjoint's second pass;
//   * the LUTs widened to 64*NJ samples per row with zeros, NJ = ceil(nS/64),
//     so a row is NJ unpredicated coalesced loads.
//
// Kernels: one wavefront per voxel, the voxel's samples spread over the 64
// lanes (NJ per lane).  Optionally (plan option, chosen by the autotuner)
// work-groups are renumbered so that each XCD (own L2) processes one
// contiguous voxel range; on the 1M-fibre head the hardware order is faster.
//   A x   : NJ FP64 accumulators per lane.  Entry records are read one per
//           lane and walked in order through a 3-stage software pipeline:
//           while the row of entry t is applied, the rows of t+1 and t+2 are in
//           flight.  Zero coefficients are dropped with a ballot.
//   A^T y : pass 1 keeps the voxel's y row in registers; each entry's per-lane
//           partial dot product goes to an LDS ring and every 8 entries are
//           reduced together (transposed read + 3 shuffles).  IC products go
//           to a scratch vector in entry order; EC and ISO columns belong to a
//           single voxel and are written straight to x.  Pass 2 sums every
//           fibre's products in ascending voxel order (the order the
//           reference's owning thread uses) — no atomics, bitwise repeatable.
//
// Entry points are declared in include/commit_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/commit_hip.h"
#include "common.h"
#include "plan.h"
#include "rows.h"

namespace commit_internal {
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
}  // namespace commit_internal

using commit_internal::fail;

namespace {

constexpr int kMaxNJ = 8;            // nS <= 512 samples per voxel
constexpr int kMaxRF = 20;           // the reference's own limit per compartment
using commit_internal::Comp;
using commit_internal::DevDict;

// Voxel of this wavefront (one wavefront per voxel).
__device__ __forceinline__ int voxel_id(const DevDict &d) { return commit_internal::wave_index(d.xcd_remap); }

// xic: the IC block of x, or x*eval when eval is not all ones.
// T: storage type of x and y (double, or float for the FP32 mode); the
// accumulation is FP64 in both (exact products of the FP32 LUT entries).
template <int NJ, typename T>
__global__ __launch_bounds__(kBlock) void k_A(DevDict d, const T *__restrict__ x, const T *__restrict__ xic,
                                              AOut<T> out)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int v = voxel_id(d);
    if (v >= d.nVOX) return;
    double acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = 0.0;
    const size_t lut_rf = (size_t)d.ndirs * d.nSp;
    const bool tail = Lanes<NJ, kQuads<T>>::tail(lane, d.nS);
    a_comp<NJ>(acc, d.ic, d.ic.ptr[v], d.ic.ptr[v + 1], xic, (size_t)d.nF, d.nIC, d.wmr, lut_rf, d.nSp, lane, tail);
    if (d.nEC > 0)
        a_comp<NJ>(acc, d.ec, d.ec.ptr[v], d.ec.ptr[v + 1], x + (size_t)d.nIC * d.nF, (size_t)d.nE, d.nEC,
                   d.wmh, lut_rf, d.nSp, lane, tail);
    if (d.nISO > 0)
        a_comp<NJ>(acc, d.iso, d.iso.ptr[v], d.iso.ptr[v + 1], x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE,
                   (size_t)d.nI, d.nISO, d.isol, (size_t)d.nSp, 0, lane, tail);
    a_store<NJ, T>(acc, d, v, out, lane);
}

// x*eval for the IC block (only when eval is not all ones)
template <typename T>
__global__ void k_x_eval(const T *__restrict__ x, const uint32_t *__restrict__ ev, T *__restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = ev[i] ? x[i] : (T)0;
}

// ------------------------------------------------------------ x = A^T y -----
// Per-wave LDS ring of per-lane partial dot products, kRing entries x 64 lanes;
// rows padded so that the transposed reads of a flush are bank-conflict free.
constexpr int kRing = 8;
constexpr int kRingStride = 66;     // doubles per row (64 lanes + pad, 16-byte aligned)

template <int NJ>
__device__ __forceinline__ double dot_row(const float (&raw)[NJ], const double (&yv)[NJ])
{
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NJ; j++) a = fma((double)raw[j], yv[j], a);
    return a;
}

// Reduce ring entries [tb, tb+8) over the 64 lanes; store the ones < cnt:
//   IC     : z[i*m + zpos] = l * dot   (the entry's fibre-major position, so that
//            pass 2 reads each fibre's products contiguously; eval is applied
//            per fibre in pass 2)
//   EC/ISO : out[col + i*xs] = dot
// colk holds zpos for IC entries, the x column otherwise.
template <typename T>
__device__ __forceinline__ void flush8(const double *ring, int tb, int cnt, int lane, double l,
                                       uint32_t colk, bool ic, uint32_t m, double *__restrict__ z, T *__restrict__ out,
                                       size_t xs, int i)
{
    const int te = tb + (lane >> 3);
    const double2 *row = reinterpret_cast<const double2 *>(ring + (size_t)(te % kRing) * kRingStride + 8 * (lane & 7));
    double q = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const double2 v = row[u];
        q += v.x;
        q += v.y;
    }
    q += __shfl_xor(q, 1);
    q += __shfl_xor(q, 2);
    q += __shfl_xor(q, 4);
    const int src = te < kWave ? te : 0;
    const double le = __shfl(l, src);
    const uint32_t ce = (uint32_t)__shfl((int)colk, src);
    if ((lane & 7) == 0 && te < cnt) {
        if (ic)
            z[(size_t)i * m + ce] = le * q;                     // (setup_operator.py:412-428)
        else
            out[(size_t)ce + (size_t)i * xs] = (T)q;            // (setup_operator.py:444-563)
    }
}

// One compartment of the voxel: per entry a partial dot product into the ring,
// flushed every 8 entries.  Rows are pipelined as in a_comp.
template <int NJ, typename T>
__device__ __forceinline__ void at_comp(const double (&yv)[NJ], const Comp &cp, const uint32_t *__restrict__ zpos,
                                        uint32_t b, uint32_t e, double *__restrict__ z, T *__restrict__ out,
                                        size_t xs, int nrf,
                                        const float *__restrict__ lut, size_t lut_rf, int nSp, int lane,
                                        bool tail, double *ring)
{
    const bool ic = cp.len != nullptr;
    for (int i = 0; i < nrf; i++) {
        const float *L = lut + (size_t)i * lut_rf;
        for (uint32_t base = b; base < e; base += kWave) {
            const uint32_t k = base + lane;
            const int cnt = (int)min((uint32_t)kWave, e - base);
            uint32_t o = 0, colk = 0;
            double l = 0.0;
            if (k < e) {
                if (cp.o) o = stream_load(cp.o + k);
                colk = stream_load((ic ? zpos : cp.col) + k);
                if (ic) l = (double)stream_load(cp.len + k);
            }
            float ra[NJ], rb[NJ], rc[NJ];
            int t = 0;
            fetch_row<NJ, kQuads<T>>(ra, L + (size_t)__builtin_amdgcn_readlane((int)o, 0) * nSp, lane, tail);
            if (cnt > 1) fetch_row<NJ, kQuads<T>>(rb, L + (size_t)__builtin_amdgcn_readlane((int)o, 1) * nSp, lane, tail);
#define AT_STAGE(CUR, NEXT2)                                                                              \
    {                                                                                                     \
        if (t + 2 < cnt) fetch_row<NJ, kQuads<T>>(NEXT2, L + (size_t)__builtin_amdgcn_readlane((int)o, t + 2) * nSp, lane, tail); \
        ring[(size_t)(t % kRing) * kRingStride + lane] = dot_row<NJ>(CUR, yv);                            \
        if ((t & 7) == 7) {                                                                               \
            __builtin_amdgcn_wave_barrier();                                                              \
            flush8(ring, t - 7, cnt, lane, l, colk, ic, cp.m, z, out, xs, i);                             \
            __builtin_amdgcn_wave_barrier();                                                              \
        }                                                                                                 \
        if (++t >= cnt) break;                                                                            \
    }
            while (true) {
                AT_STAGE(ra, rc)
                AT_STAGE(rb, ra)
                AT_STAGE(rc, rb)
            }
#undef AT_STAGE
            if (cnt & 7) {                        // last partial group of 8
                __builtin_amdgcn_wave_barrier();
                flush8(ring, cnt & ~7, cnt, lane, l, colk, ic, cp.m, z, out, xs, i);
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

template <int NJ, typename T>
__global__ __launch_bounds__(kBlock) void k_At(DevDict d, const T *__restrict__ y, const T *__restrict__ sqrtw,
                                               double *__restrict__ z, T *__restrict__ x)
{
    __shared__ __align__(16) double rings[kWavesPerBlock][kRing * kRingStride];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int v = voxel_id(d);
    if (v >= d.nVOX) return;
    double *ring = rings[w];
    double yv[NJ];
    const T *yr = y + (size_t)v * d.nS;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int s = Lanes<NJ, kQuads<T>>::sample(lane, j);
        yv[j] = (s < d.nS) ? (double)stream_load(yr + s) : 0.0;
        if (sqrtw && s < d.nS) yv[j] *= (double)stream_load(sqrtw + (size_t)v * d.nS + s);   // (solvers.py:267)
    }
    const size_t lut_rf = (size_t)d.ndirs * d.nSp;
    const bool tail = Lanes<NJ, kQuads<T>>::tail(lane, d.nS);
    at_comp<NJ>(yv, d.ic, d.zpos, d.ic.ptr[v], d.ic.ptr[v + 1], z, x, 0, d.nIC, d.wmr, lut_rf, d.nSp, lane, tail,
                ring);
    if (d.nEC > 0)
        at_comp<NJ>(yv, d.ec, d.zpos, d.ec.ptr[v], d.ec.ptr[v + 1], z, x + (size_t)d.nIC * d.nF, (size_t)d.nE,
                    d.nEC, d.wmh, lut_rf, d.nSp, lane, tail, ring);
    if (d.nISO > 0)
        at_comp<NJ>(yv, d.iso, d.zpos, d.iso.ptr[v], d.iso.ptr[v + 1], z,
                    x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE, (size_t)d.nI, d.nISO, d.isol, (size_t)d.nSp, 0,
                    lane, tail, ring);
}

// A^T y pass 2: x[f + i nF] = eval * sum of the fibre's products (pass 1 stored
// them fibre-major, ascending voxel, so the run is contiguous).  16 lanes (one
// DPP row) per (fibre, RF): lane j sums entries j, j+16, ... in ascending
// order, 4 loads in flight, and the 16 partials are combined in a fixed tree
// of row moves, so each step of the group reads one contiguous 128-byte line
// and the result is bitwise repeatable.  eval is 0 or 1, so applying it once
// per fibre equals the reference's per-segment product (setup_operator.py:428).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_At_fibre(DevDict d, const double *__restrict__ z, T *__restrict__ x)
{
    const size_t total = (size_t)d.nF * d.nIC;
    const uint32_t j = threadIdx.x & 15;
    const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 4;
    for (size_t g = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; g < total; g += stride) {
        const int i = (int)(g / d.nF);
        const uint32_t f = (uint32_t)(g - (size_t)i * d.nF);
        const double *zi = z + (size_t)i * d.n;
        const uint32_t kb = d.fib_ptr[f], ke = d.fib_ptr[f + 1];
        double acc = 0.0;
        uint32_t k = kb + j;
        for (; k + 48 < ke; k += 64) {
            const double a = stream_load(zi + k), b = stream_load(zi + k + 16);
            const double c = stream_load(zi + k + 32), e = stream_load(zi + k + 48);
            acc += a;
            acc += b;
            acc += c;
            acc += e;
        }
        for (; k < ke; k += 16) acc += stream_load(zi + k);
        acc += dpp_f64<kDppQuadXor1>(acc);          // all 16 lanes of the group are active here
        acc += dpp_f64<kDppQuadXor2>(acc);
        acc += dpp_f64<kDppHalfMirror>(acc);
        acc += dpp_f64<kDppRowMirror>(acc);
        if (j == 0) x[g] = (T)(acc * (double)d.eval[g]);
    }
}

template <typename T>
__global__ void k_zero(T *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (T)0;
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

// LUT with rows widened from nS to nSp samples (zeros)
std::vector<float> pad_rows(const float *src, size_t rows, int nS, int nSp)
{
    std::vector<float> out(rows * (size_t)nSp, 0.0f);
    for (size_t r = 0; r < rows; r++) std::memcpy(&out[r * nSp], src + r * nS, (size_t)nS * sizeof(float));
    return out;
}

}  // namespace

int commit_internal::dev_alloc(commit_hip_plan *p, size_t bytes, void **out)
{
    *out = nullptr;
    if (bytes == 0) return 0;
    if (hipMalloc(out, bytes) != hipSuccess)
        return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    p->allocs.push_back(*out);
    p->bytes += (int64_t)bytes;
    return 0;
}

namespace {

using commit_internal::dev_alloc;

template <typename T>
int upload(commit_hip_plan *p, const T *host, size_t count, const T **dst)
{
    *dst = nullptr;
    void *ptr = nullptr;
    int rc = dev_alloc(p, count * sizeof(T), &ptr);
    if (rc || !ptr) return rc;
    HIP_TRY(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<const T *>(ptr);
    return 0;
}

void release(commit_hip_plan *p)
{
    for (void *a : p->allocs) (void)hipFree(a);
    p->allocs.clear();
}

unsigned vox_blocks(const DevDict &d) { return (unsigned)((d.nVOX + kWavesPerBlock - 1) / kWavesPerBlock); }

template <typename T>
int launch_A(commit_hip_plan *p, const T *x, AOut<T> out, hipStream_t st)
{
    const DevDict &d = p->d;
    const unsigned grid = vox_blocks(d);
    if (grid == 0) return 0;
    const T *xic = x;
    if (!p->eval_ones && d.nF > 0 && d.nIC > 0) {
        const size_t n = (size_t)d.nF * d.nIC;
        T *xe = reinterpret_cast<T *>(p->xic);
        hipLaunchKernelGGL(k_x_eval<T>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, st,
                           x, d.eval, xe, n);
        HIP_TRY(hipGetLastError());
        xic = xe;
    }
    if (p->a_dense) {
        const int rc = commit_internal::launch_A_dense<T>(p, x, xic, out.y, out.yobs, out.sqrtw, out.rpart, st);
        if (rc != COMMIT_HIP_EUNSUPPORTED) return rc;
    }
    if (p->slab) {
        const int rc = commit_internal::launch_A_slab<T>(p, x, xic, out, st);
        if (rc != COMMIT_HIP_EUNSUPPORTED) return rc;
    }
    switch (d.nSp / kWave) {
#define CASE(NJ) \
    case NJ: hipLaunchKernelGGL((k_A<NJ, T>), dim3(grid), dim3(kBlock), 0, st, d, x, xic, out); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds %d", d.nS, kMaxNJ * kWave);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
int launch_At(commit_hip_plan *p, const T *y, const T *sqrtw, T *x, hipStream_t st, int pass = 3)
{
    const DevDict &d = p->d;
    const unsigned grid = vox_blocks(d);
    int slab_rc = COMMIT_HIP_EUNSUPPORTED;
    if (grid > 0 && (pass & 1) && p->slab) {
        slab_rc = commit_internal::launch_At_slab<T>(p, y, sqrtw, x, st);
        if (slab_rc && slab_rc != COMMIT_HIP_EUNSUPPORTED) return slab_rc;
    }
    if (grid > 0 && (pass & 1) && slab_rc == COMMIT_HIP_EUNSUPPORTED) {
        switch (d.nSp / kWave) {
#define CASE(NJ) \
    case NJ: hipLaunchKernelGGL((k_At<NJ, T>), dim3(grid), dim3(kBlock), 0, st, d, y, sqrtw, p->zbuf, x); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds %d", d.nS, kMaxNJ * kWave);
        }
        HIP_TRY(hipGetLastError());
    }
    const size_t total = (size_t)d.nF * d.nIC;
    if (total > 0 && (pass & 2)) {
        const unsigned g2 = (unsigned)std::min<size_t>((total * 16 + kBlock - 1) / kBlock, 1u << 20);
        hipLaunchKernelGGL(k_At_fibre<T>, dim3(g2), dim3(kBlock), 0, st, d, p->zbuf, x);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// Voxel-major CSR of one compartment: entries stably sorted by voxel and, with
// by_o, by orientation inside each voxel (the LDS-slab kernels need a voxel's
// entries of one orientation slab to be contiguous; every other kernel is
// indifferent to the order inside a voxel up to rounding).
// col_in null: the column is the entry's input position.
int build_comp(commit_hip_plan *p, const uint32_t *vox, size_t m, uint32_t nVOX, const uint16_t *o,
               const uint32_t *col_in, const float *len, Comp &c, std::vector<uint32_t> *perm_out, bool by_o = false)
{
    std::vector<uint32_t> perm, ptr;
    if (by_o && o) {                                  // LSD: orientation, then voxel (both stable)
        std::vector<uint32_t> key(m), p1, p2, ptr1;
        for (size_t k = 0; k < m; k++) key[k] = o[k];
        counting_sort(key.data(), m, 65536u, p1, ptr1);
        for (size_t k = 0; k < m; k++) key[k] = vox[p1[k]];
        if (!counting_sort(key.data(), m, nVOX, p2, ptr)) return fail(COMMIT_HIP_EINVAL, "voxel index >= nVOX=%u", nVOX);
        perm.resize(m);
        for (size_t k = 0; k < m; k++) perm[k] = p1[p2[k]];
    } else if (!counting_sort(vox, m, nVOX, perm, ptr)) {
        return fail(COMMIT_HIP_EINVAL, "voxel index >= nVOX=%u", nVOX);
    }
    std::vector<uint32_t> col(m);
    for (size_t k = 0; k < m; k++) col[k] = col_in ? col_in[perm[k]] : perm[k];
    int rc;
    c.m = (uint32_t)m;
    if ((rc = upload(p, ptr.data(), ptr.size(), &c.ptr))) return rc;
    if ((rc = upload(p, col.data(), m, &c.col))) return rc;
    c.o = nullptr;
    if (o) {
        std::vector<uint16_t> oo(m);
        for (size_t k = 0; k < m; k++) oo[k] = o[perm[k]];
        if ((rc = upload(p, oo.data(), m, &c.o))) return rc;
    }
    c.len = nullptr;
    if (len) {
        std::vector<float> ll(m);
        for (size_t k = 0; k < m; k++) ll[k] = len[perm[k]];
        if ((rc = upload(p, ll.data(), m, &c.len))) return rc;
    }
    if (perm_out) perm_out->swap(perm);
    return 0;
}

}  // namespace

extern "C" {

const char *commit_hip_last_error(void) { return commit_internal::g_err.c_str(); }

const char *commit_hip_version(void) { return "commit_hip 0.3 (gfx950)"; }

int commit_hip_plan_create(commit_hip_plan **plan, int nS, int ndirs, int nF, int n, int nE, int nV,
                           int nVOX, const uint32_t *ICf, const uint32_t *ICeval, const uint32_t *ICv,
                           const uint16_t *ICo, const float *ICl, const uint32_t *ECv,
                           const uint16_t *ECo, const uint32_t *ISOv, const float *wmrSFP,
                           const float *wmhSFP, const float *isoSFP, uint32_t nIC, uint32_t nEC,
                           uint32_t nISO, int nolut, int precision, int device)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "plan pointer is NULL");
    *plan = nullptr;
    if (precision != COMMIT_HIP_FP64 && precision != COMMIT_HIP_FP32)
        return fail(COMMIT_HIP_EINVAL, "unknown precision %d", precision);
    if (nS < 1 || ndirs < 1 || nF < 0 || n < 0 || nE < 0 || nV < 0 || nVOX < 0)
        return fail(COMMIT_HIP_EINVAL, "negative or zero size (nS=%d ndirs=%d nF=%d n=%d nE=%d nV=%d nVOX=%d)",
                    nS, ndirs, nF, n, nE, nV, nVOX);
    if (nIC > (uint32_t)kMaxRF || nEC > (uint32_t)kMaxRF || nISO > (uint32_t)kMaxRF)
        return fail(COMMIT_HIP_EUNSUPPORTED, "more than %d response functions in a compartment", kMaxRF);
    if (nolut && (nIC != 1 || nEC != 0 || nISO > 1))
        return fail(COMMIT_HIP_EUNSUPPORTED, "LUT-free operator needs nIC=1, nEC=0, nISO<=1");
    if (!nolut && nS > kMaxNJ * kWave)
        return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d exceeds %d", nS, kMaxNJ * kWave);
    if ((n > 0 && nIC == 0) || (n > 0 && (!ICf || !ICv || !ICo || !ICl)) || (nF > 0 && nIC > 0 && !ICeval))
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
    for (size_t q = 0; q < (size_t)nF * nIC; q++)
        if (ICeval[q] > 1u) return fail(COMMIT_HIP_EINVAL, "ICeval[%zu]=%u is not 0 or 1", q, ICeval[q]);

    HIP_TRY(hipSetDevice(device));
    commit_hip_plan *p = new commit_hip_plan();
    p->device = device;
    p->precision = precision;
    p->nolut = nolut;
    p->nS_ref = nS;
    DevDict &d = p->d;
    d.nS = nolut ? 1 : nS;
    d.nSp = ((d.nS + kWave - 1) / kWave) * kWave;
    d.ndirs = nolut ? 1 : ndirs;
    d.nF = nF;
    d.n = n;
    d.nE = (nEC > 0) ? nE : 0;
    d.nI = (nISO > 0) ? nV : 0;
    d.nVOX = nVOX;
    d.nIC = (int)nIC;
    d.nEC = (int)nEC;
    d.nISO = (int)nISO;
    d.xcd_remap = 0;       // hardware order: measured faster on the 1M-fibre head; the autotuner may switch
    int rc = 0;
    {
        std::vector<uint32_t> perm;
        std::vector<uint16_t> zero_o;
        const uint16_t *o = ICo;
        if (nolut) {
            zero_o.assign((size_t)n, 0);
            o = zero_o.data();
        }
        if ((rc = build_comp(p, ICv, (size_t)n, (uint32_t)nVOX, o, ICf, ICl, d.ic, &perm, true))) goto done;
        // fibre-major permutation of the voxel-sorted entries
        std::vector<uint32_t> f(n), fperm, fptr;
        for (int k = 0; k < n; k++) f[k] = ICf[perm[k]];
        counting_sort(f.data(), (size_t)n, (uint32_t)nF, fperm, fptr);
        if ((rc = upload(p, fptr.data(), fptr.size(), &d.fib_ptr))) goto done;
        if ((rc = upload(p, fperm.data(), (size_t)n, &d.fib_perm))) goto done;
        std::vector<uint32_t> zp((size_t)n);
        for (int k = 0; k < n; k++) zp[fperm[k]] = (uint32_t)k;
        if ((rc = upload(p, zp.data(), (size_t)n, &d.zpos))) goto done;
        if ((rc = upload(p, ICeval, (size_t)nF * nIC, &d.eval))) goto done;
        for (size_t q = 0; q < (size_t)nF * nIC; q++)
            if (ICeval[q] != 1u) { p->eval_ones = false; break; }
    }
    if (nEC > 0 && (rc = build_comp(p, ECv, (size_t)d.nE, (uint32_t)nVOX, ECo, nullptr, nullptr, d.ec, nullptr)))
        goto done;
    if (nISO > 0 && (rc = build_comp(p, ISOv, (size_t)d.nI, (uint32_t)nVOX, nullptr, nullptr, nullptr, d.iso, nullptr)))
        goto done;
    if (nolut) {
        std::vector<float> row((size_t)d.nSp, 0.0f);
        row[0] = 1.0f;
        if ((rc = upload(p, row.data(), row.size(), &d.wmr))) goto done;
        if (nISO && (rc = upload(p, row.data(), row.size(), &d.isol))) goto done;
    } else {
        if (nIC) {
            // three zero rows after the last RF: the dense-tile kernel reads
            // orientations in steps of 4 past ndirs (against zero weights)
            std::vector<float> t = pad_rows(wmrSFP, (size_t)nIC * ndirs, nS, d.nSp);
            t.resize(t.size() + 3 * (size_t)d.nSp, 0.0f);
            if ((rc = upload(p, t.data(), t.size(), &d.wmr))) goto done;
        }
        if (nEC) {
            const std::vector<float> t = pad_rows(wmhSFP, (size_t)nEC * ndirs, nS, d.nSp);
            if ((rc = upload(p, t.data(), t.size(), &d.wmh))) goto done;
        }
        if (nISO) {
            const std::vector<float> t = pad_rows(isoSFP, (size_t)nISO, nS, d.nSp);
            if ((rc = upload(p, t.data(), t.size(), &d.isol))) goto done;
        }
    }
    if (n > 0 && nIC > 0) {
        void *z = nullptr;
        if ((rc = dev_alloc(p, (size_t)n * nIC * sizeof(double), &z))) goto done;
        p->zbuf = static_cast<double *>(z);
    }
    {
        void *rp = nullptr;
        if ((rc = dev_alloc(p, ((size_t)nVOX + kSumBlocks + kCombBlocks) * sizeof(double), &rp))) goto done;
        p->rpart = static_cast<double *>(rp);
    }
    if (nF > 0 && nIC > 0) {
        void *xi = nullptr;
        if ((rc = dev_alloc(p, (size_t)nF * nIC * sizeof(double), &xi))) goto done;
        p->xic = static_cast<double *>(xi);
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
    for (size_t q = 0; q < cnt; q++)
        if (ICeval[q] > 1u) return fail(COMMIT_HIP_EINVAL, "ICeval[%zu]=%u is not 0 or 1", q, ICeval[q]);
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipMemcpy(const_cast<uint32_t *>(plan->d.eval), ICeval, cnt * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    plan->eval_ones = true;
    for (size_t q = 0; q < cnt; q++)
        if (ICeval[q] != 1u) { plan->eval_ones = false; break; }
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

int commit_hip_plan_set_precision(commit_hip_plan *plan, int precision)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (precision != COMMIT_HIP_FP64 && precision != COMMIT_HIP_FP32)
        return fail(COMMIT_HIP_EINVAL, "unknown precision %d", precision);
    plan->precision = precision;
    return 0;
}

int commit_hip_plan_precision(const commit_hip_plan *plan) { return plan ? plan->precision : -1; }

int commit_hip_plan_set_option(commit_hip_plan *plan, int option, int value)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    switch (option) {
        case COMMIT_HIP_OPT_XCD_REMAP:
            if (value != 0 && value != 1) return fail(COMMIT_HIP_EINVAL, "XCD remap is 0 or 1");
            plan->d.xcd_remap = value;
            return 0;
        case COMMIT_HIP_OPT_MULTI_A_WAVES:
            if (value != 1 && value != 2) return fail(COMMIT_HIP_EINVAL, "wavefronts per voxel are 1 or 2");
            plan->multi_a_waves = value;
            return 0;
        case COMMIT_HIP_OPT_A_DENSE:
            if (value != 0 && value != 1) return fail(COMMIT_HIP_EINVAL, "dense A is 0 or 1");
            plan->a_dense = value;
            return 0;
        case COMMIT_HIP_OPT_SLAB:
            if (value < 0 || value > 2) return fail(COMMIT_HIP_EINVAL, "slab mode is 0, 1 or 2");
            if (value == plan->slab) return 0;
            HIP_TRY(hipSetDevice(plan->device));
            return commit_internal::slab_configure(plan, value);
        default: return fail(COMMIT_HIP_EINVAL, "unknown option %d", option);
    }
}

int commit_hip_plan_get_option(const commit_hip_plan *plan, int option)
{
    if (!plan) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    switch (option) {
        case COMMIT_HIP_OPT_XCD_REMAP: return plan->d.xcd_remap;
        case COMMIT_HIP_OPT_MULTI_A_WAVES: return plan->multi_a_waves;
        case COMMIT_HIP_OPT_A_DENSE: return plan->a_dense;
        case COMMIT_HIP_OPT_SLAB: return plan->slab;
        default: return fail(COMMIT_HIP_EINVAL, "unknown option %d", option);
    }
}

}  // extern "C"

namespace {

template <typename T>
int do_A(commit_hip_plan *plan, const void *x, void *y, hipStream_t st)
{
    T *yd = static_cast<T *>(y);
    int rc = launch_A<T>(plan, static_cast<const T *>(x), AOut<T>{yd, nullptr, nullptr, nullptr}, st);
    if (rc) return rc;
    if (plan->nolut && plan->nS_ref > 1) {
        // LUT-free rows are indexed by voxel only (setup_operator.py:726-729);
        // the remaining rows of the reference-sized output stay zero.
        const size_t tail = (size_t)plan->d.nVOX * (plan->nS_ref - 1);
        hipLaunchKernelGGL(k_zero<T>, dim3(1024), dim3(256), 0, st, yd + plan->d.nVOX, tail);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

template <typename T>
int do_At(commit_hip_plan *plan, const void *y, const void *sqrtw, void *x, hipStream_t st, int pass)
{
    return launch_At<T>(plan, static_cast<const T *>(y), static_cast<const T *>(sqrtw), static_cast<T *>(x), st,
                        pass);
}

template <typename T>
int do_A_residual(commit_hip_plan *plan, const void *x, const void *yobs, const void *sqrtw, void *res,
                  double *rsq, hipStream_t st)
{
    const size_t nv = (size_t)plan->d.nVOX;
    AOut<T> out{static_cast<T *>(res), static_cast<const T *>(yobs), static_cast<const T *>(sqrtw), plan->rpart};
    int rc = launch_A<T>(plan, static_cast<const T *>(x), out, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_stage1, dim3(kSumBlocks), dim3(kBlock), 0, st, plan->rpart, nv, plan->rpart + nv);
    hipLaunchKernelGGL(k_sum_stage2, dim3(1), dim3(kBlock), 0, st, plan->rpart + nv, kSumBlocks, rsq);
    HIP_TRY(hipGetLastError());
    return 0;
}

// r = a r1 + b r2 over n rows, with per-block sums of r^2 (FP64, from the
// unrounded values): block k owns a fixed contiguous range (a multiple of 16
// bytes long), read with 16-byte loads and reduced in a fixed order, so the
// norm is repeatable.
template <typename T>
__device__ __forceinline__ void combine_vec(const T (&u)[16 / sizeof(T)], const T (&w)[16 / sizeof(T)], double a,
                                            double b, T (&o)[16 / sizeof(T)], double &ss)
{
#pragma unroll
    for (int k = 0; k < (int)(16 / sizeof(T)); k++) {
        const double v = a * (double)u[k] + b * (double)w[k];
        o[k] = (T)v;
        ss += v * v;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(size_t n, double a, const T *__restrict__ r1, double b,
                                                    const T *__restrict__ r2, T *__restrict__ r,
                                                    double *__restrict__ part)
{
    constexpr int VEC = 16 / sizeof(T);
    typedef T Vec[VEC];
    __shared__ double sh[kBlock / kWave];
    const size_t chunk = ((n + kCombBlocks - 1) / kCombBlocks + VEC - 1) / VEC * VEC;
    const size_t lo = (size_t)blockIdx.x * chunk;
    const size_t hi = lo + chunk < n ? lo + chunk : n;
    double ss = 0.0;
    if (lo < hi) {
        const size_t nv = (hi - lo) / VEC;                     // whole 16-byte vectors of the range
        const Vec *u = reinterpret_cast<const Vec *>(r1 + lo);
        const Vec *w = reinterpret_cast<const Vec *>(r2 + lo);
        Vec *o = reinterpret_cast<Vec *>(r + lo);
        size_t i = threadIdx.x;
        for (; i + kBlock < nv; i += 2 * kBlock) {             // 4 16-byte loads in flight per thread
            Vec u0, w0, u1, w1, o0, o1;
#pragma unroll
            for (int k = 0; k < VEC; k++) {
                u0[k] = u[i][k];
                w0[k] = w[i][k];
                u1[k] = u[i + kBlock][k];
                w1[k] = w[i + kBlock][k];
            }
            combine_vec<T>(u0, w0, a, b, o0, ss);
            combine_vec<T>(u1, w1, a, b, o1, ss);
#pragma unroll
            for (int k = 0; k < VEC; k++) {
                o[i][k] = o0[k];
                o[i + kBlock][k] = o1[k];
            }
        }
        for (; i < nv; i += kBlock) {
            Vec u0, w0, o0;
#pragma unroll
            for (int k = 0; k < VEC; k++) {
                u0[k] = u[i][k];
                w0[k] = w[i][k];
            }
            combine_vec<T>(u0, w0, a, b, o0, ss);
#pragma unroll
            for (int k = 0; k < VEC; k++) o[i][k] = o0[k];
        }
        const size_t t = lo + nv * VEC + threadIdx.x;          // the range's last elements (the end of r)
        if (t < hi) {
            const double v = a * (double)r1[t] + b * (double)r2[t];
            r[t] = (T)v;
            ss += v * v;
        }
    }
    ss = block_sum_256(ss, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = ss;
}

template <typename T>
int do_combine(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2, void *r, double *rsq,
               hipStream_t st)
{
    const size_t n = (size_t)plan->d.nVOX * plan->nS_ref;
    for (const void *q : {r1, r2, static_cast<const void *>(r)})
        if (reinterpret_cast<uintptr_t>(q) % 16 != 0) return fail(COMMIT_HIP_EINVAL, "residual vectors must be 16-byte aligned");
    double *part = plan->rpart + plan->d.nVOX + kSumBlocks;
    hipLaunchKernelGGL(k_combine<T>, dim3(kCombBlocks), dim3(kBlock), 0, st, n, a, static_cast<const T *>(r1), b,
                       static_cast<const T *>(r2), static_cast<T *>(r), part);
    hipLaunchKernelGGL(k_sum_stage2, dim3(1), dim3(kBlock), 0, st, part, kCombBlocks, rsq);
    HIP_TRY(hipGetLastError());
    return 0;
}

// A^T (sqrtw .* (a r1 + b r2)) and ||a r1 + b r2||^2: fused on the slab path,
// else the combination into r and the plain A^T.
template <typename T>
int do_At_combined(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2, const void *sqrtw,
                   void *r, void *x, double *rsq, hipStream_t st)
{
    const DevDict &d = plan->d;
    const T *p1 = static_cast<const T *>(r1), *p2 = static_cast<const T *>(r2), *w = static_cast<const T *>(sqrtw);
    T *xo = static_cast<T *>(x);
    if (vox_blocks(d) > 0 && plan->slab) {
        int rc = commit_internal::launch_At_slab_comb<T>(plan, p1, p2, a, b, w, xo, st);
        if (rc == 0) {
            if ((rc = launch_At<T>(plan, p1, w, xo, st, 2))) return rc;          // pass 2 only
            const size_t nv = (size_t)d.nVOX;
            hipLaunchKernelGGL(k_sum_stage1, dim3(kSumBlocks), dim3(kBlock), 0, st, plan->rpart, nv, plan->rpart + nv);
            hipLaunchKernelGGL(k_sum_stage2, dim3(1), dim3(kBlock), 0, st, plan->rpart + nv, kSumBlocks, rsq);
            HIP_TRY(hipGetLastError());
            return 0;
        }
        if (rc != COMMIT_HIP_EUNSUPPORTED) return rc;
    }
    int rc = do_combine<T>(plan, a, r1, b, r2, r, rsq, st);
    if (rc) return rc;
    return do_At<T>(plan, r, sqrtw, x, st, 3);
}

}  // namespace

extern "C" {

int commit_hip_At_combined(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2,
                           const void *sqrtw, void *r, void *x, double *rsq, void *stream)
{
    if (!plan || !r1 || !r2 || !r || !x || !rsq) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (plan->nolut && plan->nS_ref > 1 && sqrtw)
        return fail(COMMIT_HIP_EUNSUPPORTED, "row weights on a LUT-free operator with nS > 1");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    return plan->precision == COMMIT_HIP_FP32 ? do_At_combined<float>(plan, a, r1, b, r2, sqrtw, r, x, rsq, st)
                                              : do_At_combined<double>(plan, a, r1, b, r2, sqrtw, r, x, rsq, st);
}

int commit_hip_residual_combine(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2, void *r,
                                double *rsq, void *stream)
{
    if (!plan || !r1 || !r2 || !r || !rsq) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    return plan->precision == COMMIT_HIP_FP32 ? do_combine<float>(plan, a, r1, b, r2, r, rsq, st)
                                              : do_combine<double>(plan, a, r1, b, r2, r, rsq, st);
}

int commit_hip_A(commit_hip_plan *plan, const void *x, void *y, void *stream)
{
    if (!plan || !x || !y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    return plan->precision == COMMIT_HIP_FP32 ? do_A<float>(plan, x, y, st) : do_A<double>(plan, x, y, st);
}

int commit_hip_At(commit_hip_plan *plan, const void *y, void *x, void *stream)
{
    return commit_hip_At_weighted(plan, y, nullptr, x, stream);
}

int commit_hip_At_weighted(commit_hip_plan *plan, const void *y, const void *sqrtw, void *x, void *stream)
{
    return commit_hip_At_pass(plan, y, sqrtw, x, 3, stream);
}

int commit_hip_At_pass(commit_hip_plan *plan, const void *y, const void *sqrtw, void *x, int pass, void *stream)
{
    if (!plan || !x || !y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (pass < 1 || pass > 3) return fail(COMMIT_HIP_EINVAL, "pass must be 1, 2 or 3");
    if (plan->nolut && plan->nS_ref > 1 && sqrtw)
        return fail(COMMIT_HIP_EUNSUPPORTED, "row weights on a LUT-free operator with nS > 1");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    return plan->precision == COMMIT_HIP_FP32 ? do_At<float>(plan, y, sqrtw, x, st, pass)
                                              : do_At<double>(plan, y, sqrtw, x, st, pass);
}

int commit_hip_A_residual(commit_hip_plan *plan, const void *x, const void *yobs, const void *sqrtw,
                          void *res, double *rsq, void *stream)
{
    if (!plan || !x || !rsq) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (plan->nolut && plan->nS_ref > 1)
        return fail(COMMIT_HIP_EUNSUPPORTED, "residual of a LUT-free operator with nS > 1");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    return plan->precision == COMMIT_HIP_FP32 ? do_A_residual<float>(plan, x, yobs, sqrtw, res, rsq, st)
                                              : do_A_residual<double>(plan, x, yobs, sqrtw, res, rsq, st);
}

}  // extern "C"
