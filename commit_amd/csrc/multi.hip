// Multi-right-hand-side products on the matrix cores: Y = A X and X = A^T Y
// for R columns at once (lambda sweeps, bootstrap signals), FP32 storage.
//
// Reference: the single-vector products COMMIT_A / COMMIT_At
// (setup_operator.py:48-365, 368-710); the reference has no batched form and
// runs one A.dot per right-hand side (solvers.py:266-270).
//
// Per voxel v with entries e (orientation o_e, coefficient c_e[r]) the
// products are two small dense GEMMs, mapped onto v_mfma_f32_16x16x4f32
// (D[16x16] += A[16x4] * B[4x16]; lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives D[4(l/16)+k][l%16], k = 0..3):
//   A X   : Y_v[s][r]  = sum_e LUT[o_e][s] * c_e[r]
//           M = samples (NT tiles of 16), N = 16 RHS, K = entries (4 per step).
//           Tile m, row i holds sample i*NT + m, so a lane reads its NT
//           samples of a LUT row as one contiguous vector.
//   A^T Y : Z[e][r]    = sum_s LUT[o_e][s] * Y_v[s][r]
//           M = entries (16 per tile), N = 16 RHS, K = samples; lane group g
//           takes samples [g*4NT, g*4NT+4NT), again contiguous in the row.
//           IC products go to a per-entry buffer that a second pass sums per
//           fibre in ascending voxel order (deterministic, no atomics); EC and
//           ISO columns belong to one voxel and are written directly.
// One wavefront per (voxel, group of 16 RHS); the LUT row of an entry is read
// once for 16 RHS instead of once per RHS.  Accumulation is FP32 inside the
// MFMA over one voxel (A) or one LUT row (A^T); the per-fibre sums of A^T are
// FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/commit_hip.h"
#include "common.h"
#include "plan.h"
#include "rows.h"

using commit_internal::fail;

namespace {

using commit_internal::Comp;
using commit_internal::DevDict;
using commit_internal::wave_index;

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kMaxNT = 20;          // nS <= 320 for the multi-RHS kernels
constexpr int kTileR = 16;          // RHS per wavefront (MFMA N)

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// NT consecutive floats at p (p is 4*NT-byte aligned when NT is even)
template <int NT>
__device__ __forceinline__ void load_run(float (&a)[NT], const float *__restrict__ p)
{
    if constexpr (NT % 4 == 0) {
#pragma unroll
        for (int m = 0; m < NT / 4; m++) {
            const float4 v = reinterpret_cast<const float4 *>(p)[m];
            a[4 * m] = v.x, a[4 * m + 1] = v.y, a[4 * m + 2] = v.z, a[4 * m + 3] = v.w;
        }
    } else if constexpr (NT % 2 == 0) {
#pragma unroll
        for (int m = 0; m < NT / 2; m++) {
            const float2 v = reinterpret_cast<const float2 *>(p)[m];
            a[2 * m] = v.x, a[2 * m + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int m = 0; m < NT; m++) a[m] = p[m];
    }
}

// The rows of voxel v's product in one list: IC entries for each IC RF, then
// EC peaks for each EC RF, then ISO entries for each ISO RF.
struct VRows {
    uint32_t icb, nic, ecb, nec, isob, niso;   // the voxel's entry ranges
    uint32_t n1, n2, total;                    // row counts: IC, IC+EC, all
    const uint32_t *uhead;                     // A^T: IC rows are heads uhead[icb..icb+nic)
};

__device__ __forceinline__ VRows vrows(const DevDict &d, int v)
{
    VRows vr;
    vr.icb = d.nIC > 0 ? d.ic.ptr[v] : 0u;
    vr.nic = d.nIC > 0 ? d.ic.ptr[v + 1] - vr.icb : 0u;
    vr.ecb = d.nEC > 0 ? d.ec.ptr[v] : 0u;
    vr.nec = d.nEC > 0 ? d.ec.ptr[v + 1] - vr.ecb : 0u;
    vr.isob = d.nISO > 0 ? d.iso.ptr[v] : 0u;
    vr.niso = d.nISO > 0 ? d.iso.ptr[v + 1] - vr.isob : 0u;
    vr.n1 = (uint32_t)d.nIC * vr.nic;
    vr.n2 = vr.n1 + (uint32_t)d.nEC * vr.nec;
    vr.total = vr.n2 + (uint32_t)d.nISO * vr.niso;
    vr.uhead = nullptr;
    return vr;
}

// The A^T list of voxel v: its IC rows are the heads of its runs of equal
// orientation (mm_uhead[uptr[v]..uptr[v+1])), then EC and ISO rows as in A X.
__device__ __forceinline__ VRows vrows_at(const DevDict &d, int v, const uint32_t *__restrict__ uptr,
                                          const uint32_t *__restrict__ uhead)
{
    VRows vr = vrows(d, v);
    if (d.nIC > 0) {
        vr.icb = uptr[v];
        vr.nic = uptr[v + 1] - vr.icb;
        vr.n1 = (uint32_t)d.nIC * vr.nic;
        vr.n2 = vr.n1 + (uint32_t)d.nEC * vr.nec;
        vr.total = vr.n2 + (uint32_t)d.nISO * vr.niso;
    }
    vr.uhead = uhead;
    return vr;
}

// ------------------------------------------------------------- Y = A X ------
template <int NT>
__device__ __forceinline__ void a_mma(f4 (&acc)[NT], const float (&a)[NT], float b)
{
#pragma unroll
    for (int m = 0; m < NT; m++) acc[m] = mfma4(a[m], b, acc[m]);
}

// Start of this lane's run in the LUT row coded as in entry_b.
__device__ __forceinline__ const float *row_code(const DevDict &d, uint32_t c, size_t lane_off)
{
    const uint32_t tb = c >> 30, row = c & 0x3fffffffu;
    const float *w0 = d.wmr, *w1 = d.wmh, *w2 = d.isol;
    const float *L = tb == 0u ? w0 : tb == 1u ? w1 : w2;
    return L + (size_t)row * d.nSp + lane_off;
}

// Entry r (< vr.total) of the voxel's list for A X: its LUT row as a code
// (table in the top 2 bits: 0 IC, 1 EC, 2 ISO; row index below), its x
// column and its scale (IC: length x eval; EC, ISO: 1).  entry_a loads the
// entry's fields without per-lane branches (every lane loads from each
// compartment present, index 0 for lanes of another kind), so the loads go
// out together and need no wait until entry_b combines them; entry_b also
// reads eval (EV), which needs the column.
struct EntA {
    uint32_t o, col, eo, ecol, icol, i;
    float len;
    int kind;                                      // 0 IC, 1 EC, 2 ISO
};

__device__ __forceinline__ EntA entry_a(const DevDict &d, const VRows &vr, uint32_t r)
{
    // (members copied to scalars before the selects, so the selects stay on values)
    const uint32_t n1 = vr.n1, n2 = vr.n2, nic = vr.nic, nec = vr.nec, niso = vr.niso;
    const uint32_t icb = vr.icb, ecb = vr.ecb, isob = vr.isob;
    const int rfic = d.nIC, rfec = d.nEC, rfiso = d.nISO;
    r = min(r, vr.total - 1);
    const bool ic = r < n1, ec = !ic && r < n2;
    const uint32_t t = ic ? r : (ec ? r - n1 : r - n2);
    const uint32_t m = ic ? nic : (ec ? nec : niso);
    const int nrf = ic ? rfic : (ec ? rfec : rfiso);
    EntA e;
    e.i = nrf == 1 ? 0u : t / m;
    const uint32_t k = (ic ? icb : (ec ? ecb : isob)) + (t - e.i * m);
    e.kind = ic ? 0 : ec ? 1 : 2;
    e.o = e.col = e.eo = e.ecol = e.icol = 0u;
    e.len = 0.0f;
    if (d.nIC > 0 && d.n > 0) {
        const uint32_t kk = ic ? k : 0u;
        e.o = d.ic.o[kk];
        e.col = d.ic.col[kk];
        e.len = d.ic.len[kk];
    }
    if (d.nEC > 0 && d.nE > 0) {
        const uint32_t kk = ec ? k : 0u;
        e.eo = d.ec.o[kk];
        e.ecol = d.ec.col[kk];
    }
    if (d.nISO > 0 && d.nI > 0) e.icol = d.iso.col[!ic && !ec ? k : 0u];
    return e;
}

template <bool EV>
__device__ __forceinline__ void entry_b(const DevDict &d, const EntA &e, const uint32_t *eval, uint32_t &code,
                                        uint32_t &xcol, float &scale)
{
    const uint32_t base_ec = (uint32_t)d.nIC * (uint32_t)d.nF;
    const uint32_t base_iso = base_ec + (uint32_t)d.nEC * (uint32_t)d.nE;
    code = e.kind == 0 ? e.i * (uint32_t)d.ndirs + e.o
         : e.kind == 1 ? (1u << 30) | (e.i * (uint32_t)d.ndirs + e.eo)
                       : (2u << 30) | e.i;
    xcol = e.kind == 0 ? e.i * (uint32_t)d.nF + e.col
         : e.kind == 1 ? base_ec + e.i * (uint32_t)d.nE + e.ecol
                       : base_iso + e.i * (uint32_t)d.nI + e.icol;
    scale = e.kind == 0 ? e.len : 1.0f;
    if (EV && e.kind == 0 && eval[xcol] == 0u) scale = 0.0f;
}

// A X for one voxel (setup_operator.py:48-242): all its entries (IC for each
// RF, EC, ISO) in chunks of 64, so a voxel pays one start-up (records,
// coefficient gathers, first rows) per chunk instead of one per compartment.
// K-step t of a chunk covers entries 4t..4t+3; lane group g owns entry 4t+g.
// At a chunk start every lane gathers the coefficients of its 16 K-steps and
// the next chunk's entry records are requested; the LUT rows stream two steps
// ahead of the matrix-core work.  NH: the sample tiles this wavefront owns
// (of NT = NH * MS); soff: offset of its first sample in each lane's run.
template <int NH, int NT, bool EV>
__device__ __forceinline__ void mm_vox_A(f4 (&acc)[NH], const DevDict &d, const VRows &vr,
                                         const float *__restrict__ X, int R, int rc, bool rhs_ok,
                                         const uint32_t *__restrict__ eval, int lane, int soff)
{
    if (vr.total == 0) return;
    const int g = lane >> 4, q = lane & 15;
    const size_t lane_off = (size_t)q * NT + soff;
    uint32_t code, xcol;
    float scale;
    entry_b<EV>(d, entry_a(d, vr, (uint32_t)lane), eval, code, xcol, scale);
    for (uint32_t base = 0; base < vr.total; base += kWave) {
        const int cnt = (int)min((uint32_t)kWave, vr.total - base);
        const int nsteps = (cnt + 3) >> 2;
        const uint32_t cc = code;
        // a ring of four LUT row buffers: step t's rows are requested three
        // steps before its MFMAs; requests past the chunk's last step repeat
        // its rows (cached, never used), so every load is unconditional and
        // each wait covers one step's rows only
        float b0[NH], b1[NH], b2[NH], b3[NH];
#define ROW(t) row_code(d, (uint32_t)__shfl((int)cc, 4 * min((t), nsteps - 1) + g), lane_off)
        // the first four steps' rows are requested before the coefficient
        // gathers, so that all are in flight together
        load_run<NH>(b0, ROW(0));
        load_run<NH>(b1, ROW(1));
        load_run<NH>(b2, ROW(2));
        load_run<NH>(b3, ROW(3));
        // the scheduler would otherwise sink each row load to its first use
        __builtin_amdgcn_sched_barrier(0);
        // coefficients scale * x of this lane's entries, all 16 gathers in flight
        float xv[16];
        {
            float sv[16];
            bool okv[16];
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const int ent = 4 * t + g;
                const uint32_t ce = (uint32_t)__shfl((int)xcol, ent);
                sv[t] = __shfl(scale, ent);
                okv[t] = ent < cnt && rhs_ok;
                xv[t] = X[(size_t)(okv[t] ? ce : 0u) * (size_t)R + rc];
            }
#pragma unroll
            for (int t = 0; t < 16; t++) xv[t] = okv[t] ? sv[t] * xv[t] : 0.0f;
        }
        const bool more = base + kWave < vr.total;
        EntA en;
        if (more) en = entry_a(d, vr, base + kWave + lane);     // combined after this chunk's products
        // the 16 steps written out (buffer t % 4, static indices throughout)
#define STEP(B, T)                                                                                                 \
    if ((T) >= nsteps) goto chunk_done;                                                                            \
    a_mma<NH>(acc, B, xv[T]);                                                                                      \
    load_run<NH>(B, ROW((T) + 4));                                                                                 \
    __builtin_amdgcn_sched_barrier(0);
        STEP(b0, 0) STEP(b1, 1) STEP(b2, 2) STEP(b3, 3) STEP(b0, 4) STEP(b1, 5) STEP(b2, 6) STEP(b3, 7)
        STEP(b0, 8) STEP(b1, 9) STEP(b2, 10) STEP(b3, 11) STEP(b0, 12) STEP(b1, 13) STEP(b2, 14) STEP(b3, 15)
    chunk_done:
#undef STEP
#undef ROW
        if (more) entry_b<EV>(d, en, eval, code, xcol, scale);
    }
}

// The fused residual of the batched fits (RES): instead of A X the output
// stage writes res = sqrtw .* (A X - y) (when res is set) and the wavefront's
// 16 column sums of res^2, in FP64, at part[((v * MS + h) * nrg + rg) * 16 + q].
struct ResArgs {
    const float *y;       // [rows] or [rows][R] (ystride 1); NULL: zero
    int ystride;
    const float *w;       // sqrtw [rows] or NULL
    float *res;           // [rows][R] or NULL
    double *part;
    float *res1;          // with sel: the output is (*sel ? res : res1), the buffer not holding r(prev_x)
    const int32_t *sel;
};

// Output stage of the fused residual for one lane (column rhs of voxel v,
// rows rb + s): res = sqrtw (A X - y) stored when ra.res is set; returns the
// lane's sum of res^2.  YM 0: y from the staged LDS row; 1: y per column from
// memory; 2: y = 0.  The loop issues no memory load for YM 0 and 2, so the
// stores never wait for one another.
template <int NH, int NT, int YM>
__device__ __forceinline__ double res_out(const f4 (&acc)[NH], const ResArgs &ra, float *res, const float *sy,
                                          const float *sw, size_t rb, int nS, int g, int soff, int R, int rhs)
{
    double ss = 0.0;
    float *rp = res ? res + rb * R + rhs : nullptr;
#pragma unroll
    for (int m = 0; m < NH; m++) {
        float yv[4], wv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int s = min((4 * g + k) * NT + soff + m, nS - 1);
            yv[k] = YM == 0 ? sy[s] : YM == 1 ? ra.y[(rb + s) * R + rhs] : 0.0f;
            wv[k] = sw[s];                      // (not staged, and unused, without sqrtw)
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int s = (4 * g + k) * NT + soff + m;
            if (s < nS) {
                double t = (double)acc[m][k] - (double)yv[k];
                if (ra.w) t *= (double)wv[k];
                if (rp) __builtin_nontemporal_store((float)t, rp + s * R);
                ss = fma(t, t, ss);
            }
        }
    }
    return ss;
}

// One wavefront per (voxel, RHS group, sample half h < MS): tiles
// [h*NH, h*NH + NH) of the NT = NH * MS, i.e. samples i*NT + h*NH + m.
template <int NH, int MS, bool RES>
__global__ __launch_bounds__(kBlock) void k_A_mm(DevDict d, const float *__restrict__ X, float *__restrict__ Y,
                                                 int R, int nrg, const uint32_t *__restrict__ eval, ResArgs ra)
{
    constexpr int NT = NH * MS;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_index(d.xcd_remap);
    const int v = w / (nrg * MS), rem = w - v * (nrg * MS);
    const int rg = rem / MS, h = rem - rg * MS;
    if (v >= d.nVOX) return;
    const int g = lane >> 4, q = lane & 15;
    const int rhs = rg * kTileR + q;
    const int soff = h * NH;
    // RES: the voxel's rows of a shared y and of sqrtw go from memory into
    // this wavefront's LDS now and are read by the output stage, so that
    // stage waits on no memory.
    __shared__ float stage[RES ? kWavesPerBlock : 1][RES ? 2 * kMaxNT * 16 : 1];
    float *sy = nullptr, *sw = nullptr;
    if constexpr (RES) {
        sy = stage[threadIdx.x >> 6];
        sw = sy + kMaxNT * 16;
        const size_t rb = (size_t)v * d.nS;
        for (int j = 0; j * kWave < d.nS; j++) {
            const size_t i = rb + (size_t)min(j * kWave + lane, d.nS - 1);
            if (ra.y && !ra.ystride) __builtin_amdgcn_global_load_lds(ra.y + i, (lds_void *)(sy + kWave * j), 4, 0, 0);
            if (ra.w) __builtin_amdgcn_global_load_lds(ra.w + i, (lds_void *)(sw + kWave * j), 4, 0, 0);
        }
    }
    f4 acc[NH];
#pragma unroll
    for (int m = 0; m < NH; m++) acc[m] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int rc = rhs < R ? rhs : 0;                      // columns past R read column 0 (zeroed)
    const VRows vr = vrows(d, v);
    if (eval)
        mm_vox_A<NH, NT, true>(acc, d, vr, X, R, rc, rhs < R, eval, lane, soff);
    else
        mm_vox_A<NH, NT, false>(acc, d, vr, X, R, rc, rhs < R, eval, lane, soff);
    if constexpr (!RES) {
        if (rhs >= R) return;
        float *yv = Y + (size_t)v * d.nS * R + rhs;
#pragma unroll
        for (int m = 0; m < NH; m++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int s = (4 * g + k) * NT + soff + m;
                if (s < d.nS) __builtin_nontemporal_store(acc[m][k], yv + (size_t)s * R);
            }
        }
    } else {
        // (solvers.py:306-308) per element as commit_hip_multi_residual: FP64
        // difference of the FP32 product and y, times sqrtw, squared and summed
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");      // the staged rows are in LDS
        double ss = 0.0;
        if (rhs < R) {
            const size_t rb = (size_t)v * d.nS;
            float *res = ra.sel ? (*ra.sel ? ra.res : ra.res1) : ra.res;
            if (!ra.y)
                ss = res_out<NH, NT, 2>(acc, ra, res, sy, sw, rb, d.nS, g, soff, R, rhs);
            else if (ra.ystride)
                ss = res_out<NH, NT, 1>(acc, ra, res, sy, sw, rb, d.nS, g, soff, R, rhs);
            else
                ss = res_out<NH, NT, 0>(acc, ra, res, sy, sw, rb, d.nS, g, soff, R, rhs);
        }
        ss = swap32_add(ss, ss);           // lanes q, q+16, q+32, q+48 hold column q
        ss = swap16_add(ss, ss);
        if (g == 0) ra.part[(((size_t)v * MS + h) * nrg + rg) * kTileR + q] = ss;
    }
}

// Sums of the wavefront partials per row piece: piece p (rows [b, e) of one
// problem) owns the voxels whose first row lies in it, [ceil(b/nS), ceil(e/nS));
// out[p * R + r] = sum over those voxels and their sample halves, in order.
__global__ __launch_bounds__(kBlock) void k_res_pieces(const double *__restrict__ part, const int64_t *__restrict__ pieces,
                                                       int nS, int nVOX, int MS, int nrg, int R,
                                                       double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    const int64_t b = pieces[2 * blockIdx.x], e = pieces[2 * blockIdx.x + 1];
    const int64_t nv = nVOX, vb0 = (b + nS - 1) / nS, ve0 = (e + nS - 1) / nS;
    const int64_t vb = vb0 < nv ? vb0 : nv, ve = ve0 < nv ? ve0 : nv;
    const int q = threadIdx.x & 15, j = threadIdx.x >> 4;
    for (int rg = 0; rg < nrg; rg++) {
        double s = 0.0;
        for (int64_t i = vb * MS + j; i < ve * MS; i += kBlock / 16) s += part[((size_t)i * nrg + rg) * kTileR + q];
        sh[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x < 16 && rg * kTileR + q < R) {
            double t = 0.0;
            for (int k = 0; k < kBlock / 16; k++) t += sh[k * 16 + q];
            out[(size_t)blockIdx.x * R + rg * kTileR + q] = t;
        }
        __syncthreads();
    }
}

// rsq[s * R + r] = sum of the piece sums of problem s: one wavefront per
// (problem, column), lane l adding pieces l, l + 64, ... in order, then a
// fixed butterfly over the 64 lanes
__global__ __launch_bounds__(kBlock) void k_res_problems(const double *__restrict__ pp, const int32_t *__restrict__ pseg,
                                                         int npieces, int S, int R, double *__restrict__ rsq)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int t = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (t >= S * R) return;
    const int r = t % R, s = t / R;
    double v = 0.0;
    for (int p = lane; p < npieces; p += kWave)
        if (pseg[p] == s) v += pp[(size_t)p * R + r];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) rsq[t] = v;
}

// ------------------------------------------------------------ X = A^T Y -----
// Every row of the voxel's A^T product is one LUT row dotted with Y_v: its IC
// entries for each IC RF, then its EC peaks for each EC RF, then its ISO
// entries for each ISO RF.  They are taken as one list, 16 per M-tile, so the
// few EC and ISO rows fill the free rows of the voxel's last IC tile instead
// of a tile each.  The K order interleaves the lane groups: MFMA step 4u + j
// takes sample 16u + 4g + j from lane group g, so a lane's A operand for
// steps 4u..4u+3 is chunk 4u + g (4 floats) of its tile row q, and one
// 16-byte load per u, issued by all 64 lanes, covers 64 contiguous bytes of
// each of the 16 rows.  The A operand goes straight from memory into
// registers: as each chunk's 4 MFMAs are issued, the same registers are
// reloaded with the next tile's chunk, so every load has a whole tile of
// matrix-core work to arrive in.  (Before, tiles were staged through a
// wavefront-private LDS tile that could hold one tile only: a tile waited for
// its rows, requested when the previous tile's MFMAs began.)  The row records
// (LUT row, output slot, scale) come from registers: each lane holds the
// records of one row of a 64-row window, the next window loads four tiles
// ahead, and a tile's row pointers and slots are shuffled out of the window.
// yb: this lane's B operand, Y_v[16u + 4g + j][rhs] at yb[4u + j].
//   IC row (RF i, entry k) : z[(i*n + k)*16 + q]                        = len * dot
//   EC row (RF i, peak k)  : X[(nIC*nF + i*nE + col)*R + rhs]            = dot
//   ISO row (RF i, entry k): X[(nIC*nF + nEC*nE + i*nI + col)*R + rhs]   = dot


// Sample of K step t (< 4NT) of lane group g in A^T.
__device__ __forceinline__ int at_sample(int t, int g) { return 16 * (t >> 2) + 4 * g + (t & 3); }

// Record of row r (< vr.total) of the voxel's A^T list: LUT row code (table
// in the top 2 bits: 0 IC, 1 EC, 2 ISO), output slot and a sign (IC: slot
// i*n + h of z for head entry h, 1; EC/ISO: x column, -1).  It is read in two
// phases: at_rec_a loads what the voxel's ranges address (the head h of an IC
// row, the peak's orientation and column, the ISO column), at_rec_b the IC
// head's orientation, which needs h.
struct RecA {
    uint32_t h, eo, ecol, icol;                    // as loaded; selected by kind in phase b
    uint32_t i;
    int kind;                                      // 0 IC, 1 EC, 2 ISO, 3 past the voxel's rows
};

// Both phases load without per-lane branches: every lane loads from each
// compartment present (index 0 for lanes of another kind) and selects, so the
// loads go out together.  With a branch per kind the three kinds' loads share
// destination registers and each branch waited for the previous one's loads.
__device__ __forceinline__ RecA at_rec_a(const DevDict &d, const VRows &vr, uint32_t r)
{
    const uint32_t n1 = vr.n1, n2 = vr.n2, nic = vr.nic, nec = vr.nec, niso = vr.niso;
    const uint32_t icb = vr.icb, ecb = vr.ecb, isob = vr.isob;
    const int rfic = d.nIC, rfec = d.nEC, rfiso = d.nISO;
    r = min(r, vr.total - 1);
    const bool ic = r < n1, ec = !ic && r < n2;
    const uint32_t t = ic ? r : (ec ? r - n1 : r - n2);
    const uint32_t m = ic ? nic : (ec ? nec : niso);
    const int nrf = ic ? rfic : (ec ? rfec : rfiso);
    RecA a;
    a.i = nrf == 1 ? 0u : t / m;
    const uint32_t k = (ic ? icb : (ec ? ecb : isob)) + (t - a.i * m);
    a.kind = ic ? 0 : ec ? 1 : 2;
    a.h = a.eo = a.ecol = a.icol = 0u;
    if (d.nIC > 0 && d.n > 0) a.h = vr.uhead[ic ? k : 0u];
    if (d.nEC > 0 && d.nE > 0) {
        a.eo = d.ec.o[ec ? k : 0u];
        a.ecol = d.ec.col[ec ? k : 0u];
    }
    if (d.nISO > 0 && d.nI > 0) a.icol = d.iso.col[!ic && !ec ? k : 0u];
    return a;
}

// Phase b: the IC head's orientation (loaded by rec_o, so that the loads of
// both record windows are in flight together) and the record itself.
__device__ __forceinline__ uint32_t rec_o(const DevDict &d, const RecA &a)
{
    return d.n > 0 && d.nIC > 0 ? (uint32_t)d.ic.o[a.kind == 0 ? a.h : 0u] : 0u;
}

__device__ __forceinline__ void at_rec_b(const DevDict &d, const RecA &a, uint32_t o, uint32_t &code, uint32_t &slot,
                                         float &scale)
{
    const uint32_t base_ec = (uint32_t)d.nIC * (uint32_t)d.nF;
    const uint32_t base_iso = base_ec + (uint32_t)d.nEC * (uint32_t)d.nE;
    code = a.kind == 0 ? a.i * (uint32_t)d.ndirs + o
         : a.kind == 1 ? (1u << 30) | (a.i * (uint32_t)d.ndirs + a.eo)
                       : (2u << 30) | a.i;
    slot = a.kind == 0 ? a.i * (uint32_t)d.n + a.h
         : a.kind == 1 ? base_ec + a.i * (uint32_t)d.nE + a.ecol
                       : base_iso + a.i * (uint32_t)d.nI + a.icol;
    scale = a.kind == 0 ? 1.0f : -1.0f;
}

// One row record of the voxel's A^T list (see commit_hip_plan::mm_rec).
__device__ __forceinline__ uint64_t rec_pack(uint32_t code, uint32_t slot) { return (uint64_t)code | ((uint64_t)slot << 32); }

// A^T of one voxel over its rows [rb, rb + total) of the flat record list.
// Lane (g, q) takes row e0 + q of tile e0 for its A operand (chunk g of each
// 64-byte group) and D rows e0 + 4g + k for its outputs.  The records of the
// next two tiles are always in flight (one independent 8-byte load per lane
// per tile, clamped to the last row), so no load inside the tile loop waits
// on another.
template <int NT>
__device__ __forceinline__ void mm_vox_At(const float (&yb)[4 * NT], const DevDict &d,
                                          const uint64_t *__restrict__ rec, uint32_t rb, uint32_t total, int lane,
                                          float *__restrict__ z, float *__restrict__ Xr, int R, bool rhs_ok)
{
    const int g = lane >> 4, q = lane & 15;
    if (total == 0) return;
    const uint64_t *rr = rec + rb;
    const uint32_t last = total - 1;
    uint64_t r0 = rr[min((uint32_t)q, last)], r1 = rr[min(16u + q, last)];
    float4 a[NT];
    {
        const float *src = row_code(d, (uint32_t)r0, 0) + 4 * g;
#pragma unroll
        for (int u = 0; u < NT; u++) a[u] = *reinterpret_cast<const float4 *>(src + 16 * u);
    }
    for (uint32_t e0 = 0; e0 < total; e0 += 16) {
        const uint64_t r2 = rr[min(e0 + 32u + q, last)];              // tile e0 + 32's records
        // the next tile's rows; after the last tile the reloads read row 0 of
        // the IC table (a valid, cached address) and are never used
        const float *nsrc = e0 + 16 < total ? row_code(d, (uint32_t)r1, 0) + 4 * g : d.wmr + 4 * g;
        f4 acc0 = f4{0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
        for (int u = 0; u < NT; u++) {
            acc0 = mfma4(a[u].x, yb[4 * u + 0], acc0);
            acc1 = mfma4(a[u].y, yb[4 * u + 1], acc1);
            acc0 = mfma4(a[u].z, yb[4 * u + 2], acc0);
            acc1 = mfma4(a[u].w, yb[4 * u + 3], acc1);
            a[u] = *reinterpret_cast<const float4 *>(nsrc + 16 * u);
            // keep each reload behind its chunk's MFMAs: hoisted loads would
            // need a second register set for the A operand
            __builtin_amdgcn_sched_barrier(0);
        }
        const f4 acc = acc0 + acc1;
#pragma unroll
        for (int k = 0; k < 4; k++) {                    // D row 4g+k = row e0+4g+k, column q
            const uint64_t rk = (uint64_t)__shfl((long long)r0, 4 * g + k);
            const uint32_t slot = (uint32_t)(rk >> 32);
            if (e0 + 4 * g + k < total) {
                if ((uint32_t)rk >> 30 == 0u)
                    z[(size_t)slot * kTileR + q] = acc[k];                  // (setup_operator.py:412-428)
                else if (rhs_ok)
                    Xr[(size_t)slot * R] = acc[k];                          // (setup_operator.py:444-563)
            }
        }
        r0 = r1;
        r1 = r2;
    }
}

// The momentum point's residual of the batched fits, formed while A^T loads
// it (MOM): r(xhat) = (1 + coef) r(x) - coef r(prev_x) for xhat = x + coef (x -
// prev_x), since the residual is affine in x.  r(prev_x) is in r[*sel], the
// trial point's r(x) in r[1 - *sel].  Only the columns of fits that advanced
// are formed and written (col[problem * R + column].advance); their sums of
// rhat^2 go to part[v * 16 + q] (FP64, from the unrounded combination).
struct MomArgs {
    const float *r0, *r1;
    const int32_t *sel;
    const int32_t *vseg;                  // [nVOX] problem of each voxel
    const commit_hip_fista_col *col;
    double *part;
};
// (sel, vseg and col also go to the kernel as separate restrict arguments, so
// that the wave-uniform loads of sel and vseg[v] can be scalar loads)

// Two wavefronts per SIMD: the register budget (256) holds the B operand,
// the A operand of one tile and the records, with no LDS.
template <int NT, bool MOM>
__global__ __launch_bounds__(kBlock, 2) void k_At_mm(DevDict d, const float *__restrict__ Y,
                                                  const float *__restrict__ sqrtw, float *__restrict__ z,
                                                  float *__restrict__ X, int R, int rg,
                                                  const uint32_t *__restrict__ rptr,
                                                  const uint64_t *__restrict__ rec, MomArgs ma,
                                                  const int32_t *__restrict__ sel,
                                                  const int32_t *__restrict__ vseg,
                                                  const commit_hip_fista_col *__restrict__ col)
{
    // MOM: r(prev_x) of the voxel and its sqrtw row, staged per wavefront
    constexpr int kW = (16 * NT + kWave - 1) / kWave * kWave;             // whole 64-lane copies of nS <= 16 NT
    __shared__ __align__(16) float tiles[MOM ? kWavesPerBlock : 1][MOM ? 256 * NT : 1];
    __shared__ float wst[MOM ? kWavesPerBlock : 1][MOM ? kW : 1];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int v = wave_index(d.xcd_remap);
    if (v >= d.nVOX) return;
    const int g = lane >> 4, q = lane & 15;
    const int rhs = rg * kTileR + q;
    bool rhs_ok = rhs < R;
    // B operand: samples past nS and RHS columns past R read the voxel's last
    // sample / column 0 and are zeroed, so the loads are unconditional and in
    // flight together.  For a fixed K step the 64 lanes read 4 samples x 16
    // consecutive columns.  The row records follow (both phases of both
    // windows together).
    float yb[4 * NT];
    const size_t row0 = (size_t)v * d.nS;
    const int rc = rhs_ok ? rhs : 0;
    // at_sample(t, g) < nS  <=>  16 (t >> 2) + (t & 3) < L: one register, not
    // one sample index per K step
    const int L = d.nS - 4 * g;
    const uint32_t rb = rptr[v], total = rptr[v + 1] - rb;
    if constexpr (!MOM) {
#pragma unroll
        for (int t = 0; t < 4 * NT; t++) {
            const int s = at_sample(t, g);
            yb[t] = Y[(row0 + (s < d.nS ? s : d.nS - 1)) * R + rc];
        }
        if (sqrtw) {
            float wv[4 * NT];
#pragma unroll
            for (int t = 0; t < 4 * NT; t++) {
                const int s = at_sample(t, g);
                wv[t] = sqrtw[row0 + (s < d.nS ? s : d.nS - 1)];
            }
#pragma unroll
            for (int t = 0; t < 4 * NT; t++) yb[t] *= wv[t];
        }
#pragma unroll
        for (int t = 0; t < 4 * NT; t++)
            if (!(16 * (t >> 2) + (t & 3) < L && rhs_ok)) yb[t] = 0.0f;
    } else {
        const int sl = *sel;
        const float *rx = sl ? ma.r0 : ma.r1, *rp = sl ? ma.r1 : ma.r0;
        const commit_hip_fista_col &c = col[(size_t)vseg[v] * R + rc];
        const int adv = c.advance;                                 // requested before the rows
        const double coef = c.coef;
        // r(prev_x) of the voxel (nS * R <= 256 NT floats) and its sqrtw row
        // go to this wavefront's LDS by DMA (no registers held); r(x) goes to
        // registers.
        const int cnt = d.nS * R;
        const float *src = rp + row0 * R;
        float *T = tiles[w], *Wl = wst[w];
        if ((cnt & 3) == 0) {          // 16-byte copies (the rows start 16-byte aligned): 2% faster than 4-byte
            for (int j = 0; j * 4 * kWave < cnt; j++)
                __builtin_amdgcn_global_load_lds(src + min(j * 4 * kWave + 4 * lane, cnt - 4),
                                                 (lds_void *)(T + 4 * kWave * j), 16, 0, 0);
        } else {
            for (int j = 0; j * kWave < cnt; j++)
                __builtin_amdgcn_global_load_lds(src + min(j * kWave + lane, cnt - 1), (lds_void *)(T + kWave * j), 4,
                                                 0, 0);
        }
        if (sqrtw)
            for (int j = 0; j * kWave < d.nS; j++)
                __builtin_amdgcn_global_load_lds(sqrtw + row0 + min(j * kWave + lane, d.nS - 1),
                                                 (lds_void *)(Wl + kWave * j), 4, 0, 0);
        // Samples past nS (last lane group only) read the last sample's row
        // and stale LDS and are zeroed below.
#pragma unroll
        for (int t = 0; t < 4 * NT; t++) {
            const int s = at_sample(t, g);
            yb[t] = rx[(row0 + (s < d.nS ? s : d.nS - 1)) * R + rc];
        }
        // r(prev_x) and sqrtw are in LDS.  The builtin (vmcnt 0, the other
        // counters unconstrained), unlike inline asm, also tells the
        // compiler's wait insertion that every load so far is done.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        rhs_ok = rhs_ok && adv != 0;
        const double a = 1.0 + coef, b = -coef;
        const float *Tq = T + rc;                                  // sample s: Tq[s * R] (s * R + rc < 256 NT)
        double ss = 0.0;
#pragma unroll
        for (int t = 0; t < 4 * NT; t++) {
            const bool in = 16 * (t >> 2) + (t & 3) < L;
            const int s = in ? at_sample(t, g) : 0;
            const double u = a * (double)yb[t] + b * (double)Tq[s * R];   // as commit_hip_residual_combine
            ss = fma(in ? u : 0.0, u, ss);
            yb[t] = in && rhs_ok ? (sqrtw ? (float)u * Wl[s] : (float)u) : 0.0f;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        ss = swap32_add(ss, ss);           // lanes q, q+16, q+32, q+48 hold column q
        ss = swap16_add(ss, ss);
        if (g == 0) ma.part[(size_t)v * kTileR + q] = rhs_ok ? ss : 0.0;
    }
    mm_vox_At<NT>(yb, d, rec, rb, total, lane, z, X + rhs, R, rhs_ok);
}

// Pass 2: X[(f + i nF)*R + rhs] = eval * sum over the fibre's entries (ascending
// voxel order) of len * z[head], in FP64.  16 lanes per (fibre, RF): one
// 64-byte row of z per entry; heads and lengths are read in fibre order.
// With col (the momentum form), only the columns of fits that advanced are written.
__global__ __launch_bounds__(kBlock) void k_At_fibre_mm(DevDict d, const float *__restrict__ z,
                                                        float *__restrict__ X, int R, int rg,
                                                        const uint32_t *__restrict__ hpos,
                                                        const float *__restrict__ lenF,
                                                        const commit_hip_fista_col *__restrict__ col,
                                                        const int32_t *__restrict__ segx)
{
    const size_t total = (size_t)d.nF * d.nIC * kTileR;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(t & (kTileR - 1));
        const size_t gcol = t / kTileR;                  // f + i*nF
        const int i = (int)(gcol / d.nF);
        const uint32_t f = (uint32_t)(gcol - (size_t)i * d.nF);
        const float *zi = z + (size_t)i * d.n * kTileR + q;
        const uint32_t kb = d.fib_ptr[f], ke = d.fib_ptr[f + 1];
        double acc = 0.0;
        uint32_t k = kb;
        for (; k + 8 <= ke; k += 8) {
            float v[8], l[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                v[u] = zi[(size_t)hpos[k + u] * kTileR];
                l[u] = lenF[k + u];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc += (double)l[u] * (double)v[u];
        }
        for (; k < ke; k++) acc += (double)lenF[k] * (double)zi[(size_t)hpos[k] * kTileR];
        const int rhs = rg * kTileR + q;
        if (rhs < R && (!col || col[(size_t)segx[gcol] * R + rhs].advance))
            X[gcol * R + rhs] = d.eval[gcol] ? (float)acc : 0.0f;
    }
}

// Run heads of the multi-RHS A^T (built once per plan).  Entry k of voxel v
// heads a run when it is the voxel's first or its orientation differs from
// entry k-1's (a voxel's entries are sorted by orientation).
__global__ __launch_bounds__(kBlock) void k_mm_heads(const uint32_t *__restrict__ ptr, const uint16_t *__restrict__ o,
                                                     int nVOX, uint32_t *__restrict__ headof,
                                                     uint32_t *__restrict__ cnt)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nVOX) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    uint32_t h = b, c = 0;
    for (uint32_t k = b; k < e; k++) {
        if (k == b || o[k] != o[k - 1]) {
            h = k;
            c++;
        }
        headof[k] = h;
    }
    cnt[v] = c;
}

__global__ __launch_bounds__(kBlock) void k_mm_uhead(const uint32_t *__restrict__ ptr, const uint16_t *__restrict__ o,
                                                     const uint32_t *__restrict__ uptr, int nVOX,
                                                     uint32_t *__restrict__ uhead)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nVOX) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    uint32_t j = uptr[v];
    for (uint32_t k = b; k < e; k++)
        if (k == b || o[k] != o[k - 1]) uhead[j++] = k;
}

__global__ __launch_bounds__(kBlock) void k_mm_fibre_order(const uint32_t *__restrict__ fib_perm,
                                                           const uint32_t *__restrict__ headof,
                                                           const float *__restrict__ len, uint32_t n,
                                                           uint32_t *__restrict__ hpos, float *__restrict__ lenF)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t e = fib_perm[k];
        hpos[k] = headof[e];
        lenF[k] = len[e];
    }
}

// Rows of each voxel's A^T list: IC heads for each IC RF, EC peaks for each
// EC RF, ISO entries for each ISO RF (setup_operator.py:370-563).
__global__ __launch_bounds__(kBlock) void k_mm_rows_count(DevDict d, const uint32_t *__restrict__ uptr,
                                                          uint32_t *__restrict__ cnt)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.nVOX) return;
    const VRows vr = uptr ? vrows_at(d, v, uptr, nullptr) : vrows(d, v);
    cnt[v] = vr.total;
}

// Each voxel's records (LUT row code | output slot << 32), in list order.
__global__ __launch_bounds__(kBlock) void k_mm_rows_fill(DevDict d, const uint32_t *__restrict__ uptr,
                                                         const uint32_t *__restrict__ uhead,
                                                         const uint32_t *__restrict__ rptr,
                                                         uint64_t *__restrict__ rec)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.nVOX) return;
    const VRows vr = uptr ? vrows_at(d, v, uptr, uhead) : vrows(d, v);
    uint64_t *out = rec + rptr[v];
    for (uint32_t r = 0; r < vr.total; r++) {
        const RecA a = at_rec_a(d, vr, r);
        uint32_t code, slot;
        float scale;
        at_rec_b(d, a, rec_o(d, a), code, slot, scale);
        out[r] = rec_pack(code, slot);
    }
}

int build_mm_rows(commit_hip_plan *p)
{
    const DevDict &d = p->d;
    const int nV = d.nVOX;
    void *a = nullptr, *b = nullptr;
    uint32_t *cnt = nullptr;
    int rc;
    if ((rc = commit_internal::dev_alloc(p, (size_t)(nV + 1) * sizeof(uint32_t), &a))) return rc;
    if (hipMalloc((void **)&cnt, (size_t)(nV + 1) * sizeof(uint32_t)) != hipSuccess)
        return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", (size_t)(nV + 1) * sizeof(uint32_t));
    std::vector<uint32_t> h((size_t)nV + 1, 0u);
    const unsigned gv = (unsigned)((nV + kBlock - 1) / kBlock);
    const uint32_t *uptr = d.n > 0 && d.nIC > 0 ? p->mm_uptr : nullptr;
    if (gv > 0) hipLaunchKernelGGL(k_mm_rows_count, dim3(gv), dim3(kBlock), 0, nullptr, d, uptr, cnt);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(h.data(), cnt, (size_t)nV * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(cnt);
    uint64_t run = 0;
    if (err == hipSuccess) {
        for (int v = 0; v < nV; v++) {                     // exclusive scan: first row of each voxel
            const uint32_t cv = h[v];
            h[v] = (uint32_t)run;
            run += cv;
        }
        if (run >= 0xFFFFFFFFull) return fail(COMMIT_HIP_EUNSUPPORTED, "%llu A^T rows", (unsigned long long)run);
        h[nV] = (uint32_t)run;
        err = hipMemcpy(a, h.data(), (size_t)(nV + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) return fail(COMMIT_HIP_EDEVICE, "A^T rows: %s", hipGetErrorString(err));
    if ((rc = commit_internal::dev_alloc(p, (size_t)std::max<uint64_t>(run, 1u) * sizeof(uint64_t), &b))) return rc;
    if (gv > 0)
        hipLaunchKernelGGL(k_mm_rows_fill, dim3(gv), dim3(kBlock), 0, nullptr, d, uptr, p->mm_uhead,
                           static_cast<uint32_t *>(a), static_cast<uint64_t *>(b));
    err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(nullptr);
    if (err != hipSuccess) return fail(COMMIT_HIP_EDEVICE, "A^T rows: %s", hipGetErrorString(err));
    p->mm_rptr = static_cast<uint32_t *>(a);
    p->mm_rec = static_cast<uint64_t *>(b);             // set last: a failed build is redone by the next call
    return 0;
}

// Builds the head lists on the null stream and waits for them; mm_uhead is set
// last, so a failed build is redone by the next call.
int build_mm_heads(commit_hip_plan *p)
{
    const DevDict &d = p->d;
    const int nV = d.nVOX;
    const uint32_t n = (uint32_t)d.n;
    void *a = nullptr, *b = nullptr, *c = nullptr, *e = nullptr;
    int rc;
    if ((rc = commit_internal::dev_alloc(p, (size_t)(nV + 1) * sizeof(uint32_t), &a))) return rc;
    if ((rc = commit_internal::dev_alloc(p, (size_t)n * sizeof(uint32_t), &b))) return rc;
    if ((rc = commit_internal::dev_alloc(p, (size_t)n * sizeof(float), &c))) return rc;
    p->mm_uptr = static_cast<uint32_t *>(a);
    p->mm_hpos = static_cast<uint32_t *>(b);
    p->mm_lenF = static_cast<float *>(c);
    uint32_t *headof = nullptr, *cnt = nullptr;
    if (hipMalloc((void **)&headof, (size_t)n * sizeof(uint32_t)) != hipSuccess)
        return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", (size_t)n * sizeof(uint32_t));
    if (hipMalloc((void **)&cnt, (size_t)(nV + 1) * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(headof);
        return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed", (size_t)(nV + 1) * sizeof(uint32_t));
    }
    std::vector<uint32_t> h((size_t)nV + 1, 0u);
    const unsigned gv = (unsigned)((nV + kBlock - 1) / kBlock);
    if (gv > 0) hipLaunchKernelGGL(k_mm_heads, dim3(gv), dim3(kBlock), 0, nullptr, d.ic.ptr, d.ic.o, nV, headof, cnt);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(h.data(), cnt, (size_t)nV * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (err == hipSuccess) {
        uint64_t run = 0;                                  // exclusive scan: first head of each voxel
        for (int v = 0; v < nV; v++) {
            const uint32_t cv = h[v];
            h[v] = (uint32_t)run;
            run += cv;
        }
        h[nV] = (uint32_t)run;
        err = hipMemcpy(p->mm_uptr, h.data(), (size_t)(nV + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (err == hipSuccess) {
        if ((rc = commit_internal::dev_alloc(p, (size_t)std::max<uint32_t>(h[nV], 1u) * sizeof(uint32_t), &e))) {
            (void)hipFree(headof);
            (void)hipFree(cnt);
            return rc;
        }
        if (gv > 0)
            hipLaunchKernelGGL(k_mm_uhead, dim3(gv), dim3(kBlock), 0, nullptr, d.ic.ptr, d.ic.o, p->mm_uptr, nV,
                               static_cast<uint32_t *>(e));
        const unsigned gn = (unsigned)std::min<uint32_t>((n + kBlock - 1) / kBlock, 1u << 16);
        if (n > 0)
            hipLaunchKernelGGL(k_mm_fibre_order, dim3(gn), dim3(kBlock), 0, nullptr, d.fib_perm, headof, d.ic.len, n,
                               p->mm_hpos, p->mm_lenF);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipStreamSynchronize(nullptr);
    }
    (void)hipFree(headof);
    (void)hipFree(cnt);
    if (err != hipSuccess) return fail(COMMIT_HIP_EDEVICE, "run heads: %s", hipGetErrorString(err));
    p->mm_uhead = static_cast<uint32_t *>(e);
    return 0;
}

int nt_of(const DevDict &d) { return (d.nS + 15) / 16; }

int check_multi(commit_hip_plan *p, int nrhs, int dtype)
{
    if (dtype != COMMIT_HIP_FP32) return fail(COMMIT_HIP_EUNSUPPORTED, "multi-RHS products are built for FP32 only");
    if (nrhs < 1) return fail(COMMIT_HIP_EINVAL, "nrhs=%d", nrhs);
    if (p->nolut) return fail(COMMIT_HIP_EUNSUPPORTED, "multi-RHS products of a LUT-free operator");
    if (nt_of(p->d) > kMaxNT)
        return fail(COMMIT_HIP_EUNSUPPORTED, "multi-RHS products need nS <= %d (nS=%d)", kMaxNT * 16, p->d.nS);
    return 0;
}

// one or two wavefronts per (voxel, RHS group); with two, each owns half of
// the sample tiles (NT rounded up to even; the padded LUT rows hold zeros
// there).  The choice is a plan option (commit_hip_plan_set_option).
template <bool RES>
int launch_A_mm(commit_hip_plan *plan, const float *x, float *y, int nrhs, const ResArgs &ra, hipStream_t st)
{
    const DevDict &d = plan->d;
    const int nrg = (nrhs + kTileR - 1) / kTileR;
    const size_t waves = (size_t)d.nVOX * nrg * plan->multi_a_waves;
    if (waves == 0) return 0;
    if (waves > (size_t)INT32_MAX - kWavesPerBlock) return fail(COMMIT_HIP_EUNSUPPORTED, "too many voxels x RHS groups");
    const unsigned grid = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    const uint32_t *ev = plan->eval_ones ? nullptr : d.eval;
    if (plan->multi_a_waves == 2) {
        switch ((nt_of(d) + 1) / 2) {
#define CASE(NH)                                                                                            \
    case NH:                                                                                                \
        hipLaunchKernelGGL((k_A_mm<NH, 2, RES>), dim3(grid), dim3(kBlock), 0, st, d, x, y, nrhs, nrg, ev, ra); \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
            default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d", d.nS);
        }
    } else {
        switch (nt_of(d)) {
#define CASE(NT)                                                                                            \
    case NT:                                                                                                \
        hipLaunchKernelGGL((k_A_mm<NT, 1, RES>), dim3(grid), dim3(kBlock), 0, st, d, x, y, nrhs, nrg, ev, ra); \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
            CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17) CASE(18) CASE(19) CASE(20)
#undef CASE
            default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d", d.nS);
        }
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int check_layout(const commit_hip_plan *plan, const commit_hip_multi_layout *L)
{
    const DevDict &d = plan->d;
    if (L->m != (int64_t)d.nVOX * d.nS) return fail(COMMIT_HIP_EINVAL, "layout rows %lld != plan rows", (long long)L->m);
    if (L->nprob < 1 || L->nypieces < 1 || L->nypieces > INT32_MAX || !L->ypieces || !L->ypseg)
        return fail(COMMIT_HIP_EINVAL, "layout pieces missing");
    return 0;
}

// The fused A X residual (res chosen by ra) and its per-fit norms.
int residual_norms(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *X, ResArgs ra, double *rsq,
                   double *work, hipStream_t st)
{
    const DevDict &d = plan->d;
    const int R = L->nrhs, nrg = (R + kTileR - 1) / kTileR;
    double *part = work, *pp = work + (size_t)d.nVOX * 2 * nrg * kTileR;
    ra.part = part;
    int rc;
    if ((rc = launch_A_mm<true>(plan, X, nullptr, R, ra, st))) return rc;
    hipLaunchKernelGGL(k_res_pieces, dim3((unsigned)L->nypieces), dim3(kBlock), 0, st, part, L->ypieces, d.nS, d.nVOX,
                       plan->multi_a_waves, nrg, R, pp);
    hipLaunchKernelGGL(k_res_problems, dim3((unsigned)((L->nprob * R + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(kBlock), 0, st, pp, L->ypseg, (int)L->nypieces, L->nprob, R, rsq);
    HIP_TRY(hipGetLastError());
    return 0;
}

// A^T over 16-RHS groups: pass 1 (k_At_mm) then the per-fibre sums.  MOM: the
// momentum form (Y formed from ma, columns of fits that advanced only; nrhs <= 16).
template <bool MOM>
int launch_At_mm(commit_hip_plan *plan, const float *y, const float *w, float *x, int nrhs, const MomArgs &ma,
                 const int32_t *segx, hipStream_t st)
{
    const DevDict &d = plan->d;
    int rc;
    if (!plan->zmulti && d.n > 0 && d.nIC > 0) {
        void *zp = nullptr;
        if ((rc = commit_internal::dev_alloc(plan, (size_t)d.n * d.nIC * kTileR * sizeof(float), &zp))) return rc;
        plan->zmulti = static_cast<float *>(zp);
    }
    if (!plan->mm_uhead && d.n > 0 && d.nIC > 0 && (rc = build_mm_heads(plan))) return rc;
    if (!plan->mm_rec && (rc = build_mm_rows(plan))) return rc;
    const int nrg = (nrhs + kTileR - 1) / kTileR;
    const unsigned grid = (unsigned)((d.nVOX + kWavesPerBlock - 1) / kWavesPerBlock);
    const size_t total = (size_t)d.nF * d.nIC * kTileR;
    const unsigned g2 = (unsigned)std::min<size_t>((total + kBlock - 1) / kBlock, 1u << 20);
    for (int rg = 0; rg < nrg; rg++) {
        if (grid > 0) {
            switch (nt_of(d)) {
#define CASE(NT)                                                                                            \
    case NT:                                                                                                \
        hipLaunchKernelGGL((k_At_mm<NT, MOM>), dim3(grid), dim3(kBlock), 0, st, d, y, w, plan->zmulti, x, nrhs, \
                           rg, plan->mm_rptr, plan->mm_rec, ma, ma.sel, ma.vseg, ma.col);                   \
        break;
                CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
                CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17) CASE(18) CASE(19) CASE(20)
#undef CASE
                default: return fail(COMMIT_HIP_EUNSUPPORTED, "nS=%d", d.nS);
            }
            HIP_TRY(hipGetLastError());
        }
        if (total > 0) {
            hipLaunchKernelGGL(k_At_fibre_mm, dim3(g2), dim3(kBlock), 0, st, d, plan->zmulti, x, nrhs, rg, plan->mm_hpos,
                               plan->mm_lenF, MOM ? ma.col : nullptr, MOM ? segx : nullptr);
            HIP_TRY(hipGetLastError());
        }
    }
    return 0;
}

// Roles of the two residual buffers after a momentum step: r(prev_x) of every
// unfinished fit must be in r[*sel].  Fits that advanced have it in r[1 - *sel]
// (the trial point became prev_x), fits that backtracked in r[*sel].  When
// both kinds exist the advanced columns are copied into r[*sel]; when no
// unfinished fit backtracked the roles swap (k_roles_flip).  Finished fits are
// ignored.
__global__ __launch_bounds__(kBlock) void k_roles_copy(int64_t m, int R, int nS, const int32_t *__restrict__ vseg,
                                                       const int32_t *__restrict__ sel,
                                                       const commit_hip_fista_col *__restrict__ col, int nq,
                                                       float *__restrict__ r0, float *__restrict__ r1)
{
    int adv = 0, back = 0;
    for (int q = threadIdx.x; q < nq; q += kBlock) {
        adv |= col[q].advance != 0;
        back |= !col[q].done && !col[q].advance;
    }
    adv = __syncthreads_or(adv);
    back = __syncthreads_or(back);
    if (!(adv && back)) return;
    const int sl = *sel;
    const float *src = sl ? r0 : r1;
    float *dst = sl ? r1 : r0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBlock) {
        const commit_hip_fista_col *c = col + (size_t)vseg[i / nS] * R;
        for (int r = 0; r < R; r++)
            if (c[r].advance) dst[i * R + r] = src[i * R + r];
    }
}

__global__ __launch_bounds__(kBlock) void k_roles_flip(const commit_hip_fista_col *__restrict__ col, int nq,
                                                       int32_t *__restrict__ sel)
{
    int back = 0;
    for (int q = threadIdx.x; q < nq; q += kBlock) back |= !col[q].done && !col[q].advance;
    back = __syncthreads_or(back);
    if (!back && threadIdx.x == 0) *sel = 1 - *sel;
}

}  // namespace

extern "C" {

int commit_hip_A_multi(commit_hip_plan *plan, const void *X, void *Y, int nrhs, int dtype, void *stream)
{
    if (!plan || !X || !Y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    int rc = check_multi(plan, nrhs, dtype);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(plan->device));
    return launch_A_mm<false>(plan, static_cast<const float *>(X), static_cast<float *>(Y), nrhs, ResArgs{},
                              static_cast<hipStream_t>(stream));
}

int64_t commit_hip_A_multi_residual_work(const commit_hip_plan *plan, const commit_hip_multi_layout *L)
{
    if (!plan || !L || L->nrhs < 1 || L->nypieces < 1) return -1;
    const int64_t nrg = (L->nrhs + kTileR - 1) / kTileR;
    return (int64_t)plan->d.nVOX * 2 * nrg * kTileR + L->nypieces * L->nrhs;
}

int commit_hip_A_multi_residual(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *X,
                                const float *y, int y_per_column, const float *sqrtw, float *res, double *rsq,
                                double *work, void *stream)
{
    if (!plan || !L || !X || !rsq || !work) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    int rc = check_multi(plan, L->nrhs, COMMIT_HIP_FP32);
    if (rc) return rc;
    if (L->nrhs > kTileR) return fail(COMMIT_HIP_EUNSUPPORTED, "batched fits take 1..%d columns", kTileR);
    if ((rc = check_layout(plan, L))) return rc;
    HIP_TRY(hipSetDevice(plan->device));
    ResArgs ra{y, y_per_column ? 1 : 0, sqrtw, res, nullptr, nullptr, nullptr};
    return residual_norms(plan, L, X, ra, rsq, work, static_cast<hipStream_t>(stream));
}

int commit_hip_At_multi(commit_hip_plan *plan, const void *Y, const void *sqrtw, void *X, int nrhs, int dtype,
                        void *stream)
{
    if (!plan || !X || !Y) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    int rc = check_multi(plan, nrhs, dtype);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(plan->device));
    return launch_At_mm<false>(plan, static_cast<const float *>(Y), static_cast<const float *>(sqrtw),
                               static_cast<float *>(X), nrhs, MomArgs{}, nullptr, static_cast<hipStream_t>(stream));
}

int commit_hip_A_multi_trial(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *X, const float *y,
                             int y_per_column, const float *sqrtw, float *r0, float *r1, const int32_t *sel,
                             double *rsq, double *work, void *stream)
{
    if (!plan || !L || !X || !r0 || !r1 || !sel || !rsq || !work) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    int rc = check_multi(plan, L->nrhs, COMMIT_HIP_FP32);
    if (rc) return rc;
    if (L->nrhs > kTileR) return fail(COMMIT_HIP_EUNSUPPORTED, "batched fits take 1..%d columns", kTileR);
    if ((rc = check_layout(plan, L))) return rc;
    HIP_TRY(hipSetDevice(plan->device));
    ResArgs ra{y, y_per_column ? 1 : 0, sqrtw, r0, nullptr, r1, sel};
    return residual_norms(plan, L, X, ra, rsq, work, static_cast<hipStream_t>(stream));
}

int commit_hip_At_multi_momentum(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *r0,
                                 const float *r1, const int32_t *sel, const float *sqrtw,
                                 const commit_hip_fista_col *col, float *grad, double *rsq, double *work,
                                 void *stream)
{
    if (!plan || !L || !r0 || !r1 || !sel || !col || !grad || !rsq || !work)
        return fail(COMMIT_HIP_EINVAL, "NULL argument");
    int rc = check_multi(plan, L->nrhs, COMMIT_HIP_FP32);
    if (rc) return rc;
    if (L->nrhs > kTileR) return fail(COMMIT_HIP_EUNSUPPORTED, "batched fits take 1..%d columns", kTileR);
    if ((rc = check_layout(plan, L))) return rc;
    if (!L->vseg || !L->segx) return fail(COMMIT_HIP_EINVAL, "layout: vseg and segx are needed");
    if (reinterpret_cast<uintptr_t>(r0) % 16 || reinterpret_cast<uintptr_t>(r1) % 16)
        return fail(COMMIT_HIP_EINVAL, "residual buffers must be 16-byte aligned");
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const DevDict &d = plan->d;
    const int R = L->nrhs;
    double *part = work, *pp = work + (size_t)d.nVOX * kTileR;
    const MomArgs ma{r0, r1, sel, L->vseg, col, part};
    if ((rc = launch_At_mm<true>(plan, nullptr, sqrtw, grad, R, ma, L->segx, st))) return rc;
    hipLaunchKernelGGL(k_res_pieces, dim3((unsigned)L->nypieces), dim3(kBlock), 0, st, part, L->ypieces, d.nS, d.nVOX,
                       1, 1, R, pp);
    hipLaunchKernelGGL(k_res_problems, dim3((unsigned)((L->nprob * R + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(kBlock), 0, st, pp, L->ypseg, (int)L->nypieces, L->nprob, R, rsq);
    HIP_TRY(hipGetLastError());
    return 0;
}

int commit_hip_fista_multi_roles(const commit_hip_plan *plan, const commit_hip_multi_layout *L, float *r0, float *r1,
                                 int32_t *sel, const commit_hip_fista_col *col, void *stream)
{
    if (!plan || !L || !r0 || !r1 || !sel || !col) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (L->nrhs < 1 || L->nrhs > kTileR || L->nprob < 1) return fail(COMMIT_HIP_EINVAL, "layout sizes");
    if (!L->vseg) return fail(COMMIT_HIP_EINVAL, "layout: vseg is needed");
    const DevDict &d = plan->d;
    if (L->m != (int64_t)d.nVOX * d.nS) return fail(COMMIT_HIP_EINVAL, "layout rows %lld != plan rows", (long long)L->m);
    HIP_TRY(hipSetDevice(plan->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nq = L->nprob * L->nrhs;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((L->m + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL(k_roles_copy, dim3(grid), dim3(kBlock), 0, st, L->m, L->nrhs, d.nS, L->vseg, sel, col, nq, r0, r1);
    hipLaunchKernelGGL(k_roles_flip, dim3(1), dim3(kBlock), 0, st, col, nq, sel);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
