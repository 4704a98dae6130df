// LDS-slab products: y = A x and pass 1 of x = A^T y with the IC table served
// from LDS instead of L2.
//
// Reference algorithm: setup_operator.py:48-109 (IC part of COMMIT_A__block)
// and :370-442 (IC part of COMMIT_At__block).
//
// The per-voxel kernels of commit_hip.hip read one padded LUT row (nSp floats)
// per IC segment from L2; on the 1M-fibre head that is ~100 GB per product,
// and the L2-to-CU path bounds them.  Here a work-group of W wavefronts owns
// W*V voxels (V per wavefront, accumulators or y rows in registers) and walks
// the IC table in slabs of S orientations: the slab is copied once into LDS
// (global_load_lds, 16 bytes per lane), then every wavefront applies the
// entries of its voxels whose orientation falls in the slab, reading rows from
// LDS.  The plan keeps each voxel's IC entries sorted by orientation
// (stable), so a voxel's entries of one slab are a contiguous range; their
// offsets are a [nVOX][nsl+1] table.  Voxels are dealt to the wavefronts of a
// work-group by their work in each slab (distinct orientations there), so the
// wavefronts reach each slab barrier together; work-groups are numbered
// heaviest first.
// EC and ISO (a few entries per voxel) use rows from global memory after the
// slab walk: the items of all of a wavefront's voxels are decoded lane-parallel
// and each voxel's rows are fetched together.
//
//   A x   : per run of entries of equal orientation in a voxel (their
//           coefficients added first), NJ FP64 FMAs per lane from an LDS
//           row; zero coefficients skipped with a ballot
//           (setup_operator.py:88).
//   A^T y : per distinct (voxel, orientation), a per-lane partial dot
//           product; 8 of them are reduced together over the 64 lanes with
//           VALU cross-lane moves (permlane32/16 swaps and DPP).  Entries of
//           equal orientation in a voxel share the dot.  The IC product
//           l * dot goes to its fibre-major scratch slot, summed per fibre by
//           pass 2 (k_At_fibre).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/commit_hip.h"
#include "common.h"
#include "plan.h"
#include "rows.h"

using commit_internal::fail;

namespace {

using commit_internal::SlabArgs;

typedef __attribute__((address_space(3))) void lds_void;

// LDS floats per work-group: 160 KiB shared by the 16 wavefronts of a CU.
// The last W*kBcastFloats floats hold one broadcast table per wavefront; the
// rest holds the slab (S rows packed at rs = nS rounded up to 4 floats) and a
// zeroed pad of nSp - rs floats after it.
template <int W>
struct SlabLds {
    static constexpr int floats = 40960 * W / 16;
};

// Per-wavefront table of the current chunk: coefficients (A x) or dot
// products (A^T y) and row offsets inside the slab, read back by the whole
// wavefront at one address (an LDS broadcast) instead of lane reads on the
// VALU.  C: the value type (float for the float32 A x, double otherwise).
template <typename C>
struct Bcast {
    C c[kWave];
    int32_t off[kWave];
};
// LDS floats reserved for each wavefront's table (room for the double one)
constexpr int kBcastFloats = (int)(sizeof(Bcast<double>) / sizeof(float));
static_assert(sizeof(Bcast<double>) % 16 == 0 && sizeof(Bcast<float>) % 16 == 0,
              "broadcast tables must stay 16-byte aligned");

template <int W, typename C>
__device__ __forceinline__ Bcast<C> *bcast_table(float *L, int w)
{
    return reinterpret_cast<Bcast<C> *>(L + SlabLds<W>::floats - (w + 1) * kBcastFloats);
}

// Copy rows LUT rows (stride nSp floats in memory, from src) into L packed at
// stride rs floats (nS rounded up to 4): 1 KiB pieces of the packed slab,
// 16 bytes per lane, wavefront w takes pieces w, w+W, ...; the caller waits
// and synchronises.
template <int W>
__device__ __forceinline__ void slab_load(float *L, const float *__restrict__ src, int rows, int rs, int nSp, int w,
                                          int lane)
{
    const int nfloats = rows * rs;
    const int npieces = (nfloats + 255) / 256;
    for (int pc = w; pc < npieces; pc += W) {
        const int f = pc * 256 + 4 * lane;
        const int r = f / rs;
        if (f < nfloats)
            __builtin_amdgcn_global_load_lds(src + (size_t)r * nSp + (f - r * rs), (lds_void *)(L + pc * 256), 16, 0,
                                             0);
    }
}

// The lanes of a row's last samples read up to nSp - rs floats past the row
// (rows are packed at rs floats in LDS): past the slab's last row they read
// this pad, zeroed once per work-group, so those lanes see finite values.
// Products of padded samples are discarded (A x stores samples < nS only) or
// multiplied by a zero y (A^T y).
template <int W>
__device__ __forceinline__ void slab_pad_zero(float *L, const SlabArgs &sa, int nSp)
{
    for (int k = (int)threadIdx.x; k < nSp - sa.rs; k += W * kWave) L[sa.S * sa.rs + k] = 0.0f;
}

// One LUT row of the slab in the lane layout (Lanes<NJ, Q>): Lq is the row's
// sample 0 plus 4 lanes, Ls the same plus one lane; a 16-byte LDS read per
// quad, a 4-byte read per single sample (padded samples are zero).
template <int NJ, bool Q>
__device__ __forceinline__ void lds_row(float (&raw)[NJ], const float *Lq, const float *Ls)
{
    constexpr int NQ = Lanes<NJ, Q>::NQ, NR = Lanes<NJ, Q>::NR;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const float4 v = *reinterpret_cast<const float4 *>(Lq + 256 * q);
        raw[4 * q] = v.x, raw[4 * q + 1] = v.y, raw[4 * q + 2] = v.z, raw[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) raw[4 * NQ + r] = Ls[256 * NQ + 64 * r];
}

// Entry records of the first chunk (up to 64 entries) of a range [b, e): one
// per lane, clamped to a valid entry (b, or entry 0 for an empty range).
struct Rec {
    uint32_t o, col;      // orientation; x column (A x) or fibre-major slot (A^T y)
    float len;
};

__device__ __forceinline__ Rec rec_load(const Comp &cp, const uint32_t *__restrict__ colv, uint32_t b, uint32_t e,
                                        int lane)
{
    const uint32_t k = b + lane;
    const uint32_t ks = k < e ? k : (b < e ? b : 0u);
    return Rec{stream_load(cp.o + ks), stream_load(colv + ks), stream_load(cp.len + ks)};
}

// Nonzero coefficients in a chunk from which A x looks for runs of one
// orientation to merge.
constexpr int kMergeMin = 12;

// A x: apply up to 64 entries of one voxel and one slab, rows in LDS (Lq, Ls:
// this lane's bases in the slab's first orientation o0, see lds_row); lane l
// holds entry b+l (record r, x value xv), valid when ok.  The nonzero
// coefficients and their row offsets go, in order, to slots 0..nz-1 of the
// wavefront's broadcast table (zero coefficients are skipped,
// setup_operator.py:88); the other lanes fill the slots past nz with zero
// coefficients and valid rows.  Entries are then applied in groups of 4 with
// all four rows in flight; a group's tail adds exact zeros.  T = double: FP64
// FMAs of the FP32 rows.  T = float (blocked FP32): each group's four
// products are summed with FP32 FMAs and the group sum is added to the FP64
// accumulators, so the FP32 rounding is that of a 4-term sum.
template <int NJ, typename T>
__device__ __forceinline__ void slab_a_chunk(double (&acc)[NJ], const Rec &r, T xv, bool ok, const float *Lq,
                                             const float *Ls, Bcast<T> *bt, int o0, int rs, int lane)
{
    // IC coefficient l * x * eval                                  (setup_operator.py:63-96)
    double c = (double)xv;
    c *= (double)r.len;
    if (!ok) c = 0.0;
    // A voxel's entries are sorted by orientation, so entries sharing a LUT
    // row are adjacent: the first of each run of nonzero coefficients of one
    // orientation adds the run's coefficients in entry order and applies the
    // row once; the others drop out (one step per entry of the longest run).
    // Chunks with few nonzero coefficients (a sparse x) rarely hold such runs
    // and skip the search.
    const unsigned long long m0 = __ballot(c != 0.0);
    if (m0 == 0) return;
    unsigned long long cm = 0;
    if (__popcll(m0) >= kMergeMin) {
        const uint32_t on = (uint32_t)__shfl((int)r.o, lane < kWave - 1 ? lane + 1 : lane);
        cm = __ballot(lane < kWave - 1 && ((m0 >> lane) & 1ull) && ((m0 >> (lane + 1)) & 1ull) && on == r.o);
    }
    if (cm != 0) {
        const bool cont = (cm >> lane) & 1ull;                        // entry b+l+1 is in my run
        const bool head = lane == 0 || !((cm >> (lane - 1)) & 1ull);
        bool alive = head && cont;
        double s = c;
        for (int k = 1; __ballot(alive) != 0; k++) {
            const double ck = __shfl(c, lane + k < kWave ? lane + k : lane);
            if (alive) s += ck;
            alive = alive && lane + k < kWave && ((cm >> (lane + k)) & 1ull);
        }
        c = head ? s : 0.0;
    }
    const bool nzl = c != 0.0;
    const unsigned long long m = cm != 0 ? __ballot(nzl) : m0;
    const int nz = __popcll(m);
    if (nz == 0) return;
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const int dst = nzl ? below : nz + (lane - below);
    bt->c[dst] = (T)c;
    bt->off[dst] = ((int)r.o - o0) * rs;
    for (int t = 0; t < nz; t += 4) {
        const int4 of = *reinterpret_cast<const int4 *>(&bt->off[t]);
        float ra[NJ], rb[NJ], rc[NJ], rd[NJ];
        lds_row<NJ, kQuads<T>>(ra, Lq + of.x, Ls + of.x);
        lds_row<NJ, kQuads<T>>(rb, Lq + of.y, Ls + of.y);
        lds_row<NJ, kQuads<T>>(rc, Lq + of.z, Ls + of.z);
        lds_row<NJ, kQuads<T>>(rd, Lq + of.w, Ls + of.w);
        if constexpr (sizeof(T) == sizeof(float)) {
            const float4 cf = *reinterpret_cast<const float4 *>(&bt->c[t]);
            float g[NJ];
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                g[j] = cf.x * ra[j];
                g[j] = fmaf(cf.y, rb[j], g[j]);
                g[j] = fmaf(cf.z, rc[j], g[j]);
                g[j] = fmaf(cf.w, rd[j], g[j]);
            }
#pragma unroll
            for (int j = 0; j < NJ; j++) acc[j] += (double)g[j];
        } else {
            const double2 c01 = *reinterpret_cast<const double2 *>(&bt->c[t]);
            const double2 c23 = *reinterpret_cast<const double2 *>(&bt->c[t + 2]);
            fma_row<NJ>(acc, c01.x, ra);
            fma_row<NJ>(acc, c01.y, rb);
            fma_row<NJ>(acc, c23.x, rc);
            fma_row<NJ>(acc, c23.y, rd);
        }
    }
}

// One slab pass of A x over the wavefront's V voxels (ranges rb[u]..re[u]).
// The first chunk of every voxel is software-pipelined: the entry records of
// voxel u+2 and the x gathers of voxel u+1 are in flight while voxel u is
// applied.  Later chunks of a long range are loaded in place.
template <int NJ, int V, typename T>
__device__ __forceinline__ void slab_a_pass(double (&acc)[V][NJ], const Comp &cp, const uint32_t (&rb)[V],
                                            const uint32_t (&re)[V], const T *__restrict__ xb, const float *Lq,
                                            const float *Ls, Bcast<T> *bt, int o0, int rs, int lane, Rec r0, Rec r1)
{
    T x0 = xb[r0.col];
#pragma unroll
    for (int u = 0; u < V; u++) {
        Rec r2 = r1;
        if (u + 2 < V) r2 = rec_load(cp, cp.col, rb[u + 2 < V ? u + 2 : 0], re[u + 2 < V ? u + 2 : 0], lane);
        T x1 = x0;
        if (u + 1 < V) x1 = xb[r1.col];
        if (rb[u] < re[u]) {
            slab_a_chunk<NJ, T>(acc[u], r0, x0, rb[u] + lane < re[u], Lq, Ls, bt, o0, rs, lane);
            for (uint32_t base = rb[u] + kWave; base < re[u]; base += kWave) {
                const Rec r = rec_load(cp, cp.col, base, re[u], lane);
                slab_a_chunk<NJ, T>(acc[u], r, xb[r.col], base + lane < re[u], Lq, Ls, bt, o0, rs, lane);
            }
        }
        r0 = r1;
        r1 = r2;
        x0 = x1;
    }
}

// EC and ISO items (entry x RF) of the wavefront's voxels: lane 8u + j holds
// item j of voxel u (EC items RF-major, then ISO items RF-major), so the
// records and x values of all of them load together, and each voxel's rows
// are then fetched together instead of one dependent chain per entry.  A
// voxel with more than kItemLanes items takes the per-entry path.
constexpr int kItemLanes = 8;

struct Item {
    const float *row;   // the item's LUT row (sample 0)
    uint32_t col;       // the item's x column
    int m;              // items of this lane's voxel
};

template <int V>
__device__ __forceinline__ Item item_decode(const DevDict &d, const int (&vox)[V], size_t lut_rf, int lane)
{
    const int u = lane >> 3, j = lane & (kItemLanes - 1);
    int v = -1;
#pragma unroll
    for (int uu = 0; uu < V; uu++)
        if (u == uu) v = vox[uu];
    Item it{d.wmr, 0u, 0};
    if (v < 0) return it;
    const uint32_t eb = d.nEC > 0 ? d.ec.ptr[v] : 0u, ne = d.nEC > 0 ? d.ec.ptr[v + 1] - eb : 0u;
    const uint32_t ib = d.nISO > 0 ? d.iso.ptr[v] : 0u, ni = d.nISO > 0 ? d.iso.ptr[v + 1] - ib : 0u;
    const uint32_t nec = (uint32_t)d.nEC * ne;
    it.m = (int)(nec + (uint32_t)d.nISO * ni);
    if (j >= it.m) return it;
    if ((uint32_t)j < nec) {                      // EC RF i, peak k     (setup_operator.py:111-165)
        const uint32_t i = (uint32_t)j / ne, k = eb + ((uint32_t)j - i * ne);
        it.row = d.wmh + (size_t)i * lut_rf + (size_t)d.ec.o[k] * d.nSp;
        it.col = (uint32_t)d.nIC * (uint32_t)d.nF + i * (uint32_t)d.nE + d.ec.col[k];
    } else {                                      // ISO RF i, entry k    (setup_operator.py:167-218)
        const uint32_t t = (uint32_t)j - nec, i = t / ni, k = ib + (t - i * ni);
        it.row = d.isol + (size_t)i * d.nSp;
        it.col = (uint32_t)d.nIC * (uint32_t)d.nF + (uint32_t)d.nEC * (uint32_t)d.nE + i * (uint32_t)d.nI +
                 d.iso.col[k];
    }
    return it;
}

__device__ __forceinline__ const float *readlane_ptr(const float *p, int l)
{
    const uint64_t b = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return reinterpret_cast<const float *>(((uint64_t)hi << 32) | lo);
}

// RES: the residual output stage (commit_hip_A_residual); a separate
// instantiation, so the plain product keeps its own register allocation.
template <int NJ, int V, int W, typename T, bool RES>
__device__ __forceinline__ void a_slab(const DevDict &d, const T *__restrict__ x, const T *__restrict__ xic,
                                       const AOut<T> &out, const SlabArgs &sa)
{
    __shared__ __align__(16) float L[SlabLds<W>::floats];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    Bcast<T> *bt = bcast_table<W, T>(L, w);
    const int32_t *vl = sa.vlist + ((size_t)blockIdx.x * W + w) * V;
    int vox[V];
#pragma unroll
    for (int u = 0; u < V; u++) vox[u] = __builtin_amdgcn_readfirstlane(vl[u]);
    slab_pad_zero<W>(L, sa, d.nSp);
    double acc[V][NJ];
#pragma unroll
    for (int u = 0; u < V; u++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[u][j] = 0.0;
    const size_t lut_rf = (size_t)d.ndirs * d.nSp;
    for (int i = 0; i < d.nIC; i++) {
        const T *xb = xic + (size_t)i * d.nF;
        for (int s = 0; s < sa.nsl; s++) {
            const int o0 = s * sa.S, rows = min(sa.S, d.ndirs - o0);
            uint32_t rb[V], re[V];
#pragma unroll
            for (int u = 0; u < V; u++) {
                const uint32_t *of = sa.ofs + (size_t)(vox[u] < 0 ? 0 : vox[u]) * (sa.nsl + 1) + s;
                rb[u] = vox[u] < 0 ? 0u : of[0];
                re[u] = vox[u] < 0 ? 0u : of[1];
            }
            // records of the first two voxels load while the slab is copied
            const Rec r0 = rec_load(d.ic, d.ic.col, rb[0], re[0], lane);
            const Rec r1 = rec_load(d.ic, d.ic.col, rb[V > 1 ? 1 : 0], re[V > 1 ? 1 : 0], lane);
            __syncthreads();                                         // previous slab consumed
            slab_load<W>(L, d.wmr + (size_t)i * lut_rf + (size_t)o0 * d.nSp, rows, sa.rs, d.nSp, w, lane);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();                                         // slab in LDS
            slab_a_pass<NJ, V, T>(acc, d.ic, rb, re, xb, L + 4 * lane, L + lane, bt, o0, sa.rs, lane, r0, r1);
        }
    }
    const bool tail = Lanes<NJ, kQuads<T>>::tail(lane, d.nS);
    // EC and ISO: coefficient x of every item of the wavefront's voxels at once
    const Item it = item_decode<V>(d, vox, lut_rf, lane);
    double cit = 0.0;
    if ((lane & (kItemLanes - 1)) < it.m) cit = (double)x[it.col];
#pragma unroll
    for (int u = 0; u < V; u++) {
        const int v = vox[u];
        if (v < 0) continue;
        const int m = __builtin_amdgcn_readlane(it.m, kItemLanes * u);
        if (m <= kItemLanes) {
            for (int j0 = 0; j0 < m; j0 += 4) {                  // up to 4 rows in flight
                float r[4][NJ];
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    if (j0 + jj < m)
                        fetch_row<NJ, kQuads<T>>(r[jj], readlane_ptr(it.row, kItemLanes * u + j0 + jj), lane, tail);
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    if (j0 + jj < m) fma_row<NJ>(acc[u], readlane_f64(cit, kItemLanes * u + j0 + jj), r[jj]);
            }
        } else {
            if (d.nEC > 0)
                a_comp<NJ>(acc[u], d.ec, d.ec.ptr[v], d.ec.ptr[v + 1], x + (size_t)d.nIC * d.nF, (size_t)d.nE,
                           d.nEC, d.wmh, lut_rf, d.nSp, lane, tail);
            if (d.nISO > 0)
                a_comp<NJ>(acc[u], d.iso, d.iso.ptr[v], d.iso.ptr[v + 1],
                           x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE, (size_t)d.nI, d.nISO, d.isol,
                           (size_t)d.nSp, 0, lane, tail);
        }
        if constexpr (!RES) a_store<NJ, T, false>(acc[u], d, v, out, lane);
    }
    if constexpr (RES) {
        // residual: stores after every voxel's items, so no row fetch above
        // waits on a store, and the next voxel's observations and weights are
        // requested before this voxel's stores
        T yo[NJ] = {}, wv[NJ] = {};
        if (vox[0] >= 0) res_load<NJ, T>(yo, wv, d, vox[0], out, lane);
#pragma unroll
        for (int u = 0; u < V; u++) {
            T yn[NJ] = {}, wn[NJ] = {};
            const int vn = vox[u + 1 < V ? u + 1 : 0];
            if (u + 1 < V && vn >= 0) res_load<NJ, T>(yn, wn, d, vn, out, lane);
            if (vox[u] >= 0) res_store<NJ, T>(acc[u], yo, wv, d, vox[u], out, lane);
#pragma unroll
            for (int j = 0; j < NJ; j++) yo[j] = yn[j], wv[j] = wn[j];
        }
    }
}

template <int NJ, int V, int W, typename T>
__global__ __launch_bounds__(W * kWave, 4) void k_A_slab(DevDict d, const T *__restrict__ x, const T *__restrict__ xic,
                                                     AOut<T> out, SlabArgs sa)
{
    a_slab<NJ, V, W, T, false>(d, x, xic, out, sa);
}

template <int NJ, int V, int W, typename T>
__global__ __launch_bounds__(W * kWave, 4) void k_A_slab_res(DevDict d, const T *__restrict__ x,
                                                         const T *__restrict__ xic, AOut<T> out, SlabArgs sa)
{
    a_slab<NJ, V, W, T, true>(d, x, xic, out, sa);
}

// This lane's part of LUT[o] . y[v]: FP64 FMAs for a double y row; for a
// float y row, FP32 FMAs over the lane's NJ samples, widened once (the 64
// partials are then added in FP64).
template <int NJ>
__device__ __forceinline__ double dot_row(const float (&raw)[NJ], const double (&yv)[NJ])
{
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NJ; j++) a = fma((double)raw[j], yv[j], a);
    return a;
}

template <int NJ>
__device__ __forceinline__ double dot_row(const float (&raw)[NJ], const float (&yv)[NJ])
{
    float a = raw[0] * yv[0];
#pragma unroll
    for (int j = 1; j < NJ; j++) a = fmaf(raw[j], yv[j], a);
    return (double)a;
}

// A^T y pass 1: up to 64 IC entries of one voxel and one slab, lane l holding
// record r of entry b+l, cnt of them valid; z[zpos] = l * dot.  A voxel's
// entries are sorted by orientation, so entries of equal orientation are
// adjacent and share one dot product LUT[o].y[v]: only the first entry of each
// run (a head) computes it.  The heads' row offsets go, in order, to slots
// 0..nu-1 of the wavefront's broadcast table (the other lanes fill the slots
// past nu with valid rows).  Groups of 8 heads: per-lane partial dots, then
// reduce8 (head t0+g ends in lanes 8g..8g+7; lane 8g puts it in the table's
// c[t0+g]).  Slots past nu in the last group read valid rows and are not
// kept.  Each entry then reads its run's dot from the table and stores
// l * dot.                                            (setup_operator.py:384-428)
template <int NJ, typename Y>
__device__ __forceinline__ void slab_at_chunk(const Y (&yv)[NJ], const Rec &r, int cnt, double *__restrict__ zi,
                                              const float *Lq, const float *Ls, Bcast<double> *bt, int o0, int rs,
                                              int lane)
{
    constexpr bool Q = kQuads<Y>;
    const uint32_t op = (uint32_t)__shfl((int)r.o, lane > 0 ? lane - 1 : 0);
    const bool head = lane < cnt && (lane == 0 || r.o != op);
    const unsigned long long m = __ballot(head);
    const int nu = __popcll(m);
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const int run = head ? below : below - 1;                 // this entry's head, counted in head order
    bt->off[head ? below : nu + (lane - below)] = ((int)r.o - o0) * rs;
    int4 oa = *reinterpret_cast<const int4 *>(&bt->off[0]);
    int4 ob = *reinterpret_cast<const int4 *>(&bt->off[4]);
    int t0 = 0;
    for (; t0 + 4 < nu; t0 += 8) {
        const int of[8] = {oa.x, oa.y, oa.z, oa.w, ob.x, ob.y, ob.z, ob.w};
        double p[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            float raw[NJ];
            lds_row<NJ, Q>(raw, Lq + of[q], Ls + of[q]);
            p[q] = dot_row<NJ>(raw, yv);
        }
        const int tn = t0 + 8 < kWave ? t0 + 8 : t0;      // next group's offsets load under the reduction
        oa = *reinterpret_cast<const int4 *>(&bt->off[tn]);
        ob = *reinterpret_cast<const int4 *>(&bt->off[tn + 4]);
        const double sum = reduce8(p, lane);
        const int te = t0 + (lane >> 3);
        if ((lane & 7) == 0 && te < nu) bt->c[te] = sum;
    }
    if (t0 < nu) {                           // a tail of at most 4 heads: reduce4 (head t0+j in row j)
        const int of[4] = {oa.x, oa.y, oa.z, oa.w};
        double p[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float raw[NJ];
            lds_row<NJ, Q>(raw, Lq + of[q], Ls + of[q]);
            p[q] = dot_row<NJ>(raw, yv);
        }
        const double sum = reduce4(p);
        const int te = t0 + (lane >> 4);
        if ((lane & 15) == 0 && te < nu) bt->c[te] = sum;
    }
    if (lane < cnt) zi[r.col] = (double)r.len * bt->c[run];
}

// EC / ISO entries [b, e) of one voxel, rows from global memory:
// out[col + i*xs] = dot                                 (setup_operator.py:444-563)
template <int NJ, typename T>
__device__ __forceinline__ void glob_at(const T (&yv)[NJ], const Comp &cp, uint32_t b, uint32_t e,
                                        T *__restrict__ out, size_t xs, int nrf, const float *__restrict__ lut,
                                        size_t lut_rf, int nSp, int lane, bool tail)
{
    for (int i = 0; i < nrf; i++) {
        const float *Li = lut + (size_t)i * lut_rf;
        for (uint32_t k = b; k < e; k++) {
            const uint32_t o = cp.o ? (uint32_t)cp.o[k] : 0u;
            float raw[NJ];
            fetch_row<NJ, kQuads<T>>(raw, Li + (size_t)o * nSp, lane, tail);
            const double sum = wave_sum(dot_row<NJ>(raw, yv));
            if (lane == 0) out[(size_t)cp.col[k] + (size_t)i * xs] = (T)sum;
        }
    }
}

// COMB: the y rows are formed from two residuals, y = a r1 + b r2 in FP64
// (as commit_hip_residual_combine forms them, rounded to the vectors' type),
// and each voxel's sum of the unrounded values' squares goes to rpart[v].
template <typename T>
struct CombIn {
    const T *r2;
    double a, b;
    double *rpart;
};

template <int NJ, int V, int W, typename T, bool COMB>
__device__ __forceinline__ void at_slab(const DevDict &d, const T *__restrict__ y, const T *__restrict__ sqrtw,
                                        double *__restrict__ z, T *__restrict__ x, const CombIn<T> &cb,
                                        const SlabArgs &sa)
{
    __shared__ __align__(16) float L[SlabLds<W>::floats];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    Bcast<double> *bt = bcast_table<W, double>(L, w);
    slab_pad_zero<W>(L, sa, d.nSp);
    const int32_t *vl = sa.vlist + ((size_t)blockIdx.x * W + w) * V;
    int vox[V];
#pragma unroll
    for (int u = 0; u < V; u++) vox[u] = __builtin_amdgcn_readfirstlane(vl[u]);
    // y rows of the wavefront's voxels (times sqrtw), in the vectors' type:
    // FP64 dot products for double, FP32 lane partials for float (solvers.py:267)
    T yv[V][NJ];
#pragma unroll
    for (int u = 0; u < V; u++) {
        const size_t row = (size_t)(vox[u] < 0 ? 0 : vox[u]) * d.nS;
        T raw[NJ], wv[NJ], r2v[NJ];
        if constexpr (kQuads<T>) {
            row_load<NJ, T>(raw, y + row, d.nS, lane);
            if constexpr (COMB) row_load<NJ, T>(r2v, cb.r2 + row, d.nS, lane);
            if (sqrtw) row_load<NJ, T>(wv, sqrtw + row, d.nS, lane);
            else
#pragma unroll
                for (int j = 0; j < NJ; j++) wv[j] = (T)1;
        } else {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int s = Lanes<NJ, false>::sample(lane, j);
                raw[j] = stream_load(y + row + (s < d.nS ? s : 0));
                if constexpr (COMB) r2v[j] = stream_load(cb.r2 + row + (s < d.nS ? s : 0));
                wv[j] = sqrtw ? stream_load(sqrtw + row + (s < d.nS ? s : 0)) : (T)1;
            }
        }
        double ss = 0.0;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const bool in = vox[u] >= 0 && Lanes<NJ, kQuads<T>>::sample(lane, j) < d.nS;
            T t = raw[j];
            if constexpr (COMB) {
                const double v = cb.a * (double)raw[j] + cb.b * (double)r2v[j];
                t = (T)v;
                if (in) ss += v * v;
            }
            if (sqrtw) t *= wv[j];
            yv[u][j] = in ? t : (T)0;
        }
        if constexpr (COMB) {                    // the voxel's sum, in a fixed order
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) ss += __shfl_xor(ss, off);
            if (lane == 0 && vox[u] >= 0) cb.rpart[vox[u]] = ss;
        }
    }
    const size_t lut_rf = (size_t)d.ndirs * d.nSp;
    for (int i = 0; i < d.nIC; i++) {
        double *zi = z + (size_t)i * d.n;
        for (int s = 0; s < sa.nsl; s++) {
            const int o0 = s * sa.S, rows = min(sa.S, d.ndirs - o0);
            uint32_t rb[V], re[V];
#pragma unroll
            for (int u = 0; u < V; u++) {
                const uint32_t *of = sa.ofs + (size_t)(vox[u] < 0 ? 0 : vox[u]) * (sa.nsl + 1) + s;
                rb[u] = vox[u] < 0 ? 0u : of[0];
                re[u] = vox[u] < 0 ? 0u : of[1];
            }
            Rec r0 = rec_load(d.ic, d.zpos, rb[0], re[0], lane);       // loads while the slab is copied
            __syncthreads();
            slab_load<W>(L, d.wmr + (size_t)i * lut_rf + (size_t)o0 * d.nSp, rows, sa.rs, d.nSp, w, lane);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            const float *Lq = L + 4 * lane, *Ls = L + lane;
#pragma unroll
            for (int u = 0; u < V; u++) {
                Rec r1 = r0;                                            // next voxel's records in flight
                if (u + 1 < V) r1 = rec_load(d.ic, d.zpos, rb[u + 1 < V ? u + 1 : 0], re[u + 1 < V ? u + 1 : 0], lane);
                if (rb[u] < re[u]) {
                    slab_at_chunk<NJ, T>(yv[u], r0, (int)min((uint32_t)kWave, re[u] - rb[u]), zi, Lq, Ls, bt, o0,
                                         sa.rs, lane);
                    for (uint32_t base = rb[u] + kWave; base < re[u]; base += kWave)
                        slab_at_chunk<NJ, T>(yv[u], rec_load(d.ic, d.zpos, base, re[u], lane),
                                             (int)min((uint32_t)kWave, re[u] - base), zi, Lq, Ls, bt, o0, sa.rs, lane);
                }
                r0 = r1;
            }
        }
    }
    const bool tail = Lanes<NJ, kQuads<T>>::tail(lane, d.nS);
    // EC and ISO: a voxel's items (at most 8) reduced together by reduce8;
    // item j ends in lanes 8j..8j+7 and lane 8j stores it (setup_operator.py:444-563)
    const Item it = item_decode<V>(d, vox, lut_rf, lane);
#pragma unroll
    for (int u = 0; u < V; u++) {
        const int v = vox[u];
        if (v < 0) continue;
        const int m = __builtin_amdgcn_readlane(it.m, kItemLanes * u);
        if (m == 0) continue;
        if (m <= kItemLanes) {
            double p[8];
#pragma unroll
            for (int jj = 0; jj < 8; jj++) {
                p[jj] = 0.0;
                if (jj < m) {
                    float raw[NJ];
                    fetch_row<NJ, kQuads<T>>(raw, readlane_ptr(it.row, kItemLanes * u + jj), lane, tail);
                    p[jj] = dot_row<NJ>(raw, yv[u]);
                }
            }
            const double sum = reduce8(p, lane);
            const int jj = lane >> 3;
            const uint32_t col = (uint32_t)__shfl((int)it.col, kItemLanes * u + jj);
            if ((lane & 7) == 0 && jj < m) x[col] = (T)sum;
        } else {
            if (d.nEC > 0)
                glob_at<NJ>(yv[u], d.ec, d.ec.ptr[v], d.ec.ptr[v + 1], x + (size_t)d.nIC * d.nF, (size_t)d.nE,
                            d.nEC, d.wmh, lut_rf, d.nSp, lane, tail);
            if (d.nISO > 0)
                glob_at<NJ>(yv[u], d.iso, d.iso.ptr[v], d.iso.ptr[v + 1],
                            x + (size_t)d.nIC * d.nF + (size_t)d.nEC * d.nE, (size_t)d.nI, d.nISO, d.isol,
                            (size_t)d.nSp, 0, lane, tail);
        }
    }
}

template <int NJ, int V, int W, typename T>
__global__ __launch_bounds__(W * kWave, 4) void k_At_slab(DevDict d, const T *__restrict__ y, const T *__restrict__ sqrtw,
                                                      double *__restrict__ z, T *__restrict__ x, SlabArgs sa)
{
    at_slab<NJ, V, W, T, false>(d, y, sqrtw, z, x, CombIn<T>{nullptr, 0.0, 0.0, nullptr}, sa);
}

template <int NJ, int V, int W, typename T>
__global__ __launch_bounds__(W * kWave, 4) void k_At_slab_comb(DevDict d, const T *__restrict__ y,
                                                           const T *__restrict__ sqrtw, double *__restrict__ z,
                                                           T *__restrict__ x, CombIn<T> cb, SlabArgs sa)
{
    at_slab<NJ, V, W, T, true>(d, y, sqrtw, z, x, cb, sa);
}

// ofs[v*(nsl+1) + s] = first IC entry of voxel v with orientation >= s*S
// (entries of a voxel are sorted by orientation); ofs[v*(nsl+1) + nsl] = end.
__global__ void k_slab_ofs(const uint32_t *__restrict__ ptr, const uint16_t *__restrict__ o, int nVOX, int nsl,
                           int S, uint32_t *__restrict__ ofs)
{
    const size_t total = (size_t)nVOX * (nsl + 1);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t v = t / (nsl + 1);
        const int s = (int)(t - v * (nsl + 1));
        uint32_t lo = ptr[v], hi = ptr[v + 1];
        if (s == nsl) {
            ofs[t] = hi;
            continue;
        }
        const uint32_t key = (uint32_t)(s * S);
        while (lo < hi) {                           // lower bound of key in o[lo, hi)
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((uint32_t)o[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        ofs[t] = lo;
    }
}

// Work of voxel v in slab s, in half head-slots: the (voxel, slab) range
// [ofs[v][s], ofs[v][s+1]) holds nu distinct orientations (heads), applied in
// groups of 8 with a tail of at most 4 (as slab_at_chunk runs them), plus a
// fixed start-up per non-empty range and half a slot per entry.
__global__ void k_slab_cost(const uint32_t *__restrict__ ofs, const uint16_t *__restrict__ o, int nVOX, int nsl,
                            uint32_t *__restrict__ cost)
{
    const size_t total = (size_t)nVOX * nsl;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t v = t / nsl;
        const int s = (int)(t - v * nsl);
        const uint32_t b = ofs[v * (nsl + 1) + s], e = ofs[v * (nsl + 1) + s + 1];
        uint32_t nu = 0;
        for (uint32_t k = b; k < e; k++) nu += (k == b || o[k] != o[k - 1]) ? 1u : 0u;
        const uint32_t g = nu > 4 ? (nu - 4 + 7) / 8 : 0u;
        const uint32_t slots = 8 * g + (8 * g < nu ? 4u : 0u);
        cost[t] = 2 * slots + (e > b ? 16u : 0u) + (e - b);
    }
}

}  // namespace

namespace commit_internal {

namespace {

// One work-group's voxels dealt to its W wavefronts (V each).  A wavefront
// runs the slabs in order with a barrier after each, so a work-group takes
// about the sum over slabs of its slowest wavefront's work in that slab;
// c[i*nsl + s] is voxel i's work in slab s.  Voxels go longest first to the
// wavefront (with room) that raises that sum least, ties to the lighter one;
// then one pass of pairwise swaps between wavefronts keeps each swap that
// lowers it.  Returns the sum; slot[i] = wavefront * V + place.
uint64_t deal_group(const uint32_t *c, int nv, int nsl, int W, int V, int *slot)
{
    std::vector<int> order(nv), wave(nv);
    std::vector<uint64_t> tot(nv, 0), L((size_t)W * nsl, 0), mx(nsl);
    std::vector<int> used(W, 0);
    for (int i = 0; i < nv; i++)
        for (int s = 0; s < nsl; s++) tot[i] += c[(size_t)i * nsl + s];
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tot[a] > tot[b]; });
    for (int i : order) {
        const uint32_t *ci = c + (size_t)i * nsl;
        for (int s = 0; s < nsl; s++) {
            mx[s] = 0;
            for (int w = 0; w < W; w++) mx[s] = std::max(mx[s], L[(size_t)w * nsl + s]);
        }
        int best = -1;
        uint64_t best_inc = 0, best_row = 0;
        for (int w = 0; w < W; w++) {
            if (used[w] >= V) continue;
            uint64_t inc = 0, row = 0;
            for (int s = 0; s < nsl; s++) {
                const uint64_t nl = L[(size_t)w * nsl + s] + ci[s];
                inc += nl > mx[s] ? nl - mx[s] : 0;
                row += nl;
            }
            if (best < 0 || inc < best_inc || (inc == best_inc && row < best_row)) {
                best = w;
                best_inc = inc;
                best_row = row;
            }
        }
        wave[i] = best;
        used[best]++;
        for (int s = 0; s < nsl; s++) L[(size_t)best * nsl + s] += ci[s];
    }
    // per slab, the three heaviest wavefronts: the largest load outside any
    // two wavefronts is among them
    std::vector<int> top((size_t)nsl * 3);
    auto rank = [&]() {
        for (int s = 0; s < nsl; s++) {
            int *t = &top[(size_t)s * 3];
            t[0] = t[1] = t[2] = -1;
            for (int w = 0; w < W; w++) {
                const uint64_t v = L[(size_t)w * nsl + s];
                if (t[0] < 0 || v > L[(size_t)t[0] * nsl + s]) { t[2] = t[1]; t[1] = t[0]; t[0] = w; }
                else if (t[1] < 0 || v > L[(size_t)t[1] * nsl + s]) { t[2] = t[1]; t[1] = w; }
                else if (t[2] < 0 || v > L[(size_t)t[2] * nsl + s]) t[2] = w;
            }
        }
    };
    auto objective = [&]() {
        uint64_t m = 0;
        for (int s = 0; s < nsl; s++) {
            uint64_t v = 0;
            for (int w = 0; w < W; w++) v = std::max(v, L[(size_t)w * nsl + s]);
            m += v;
        }
        return m;
    };
    rank();
    uint64_t cur = objective();
    for (int i = 0; i < nv; i++)
        for (int j = i + 1; j < nv; j++) {
            const int wi = wave[i], wj = wave[j];
            if (wi == wj) continue;
            const uint32_t *ci = c + (size_t)i * nsl, *cj = c + (size_t)j * nsl;
            uint64_t val = 0;
            for (int s = 0; s < nsl && val < cur; s++) {
                uint64_t other = 0;
                for (int k = 0; k < 3; k++) {
                    const int w = top[(size_t)s * 3 + k];
                    if (w >= 0 && w != wi && w != wj) {
                        other = L[(size_t)w * nsl + s];
                        break;
                    }
                }
                const uint64_t li = L[(size_t)wi * nsl + s] - ci[s] + cj[s];
                const uint64_t lj = L[(size_t)wj * nsl + s] - cj[s] + ci[s];
                val += std::max(other, std::max(li, lj));
            }
            if (val >= cur) continue;
            for (int s = 0; s < nsl; s++) {
                L[(size_t)wi * nsl + s] += (uint64_t)cj[s] - ci[s];
                L[(size_t)wj * nsl + s] += (uint64_t)ci[s] - cj[s];
            }
            std::swap(wave[i], wave[j]);
            cur = val;
            rank();
        }
    std::fill(used.begin(), used.end(), 0);
    for (int i = 0; i < nv; i++) slot[i] = wave[i] * V + used[wave[i]]++;
    return cur;
}

// Voxels dealt to the W wavefronts of each work-group (W*V consecutive voxels
// per group) by deal_group; work-groups numbered heaviest first.  cost:
// [nVOX][nsl] per-slab work (k_slab_cost).
int build_vlist(commit_hip_plan *p, const std::vector<uint32_t> &cost, int nsl, int W, int V, int32_t **out,
                unsigned *blocks)
{
    const int nVOX = p->d.nVOX, per = W * V;
    const size_t nb = ((size_t)nVOX + per - 1) / per;
    std::vector<int32_t> vl(nb * per, -1);
    std::vector<uint64_t> gcost(nb);
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t]() {
            std::vector<int> slot(per);
            for (size_t b = t; b < nb; b += nth) {
                const int v0 = (int)(b * per), nv = std::min(per, nVOX - v0);
                gcost[b] = deal_group(cost.data() + (size_t)v0 * nsl, nv, nsl, W, V, slot.data());
                for (int i = 0; i < nv; i++) vl[b * per + (size_t)slot[i]] = v0 + i;
            }
        });
    for (auto &th : pool) th.join();
    // work-groups are dispatched in index order: the heaviest first, so that
    // the last ones to start are short and the launch ends evenly
    std::vector<uint32_t> bo(nb);
    std::iota(bo.begin(), bo.end(), 0u);
    std::stable_sort(bo.begin(), bo.end(), [&](uint32_t a, uint32_t c) { return gcost[a] > gcost[c]; });
    std::vector<int32_t> sorted(vl.size());
    for (size_t b = 0; b < nb; b++)
        std::copy(vl.begin() + (size_t)bo[b] * per, vl.begin() + (size_t)(bo[b] + 1) * per, sorted.begin() + b * per);
    vl.swap(sorted);
    void *mem = nullptr;
    const int rc = dev_alloc(p, vl.size() * sizeof(int32_t), &mem);
    if (rc) return rc;
    *out = static_cast<int32_t *>(mem);
    HIP_TRY(hipMemcpy(mem, vl.data(), vl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    *blocks = (unsigned)nb;
    return 0;
}

}  // namespace

int slab_configure(commit_hip_plan *p, int mode)
{
    slab_release(p);
    if (mode == 0) return 0;
    const DevDict &d = p->d;
    if (mode != 1 && mode != 2) return fail(COMMIT_HIP_EINVAL, "slab mode %d is not 0, 1 or 2", mode);
    p->slab = mode;                                 // with no IC entries the launches fall back
    if (d.n == 0 || d.nIC == 0 || d.nVOX == 0 || d.nSp / kWave > kSlabMaxNJ) return 0;
    const int W = mode == 1 ? 8 : 16;
    const int ldsf = (mode == 1 ? SlabLds<8>::floats : SlabLds<16>::floats);
    const int rs = slab_row_stride(d);                // rows packed at nS rounded up to 4 floats
    const int Smax = std::min(d.ndirs, (ldsf - W * kBcastFloats - (d.nSp - rs)) / rs);
    const int nsl = (d.ndirs + Smax - 1) / Smax;
    const int S = (d.ndirs + nsl - 1) / nsl;           // equal slabs: no short last step
    int rc;
    const size_t nofs = (size_t)d.nVOX * (nsl + 1);
    void *mem = nullptr;
    if ((rc = dev_alloc(p, nofs * sizeof(uint32_t), &mem))) return rc;
    p->slab_ofs = static_cast<uint32_t *>(mem);
    const unsigned g = (unsigned)std::min<size_t>((nofs + 255) / 256, 65536);
    hipLaunchKernelGGL(k_slab_ofs, dim3(g), dim3(256), 0, 0, d.ic.ptr, d.ic.o, d.nVOX, nsl, S, p->slab_ofs);
    HIP_TRY(hipGetLastError());
    // per-slab work of every voxel, for the dealing
    const size_t ncost = (size_t)d.nVOX * nsl;
    uint32_t *dcost = nullptr;
    HIP_TRY(hipMalloc(&dcost, ncost * sizeof(uint32_t)));
    const unsigned gc = (unsigned)std::min<size_t>((ncost + 255) / 256, 65536);
    hipLaunchKernelGGL(k_slab_cost, dim3(gc), dim3(256), 0, 0, p->slab_ofs, d.ic.o, d.nVOX, nsl, dcost);
    std::vector<uint32_t> cost(ncost);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(cost.data(), dcost, ncost * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(dcost);
    HIP_TRY(e);
    if ((rc = build_vlist(p, cost, nsl, W, kSlabVA, &p->slab_vlist, &p->slab_blocks))) return rc;
    if ((rc = build_vlist(p, cost, nsl, W, kSlabVAt, &p->slab_vlist_at, &p->slab_blocks_at))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    p->slab_S = S;
    p->slab_nsl = nsl;
    return 0;
}

void slab_release(commit_hip_plan *p)
{
    for (void *q : {static_cast<void *>(p->slab_vlist), static_cast<void *>(p->slab_vlist_at),
                    static_cast<void *>(p->slab_ofs)}) {
        if (!q) continue;
        for (size_t k = 0; k < p->allocs.size(); k++)
            if (p->allocs[k] == q) {
                p->allocs.erase(p->allocs.begin() + (long)k);
                break;
            }
        (void)hipFree(q);
    }
    p->slab_vlist = nullptr;
    p->slab_vlist_at = nullptr;
    p->slab_ofs = nullptr;
    p->slab = 0;
    p->slab_blocks = p->slab_blocks_at = 0;
}

#define SLAB_SWITCH(KERNEL, kSlabV, VLIST, BLOCKS, ...)                                                           \
    do {                                                                                                          \
        const SlabArgs sa{p->slab_ofs, VLIST, p->slab_S, p->slab_nsl, slab_row_stride(d)};                                            \
        const dim3 grid(BLOCKS);                                                                                  \
        const int NJv = d.nSp / kWave;                                                                            \
        if (p->slab == 1) {                                                                                       \
            switch (NJv) {                                                                                        \
                case 1: hipLaunchKernelGGL((KERNEL<1, kSlabV, 8, T>), grid, dim3(8 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 2: hipLaunchKernelGGL((KERNEL<2, kSlabV, 8, T>), grid, dim3(8 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 3: hipLaunchKernelGGL((KERNEL<3, kSlabV, 8, T>), grid, dim3(8 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 4: hipLaunchKernelGGL((KERNEL<4, kSlabV, 8, T>), grid, dim3(8 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 5: hipLaunchKernelGGL((KERNEL<5, kSlabV, 8, T>), grid, dim3(8 * kWave), 0, st, __VA_ARGS__, sa); break; \
                default: return COMMIT_HIP_EUNSUPPORTED;                                                          \
            }                                                                                                     \
        } else {                                                                                                  \
            switch (NJv) {                                                                                        \
                case 1: hipLaunchKernelGGL((KERNEL<1, kSlabV, 16, T>), grid, dim3(16 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 2: hipLaunchKernelGGL((KERNEL<2, kSlabV, 16, T>), grid, dim3(16 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 3: hipLaunchKernelGGL((KERNEL<3, kSlabV, 16, T>), grid, dim3(16 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 4: hipLaunchKernelGGL((KERNEL<4, kSlabV, 16, T>), grid, dim3(16 * kWave), 0, st, __VA_ARGS__, sa); break; \
                case 5: hipLaunchKernelGGL((KERNEL<5, kSlabV, 16, T>), grid, dim3(16 * kWave), 0, st, __VA_ARGS__, sa); break; \
                default: return COMMIT_HIP_EUNSUPPORTED;                                                          \
            }                                                                                                     \
        }                                                                                                         \
        HIP_TRY(hipGetLastError());                                                                               \
    } while (0)

template <typename T>
int launch_A_slab(commit_hip_plan *p, const T *x, const T *xic, const AOut<T> &out, hipStream_t st)
{
    const DevDict &d = p->d;
    if (!p->slab || !p->slab_vlist || d.nSp / kWave > kSlabMaxNJ) return COMMIT_HIP_EUNSUPPORTED;
    if (out.rpart || out.yobs || out.sqrtw)
        SLAB_SWITCH(k_A_slab_res, kSlabVA, p->slab_vlist, p->slab_blocks, d, x, xic, out);
    else
        SLAB_SWITCH(k_A_slab, kSlabVA, p->slab_vlist, p->slab_blocks, d, x, xic, out);
    return 0;
}

template <typename T>
int launch_At_slab(commit_hip_plan *p, const T *y, const T *sqrtw, T *x, hipStream_t st)
{
    const DevDict &d = p->d;
    if (!p->slab || !p->slab_vlist_at || d.nSp / kWave > kSlabMaxNJ) return COMMIT_HIP_EUNSUPPORTED;
    SLAB_SWITCH(k_At_slab, kSlabVAt, p->slab_vlist_at, p->slab_blocks_at, d, y, sqrtw, p->zbuf, x);
    return 0;
}

template <typename T>
int launch_At_slab_comb(commit_hip_plan *p, const T *r1, const T *r2, double a, double b, const T *sqrtw, T *x,
                        hipStream_t st)
{
    const DevDict &d = p->d;
    if (!p->slab || !p->slab_vlist_at || d.nSp / kWave > kSlabMaxNJ) return COMMIT_HIP_EUNSUPPORTED;
    const CombIn<T> cb{r2, a, b, p->rpart};
    SLAB_SWITCH(k_At_slab_comb, kSlabVAt, p->slab_vlist_at, p->slab_blocks_at, d, r1, sqrtw, p->zbuf, x, cb);
    return 0;
}

#undef SLAB_SWITCH

template int launch_A_slab<double>(commit_hip_plan *, const double *, const double *, const AOut<double> &,
                                   hipStream_t);
template int launch_A_slab<float>(commit_hip_plan *, const float *, const float *, const AOut<float> &, hipStream_t);
template int launch_At_slab<double>(commit_hip_plan *, const double *, const double *, double *, hipStream_t);
template int launch_At_slab<float>(commit_hip_plan *, const float *, const float *, float *, hipStream_t);
template int launch_At_slab_comb<double>(commit_hip_plan *, const double *, const double *, double, double,
                                         const double *, double *, hipStream_t);
template int launch_At_slab_comb<float>(commit_hip_plan *, const float *, const float *, double, double,
                                        const float *, float *, hipStream_t);

}  // namespace commit_internal
