// Device helpers shared by the per-voxel kernels (commit_hip.hip) and the
// LDS-slab kernels (slab.hip): LUT row fetches, the A x entry pipeline, the
// output stage of A and lane reductions built from VALU cross-lane moves.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "plan.h"

namespace {

using commit_internal::AOut;
using commit_internal::Comp;
using commit_internal::DevDict;

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
__device__ __forceinline__ T stream_load(const T *p)
{
    return __builtin_nontemporal_load(p);    // read once per product: keep L2 for LUT rows and x
}

// Lane <-> sample layout of the NJ = nSp / 64 samples a lane holds for one
// voxel (a LUT row, an accumulator row, a y row).  Q false: sample 64 j + l
// for j < NJ.  Q true (the float32-vector kernels): the first 4 (NJ / 4) are
// quads, quad q of lane l being samples 256 q + 4 l .. 256 q + 4 l + 3, read
// as one 16-byte access (the LDS serves 16-byte reads at twice the bandwidth
// of 4- and 8-byte ones); the remaining NJ % 4 are single samples
// 256 (NJ / 4) + 64 r + l.  The FP64 kernels keep Q false: with their FP64
// y rows and accumulators in registers they measured faster with the strided
// reads.  Padded LUT samples (>= nS) are zero.
template <int NJ, bool Q>
struct Lanes {
    static constexpr int NQ = Q ? NJ / 4 : 0, NR = NJ - 4 * NQ;
    static __device__ __forceinline__ int sample(int lane, int j)
    {
        return j < 4 * NQ ? 256 * (j / 4) + 4 * lane + (j % 4) : 256 * NQ + 64 * (j - 4 * NQ) + lane;
    }
    // the lane's last single sample is < nS (a lane past nS skips that load;
    // quads are always read whole)
    static __device__ __forceinline__ bool tail(int lane, int nS)
    {
        return NR == 0 || 256 * NQ + 64 * (NR - 1) + lane < nS;
    }
};

// the layout of the kernels whose vectors are of type T
template <typename T>
constexpr bool kQuads = sizeof(T) == sizeof(float);

// One padded LUT row (row: its sample 0, 16-byte aligned) in the lane layout:
// a 16-byte load per quad, a 4-byte load per single sample; without tail the
// lane's last single sample is past nS and is not loaded (it is zero).
template <int NJ, bool Q>
__device__ __forceinline__ void fetch_row(float (&raw)[NJ], const float *__restrict__ row, int lane, bool tail)
{
    constexpr int NQ = Lanes<NJ, Q>::NQ, NR = Lanes<NJ, Q>::NR;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const float4 v = *reinterpret_cast<const float4 *>(row + 256 * q + 4 * lane);
        raw[4 * q] = v.x, raw[4 * q + 1] = v.y, raw[4 * q + 2] = v.z, raw[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < NR; r++)
        raw[4 * NQ + r] = (r < NR - 1 || tail) ? row[256 * NQ + 64 * r + lane] : 0.0f;
}

// next set bit of m (consumed), or -1
__device__ __forceinline__ int take(unsigned long long &m)
{
    if (!m) return -1;
    const int t = __builtin_ctzll(m);
    m &= m - 1;
    return t;
}

// ---------------------------------------------------------------- y = A x ---
template <int NJ>
__device__ __forceinline__ void fma_row(double (&acc)[NJ], double c, const float (&raw)[NJ])
{
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = fma(c, (double)raw[j], acc[j]);
}

// Accumulate the entries [b, e) of one compartment into acc.
//   xb: x (or x*eval) block of RF 0, xs: stride between RF blocks;
//   lut: padded RF 0 table, lut_rf: stride between RF tables.
template <int NJ, typename T>
__device__ __forceinline__ void a_comp(double (&acc)[NJ], const Comp &cp, uint32_t b, uint32_t e,
                                       const T *__restrict__ xb, size_t xs, int nrf,
                                       const float *__restrict__ lut, size_t lut_rf, int nSp, int lane, bool tail)
{
    for (int i = 0; i < nrf; i++) {
        const float *L = lut + (size_t)i * lut_rf;
        for (uint32_t base = b; base < e; base += kWave) {
            const uint32_t k = base + lane;
            double c = 0.0;
            uint32_t o = 0;
            if (k < e) {
                if (cp.o) o = stream_load(cp.o + k);
                c = (double)xb[(size_t)stream_load(cp.col + k) + (size_t)i * xs];
                // IC coefficient l * x * eval                      (setup_operator.py:63-96)
                if (cp.len) c *= (double)stream_load(cp.len + k);
            }
            // zero coefficients are skipped                      (setup_operator.py:88)
            unsigned long long m = __ballot(c != 0.0);
            float ra[NJ], rb[NJ], rc[NJ];
            int ta = take(m), tb = -1, tc = -1;
            if (ta >= 0) fetch_row<NJ, kQuads<T>>(ra, L + (size_t)__builtin_amdgcn_readlane((int)o, ta) * nSp, lane, tail);
            if ((tb = take(m)) >= 0) fetch_row<NJ, kQuads<T>>(rb, L + (size_t)__builtin_amdgcn_readlane((int)o, tb) * nSp, lane, tail);
            // three rows rotate through the pipeline; entries keep their order
            while (true) {
                if (ta < 0) break;
                if ((tc = take(m)) >= 0) fetch_row<NJ, kQuads<T>>(rc, L + (size_t)__builtin_amdgcn_readlane((int)o, tc) * nSp, lane, tail);
                fma_row<NJ>(acc, readlane_f64(c, ta), ra);
                if (tb < 0) break;
                if ((ta = take(m)) >= 0) fetch_row<NJ, kQuads<T>>(ra, L + (size_t)__builtin_amdgcn_readlane((int)o, ta) * nSp, lane, tail);
                fma_row<NJ>(acc, readlane_f64(c, tb), rb);
                if (tc < 0) break;
                if ((tb = take(m)) >= 0) fetch_row<NJ, kQuads<T>>(rb, L + (size_t)__builtin_amdgcn_readlane((int)o, tb) * nSp, lane, tail);
                fma_row<NJ>(acc, readlane_f64(c, tc), rc);
            }
        }
    }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// A row of T (p: its sample 0) moved in the lane layout Lanes<NJ, kQuads<T>>.
// With the float layout a lane's quad is 4 consecutive samples: moved as 4
// separate 4-byte accesses, each instruction touches the same 1 KiB of lines
// at a 16-byte lane stride, and with the streaming hint the lines can leave
// L2 between the quad's four accesses, so the row is fetched (or written back)
// several times.  When the row is 16-byte aligned (nS a multiple of 4 and p
// aligned) a quad is one 16-byte access.
__device__ __forceinline__ bool quad_ok(const void *p, int nS)
{
    return (nS & 3) == 0 && ((uintptr_t)p & 15) == 0;
}

// the lane's NJ samples of the row; a sample past nS reads sample nS - 1
// (the caller masks it)
template <int NJ, typename T>
__device__ __forceinline__ void row_load(T (&v)[NJ], const T *__restrict__ p, int nS, int lane)
{
    constexpr bool Q = kQuads<T>;
    constexpr int NQ = Lanes<NJ, Q>::NQ;
    const bool vec = Q && quad_ok(p, nS);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int s0 = 256 * q + 4 * lane;
        if (vec && s0 + 3 < nS) {
            const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p + s0));
            v[4 * q] = (T)t.x, v[4 * q + 1] = (T)t.y, v[4 * q + 2] = (T)t.z, v[4 * q + 3] = (T)t.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) v[4 * q + e] = stream_load(p + (s0 + e < nS ? s0 + e : nS - 1));
        }
    }
#pragma unroll
    for (int j = 4 * NQ; j < NJ; j++) {
        const int s = Lanes<NJ, Q>::sample(lane, j);
        v[j] = stream_load(p + (s < nS ? s : nS - 1));
    }
}

// the lane's samples below nS of the row, stored
template <int NJ, typename T>
__device__ __forceinline__ void row_store(const T (&v)[NJ], T *__restrict__ p, int nS, int lane)
{
    constexpr bool Q = kQuads<T>;
    constexpr int NQ = Lanes<NJ, Q>::NQ;
    const bool vec = Q && quad_ok(p, nS);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int s0 = 256 * q + 4 * lane;
        if (vec && s0 + 3 < nS) {
            const f4v t = {(float)v[4 * q], (float)v[4 * q + 1], (float)v[4 * q + 2], (float)v[4 * q + 3]};
            __builtin_nontemporal_store(t, reinterpret_cast<f4v *>(p + s0));
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (s0 + e < nS) __builtin_nontemporal_store(v[4 * q + e], p + s0 + e);
        }
    }
#pragma unroll
    for (int j = 4 * NQ; j < NJ; j++) {
        const int s = Lanes<NJ, Q>::sample(lane, j);
        if (s < nS) __builtin_nontemporal_store(v[j], p + s);
    }
}

// The residual output stage split in two, so that a wavefront with several
// voxels can request the next voxel's observations and weights before it
// stores the current one: res_load, then res_store.
template <int NJ, typename T>
__device__ __forceinline__ void res_load(T (&yo)[NJ], T (&wv)[NJ], const DevDict &d, int v, const AOut<T> &out,
                                         int lane)
{
    const size_t row = (size_t)v * d.nS;
    if constexpr (kQuads<T>) {
        if (out.yobs) row_load<NJ, T>(yo, out.yobs + row, d.nS, lane);
        else
#pragma unroll
            for (int j = 0; j < NJ; j++) yo[j] = (T)0;
        if (out.sqrtw) row_load<NJ, T>(wv, out.sqrtw + row, d.nS, lane);
        else
#pragma unroll
            for (int j = 0; j < NJ; j++) wv[j] = (T)1;
    } else {
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int s = Lanes<NJ, false>::sample(lane, j);
            const size_t i = row + (size_t)(s < d.nS ? s : d.nS - 1);
            yo[j] = out.yobs ? stream_load(out.yobs + i) : (T)0;
            wv[j] = out.sqrtw ? stream_load(out.sqrtw + i) : (T)1;
        }
    }
}

template <int NJ, typename T>
__device__ __forceinline__ void res_store(const double (&acc)[NJ], const T (&yo)[NJ], const T (&wv)[NJ],
                                          const DevDict &d, int v, const AOut<T> &out, int lane)
{
    const size_t row = (size_t)v * d.nS;
    double sq = 0.0;
    if constexpr (kQuads<T>) {
        T rv[NJ];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int s = Lanes<NJ, true>::sample(lane, j);
            double r = acc[j];
            if (out.yobs) r -= (double)yo[j];
            if (out.sqrtw) r *= (double)wv[j];
            rv[j] = (T)r;
            if (s < d.nS) sq = fma(r, r, sq);
        }
        if (out.y) row_store<NJ, T>(rv, out.y + row, d.nS, lane);
    } else {
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int s = Lanes<NJ, false>::sample(lane, j);
            if (s < d.nS) {
                double r = acc[j];
                if (out.yobs) r -= (double)yo[j];
                if (out.sqrtw) r *= (double)wv[j];
                if (out.y) __builtin_nontemporal_store((T)r, out.y + row + s);
                sq = fma(r, r, sq);
            }
        }
    }
    if (out.rpart) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off);
        if (lane == 0) out.rpart[v] = sq;
    }
}

// Output stage of A for voxel v: y row (or the residual sqrtw .* (A x - yobs))
// and the voxel's partial of ||r||^2.                  (solvers.py:306-308)
// RES false: a plain product (out.yobs, out.sqrtw and out.rpart are null).
template <int NJ, typename T, bool RES = true>
__device__ __forceinline__ void a_store(const double (&acc)[NJ], const DevDict &d, int v, const AOut<T> &out, int lane)
{
    if constexpr (!RES) {
        if constexpr (kQuads<T>) {
            T yv[NJ];
#pragma unroll
            for (int j = 0; j < NJ; j++) yv[j] = (T)acc[j];
            row_store<NJ, T>(yv, out.y + (size_t)v * d.nS, d.nS, lane);
        } else {
            const size_t row = (size_t)v * d.nS;
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int s = Lanes<NJ, false>::sample(lane, j);
                if (s < d.nS) __builtin_nontemporal_store((T)acc[j], out.y + row + s);
            }
        }
        return;
    }
    // the observations and weights are all requested before the first store:
    // a load issued after a store would wait for that store as well
    T yo[NJ], wv[NJ];
    res_load<NJ, T>(yo, wv, d, v, out, lane);
    res_store<NJ, T>(acc, yo, wv, d, v, out, lane);
}

// ---- cross-lane moves of doubles on the VALU (no LDS traffic) --------------
// DPP controls: quad_perm [1,0,3,2] and [2,3,0,1], row_mirror (lane i <- 15-i
// inside each row of 16), row_half_mirror (lane i <- 7-i inside each 8).
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppRowMirror = 0x140;
constexpr int kDppHalfMirror = 0x141;

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v_permlane32_swap on a pair of doubles (a, b): lanes 32-63 of a trade places
// with lanes 0-31 of b; returns a' + b', i.e. lanes 0-31: a(l) + a(l+32),
// lanes 32-63: b(l-32) + b(l).
__device__ __forceinline__ double swap32_add(double a, double b)
{
    const long long ab = __double_as_longlong(a), bb = __double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)(ab & 0xffffffffll), (unsigned)(bb & 0xffffffffll),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ab >> 32), (unsigned)(bb >> 32), false, false);
    const double a2 = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    const double b2 = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
    return a2 + b2;
}

// v_permlane16_swap likewise on rows of 16 lanes: odd rows of a trade places
// with even rows of b; returns a' + b' (even rows: a(l) + a(l+16), odd rows:
// b(l-16) + b(l)).
__device__ __forceinline__ double swap16_add(double a, double b)
{
    const long long ab = __double_as_longlong(a), bb = __double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)(ab & 0xffffffffll), (unsigned)(bb & 0xffffffffll),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ab >> 32), (unsigned)(bb >> 32), false, false);
    const double a2 = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    const double b2 = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
    return a2 + b2;
}

// Sum of 8 per-lane values over the 64 lanes, transposed: on return every lane
// of the group of 8 lanes g = lane/8 holds the wave total of p[g].  Fixed
// order, so bitwise repeatable.  30-odd VALU operations for 8 sums.
__device__ __forceinline__ double reduce8(const double (&p)[8], int lane)
{
    double q[4], r[2];
#pragma unroll
    for (int t = 0; t < 4; t++) q[t] = swap32_add(p[t], p[t + 4]);   // lanes 0-31: p[t], 32-63: p[t+4]
#pragma unroll
    for (int t = 0; t < 2; t++) r[t] = swap16_add(q[t], q[t + 2]);   // + rows 1,3: q[t+2]
    const bool hi8 = (lane & 8) != 0;                                 // lanes 8-15 of each row: r[1]
    const double keep = hi8 ? r[1] : r[0], send = hi8 ? r[0] : r[1];
    double s = keep + dpp_f64<kDppRowMirror>(send);
    s += dpp_f64<kDppQuadXor1>(s);
    s += dpp_f64<kDppQuadXor2>(s);
    s += dpp_f64<kDppHalfMirror>(s);
    return s;
}

// Sum of 4 per-lane values over the 64 lanes, transposed: on return every
// lane of the row of 16 lanes j = lane/16 holds the wave total of p[j].
__device__ __forceinline__ double reduce4(const double (&p)[4])
{
    const double q0 = swap32_add(p[0], p[2]), q1 = swap32_add(p[1], p[3]);   // lanes 32-63: p[2], p[3]
    double s = swap16_add(q0, q1);                                          // rows 1, 3: q1
    s += dpp_f64<kDppQuadXor1>(s);
    s += dpp_f64<kDppQuadXor2>(s);
    s += dpp_f64<kDppHalfMirror>(s);
    s += dpp_f64<kDppRowMirror>(s);
    return s;
}

// Sum of one per-lane value over the 64 lanes, in every lane.
__device__ __forceinline__ double wave_sum(double s)
{
    s += dpp_f64<kDppQuadXor1>(s);
    s += dpp_f64<kDppQuadXor2>(s);
    s += dpp_f64<kDppHalfMirror>(s);
    s += dpp_f64<kDppRowMirror>(s);
    s = swap16_add(s, s);
    return swap32_add(s, s);
}

}  // namespace
