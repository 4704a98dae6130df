// Tractogram voxeliser on the GPU: streamlines -> the IC/EC/ISO dictionary
// that trk2dictionary() exports (trk2dictionary_c.cpp:111-869), for
// streamlines already read into memory.
//
// Pipeline (one stream, all device-resident):
//   (one replica: k_vox_bound sizes each streamline's region, and a single
//   k_vox_walk<true, .., BOUNDED> records and counts; a streamline that
//   outgrows its region sends the tractogram through the two walks below)
//   k_vox_walk<false>   one thread per streamline runs fiberForwardModel
//                       (:584-755) with its replicas, rayBoxIntersection
//                       (:814-869) and segmentForwardModel (:761-808), and
//                       counts the pieces it would record, with FiberLen and
//                       FiberLenTot (float sums in recording order, :804-807)
//   exclusive scan      piece offsets per streamline
//   k_vox_walk<true>    the same walk again, storing each piece as
//                       key = o * nvox + v (v = x + X(y + Y z), so ascending
//                       keys are the std::map<segKey> order (o, z, y, x), :38-41)
//                       and its value w * len, in recording order (32-bit keys
//                       when they fit, float values when w is 1: ic_segments)
//   segmented sort      stable radix sort of each streamline's pieces by key
//   k_vox_merge         equal keys summed into a float in recording order
//                       (the map's `+=`, :804), kept flag (:510-515)
//   segmented sort      the merged segments of each streamline by voxel
//                       (stable, so each voxel's segments stay in map order)
//   k_vox_norm          TRK_norm: per-voxel float sums, float sum of their
//                       squares in (z, y, x) order, float square root (:530-537),
//                       a wavefront per streamline
//   scans + k_vox_write kept streamlines renumbered (trk2dictionary.pyx:551-563)
//                       and their segments written in streamline order
//   radix sort + k_vox_tdi   compute_tdi (pyx:50-59): float sums per voxel in
//                       segment order, and which voxels hold an IC segment
//   k_vox_ec_*          ECSegments (:289-397), voxels in (z, y, x) order
//   k_vox_iso_*         ISOcompartments (:400-437)
//
// The arithmetic is the reference's, operation by operation, in IEEE double
// with its float roundings, and this file is compiled without contraction
// (no fused multiply-adds), so every output is bit-for-bit the host
// restatement's.  The sines and cosines of the blur quaternions depend only on
// the replica and are taken on the host.  The one transcendental left on the
// device, atan2, feeds round(degrees): a last-place difference from the host
// libm changes an orientation only for an angle within 1e-13 degrees of a
// half degree.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "commit_hip.h"
#include "common.h"

using commit_internal::fail;

namespace {

constexpr int kVoxBlock = 256;
constexpr int kMaxFibLen = 10000;      // MAX_FIB_LEN, trk2dictionary_c.cpp:16

struct VoxArgs {
    int X, Y, Z;
    float px, py, pz;
    double rpx, rpy, rpz;              // 1.0 / (double)pixdim, correctly rounded (floor_div)
    float sx, sy, sz;                  // shift in mm: fiber_shift * pixdim, float (:125-127)
    float min_seg, min_fib, max_fib;
    int intersect;
    int n_rep;
    uint64_t nvox;
    const float *mask;                 // C order [X][Y][Z] or null
    const uint64_t *mbits;             // mask != 0, one bit per voxel in the same order (with mask)
    const short *hash;                 // 181 * 181
    const double *rho, *sinh, *cosh, *wt;   // per replica
    const uint8_t *apply;              // per streamline or null
};

struct D3 {
    double x, y, z;
};

// Vector<double>::norm (Vector.h:44-47)
__device__ __forceinline__ float vnorm(D3 a) { return (float)sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Vector<double>::Normalize (Vector.h:50-57)
__device__ __forceinline__ D3 vnormalize(D3 a)
{
    float len = (float)sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (len == 0) return a;
    a.x = a.x / len;
    a.y = a.y / len;
    a.z = a.z / len;
    return a;
}

// floor(x / p) with x / p the correctly rounded IEEE quotient, from r =
// RN(1 / p).  q0 = RN(x r) is within 3.0001 u |x / p| of RN(x / p) (u =
// 2^-53: one rounding in r, one in the product, one in the quotient), so when
// q0 lies farther than 2^-50 |q0| from every integer the two have the same
// floor; otherwise (about one case in 10^13 for these magnitudes) the quotient
// is formed by IEEE division.  q0 - floor(q0) and ceil(q0) - q0 are exact
// whenever either is below 1/2.  Needs no overflow or underflow: positions
// and voxel sizes are millimetres.
__device__ __forceinline__ double floor_div(double x, double p, double r)
{
    const double q0 = x * r;
    const double f = floor(q0), eps = fabs(q0) * 0x1p-50;
    if (q0 - f > eps && ceil(q0) - q0 > eps) return f;
    return floor(x / p);
}

// round(x / M_PI * 180.0) the same way: m0 = RN(RN(x / pi) 180) is within
// about 5 u |m| of the IEEE chain's m, and round() changes value only at
// half-integers; near one (within 2^-48 |m0|) the chain itself is evaluated.
__device__ __forceinline__ double round_deg(double x)
{
    constexpr double kInvPi = 1.0 / M_PI;
    const double m0 = x * kInvPi * 180.0;
    const double fr = m0 - floor(m0), eps = fabs(m0) * 0x1p-48;
    if (fabs(fr - 0.5) > eps) return round(m0);
    return round(x / M_PI * 180.0);
}

// ptrHashTable[ox*181 + oy] for angles already rounded to degrees
__device__ __forceinline__ int hash_lookup(const short *hash, int ox, int oy)
{
    long long idx = (long long)ox * 181 + oy;
    if (idx < 0) idx = 0;              // the reference reads before its table here
    if (idx >= 181 * 181) idx = 181 * 181 - 1;
    return hash[idx];
}

// ptrHashTable[ox*181 + oy] with the angles rounded to degrees as
// segmentForwardModel (:799-802) and ECSegments (:371-374) round them
__device__ __forceinline__ int hash_index(const short *hash, double colat, double lon)
{
    return hash_lookup(hash, (int)round_deg(colat), (int)round_deg(lon));
}

// The two rounded angles of segmentForwardModel (:799-802) for a direction
// with d.y >= 0, from FP32 arithmetic: the components rounded to float move
// each angle by at most about 1e-7 rad, atan2f and the float square root add
// a few ulp (under 1e-6 rad), and the scaling to degrees under 2e-5 degrees,
// so an angle farther than 1e-3 degrees from a half degree rounds as the FP64
// one does.  Returns false, leaving the FP64 path to the caller, when either
// angle is nearer than that (about 0.4% of pieces).
__device__ __forceinline__ bool degrees_f32(D3 d, int &ox, int &oy)
{
    const float x = (float)d.x, y = (float)d.y, z = (float)d.z;
    const float kDeg = 57.29577951308232f;           // 180 / pi
    const float mx = atan2f(sqrtf(x * x + y * y), z) * kDeg, my = atan2f(y, x) * kDeg;
    const float fx = mx - floorf(mx), fy = my - floorf(my);
    if (!(fabsf(fx - 0.5f) > 1e-3f && fabsf(fy - 0.5f) > 1e-3f)) return false;
    ox = (int)floorf(mx + 0.5f);
    oy = (int)floorf(my + 0.5f);
    return true;
}

// cap of one sub-step's voxel-crossing loop (the reference loops unbounded)
__device__ __forceinline__ long long max_cross(double len, const VoxArgs &a)
{
    double m = a.px < a.py ? a.px : a.py;
    if (a.pz < m) m = a.pz;
    double c = 3.0 * len / m;
    if (!(c < 1e7)) c = 1e7;
    return 64 + 4 * (long long)c;
}

// rayBoxIntersection (trk2dictionary_c.cpp:814-869); inv = {1 / d.x, 1 / d.y,
// 1 / d.z} of the ray direction, which is fixed for a whole polyline step and
// is formed once per step by the caller
__device__ __forceinline__ bool raybox(D3 o, D3 inv, D3 vmin, D3 vmax, double &t)
{
    double tmin, tmax, tymin, tymax, tzmin, tzmax;
    if (inv.x >= 0) { tmin = (vmin.x - o.x) * inv.x; tmax = (vmax.x - o.x) * inv.x; }
    else            { tmin = (vmax.x - o.x) * inv.x; tmax = (vmin.x - o.x) * inv.x; }
    if (inv.y >= 0) { tymin = (vmin.y - o.y) * inv.y; tymax = (vmax.y - o.y) * inv.y; }
    else            { tymin = (vmax.y - o.y) * inv.y; tymax = (vmin.y - o.y) * inv.y; }
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    if (inv.z >= 0) { tzmin = (vmin.z - o.z) * inv.z; tzmax = (vmax.z - o.z) * inv.z; }
    else            { tzmin = (vmax.z - o.z) * inv.z; tzmax = (vmin.z - o.z) * inv.z; }
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    t = tmin;
    if (t <= 0) t = tmax;
    return true;
}

// Pieces a lane holds in LDS before it writes them out.  Every lane writes
// its own streamline's region, so one store instruction touches 64 cache
// lines; stored one piece per crossing iteration, those lines were evicted
// from L2 between consecutive pieces and written back many times over.
// Written in runs of up to kRecBuf consecutive pieces, each line is filled in
// one burst.
constexpr int kRecBuf = 16;

// One streamline's recorder: counts pieces (EMIT = false) or stores them,
// keys as KT (32 bits when every key fits) and values w * len as VT (float
// when w is 1, so that w * len is the float length itself, exactly).
// EMIT: the pieces are buffered in this lane's slots of its wavefront's LDS
// buffer (lk, lv: slot 0; slot e at e * kWave) and written out by flush().
template <bool EMIT, class KT, class VT>
struct Recorder {
    const VoxArgs &a;
    KT *keys;
    VT *vals;
    KT *lk;
    VT *lv;
    uint32_t limit;                    // the count pass's total: stores never pass it
    uint32_t n = 0;
    uint32_t nb = 0, nf = 0;           // pieces buffered, pieces written
    float flen = 0.0f, flentot = 0.0f;

    __device__ Recorder(const VoxArgs &a_, KT *k, VT *v, KT *lk_, VT *lv_, uint32_t lim)
        : a(a_), keys(k), vals(v), lk(lk_), lv(lv_), limit(lim) {}

    // every active lane writes its buffered pieces, in recording order, to
    // slots nf .. nf + nb - 1 of its region (all below limit: only pieces
    // with n < limit are buffered)
    __device__ void flush()
    {
        for (uint32_t e = 0; e < nb; e++) {
            keys[nf + e] = lk[e * kWave];
            vals[nf + e] = lv[e * kWave];
        }
        nf += nb;
        nb = 0;
    }

    // segmentForwardModel (trk2dictionary_c.cpp:761-808)
    __device__ void segment(D3 P1, D3 P2, int k, double w)
    {
        D3 dir;
        dir.y = P2.y - P1.y;
        if (dir.y >= 0) {
            dir.x = P2.x - P1.x;
            dir.z = P2.z - P1.z;
        } else {
            dir.x = P1.x - P2.x;
            dir.y = P1.y - P2.y;
            dir.z = P1.z - P2.z;
        }
        double len = vnorm(dir);
        if (w * len <= a.min_seg) return;
        dir = vnormalize(dir);
        int vx = (int)floor_div(0.5 * (P1.x + P2.x), a.px, a.rpx);
        int vy = (int)floor_div(0.5 * (P1.y + P2.y), a.py, a.rpy);
        int vz = (int)floor_div(0.5 * (P1.z + P2.z), a.pz, a.rpz);
        if (vx >= a.X || vx < 0 || vy >= a.Y || vy < 0 || vz >= a.Z || vz < 0) return;
        if (a.mask) {                  // ptrMASK[...] == 0, from a bit array that stays in L2
            const uint64_t mv = (uint64_t)vz + (uint64_t)a.Z * ((uint64_t)vy + (uint64_t)a.Y * vx);
            if (!((a.mbits[mv >> 6] >> (mv & 63)) & 1ull)) return;
        }
        if (EMIT && n < limit) {
            int ox, oy;
            if (!degrees_f32(dir, ox, oy)) {
                double lon = atan2(dir.y, dir.x);
                double colat = atan2(sqrt(dir.x * dir.x + dir.y * dir.y), dir.z);
                ox = (int)round_deg(colat);
                oy = (int)round_deg(lon);
            }
            unsigned short o = (unsigned short)hash_lookup(a.hash, ox, oy);
            uint64_t v = (uint64_t)vx + (uint64_t)a.X * ((uint64_t)vy + (uint64_t)a.Y * vz);
            lk[nb * kWave] = (KT)((uint64_t)o * a.nvox + v);
            lv[nb * kWave] = (VT)(w * len);
            nb++;
        }
        n++;
        flentot = (float)((double)flentot + w * len);
        if (k == 0) flen = (float)((double)flen + len);
        // a full buffer in any active lane: the active lanes write theirs
        if (EMIT && __ballot(nb == (uint32_t)kRecBuf) != 0) flush();
    }
};

// fiberForwardModel (trk2dictionary_c.cpp:584-755) for streamline f.  MULTI:
// more than one replica, their positions kept in Pbuf [nF][n_rep].
//   EMIT = false             counts the pieces (cnt, FiberLen, FiberLenTot)
//   EMIT, !BOUNDED           stores exactly the counted pieces
//   EMIT, BOUNDED            stores the pieces into a region of at most
//                            off[f+1] - off[f] slots and counts them too; a
//                            streamline with more pieces sets status[2]
// status[0] = 1: a crossing loop hit its cap, or the recording pass found
// other counts than the counting pass.
template <bool EMIT, bool MULTI, class KT, class VT, bool BOUNDED = false>
__global__ __launch_bounds__(kVoxBlock) void k_vox_walk(VoxArgs a, const float *__restrict__ pts,
                                                        const int64_t *__restrict__ fptr, int nF,
                                                        D3 *__restrict__ Pbuf, const uint64_t *__restrict__ off,
                                                        KT *__restrict__ keys, VT *__restrict__ vals,
                                                        uint32_t *__restrict__ cnt, float *__restrict__ flen,
                                                        float *__restrict__ flentot, uint32_t *__restrict__ status,
                                                        const uint32_t *__restrict__ order)
{
    // EMIT: per wavefront, kRecBuf slots of 64 lanes for keys and values
    constexpr int kBuf = EMIT ? kVoxBlock * kRecBuf : 1;
    __shared__ KT lk_s[kBuf];
    __shared__ VT lv_s[kBuf];
    const int t = blockIdx.x * kVoxBlock + threadIdx.x;
    if (t >= nF) return;
    const int f = (int)order[t];
    const int lslot = EMIT ? (int)(threadIdx.x >> 6) * kRecBuf * kWave + (int)(threadIdx.x & (kWave - 1)) : 0;
    Recorder<EMIT, KT, VT> rec(a, EMIT ? keys + off[f] : nullptr, EMIT ? vals + off[f] : nullptr, lk_s + lslot,
                               lv_s + lslot, EMIT ? (uint32_t)(off[f + 1] - off[f]) : 0u);
    const int64_t p0 = fptr[f];
    const int pts_n = (int)(fptr[f + 1] - p0);
    const float *fib = pts + 3 * p0;
    const bool blur = a.apply ? (a.apply[f] != 0) : true;
    D3 P0 = {0, 0, 0};
    D3 *Pm = MULTI ? Pbuf + (size_t)f * a.n_rep : nullptr;
    bool capped = false;

    if (pts_n > 2) {
        D3 S1 = {(double)(fib[0] + a.sx), (double)(fib[1] + a.sy), (double)(fib[2] + a.sz)};
        D3 dir2 = {(double)(fib[3] + a.sx) - S1.x, (double)(fib[4] + a.sy) - S1.y, (double)(fib[5] + a.sz) - S1.z};
        dir2 = vnormalize(dir2);
        D3 n = {dir2.y - dir2.z, dir2.z - dir2.x, dir2.x - dir2.y};
        n = vnormalize(n);
        for (int k = 0; k < a.n_rep; k++) {
            if (!blur && k > 0) continue;
            double R = a.rho[k];
            double w = a.sinh[k];
            D3 q = {dir2.x * w, dir2.y * w, dir2.z * w};
            w = a.cosh[k];
            D3 qxn = {2.0 * (q.y * n.z - q.z * n.y), 2.0 * (q.z * n.x - q.x * n.z), 2.0 * (q.x * n.y - q.y * n.x)};
            D3 qq = {q.y * qxn.z - q.z * qxn.y, q.z * qxn.x - q.x * qxn.z, q.x * qxn.y - q.y * qxn.x};
            D3 nr = {n.x + w * qxn.x + qq.x, n.y + w * qxn.y + qq.y, n.z + w * qxn.z + qq.z};
            nr = vnormalize(nr);
            D3 s = {S1.x + R * nr.x, S1.y + R * nr.y, S1.z + R * nr.z};
            if (MULTI) Pm[k] = s;
            else P0 = s;
        }
        if constexpr (MULTI) {
            for (int i = 1; i < pts_n; i++) {
                D3 S2 = {(double)(fib[3 * i] + a.sx), (double)(fib[3 * i + 1] + a.sy), (double)(fib[3 * i + 2] + a.sz)};
                D3 dir1 = {S2.x - (double)(fib[3 * i - 3] + a.sx), S2.y - (double)(fib[3 * i - 2] + a.sy),
                           S2.z - (double)(fib[3 * i - 1] + a.sz)};
                dir1 = vnormalize(dir1);
                if (i == pts_n - 1) {
                    dir2 = dir1;
                } else {
                    dir2.x = (double)(fib[3 * i + 3] + a.sx) - S2.x;
                    dir2.y = (double)(fib[3 * i + 4] + a.sy) - S2.y;
                    dir2.z = (double)(fib[3 * i + 5] + a.sz) - S2.z;
                    dir2 = vnormalize(dir2);
                }
                D3 nn = {0.5 * (dir1.x + dir2.x), 0.5 * (dir1.y + dir2.y), 0.5 * (dir1.z + dir2.z)};
                double dot = dir1.x * nn.x + dir1.y * nn.y + dir1.z * nn.z;
                nn.x /= dot;
                nn.y /= dot;
                nn.z /= dot;
                D3 inv = {0, 0, 0};             // 1 / dir1, formed at the step's first crossing
                bool have_inv = false;
                for (int k = 0; k < a.n_rep; k++) {
                    if (!blur && k > 0) continue;
                    if (a.wt[k] < 1e-3) continue;
                    const double wk = a.wt[k];
                    D3 Po = MULTI ? Pm[k] : P0;
                    double len = (S2.x - Po.x) * nn.x + (S2.y - Po.y) * nn.y + (S2.z - Po.z) * nn.z;
                    if (!(len > 0)) continue;
                    D3 Pn = {Po.x + dir1.x * len, Po.y + dir1.y * len, Po.z + dir1.z * len};
                    if (MULTI) Pm[k] = Pn;
                    else P0 = Pn;
                    if (!a.intersect) {
                        rec.segment(Po, Pn, k, wk);
                        continue;
                    }
                    D3 S1m = Po;
                    const D3 S2m = Pn;
                    long long cap = -1;
                    for (long long it = 0;; it++) {
                        double dx = S2m.x - S1m.x, dy = S2m.y - S1m.y, dz = S2m.z - S1m.z;
                        double sl = sqrt(dx * dx + dy * dy + dz * dz);
                        if (sl <= a.min_seg) break;
                        if (cap < 0) cap = max_cross(sl, a);
                        if (it >= cap) {
                            capped = true;
                            break;
                        }
                        if (floor_div(S1m.x, a.px, a.rpx) == floor_div(S2m.x, a.px, a.rpx) &&
                            floor_div(S1m.y, a.py, a.rpy) == floor_div(S2m.y, a.py, a.rpy) &&
                            floor_div(S1m.z, a.pz, a.rpz) == floor_div(S2m.z, a.pz, a.rpz)) {
                            rec.segment(S1m, S2m, k, wk);
                            break;
                        }
                        D3 vmin, vmax;
                        vmin.x = floor_div(S1m.x + 1e-6 * dir1.x, a.px, a.rpx) * a.px;
                        vmin.y = floor_div(S1m.y + 1e-6 * dir1.y, a.py, a.rpy) * a.py;
                        vmin.z = floor_div(S1m.z + 1e-6 * dir1.z, a.pz, a.rpz) * a.pz;
                        vmax.x = vmin.x + a.px;
                        vmax.y = vmin.y + a.py;
                        vmax.z = vmin.z + a.pz;
                        if (!have_inv) {
                            inv = {1.0 / dir1.x, 1.0 / dir1.y, 1.0 / dir1.z};
                            have_inv = true;
                        }
                        double t;
                        if (raybox(S1m, inv, vmin, vmax, t) && t > 0 && t < sl) {
                            D3 Pi = {S1m.x + t * dir1.x, S1m.y + t * dir1.y, S1m.z + t * dir1.z};
                            rec.segment(S1m, Pi, k, wk);
                            S1m = Pi;
                        } else {
                            rec.segment(S1m, S2m, k, wk);
                            break;
                        }
                    }
                }
            }
        } else if (!(a.wt[0] < 1e-3)) {
            // One replica: the point loop and each step's voxel-crossing loop
            // run as one loop of crossing iterations.  A lane takes its next
            // polyline step only when its current one is done, so the lanes of
            // a wavefront stay busy while their steps cross different numbers
            // of voxels.  The operations and their order per streamline are
            // those of the nested loops above.
            const double wk = a.wt[0];
            int i = 1;
            bool cross = false;
            // points i-1, i, i+1 of the next step in registers, point i+2
            // requested a step before it is needed
            float3 pa = {fib[0], fib[1], fib[2]}, pb = {fib[3], fib[4], fib[5]}, pc = {fib[6], fib[7], fib[8]};
            float3 pd = pc;
            if (pts_n > 3) pd = {fib[9], fib[10], fib[11]};
            D3 S1m = {0, 0, 0}, S2m = {0, 0, 0}, dir1 = {0, 0, 0};
            double f2x = 0, f2y = 0, f2z = 0;               // floor(S2m / pixdim), fixed for a step
            D3 inv = {0, 0, 0};                             // 1 / dir1, formed at the step's first crossing
            bool have_inv = false;
            long long it = 0, cap = -1;
            for (;;) {
                if (!cross) {
                    while (i < pts_n) {
                        D3 S2 = {(double)(pb.x + a.sx), (double)(pb.y + a.sy), (double)(pb.z + a.sz)};
                        dir1 = {S2.x - (double)(pa.x + a.sx), S2.y - (double)(pa.y + a.sy),
                                S2.z - (double)(pa.z + a.sz)};
                        dir1 = vnormalize(dir1);
                        if (i == pts_n - 1) {
                            dir2 = dir1;
                        } else {
                            dir2.x = (double)(pc.x + a.sx) - S2.x;
                            dir2.y = (double)(pc.y + a.sy) - S2.y;
                            dir2.z = (double)(pc.z + a.sz) - S2.z;
                            dir2 = vnormalize(dir2);
                        }
                        pa = pb;
                        pb = pc;
                        pc = pd;
                        if (i + 3 < pts_n) pd = {fib[3 * i + 9], fib[3 * i + 10], fib[3 * i + 11]};
                        D3 nn = {0.5 * (dir1.x + dir2.x), 0.5 * (dir1.y + dir2.y), 0.5 * (dir1.z + dir2.z)};
                        double dot = dir1.x * nn.x + dir1.y * nn.y + dir1.z * nn.z;
                        nn.x /= dot;
                        nn.y /= dot;
                        nn.z /= dot;
                        i++;
                        const D3 Po = P0;
                        double len = (S2.x - Po.x) * nn.x + (S2.y - Po.y) * nn.y + (S2.z - Po.z) * nn.z;
                        if (!(len > 0)) continue;
                        D3 Pn = {Po.x + dir1.x * len, Po.y + dir1.y * len, Po.z + dir1.z * len};
                        P0 = Pn;
                        if (!a.intersect) {
                            rec.segment(Po, Pn, 0, wk);
                            continue;
                        }
                        S1m = Po;
                        S2m = Pn;
                        f2x = floor_div(S2m.x, a.px, a.rpx);
                        f2y = floor_div(S2m.y, a.py, a.rpy);
                        f2z = floor_div(S2m.z, a.pz, a.rpz);
                        have_inv = false;
                        it = 0;
                        cap = -1;
                        cross = true;
                        break;
                    }
                    if (!cross) break;
                }
                // one iteration of the step's crossing loop
                double dx = S2m.x - S1m.x, dy = S2m.y - S1m.y, dz = S2m.z - S1m.z;
                double sl = sqrt(dx * dx + dy * dy + dz * dz);
                if (sl <= a.min_seg) {
                    cross = false;
                    continue;
                }
                if (cap < 0) cap = max_cross(sl, a);
                if (it >= cap) {
                    capped = true;
                    cross = false;
                    continue;
                }
                it++;
                D3 Pb = S2m;
                bool fin = true;
                if (!(floor_div(S1m.x, a.px, a.rpx) == f2x && floor_div(S1m.y, a.py, a.rpy) == f2y &&
                      floor_div(S1m.z, a.pz, a.rpz) == f2z)) {
                    D3 vmin, vmax;
                    vmin.x = floor_div(S1m.x + 1e-6 * dir1.x, a.px, a.rpx) * a.px;
                    vmin.y = floor_div(S1m.y + 1e-6 * dir1.y, a.py, a.rpy) * a.py;
                    vmin.z = floor_div(S1m.z + 1e-6 * dir1.z, a.pz, a.rpz) * a.pz;
                    vmax.x = vmin.x + a.px;
                    vmax.y = vmin.y + a.py;
                    vmax.z = vmin.z + a.pz;
                    if (!have_inv) {
                        inv = {1.0 / dir1.x, 1.0 / dir1.y, 1.0 / dir1.z};
                        have_inv = true;
                    }
                    double t;
                    if (raybox(S1m, inv, vmin, vmax, t) && t > 0 && t < sl) {
                        Pb = {S1m.x + t * dir1.x, S1m.y + t * dir1.y, S1m.z + t * dir1.z};
                        fin = false;
                    }
                }
                rec.segment(S1m, Pb, 0, wk);
                if (fin) cross = false;
                else S1m = Pb;
            }
        }
    }
    if constexpr (EMIT) rec.flush();
    if (!EMIT || BOUNDED) {
        cnt[f] = rec.n;
        flen[f] = rec.flen;
        flentot[f] = rec.flentot;
    }
    if (capped || (EMIT && !BOUNDED && rec.n != rec.limit)) status[0] = 1;
    if (BOUNDED && rec.n > rec.limit) status[2] = 1;
}

// Slots of streamline f for the one-pass walk of a single replica: pieces of
// one polyline step are one more than the voxel boundaries it crosses, at most
// 4 for a step under a voxel per axis; 16 more for the rest.
__global__ __launch_bounds__(kVoxBlock) void k_vox_bound(const int64_t *__restrict__ fptr, int nF,
                                                         uint32_t *__restrict__ bnd)
{
    const int f = blockIdx.x * kVoxBlock + threadIdx.x;
    if (f >= nF) return;
    const int64_t n = fptr[f + 1] - fptr[f];
    bnd[f] = n > 2 ? (uint32_t)(4 * n + 16) : 0u;
}

// bits[w] bit l = (mask[64 w + l] != 0): the mask as one bit per voxel, a
// wavefront per 64-bit word
__global__ __launch_bounds__(kVoxBlock) void k_vox_mask_bits(const float *__restrict__ mask, uint64_t nvox,
                                                             uint64_t *__restrict__ bits)
{
    const uint64_t v = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    const uint64_t m = __ballot(v < nvox && mask[v < nvox ? v : 0] != 0.0f);
    if ((threadIdx.x & (kWave - 1)) == 0 && v < nvox) bits[v >> 6] = m;
}

// end[f] = off[f] + cnt[f]: the recorded part of streamline f's region
__global__ __launch_bounds__(kVoxBlock) void k_vox_ends(const uint64_t *__restrict__ off,
                                                        const uint32_t *__restrict__ cnt, int nF,
                                                        uint64_t *__restrict__ end)
{
    const int f = blockIdx.x * kVoxBlock + threadIdx.x;
    if (f < nF) end[f] = off[f] + cnt[f];
}

// Walk order: the streamlines of one wavefront take similar time when they
// have similar point counts, so threads take them by descending count
// (a stable radix sort of these keys; ties keep streamline order).
__global__ __launch_bounds__(kVoxBlock) void k_vox_order_keys(const int64_t *__restrict__ fptr, int nF,
                                                              uint32_t *__restrict__ key, uint32_t *__restrict__ idx)
{
    const int f = blockIdx.x * kVoxBlock + threadIdx.x;
    if (f >= nF) return;
    const int64_t n = fptr[f + 1] - fptr[f];
    key[f] = (uint32_t)(16383 - (n < 16383 ? n : 16383));
    idx[f] = (uint32_t)f;
}

// Equal keys of a streamline's sorted pieces [off[f], end[f]) summed into a
// float in recording order; m[f] merged segments stored from the streamline's
// first slot.  One
// wavefront per streamline: each lane of a 64-piece chunk that starts a run
// of equal keys adds the run in order and stores it at its rank among the
// run heads.
template <class KT, class VT>
__global__ __launch_bounds__(kVoxBlock) void k_vox_merge(const uint64_t *__restrict__ off,
                                                         const uint64_t *__restrict__ end, int nF,
                                                         const KT *__restrict__ skey, const VT *__restrict__ sval,
                                                         uint64_t nvox, const float *__restrict__ flen, float min_fib,
                                                         float max_fib, KT *__restrict__ mkey,
                                                         float *__restrict__ mlen, uint32_t *__restrict__ vkey,
                                                         uint32_t *__restrict__ mcnt, uint64_t *__restrict__ mend,
                                                         uint32_t *__restrict__ kept)
{
    const int f = blockIdx.x * (kVoxBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (f >= nF) return;
    const uint64_t b = off[f], e = end[f];
    uint64_t m = 0;
    for (uint64_t base = b; base < e; base += kWave) {
        const uint64_t i = base + lane;
        const bool valid = i < e;
        const KT k = valid ? skey[i] : (KT)0;
        const bool head = valid && (i == b || skey[i - 1] != k);
        const uint64_t heads = __ballot(head);
        if (head) {
            float acc = 0.0f;
            for (uint64_t j = i; j < e && skey[j] == k; j++) acc = (float)((double)acc + (double)sval[j]);
            const uint64_t pos = b + m + __popcll(heads & ((1ull << lane) - 1ull));
            mkey[pos] = k;
            mlen[pos] = acc;
            vkey[pos] = (uint32_t)((uint64_t)k % nvox);
        }
        m += __popcll(heads);
    }
    if (lane != 0) return;
    const float fl = flen[f];
    const uint32_t kf = (m > 0 && fl > min_fib && fl < max_fib) ? 1u : 0u;
    mcnt[f] = kf ? (uint32_t)m : 0u;
    mend[f] = kf ? b + m : b;
    kept[f] = kf;
}

__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const long long u = __double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// TRK_norm of each kept streamline from its segments sorted by voxel,
// [off[f], mend[f]) (:530-537): per voxel the float sum of its lengths in
// order, then the float sum of their squares in voxel order and a float
// square root.  A wavefront per streamline: the lane that starts a voxel's
// run adds the run in order and squares it; the squares are then added one
// by one in run order (readlane).
__global__ __launch_bounds__(kVoxBlock) void k_vox_norm(const uint64_t *__restrict__ off,
                                                        const uint64_t *__restrict__ mend, int nF,
                                                        const uint32_t *__restrict__ vsort,
                                                        const float *__restrict__ lsort, float *__restrict__ norm)
{
    const int f = blockIdx.x * (kVoxBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (f >= nF) return;
    const uint64_t b = off[f], e = mend[f];
    float fn = 0.0f;
    for (uint64_t base = b; base < e; base += kWave) {
        const uint64_t i = base + lane;
        const bool valid = i < e;
        const uint32_t v = valid ? vsort[i] : 0u;
        const bool head = valid && (i == b || vsort[i - 1] != v);
        double sq = 0.0;
        if (head) {
            float r = 0.0f;
            for (uint64_t j = i; j < e && vsort[j] == v; j++) r = r + lsort[j];
            sq = (double)r * (double)r;
        }
        unsigned long long hm = __ballot(head);
        while (hm) {
            const int l = (int)__builtin_ctzll(hm);
            hm &= hm - 1ull;
            fn = (float)((double)fn + readlane_f64(sq, l));
        }
    }
    if (lane == 0) norm[f] = (float)sqrt((double)fn);
}

// Streamlines of at most kSortCap pieces take k_vox_sortmerge, in launches
// by size class: at most kSortSmall[0] pieces, at most kSortSmall[1], and the
// rest, each with LDS slices sized to its class.
constexpr int kSortCap = 2048;
constexpr int kSortSmall[2] = {256, 512};

// One wavefront's LDS writes done and visible to all its lanes.
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// Ascending bitonic sort of sk[0, N) (N a power of two, >= 64) by one
// wavefront, in LDS.
__device__ __forceinline__ void bitonic_u64(uint64_t *sk, int N, int lane)
{
    for (int k = 2; k <= N; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < (N >> 1); t += kWave) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int l = i | j;
                const uint64_t x = sk[i], y = sk[l];
                if ((x > y) == ((i & k) == 0)) {
                    sk[i] = y;
                    sk[l] = x;
                }
            }
            wave_lds_sync();
        }
}

// Ascending bitonic sort of N = 64 E words held by one wavefront in
// registers, word p = 64-bit v[e] of lane p / E, e = p % E: stages whose
// pairs lie inside a lane compare in registers, the others exchange with the
// partner lane (lane ^ j / E) by shuffles.
template <int E, int J>
__device__ __forceinline__ void bitonic_lane_stage(uint64_t (&v)[8], int lane, int k)
{
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int pe = e ^ J;
        if (pe > e) {
            const bool up = ((lane * E + e) & k) == 0;
            const uint64_t x = v[e], y = v[pe];
            const bool sw = (x > y) == up;
            v[e] = sw ? y : x;
            v[pe] = sw ? x : y;
        }
    }
}

template <int E>
__device__ __forceinline__ void bitonic_reg(uint64_t (&v)[8], int lane)
{
    constexpr int N = kWave * E;
    for (int k = 2; k <= N; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= E) {
                const int lm = j / E;
                const bool lower = (lane & lm) == 0;
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const uint64_t o = (uint64_t)__shfl_xor((unsigned long long)v[e], lm);
                    const bool up = ((lane * E + e) & k) == 0;
                    const bool keep_min = lower == up;
                    v[e] = keep_min ? (v[e] < o ? v[e] : o) : (v[e] > o ? v[e] : o);
                }
            } else if (j == 4) {
                if constexpr (E > 4) bitonic_lane_stage<E, 4>(v, lane, k);
            } else if (j == 2) {
                if constexpr (E > 2) bitonic_lane_stage<E, 2>(v, lane, k);
            } else {
                if constexpr (E > 1) bitonic_lane_stage<E, 1>(v, lane, k);
            }
        }
}

// Sort sk[0, N) (N = 64 E, E in 1, 2, 4, 8): in registers, the result back
// in sk.  Entries past the valid ones are all-ones words (they sort last).
__device__ __forceinline__ void sort_words(uint64_t *sk, int N, int lane)
{
    uint64_t v[8];
    const int E = N / kWave;
#define SORT_E(EE)                                                   \
    {                                                                \
        _Pragma("unroll") for (int e = 0; e < EE; e++) v[e] = sk[lane * EE + e]; \
        bitonic_reg<EE>(v, lane);                                    \
        _Pragma("unroll") for (int e = 0; e < EE; e++) sk[lane * EE + e] = v[e]; \
    }
    if (E == 1) SORT_E(1) else if (E == 2) SORT_E(2) else if (E == 4) SORT_E(4) else SORT_E(8)
#undef SORT_E
    wave_lds_sync();
}

// The IC stage between the walk and the output for 32-bit keys, one
// wavefront per streamline, in LDS: the pieces [off[f], end[f]) taken as runs
// of consecutive pieces of one key (nearly every key's pieces are one run:
// runs are 46% of pieces and 100.4% of merged segments on the brain
// tractogram), the runs sorted by (key, start), each key's pieces summed into
// a float in recording order (run by run, piece by piece) and stored from the
// streamline's first slot (keys over the streamline's own unsorted keys,
// lengths to mlen), the kept flag, and
// for a kept streamline its merged segments sorted by (voxel, rank) and
// TRK_norm: per voxel the float sum of its lengths in order, the float sum of
// their squares in voxel order, a float square root.  The operations and
// their order are those of the segmented sorts, k_vox_merge and k_vox_norm.
// A launch takes the streamlines of nlo < pieces <= nmax (a power of two).
// Each wavefront of a work-group takes its own streamline and its own LDS
// slice of nmax x (8-byte sort word, piece value, merged key, merged length).
template <class VT>
__global__ __launch_bounds__(kVoxBlock) void k_vox_sortmerge(const uint64_t *__restrict__ off,
                                                         const uint64_t *__restrict__ end, int nF, uint32_t *keys,
                                                         const VT *__restrict__ vals, uint64_t nvox,
                                                         const float *__restrict__ flen, float min_fib,
                                                         float max_fib, int nmax, int nlo, float *__restrict__ mlen,
                                                         uint32_t *__restrict__ mcnt, uint64_t *__restrict__ mend,
                                                         uint32_t *__restrict__ kept, float *__restrict__ norm)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & (kWave - 1));
    // this wavefront's slice of the work-group's LDS
    uint64_t *sk = reinterpret_cast<uint64_t *>(smem + (size_t)w * nmax * (sizeof(uint64_t) + sizeof(VT) + 8));
    VT *sv = reinterpret_cast<VT *>(sk + nmax);
    uint32_t *mk = reinterpret_cast<uint32_t *>(sv + nmax);
    float *ml = reinterpret_cast<float *>(mk + nmax);
    const int f = (int)(blockIdx.x * (blockDim.x >> 6)) + w;
    if (f >= nF) return;
    const uint64_t b = off[f];
    const int n = (int)(end[f] - b);
    if (n <= nlo || n > nmax) return;                 // another launch's size class
    // runs: sk[j] = key << 32 | start, in recording order, then start << 11 |
    // (length - 1) in the low word (start < 2048 and length <= 2048 = the
    // largest n), so that runs sort by (key, start)
    int nr = 0;
    for (int base = 0; base < n; base += kWave) {
        const int i = base + lane;
        const bool valid = i < n;
        const uint32_t k = valid ? keys[b + i] : 0u;
        const bool head = valid && (i == 0 || keys[b + i - 1] != k);
        if (valid) sv[i] = vals[b + i];
        const unsigned long long hm = __ballot(head);
        if (head) sk[nr + __popcll(hm & ((1ull << lane) - 1ull))] = ((uint64_t)k << 32) | (uint32_t)i;
        nr += __popcll(hm);
    }
    wave_lds_sync();
    for (int base = 0; base < nr; base += kWave) {
        const int j = base + lane;
        uint64_t w = 0;
        uint32_t next = (uint32_t)n;
        if (j < nr) {
            w = sk[j];
            if (j + 1 < nr) next = (uint32_t)sk[j + 1];
        }
        wave_lds_sync();                              // the chunk's reads before its writes
        if (j < nr) {
            const uint32_t st = (uint32_t)w;
            sk[j] = (w & 0xffffffff00000000ull) | (st << 11) | (next - st - 1u);
        }
        wave_lds_sync();
    }
    int N = kWave;
    while (N < nr) N <<= 1;
    for (int i = nr + lane; i < N; i += kWave) sk[i] = ~0ull;
    wave_lds_sync();
    if (N <= kWave * 8) sort_words(sk, N, lane);
    else bitonic_u64(sk, N, lane);
    // a key's runs, adjacent and by start: the first run's lane adds the key's
    // pieces into a float in recording order (the map's +=) and stores the
    // sum at its rank
    int m = 0;
    for (int base = 0; base < nr; base += kWave) {
        const int p = base + lane;
        const bool valid = p < nr;
        const uint32_t k = valid ? (uint32_t)(sk[p] >> 32) : 0u;
        const bool head = valid && (p == 0 || (uint32_t)(sk[p - 1] >> 32) != k);
        const unsigned long long heads = __ballot(head);
        if (head) {
            float acc = 0.0f;
            for (int q = p; q < nr && (uint32_t)(sk[q] >> 32) == k; q++) {
                const uint32_t lo = (uint32_t)sk[q], st = lo >> 11, len = (lo & 0x7ffu) + 1u;
                for (uint32_t t = st; t < st + len; t++) acc = (float)((double)acc + (double)sv[t]);
            }
            const int r = m + __popcll(heads & ((1ull << lane) - 1ull));
            keys[b + r] = k;
            mlen[b + r] = acc;
            mk[r] = k;
            ml[r] = acc;
        }
        m += __popcll(heads);
    }
    const float fl = flen[f];
    const bool kf = m > 0 && fl > min_fib && fl < max_fib;
    if (lane == 0) {
        mcnt[f] = kf ? (uint32_t)m : 0u;
        mend[f] = kf ? b + m : b;
        kept[f] = kf ? 1u : 0u;
    }
    if (!kf) {
        if (lane == 0) norm[f] = 0.0f;
        return;
    }
    wave_lds_sync();                                  // the merge's reads of sk are done
    int N2 = kWave;
    while (N2 < m) N2 <<= 1;
    for (int i = lane; i < N2; i += kWave)
        sk[i] = i < m ? ((uint64_t)(uint32_t)((uint64_t)mk[i] % nvox) << 32) | (uint32_t)i : ~0ull;
    wave_lds_sync();
    if (N2 <= kWave * 8) sort_words(sk, N2, lane);
    else bitonic_u64(sk, N2, lane);
    float fn = 0.0f;
    for (int base = 0; base < m; base += kWave) {
        const int p = base + lane;
        const bool valid = p < m;
        const uint32_t v = valid ? (uint32_t)(sk[p] >> 32) : 0u;
        const bool head = valid && (p == 0 || (uint32_t)(sk[p - 1] >> 32) != v);
        double sq = 0.0;
        if (head) {
            float r = 0.0f;
            for (int q = p; q < m && (uint32_t)(sk[q] >> 32) == v; q++) r = r + ml[(uint32_t)sk[q]];
            sq = (double)r * (double)r;
        }
        unsigned long long hm = __ballot(head);
        while (hm) {
            const int l = (int)__builtin_ctzll(hm);
            hm &= hm - 1ull;
            fn = (float)((double)fn + readlane_f64(sq, l));
        }
    }
    if (lane == 0) norm[f] = (float)sqrt((double)fn);
}

// Kept streamlines' segments and per-streamline values, in streamline order.
template <class KT>
__global__ __launch_bounds__(kVoxBlock) void k_vox_write(
    const uint64_t *__restrict__ off, int nF, const uint32_t *__restrict__ kept, const uint64_t *__restrict__ kidx,
    const uint64_t *__restrict__ spos, const uint32_t *__restrict__ mcnt, const KT *__restrict__ mkey,
    const float *__restrict__ mlen, uint64_t nvox, const float *__restrict__ norm, const float *__restrict__ flen,
    const float *__restrict__ flentot, uint32_t *__restrict__ IC_f, uint32_t *__restrict__ IC_v,
    uint16_t *__restrict__ IC_o, float *__restrict__ IC_len, float *__restrict__ TRK_norm,
    float *__restrict__ TRK_len, float *__restrict__ TRK_lenTot, uint8_t *__restrict__ TRK_kept)
{
    const int f = blockIdx.x * (kVoxBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (f >= nF) return;
    const uint32_t kf = kept[f];
    if (lane == 0) TRK_kept[f] = (uint8_t)kf;
    if (!kf) return;
    const uint64_t kk = kidx[f], s = spos[f], b = off[f];
    const uint32_t m = mcnt[f];
    for (uint32_t i = lane; i < m; i += kWave) {
        const uint64_t key = (uint64_t)mkey[b + i];
        const uint64_t o = key / nvox;
        IC_f[s + i] = (uint32_t)kk;
        IC_v[s + i] = (uint32_t)(key - o * nvox);
        IC_o[s + i] = (uint16_t)o;
        IC_len[s + i] = mlen[b + i];
    }
    if (lane != 0) return;
    TRK_norm[kk] = norm[f];
    TRK_len[kk] = flen[f];
    TRK_lenTot[kk] = flentot[f];
}

// Each voxel's range [rs[v], re[v]) in the segments sorted by voxel (stable,
// so each voxel's lengths are in segment order); voxels without segments keep
// rs = re = 0.
__global__ __launch_bounds__(kVoxBlock) void k_vox_runs(const uint32_t *__restrict__ vs, uint64_t n,
                                                        uint32_t *__restrict__ rs, uint32_t *__restrict__ re)
{
    const uint64_t i = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = vs[i];
    if (i == 0 || vs[i - 1] != v) rs[v] = (uint32_t)i;
    if (i == n - 1 || vs[i + 1] != v) re[v] = (uint32_t)(i + 1);
}

// compute_tdi (trk2dictionary.pyx:50-59): a wavefront per voxel loads its
// lengths 64 at a time and adds them one by one in segment order (the float
// sum of the reference's loop); marks the voxels holding an IC segment.
__global__ __launch_bounds__(kVoxBlock) void k_vox_tdi(const uint32_t *__restrict__ rs,
                                                       const uint32_t *__restrict__ re,
                                                       const float *__restrict__ ls, uint64_t nvox,
                                                       float *__restrict__ tdi, uint8_t *__restrict__ has)
{
    const uint64_t v = (uint64_t)blockIdx.x * (kVoxBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (v >= nvox) return;
    const uint32_t b = rs[v], e = re[v];
    if (b == e) return;
    float s = 0.0f;
    for (uint32_t base = b; base < e; base += kWave) {
        const float x = base + lane < e ? ls[base + lane] : 0.0f;
        const int cnt = (int)(e - base < (uint32_t)kWave ? e - base : (uint32_t)kWave);
        for (int k = 0; k < cnt; k++)
            s = s + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
    }
    if (lane == 0) {
        tdi[v] = s;
        has[v] = 1;
    }
}

// fib_ptr checks on the device: res[0] bit 0 a decreasing offset, bit 1 a
// streamline of kMaxFibLen points or more, bit 2 bad first or last offset;
// res[1] the first streamline of bit 1.
__global__ __launch_bounds__(kVoxBlock) void k_vox_check(const int64_t *__restrict__ fptr, int nF, int64_t npts,
                                                         uint32_t *__restrict__ res)
{
    const int f = blockIdx.x * kVoxBlock + threadIdx.x;
    if (f == 0 && (fptr[0] != 0 || fptr[nF] != npts)) atomicOr(&res[0], 4u);
    if (f >= nF) return;
    const int64_t n = fptr[f + 1] - fptr[f];
    if (n < 0) atomicOr(&res[0], 1u);
    if (n >= kMaxFibLen) {
        atomicOr(&res[0], 2u);
        atomicMin(&res[1], (uint32_t)f);
    }
}

__global__ __launch_bounds__(kVoxBlock) void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kVoxBlock) void k_fill_f32(float *p, uint64_t n, float v)
{
    const uint64_t i = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kVoxBlock) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

struct EcArgs {
    int X, Y, Z, Np;
    float vf_thr;
    int fx, fy, fz;
    double A[9];
    const float *peaks;               // C order [X][Y][Z][3 Np]
    const short *hash;
};

// ECSegments (trk2dictionary_c.cpp:289-397) for voxel v (x fastest, so the
// reference's z, y, x loops visit ascending v).  WRITE = false counts.
template <bool WRITE>
__global__ __launch_bounds__(kVoxBlock) void k_vox_ec(EcArgs a, uint64_t nvox, const uint8_t *__restrict__ has,
                                                      uint32_t *__restrict__ cnt, const uint64_t *__restrict__ eoff,
                                                      uint32_t *__restrict__ EC_v, uint16_t *__restrict__ EC_o)
{
    const uint64_t v = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (v >= nvox) return;
    uint32_t c = 0;
    if (has[v]) {
        const uint64_t ix = v % a.X, iy = (v / a.X) % a.Y, iz = v / ((uint64_t)a.X * a.Y);
        const float *base = a.peaks + 3 * (size_t)a.Np * (iz + (size_t)a.Z * (iy + (size_t)a.Y * ix));
        float peakMax = -1;
        for (int id = 0; id < a.Np; id++) {
            D3 d = {base[3 * id], base[3 * id + 1], base[3 * id + 2]};
            float nm = vnorm(d);
            if (nm > peakMax) peakMax = nm;
        }
        if (peakMax > 0) {
            const uint64_t o0 = WRITE ? eoff[v] : 0;
            for (int id = 0; id < a.Np; id++) {
                const float *q = base + 3 * id;
                D3 d0 = {q[0], q[1], q[2]};
                float nm = vnorm(d0);
                if (nm == 0 || nm < a.vf_thr * peakMax) continue;
                if (WRITE) {
                    D3 d;
                    d.x = q[0] * a.A[0] + q[1] * a.A[1] + q[2] * a.A[2];
                    d.y = q[0] * a.A[3] + q[1] * a.A[4] + q[2] * a.A[5];
                    d.z = q[0] * a.A[6] + q[1] * a.A[7] + q[2] * a.A[8];
                    d.x *= a.fx;
                    d.y *= a.fy;
                    d.z *= a.fz;
                    if (d.y < 0) {
                        d.x = -d.x;
                        d.y = -d.y;
                        d.z = -d.z;
                    }
                    double colat = atan2(sqrt(d.x * d.x + d.y * d.y), d.z);
                    double lon = atan2(d.y, d.x);
                    EC_v[o0 + c] = (uint32_t)v;
                    EC_o[o0 + c] = (uint16_t)hash_index(a.hash, colat, lon);
                }
                c++;
            }
        }
    }
    if (!WRITE) cnt[v] = c;
}

// ISOcompartments (trk2dictionary_c.cpp:400-437).  WRITE = false flags.
template <bool WRITE>
__global__ __launch_bounds__(kVoxBlock) void k_vox_iso(int X, int Y, int Z, uint64_t nvox,
                                                       const float *__restrict__ iso, const float *__restrict__ mask,
                                                       const uint8_t *__restrict__ has, uint32_t *__restrict__ flag,
                                                       const uint64_t *__restrict__ ioff, uint32_t *__restrict__ ISO_v)
{
    const uint64_t v = (uint64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (v >= nvox) return;
    if (WRITE) {
        if (flag[v]) ISO_v[ioff[v]] = (uint32_t)v;
        return;
    }
    const uint64_t ix = v % X, iy = (v / X) % Y, iz = v / ((uint64_t)X * Y);
    const size_t c = iz + (size_t)Z * (iy + (size_t)Y * ix);
    uint32_t ok = has[v] ? 1u : 0u;
    if (iso && iso[c] == 0) ok = 0;
    if (mask && mask[c] == 0) ok = 0;
    flag[v] = ok;
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kVoxBlock - 1) / kVoxBlock); }

inline int bit_length(uint64_t x)
{
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

}  // namespace

// ---------------------------------------------------------------- context --
struct commit_hip_vox {
    int device = 0;
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
    };
    enum {
        PTS, FPTR, MASK, ISO, PEAKS, HASH, REP, APPLY, PBUF, CNT, FLEN, FLENTOT, OFF, K1, V1, K2, V2, L1,
        MCNT, MEND, KEPT, KIDX, SPOS, NORM, STATUS, TMP,
        IC_F, IC_V, IC_O, IC_LEN, TRK_NORM, TRK_LEN, TRK_LENTOT, TRK_KEPT, TDI_V, TDI_L, TDI, HAS,
        ECNT, EOFF, EC_V, EC_O, IFLAG, IOFF, ISO_V, OKEY, OIDX, OKEY2, ORDER, END, POFF, MBITS, NBUF
    };
    Buf buf[NBUF];
    int64_t n_ic = 0, n_kept = 0, n_ec = 0, n_iso = 0, n_count = 0, n_pieces = 0;
    uint64_t nvox = 0;
    bool valid = false;

    ~commit_hip_vox()
    {
        for (auto &b : buf)
            if (b.p) (void)hipFree(b.p);
    }
    // grow buffer i to at least `bytes` (contents not kept)
    int need(int i, size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        if (buf[i].bytes >= bytes) return 0;
        if (buf[i].p) (void)hipFree(buf[i].p);
        buf[i].p = nullptr;
        buf[i].bytes = 0;
        if (hipMalloc(&buf[i].p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(COMMIT_HIP_ENOMEM, "hipMalloc of %zu bytes failed (voxeliser buffer %d)", bytes, i);
        }
        buf[i].bytes = bytes;
        return 0;
    }
    template <class T>
    T *at(int i) { return (T *)buf[i].p; }
};

namespace {

// temporary storage of the rocprim calls, grown on demand
int ensure_tmp(commit_hip_vox *V, size_t bytes) { return V->need(commit_hip_vox::TMP, bytes); }

int exscan_u32(commit_hip_vox *V, const uint32_t *in, uint64_t *out, size_t n, hipStream_t s)
{
    size_t bytes = 0;
    HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
    int rc = ensure_tmp(V, bytes);
    if (rc) return rc;
    bytes = V->buf[commit_hip_vox::TMP].bytes;
    HIP_TRY(rocprim::exclusive_scan(V->buf[commit_hip_vox::TMP].p, bytes, in, out, (uint64_t)0, n,
                                    rocprim::plus<uint64_t>(), s));
    return 0;
}

template <class K, class Vt>
int segsort(commit_hip_vox *V, const K *ki, K *ko, const Vt *vi, Vt *vo, size_t n, int nseg, const uint64_t *b,
            const uint64_t *e, int bits, hipStream_t s)
{
    if (n == 0 || nseg == 0) return 0;
    size_t bytes = 0;
    HIP_TRY(rocprim::segmented_radix_sort_pairs(nullptr, bytes, ki, ko, vi, vo, (unsigned)n, (unsigned)nseg, b, e,
                                                0u, (unsigned)bits, s));
    int rc = ensure_tmp(V, bytes);
    if (rc) return rc;
    bytes = V->buf[commit_hip_vox::TMP].bytes;
    HIP_TRY(rocprim::segmented_radix_sort_pairs(V->buf[commit_hip_vox::TMP].p, bytes, ki, ko, vi, vo, (unsigned)n,
                                                (unsigned)nseg, b, e, 0u, (unsigned)bits, s));
    return 0;
}

template <class Vt>
int sort_pairs_u32(commit_hip_vox *V, const uint32_t *ki, uint32_t *ko, const Vt *vi, Vt *vo, size_t n, int bits,
                   hipStream_t s)
{
    if (n == 0) return 0;
    size_t bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, ki, ko, vi, vo, (unsigned)n, 0u, (unsigned)bits, s));
    int rc = ensure_tmp(V, bytes);
    if (rc) return rc;
    bytes = V->buf[commit_hip_vox::TMP].bytes;
    HIP_TRY(rocprim::radix_sort_pairs(V->buf[commit_hip_vox::TMP].p, bytes, ki, ko, vi, vo, (unsigned)n, 0u,
                                      (unsigned)bits, s));
    return 0;
}

template <class T>
int upload(commit_hip_vox *V, int i, const T *host, size_t count, hipStream_t s)
{
    int rc = V->need(i, count * sizeof(T));
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpyAsync(V->buf[i].p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    return 0;
}

// ic_segments with a bounded walk: some streamline had more pieces than its
// region holds, and the counting walk must size the regions (not an API code)
constexpr int kRetryExact = 1;

// Arguments of the IC stage (ic_segments).
struct IcArgs {
    VoxArgs a;
    const float *pts;
    const int64_t *ptr;
    int n_count;
    D3 *P;
    const uint64_t *off;               // streamline regions start here
    bool bounded;                      // one-pass walk into regions of off[f+1] - off[f] slots
    uint32_t *cnt;                     // pieces per streamline (written by the one-pass walk)
    float *flen, *flentot;
    uint32_t *status;
    const uint32_t *order;
    size_t np;                         // slots of all regions
    int kbits, vbits;
    float min_fib, max_fib;
    hipStream_t s;
};

// The IC stage from the recording walk to the written segments, with piece
// keys as KT and piece values as VT (see Recorder): walk, sort of each
// streamline's pieces by key, merge, sort of each kept streamline's segments
// by voxel, TRK_norm, renumbering and the IC / TRK outputs.
template <class KT, class VT>
int ic_segments(commit_hip_vox *V, const IcArgs &ia, size_t *n_ic_out, size_t *n_kept_out)
{
    using B = commit_hip_vox;
    const VoxArgs &a = ia.a;
    const hipStream_t s = ia.s;
    const int n_count = ia.n_count;
    const size_t nF = (size_t)n_count, np = ia.np;
    const uint64_t nvox = a.nvox;
    const uint64_t *d_off = ia.off;
    const bool multi = a.n_rep > 1;
    const unsigned gF = grid_for(nF ? nF : 1);
    const unsigned gW = (unsigned)((nF + kVoxBlock / kWave - 1) / (kVoxBlock / kWave));   // a wavefront per streamline
    // V1 later holds the 32-bit voxel keys, V2 the float lengths
    const size_t vb = np * (sizeof(VT) > 4 ? sizeof(VT) : 4);
    int rc;
    if ((rc = V->need(B::K1, np * sizeof(KT))) || (rc = V->need(B::V1, vb)) || (rc = V->need(B::K2, np * sizeof(KT))) ||
        (rc = V->need(B::V2, vb)) || (rc = V->need(B::L1, np * 4)))
        return rc;
    KT *K1 = V->at<KT>(B::K1), *K2 = V->at<KT>(B::K2);
    VT *V1 = V->at<VT>(B::V1), *V2 = V->at<VT>(B::V2);
    float *L1 = V->at<float>(B::L1);
    if (nF) {
        if (ia.bounded)
            k_vox_walk<true, false, KT, VT, true><<<gF, kVoxBlock, 0, s>>>(a, ia.pts, ia.ptr, n_count, nullptr,
                                                                           d_off, K1, V1, ia.cnt, ia.flen,
                                                                           ia.flentot, ia.status, ia.order);
        else if (multi)
            k_vox_walk<true, true, KT, VT><<<gF, kVoxBlock, 0, s>>>(a, ia.pts, ia.ptr, n_count, ia.P, d_off, K1, V1,
                                                                    nullptr, nullptr, nullptr, ia.status, ia.order);
        else
            k_vox_walk<true, false, KT, VT><<<gF, kVoxBlock, 0, s>>>(a, ia.pts, ia.ptr, n_count, nullptr, d_off, K1,
                                                                     V1, nullptr, nullptr, nullptr, ia.status,
                                                                     ia.order);
        HIP_TRY(hipGetLastError());
    }
    uint32_t hst[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(hst, ia.status, 12, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint64_t *d_end = d_off + 1;
    if (ia.bounded) {
        if (hst[0]) return fail(COMMIT_HIP_EUNSUPPORTED, "a polyline step crossed more voxels than its cap allows");
        if (hst[2]) return kRetryExact;
        // each region's recorded part, and the pieces in all
        uint64_t *end = V->at<uint64_t>(B::END), *poff = V->at<uint64_t>(B::POFF);
        if (nF) {
            k_vox_ends<<<gF, kVoxBlock, 0, s>>>(d_off, ia.cnt, n_count, end);
            HIP_TRY(hipGetLastError());
        }
        if ((rc = exscan_u32(V, ia.cnt, poff, nF + 1, s))) return rc;
        uint64_t npieces = 0;
        HIP_TRY(hipMemcpyAsync(&npieces, poff + nF, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        V->n_pieces = (int64_t)npieces;
        d_end = end;
    } else if (hst[0]) {
        return fail(COMMIT_HIP_EDEVICE, "the recording pass differs from the counting pass");
    }
    uint32_t *d_mcnt = V->at<uint32_t>(B::MCNT), *d_kept = V->at<uint32_t>(B::KEPT);
    uint64_t *d_mend = V->at<uint64_t>(B::MEND);
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_mcnt + nF, 1, 0u);
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_kept + nF, 1, 0u);
    bool fused = false;
    if constexpr (sizeof(KT) == 4) {
        if (nF) {
            // the longest streamline's pieces decide whether k_vox_sortmerge fits
            uint32_t *dmax = ia.status + 3;
            size_t bytes = 0;
            HIP_TRY(rocprim::reduce(nullptr, bytes, (const uint32_t *)ia.cnt, dmax, 0u, nF,
                                    rocprim::maximum<uint32_t>(), s));
            if ((rc = ensure_tmp(V, bytes))) return rc;
            bytes = V->buf[B::TMP].bytes;
            HIP_TRY(rocprim::reduce(V->buf[B::TMP].p, bytes, (const uint32_t *)ia.cnt, dmax, 0u, nF,
                                    rocprim::maximum<uint32_t>(), s));
            uint32_t maxn = 0;
            HIP_TRY(hipMemcpyAsync(&maxn, dmax, 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (maxn <= (uint32_t)kSortCap) {
                // size classes, so that the common short streamlines get
                // small LDS buffers and many wavefronts per CU
                int ntop = kWave;
                while (ntop < (int)maxn) ntop <<= 1;
                const int n1 = ntop < kSortSmall[0] ? ntop : kSortSmall[0];
                const int n2 = ntop < kSortSmall[1] ? ntop : kSortSmall[1];
                const size_t slot = sizeof(uint64_t) + sizeof(VT) + 2 * sizeof(uint32_t);
                for (const int2 c : {int2{-1, n1}, int2{n1, n2}, int2{n2, ntop}}) {
                    if (c.x >= c.y) continue;
                    const int wpb = 1;                        // wavefronts per work-group
                    k_vox_sortmerge<VT><<<(unsigned)((nF + wpb - 1) / wpb), wpb * kWave, (size_t)wpb * c.y * slot, s>>>(
                        d_off, d_end, n_count, (uint32_t *)K1, V1, nvox, V->at<float>(B::FLEN), ia.min_fib,
                        ia.max_fib, c.y, c.x, L1, d_mcnt, d_mend, d_kept, V->at<float>(B::NORM));
                    HIP_TRY(hipGetLastError());
                }
                fused = true;
            }
        }
    }
    if (!fused) {
        if ((rc = segsort(V, (const KT *)K1, K2, (const VT *)V1, V2, np, n_count, d_off, d_end, ia.kbits, s)))
            return rc;
        // merged segments: keys to K1, float lengths to L1, voxel keys into V1's memory
        uint32_t *VK = (uint32_t *)V1;
        if (nF) {
            k_vox_merge<KT, VT><<<gW, kVoxBlock, 0, s>>>(d_off, d_end, n_count, K2, V2, nvox, V->at<float>(B::FLEN),
                                                         ia.min_fib, ia.max_fib, K1, L1, VK, d_mcnt, d_mend, d_kept);
            HIP_TRY(hipGetLastError());
        }
        // per-voxel order inside each kept streamline: voxel keys into K2's
        // memory, lengths into V2's
        uint32_t *VS = (uint32_t *)K2;
        float *L2 = (float *)V2;
        if ((rc = segsort(V, (const uint32_t *)VK, VS, (const float *)L1, L2, np, n_count, d_off, d_mend, ia.vbits,
                          s)))
            return rc;
        if (nF) {
            k_vox_norm<<<gW, kVoxBlock, 0, s>>>(d_off, d_mend, n_count, VS, L2, V->at<float>(B::NORM));
            HIP_TRY(hipGetLastError());
        }
    }
    uint64_t *d_kidx = V->at<uint64_t>(B::KIDX), *d_spos = V->at<uint64_t>(B::SPOS);
    if ((rc = exscan_u32(V, d_kept, d_kidx, nF + 1, s))) return rc;
    if ((rc = exscan_u32(V, d_mcnt, d_spos, nF + 1, s))) return rc;
    uint64_t tot[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&tot[0], d_kidx + nF, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&tot[1], d_spos + nF, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const size_t n_kept = (size_t)tot[0], n_ic = (size_t)tot[1];
    if ((rc = V->need(B::IC_F, n_ic * 4)) || (rc = V->need(B::IC_V, n_ic * 4)) || (rc = V->need(B::IC_O, n_ic * 2)) ||
        (rc = V->need(B::IC_LEN, n_ic * 4)) || (rc = V->need(B::TRK_NORM, n_kept * 4)) ||
        (rc = V->need(B::TRK_LEN, n_kept * 4)) || (rc = V->need(B::TRK_LENTOT, n_kept * 4)) ||
        (rc = V->need(B::TRK_KEPT, nF)) || (rc = V->need(B::TDI_V, n_ic * 4)) || (rc = V->need(B::TDI_L, n_ic * 4)) ||
        (rc = V->need(B::TDI, nvox * 4)) || (rc = V->need(B::HAS, nvox)))
        return rc;
    if (nF) {
        k_vox_write<KT><<<gW, kVoxBlock, 0, s>>>(
            d_off, n_count, d_kept, d_kidx, d_spos, d_mcnt, K1, L1, nvox, V->at<float>(B::NORM), V->at<float>(B::FLEN),
            V->at<float>(B::FLENTOT), V->at<uint32_t>(B::IC_F), V->at<uint32_t>(B::IC_V), V->at<uint16_t>(B::IC_O),
            V->at<float>(B::IC_LEN), V->at<float>(B::TRK_NORM), V->at<float>(B::TRK_LEN), V->at<float>(B::TRK_LENTOT),
            V->at<uint8_t>(B::TRK_KEPT));
        HIP_TRY(hipGetLastError());
    }
    *n_ic_out = n_ic;
    *n_kept_out = n_kept;
    return 0;
}

}  // namespace

extern "C" int commit_hip_vox_create(commit_hip_vox **out, int device)
{
    if (!out) return fail(COMMIT_HIP_EINVAL, "commit_hip_vox_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(COMMIT_HIP_EINVAL, "device %d out of range (%d devices)", device, ndev);
    commit_hip_vox *V = new (std::nothrow) commit_hip_vox();
    if (!V) return fail(COMMIT_HIP_ENOMEM, "host allocation failed");
    V->device = device;
    *out = V;
    return 0;
}

extern "C" int commit_hip_vox_destroy(commit_hip_vox *V)
{
    if (!V) return 0;
    (void)hipSetDevice(V->device);
    delete V;
    return 0;
}

extern "C" int commit_hip_trk2dictionary(commit_hip_vox *V, const float *fibers, const int64_t *fib_ptr,
                                         int n_count, int64_t n_points, int inputs_on_device, int Nx, int Ny, int Nz,
                                         float Px, float Py, float Pz, float fiber_shiftX, float fiber_shiftY,
                                         float fiber_shiftZ, float min_seg_len, float min_fiber_len,
                                         float max_fiber_len, const float *ptrPEAKS, int Np, float vf_THR, int ECix,
                                         int ECiy, int ECiz, const float *ptrMASK, const float *ptrISO, int c,
                                         const double *ptrPeaksAffine, int nReplicas, const double *ptrBlurRho,
                                         const double *ptrBlurAngle, const double *ptrBlurWeights,
                                         const uint8_t *ptrBlurApplyTo, const short *ptrHashTable, void *stream)
{
    using B = commit_hip_vox;
    if (!V) return fail(COMMIT_HIP_EINVAL, "null voxeliser handle");
    V->valid = false;
    if (n_count < 0 || n_points < 0) return fail(COMMIT_HIP_EINVAL, "negative streamline or point count");
    if (Nx <= 0 || Ny <= 0 || Nz <= 0 || Nx >= 65536 || Ny >= 65536 || Nz >= 65536)
        return fail(COMMIT_HIP_EINVAL, "grid %d x %d x %d outside 1..65535 per axis", Nx, Ny, Nz);
    if (!(Px > 0) || !(Py > 0) || !(Pz > 0)) return fail(COMMIT_HIP_EINVAL, "voxel sizes must be positive");
    const uint64_t nvox = (uint64_t)Nx * Ny * Nz;
    if (nvox > 0xFFFFFFFFull) return fail(COMMIT_HIP_EUNSUPPORTED, "grid of %llu voxels: IC_v is 32-bit", (unsigned long long)nvox);
    if (!ptrHashTable) return fail(COMMIT_HIP_EINVAL, "null orientation hash table");
    if (nReplicas < 1 || !ptrBlurRho || !ptrBlurAngle || !ptrBlurWeights)
        return fail(COMMIT_HIP_EINVAL, "need at least one replica with its rho, angle and weight");
    if (ptrPEAKS && (Np <= 0 || !ptrPeaksAffine)) return fail(COMMIT_HIP_EINVAL, "peaks need Np > 0 and an affine");
    if (n_count > 0 && (!fibers || !fib_ptr)) return fail(COMMIT_HIP_EINVAL, "null streamline arrays");
    HIP_TRY(hipSetDevice(V->device));
    hipStream_t s = (hipStream_t)stream;

    if (!inputs_on_device && n_count > 0 && (fib_ptr[0] != 0 || fib_ptr[n_count] != n_points))
        return fail(COMMIT_HIP_EINVAL, "fib_ptr must run from 0 to n_points (%lld)", (long long)n_points);
    int maxo = 0;
    for (int i = 0; i < 181 * 181; i++) {
        if (ptrHashTable[i] < 0) return fail(COMMIT_HIP_EINVAL, "negative orientation in the hash table");
        if (ptrHashTable[i] > maxo) maxo = ptrHashTable[i];
    }

    int rc;
    const float *d_pts;
    const int64_t *d_ptr;
    if (inputs_on_device) {
        d_pts = fibers;
        d_ptr = fib_ptr;
    } else {
        if (n_count > 0) {
            if ((rc = upload(V, B::PTS, fibers, (size_t)n_points * 3, s))) return rc;
            if ((rc = upload(V, B::FPTR, fib_ptr, (size_t)n_count + 1, s))) return rc;
        }
        d_pts = V->at<float>(B::PTS);
        d_ptr = V->at<int64_t>(B::FPTR);
    }
    // streamline offsets (MAX_FIB_LEN, trk2dictionary_c.cpp:195 and :878)
    if ((rc = V->need(B::STATUS, 16))) return rc;
    if (n_count > 0) {
        uint32_t *st = V->at<uint32_t>(B::STATUS);
        k_fill_u32<<<1, kVoxBlock, 0, s>>>(st, 1, 0u);
        k_fill_u32<<<1, kVoxBlock, 0, s>>>(st + 1, 1, 0xFFFFFFFFu);
        k_vox_check<<<grid_for((size_t)n_count + 1), kVoxBlock, 0, s>>>(d_ptr, n_count, n_points, st);
        HIP_TRY(hipGetLastError());
        uint32_t chk[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(chk, st, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (chk[0] & 4u) return fail(COMMIT_HIP_EINVAL, "fib_ptr must run from 0 to n_points (%lld)", (long long)n_points);
        if (chk[0] & 1u) return fail(COMMIT_HIP_EINVAL, "fib_ptr decreases");
        if (chk[0] & 2u)
            return fail(COMMIT_HIP_EUNSUPPORTED, "streamline %u has %d points or more (the reference's limit is %d)",
                        chk[1], kMaxFibLen, kMaxFibLen - 1);
    }
    if ((rc = upload(V, B::HASH, ptrHashTable, 181 * 181, s))) return rc;
    // replicas: rho, sin(angle/2), cos(angle/2), weight (:615-623)
    std::vector<double> rep(4 * (size_t)nReplicas);
    for (int k = 0; k < nReplicas; k++) {
        rep[k] = ptrBlurRho[k];
        rep[nReplicas + k] = std::sin(ptrBlurAngle[k] / 2.0);
        rep[2 * nReplicas + k] = std::cos(ptrBlurAngle[k] / 2.0);
        rep[3 * nReplicas + k] = ptrBlurWeights[k];
    }
    if ((rc = upload(V, B::REP, rep.data(), rep.size(), s))) return rc;
    const float *d_mask = nullptr, *d_iso = nullptr;
    if (ptrMASK && inputs_on_device) {
        d_mask = ptrMASK;
    } else if (ptrMASK) {
        if ((rc = upload(V, B::MASK, ptrMASK, nvox, s))) return rc;
        d_mask = V->at<float>(B::MASK);
    }
    if (ptrISO && inputs_on_device) {
        d_iso = ptrISO;
    } else if (ptrISO) {
        if ((rc = upload(V, B::ISO, ptrISO, nvox, s))) return rc;
        d_iso = V->at<float>(B::ISO);
    }
    const uint8_t *d_apply = nullptr;
    if (ptrBlurApplyTo && nReplicas > 1) {
        if ((rc = upload(V, B::APPLY, ptrBlurApplyTo, (size_t)n_count, s))) return rc;
        d_apply = V->at<uint8_t>(B::APPLY);
    }

    VoxArgs a;
    a.X = Nx; a.Y = Ny; a.Z = Nz;
    a.px = Px; a.py = Py; a.pz = Pz;
    a.rpx = 1.0 / (double)a.px; a.rpy = 1.0 / (double)a.py; a.rpz = 1.0 / (double)a.pz;
    a.sx = fiber_shiftX * Px; a.sy = fiber_shiftY * Py; a.sz = fiber_shiftZ * Pz;
    a.min_seg = min_seg_len; a.min_fib = min_fiber_len; a.max_fib = max_fiber_len;
    a.intersect = c > 0;
    a.n_rep = nReplicas;
    a.nvox = nvox;
    a.mask = d_mask;
    a.mbits = nullptr;
    if (d_mask) {
        if ((rc = V->need(B::MBITS, (nvox + 63) / 64 * 8))) return rc;
        uint64_t *bits = V->at<uint64_t>(B::MBITS);
        a.mbits = bits;
        k_vox_mask_bits<<<grid_for(nvox), kVoxBlock, 0, s>>>(d_mask, nvox, bits);
        HIP_TRY(hipGetLastError());
    }
    a.hash = V->at<short>(B::HASH);
    const double *d_rep = V->at<double>(B::REP);
    a.rho = d_rep;
    a.sinh = d_rep + nReplicas;
    a.cosh = d_rep + 2 * nReplicas;
    a.wt = d_rep + 3 * nReplicas;
    a.apply = d_apply;

    const size_t nF = (size_t)n_count;
    if ((rc = V->need(B::CNT, (nF + 1) * 4)) || (rc = V->need(B::FLEN, nF * 4)) ||
        (rc = V->need(B::FLENTOT, nF * 4)) || (rc = V->need(B::OFF, (nF + 1) * 8)) ||
        (rc = V->need(B::MCNT, (nF + 1) * 4)) ||
        (rc = V->need(B::MEND, nF * 8)) || (rc = V->need(B::KEPT, (nF + 1) * 4)) ||
        (rc = V->need(B::KIDX, (nF + 1) * 8)) || (rc = V->need(B::SPOS, (nF + 1) * 8)) ||
        (rc = V->need(B::NORM, nF * 4)))
        return rc;
    const bool multi = nReplicas > 1;
    if (multi && (rc = V->need(B::PBUF, nF * (size_t)nReplicas * sizeof(D3)))) return rc;
    uint32_t *d_status = V->at<uint32_t>(B::STATUS);
    uint32_t *d_cnt = V->at<uint32_t>(B::CNT);
    uint64_t *d_off = V->at<uint64_t>(B::OFF);
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_status, 1, 0u);
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_status + 2, 1, 0u);
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_cnt + nF, 1, 0u);
    D3 *d_P = multi ? V->at<D3>(B::PBUF) : nullptr;
    const unsigned gF = grid_for(nF ? nF : 1);
    if ((rc = V->need(B::OKEY, nF * 4)) || (rc = V->need(B::OIDX, nF * 4)) || (rc = V->need(B::OKEY2, nF * 4)) ||
        (rc = V->need(B::ORDER, nF * 4)) || (rc = V->need(B::END, nF * 8)) || (rc = V->need(B::POFF, (nF + 1) * 8)))
        return rc;
    const uint32_t *d_order = V->at<uint32_t>(B::ORDER);
    if (nF) {
        k_vox_order_keys<<<gF, kVoxBlock, 0, s>>>(d_ptr, n_count, V->at<uint32_t>(B::OKEY), V->at<uint32_t>(B::OIDX));
        HIP_TRY(hipGetLastError());
        if ((rc = sort_pairs_u32(V, (const uint32_t *)V->at<uint32_t>(B::OKEY), V->at<uint32_t>(B::OKEY2),
                                 (const uint32_t *)V->at<uint32_t>(B::OIDX), V->at<uint32_t>(B::ORDER), nF, 14, s)))
            return rc;
    }
    const int kbits = bit_length((uint64_t)(maxo + 1) * nvox - 1);
    const int vbits = bit_length(nvox - 1);
    // 32-bit keys when every key fits; float piece values when the one replica
    // has weight 1 (then w * len is the float length, exactly)
    const bool narrow = kbits <= 32;
    const bool fval = nReplicas == 1 && ptrBlurWeights[0] == 1.0;
    size_t n_ic = 0, n_kept = 0;
    IcArgs ia{a, d_pts, d_ptr, n_count, d_P, d_off, false, d_cnt, V->at<float>(B::FLEN), V->at<float>(B::FLENTOT),
              d_status, d_order, 0, kbits > 0 ? kbits : 1, vbits > 0 ? vbits : 1, min_fiber_len, max_fiber_len, s};
    auto ic = [&]() {
        return narrow ? (fval ? ic_segments<uint32_t, float>(V, ia, &n_ic, &n_kept)
                              : ic_segments<uint32_t, double>(V, ia, &n_ic, &n_kept))
                      : (fval ? ic_segments<uint64_t, float>(V, ia, &n_ic, &n_kept)
                              : ic_segments<uint64_t, double>(V, ia, &n_ic, &n_kept));
    };
    // One replica: a single walk into regions sized by k_vox_bound, when they
    // fit in half the free memory; a streamline that outgrows its region sends
    // the whole tractogram through the counting walk instead.
    bool done = false;
    if (!multi && nF) {
        k_vox_bound<<<gF, kVoxBlock, 0, s>>>(d_ptr, n_count, d_cnt);
        HIP_TRY(hipGetLastError());
        if ((rc = exscan_u32(V, d_cnt, d_off, nF + 1, s))) return rc;
        uint64_t slots = 0;
        HIP_TRY(hipMemcpyAsync(&slots, d_off + nF, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        const size_t per_slot = 2 * (narrow ? 4 : 8) + 2 * (fval ? 4 : 8) + 4;
        if (slots < 0xFFFFFFFFull && slots * per_slot <= free_b / 2) {
            ia.bounded = true;
            ia.np = (size_t)slots;
            rc = ic();
            if (rc == 0) done = true;
            else if (rc != kRetryExact) return rc;
            ia.bounded = false;
        }
    }
    if (!done) {
        k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_status, 1, 0u);
        k_fill_u32<<<1, kVoxBlock, 0, s>>>(d_status + 2, 1, 0u);
        if (nF) {
            if (multi)
                k_vox_walk<false, true, uint64_t, double><<<gF, kVoxBlock, 0, s>>>(
                    a, d_pts, d_ptr, n_count, d_P, nullptr, nullptr, nullptr, d_cnt, V->at<float>(B::FLEN),
                    V->at<float>(B::FLENTOT), d_status, d_order);
            else
                k_vox_walk<false, false, uint64_t, double><<<gF, kVoxBlock, 0, s>>>(
                    a, d_pts, d_ptr, n_count, nullptr, nullptr, nullptr, nullptr, d_cnt, V->at<float>(B::FLEN),
                    V->at<float>(B::FLENTOT), d_status, d_order);
            HIP_TRY(hipGetLastError());
        }
        if ((rc = exscan_u32(V, d_cnt, d_off, nF + 1, s))) return rc;
        uint64_t npieces = 0;
        uint32_t hstatus = 0;
        HIP_TRY(hipMemcpyAsync(&npieces, d_off + nF, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&hstatus, d_status, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (hstatus) return fail(COMMIT_HIP_EUNSUPPORTED, "a polyline step crossed more voxels than its cap allows");
        if (npieces >= 0xFFFFFFFFull)
            return fail(COMMIT_HIP_EUNSUPPORTED, "%llu pieces exceed the 32-bit sort size", (unsigned long long)npieces);
        V->n_pieces = (int64_t)npieces;
        ia.np = (size_t)npieces;
        if ((rc = ic())) return rc;
    }
    // TDI and the voxels holding IC segments
    k_fill_f32<<<grid_for(nvox), kVoxBlock, 0, s>>>(V->at<float>(B::TDI), nvox, 0.0f);
    k_fill_u8<<<grid_for(nvox), kVoxBlock, 0, s>>>(V->at<uint8_t>(B::HAS), nvox, 0);
    if ((rc = sort_pairs_u32(V, V->at<uint32_t>(B::IC_V), V->at<uint32_t>(B::TDI_V), V->at<float>(B::IC_LEN),
                             V->at<float>(B::TDI_L), n_ic, vbits > 0 ? vbits : 1, s)))
        return rc;
    if (n_ic) {
        // run bounds per voxel into the (free) EC count and ISO flag buffers
        if ((rc = V->need(B::ECNT, (nvox + 1) * 4)) || (rc = V->need(B::IFLAG, (nvox + 1) * 4))) return rc;
        uint32_t *rs = V->at<uint32_t>(B::ECNT), *re = V->at<uint32_t>(B::IFLAG);
        k_fill_u32<<<grid_for(nvox), kVoxBlock, 0, s>>>(rs, nvox, 0u);
        k_fill_u32<<<grid_for(nvox), kVoxBlock, 0, s>>>(re, nvox, 0u);
        k_vox_runs<<<grid_for(n_ic), kVoxBlock, 0, s>>>(V->at<uint32_t>(B::TDI_V), n_ic, rs, re);
        const unsigned gV = (unsigned)((nvox + kVoxBlock / kWave - 1) / (kVoxBlock / kWave));
        k_vox_tdi<<<gV, kVoxBlock, 0, s>>>(rs, re, V->at<float>(B::TDI_L), nvox, V->at<float>(B::TDI),
                                           V->at<uint8_t>(B::HAS));
        HIP_TRY(hipGetLastError());
    }
    // EC
    size_t n_ec = 0;
    if (ptrPEAKS) {
        if (!inputs_on_device && (rc = upload(V, B::PEAKS, ptrPEAKS, nvox * 3 * (size_t)Np, s))) return rc;
        if ((rc = V->need(B::ECNT, (nvox + 1) * 4)) || (rc = V->need(B::EOFF, (nvox + 1) * 8))) return rc;
        EcArgs e;
        e.X = Nx; e.Y = Ny; e.Z = Nz; e.Np = Np;
        e.vf_thr = vf_THR;
        e.fx = ECix; e.fy = ECiy; e.fz = ECiz;
        for (int i = 0; i < 9; i++) e.A[i] = ptrPeaksAffine[i];
        e.peaks = inputs_on_device ? ptrPEAKS : V->at<float>(B::PEAKS);
        e.hash = a.hash;
        k_fill_u32<<<1, kVoxBlock, 0, s>>>(V->at<uint32_t>(B::ECNT) + nvox, 1, 0u);
        k_vox_ec<false><<<grid_for(nvox), kVoxBlock, 0, s>>>(e, nvox, V->at<uint8_t>(B::HAS), V->at<uint32_t>(B::ECNT),
                                                             nullptr, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        if ((rc = exscan_u32(V, V->at<uint32_t>(B::ECNT), V->at<uint64_t>(B::EOFF), nvox + 1, s))) return rc;
        uint64_t t = 0;
        HIP_TRY(hipMemcpyAsync(&t, V->at<uint64_t>(B::EOFF) + nvox, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        n_ec = (size_t)t;
        if ((rc = V->need(B::EC_V, n_ec * 4)) || (rc = V->need(B::EC_O, n_ec * 2))) return rc;
        k_vox_ec<true><<<grid_for(nvox), kVoxBlock, 0, s>>>(e, nvox, V->at<uint8_t>(B::HAS), nullptr,
                                                            V->at<uint64_t>(B::EOFF), V->at<uint32_t>(B::EC_V),
                                                            V->at<uint16_t>(B::EC_O));
        HIP_TRY(hipGetLastError());
    }
    // ISO
    if ((rc = V->need(B::IFLAG, (nvox + 1) * 4)) || (rc = V->need(B::IOFF, (nvox + 1) * 8))) return rc;
    k_fill_u32<<<1, kVoxBlock, 0, s>>>(V->at<uint32_t>(B::IFLAG) + nvox, 1, 0u);
    k_vox_iso<false><<<grid_for(nvox), kVoxBlock, 0, s>>>(Nx, Ny, Nz, nvox, d_iso, d_mask, V->at<uint8_t>(B::HAS),
                                                          V->at<uint32_t>(B::IFLAG), nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    if ((rc = exscan_u32(V, V->at<uint32_t>(B::IFLAG), V->at<uint64_t>(B::IOFF), nvox + 1, s))) return rc;
    uint64_t ti = 0;
    HIP_TRY(hipMemcpyAsync(&ti, V->at<uint64_t>(B::IOFF) + nvox, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const size_t n_iso = (size_t)ti;
    if ((rc = V->need(B::ISO_V, n_iso * 4))) return rc;
    k_vox_iso<true><<<grid_for(nvox), kVoxBlock, 0, s>>>(Nx, Ny, Nz, nvox, d_iso, d_mask, V->at<uint8_t>(B::HAS),
                                                         V->at<uint32_t>(B::IFLAG), V->at<uint64_t>(B::IOFF),
                                                         V->at<uint32_t>(B::ISO_V));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    V->n_ic = (int64_t)n_ic;
    V->n_kept = (int64_t)n_kept;
    V->n_ec = (int64_t)n_ec;
    V->n_iso = (int64_t)n_iso;
    V->n_count = n_count;
    V->nvox = nvox;
    V->valid = true;
    return 0;
}

extern "C" int commit_hip_vox_sizes(const commit_hip_vox *V, int64_t *n_ic, int64_t *n_kept, int64_t *n_ec,
                                    int64_t *n_iso, int64_t *n_pieces)
{
    if (!V) return fail(COMMIT_HIP_EINVAL, "null voxeliser handle");
    if (!V->valid) return fail(COMMIT_HIP_EINVAL, "no successful commit_hip_trk2dictionary on this handle");
    if (n_ic) *n_ic = V->n_ic;
    if (n_kept) *n_kept = V->n_kept;
    if (n_ec) *n_ec = V->n_ec;
    if (n_iso) *n_iso = V->n_iso;
    if (n_pieces) *n_pieces = V->n_pieces;
    return 0;
}

extern "C" int commit_hip_vox_fetch(commit_hip_vox *V, uint32_t *IC_f, uint32_t *IC_v, uint16_t *IC_o,
                                    float *IC_len, float *TRK_norm, float *TRK_len, float *TRK_lenTot,
                                    uint8_t *TRK_kept, uint32_t *EC_v, uint16_t *EC_o, uint32_t *ISO_v, float *tdi)
{
    using B = commit_hip_vox;
    if (!V) return fail(COMMIT_HIP_EINVAL, "null voxeliser handle");
    if (!V->valid) return fail(COMMIT_HIP_EINVAL, "no successful commit_hip_trk2dictionary on this handle");
    HIP_TRY(hipSetDevice(V->device));
    struct Item {
        void *dst;
        int buf;
        size_t bytes;
    } items[] = {
        {IC_f, B::IC_F, (size_t)V->n_ic * 4},        {IC_v, B::IC_V, (size_t)V->n_ic * 4},
        {IC_o, B::IC_O, (size_t)V->n_ic * 2},        {IC_len, B::IC_LEN, (size_t)V->n_ic * 4},
        {TRK_norm, B::TRK_NORM, (size_t)V->n_kept * 4}, {TRK_len, B::TRK_LEN, (size_t)V->n_kept * 4},
        {TRK_lenTot, B::TRK_LENTOT, (size_t)V->n_kept * 4}, {TRK_kept, B::TRK_KEPT, (size_t)V->n_count},
        {EC_v, B::EC_V, (size_t)V->n_ec * 4},        {EC_o, B::EC_O, (size_t)V->n_ec * 2},
        {ISO_v, B::ISO_V, (size_t)V->n_iso * 4},     {tdi, B::TDI, (size_t)V->nvox * 4},
    };
    for (const Item &it : items)
        if (it.dst && it.bytes) HIP_TRY(hipMemcpy(it.dst, V->buf[it.buf].p, it.bytes, hipMemcpyDeviceToHost));
    return 0;
}
