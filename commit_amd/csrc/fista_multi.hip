// Batched FISTA: S x R independent regularised fits advanced together.  R
// columns of X (a lambda sweep or bootstrap signals) times S problems whose
// unknowns and rows are disjoint ranges of one block-diagonal plan (several
// tractograms of uneven size in one launch).  Every decision of the
// reference's loop (solvers.py:244-391) is taken per problem ON THE DEVICE,
// so that one trial step is a fixed sequence of kernels with no host round
// trip and can be captured in a HIP graph and replayed.
//
// Vectors are float32 [rows][R] (the RHS index fastest, as the multi-RHS
// products expect); every sum and every scalar of the control flow is FP64.
// Reductions are deterministic: rows are cut into pieces that never cross a
// problem boundary (host-built, fixed), each piece is reduced in a fixed
// order, and the pieces of a problem are added in piece order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/commit_hip.h"
#include "common.h"

using commit_internal::fail;

namespace {

constexpr double kEps = 2.220446049250313e-16;   // numpy float64 eps (solvers.py:10)
constexpr int kMaxR = 16;
constexpr int kNSum = 8;                          // sums per problem
constexpr int kMaxProblems = 1024;                // S * R

using Col = commit_hip_fista_col;

struct MProx {
    int64_t start[3], end[3];
    int reg[3], nn[3], weighted[3];
    const double *w;
    int64_t ngroups;
    const int32_t *gptr, *gidx, *gseg;
    const double *gw;
};

// Problem of every unknown (segx[i]); S problems, R columns.
struct Seg {
    const int32_t *segx;
    int S, R;
};

// t = i R + r (0 <= r < R) for the elementwise kernels: shifts when R is a
// power of two (sh = log2 R, the usual 16 columns), a 32-bit division while
// t fits, a 64-bit one otherwise.  A 64-bit division by a run-time R per
// element made these kernels several times slower than their memory traffic.
__device__ __forceinline__ int col_shift(int R) { return (R & (R - 1)) == 0 ? __builtin_ctz((unsigned)R) : -1; }

__device__ __forceinline__ void split_rc(int64_t t, int R, int sh, int64_t &i, int &r)
{
    if (sh >= 0) {
        i = t >> sh;
        r = (int)(t & (int64_t)(R - 1));
    } else if (t < 0x100000000ll) {
        const uint32_t q = (uint32_t)t / (uint32_t)R;
        i = (int64_t)q;
        r = (int)((uint32_t)t - q * (uint32_t)R);
    } else {
        i = t / R;
        r = (int)(t - i * R);
    }
}

__device__ __forceinline__ double soft(double v, double t)
{
    if (v > t) return v - t;                      // proximals.pyx:25-56
    if (v < -t) return v + t;
    return 0.0;
}

// The fields of the fit records the elementwise kernels read, staged in LDS
// by each work-group (S * R records, at most kMaxProblems): an element then
// waits on one global load (its problem) and one LDS read, not on a chain of
// global loads into the 152-byte records.
struct ProxRec {
    double step, lam[3];
    int32_t done, pad_;
};
struct UpdRec {
    double coef;
    int32_t advance, pad_;
};

// x = prox(xhat - step*grad, step) elementwise for the active problems
// (soft-thresholding with the problem's lambdas, non-negativity where no
// group penalty follows).  Finished problems keep x.
__global__ __launch_bounds__(kBlock) void k_mprox_elem(MProx P, Seg G, int64_t n, const float *__restrict__ xhat,
                                                       const float *__restrict__ grad, float *__restrict__ x,
                                                       const Col *__restrict__ col)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ProxRec *cs = reinterpret_cast<ProxRec *>(smem);
    const int R = G.R, sh = col_shift(R);
    for (int q = threadIdx.x; q < G.S * R; q += blockDim.x) {
        const Col &c = col[q];
        cs[q] = ProxRec{c.step, {c.lam[0], c.lam[1], c.lam[2]}, c.done, 0};
    }
    __syncthreads();
    const int64_t total = n * R;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i;
        int r;
        split_rc(t, R, sh, i, r);
        const ProxRec &c = cs[G.segx[i] * R + r];
        if (c.done) continue;
        const double step = c.step;
        double v = (double)xhat[t] - step * (double)grad[t];             // (solvers.py:299)
        for (int k = 0; k < 3; k++) {
            if (i < P.start[k] || i >= P.end[k]) continue;
            const int rg = P.reg[k];
            if (rg == COMMIT_HIP_REG_LASSO || rg == COMMIT_HIP_REG_SPARSE_GROUP_LASSO) {
                const double thr = c.lam[k] * step;
                v = soft(v, P.weighted[k] ? P.w[i] * thr : thr);
            }
            if (P.nn[k] && rg != COMMIT_HIP_REG_GROUP_LASSO && rg != COMMIT_HIP_REG_SPARSE_GROUP_LASSO)
                v = (v <= 0.0) ? 0.0 : v;                                  // (proximals.pyx:13-22)
            break;
        }
        x[t] = (float)v;
    }
}

// Group shrinkage (proximals.pyx:103-134) and group norms (proximals.pyx:80-99):
// one block of kGroupBlock threads per group; thread t takes column t % 16 of
// the members t/16, t/16 + 64, ... (each member's row of R floats is read
// coalesced, four members per round in flight), the 64 partial sums of a
// column are added in order in LDS.  Groups are disjoint (checked on the
// host), each inside one problem (gseg).
constexpr int kGroupBlock = 1024;
constexpr int kGroupSubs = kGroupBlock / 16;

// sum over the group's members of x[member][r]^2 for r = threadIdx.x % 16,
// returned to every thread of that column
__device__ __forceinline__ double group_sumsq(const MProx &P, int64_t g, int R, const float *__restrict__ x,
                                              double (*sh)[16])
{
    const int r = threadIdx.x & 15, sub = threadIdx.x >> 4;
    const int rr = r < R ? r : 0;
    const int b = P.gptr[g], e = P.gptr[g + 1];
    double s = 0.0;
    int k = b + sub;
    for (; k + 3 * kGroupSubs < e; k += 4 * kGroupSubs) {
        int id[4];
#pragma unroll
        for (int j = 0; j < 4; j++) id[j] = P.gidx[k + j * kGroupSubs];
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = x[(size_t)id[j] * R + rr];
#pragma unroll
        for (int j = 0; j < 4; j++) s = fma((double)v[j], (double)v[j], s);
    }
    for (; k < e; k += kGroupSubs) {
        const double v = (double)x[(size_t)P.gidx[k] * R + rr];
        s = fma(v, v, s);
    }
    sh[sub][r] = s;
    __syncthreads();
    double t = 0.0;
    for (int j = 0; j < kGroupSubs; j++) t += sh[j][r];
    return t;
}

__global__ __launch_bounds__(kGroupBlock) void k_mprox_group(MProx P, Seg G, float *__restrict__ x,
                                                             const Col *__restrict__ col)
{
    __shared__ double sh[kGroupSubs][16];
    const int64_t g = blockIdx.x;
    const int R = G.R;
    const int r = threadIdx.x & 15, sub = threadIdx.x >> 4;
    const double s = group_sumsq(P, g, R, x, sh);
    if (r >= R) return;
    const Col *c = col + (size_t)P.gseg[g] * R + r;
    if (c->done) return;
    const double lam = c->lam_group * c->step;
    if (lam == 0.0) return;                                               // (proximals.pyx:116)
    const double nrm = sqrt(s);
    const double wl = P.gw[g] * lam;
    const double f = (nrm <= wl) ? 0.0 : (nrm - wl) / nrm;
    for (int k = P.gptr[g] + sub; k < P.gptr[g + 1]; k += kGroupSubs) {
        float *p = x + (size_t)P.gidx[k] * R + r;
        *p = (f == 0.0) ? 0.0f : (float)((double)*p * f);
    }
}

__global__ __launch_bounds__(kBlock) void k_mnn_ic(MProx P, Seg G, float *__restrict__ x, const Col *__restrict__ col)
{
    const int R = G.R, sh = col_shift(R);
    const int64_t b = P.start[0], total = (P.end[0] - P.start[0]) * R;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i;
        int r;
        split_rc(t, R, sh, i, r);
        i += b;
        if (col[G.segx[i] * R + r].done) continue;
        float *p = x + b * R + t;
        *p = (*p <= 0.0f) ? 0.0f : *p;
    }
}

// Sums over the unknowns, stage 1: block (p, j) reduces part j of kSumSplit
// of piece p (rows [pb, pe) of one problem) for all R columns; thread t takes
// column t % R and rows b + t/R, b + t/R + L, ... (L = 256 / R, four rows per
// round in flight); the L partials of a column are then added in order.
//   s0 = sum (x - xhat) grad   s1 = sum (x - xhat)^2   s2 = sum (x - prev)^2
//   s3 = sum x^2               s4..s6 = l1 terms of the three compartments
constexpr int kSumSplit = 8;

struct MSum {
    const float *x, *xhat, *grad, *prev;
    const double *w;
    int64_t l1b[3], l1e[3];
    int l1w[3];
};

__device__ __forceinline__ void msum_row(const MSum &a, int64_t i, float xf, float hf, float gf, float pf,
                                         double (&s)[7])
{
    const double xi = (double)xf;
    if (a.xhat) {
        const double d = xi - (double)hf;
        s[0] = fma(d, (double)gf, s[0]);
        s[1] = fma(d, d, s[1]);
    }
    if (a.prev) {
        const double d = xi - (double)pf;
        s[2] = fma(d, d, s[2]);
    }
    s[3] = fma(xi, xi, s[3]);
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (i >= a.l1b[k] && i < a.l1e[k]) s[4 + k] += fabs(a.l1w[k] ? a.w[i] * xi : xi);
}

__global__ __launch_bounds__(kBlock) void k_msums_stage1(MSum a, int R, const int64_t *__restrict__ pieces,
                                                         double *__restrict__ part)
{
    __shared__ double sh[kBlock][7];
    const int L = kBlock / R;
    const int tid = threadIdx.x;
    const int r = tid % R, lane0 = tid / R;
    const int64_t pb = pieces[2 * blockIdx.x], pe = pieces[2 * blockIdx.x + 1];
    const int64_t per = (pe - pb + kSumSplit - 1) / kSumSplit;
    const int64_t b = pb + per * blockIdx.y, e = b + per < pe ? b + per : pe;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    if (lane0 < L) {
        int64_t i = b + lane0;
        for (; i + 3 * L < e; i += 4 * L) {
            float xv[4], hv[4] = {0, 0, 0, 0}, gv[4] = {0, 0, 0, 0}, pv[4] = {0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const size_t t = (size_t)(i + j * L) * R + r;
                xv[j] = a.x[t];
                if (a.xhat) hv[j] = a.xhat[t], gv[j] = a.grad[t];
                if (a.prev) pv[j] = a.prev[t];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) msum_row(a, i + j * L, xv[j], hv[j], gv[j], pv[j], s);
        }
        for (; i < e; i += L) {
            const size_t t = (size_t)i * R + r;
            msum_row(a, i, a.x[t], a.xhat ? a.xhat[t] : 0.0f, a.xhat ? a.grad[t] : 0.0f,
                     a.prev ? a.prev[t] : 0.0f, s);
        }
    }
#pragma unroll
    for (int k = 0; k < 7; k++) sh[tid][k] = s[k];
    __syncthreads();
    if (tid < R * 7) {
        const int rr = tid % R, k = tid / R;
        double v = 0.0;
        for (int l = 0; l < L; l++) v += sh[l * R + rr][k];
        part[(((size_t)blockIdx.x * kSumSplit + blockIdx.y) * R + rr) * 7 + k] = v;
    }
}

// gn[g*R + r] = group_w[g] * ||x_g[:, r]||   (proximals.pyx:80-99)
__global__ __launch_bounds__(kGroupBlock) void k_mgroup_norm(MProx P, int R, const float *__restrict__ x,
                                                             double *__restrict__ gn)
{
    __shared__ double sh[kGroupSubs][16];
    const int64_t g = blockIdx.x;
    const double s = group_sumsq(P, g, R, x, sh);
    if (threadIdx.x < 16 && (int)threadIdx.x < R) gn[(size_t)g * R + threadIdx.x] = P.gw[g] * sqrt(s);
}

// Stage 2: out[(s*R + r)*8 + k], one wavefront per (problem, column, term):
// lane l adds the parts l, l + 64, ... of problem s (pieces x splits, k < 7)
// or its groups (k = 7) in order, then the 64 lane sums are added by a fixed
// butterfly.
__global__ __launch_bounds__(kBlock) void k_msums_stage2(const double *__restrict__ part,
                                                         const int32_t *__restrict__ pseg, int npieces,
                                                         const double *__restrict__ gn,
                                                         const int32_t *__restrict__ gseg, int64_t ngroups, int S,
                                                         int R, double *__restrict__ out)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int t = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (t >= S * R * kNSum) return;
    const int k = t % kNSum, sr = t / kNSum, r = sr % R, s = sr / R;
    double v = 0.0;
    if (k < 7) {
        for (int q = lane; q < npieces * kSumSplit; q += kWave)
            if (pseg[q / kSumSplit] == s) v += part[((size_t)q * R + r) * 7 + k];
    } else {
        for (int64_t g = lane; g < ngroups; g += kWave)
            if (gseg[g] == s) v += gn[(size_t)g * R + r];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) out[t] = v;
}

// res = sqrtw .* (AX - y) over the row pieces (y: [m] shared by the columns
// when ystride == 0, else [m][R]); part[p][r] = sum of res^2 over piece p.
// res may be null when only the norms are needed.
__global__ __launch_bounds__(kBlock) void k_mresid(const float *__restrict__ AX, const float *__restrict__ y,
                                                   int ystride, const float *__restrict__ sqrtw, int R,
                                                   const int64_t *__restrict__ pieces, float *__restrict__ res,
                                                   double *__restrict__ part)
{
    __shared__ double sh[kBlock];
    const int L = kBlock / R;
    const int tid = threadIdx.x;
    const int r = tid % R, lane0 = tid / R;
    const int64_t b = pieces[2 * blockIdx.x], e = pieces[2 * blockIdx.x + 1];
    double s = 0.0;
    if (lane0 < L) {
        for (int64_t i = b + lane0; i < e; i += L) {
            const size_t t = (size_t)i * R + r;
            double v = (double)AX[t] - (double)(ystride ? y[t] : y[i]);          // (solvers.py:308)
            if (sqrtw) v *= (double)sqrtw[i];                                      // (solvers.py:306)
            if (res) res[t] = (float)v;
            s = fma(v, v, s);
        }
    }
    sh[tid] = s;
    __syncthreads();
    if (tid < R) {
        double v = 0.0;
        for (int l = 0; l < L; l++) v += sh[l * R + tid];
        part[(size_t)blockIdx.x * R + tid] = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_mresid_stage2(const double *__restrict__ part,
                                                          const int32_t *__restrict__ pseg, int npieces, int S,
                                                          int R, double *__restrict__ out)
{
    for (int t = threadIdx.x; t < S * R; t += kBlock) {
        const int r = t % R, s = t / R;
        double v = 0.0;
        for (int p = 0; p < npieces; p++)
            if (pseg[p] == s) v += part[(size_t)p * R + r];
        out[t] = v;
    }
}

// Omega of one problem from its sums (solvers.py:200 and the compartment omegas)
__device__ double omega_of(const Col &c, const double *s, const int *reg)
{
    double om = 0.0;
    if (reg[0] == COMMIT_HIP_REG_LASSO) om += c.lam[0] * s[4];
    else if (reg[0] == COMMIT_HIP_REG_GROUP_LASSO) om += (c.lam_group != 0.0) ? c.lam_group * s[7] : 0.0;
    else if (reg[0] == COMMIT_HIP_REG_SPARSE_GROUP_LASSO) {
        if (c.lam[0] != 0.0 || c.lam_group != 0.0) om += c.lam[0] * s[4] + c.lam_group * s[7];
    }
    if (reg[1] == COMMIT_HIP_REG_LASSO) om += c.lam[1] * s[5];
    if (reg[2] == COMMIT_HIP_REG_LASSO) om += c.lam[2] * s[6];
    return om;
}

struct Ctl {
    int reg[3];
    double tol_fun, tol_x, beta;
    int max_iter;
};

// Decision after a trial step, per problem (solvers.py:299-357).
//   sums: the trial x against xhat, grad and prev_x;  rsq: ||sqrtw (A x - y)||^2.
// Sets advance = 1 when x is accepted and the fit continues.
__global__ __launch_bounds__(kBlock) void k_mcontrol(Ctl C, int nprob, const double *__restrict__ sums,
                                                     const double *__restrict__ rsq, Col *__restrict__ col)
{
    for (int q = threadIdx.x; q < nprob; q += kBlock) {
        Col c = col[q];
        c.advance = 0;
        if (c.done) {
            col[q] = c;
            continue;
        }
        const double *s = sums + (size_t)q * kNSum;
        const double reg_x = omega_of(c, s, C.reg);
        const double qv = c.qfval + s[0] + 0.5 / c.step * s[1] + reg_x;  // (solvers.py:311)
        const double rn = sqrt(rsq[q]);
        const double curr = 0.5 * rn * rn + reg_x;
        c.trials += 1;
        if (curr > qv) {                                                  // backtracking (solvers.py:315-332)
            c.step *= C.beta;
            col[q] = c;
            continue;
        }
        c.regterm = reg_x;
        c.residual = 0.5 * rn * rn;
        c.cost = curr;
        c.abs_cost = fabs(curr - c.prev_obj);
        c.rel_cost = c.abs_cost / curr;
        c.abs_x = sqrt(s[2]);
        c.rel_x = c.abs_x / (sqrt(s[3]) + kEps);
        int crit = 0;                                                     // (solvers.py:342-356)
        if (c.abs_cost < kEps) crit = 1;
        else if (c.rel_cost < C.tol_fun) crit = 2;
        else if (c.abs_x < kEps) crit = 3;
        else if (c.rel_x < C.tol_x) crit = 4;
        else if (c.iterations >= C.max_iter) crit = 5;
        if (crit) {
            c.done = 1;
            c.criterion = crit;
            col[q] = c;
            continue;
        }
        const double t = 0.5 * (1.0 + sqrt(1.0 + 4.0 * c.told * c.told));   // (solvers.py:359-375)
        c.coef = (c.told - 1.0) / t;
        c.told = t;
        c.prev_obj = curr;
        c.iterations += 1;
        c.advance = 1;
        col[q] = c;
    }
}

// Accepted problems: xhat = x + coef (x - prev_x); prev_x = x.
__global__ __launch_bounds__(kBlock) void k_mupdate(Seg G, int64_t n, const float *__restrict__ x,
                                                    float *__restrict__ xhat, float *__restrict__ prev,
                                                    const Col *__restrict__ col)
{
    extern __shared__ __align__(16) unsigned char smem[];
    UpdRec *cs = reinterpret_cast<UpdRec *>(smem);
    const int R = G.R, sh = col_shift(R);
    for (int q = threadIdx.x; q < G.S * R; q += blockDim.x) cs[q] = UpdRec{col[q].coef, col[q].advance, 0};
    __syncthreads();
    const int64_t total = n * R;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i;
        int r;
        split_rc(t, R, sh, i, r);
        const UpdRec &c = cs[G.segx[i] * R + r];
        if (!c.advance) continue;
        const double xv = (double)x[t];
        xhat[t] = (float)(xv + c.coef * (xv - (double)prev[t]));
        prev[t] = x[t];
    }
}

// After the residual at the new xhat: qfval = 0.5 ||res||^2 for the problems that advanced.
__global__ __launch_bounds__(kBlock) void k_mqf(int nprob, const double *__restrict__ rsq, Col *__restrict__ col)
{
    for (int q = threadIdx.x; q < nprob; q += kBlock) {
        if (!col[q].advance) continue;
        const double rn = sqrt(rsq[q]);
        col[q].qfval = 0.5 * rn * rn;
    }
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

int check_layout(const commit_hip_multi_layout *L)
{
    if (!L) return fail(COMMIT_HIP_EINVAL, "NULL layout");
    if (L->nrhs < 1 || L->nrhs > kMaxR) return fail(COMMIT_HIP_EUNSUPPORTED, "batched fits take 1..%d columns", kMaxR);
    if (L->nprob < 1 || L->nprob * L->nrhs > kMaxProblems)
        return fail(COMMIT_HIP_EUNSUPPORTED, "problems x columns must be 1..%d", kMaxProblems);
    if (L->n < 0 || L->m < 0 || (L->n > 0 && !L->segx)) return fail(COMMIT_HIP_EINVAL, "bad layout");
    if (L->nxpieces < 1 || !L->xpieces || !L->xpseg || L->nypieces < 1 || !L->ypieces || !L->ypseg)
        return fail(COMMIT_HIP_EINVAL, "layout pieces missing");
    return 0;
}

Seg seg_of(const commit_hip_multi_layout *L)
{
    Seg G{};
    G.segx = L->segx, G.S = L->nprob, G.R = L->nrhs;
    return G;
}

int make_mprox(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, MProx &a)
{
    const int64_t n = L->n;
    for (int c = 0; c < 3; c++) {
        a.start[c] = p->start[c];
        a.end[c] = p->start[c] + p->size[c];
        if (a.start[c] < 0 || a.end[c] > n || p->size[c] < 0)
            return fail(COMMIT_HIP_EINVAL, "compartment %d range outside x", c);
        a.reg[c] = p->reg[c];
        if (c > 0 && (a.reg[c] == COMMIT_HIP_REG_GROUP_LASSO || a.reg[c] == COMMIT_HIP_REG_SPARSE_GROUP_LASSO))
            return fail(COMMIT_HIP_EUNSUPPORTED, "group penalties exist only for the IC compartment");
        a.nn[c] = p->nn[c];
        a.weighted[c] = p->weighted[c];
        if (a.weighted[c] && !p->w) return fail(COMMIT_HIP_EINVAL, "weighted lasso without weights");
    }
    a.w = p->w;
    const bool grp = a.reg[0] == COMMIT_HIP_REG_GROUP_LASSO || a.reg[0] == COMMIT_HIP_REG_SPARSE_GROUP_LASSO;
    a.ngroups = grp ? p->ngroups : 0;
    a.gptr = p->group_ptr;
    a.gidx = p->group_idx;
    a.gw = p->group_w;
    a.gseg = L->gseg;
    if (a.ngroups > 0 && (!a.gptr || !a.gidx || !a.gw || !a.gseg))
        return fail(COMMIT_HIP_EINVAL, "group arrays missing");
    if (a.ngroups > 0 && p->groups_overlap)
        return fail(COMMIT_HIP_EUNSUPPORTED, "batched fits need non-overlapping groups");
    return 0;
}

}  // namespace

extern "C" {

int64_t commit_hip_fista_multi_work(const commit_hip_multi_layout *L, int64_t ngroups)
{
    if (!L) return -1;
    const int64_t R = L->nrhs;
    return L->nxpieces * kSumSplit * R * 7 + ngroups * R;
}

int commit_hip_prox_multi(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, const float *xhat,
                          const float *grad, float *x, const commit_hip_fista_col *col, void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!p || !xhat || !grad || !x || !col) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    MProx a{};
    if ((rc = make_mprox(p, L, a))) return rc;
    const Seg G = seg_of(L);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (L->n == 0) return 0;
    hipLaunchKernelGGL(k_mprox_elem, dim3(elem_grid(L->n * L->nrhs)), dim3(kBlock),
                       (size_t)G.S * G.R * sizeof(ProxRec), st, a, G, L->n, xhat, grad, x, col);
    HIP_TRY(hipGetLastError());
    const int r0 = a.reg[0];
    if (r0 == COMMIT_HIP_REG_GROUP_LASSO || r0 == COMMIT_HIP_REG_SPARSE_GROUP_LASSO) {
        if (a.ngroups > 0)
            hipLaunchKernelGGL(k_mprox_group, dim3((unsigned)a.ngroups), dim3(kGroupBlock), 0, st, a, G, x, col);
        if (a.nn[0] && a.end[0] > a.start[0])
            hipLaunchKernelGGL(k_mnn_ic, dim3(elem_grid((a.end[0] - a.start[0]) * L->nrhs)), dim3(kBlock), 0, st, a,
                               G, x, col);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int commit_hip_fista_multi_sums(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, const float *x,
                                const float *xhat, const float *grad, const float *prev, const int64_t *l1_range,
                                const int *l1_weighted, double *out, double *work, void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!p || !x || !out || !work) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if ((xhat == nullptr) != (grad == nullptr)) return fail(COMMIT_HIP_EINVAL, "xhat and grad go together");
    MProx a{};
    if ((rc = make_mprox(p, L, a))) return rc;
    MSum s{};
    s.x = x, s.xhat = xhat, s.grad = grad, s.prev = prev, s.w = p->w;
    for (int t = 0; t < 3; t++) {
        s.l1b[t] = l1_range ? l1_range[2 * t] : 0;
        s.l1e[t] = l1_range ? l1_range[2 * t + 1] : 0;
        s.l1w[t] = l1_weighted ? l1_weighted[t] : 0;
        if (s.l1w[t] && !p->w) return fail(COMMIT_HIP_EINVAL, "weighted l1 term without weights");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int R = L->nrhs;
    double *part = work;
    double *gn = part + (size_t)L->nxpieces * kSumSplit * R * 7;
    hipLaunchKernelGGL(k_msums_stage1, dim3((unsigned)L->nxpieces, kSumSplit), dim3(kBlock), 0, st, s, R, L->xpieces,
                       part);
    if (a.ngroups > 0)
        hipLaunchKernelGGL(k_mgroup_norm, dim3((unsigned)a.ngroups), dim3(kGroupBlock), 0, st, a, R, x, gn);
    hipLaunchKernelGGL(k_msums_stage2, dim3(wave_grid((int64_t)L->nprob * R * kNSum)), dim3(kBlock), 0, st, part,
                       L->xpseg, (int)L->nxpieces, gn, L->gseg, a.ngroups, L->nprob, R, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

int commit_hip_multi_residual(const commit_hip_multi_layout *L, const float *AX, const float *y, int y_per_column,
                              const float *sqrtw, float *res, double *rsq, double *work, void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!AX || !y || !rsq || !work) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int R = L->nrhs;
    double *part = work;
    hipLaunchKernelGGL(k_mresid, dim3((unsigned)L->nypieces), dim3(kBlock), 0, st, AX, y, y_per_column ? 1 : 0, sqrtw,
                       R, L->ypieces, res, part);
    hipLaunchKernelGGL(k_mresid_stage2, dim3(1), dim3(kBlock), 0, st, part, L->ypseg, (int)L->nypieces, L->nprob, R,
                       rsq);
    HIP_TRY(hipGetLastError());
    return 0;
}

int commit_hip_fista_multi_control(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, double tol_fun,
                                   double tol_x, int max_iter, const double *sums, const double *rsq,
                                   commit_hip_fista_col *col, void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!p || !sums || !rsq || !col) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    Ctl c{};
    for (int k = 0; k < 3; k++) c.reg[k] = p->reg[k];
    c.tol_fun = tol_fun, c.tol_x = tol_x, c.beta = 0.9, c.max_iter = max_iter;
    hipLaunchKernelGGL(k_mcontrol, dim3(1), dim3(kBlock), 0, static_cast<hipStream_t>(stream), c,
                       L->nprob * L->nrhs, sums, rsq, col);
    HIP_TRY(hipGetLastError());
    return 0;
}

int commit_hip_fista_multi_update(const commit_hip_multi_layout *L, const float *x, float *xhat, float *prev,
                                  const commit_hip_fista_col *col, void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!x || !xhat || !prev || !col) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    if (L->n == 0) return 0;
    const Seg G = seg_of(L);
    hipLaunchKernelGGL(k_mupdate, dim3(elem_grid(L->n * L->nrhs)), dim3(kBlock), (size_t)G.S * G.R * sizeof(UpdRec),
                       static_cast<hipStream_t>(stream), G, L->n, x, xhat, prev, col);
    HIP_TRY(hipGetLastError());
    return 0;
}

int commit_hip_fista_multi_qf(const commit_hip_multi_layout *L, const double *rsq, commit_hip_fista_col *col,
                              void *stream)
{
    int rc = check_layout(L);
    if (rc) return rc;
    if (!rsq || !col) return fail(COMMIT_HIP_EINVAL, "NULL argument");
    hipLaunchKernelGGL(k_mqf, dim3(1), dim3(kBlock), 0, static_cast<hipStream_t>(stream), L->nprob * L->nrhs, rsq, col);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
