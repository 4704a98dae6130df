/*
 * commit_hip.h — C ABI of libcommit_hip.so, the MI355X (gfx950) implementation
 * of the COMMIT implicit operator A (forward A*x, adjoint A^T*y) and of the
 * vector kernels of its FISTA solver.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative COMMIT_HIP_E* code;
 *     commit_hip_last_error() then returns a message (thread-local).
 *   - A plan owns the device-resident dictionary.  Different plans may be used
 *     concurrently from different host threads (no global state); calls on the
 *     same plan must be serialised by the caller (they share scratch memory).
 *   - Vectors passed to the products are DEVICE pointers; `stream` is a
 *     hipStream_t (NULL = default stream).  Products are asynchronous.
 *   - The dictionary arrays given to commit_hip_plan_create are HOST pointers,
 *     exactly the numpy arrays the reference hands to its C operator.
 */
#ifndef COMMIT_HIP_H
#define COMMIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COMMIT_HIP_OK 0
#define COMMIT_HIP_EINVAL -1      /* bad argument / out-of-range index in the dictionary */
#define COMMIT_HIP_ENOMEM -2      /* device allocation failed */
#define COMMIT_HIP_EDEVICE -3     /* HIP runtime error */
#define COMMIT_HIP_EUNSUPPORTED -4 /* size outside what the kernels are built for */

/* precision of x, y and the accumulation */
#define COMMIT_HIP_FP64 0         /* drop-in: x, y float64; FP64 accumulation */
#define COMMIT_HIP_FP32 1         /* x, y (and FISTA vectors) float32.  LDS-slab products in FP32
                                     arithmetic: A x sums groups of 4 entries in FP32 into FP64
                                     accumulators, A^T y forms per-lane FP32 partial dot products and
                                     reduces them in FP64.  The per-voxel and dense-tile products
                                     accumulate in FP64.  Outputs rounded once on store. */

typedef struct commit_hip_plan commit_hip_plan;

/*
 * Build a device-resident operator from the dictionary.
 * Replaces the pointer capture of LinearOperator.__init__ (operator.pyx:40-94)
 * plus the thread partition of Evaluation.set_threads (core.pyx:586-724): the
 * plan derives its own voxel-major and fibre-major layouts on the host and
 * uploads them once.  Argument names and meanings follow COMMIT_A/COMMIT_At
 * (setup_operator.py:314-324, 660-669):
 *   nS     samples per voxel        ndirs  LUT orientations
 *   nF     fibres (IC columns/RF)   n      IC segments
 *   nE     EC peaks                 nV     ISO voxels (the ISO x stride)
 *   nVOX   voxels in the fit mask (rows of A are nVOX*nS)
 *   ICf/ICv/ICo/ICl  segment fibre, compact voxel, orientation, length [n]
 *   ICeval           per-column on/off mask [nF*nIC]
 *   ECv/ECo [nE], ISOv [nV]
 *   wmrSFP [nIC][ndirs][nS], wmhSFP [nEC][ndirs][nS], isoSFP [nISO][nS]  (float32)
 *   nolut  1: LUT-free operator (COMMIT_A_nolut, setup_operator.py:713-813):
 *          nS = 1, LUT = 1, no EC compartment.
 */
int commit_hip_plan_create(commit_hip_plan **plan, int nS, int ndirs, int nF, int n, int nE, int nV,
                           int nVOX, const uint32_t *ICf, const uint32_t *ICeval, const uint32_t *ICv,
                           const uint16_t *ICo, const float *ICl, const uint32_t *ECv,
                           const uint16_t *ECo, const uint32_t *ISOv, const float *wmrSFP,
                           const float *wmhSFP, const float *isoSFP, uint32_t nIC, uint32_t nEC,
                           uint32_t nISO, int nolut, int precision, int device);

/* Free every device buffer of the plan. */
int commit_hip_plan_destroy(commit_hip_plan *plan);

/* Replace ICeval (debias mask, core.pyx:1482-1487); host pointer, nF*nIC entries. */
int commit_hip_plan_set_eval(commit_hip_plan *plan, const uint32_t *ICeval);

/* Sizes: rows = nVOX*nS, cols = nF*nIC + nE*nEC + nV*nISO (operator.pyx:69-70). */
int commit_hip_plan_shape(const commit_hip_plan *plan, int64_t *nrows, int64_t *ncols);

/* Device bytes held by the plan (dictionary + scratch). */
int64_t commit_hip_plan_device_bytes(const commit_hip_plan *plan);

/*
 * Storage type of the vectors the single-vector entry points (commit_hip_A,
 * _At, _At_weighted, _At_pass, _A_residual) read and write: COMMIT_HIP_FP64
 * or COMMIT_HIP_FP32.  The dictionary is shared; switching is free, so a
 * mixed-precision fit can take its Lipschitz estimate and final residual in
 * FP64 on the same plan.  No counterpart in the reference (FP64 only).
 */
int commit_hip_plan_set_precision(commit_hip_plan *plan, int precision);
int commit_hip_plan_precision(const commit_hip_plan *plan);

/*
 * Launch options, set by the autotuner (commit_amd.autotune) after timing the
 * variants on the plan's own data; no counterpart in the reference.
 *   COMMIT_HIP_OPT_XCD_REMAP     1: blocks renumbered so each XCD walks a
 *                                contiguous voxel range; 0 (default): hardware order
 *   COMMIT_HIP_OPT_MULTI_A_WAVES 1 or 2 (default): wavefronts per (voxel, 16 RHS)
 *                                in commit_hip_A_multi
 *   COMMIT_HIP_OPT_A_DENSE       0 (default): A x per segment; 1: per tile of 16
 *                                voxels as a dense FP64 matrix-core product over
 *                                all orientations (ndirs <= 512, nS <= 320)
 *   COMMIT_HIP_OPT_SLAB          0 (default): IC rows of A x and A^T y read from
 *                                L2 per entry; 1 or 2: IC table staged in LDS in
 *                                slabs of orientations by work-groups of 8 or 16
 *                                wavefronts (80 or 160 KiB; nS <= 320); setting
 *                                it builds the slab tables (a host copy of the
 *                                voxel pointers and one small kernel)
 * get returns the value, or a negative code.
 */
#define COMMIT_HIP_OPT_XCD_REMAP 1
#define COMMIT_HIP_OPT_MULTI_A_WAVES 2
#define COMMIT_HIP_OPT_A_DENSE 3
#define COMMIT_HIP_OPT_SLAB 4
int commit_hip_plan_set_option(commit_hip_plan *plan, int option, int value);
int commit_hip_plan_get_option(const commit_hip_plan *plan, int option);

/*
 * y = A x.  Replaces COMMIT_A (setup_operator.py:314-365) and, for a LUT-free
 * plan, COMMIT_A_nolut (setup_operator.py:779-813).  x: cols, y: rows, device
 * pointers of the plan's precision.  y is fully overwritten (no pre-zeroing
 * needed, unlike the reference which accumulates into a zeroed buffer,
 * operator.pyx:133).
 */
int commit_hip_A(commit_hip_plan *plan, const void *x, void *y, void *stream);

/*
 * x = A^T y.  Replaces COMMIT_At (setup_operator.py:660-710) and, for a
 * LUT-free plan, COMMIT_At_nolut (setup_operator.py:880-915).  Deterministic:
 * no atomics; each fibre's segment contributions are summed in ascending
 * voxel order, the order the reference's owning thread uses.
 */
int commit_hip_At(commit_hip_plan *plan, const void *y, void *x, void *stream);

/* ------------------------------------------------------------------------
 * FISTA building blocks (solvers.py:244-391, proximals.pyx:13-186).
 * Vectors are DEVICE arrays of the plan's precision for the products;
 * commit_hip_prox / _fista_sums / _axpby take float64 vectors and their _f32
 * twins float32 ones (weights stay float64; every sum accumulates in FP64).
 * Reductions are deterministic (fixed work split and order, no atomics).
 * ---------------------------------------------------------------------- */

/*
 * res = sqrtw .* (A x - yobs) and *rsq = ||res||^2 (rsq: one device double).
 * yobs may be NULL (res = A x, *rsq = ||A x||^2, the Lipschitz quotient of
 * solvers.py:282-286); sqrtw may be NULL (no confidence weights,
 * solvers.py:238-239); res may be NULL (the norm only, as a FISTA trial point
 * needs).  Fuses the reference's A.dot(x) - y and norm (solvers.py:265-270,
 * 308-311).
 */
int commit_hip_A_residual(commit_hip_plan *plan, const void *x, const void *yobs, const void *sqrtw,
                          void *res, double *rsq, void *stream);

/*
 * r = a r1 + b r2 over the operator's nrows rows, in the plan's precision, and
 * *rsq = ||r||^2 (FP64, from the unrounded values, fixed order).  The FISTA
 * momentum point's residual: the residual is affine in x, so
 * r(xhat) = (1 + beta) r(x) - beta r(prev_x) for xhat = x + beta (x - prev_x),
 * one pass over three row vectors instead of the product A xhat that
 * solvers.py:363-368 recomputes (no counterpart in the reference).  r1, r2
 * and r must be 16-byte aligned (as hipMalloc and torch allocations are).
 */
int commit_hip_residual_combine(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2, void *r,
                                double *rsq, void *stream);

/* x = A^T (sqrtw .* y); sqrtw may be NULL (solvers.py:266-267, 365-366). */
int commit_hip_At_weighted(commit_hip_plan *plan, const void *y, const void *sqrtw, void *x, void *stream);

/*
 * The FISTA gradient at the momentum point in one call: with
 * rhat = a r1 + b r2 (FP64, rounded to the plan's vectors, as
 * commit_hip_residual_combine forms it), x = A^T (sqrtw .* rhat) and
 * *rsq = ||rhat||^2 (FP64, fixed order).  On the LDS-slab path the combination
 * is formed while A^T loads its rows and rhat is never stored; otherwise the
 * call is commit_hip_residual_combine into r, then commit_hip_At_weighted.  r:
 * a row vector the call may overwrite.  Replaces the combination of
 * solvers.py:363-368 followed by the gradient's A^T (no counterpart in the
 * reference, which recomputes A xhat).
 */
int commit_hip_At_combined(commit_hip_plan *plan, double a, const void *r1, double b, const void *r2,
                           const void *sqrtw, void *r, void *x, double *rsq, void *stream);

/*
 * One pass of A^T y, for timing and profiling each kernel on its own (no
 * counterpart in the reference): pass 1 = voxel-major products (and the EC/ISO
 * columns of x), pass 2 = per-fibre sums into the IC columns of x; pass 3 = both
 * (commit_hip_At_weighted).  Pass 2 reads the scratch pass 1 left in the plan.
 */
int commit_hip_At_pass(commit_hip_plan *plan, const void *y, const void *sqrtw, void *x, int pass,
                       void *stream);

/*
 * Multi-right-hand-side products (no counterpart in the reference, which runs
 * one COMMIT_A / COMMIT_At per vector, solvers.py:266-270): R = nrhs columns
 * at once, row-major device matrices with the RHS index fastest:
 *   commit_hip_A_multi  : Y[nrows][R] = A X[ncols][R]
 *   commit_hip_At_multi : X[ncols][R] = A^T (sqrtw .* Y)   (sqrtw [nrows] or NULL)
 * dtype must be COMMIT_HIP_FP32 (float32 X, Y, sqrtw; FP32 matrix-core
 * products, FP64 per-fibre sums in A^T).  Needs nS <= 320 and a LUT-based
 * plan.  Outputs are fully overwritten.  A^T allocates its per-entry scratch
 * (16 floats per IC segment and RF) and builds its run-head lists (12 bytes per
 * IC segment; entries of equal orientation in a voxel share one row product)
 * on first use, synchronously, and keeps them in the plan: make the first call
 * outside graph capture.
 */
int commit_hip_A_multi(commit_hip_plan *plan, const void *X, void *Y, int nrhs, int dtype, void *stream);
int commit_hip_At_multi(commit_hip_plan *plan, const void *Y, const void *sqrtw, void *X, int nrhs, int dtype,
                        void *stream);

/* Regularisation of one fit, the device image of solvers.init_regularisation
 * (solvers.py:16-206).  Compartments: 0 = IC, 1 = EC, 2 = ISO. */
#define COMMIT_HIP_REG_NONE 0
#define COMMIT_HIP_REG_LASSO 1
#define COMMIT_HIP_REG_GROUP_LASSO 2
#define COMMIT_HIP_REG_SPARSE_GROUP_LASSO 3
typedef struct {
    int64_t start[3], size[3];  /* compartment ranges in x */
    int reg[3];                 /* COMMIT_HIP_REG_*; group types only for IC */
    int nn[3];                  /* non-negativity per compartment */
    int weighted[3];            /* lasso (or the l1 part of sparse GL) uses w */
    double lam[3];              /* lasso lambda (lambda_1 for sparse GL) */
    double lam_group;           /* group-lasso lambda (IC) */
    const double *w;            /* [n] per-coefficient weights, absolute index */
    int64_t ngroups;
    const int32_t *group_ptr;   /* [ngroups+1] offsets into group_idx */
    const int32_t *group_idx;   /* member indices into x */
    const double *group_w;      /* [ngroups] */
    int groups_overlap;         /* 1: apply groups one after another, in order */
} commit_hip_prox_params;

/*
 * x = prox(xin - step * grad, scale) with the reference's composition
 * proxIC(proxEC(proxISO(.))) (solvers.py:200-201): soft-thresholding, group
 * shrinkage (proximals.pyx:103-134, groups in the given order), then
 * non-negativity (proximals.pyx:13-22).  grad may be NULL (x = prox(xin));
 * xin may equal x.
 */
int commit_hip_prox(const commit_hip_prox_params *p, int64_t n, const double *xin, const double *grad,
                    double step, double scale, double *x, void *stream);

/* doubles of workspace commit_hip_fista_sums needs, for ngroups groups */
#define COMMIT_HIP_SUMS_WORK(ngroups) (8 * 256 + (ngroups))

/*
 * Deterministic sums for one FISTA step (device out[8]; device workspace
 * `work` of COMMIT_HIP_SUMS_WORK(ngroups) doubles):
 *   out[0] = sum (x - xhat) * grad      out[1] = sum (x - xhat)^2
 *   out[2] = sum (x - prev)^2           out[3] = sum x^2
 *   out[4..6] = sum |w_i x_i| over l1_range[2t]..l1_range[2t+1] (w_i = 1
 *               unless l1_weighted[t])
 *   out[7] = sum_g group_w[g] * ||x_g||
 * xhat/grad or prev may be NULL (their sums are 0); ngroups may be 0.
 */
int commit_hip_fista_sums(int64_t n, const double *x, const double *xhat, const double *grad,
                          const double *prev, const int64_t *l1_range, const int *l1_weighted,
                          const double *w, int64_t ngroups, const int32_t *group_ptr,
                          const int32_t *group_idx, const double *group_w, double *out, double *work,
                          void *stream);

/* Most groups commit_hip_fista_sums_noalloc accepts. */
#define COMMIT_HIP_NOALLOC_MAX_GROUPS 65536

/*
 * commit_hip_fista_sums without the `work` argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels, split and order as commit_hip_fista_sums, so the sums are
 * bitwise equal to it.  The workspace is a fixed device array of
 * COMMIT_HIP_SUMS_WORK(COMMIT_HIP_NOALLOC_MAX_GROUPS) doubles (about 540 KB)
 * that every GPU the library runs on holds until the process ends; the call
 * allocates nothing (no hipMalloc, no hipFree).  Runs on the current device.
 * ngroups > COMMIT_HIP_NOALLOC_MAX_GROUPS returns COMMIT_HIP_EUNSUPPORTED.
 * All calls on one GPU share that array: calls that overlap in time on one GPU
 * (on different streams, or from different host threads) race; calls queued
 * on one stream are ordered and safe.
 */
int commit_hip_fista_sums_noalloc(int64_t n, const double *x, const double *xhat, const double *grad,
                                  const double *prev, const int64_t *l1_range, const int *l1_weighted,
                                  const double *w, int64_t ngroups, const int32_t *group_ptr,
                                  const int32_t *group_idx, const double *group_w, double *out, void *stream);

/* out = a * x + b * y (FISTA extrapolation, solvers.py:360). */
int commit_hip_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out,
                     void *stream);

/* float32 twins of the three above (same arguments, x-like vectors float32) */
int commit_hip_prox_f32(const commit_hip_prox_params *p, int64_t n, const float *xin, const float *grad,
                        double step, double scale, float *x, void *stream);
int commit_hip_fista_sums_f32(int64_t n, const float *x, const float *xhat, const float *grad,
                              const float *prev, const int64_t *l1_range, const int *l1_weighted,
                              const double *w, int64_t ngroups, const int32_t *group_ptr,
                              const int32_t *group_idx, const double *group_w, double *out, double *work,
                              void *stream);
int commit_hip_axpby_f32(int64_t n, double a, const float *x, double b, const float *y, float *out,
                         void *stream);

/*
 * Differentiable operator support (no counterpart in the reference, whose
 * response-function tables are fixed inputs, setup_operator.py:93,151,299).
 *
 * commit_hip_plan_set_lut: replace the plan's tables from DEVICE float32
 * arrays wmr [nIC][ndirs][nS], wmh [nEC][ndirs][nS], iso [nISO][nS] (NULL:
 * keep that table).
 *
 * commit_hip_lut_grad: gradient of L(y) with y = A x with respect to the
 * tables, given x and gy = dL/dy (device vectors of the plan's precision):
 *   gwmr[i][o][s] = sum_{IC segments (f,v,o,l)} l * x[i*nF+f] * eval * gy[v*nS+s]
 *   gwmh[i][o][s] = sum_{EC peaks (v,o), column e} x[nIC*nF + i*nE + e] * gy[v*nS+s]
 *   giso[i][s]    = sum_{ISO voxels v, column k} x[nIC*nF + nEC*nE + i*nV + k] * gy[v*nS+s]
 * FP64 device outputs (NULL: skip that table).  Deterministic (LDS-private
 * tables per voxel slab, added in slab order).  Keeps a scratch buffer in the
 * plan after the first call.
 */
int commit_hip_plan_set_lut(commit_hip_plan *plan, const float *wmr, const float *wmh, const float *iso,
                            void *stream);
int commit_hip_lut_grad(commit_hip_plan *plan, const void *x, const void *gy, double *gwmr, double *gwmh,
                        double *giso, void *stream);

/* ------------------------------------------------------------------------
 * Batched FISTA (no counterpart in the reference, which fits one problem per
 * solve call, solvers.py:218-391): nprob x nrhs independent fits advanced
 * together, nrhs columns of float32 [rows][nrhs] matrices (lambda sweeps,
 * bootstrap signals) times nprob problems occupying disjoint unknowns and
 * rows of one block-diagonal plan (tractograms of uneven size).  Every
 * decision of the reference's loop (step acceptance and backtracking,
 * stopping tests in order, momentum) is taken per problem on the device, so
 * the kernels of one trial step can be captured in a HIP graph and replayed.
 * All sums and scalars are FP64; reductions are deterministic.
 * ---------------------------------------------------------------------- */

/* State of one (problem, column) fit; device array indexed problem * nrhs + column. */
typedef struct {
    double step, told, prev_obj, qfval, coef;
    double cost, regterm, residual, abs_cost, rel_cost, abs_x, rel_x;
    double lam[3], lam_group;   /* the penalty of this fit (lambda * its weights' scale) */
    int32_t iterations, done, criterion, advance, trials, pad_;
} commit_hip_fista_col;

/* criterion codes: 1 abs. objective, 2 rel. objective, 3 abs. unknowns, 4 rel. unknowns, 5 max iterations */

typedef struct {
    int64_t n, m;                 /* unknowns (cols of A) and rows */
    int nrhs, nprob;
    const int32_t *segx;          /* [n] problem of each unknown (device) */
    const int64_t *xpieces;       /* [2*nxpieces] [begin, end) unknown ranges, each inside one problem */
    const int32_t *xpseg;         /* [nxpieces] their problem */
    int64_t nxpieces;
    const int64_t *ypieces;       /* [2*nypieces] row ranges, each inside one problem */
    const int32_t *ypseg;
    int64_t nypieces;
    const int32_t *gseg;          /* [ngroups] problem of each group (group penalties), or NULL */
    const int32_t *vseg;          /* [voxels] problem of each voxel (device): the momentum entry points, or NULL */
} commit_hip_multi_layout;

/* doubles of workspace commit_hip_fista_multi_sums needs; the residual needs nypieces * nrhs */
int64_t commit_hip_fista_multi_work(const commit_hip_multi_layout *L, int64_t ngroups);

/* x = prox(xhat - step * grad, step) per fit (step and lambdas from col); finished fits keep x */
int commit_hip_prox_multi(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, const float *xhat,
                          const float *grad, float *x, const commit_hip_fista_col *col, void *stream);

/* out[(problem * nrhs + column) * 8 + k]: the eight sums of commit_hip_fista_sums per fit */
int commit_hip_fista_multi_sums(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, const float *x,
                                const float *xhat, const float *grad, const float *prev, const int64_t *l1_range,
                                const int *l1_weighted, double *out, double *work, void *stream);

/* res = sqrtw .* (AX - y) (res may be NULL) and rsq[problem * nrhs + column] = ||res||^2;
 * y is [m] shared by the columns, or [m][nrhs] when y_per_column */
int commit_hip_multi_residual(const commit_hip_multi_layout *L, const float *AX, const float *y, int y_per_column,
                              const float *sqrtw, float *res, double *rsq, double *work, void *stream);

/* commit_hip_A_multi and commit_hip_multi_residual in one pass over the rows:
 * res = sqrtw .* (A X - y) (res may be NULL: norms only; y NULL: zero;
 * sqrtw NULL: one), rsq[problem * nrhs + column] = ||res||^2, without the
 * [rows][nrhs] product A X in memory.  res equals the two-call result bit for
 * bit; rsq is summed per voxel first (FP64, deterministic).  A voxel counts
 * towards the problem holding its first row, so problem row ranges must begin
 * and end on voxel boundaries.  work: commit_hip_A_multi_residual_work doubles. */
int64_t commit_hip_A_multi_residual_work(const commit_hip_plan *plan, const commit_hip_multi_layout *L);
int commit_hip_A_multi_residual(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *X,
                                const float *y, int y_per_column, const float *sqrtw, float *res, double *rsq,
                                double *work, void *stream);

/* accept / backtrack / stop each fit after a trial (sums of the trial x, rsq of A x - y) */
int commit_hip_fista_multi_control(const commit_hip_prox_params *p, const commit_hip_multi_layout *L, double tol_fun,
                                   double tol_x, int max_iter, const double *sums, const double *rsq,
                                   commit_hip_fista_col *col, void *stream);

/* accepted fits: xhat = x + coef (x - prev), prev = x */
int commit_hip_fista_multi_update(const commit_hip_multi_layout *L, const float *x, float *xhat, float *prev,
                                  const commit_hip_fista_col *col, void *stream);

/* accepted fits: qfval = ||res at the new xhat||^2 / 2 */
int commit_hip_fista_multi_qf(const commit_hip_multi_layout *L, const double *rsq, commit_hip_fista_col *col,
                              void *stream);

/* Momentum residual of the batched fits, without the product A xhat that the
 * reference recomputes after every accepted step (solvers.py:362-368).  The
 * residual is affine in x, so for xhat = x + coef (x - prev_x),
 * r(xhat) = (1 + coef) r(x) - coef r(prev_x).  Two [rows][nrhs] float32
 * buffers r0, r1 hold r(prev_x) of every unfinished fit (in r[*sel]) and the
 * trial point's residual (in r[1 - *sel]); *sel is an int32 in device memory,
 * so a captured trial step keeps working when the roles swap.  All three need
 * L->vseg; work: commit_hip_A_multi_residual_work doubles.
 *
 * commit_hip_A_multi_trial: commit_hip_A_multi_residual with res = r[1 - *sel]. */
int commit_hip_A_multi_trial(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *X, const float *y,
                             int y_per_column, const float *sqrtw, float *r0, float *r1, const int32_t *sel,
                             double *rsq, double *work, void *stream);

/* After commit_hip_fista_multi_control and _update: for every fit with advance
 * set, rhat = (1 + coef) r[1 - *sel] - coef r[*sel] (FP64, rounded to float as
 * commit_hip_residual_combine), grad = A^T (sqrtw .* rhat) (the fit's columns
 * of the [n][nrhs] matrix grad) and rsq = ||rhat||^2.  Columns of the other
 * fits of grad keep their values; their rsq entries are unspecified. */
int commit_hip_At_multi_momentum(commit_hip_plan *plan, const commit_hip_multi_layout *L, const float *r0,
                                 const float *r1, const int32_t *sel, const float *sqrtw,
                                 const commit_hip_fista_col *col, float *grad, double *rsq, double *work,
                                 void *stream);

/* After commit_hip_At_multi_momentum: moves r(prev_x) of every unfinished fit
 * into r[*sel] (copies the advanced fits' columns when some unfinished fit
 * backtracked, else swaps the roles by flipping *sel). */
int commit_hip_fista_multi_roles(const commit_hip_plan *plan, const commit_hip_multi_layout *L, float *r0, float *r1,
                                 int32_t *sel, const commit_hip_fista_col *col, void *stream);

/* ------------------------------------------------------------------------
 * Tractogram voxeliser: streamlines -> dictionary, on the GPU.
 * ---------------------------------------------------------------------- */

/* A voxeliser handle keeps its device buffers between runs (they grow on
 * demand) and the dictionary of its last successful run.  Calls on one handle
 * must be serialised by the caller. */
typedef struct commit_hip_vox commit_hip_vox;
int commit_hip_vox_create(commit_hip_vox **vox, int device);
int commit_hip_vox_destroy(commit_hip_vox *vox);

/*
 * Replaces trk2dictionary() (trk2dictionary_c.cpp:111-286, called from
 * trk2dictionary.pyx:534-539) for streamlines already read into memory: the
 * IC segments (ICSegments + fiberForwardModel + segmentForwardModel +
 * rayBoxIntersection, :444-848), the EC peaks (ECSegments, :266-383), the ISO
 * voxels (ISOcompartments, :386-436), the kept-streamline renumbering of
 * IC_f (trk2dictionary.pyx:551-563) and the TDI map (compute_tdi, pyx:50-59).
 * Arguments keep the reference's names and meanings, except:
 *   fibers      float32 [n_points][3] streamline coordinates in voxmm, as
 *               read_fiberTRK / read_fiberTCK return them (:873-911), and
 *   fib_ptr     int64 [n_count+1] first point of each streamline, in place of
 *               the file name, data offset, scalars and properties;
 *   inputs_on_device  0: fibers, fib_ptr, ptrPEAKS, ptrMASK and ptrISO are
 *               HOST pointers; 1: they are DEVICE pointers;
 *   fiber_shift in voxel units (the reference scales it by the voxel size);
 *   ptrTDI, path_out, threads_count, verbose are dropped: the results stay in
 *               the handle (commit_hip_vox_sizes / _fetch);
 *   ptrBlurApplyTo is uint8 per streamline (NULL: blur all);
 *   stream      hipStream_t of the work (the call returns when it is done).
 * Masks and peaks are float32 arrays in C order [Nx][Ny][Nz](...), as the
 * reference receives them; ptrPEAKS / ptrMASK / ptrISO may be NULL.  The
 * hash table, replicas and affine are HOST arrays.
 * ptrPeaksAffine: 3x3 row-major (the identity unless peaks_use_affine).
 * Returns COMMIT_HIP_EUNSUPPORTED for a streamline of >= 10000 points (the
 * reference's MAX_FIB_LEN, where it generates no dictionary), for a grid of
 * more than 2^32 voxels, or when a polyline step crosses more voxels than a
 * cap of 64 + 12 * length / (smallest voxel size) (the reference has no cap).
 * Every output is bit-for-bit that of the reference's arithmetic.
 */
int commit_hip_trk2dictionary(commit_hip_vox *vox, const float *fibers, const int64_t *fib_ptr, int n_count,
                              int64_t n_points, int inputs_on_device, int Nx, int Ny, int Nz, float Px, float Py,
                              float Pz, float fiber_shiftX, float fiber_shiftY, float fiber_shiftZ,
                              float min_seg_len, float min_fiber_len, float max_fiber_len, const float *ptrPEAKS,
                              int Np, float vf_THR, int ECix, int ECiy, int ECiz, const float *ptrMASK,
                              const float *ptrISO, int c, const double *ptrPeaksAffine, int nReplicas,
                              const double *ptrBlurRho, const double *ptrBlurAngle, const double *ptrBlurWeights,
                              const uint8_t *ptrBlurApplyTo, const short *ptrHashTable, void *stream);

/* Sizes of the last run: IC segments, kept streamlines, EC peaks, ISO voxels,
 * and the pieces recorded before aggregation (any pointer may be NULL). */
int commit_hip_vox_sizes(const commit_hip_vox *vox, int64_t *n_ic, int64_t *n_kept, int64_t *n_ec,
                         int64_t *n_iso, int64_t *n_pieces);

/* Copies the last run's dictionary to HOST arrays (NULL skips one): the
 * contents of dictionary_{IC_f, IC_v, IC_o, IC_len, TRK_norm, TRK_len,
 * TRK_lenTot, TRK_kept, EC_v, EC_o, ISO_v}.dict (trk2dictionary.pyx:565-612)
 * and the TDI map, float32 [Nx*Ny*Nz] with x fastest (pyx:616-620). */
int commit_hip_vox_fetch(commit_hip_vox *vox, uint32_t *IC_f, uint32_t *IC_v, uint16_t *IC_o, float *IC_len,
                         float *TRK_norm, float *TRK_len, float *TRK_lenTot, uint8_t *TRK_kept, uint32_t *EC_v,
                         uint16_t *EC_o, uint32_t *ISO_v, float *tdi);

/* Thread-local message describing the last failure of this thread. */
const char *commit_hip_last_error(void);

/* Library version string. */
const char *commit_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif /* COMMIT_HIP_H */
