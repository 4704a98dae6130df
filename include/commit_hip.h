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
#define COMMIT_HIP_FP32 1         /* x, y float32; compensated FP32 accumulation */

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

/* Replace ICeval (debias mask, core.pyx:1481-1487); host pointer, nF*nIC entries. */
int commit_hip_plan_set_eval(commit_hip_plan *plan, const uint32_t *ICeval);

/* Sizes: rows = nVOX*nS, cols = nF*nIC + nE*nEC + nV*nISO (operator.pyx:69-70). */
int commit_hip_plan_shape(const commit_hip_plan *plan, int64_t *nrows, int64_t *ncols);

/* Device bytes held by the plan (dictionary + scratch). */
int64_t commit_hip_plan_device_bytes(const commit_hip_plan *plan);

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

/* Thread-local message describing the last failure of this thread. */
const char *commit_hip_last_error(void);

/* Library version string. */
const char *commit_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif /* COMMIT_HIP_H */
