// Device-resident dictionary and plan, shared by the translation units of
// libcommit_hip.so.  Built by commit_hip_plan_create (commit_hip.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"

namespace commit_internal {

constexpr int kXcds = 8;             // MI355X: 8 XCDs, work-groups dealt round-robin

struct Comp {                       // one compartment, voxel-major CSR
    const uint32_t *ptr;            // [nVOX+1]
    const uint16_t *o;              // [m] orientation (null: ISO, a single row)
    const uint32_t *col;            // [m] x column inside one RF block
    const float *len;               // [m] segment length (IC only)
    uint32_t m;
};

struct DevDict {
    Comp ic, ec, iso;
    const uint32_t *eval;           // [nF*nIC]
    const uint32_t *fib_ptr;        // [nF+1]
    const uint32_t *fib_perm;       // [n] IC entry index, fibre-major, ascending voxel
    const uint32_t *zpos;           // [n] inverse of fib_perm: fibre-major position of each IC entry
    const float *wmr, *wmh, *isol;  // padded LUTs: rows of nSp floats
    int nS, nSp, ndirs, nF, n, nE, nI, nVOX, nIC, nEC, nISO;
    int xcd_remap;                  // 1: XCD-contiguous work ranges (wave_index)
};

// Global wavefront index.  Blocks are dealt to the XCDs round-robin; with
// remap, block b is renumbered so that XCD (b % 8) walks a contiguous range of
// the work (a spatial slab of voxels, whose fibres, LUT rows and x entries
// then share that XCD's L2).  The result is made scalar: the compiler cannot
// prove threadIdx.x >> 6 uniform.
__device__ __forceinline__ int wave_index(int remap)
{
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
    const int lb = remap ? x * q + min(x, r) + b / kXcds : b;
    return __builtin_amdgcn_readfirstlane(lb * kWavesPerBlock + (int)(threadIdx.x >> 6));
}

// Output stage of A: y = A x, or the FISTA residual r = sqrtw .* (A x - yobs)
// with a per-voxel partial of ||r||^2 (reduced later in voxel order).
template <typename T>
struct AOut {
    T *y;
    const T *yobs;        // null: plain product
    const T *sqrtw;       // null: unweighted
    double *rpart;        // [nVOX] partial sums of r^2, or null
};

// One compartment's entries listed by orientation (voxel order inside one
// orientation), cut into chunks of one orientation each, for the LUT gradient.
struct OmList {
    uint32_t *ent = nullptr;     // [m] entry index
    uint32_t *vox = nullptr;     // [m] voxel of each listed entry
    uint32_t *chunk = nullptr;   // [nchunks][2] first, end (positions in ent)
    uint32_t *cptr = nullptr;    // [ndirs+1] first chunk of each orientation; null: not built
    double *part = nullptr;      // [nchunks][nS] partial rows, one per chunk
    int nchunks = 0;
};

}  // namespace commit_internal

struct commit_hip_plan {
    int device = 0;
    int precision = 0;
    int nolut = 0;
    int nS_ref = 0;            // nS as the reference reports it (rows = nVOX * nS_ref)
    commit_internal::DevDict d{};
    std::vector<void *> allocs;
    double *zbuf = nullptr;    // [nIC][n] per-entry products of A^T pass 1
    double *xic = nullptr;     // [nF*nIC] x*eval, when eval is not all ones
    double *rpart = nullptr;   // [nVOX + kSumBlocks] residual partials
    float *zmulti = nullptr;   // [nIC][n][16] per-entry products of the multi-RHS A^T pass 1
    // multi-RHS A^T: a voxel's entries of equal orientation share one row
    // product, computed for the first entry of each run (its head); built
    // with zmulti
    uint32_t *mm_uptr = nullptr;   // [nVOX+1] first head of each voxel in mm_uhead
    uint32_t *mm_uhead = nullptr;  // [heads] entry index of each head, voxel-major
    uint32_t *mm_hpos = nullptr;   // [n] head of each entry, in fibre order (as fib_perm)
    float *mm_lenF = nullptr;      // [n] length of each entry, in fibre order
    // the rows of each voxel's A^T product as one flat list (IC heads for each
    // IC RF, EC peaks for each EC RF, ISO entries for each ISO RF): record =
    // LUT row code (low 32 bits) | output slot (high 32 bits); built on first use
    uint32_t *mm_rptr = nullptr;   // [nVOX+1] first row of each voxel
    uint64_t *mm_rec = nullptr;    // [rows]
    // orientation-major lists of the IC, EC and ISO entries for the LUT gradient
    // (lut_grad.hip; built on first use)
    commit_internal::OmList lg_om[3];
    bool eval_ones = true;
    int multi_a_waves = 2;     // wavefronts per (voxel, 16 RHS) in the multi-RHS A x (1 or 2)
    int a_dense = 0;           // 1: A x by the dense-tile matrix-core kernel (dense.hip)
    // LDS-slab kernels (slab.hip): 0 off, 1: 8 wavefronts x 80 KiB per work-group,
    // 2: 16 wavefronts x 160 KiB; S orientations per slab, nsl slabs
    int slab = 0;
    int slab_S = 0, slab_nsl = 0;
    unsigned slab_blocks = 0;
    unsigned slab_blocks_at = 0;
    int32_t *slab_vlist = nullptr;     // A x:   [slab_blocks][W][kSlabVA] voxel ids (-1: none)
    int32_t *slab_vlist_at = nullptr;  // A^T y: [slab_blocks_at][W][kSlabVAt]
    uint32_t *slab_ofs = nullptr;      // [nVOX][nsl+1] first IC entry of each slab
    int64_t bytes = 0;
};

namespace commit_internal {
// hipMalloc tracked by the plan (freed by commit_hip_plan_destroy)
int dev_alloc(commit_hip_plan *p, size_t bytes, void **out);

// LDS-slab kernels (slab.hip): voxels per wavefront (A x holds NJ FP64
// accumulators per voxel, A^T y an FP64 y row) and the largest NJ.
constexpr int kSlabVA = 6;
constexpr int kSlabVAt = 6;
constexpr int kSlabMaxNJ = 5;

struct SlabArgs {
    const uint32_t *ofs;    // [nVOX][nsl+1]
    const int32_t *vlist;   // [blocks][W][V]
    int S, nsl;
    int rs;                 // LDS row stride of the slab in floats: nS rounded up to 4
};

// LDS row stride of the slab kernels (16-byte aligned rows, no padding to nSp)
inline int slab_row_stride(const DevDict &d) { return (d.nS + 3) & ~3; }

// (Re)build the slab tables for mode 0 (off), 1 or 2; the plan's IC entries
// must be sorted by orientation inside each voxel.
int slab_configure(commit_hip_plan *p, int mode);
void slab_release(commit_hip_plan *p);
// COMMIT_HIP_EUNSUPPORTED when slabs are off or the sizes are outside the build.
template <typename T>
int launch_A_slab(commit_hip_plan *p, const T *x, const T *xic, const AOut<T> &out, hipStream_t st);
template <typename T>
int launch_At_slab(commit_hip_plan *p, const T *y, const T *sqrtw, T *x, hipStream_t st);
// Pass 1 of A^T (sqrtw .* (a r1 + b r2)) on the LDS slabs, with each voxel's
// sum of squares of a r1 + b r2 (FP64) in p->rpart[v]; EUNSUPPORTED when the
// plan has no slab tables.
template <typename T>
int launch_At_slab_comb(commit_hip_plan *p, const T *r1, const T *r2, double a, double b, const T *sqrtw, T *x,
                        hipStream_t st);

// Dense-tile A x (dense.hip); COMMIT_HIP_EUNSUPPORTED when the plan's sizes
// are outside what it is built for.
template <typename T>
int launch_A_dense(commit_hip_plan *p, const T *x, const T *xic, T *y, const T *yobs, const T *sqrtw, double *rpart,
                   hipStream_t st);
}  // namespace commit_internal
