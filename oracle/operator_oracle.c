/*
 * ORACLE — test infrastructure only. Never shipped, never on the product path.
 *
 * CPU restatement, in plain C99 + pthreads, of the COMMIT implicit operator:
 * the forward product y = A x and the adjoint x = A^T y over the
 * intra-cellular (IC), extra-cellular (EC) and isotropic (ISO) compartments,
 * with and without a kernel look-up table (LUT).
 *
 * Followed algorithm (reference file:line):
 *   oracle_A          <- COMMIT_A / COMMIT_A__block        setup_operator.py:48-242, 245-365
 *   oracle_At         <- COMMIT_At / COMMIT_At__block      setup_operator.py:368-588, 591-710
 *   oracle_A_nolut    <- COMMIT_A_nolut (+ __block)        setup_operator.py:713-813
 *   oracle_At_nolut   <- COMMIT_At_nolut (+ __block)       setup_operator.py:816-915
 *
 * The argument lists are the ones COMMIT_A / COMMIT_At take (operator.pyx:189-227),
 * including the host-thread partitions built by Evaluation.set_threads
 * (core.pyx:586-724), so that the summation order per output element is the
 * reference's: A x walks each thread's voxel-sorted range in order and
 * accumulates every sample in double; A^T y accumulates each dot product in
 * double over ascending samples and adds it to x in segment order within
 * the owning thread.  Unlike the reference there is no process-global state:
 * every call packs its arguments into a struct handed to the workers.
 *
 * This file is also the CPU baseline timed next to the GPU ("port" kind):
 * it reproduces the reference's algorithm and threading, not its source.
 *
 * Parity: pinned by tests/test_oracle.py against a dense FP64 assembly of A
 * from the closed-form entry formula (brute force on tiny inputs), against
 * adjointness <Ax,y> = <x,A^T y>, the single-segment closed form, and the
 * all-ones-LUT (ScalarMap) incidence special case.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_MAX_THREADS 255
#define ORACLE_MAX_RF 20

typedef struct {
    const double *in;
    double *out;
    int nS, ndirs, nF, n, nE, nV;
    const uint32_t *ICf, *ICeval, *ICv, *ECv, *ISOv;
    const uint16_t *ICo, *ECo;
    const float *ICl;
    const float *wmr, *wmh, *iso;
    const uint32_t *ICthreads, *ECthreads, *ISOthreads;
    const uint8_t *ICthreadsT;
    const uint32_t *ECthreadsT, *ISOthreadsT;
    uint32_t nIC, nEC, nISO;
} op_args;

typedef struct {
    const op_args *a;
    int id;
} worker_arg;

/* ---------------------------------------------------------------- A x --- */
static void *A_worker(void *p)
{
    const worker_arg *w = (const worker_arg *)p;
    const op_args *a = w->a;
    const int id = w->id;
    const int nS = a->nS;
    const size_t lut_stride = (size_t)a->ndirs * nS;
    double xv[ORACLE_MAX_RF];

    /* IC: y[v,s] += l * sum_i x_i * LUT_IC_i[o,s], x_i = x[f + i nF] * eval  (:50-109) */
    if (a->nIC > 0 && a->ICthreads) {
        for (uint32_t k = a->ICthreads[id]; k < a->ICthreads[id + 1]; k++) {
            const uint32_t f = a->ICf[k];
            int any = 0;
            for (uint32_t i = 0; i < a->nIC; i++) {
                const size_t c = f + (size_t)i * a->nF;
                xv[i] = a->in[c] * (double)a->ICeval[c];
                any |= (xv[i] != 0);
            }
            if (!any) continue;
            double *y = a->out + (size_t)nS * a->ICv[k];
            const double l = (double)a->ICl[k];
            const size_t off = (size_t)nS * a->ICo[k];
            for (int s = 0; s < nS; s++) {
                double t = 0.0;
                for (uint32_t i = 0; i < a->nIC; i++)
                    t += xv[i] * (double)a->wmr[i * lut_stride + off + s];
                y[s] += l * t;
            }
        }
    }

    /* EC: y[v_e,s] += sum_i x[nIC nF + i nE + e] * LUT_EC_i[o_e,s]  (:111-165) */
    if (a->nEC > 0 && a->ECthreads) {
        const double *xb = a->in + (size_t)a->nIC * a->nF;
        for (uint32_t e = a->ECthreads[id]; e < a->ECthreads[id + 1]; e++) {
            int any = 0;
            for (uint32_t i = 0; i < a->nEC; i++) {
                xv[i] = xb[e + (size_t)i * a->nE];
                any |= (xv[i] != 0);
            }
            if (!any) continue;
            double *y = a->out + (size_t)nS * a->ECv[e];
            const size_t off = (size_t)nS * a->ECo[e];
            for (int s = 0; s < nS; s++) {
                double t = 0.0;
                for (uint32_t i = 0; i < a->nEC; i++)
                    t += xv[i] * (double)a->wmh[i * lut_stride + off + s];
                y[s] += t;
            }
        }
    }

    /* ISO: y[v_k,s] += sum_i x[nIC nF + nEC nE + i nV + k] * LUT_ISO_i[s]  (:167-218) */
    if (a->nISO > 0 && a->ISOthreads) {
        const double *xb = a->in + (size_t)a->nIC * a->nF + (size_t)a->nEC * a->nE;
        for (uint32_t k = a->ISOthreads[id]; k < a->ISOthreads[id + 1]; k++) {
            int any = 0;
            for (uint32_t i = 0; i < a->nISO; i++) {
                xv[i] = xb[k + (size_t)i * a->nV];
                any |= (xv[i] != 0);
            }
            if (!any) continue;
            double *y = a->out + (size_t)nS * a->ISOv[k];
            for (int s = 0; s < nS; s++) {
                double t = 0.0;
                for (uint32_t i = 0; i < a->nISO; i++)
                    t += xv[i] * (double)a->iso[(size_t)i * nS + s];
                y[s] += t;
            }
        }
    }
    return NULL;
}

/* -------------------------------------------------------------- A^T y --- */
static void *At_worker(void *p)
{
    const worker_arg *w = (const worker_arg *)p;
    const op_args *a = w->a;
    const int id = w->id;
    const int nS = a->nS;
    const size_t lut_stride = (size_t)a->ndirs * nS;
    double z[ORACLE_MAX_RF];

    /* IC: every worker scans all segments and handles those it owns (:370-442) */
    if (a->nIC > 0 && a->ICthreadsT) {
        for (int k = 0; k < a->n; k++) {
            if (a->ICthreadsT[k] != id) continue;
            const double *y = a->in + (size_t)nS * a->ICv[k];
            const size_t off = (size_t)nS * a->ICo[k];
            for (uint32_t i = 0; i < a->nIC; i++) {
                const float *lut = a->wmr + i * lut_stride + off;
                double acc = 0.0;
                for (int s = 0; s < nS; s++) acc += (double)lut[s] * y[s];
                z[i] = acc;
            }
            const double l = (double)a->ICl[k];
            const uint32_t f = a->ICf[k];
            for (uint32_t i = 0; i < a->nIC; i++) {
                const size_t c = f + (size_t)i * a->nF;
                a->out[c] += l * z[i] * (double)a->ICeval[c];
            }
        }
    }

    /* EC: equal index split (:444-504) */
    if (a->nEC > 0 && a->ECthreadsT) {
        double *xb = a->out + (size_t)a->nIC * a->nF;
        for (uint32_t e = a->ECthreadsT[id]; e < a->ECthreadsT[id + 1]; e++) {
            const double *y = a->in + (size_t)nS * a->ECv[e];
            const size_t off = (size_t)nS * a->ECo[e];
            for (uint32_t i = 0; i < a->nEC; i++) {
                const float *lut = a->wmh + i * lut_stride + off;
                double acc = 0.0;
                for (int s = 0; s < nS; s++) acc += (double)lut[s] * y[s];
                xb[e + (size_t)i * a->nE] += acc;
            }
        }
    }

    /* ISO (:506-563) */
    if (a->nISO > 0 && a->ISOthreadsT) {
        double *xb = a->out + (size_t)a->nIC * a->nF + (size_t)a->nEC * a->nE;
        for (uint32_t k = a->ISOthreadsT[id]; k < a->ISOthreadsT[id + 1]; k++) {
            const double *y = a->in + (size_t)nS * a->ISOv[k];
            for (uint32_t i = 0; i < a->nISO; i++) {
                const float *lut = a->iso + (size_t)i * nS;
                double acc = 0.0;
                for (int s = 0; s < nS; s++) acc += (double)lut[s] * y[s];
                xb[k + (size_t)i * a->nV] += acc;
            }
        }
    }
    return NULL;
}

/* ----------------------------------------------------------- no LUT ----- */
static void *A_nolut_worker(void *p)
{
    const worker_arg *w = (const worker_arg *)p;
    const op_args *a = w->a;
    const int id = w->id;
    /* (:719-733) */
    for (uint32_t k = a->ICthreads[id]; k < a->ICthreads[id + 1]; k++) {
        const uint32_t f = a->ICf[k];
        const double x0 = a->in[f] * (double)a->ICeval[f];
        if (x0 != 0) a->out[a->ICv[k]] += (double)a->ICl[k] * x0;
    }
    /* (:738-752) */
    if (a->nISO > 0 && a->ISOthreads) {
        const double *xb = a->in + a->nF;
        for (uint32_t k = a->ISOthreads[id]; k < a->ISOthreads[id + 1]; k++) {
            const double x0 = xb[k];
            if (x0 != 0) a->out[a->ISOv[k]] += x0;
        }
    }
    return NULL;
}

static void *At_nolut_worker(void *p)
{
    const worker_arg *w = (const worker_arg *)p;
    const op_args *a = w->a;
    const int id = w->id;
    /* (:822-836) */
    for (int k = 0; k < a->n; k++) {
        if (a->ICthreadsT[k] != id) continue;
        const uint32_t f = a->ICf[k];
        a->out[f] += (double)a->ICl[k] * a->in[a->ICv[k]] * (double)a->ICeval[f];
    }
    /* (:841-850) */
    if (a->nISO > 0 && a->ISOthreadsT) {
        double *xb = a->out + a->nF;
        for (uint32_t k = a->ISOthreadsT[id]; k < a->ISOthreadsT[id + 1]; k++)
            xb[k] += a->in[a->ISOv[k]];
    }
    return NULL;
}

static int run_workers(void *(*fn)(void *), const op_args *a, uint32_t nThreads)
{
    pthread_t th[ORACLE_MAX_THREADS];
    worker_arg wa[ORACLE_MAX_THREADS];
    if (nThreads < 1 || nThreads > ORACLE_MAX_THREADS) return -1;
    if (a->nIC > ORACLE_MAX_RF || a->nEC > ORACLE_MAX_RF || a->nISO > ORACLE_MAX_RF) return -2;
    uint32_t started = 0;
    int rc = 0;
    for (uint32_t t = 0; t < nThreads; t++) {
        wa[t].a = a;
        wa[t].id = (int)t;
        if (pthread_create(&th[t], NULL, fn, &wa[t]) != 0) { rc = -3; break; }
        started++;
    }
    for (uint32_t t = 0; t < started; t++) pthread_join(th[t], NULL);
    return rc;
}

/* Same arguments, in the same order, as COMMIT_A (setup_operator.py:314-324). */
int oracle_A(const double *vIN, double *vOUT, int nS, int ndirs, int nF, int nE, int nV,
             const uint32_t *ICf, const uint32_t *ICeval, const uint32_t *ICv, const uint16_t *ICo,
             const float *ICl, const uint32_t *ECv, const uint16_t *ECo, const uint32_t *ISOv,
             const float *wmrSFP, const float *wmhSFP, const float *isoSFP,
             const uint32_t *ICthreads, const uint32_t *ECthreads, const uint32_t *ISOthreads,
             uint32_t nIC, uint32_t nEC, uint32_t nISO, uint32_t nThreads)
{
    op_args a;
    memset(&a, 0, sizeof a);
    a.in = vIN; a.out = vOUT; a.nS = nS; a.ndirs = ndirs; a.nF = nF; a.nE = nE; a.nV = nV;
    a.ICf = ICf; a.ICeval = ICeval; a.ICv = ICv; a.ICo = ICo; a.ICl = ICl;
    a.ECv = ECv; a.ECo = ECo; a.ISOv = ISOv;
    a.wmr = wmrSFP; a.wmh = wmhSFP; a.iso = isoSFP;
    a.ICthreads = ICthreads; a.ECthreads = ECthreads; a.ISOthreads = ISOthreads;
    a.nIC = nIC; a.nEC = nEC; a.nISO = nISO;
    return run_workers(A_worker, &a, nThreads);
}

/* Same arguments, in the same order, as COMMIT_At (setup_operator.py:660-669). */
int oracle_At(const double *vIN, double *vOUT, int nS, int ndirs, int nF, int n, int nE, int nV,
              const uint32_t *ICf, const uint32_t *ICeval, const uint32_t *ICv, const uint16_t *ICo,
              const float *ICl, const uint32_t *ECv, const uint16_t *ECo, const uint32_t *ISOv,
              const float *wmrSFP, const float *wmhSFP, const float *isoSFP,
              const uint8_t *ICthreadsT, const uint32_t *ECthreadsT, const uint32_t *ISOthreadsT,
              uint32_t nIC, uint32_t nEC, uint32_t nISO, uint32_t nThreads)
{
    op_args a;
    memset(&a, 0, sizeof a);
    a.in = vIN; a.out = vOUT; a.nS = nS; a.ndirs = ndirs; a.nF = nF; a.n = n; a.nE = nE; a.nV = nV;
    a.ICf = ICf; a.ICeval = ICeval; a.ICv = ICv; a.ICo = ICo; a.ICl = ICl;
    a.ECv = ECv; a.ECo = ECo; a.ISOv = ISOv;
    a.wmr = wmrSFP; a.wmh = wmhSFP; a.iso = isoSFP;
    a.ICthreadsT = ICthreadsT; a.ECthreadsT = ECthreadsT; a.ISOthreadsT = ISOthreadsT;
    a.nIC = nIC; a.nEC = nEC; a.nISO = nISO;
    return run_workers(At_worker, &a, nThreads);
}

/* Same arguments as COMMIT_A_nolut (setup_operator.py:779-786). */
int oracle_A_nolut(const double *vIN, double *vOUT, int nF, const uint32_t *ICf,
                   const uint32_t *ICeval, const uint32_t *ICv, const float *ICl,
                   const uint32_t *ISOv, const uint32_t *ICthreads, const uint32_t *ISOthreads,
                   uint32_t nISO, uint32_t nThreads)
{
    op_args a;
    memset(&a, 0, sizeof a);
    a.in = vIN; a.out = vOUT; a.nF = nF;
    a.ICf = ICf; a.ICeval = ICeval; a.ICv = ICv; a.ICl = ICl; a.ISOv = ISOv;
    a.ICthreads = ICthreads; a.ISOthreads = ISOthreads; a.nISO = nISO; a.nIC = 1;
    return run_workers(A_nolut_worker, &a, nThreads);
}

/* Same arguments as COMMIT_At_nolut (setup_operator.py:880-887). */
int oracle_At_nolut(const double *vIN, double *vOUT, int nF, int n, const uint32_t *ICf,
                    const uint32_t *ICeval, const uint32_t *ICv, const float *ICl,
                    const uint32_t *ISOv, const uint8_t *ICthreadsT, const uint32_t *ISOthreadsT,
                    uint32_t nISO, uint32_t nThreads)
{
    op_args a;
    memset(&a, 0, sizeof a);
    a.in = vIN; a.out = vOUT; a.nF = nF; a.n = n;
    a.ICf = ICf; a.ICeval = ICeval; a.ICv = ICv; a.ICl = ICl; a.ISOv = ISOv;
    a.ICthreadsT = ICthreadsT; a.ISOthreadsT = ISOthreadsT; a.nISO = nISO; a.nIC = 1;
    return run_workers(At_nolut_worker, &a, nThreads);
}
