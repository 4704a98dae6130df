/*
 * Synthetic tractogram -> COMMIT dictionary generator (host C, pthreads).
 *
 * Produces the IC part of a dictionary in the layout trk2dictionary writes
 * (trk2dictionary_c.cpp:517-547): per kept fibre, its segments as
 * (u32 fibre, u32 linear voxel x + X*(y + Y*z), u16 orientation, f32 length),
 * aggregated per (orientation, z, y, x) key in ascending key order
 * (trk2dictionary_c.cpp:38-41 segKey ordering), plus per-fibre
 * TRK_norm = sqrt(sum_v (sum_{seg in v} len)^2), TRK_len and TRK_lenTot
 * (trk2dictionary_c.cpp:528-547).
 *
 * Fibres are seeded random cubic B-splines with 5 control points, sampled at a
 * fixed step; each sample step contributes its length to the voxel holding its
 * midpoint (trk2dictionary_c.cpp:761-808 segment model), with the direction
 * flipped into the y >= 0 hemisphere and hashed to the nearest of `ndirs`
 * directions through a 181x181 (colatitude, longitude) degree table.
 *
 * Every fibre draws from its own counter-based random stream, so the output
 * does not depend on the number of threads.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t seed;
    int nF;                   /* fibres to keep */
    int dim[3];
    double pixdim;            /* mm per voxel (lengths are written in mm) */
    const uint8_t *mask;      /* dim[0]*dim[1]*dim[2], x fastest; NULL = whole box */
    int mode;                 /* 0: 5 control points uniform in the box; 1: length-controlled */
    double len_min_mm, len_max_mm;   /* mode 1 */
    double step_vox;          /* sampling step along the curve, in voxels */
    const short *hash;        /* 181*181 direction table */
    int nthreads;
} synth_params;

typedef struct {
    uint32_t *f, *v;
    uint16_t *o;
    float *len;
    size_t n, cap;
} seg_buf;

typedef struct {
    uint32_t *IC_f, *IC_v;
    uint16_t *IC_o;
    float *IC_len;
    float *TRK_norm, *TRK_len, *TRK_lenTot;
    size_t n;
    int nF;
} synth_result;

/* ------------------------------------------------------------ random ---- */
static inline uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double urand(uint64_t *s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
static inline double nrand(uint64_t *s)
{
    double u1 = urand(s), u2 = urand(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ----------------------------------------------------------- B-spline --- */
/* clamped uniform cubic B-spline, 5 control points, knots {0,0,0,0,.5,1,1,1,1} */
static void bspline(const double P[5][3], double u, double out[3])
{
    static const double kn[9] = {0, 0, 0, 0, 0.5, 1, 1, 1, 1};
    int k = (u < 0.5) ? 3 : 4;
    if (u >= 1.0) { u = 1.0; k = 4; }
    double d[4][3];
    for (int j = 0; j < 4; j++)
        for (int c = 0; c < 3; c++) d[j][c] = P[j + k - 3][c];
    for (int r = 1; r <= 3; r++) {
        for (int j = 3; j >= r; j--) {
            int i = j + k - 3;
            double den = kn[i + 4 - r] - kn[i];
            double a = den > 0 ? (u - kn[i]) / den : 0.0;
            for (int c = 0; c < 3; c++) d[j][c] = (1.0 - a) * d[j - 1][c] + a * d[j][c];
        }
    }
    for (int c = 0; c < 3; c++) out[c] = d[3][c];
}

/* ----------------------------------------------------------- per fibre -- */
typedef struct { uint64_t key; float len; } keyed_len;

static int cmp_key(const void *a, const void *b)
{
    uint64_t x = ((const keyed_len *)a)->key, y = ((const keyed_len *)b)->key;
    return (x > y) - (x < y);
}

static int in_mask(const synth_params *p, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= p->dim[0] || y >= p->dim[1] || z >= p->dim[2]) return 0;
    if (!p->mask) return 1;
    return p->mask[x + (size_t)p->dim[0] * (y + (size_t)p->dim[1] * z)] != 0;
}

static void walk_points(const synth_params *p, uint64_t *rs, double L, double P[5][3]);

static void control_points(const synth_params *p, uint64_t *rs, double P[5][3])
{
    if (p->mode == 0) {
        for (int k = 0; k < 5; k++)
            for (int c = 0; c < 3; c++) P[k][c] = urand(rs) * p->dim[c];
        return;
    }
    /* mode 1: centre drawn inside the mask, smooth random walk of total length L;
     * redrawn (up to 64 times) until all control points lie in the mask, so that
     * for a convex mask the whole curve (inside the control hull) does too */
    const double L = (p->len_min_mm + urand(rs) * (p->len_max_mm - p->len_min_mm)) / p->pixdim;
    for (int attempt = 0; attempt < 64; attempt++) {
        walk_points(p, rs, L, P);
        int ok = 1;
        for (int k = 0; k < 5 && ok; k++) ok = in_mask(p, (int)floor(P[k][0]), (int)floor(P[k][1]), (int)floor(P[k][2]));
        if (ok) return;
    }
}

static void walk_points(const synth_params *p, uint64_t *rs, double L, double P[5][3])
{
    double cen[3];
    for (int tries = 0;; tries++) {
        for (int c = 0; c < 3; c++) cen[c] = urand(rs) * p->dim[c];
        if (in_mask(p, (int)cen[0], (int)cen[1], (int)cen[2]) || tries > 1000) break;
    }
    double d[3], nrm = 0;
    for (int c = 0; c < 3; c++) { d[c] = nrand(rs); nrm += d[c] * d[c]; }
    nrm = sqrt(nrm);
    for (int c = 0; c < 3; c++) d[c] /= nrm;
    double pts[5][3];
    for (int c = 0; c < 3; c++) pts[0][c] = 0;
    for (int k = 1; k < 5; k++) {
        for (int c = 0; c < 3; c++) pts[k][c] = pts[k - 1][c] + 0.25 * L * d[c];
        double e[3], m = 0;
        for (int c = 0; c < 3; c++) { e[c] = d[c] + 0.35 * nrand(rs); m += e[c] * e[c]; }
        m = sqrt(m);
        for (int c = 0; c < 3; c++) d[c] = e[c] / m;
    }
    for (int c = 0; c < 3; c++) {
        double mid = pts[2][c];
        for (int k = 0; k < 5; k++) P[k][c] = pts[k][c] - mid + cen[c];
    }
}

/* Returns the number of (v,o) entries appended to buf for fibre id `f`. */
static size_t one_fibre(const synth_params *p, uint64_t stream, uint32_t f, seg_buf *buf,
                        keyed_len **tmp, size_t *tmpcap, double *norm, double *flen)
{
    uint64_t rs = stream;
    double P[5][3];
    control_points(p, &rs, P);
    double poly = 0;
    for (int k = 0; k < 4; k++) {
        double s = 0;
        for (int c = 0; c < 3; c++) s += (P[k + 1][c] - P[k][c]) * (P[k + 1][c] - P[k][c]);
        poly += sqrt(s);
    }
    int M = (int)ceil(poly / p->step_vox) + 1;
    if (M < 2) M = 2;
    if ((size_t)M > *tmpcap) {
        *tmpcap = (size_t)M * 2;
        *tmp = (keyed_len *)realloc(*tmp, *tmpcap * sizeof(keyed_len));
    }
    size_t cnt = 0;
    double a[3], b[3];
    bspline(P, 0.0, a);
    *flen = 0;
    for (int m = 1; m < M; m++) {
        bspline(P, (double)m / (M - 1), b);
        double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        int vx = (int)floor(0.5 * (a[0] + b[0]));
        int vy = (int)floor(0.5 * (a[1] + b[1]));
        int vz = (int)floor(0.5 * (a[2] + b[2]));
        double ln = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (ln > 1e-9 && in_mask(p, vx, vy, vz)) {
            if (d[1] < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; }
            double colat = atan2(sqrt(d[0] * d[0] + d[1] * d[1]), d[2]);
            double lon = atan2(d[1], d[0]);
            int ox = (int)lround(colat / M_PI * 180.0);
            int oy = (int)lround(lon / M_PI * 180.0);
            if (oy < 0) oy = 0;
            uint16_t o = (uint16_t)p->hash[ox * 181 + oy];
            uint64_t key = ((uint64_t)o << 48) | ((uint64_t)vz << 32) | ((uint64_t)vy << 16) | (uint64_t)vx;
            (*tmp)[cnt].key = key;
            (*tmp)[cnt].len = (float)(ln * p->pixdim);
            cnt++;
            *flen += ln * p->pixdim;
        }
        a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
    }
    if (cnt == 0) return 0;
    qsort(*tmp, cnt, sizeof(keyed_len), cmp_key);
    /* merge equal keys (std::map accumulation in trk2dictionary) */
    size_t u = 0;
    for (size_t i = 1; i < cnt; i++) {
        if ((*tmp)[i].key == (*tmp)[u].key) (*tmp)[u].len += (*tmp)[i].len;
        else (*tmp)[++u] = (*tmp)[i];
    }
    cnt = u + 1;
    if (buf->n + cnt > buf->cap) {
        size_t nc = (buf->cap ? buf->cap * 2 : 1u << 16);
        while (nc < buf->n + cnt) nc *= 2;
        buf->f = (uint32_t *)realloc(buf->f, nc * sizeof(uint32_t));
        buf->v = (uint32_t *)realloc(buf->v, nc * sizeof(uint32_t));
        buf->o = (uint16_t *)realloc(buf->o, nc * sizeof(uint16_t));
        buf->len = (float *)realloc(buf->len, nc * sizeof(float));
        buf->cap = nc;
    }
    /* fibre norm: per-voxel summed length, squared and summed */
    for (size_t i = 0; i < cnt; i++) {
        uint64_t k = (*tmp)[i].key;
        uint32_t x = (uint32_t)(k & 0xFFFF), y = (uint32_t)((k >> 16) & 0xFFFF), z = (uint32_t)((k >> 32) & 0xFFFF);
        buf->f[buf->n + i] = f;
        buf->v[buf->n + i] = x + (uint32_t)p->dim[0] * (y + (uint32_t)p->dim[1] * z);
        buf->o[buf->n + i] = (uint16_t)(k >> 48);
        buf->len[buf->n + i] = (*tmp)[i].len;
    }
    /* sum per voxel: sort a copy of voxel ids */
    double nrm = 0;
    {
        uint64_t *vk = (uint64_t *)malloc(cnt * sizeof(uint64_t));
        for (size_t i = 0; i < cnt; i++) vk[i] = ((uint64_t)buf->v[buf->n + i] << 32) | (uint32_t)i;
        for (size_t i = 1; i < cnt; i++) {          /* insertion sort: cnt is small */
            uint64_t t = vk[i];
            size_t j = i;
            while (j > 0 && vk[j - 1] > t) { vk[j] = vk[j - 1]; j--; }
            vk[j] = t;
        }
        size_t i = 0;
        while (i < cnt) {
            uint32_t vv = (uint32_t)(vk[i] >> 32);
            double s = 0;
            while (i < cnt && (uint32_t)(vk[i] >> 32) == vv) { s += buf->len[buf->n + (uint32_t)vk[i]]; i++; }
            nrm += s * s;
        }
        free(vk);
    }
    *norm = sqrt(nrm);
    buf->n += cnt;
    return cnt;
}

typedef struct {
    const synth_params *p;
    int f0, f1;               /* kept-fibre ids handled by this thread */
    seg_buf buf;
    float *norm, *flen;       /* global arrays, index by fibre id */
} thread_job;

static void *worker(void *arg)
{
    thread_job *j = (thread_job *)arg;
    keyed_len *tmp = NULL;
    size_t tmpcap = 0;
    for (int f = j->f0; f < j->f1; f++) {
        /* redraw until the fibre has at least one segment in the mask */
        for (uint64_t attempt = 0;; attempt++) {
            uint64_t st = j->p->seed * 0x9E3779B97F4A7C15ull ^ ((uint64_t)f << 20) ^ attempt;
            splitmix64(&st);
            double nrm, fl;
            const size_t before = j->buf.n;
            size_t c = one_fibre(j->p, st, (uint32_t)f, &j->buf, &tmp, &tmpcap, &nrm, &fl);
            /* keep fibres with segments and, in mode 1, an in-mask length inside
             * the requested range (the minFiberLen/maxFiberLen test of
             * trk2dictionary_c.cpp:517-518) */
            if (c > 0 && j->p->mode == 1 && (fl < j->p->len_min_mm || fl > j->p->len_max_mm)) {
                j->buf.n = before;
                c = 0;
            }
            if (c > 0) {
                j->norm[f] = (float)nrm;
                j->flen[f] = (float)fl;
                break;
            }
            if (attempt > 100000) { j->norm[f] = 1.0f; j->flen[f] = 0.0f; break; }
        }
    }
    free(tmp);
    return NULL;
}

/* Build the 181x181 (colatitude deg, longitude deg) -> nearest direction table. */
void synth_hash_table(const double *dirs, int ndirs, short *table)
{
    for (int ox = 0; ox <= 180; ox++) {
        for (int oy = 0; oy <= 180; oy++) {
            double th = ox * M_PI / 180.0, ph = oy * M_PI / 180.0;
            double u[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
            int best = 0;
            double bd = -1;
            for (int k = 0; k < ndirs; k++) {
                double d = fabs(u[0] * dirs[3 * k] + u[1] * dirs[3 * k + 1] + u[2] * dirs[3 * k + 2]);
                if (d > bd) { bd = d; best = k; }
            }
            table[ox * 181 + oy] = (short)best;
        }
    }
}

int synth_generate(const synth_params *p, synth_result *r)
{
    memset(r, 0, sizeof *r);
    int nt = p->nthreads < 1 ? 1 : (p->nthreads > 256 ? 256 : p->nthreads);
    if (nt > p->nF) nt = p->nF > 0 ? p->nF : 1;
    r->nF = p->nF;
    r->TRK_norm = (float *)calloc(p->nF > 0 ? p->nF : 1, sizeof(float));
    r->TRK_len = (float *)calloc(p->nF > 0 ? p->nF : 1, sizeof(float));
    r->TRK_lenTot = (float *)calloc(p->nF > 0 ? p->nF : 1, sizeof(float));
    thread_job *jobs = (thread_job *)calloc(nt, sizeof(thread_job));
    pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
    for (int t = 0; t < nt; t++) {
        jobs[t].p = p;
        jobs[t].f0 = (int)((int64_t)p->nF * t / nt);
        jobs[t].f1 = (int)((int64_t)p->nF * (t + 1) / nt);
        jobs[t].norm = r->TRK_norm;
        jobs[t].flen = r->TRK_len;
        if (pthread_create(&th[t], NULL, worker, &jobs[t]) != 0) worker(&jobs[t]), th[t] = 0;
    }
    for (int t = 0; t < nt; t++) if (th[t]) pthread_join(th[t], NULL);
    size_t n = 0;
    for (int t = 0; t < nt; t++) n += jobs[t].buf.n;
    r->n = n;
    r->IC_f = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    r->IC_v = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    r->IC_o = (uint16_t *)malloc((n ? n : 1) * sizeof(uint16_t));
    r->IC_len = (float *)malloc((n ? n : 1) * sizeof(float));
    size_t off = 0;
    for (int t = 0; t < nt; t++) {
        seg_buf *b = &jobs[t].buf;
        if (b->n) {
            memcpy(r->IC_f + off, b->f, b->n * sizeof(uint32_t));
            memcpy(r->IC_v + off, b->v, b->n * sizeof(uint32_t));
            memcpy(r->IC_o + off, b->o, b->n * sizeof(uint16_t));
            memcpy(r->IC_len + off, b->len, b->n * sizeof(float));
        }
        off += b->n;
        free(b->f); free(b->v); free(b->o); free(b->len);
    }
    memcpy(r->TRK_lenTot, r->TRK_len, (size_t)p->nF * sizeof(float));
    free(jobs);
    free(th);
    return 0;
}

void synth_free(synth_result *r)
{
    free(r->IC_f); free(r->IC_v); free(r->IC_o); free(r->IC_len);
    free(r->TRK_norm); free(r->TRK_len); free(r->TRK_lenTot);
    memset(r, 0, sizeof *r);
}

/* Stable counting sort of n keys in [0, nkeys): writes the permutation that
 * orders `keys` (ties keep input order) and the bucket offsets (nkeys+1).
 * The same order as numpy's stable argsort (kind='mergesort'). */
int synth_stable_argsort_u32(const uint32_t *keys, size_t n, uint32_t nkeys, uint32_t *perm, uint64_t *offs)
{
    memset(offs, 0, ((size_t)nkeys + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) {
        if (keys[i] >= nkeys) return -1;
        offs[keys[i] + 1]++;
    }
    for (size_t k = 0; k < nkeys; k++) offs[k + 1] += offs[k];
    uint64_t *pos = (uint64_t *)malloc(((size_t)nkeys + 1) * sizeof(uint64_t));
    if (!pos) return -2;
    memcpy(pos, offs, ((size_t)nkeys + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) perm[pos[keys[i]]++] = (uint32_t)i;
    free(pos);
    return 0;
}

/* ------------------------------------------------------ streamlines ---- */
/* The polylines themselves, for the tractogram voxeliser: streamline f is the
 * B-spline of fibre f's first draw (no redraw: the voxeliser applies the mask
 * and the length limits itself), sampled every step_vox voxels, with
 * coordinates in mm in voxmm space (voxel i spans [i, i+1) * pixdim).
 * Pass 1 (pts == NULL) fills counts[nF]; pass 2 fills pts at the offsets
 * ptr[nF+1] (float32 x, y, z). */
typedef struct {
    const synth_params *p;
    int f0, f1;
    int64_t *counts;
    const int64_t *ptr;
    float *pts;
} stream_job;

static void *stream_worker(void *arg)
{
    stream_job *j = (stream_job *)arg;
    const synth_params *p = j->p;
    for (int f = j->f0; f < j->f1; f++) {
        uint64_t st = p->seed * 0x9E3779B97F4A7C15ull ^ ((uint64_t)f << 20);
        splitmix64(&st);
        uint64_t rs = st;
        double P[5][3];
        control_points(p, &rs, P);
        double poly = 0;
        for (int k = 0; k < 4; k++) {
            double s = 0;
            for (int c = 0; c < 3; c++) s += (P[k + 1][c] - P[k][c]) * (P[k + 1][c] - P[k][c]);
            poly += sqrt(s);
        }
        int M = (int)ceil(poly / p->step_vox) + 1;
        if (M < 2) M = 2;
        if (M > 9999) M = 9999;
        if (!j->pts) {
            j->counts[f] = M;
            continue;
        }
        float *o = j->pts + 3 * j->ptr[f];
        for (int m = 0; m < M; m++) {
            double b[3];
            bspline(P, (double)m / (M - 1), b);
            for (int c = 0; c < 3; c++) o[3 * m + c] = (float)(b[c] * p->pixdim);
        }
    }
    return NULL;
}

int synth_streamlines(const synth_params *p, int64_t *counts, const int64_t *ptr, float *pts)
{
    int nt = p->nthreads < 1 ? 1 : (p->nthreads > 256 ? 256 : p->nthreads);
    if (nt > p->nF) nt = p->nF > 0 ? p->nF : 1;
    stream_job *jobs = (stream_job *)calloc(nt, sizeof(stream_job));
    pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
    if (!jobs || !th) { free(jobs); free(th); return -1; }
    for (int t = 0; t < nt; t++) {
        jobs[t].p = p;
        jobs[t].f0 = (int)((int64_t)p->nF * t / nt);
        jobs[t].f1 = (int)((int64_t)p->nF * (t + 1) / nt);
        jobs[t].counts = counts;
        jobs[t].ptr = ptr;
        jobs[t].pts = pts;
        if (pthread_create(&th[t], NULL, stream_worker, &jobs[t]) != 0) stream_worker(&jobs[t]), th[t] = 0;
    }
    for (int t = 0; t < nt; t++) if (th[t]) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
    return 0;
}
