/*
 * CPU oracle for the structured-grid stencil sweep -- TEST INFRASTRUCTURE ONLY
 * (never linked into or called by the product library).  Semantics and the
 * reference lines each function follows are listed in uhs_oracle.h.
 *
 * Build: gcc -O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math -fopenmp
 * (oracle/Makefile).  -ffp-contract=off keeps ORC_PLAIN a separate multiply
 * and add per term, as in the reference's cpu_verify; ORC_FUSED calls fma()
 * explicitly.  The OpenMP split is over rows only, so the result of every
 * point is bitwise independent of the thread count.
 */
#include "uhs_oracle.h"
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t seed, uint64_t i)
{
    uint64_t z = orc_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

/* Normalised view of a grid as (planes, rows, cols) with halo h on every
 * axis; 2D grids are one plane with no z halo. */
typedef struct {
    int64_t n0, n1, n2;      /* interior extents (n0 == 1 for 2D) */
    int64_t h0, h;           /* halo on axis 0, halo on the other axes */
    int64_t sy, sz;          /* row and plane strides (elements) */
} view_t;

static int make_view(int dims, const int64_t *ext, int halo, view_t *v)
{
    if ((dims != 2 && dims != 3) || halo < 0 || !ext) return -1;
    for (int k = 0; k < dims; k++) if (ext[k] <= 0) return -1;
    v->h = halo;
    if (dims == 3) {
        v->n0 = ext[0]; v->n1 = ext[1]; v->n2 = ext[2]; v->h0 = halo;
    } else {
        v->n0 = 1; v->n1 = ext[0]; v->n2 = ext[1]; v->h0 = 0;
    }
    v->sy = v->n2 + 2 * halo;
    v->sz = (v->n1 + 2 * halo) * v->sy;
    return 0;
}

/* Linear offset of each term relative to the updated point. */
static int term_offsets(int dims, int P, const int *off, int halo,
                        const view_t *v, int64_t *lin)
{
    for (int t = 0; t < P; t++) {
        int64_t dz = 0, dy, dx;
        if (dims == 3) { dz = off[3*t]; dy = off[3*t+1]; dx = off[3*t+2]; }
        else { dy = off[2*t]; dx = off[2*t+1]; }
        if (llabs(dz) > halo || llabs(dy) > halo || llabs(dx) > halo) return -1;
        lin[t] = dz * v->sz + dy * v->sy + dx;
    }
    return 0;
}

#define SWEEP_BODY(T, FMA)                                                     \
    view_t v;                                                                  \
    if (P < 1 || !in || !out || !off || !coef) return -1;                      \
    if (make_view(dims, ext, halo, &v)) return -1;                             \
    int64_t *lin = (int64_t *)malloc(sizeof(int64_t) * P);                     \
    if (!lin) return -1;                                                       \
    if (term_offsets(dims, P, off, halo, &v, lin)) { free(lin); return -1; }   \
    if (nthreads <= 0) nthreads = omp_get_max_threads();                       \
    const int64_t nrows = v.n0 * v.n1;                                         \
    _Pragma("omp parallel num_threads(nthreads)")                              \
    {                                                                          \
        T *acc = (T *)malloc(sizeof(T) * v.n2);                                \
        T *cmp = (T *)malloc(sizeof(T) * v.n2);                                \
        _Pragma("omp for schedule(static)")                                    \
        for (int64_t row = 0; row < nrows; row++) {                            \
            const int64_t z = row / v.n1, y = row % v.n1;                      \
            const int64_t base = (z + v.h0) * v.sz + (y + v.h) * v.sy + v.h;   \
            const T *p0 = in + base + lin[0];                                  \
            const T c0 = coef[0];                                              \
            for (int64_t x = 0; x < v.n2; x++) { acc[x] = c0 * p0[x]; cmp[x] = 0; } \
            for (int t = 1; t < P; t++) {                                      \
                const T *pt = in + base + lin[t];                              \
                const T ct = coef[t];                                          \
                if (mode == ORC_PLAIN) {                                       \
                    for (int64_t x = 0; x < v.n2; x++) {                       \
                        T prod = ct * pt[x];                                   \
                        acc[x] = acc[x] + prod;                                \
                    }                                                          \
                } else if (mode == ORC_FUSED) {                                \
                    for (int64_t x = 0; x < v.n2; x++)                         \
                        acc[x] = FMA(ct, pt[x], acc[x]);                       \
                } else {                                                       \
                    for (int64_t x = 0; x < v.n2; x++) {                       \
                        T yv = FMA(ct, pt[x], -cmp[x]);                        \
                        T u = acc[x] + yv;                                     \
                        cmp[x] = (u - acc[x]) - yv;                            \
                        acc[x] = u;                                            \
                    }                                                          \
                }                                                              \
            }                                                                  \
            memcpy(out + base, acc, sizeof(T) * v.n2);                         \
        }                                                                      \
        free(acc); free(cmp);                                                  \
    }                                                                          \
    free(lin);                                                                 \
    return 0;

int orc_sweep_f64(const double *in, double *out, int dims, const int64_t *ext,
                  int halo, int P, const int *off, const double *coef,
                  int mode, int nthreads)
{
    if (mode != ORC_PLAIN && mode != ORC_FUSED) return -1;
    SWEEP_BODY(double, fma)
}

int orc_sweep_f32(const float *in, float *out, int dims, const int64_t *ext,
                  int halo, int P, const int *off, const float *coef,
                  int mode, int nthreads)
{
    if (mode != ORC_PLAIN && mode != ORC_FUSED && mode != ORC_KAHAN) return -1;
    SWEEP_BODY(float, fmaf)
}

#define ITERATE_BODY(SWEEP)                                                    \
    if (T < 1 || !A || !B) return -1;                                          \
    for (int k = 1; k <= T; k++) {                                             \
        int rc = (k & 1) ? SWEEP(A, B, dims, ext, halo, P, off, coef, mode, nthreads) \
                         : SWEEP(B, A, dims, ext, halo, P, off, coef, mode, nthreads); \
        if (rc) return rc;                                                     \
    }                                                                          \
    return 0;

int orc_iterate_f64(double *A, double *B, int dims, const int64_t *ext,
                    int halo, int P, const int *off, const double *coef,
                    int T, int mode, int nthreads)
{
    ITERATE_BODY(orc_sweep_f64)
}

int orc_iterate_f32(float *A, float *B, int dims, const int64_t *ext,
                    int halo, int P, const int *off, const float *coef,
                    int T, int mode, int nthreads)
{
    ITERATE_BODY(orc_sweep_f32)
}

#define FILL_BODY(T)                                                           \
    view_t v;                                                                  \
    if (make_view(dims, ext, halo, &v)) return;                                \
    const int64_t planes = v.n0 + 2 * v.h0, rows = v.n1 + 2 * v.h;             \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int64_t zr = 0; zr < planes * rows; zr++) {                           \
        const int64_t z = zr / rows, y = zr % rows;                            \
        const int zin = z >= v.h0 && z < v.h0 + v.n0;                          \
        const int yin = y >= v.h && y < v.h + v.n1;                            \
        T *row = a + z * v.sz + y * v.sy;                                      \
        for (int64_t x = 0; x < v.sy; x++) {                                   \
            const int in = zin && yin && x >= v.h && x < v.h + v.n2;           \
            row[x] = in ? (T)u01(seed, (uint64_t)(z * v.sz + y * v.sy + x)) : (T)0; \
        }                                                                      \
    }

void orc_fill_uniform_f64(double *a, int dims, const int64_t *ext, int halo,
                          uint64_t seed)
{
    FILL_BODY(double)
}

void orc_fill_uniform_f32(float *a, int dims, const int64_t *ext, int halo,
                          uint64_t seed)
{
    FILL_BODY(float)
}

void orc_make_coeffs(int dims, int P, const int *off, uint64_t seed,
                     double *coef)
{
    uint64_t draws = 0;
    for (int t = 0; t < P; t++) {
        int partner = -1;
        for (int s = 0; s < t && partner < 0; s++) {
            int mirror = 1;
            for (int k = 0; k < dims; k++)
                if (off[s * dims + k] != -off[t * dims + k]) mirror = 0;
            if (mirror) partner = s;
        }
        coef[t] = partner >= 0 ? coef[partner] : u01(seed, draws++);
    }
    double sum = 0.0;
    for (int t = 0; t < P; t++) sum += coef[t];
    for (int t = 0; t < P; t++) coef[t] /= sum;
}

void orc_fill_rand50_f64(double *a, size_t n)
{
    for (size_t i = 0; i < n; i++) a[i] = rand() % 50;
}

/* ---- fused multi-kernel programs ---------------------------------------- */

/* Left-to-right sum step in `mode`: s += c*v (PLAIN: separate product and
 * add; FUSED: one fused multiply-add; KAHAN: the compensated form of the
 * FP32 sweep, e carrying the compensation). */
#define PROG_ADD(FMA, mode, s, e, c, v)                                        \
    do {                                                                       \
        if ((mode) == ORC_PLAIN) (s) = (s) + (c) * (v);                        \
        else if ((mode) == ORC_FUSED) (s) = FMA((c), (v), (s));                \
        else {                                                                 \
            const __typeof__(s) y_ = FMA((c), (v), -(e));                      \
            const __typeof__(s) u_ = (s) + y_;                                 \
            (e) = (u_ - (s)) - y_;                                             \
            (s) = u_;                                                          \
        }                                                                      \
    } while (0)

/* The value of producer k at linear cell q: its terms left to right, each
 * reading the input grid or an earlier producer evaluated at the shifted
 * cell (the reference's inlining at every access offset). */
#define PROG_VALUE_FN(T, FMA, NAME)                                            \
    struct NAME##_ctx {                                                        \
        const T *in;                                                           \
        const int *pstart, *psrc;                                              \
        const long long *poff;                                                 \
        const T *pcoef;                                                        \
        int mode;                                                              \
    };                                                                         \
    static T NAME(const struct NAME##_ctx *c, int k, long long q)              \
    {                                                                          \
        T s = 0, e = 0;                                                        \
        for (int t = c->pstart[k]; t < c->pstart[k + 1]; t++) {                \
            const int src = c->psrc ? c->psrc[t] : -1;                         \
            const T v = src < 0 ? c->in[q + c->poff[t]] : NAME(c, src, q + c->poff[t]); \
            if (t == c->pstart[k]) s = c->pcoef[t] * v;                        \
            else PROG_ADD(FMA, c->mode, s, e, c->pcoef[t], v);                 \
        }                                                                      \
        return s;                                                              \
    }

PROG_VALUE_FN(double, fma, prog_value_f64)
PROG_VALUE_FN(float, fmaf, prog_value_f32)

#define PROG_BODY(T, FMA, NAME)                                                \
    if (!A || !B || !shape || (dims != 2 && dims != 3) || np < 1 || !off || !coef || !src || \
        nprod < 0 || (nprod && (!prod_np || !prod_off || !prod_coef)) || !lo || !hi || T_ < 0) \
        return -1;                                                             \
    /* 3D view: a 2D grid is one plane. */                                     \
    long long shp[3] = {1, 1, 1}, l[3] = {0, 0, 0}, h[3] = {1, 1, 1};          \
    for (int k = 0; k < dims; k++) {                                           \
        shp[3 - dims + k] = shape[k];                                          \
        l[3 - dims + k] = lo[k];                                               \
        h[3 - dims + k] = hi[k];                                               \
        if (lo[k] < 0 || hi[k] > shape[k] || lo[k] >= hi[k]) return -1;       \
    }                                                                          \
    const long long sy = shp[2], sz = shp[1] * shp[2];                         \
    int ntot = 0;                                                              \
    for (int k = 0; k < nprod; k++) {                                          \
        if (prod_np[k] < 1) return -1;                                         \
        ntot += prod_np[k];                                                    \
    }                                                                          \
    long long *poff = (long long *)malloc(sizeof(long long) * (ntot + 1));     \
    long long *ooff = (long long *)malloc(sizeof(long long) * np);             \
    int *pstart = (int *)malloc(sizeof(int) * (nprod + 1));                    \
    long long *reach = (long long *)malloc(sizeof(long long) * 6 * (nprod + 1)); \
    if (!poff || !ooff || !pstart || !reach) {                                 \
        free(poff); free(ooff); free(pstart); free(reach);                     \
        return -1;                                                             \
    }                                                                          \
    int rc = 0, t = 0;                                                         \
    /* reach[k]: input offsets producer k reads, per axis [lo, hi]; a          \
     * producer may read only earlier producers. */                            \
    for (int k = 0; k < nprod && !rc; k++) {                                   \
        pstart[k] = t;                                                         \
        long long *r = reach + 6 * k;                                          \
        for (int a = 0; a < 3; a++) { r[a] = LLONG_MAX; r[3 + a] = LLONG_MIN; } \
        for (int i = 0; i < prod_np[k]; i++, t++) {                            \
            long long v[3] = {0, 0, 0};                                        \
            for (int a = 0; a < dims; a++) v[3 - dims + a] = prod_off[t * dims + a]; \
            poff[t] = v[0] * sz + v[1] * sy + v[2];                            \
            const int s = prod_src ? prod_src[t] : -1;                         \
            if (s < -1 || s >= k) { rc = -1; break; }                          \
            for (int a = 0; a < 3; a++) {                                      \
                const long long rl = s < 0 ? v[a] : v[a] + reach[6 * s + a];   \
                const long long rh = s < 0 ? v[a] : v[a] + reach[6 * s + 3 + a]; \
                if (rl < r[a]) r[a] = rl;                                      \
                if (rh > r[3 + a]) r[3 + a] = rh;                              \
            }                                                                  \
        }                                                                      \
    }                                                                          \
    pstart[nprod] = t;                                                         \
    /* Reads must stay inside the grid for every written cell. */              \
    for (int j = 0; j < np && !rc; j++) {                                      \
        if (src[j] < -1 || src[j] >= nprod) { rc = -1; break; }                \
        long long v[3] = {0, 0, 0};                                            \
        for (int a = 0; a < dims; a++) v[3 - dims + a] = off[j * dims + a];    \
        ooff[j] = v[0] * sz + v[1] * sy + v[2];                                \
        for (int a = 0; a < 3; a++) {                                          \
            const long long rl = src[j] < 0 ? v[a] : v[a] + reach[6 * src[j] + a]; \
            const long long rh = src[j] < 0 ? v[a] : v[a] + reach[6 * src[j] + 3 + a]; \
            if (l[a] + rl < 0 || h[a] - 1 + rh >= shp[a]) rc = -1;             \
        }                                                                      \
    }                                                                          \
    for (int step = 1; step <= T_ && !rc; step++) {                            \
        const T *in = (step & 1) ? A : B;                                      \
        T *out = (step & 1) ? B : A;                                           \
        struct NAME##_ctx c = {in, pstart, prod_src, poff, prod_coef, mode};   \
        for (long long z = l[0]; z < h[0]; z++)                                \
            for (long long y = l[1]; y < h[1]; y++)                            \
                for (long long x = l[2]; x < h[2]; x++) {                      \
                    const long long q = z * sz + y * sy + x;                   \
                    T acc = 0, e = 0;                                          \
                    for (int j = 0; j < np; j++) {                             \
                        const T xv = src[j] < 0 ? in[q + ooff[j]] : NAME(&c, src[j], q + ooff[j]); \
                        if (j == 0) acc = coef[0] * xv;                        \
                        else PROG_ADD(FMA, mode, acc, e, coef[j], xv);         \
                    }                                                          \
                    out[q] = acc;                                              \
                }                                                              \
    }                                                                          \
    free(poff); free(ooff); free(pstart); free(reach);                         \
    return rc;

int orc_program_f64(double *A, double *B, int dims, const int64_t *shape,
                    int nprod, const int *prod_np, const int *prod_off, const double *prod_coef,
                    const int *prod_src,
                    int np, const int *off, const double *coef, const int *src,
                    const int64_t *lo, const int64_t *hi, int T_, int mode)
{
    if (mode != ORC_PLAIN && mode != ORC_FUSED) return -1;
    PROG_BODY(double, fma, prog_value_f64)
}

int orc_program_f32(float *A, float *B, int dims, const int64_t *shape,
                    int nprod, const int *prod_np, const int *prod_off, const float *prod_coef,
                    const int *prod_src,
                    int np, const int *off, const float *coef, const int *src,
                    const int64_t *lo, const int64_t *hi, int T_, int mode)
{
    if (mode != ORC_PLAIN && mode != ORC_FUSED && mode != ORC_KAHAN) return -1;
    PROG_BODY(float, fmaf, prog_value_f32)
}
