/*
 * CPU oracle for the EGG-SynC solve.  TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
 * checker / baseline, never by the product path.  See egg_oracle.h for the
 * conventions and for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no -ffast-math).
 */
#include "egg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EO_MAXD 64

float eo_sqrt_threshold(float e) {
    if (!(e >= 0.0f) || isinf(e)) return e;
    float s = e * e;
    if (isinf(s)) s = FLT_MAX;
    while (s > 0.0f && sqrtf(s) > e) s = nextafterf(s, 0.0f);
    for (;;) {
        float nx = nextafterf(s, INFINITY);
        if (isinf(nx) || sqrtf(nx) > e) break;
        s = nx;
    }
    return s;
}

/* Grid geometry, EGG_SynC.cu:1249-1260 and :563,779,1045,1430. */
int eo_geometry(int64_t n, int32_t d, float eps, eo_geom *g) {
    if (d < 1 || d > EO_MAXD || n < 0 || !(eps > 0.0f) || !isfinite(eps)) return -1;
    memset(g, 0, sizeof *g);
    g->n = n;
    g->d = d;
    g->eps = eps;
    g->cs = (float)sqrt(pow((double)eps / 2., 2.) / d);          /* :1249 */
    float inv = 1 / g->cs;                                         /* :1250 */
    int w = inv >= 2147483647.0f ? 2147483647 : (int)inv;
    if (w < 1) w = 1;
    g->w = w;
    int od;                                                        /* :1256 */
    if (w == 1 || n <= 1) {
        od = 1;
    } else {
        od = (int)(log((double)n) / log((double)w));
        if (od > d) od = d;
        if (od < 1) od = 1;
    }
    g->od = od;
    int64_t O = 1;                                                 /* :1257 */
    for (int l = 0; l < od; l++) O *= w;
    g->O = O;
    g->R = (int)ceilf(eps / g->cs);                                /* :563 */
    float e2 = 2 * eps;                                            /* :1430 */
    double ex = (double)e2 - (double)eps * sqrt(15. / 16.) + (double)eps / 2. - sin((double)eps / 2.);
    g->eps_extra = (float)ex;
    g->R_extra = (int)ceilf(g->eps_extra / g->cs);                 /* :779 */
    g->R_half = (int)ceil(((double)eps / 2.) / (double)g->cs);     /* :1045 */
    g->thr_eps = eo_sqrt_threshold(eps);
    g->thr_half = eo_sqrt_threshold(eps / 2);
    g->thr_extra = eo_sqrt_threshold(g->eps_extra);
    return 0;
}

/* ---------------------------------------------------------------- FP32 -- */

/* compute_cell_id per coordinate (:67-74): FP32 division, truncation, clamp
 * of w to w-1 (the reference leaves other out-of-range values undefined). */
static inline int coord_f32(float x, float cs, int w) {
    float q = x / cs;
    if (!(q > -1.0f)) return 0;
    if (q >= (float)w) return w - 1;
    int t = (int)q;
    return t >= w ? w - 1 : t;
}

/* On-border test (:590-595): x +- 1e-5 crosses a cell boundary, in double. */
static inline int border_f32(float x, float cs) {
    double xd = (double)x;
    int idm = (int)((xd - 0.00001) / (double)cs);
    int id = (int)(x / cs);
    int idp = (int)((xd + 0.00001) / (double)cs);
    return id != idm || id != idp;
}

#define REAL float
#define SUF f32
#define SINR sinf
#define COSR cosf
#define COORD(x, G) coord_f32((x), (G)->cs, (G)->w)
#define BORDER_DIM(x, G) border_f32((x), (G)->cs)
#define LE_EPS(s) ((s) <= geo->thr_eps)
#define LE_HALF(s) ((s) <= geo->thr_half)
#define LE_EXTRA(s) ((s) <= geo->thr_extra)
#include "egg_core.inc"
#undef REAL
#undef SUF
#undef SINR
#undef COSR
#undef COORD
#undef BORDER_DIM
#undef LE_EPS
#undef LE_HALF
#undef LE_EXTRA

static int bitlen64(int64_t v) {
    int b = 0;
    while (v > 0) { b++; v >>= 1; }
    return b < 1 ? 1 : b;
}

/* The GPU's sort key: cell fields, then d split bits if they fit 256 bits,
 * then up to 16 bits of position hash in the room left in the key's last
 * 64-bit word. */
static int key_bits_of(const eo_geom *g) {
    return bitlen64(g->O - 1) + (g->d - g->od) * bitlen64((int64_t)g->w - 1);
}

static int sub_cells_on(const eo_geom *g) { return key_bits_of(g) + g->d <= 256; }

static int hash_bits_of(const eo_geom *g) {
    int used = key_bits_of(g) + (sub_cells_on(g) ? g->d : 0);
    int room = (used + 63) / 64 * 64 - used;
    return room < 16 ? room : 16;
}

static void geo_f32_from(const eo_geom *g, geo_t_f32 *G) {
    G->n = g->n; G->d = g->d; G->w = g->w; G->od = g->od; G->R = g->R;
    G->R_extra = g->R_extra; G->R_half = g->R_half; G->O = g->O;
    G->eps = g->eps; G->cs = g->cs; G->eps_extra = g->eps_extra;
    G->thr_eps = g->thr_eps; G->thr_half = g->thr_half; G->thr_extra = g->thr_extra;
    G->sub = sub_cells_on(g);
    G->hash_bits = hash_bits_of(g);
}

/* ---------------------------------------------------------------- FP64 -- */

static inline int coord_f64(double x, double cs, int w) {
    double q = x / cs;
    if (!(q > -1.0)) return 0;
    if (q >= (double)w) return w - 1;
    int t = (int)q;
    return t >= w ? w - 1 : t;
}

static inline int border_f64(double x, double cs) {
    int idm = (int)((x - 0.00001) / cs);
    int id = (int)(x / cs);
    int idp = (int)((x + 0.00001) / cs);
    return id != idm || id != idp;
}

#define REAL double
#define SUF f64
#define SINR sin
#define COSR cos
#define COORD(x, G) coord_f64((x), (G)->cs, (G)->w)
#define BORDER_DIM(x, G) border_f64((x), (G)->cs)
#define LE_EPS(s) (sqrt(s) <= geo->eps)
#define LE_HALF(s) (sqrt(s) <= geo->eps / 2.)
#define LE_EXTRA(s) (sqrt(s) <= geo->eps_extra)
#include "egg_core.inc"
#undef REAL
#undef SUF
#undef SINR
#undef COSR
#undef COORD
#undef BORDER_DIM
#undef LE_EPS
#undef LE_HALF
#undef LE_EXTRA

static int geo_f64_make(int64_t n, int d, double eps, geo_t_f64 *G) {
    if (d < 1 || d > EO_MAXD || n < 0 || !(eps > 0.0) || !isfinite(eps)) return -1;
    memset(G, 0, sizeof *G);
    G->n = n;
    G->d = d;
    G->eps = eps;
    G->cs = sqrt((eps / 2.) * (eps / 2.) / d);
    double inv = 1. / G->cs;
    int w = inv >= 2147483647.0 ? 2147483647 : (int)inv;
    if (w < 1) w = 1;
    G->w = w;
    int od;
    if (w == 1 || n <= 1) {
        od = 1;
    } else {
        od = (int)(log((double)n) / log((double)w));
        if (od > d) od = d;
        if (od < 1) od = 1;
    }
    G->od = od;
    int64_t O = 1;
    for (int l = 0; l < od; l++) O *= w;
    G->O = O;
    G->R = (int)ceil(eps / G->cs);
    G->eps_extra = 2 * eps - eps * sqrt(15. / 16.) + eps / 2. - sin(eps / 2.);
    G->R_extra = (int)ceil(G->eps_extra / G->cs);
    G->R_half = (int)ceil((eps / 2.) / G->cs);
    return 0;
}

/* ------------------------------------------------------------ public API -- */

int eo_build_grid_f32(const float *X, int64_t n, int32_t d, float eps, const int32_t *order_in,
                      int32_t *order_out, int32_t *cell_of, int32_t *cell_start, int32_t *cell_g,
                      int32_t *ob, int32_t *oe, int32_t *C_out) {
    eo_geom g;
    if (eo_geometry(n, d, eps, &g) != 0) return -1;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    grid_t_f32 G;
    if (build_grid_f32(&geo, X, order_in, &G) != 0) return -2;
    if (order_out) memcpy(order_out, G.order, sizeof(int32_t) * (size_t)n);
    if (cell_of) memcpy(cell_of, G.cell_of, sizeof(int32_t) * (size_t)n);
    if (cell_start) memcpy(cell_start, G.cell_start, sizeof(int32_t) * (size_t)(G.C + 1));
    if (cell_g) memcpy(cell_g, G.cell_g, sizeof(int32_t) * (size_t)G.C * (size_t)d);
    if (ob) memcpy(ob, G.ob, sizeof(int32_t) * (size_t)g.O);
    if (oe) memcpy(oe, G.oe, sizeof(int32_t) * (size_t)g.O);
    if (C_out) *C_out = G.C;
    grid_free_f32(&G);
    return 0;
}

int eo_egg_iteration_f32(const float *X_cur, int64_t n, int32_t d, float eps,
                         const int32_t *order_in, float *X_next, int32_t *order_out,
                         int32_t *N_out, int32_t *half_out, int32_t *border_out,
                         int32_t *synced_out, int32_t *flag_local, int32_t *flag_final) {
    eo_geom g;
    if (eo_geometry(n, d, eps, &g) != 0) return -1;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    return iteration_f32(&geo, X_cur, order_in, X_next, order_out, N_out, half_out, border_out,
                         synced_out, flag_local, flag_final);
}

int eo_extra_check_f32(const float *X, int64_t n, int32_t d, float eps) {
    eo_geom g;
    if (eo_geometry(n, d, eps, &g) != 0) return -1;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    grid_t_f32 G;
    if (build_grid_f32(&geo, X, NULL, &G) != 0) return -2;
    cell_summaries_f32(&geo, X, &G);
    int r = extra_check_f32(&geo, X, &G);
    grid_free_f32(&G);
    return r;
}

int eo_egg_solve_f32(const float *X0, int64_t n, int32_t d, float eps, int32_t max_iter,
                     int32_t *labels, float *X_final, int32_t *iters) {
    eo_geom g;
    if (eo_geometry(n, d, eps, &g) != 0) return -1;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    return solve_f32(&geo, X0, max_iter, labels, X_final, iters);
}

int eo_egg_solve_f64(const double *X0, int64_t n, int32_t d, double eps, int32_t max_iter,
                     int32_t *labels, double *X_final, int32_t *iters) {
    geo_t_f64 geo;
    if (geo_f64_make(n, d, eps, &geo) != 0) return -1;
    return solve_f64(&geo, X0, max_iter, labels, X_final, iters);
}

int eo_egg_labels_f32(const float *X, int64_t n, int32_t d, float eps, int32_t *labels) {
    eo_geom g;
    if (eo_geometry(n, d, eps, &g) != 0) return -1;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    grid_t_f32 G;
    if (build_grid_f32(&geo, X, NULL, &G) != 0) return -2;
    int k = labels_from_grid_f32(&G, labels);
    grid_free_f32(&G);
    return k;
}

/* ------------------------------------------------------------- CPU SynC -- */

/* distance (CPU_math.cpp:7-14): FP32 sequential sum of squares, sqrtf. */
static inline float distance_f32(const float *x, const float *y, int d) {
    float dist = 0.f;
    for (int l = 0; l < d; l++) {
        float diff = x[l] - y[l];
        dist += diff * diff;
    }
    return sqrtf(dist);
}

/* One point of a SynC pass: ComputeNeighborhood (SynC.cpp:16-28),
 * UpdatePoint (:30-40), ComputeLocationOrder (:42-56).  nb is scratch [n]. */
static float sync_point_f32(const float *D, int64_t n, int d, float eps, int64_t p, float *D_next,
                            int32_t *nb) {
    const float *x = &D[p * d];
    int64_t m = 0;
    for (int64_t j = 0; j < n; j++)
        if (distance_f32(x, &D[j * d], d) <= eps) nb[m++] = (int32_t)j;
    for (int l = 0; l < d; l++) {
        D_next[p * d + l] = D[p * d + l];
        float sum = 0.f;
        for (int64_t k = 0; k < m; k++) sum += sinf(D[(int64_t)nb[k] * d + l] - D[p * d + l]);
        D_next[p * d + l] += sum / (float)m;
    }
    if (m == 0) return 1.f;
    float sum = 0.f;
    for (int64_t k = 0; k < m; k++) sum += expf(-fabsf(distance_f32(x, &D[(int64_t)nb[k] * d], d)));
    return sum / (float)m;
}

int eo_sync_labels_f32(const float *X, int64_t n, int32_t d, float eps, int32_t *C) {
    for (int64_t i = 0; i < n; i++) C[i] = -1;
    int cl = 0;
    for (int64_t i = 0; i < n; i++) {
        if (C[i] != -1) continue;
        int in_cluster = 0;
        for (int64_t j = 0; j < n; j++) {
            if (i != j && distance_f32(&X[i * d], &X[j * d], d) <= eps) {
                C[j] = cl;
                in_cluster = 1;
            }
        }
        if (in_cluster) {
            C[i] = cl;
            cl += 1;
        }
    }
    return cl;
}

/* GPU_synCluster + rename_cluster (simple_GPU_SynC.cu:70-102): a point with
 * more than one eps-neighbour (itself included; FP32 distance of
 * CPU_math.cpp:7-14 / simple_GPU_SynC.cu:76-82) gets the smallest index among
 * them; labels are those representatives numbered 0..k-1 in index order
 * (inclusive scan of the representative flags, minus one). */
int eo_simple_labels_f32(const float *X, int64_t n, int32_t d, float eps, int32_t *C) {
    uint8_t *incl = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t *map = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!incl || !map) { free(incl); free(map); return -2; }
    for (int64_t p = 0; p < n; p++) {
        int64_t mn = p, count = 0;
        for (int64_t q = 0; q < n; q++) {
            if (distance_f32(&X[p * d], &X[q * d], d) <= eps) {
                count++;
                if (q < mn) mn = q;
            }
        }
        C[p] = -1;
        if (count > 1) {
            C[p] = (int32_t)mn;
            incl[mn] = 1;
        }
    }
    int32_t run = 0;
    for (int64_t i = 0; i < n; i++) {
        run += incl[i];
        map[i] = run;
    }
    for (int64_t p = 0; p < n; p++)
        if (C[p] >= 0) C[p] = map[C[p]] - 1;
    free(incl);
    free(map);
    return run;
}

/* DynamicalClustering (SynC.cpp:268-326). */
int eo_sync_solve_f32(const float *X0, int64_t n, int32_t d, float eps, float lam,
                      int32_t max_iter, int32_t *labels, float *X_final, int32_t *iters,
                      float *r_trace) {
    if (d < 1 || n < 0) return -1;
    size_t nd = (size_t)(n > 0 ? n : 1) * (size_t)d;
    float *A = (float *)malloc(sizeof(float) * nd);
    float *B = (float *)malloc(sizeof(float) * nd);
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!A || !B || !nb) { free(A); free(B); free(nb); return -2; }
    memcpy(A, X0, sizeof(float) * (size_t)n * (size_t)d);
    memcpy(B, X0, sizeof(float) * (size_t)n * (size_t)d);
    float r_local = 0.f;
    int it = 0, status = 0;
    while (r_local < lam) {
        if (max_iter > 0 && it >= max_iter) { status = 1; break; }
        r_local = 0.f;
        for (int64_t p = 0; p < n; p++) r_local += sync_point_f32(A, n, d, eps, p, B, nb);
        r_local /= (float)n;
        if (r_trace) r_trace[it] = r_local;
        it++;
        float *t = B; B = A; A = t;
    }
    if (labels) eo_sync_labels_f32(A, n, d, eps, labels);
    if (X_final) memcpy(X_final, A, sizeof(float) * (size_t)n * (size_t)d);
    if (iters) *iters = it;
    free(A); free(B); free(nb);
    return status;
}

/* Targets [t0,t1) of one SynC pass, OpenMP over targets (SynC.cpp:352-361). */
double eo_sync_iteration_range_f32(const float *X_cur, int64_t n, int32_t d, float eps,
                                   int64_t t0, int64_t t1, float *X_next, int32_t nthreads) {
    double total = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : total)
    {
        int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t p = t0; p < t1; p++) total += (double)sync_point_f32(X_cur, n, d, eps, p, X_next, nb);
        free(nb);
    }
    return total;
}

/* r_c over positions X with the FP32 eps-membership of SynC.cpp:16-28:
 * per-pair terms expf(-|x_j - x_p|) as in SynC.cpp:48, all sums in double.
 * Candidates come from the FP32 grid with Chebyshev radius R+1 in every
 * dimension, which contains every point whose FP32 distance is <= eps. */
double eo_rc_f64(const float *X, int64_t n, int32_t d, float eps) {
    eo_geom g;
    if (n <= 0 || eo_geometry(n, d, eps, &g) != 0) return NAN;
    geo_t_f32 geo;
    geo_f32_from(&g, &geo);
    grid_t_f32 G;
    if (build_grid_f32(&geo, X, NULL, &G) != 0) return NAN;
    int RR = geo.R + 1;
    double total = 0.0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        const float *xp = &X[(int64_t)G.order[i] * d];
        int c0 = G.cell_of[i];
        int go[EO_MAXD];
        for (int l = 0; l < geo.od; l++) go[l] = G.cell_g[(int64_t)c0 * d + l];
        const geo_t_f32 *gp = &geo;
        double s = 0.0;
        int64_t m = 0;
        OUTER_LOOP_BEGIN(gp, go, RR, o)
        for (int32_t c = G.ob[o]; c < G.oe[o]; c++) {
            int skip = 0;
            for (int l = geo.od; l < d; l++) {
                int dg = G.cell_g[(int64_t)c * d + l] - G.cell_g[(int64_t)c0 * d + l];
                if (dg > RR || dg < -RR) skip = 1;
            }
            if (skip) continue;
            for (int32_t j = G.cell_start[c]; j < G.cell_start[c + 1]; j++) {
                float dist = distance_f32(xp, &X[(int64_t)G.order[j] * d], d);
                if (dist <= eps) {
                    s += (double)expf(-fabsf(dist));
                    m++;
                }
            }
        }
        OUTER_LOOP_END(gp, RR)
        total += m > 0 ? s / (double)m : 1.0;
    }
    grid_free_f32(&G);
    return total / (double)n;
}
