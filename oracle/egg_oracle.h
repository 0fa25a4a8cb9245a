/*
 * CPU oracle for the EGG-SynC solve.  TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the algorithm of the reference
 * jakobrj/EGG-SynC for the hot path (data tensor, eps) -> labels:
 *   - EGG-SynC (src/algorithms/EGG_SynC.cu:28-1601): two-level grid, per-cell
 *     sin/cos summaries, full/partial neighbour update, local-sync flag, MBR
 *     extra check, cell-based labels;
 *   - CPU SynC (src/algorithms/SynC.cpp:16-129,268-399, src/utils/CPU_math.cpp:7-14):
 *     brute-force eps-neighbourhood update, order parameter r_c, synCluster_v2.
 *
 * It is the checker for the HIP implementation and the CPU baseline of the
 * benchmark.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product path never does.
 *
 * Floating-point conventions (FP32 restatement, suffix _f32):
 *   - every product and sum is rounded separately (built with
 *     -ffp-contract=off, no FMA), exactly as the reference source reads;
 *   - "sqrtf(s) <= e" is evaluated as "s <= thr(e)", where thr(e) is the
 *     largest float whose correctly rounded square root is <= e (same truth
 *     table, no sqrt);
 *   - the on-border test is done in double, as the reference does
 *     (EGG_SynC.cu:590-592);
 *   - per-cell sums are accumulated in double and rounded once to float.
 * The _f64 variants run the same algorithm entirely in double (the FP64
 * oracle used to report the error of the FP32 results).
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   eo_sync_solve_f32      bitwise against golden vectors written by the
 *                          reference's own SynC.cpp (oracle/build_ref.sh)
 *   eo_egg_iteration_f32   its position update against the reference's
 *                          ComputeNeighborhood+UpdatePoint (golden vectors),
 *                          against a brute-force FP64 update, and against
 *                          closed forms (two-point contraction d' = d - sin d);
 *                          its stop rule (N, half, border, synced,
 *                          EGG_SynC.cu:583-709) against a second, literal
 *                          restatement (FP32 sqrtf compares, no thresholds),
 *                          hand-worked edge cases and the iteration counts of
 *                          two-point solves (tests/test_stop_rule.py); planted
 *                          mistakes in it are caught
 *                          (tests/test_stop_rule_mutations.py)
 *   eo_build_grid_f32      invariants (bijection, cell boxes, sortedness,
 *                          counts) and brute-force grouping
 *   eo_egg_solve_*         closed-form known answers (single point, identical
 *                          points, separated clusters) and partition checks
 *   extra check            known-answer configurations in the
 *                          (eps, eps_extra] band
 */
#ifndef EGG_ORACLE_H
#define EGG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eo_geom {
    int64_t n;
    int32_t d;
    float eps;
    float cs;          /* inner cell side = outer cell side (EGG_SynC.cu:1249,1254) */
    int32_t w;         /* cells per dimension (EGG_SynC.cu:1250)                      */
    int32_t od;        /* outer grid dimensions (EGG_SynC.cu:1256)                    */
    int64_t O;         /* outer cells = w^od (EGG_SynC.cu:1257)                       */
    int32_t R;         /* ceil(eps/cs) (EGG_SynC.cu:563)                              */
    int32_t R_extra;   /* ceil(eps_extra/cs) (EGG_SynC.cu:779,911)                    */
    int32_t R_half;    /* ceil((eps/2.)/cs) (EGG_SynC.cu:1045)                        */
    float eps_extra;   /* EGG_SynC.cu:1430                                             */
    float thr_eps;     /* largest s with sqrtf(s) <= eps                               */
    float thr_half;    /* largest s with sqrtf(s) <= eps/2                             */
    float thr_extra;   /* largest s with sqrtf(s) <= eps_extra                         */
} eo_geom;

/* Grid geometry exactly as the reference computes it; returns 0 or -1. */
int eo_geometry(int64_t n, int32_t d, float eps, eo_geom *g);

/* Largest float s >= 0 such that sqrtf(s) <= e. */
float eo_sqrt_threshold(float e);

/* Grid build on positions X (n x d, row-major, indexed by point id) visited in
 * order order_in (NULL = identity).  Outputs, all indexed by sorted position
 * unless noted: order_out[n] (point id), cell_of[n], cell_start[C+1],
 * cell_g[C*d], ob[O], oe[O] (cell range of each outer cell), *C_out.
 * Sorting is a stable sort by (outer id, g_od, ..., g_{d-1}). */
int eo_build_grid_f32(const float *X, int64_t n, int32_t d, float eps,
                      const int32_t *order_in, int32_t *order_out, int32_t *cell_of,
                      int32_t *cell_start, int32_t *cell_g, int32_t *ob, int32_t *oe,
                      int32_t *C_out);

/* One EGG-SynC iteration (EGG_SynC.cu:1322-1491) from X_cur (point-id indexed)
 * into X_next.  order_in is the visiting order of the previous grid (NULL =
 * identity); order_out receives this iteration's sorted order.  Optional
 * per-sorted-position outputs N_out (neighbour count), half_out, border_out,
 * synced_out.  *flag_local = 1 if every point is locally synced;
 * *flag_final = flag after the extra check (= converged). */
int eo_egg_iteration_f32(const float *X_cur, int64_t n, int32_t d, float eps,
                         const int32_t *order_in, float *X_next, int32_t *order_out,
                         int32_t *N_out, int32_t *half_out, int32_t *border_out,
                         int32_t *synced_out, int32_t *flag_local, int32_t *flag_final);

/* Extra check alone (EGG_SynC.cu:767-1203) on positions X: returns 1 if no
 * (p,q) pair with eps < |p-q| <= eps_extra has its MBR within eps. */
int eo_extra_check_f32(const float *X, int64_t n, int32_t d, float eps);

/* Full EGG-SynC solve.  labels[n] (point id order; -1 = noise; clusters
 * numbered 0..k-1 in cell order), X_final[n*d] (may be NULL), *iters.
 * Returns 0 (converged), 1 (max_iter reached) or <0 on error. */
int eo_egg_solve_f32(const float *X0, int64_t n, int32_t d, float eps, int32_t max_iter,
                     int32_t *labels, float *X_final, int32_t *iters);
int eo_egg_solve_f64(const double *X0, int64_t n, int32_t d, double eps, int32_t max_iter,
                     int32_t *labels, double *X_final, int32_t *iters);

/* Cell-based labels of EGG_SynC.cu:1205-1227 on positions X. Returns #clusters. */
int eo_egg_labels_f32(const float *X, int64_t n, int32_t d, float eps, int32_t *labels);

/* CPU SynC, DynamicalClustering (SynC.cpp:268-326), serial FP32.  r_trace
 * (may be NULL, capacity max_iter) receives r_local per iteration.
 * Returns 0 (converged) or 1 (max_iter reached). */
int eo_sync_solve_f32(const float *X0, int64_t n, int32_t d, float eps, float lam,
                      int32_t max_iter, int32_t *labels, float *X_final, int32_t *iters,
                      float *r_trace);

/* One SynC iteration (SynC.cpp:281-295) for the target points [t0, t1),
 * OpenMP-parallel over targets with nthreads threads (0 = default).
 * Writes X_next rows of those targets and returns sum of r_p (FP32 per point). */
double eo_sync_iteration_range_f32(const float *X_cur, int64_t n, int32_t d, float eps,
                                   int64_t t0, int64_t t1, float *X_next, int32_t nthreads);

/* synCluster_v2 (SynC.cpp:88-129) on positions X. Returns #clusters. */
int eo_sync_labels_f32(const float *X, int64_t n, int32_t d, float eps, int32_t *labels);

/* Order parameter r_c of positions X: mean over p of mean over the FP32
 * eps-neighbourhood N_p of expf(-|x_j - x_p|) (SynC.cpp:42-56); terms in FP32,
 * all sums in double. */
double eo_rc_f64(const float *X, int64_t n, int32_t d, float eps);

#ifdef __cplusplus
}
#endif
#endif
