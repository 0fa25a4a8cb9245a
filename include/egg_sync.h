/*
 * egg_sync.h -- C ABI of the MI355X (gfx950) EGG-SynC solver.
 *
 * Plain C types only (no HIP or torch types): a HIP stream is passed as
 * `void *` (hipStream_t), device buffers as plain pointers.  Every entry point
 * returns 0 on success, a positive status for a completed solve that hit the
 * iteration cap, or a negative EGG_SYNC_E* error code; nothing calls exit().
 *
 * Reference interfaces replaced (jakobrj/EGG-SynC):
 *   egg_sync_solve           std::vector<at::Tensor> EGG_SynC(float *h_D, int n, int d, float eps)
 *                            (src/algorithms/EGG_SynC.cuh:11, src/algorithms/EGG_SynC.cu:1229),
 *                            reached from Python through EGG_SynC_wrapper (src/maps/map.cpp:31-33)
 *                            and python/algorithms.py:110-111.  With stop_rule
 *                            EGG_SYNC_STOP_RC it also replaces DynamicalClustering
 *                            (src/algorithms/SynC.h:38, src/maps/map.cpp:16-18) on the GPU.
 *   egg_sync_solve_device    the same solve on device-resident data (no reference counterpart:
 *                            the reference always copies a host array in, EGG_SynC.cu:1245).
 *   egg_sync_build_grid_device   build_the_grid (EGG_SynC.cu:248-341)
 *   egg_sync_iteration_device    one pass of the loop body (EGG_SynC.cu:1322-1491)
 *   egg_sync_rc_device           order parameter r_c (SynC.cpp:42-56, FP64 sums)
 *   egg_sync_solve_batch_device  many EGG_SynC calls at once (run_experiment.py:294-331 loops
 *                                over datasets and eps values one call at a time)
 *   egg_sync_simple_solve[_device]  std::vector<at::Tensor> GPU_DynamicalClustering(float *h_D,
 *                            int n, int d, float eps, float lam) (src/algorithms/simple_GPU_SynC.cuh:11,
 *                            simple_GPU_SynC.cu:105), reached through GPU_DynamicalClustering_wrapper
 *                            (src/maps/map.cpp:27-29) and simple_GPU_SynC (python/algorithms.py:106-107)
 *   egg_sync_simple_labels_device   GPU_synCluster + rename_cluster (simple_GPU_SynC.cu:70-102)
 * Threading: one solve at a time per process (the library caches one
 * workspace per device, like the reference's global allocation map,
 * GPU_utils.cu:176-178); calls are synchronous with respect to the host.
 */
#ifndef EGG_SYNC_H
#define EGG_SYNC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EGG_SYNC_OK 0
#define EGG_SYNC_CAPPED 1           /* solved, but max_iter stopped the loop   */
#define EGG_SYNC_E_ARG (-1)         /* bad n / d / eps / pointer               */
#define EGG_SYNC_E_RANGE (-2)       /* input outside [0,1] or not finite       */
#define EGG_SYNC_E_DIM (-3)         /* d above the supported maximum           */
#define EGG_SYNC_E_KEY (-4)         /* cell keys would need more than 256 bits */
#define EGG_SYNC_E_ALLOC (-5)       /* device allocation failed                */
#define EGG_SYNC_E_HIP (-6)         /* HIP runtime error                       */
#define EGG_SYNC_E_NODEV (-7)       /* no GPU visible                          */
#define EGG_SYNC_E_TOO_BIG (-8)     /* workspace beyond the no-alloc scratch   */
#define EGG_SYNC_E_INTERNAL (-9)    /* a bounded device-side wait expired      */

#define EGG_SYNC_STOP_EGG 0         /* local sync + MBR extra check (EGG_SynC.cu:706-709,1425-1489) */
#define EGG_SYNC_STOP_RC 1          /* r_c >= lam (SynC.cpp:277)                                    */
#define EGG_SYNC_LABELS_CELLS 0     /* EGG_SynC.cu:1205-1227 */
#define EGG_SYNC_LABELS_SYNC 1      /* synCluster_v2, SynC.cpp:88-129 */

#define EGG_SYNC_MAX_D 20

typedef struct egg_sync_options {
    int32_t stop_rule;       /* EGG_SYNC_STOP_*                                   */
    int32_t label_rule;      /* EGG_SYNC_LABELS_*                                 */
    float lam;               /* threshold of EGG_SYNC_STOP_RC (reference 1-1e-3)  */
    int32_t max_iter;        /* 0 = no cap (the reference has none)               */
    int32_t compute_rc;      /* 1: report r_c (FP64) of the final positions       */
    int32_t timing;          /* 1: per-stage / per-iteration GPU event timings    */
    int32_t iters_per_sync;  /* iterations enqueued between host checks (>= 1)    */
    int32_t validate;        /* 1: check the input lies in [0,1] (default 1)      */
    int32_t use_graph;       /* 1: capture iters_per_sync iterations in one hipGraph
                                and replay it; with timing=1 the loop time is then
                                reported as a whole in stage_ms[2]                */
    int32_t queue_blocks;    /* workgroups of the tile-queue kernels (0 = auto)   */
    int32_t queue_threads;   /* threads per such workgroup: 64, 128, 192 or 256
                                (0 = 256)                                         */
    int32_t streams;         /* batched solve: concurrent HIP streams, at most 64
                                (0 = 4)                                           */
    int32_t tile_group;      /* update tiles: neighbouring cells that differ only
                                in the last coordinate, within aligned blocks of
                                this many cells (1, 2, 4 or 8; others round down),
                                share a wavefront (0 = default).  It changes the
                                order of FP32 additions only                      */
    int32_t pair_groups;     /* cells updated point pair by point pair instead of
                                by the cell walk: 1 = groups of fewer than 8
                                distinct positions in sparse grids (fewer than 4
                                distinct positions per cell), 2 = every group,
                                0 = automatic: every group of a sparse grid from
                                d = 6 on, else as 1.  It changes the order of
                                FP32 additions only                               */
} egg_sync_options;

typedef struct egg_sync_stats {
    int32_t iterations;      /* passes of the update loop                          */
    int32_t status;          /* 0 converged, 1 capped                              */
    int32_t n_clusters;
    int32_t n_cells;         /* non-empty cells of the final grid                  */
    double rc;               /* FP64 order parameter (NaN unless computed)         */
    double stage_ms[6];      /* Allocating, Build structure, Update, Extra check,
                                Clustering, Free Memory (run_experiment.py:20)     */
    double total_ms;         /* wall time of the call                              */
    double update_kernel_ms; /* summed duration of the update kernel (timing=1)    */
    int32_t update_launches;
    int32_t pad;
    int64_t workspace_bytes; /* device bytes held for the solve ("space")          */
    float *itr_ms;           /* optional caller array, per-iteration GPU time (ms) */
    int32_t itr_ms_cap;      /* capacity of itr_ms                                 */
    int32_t trace_cap;       /* capacity of cells_trace / rc_trace                 */
    int32_t *cells_trace;    /* optional: non-empty cells C of each iteration      */
    double *rc_trace;        /* optional: r_c of each iteration (stop rule RC)     */
} egg_sync_stats;

/* Fills *opt with the defaults of EGG_SynC (stop_rule EGG, labels cells, lam 0.999). */
int egg_sync_default_options(egg_sync_options *opt);

/* Host arrays in and out: h_D is n x d row-major float32 in [0,1];
 * labels_out[n] int32 (-1 = noise).  stats_out may be NULL. */
int egg_sync_solve(const float *h_D, int64_t n, int32_t d, float eps, int32_t stop_rule,
                   int32_t *labels_out, egg_sync_stats *stats_out);

/* As egg_sync_solve with options and optional final positions (n x d, host). */
int egg_sync_solve_ex(const float *h_D, int64_t n, int32_t d, float eps,
                      const egg_sync_options *opt, int32_t *labels_out, float *final_out,
                      egg_sync_stats *stats_out);

/* Device arrays: d_D (n x d row-major), d_labels[n], d_final (n x d, may be
 * NULL), all on the current device; work is ordered on `stream`. */
int egg_sync_solve_device(const float *d_D, int64_t n, int32_t d, float eps,
                          const egg_sync_options *opt, int32_t *d_labels, float *d_final,
                          egg_sync_stats *stats_out, void *stream);

/* egg_sync_solve_device without any allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a 64 MiB device array inside the library, one per GPU, held
 * until the process ends; a workload whose workspace exceeds it returns
 * EGG_SYNC_E_TOO_BIG (about n <= 135,000 at d = 3, 70,000 at d = 8;
 * stats_out->workspace_bytes reports the bytes a call used).  The caller
 * passes inputs and outputs as for egg_sync_solve_device.  No graph, no
 * timing events (use_graph and timing are ignored), the caller's stream or
 * the null stream.  Calls share the array: two calls that overlap on one GPU
 * race.  Results equal egg_sync_solve_device's bit for bit. */
int egg_sync_solve_device_noalloc(const float *d_D, int64_t n, int32_t d, float eps,
                                  const egg_sync_options *opt, int32_t *d_labels,
                                  float *d_final, egg_sync_stats *stats_out, void *stream);

/* ---- Batched solve (no counterpart in the reference: EGG_SynC solves one
 * array per call).  Many independent solves -- datasets of uneven size,
 * eps sweeps over shared data -- in one call on device data. */
typedef struct egg_sync_instance {
    int64_t data_offset;     /* float offset of this instance's n x d rows in d_D */
    int64_t n;
    int32_t d;
    float eps;
} egg_sync_instance;

typedef struct egg_sync_instance_result {
    int32_t iterations;
    int32_t status;          /* EGG_SYNC_OK or EGG_SYNC_CAPPED                    */
    int32_t n_clusters;
    int32_t n_cells;
} egg_sync_instance_result;

/* Solves inst[0..n_inst) with the options in *opt.  Labels are packed in
 * instance order (instance k starts at row sum_{j<k} n_j of d_labels); final
 * positions, if d_final is not NULL, likewise at float offset
 * sum_{j<k} n_j d_j.  results[n_inst] and stats may be NULL.  Each instance's
 * result equals what egg_sync_solve_device returns for it alone. */
int egg_sync_solve_batch_device(const float *d_D, const egg_sync_instance *inst,
                                int32_t n_inst, const egg_sync_options *opt,
                                int32_t *d_labels, float *d_final,
                                egg_sync_instance_result *results, egg_sync_stats *stats,
                                void *stream);

/* ---- Brute-force SynC (simple_GPU_SynC.cu): every point scans all n points,
 * O(n^2) per pass.  x'_p = x_p + mean_{|x_q - x_p| <= eps} sin(x_q - x_p), FP32
 * sums in ascending q; r = mean_p mean_q exp(-|x_q - x_p|) (FP64 mean of the
 * FP32 per-point values); passes stop once r >= lam or after max_iter passes
 * (max_iter <= 0: 100, the reference's cap, simple_GPU_SynC.cu:126; returns
 * EGG_SYNC_CAPPED when the cap stopped it).  Labels: a point with another
 * point within eps gets the smallest index among its eps-neighbours, and
 * those representatives are numbered 0..k-1 in index order; others -1.
 * Positions are row-major n x d with any finite values (no grid, no [0,1]
 * requirement).  stats (may be NULL): iterations, status, n_clusters, rc = r
 * of the last pass, stage_ms 0/2/4/5, itr_ms / rc_trace if given,
 * workspace_bytes; n_cells is 0. */
int egg_sync_simple_solve_device(const float *d_D, int64_t n, int32_t d, float eps, float lam,
                                 int32_t max_iter, int32_t *d_labels, float *d_final,
                                 egg_sync_stats *stats, void *stream);

/* Host arrays in and out (final_out, n x d, may be NULL). */
int egg_sync_simple_solve(const float *h_D, int64_t n, int32_t d, float eps, float lam,
                          int32_t max_iter, int32_t *labels_out, float *final_out,
                          egg_sync_stats *stats_out);

/* The labels above for device positions d_D; *n_clusters (host, may be NULL). */
int egg_sync_simple_labels_device(const float *d_D, int64_t n, int32_t d, float eps,
                                  int32_t *d_labels, int32_t *n_clusters, void *stream);

/* Grid build on device positions d_D (n x d) visited in order d_order_in
 * (point ids; NULL = identity).  Outputs (device, may be NULL):
 * d_order_out[n] point id per sorted position, d_cell_of[n],
 * d_cell_start[n+1], d_cell_g[n*d] (cell-major, C x d), d_ob[O], d_oe[O];
 * *C_out (host) = number of non-empty cells. */
int egg_sync_build_grid_device(const float *d_D, int64_t n, int32_t d, float eps,
                               const int32_t *d_order_in, int32_t *d_order_out,
                               int32_t *d_cell_of, int32_t *d_cell_start, int32_t *d_cell_g,
                               int32_t *d_ob, int32_t *d_oe, int32_t *C_out, void *stream);

/* One iteration of the EGG loop on device positions d_D (n x d, point-id
 * order) visited in order d_order_in (NULL = identity): writes d_next (n x d,
 * point-id order), d_order_out[n] and, per sorted position, d_N, d_half,
 * d_border, d_synced (each may be NULL); *flag_local / *flag_final (host) as
 * the reference's r_local before / after the extra check. */
int egg_sync_iteration_device(const float *d_D, int64_t n, int32_t d, float eps,
                              const int32_t *d_order_in, float *d_next, int32_t *d_order_out,
                              int32_t *d_N, int32_t *d_half, int32_t *d_border,
                              int32_t *d_synced, int32_t *flag_local, int32_t *flag_final,
                              void *stream);

/* r_c of device positions d_D (n x d): FP32 eps-membership and expf terms,
 * FP64 sums (written to *rc_out on the host). */
int egg_sync_rc_device(const float *d_D, int64_t n, int32_t d, float eps, double *rc_out,
                       void *stream);

/* ---- Learned jump (no counterpart in the reference) ------------------------
 * A small MLP predicts, per point, a multi-iteration jump toward its attractor
 * from 4d features; the plain iteration then resumes and its labels decide.
 * Parameters are one flat float array W1[h][in], b1[h], W2[h][h], b2[h],
 * W3[out][h], b3[out]. */

/* Features of device positions d_D (n x d, point-id order): per point, 4d
 * floats = position, own-cell mean sin, own-cell mean cos (cell totals of the
 * summary pass / count), (x - prev) / eps (0 if d_prev is NULL). */
int egg_sync_features_device(const float *d_D, const float *d_prev, int64_t n, int32_t d,
                             float eps, float *d_feat, void *stream);

/* d_X <- clamp(d_X + scale * clamp(d_Y, -clip, clip), 0, 1), n x d (no clamp of
 * d_Y if clip <= 0). */
int egg_sync_jump_device(float *d_X, const float *d_Y, int64_t n, int32_t d, float scale,
                         float clip, void *stream);

/* Number of parameters of an in-hidden-hidden-out MLP. */
int egg_mlp_param_count(int32_t in, int32_t hidden, int32_t out, int64_t *count);

/* Device workspace (floats) a training step on `batch` samples needs. */
size_t egg_mlp_workspace_floats(int32_t in, int32_t hidden, int32_t out, int64_t batch);

/* Inference: d_Y[batch x out] = MLP(d_X[batch x in]); fused kernel for 12-128-3. */
int egg_mlp_forward_device(const float *d_params, int32_t in, int32_t hidden, int32_t out,
                           const float *d_X, int64_t batch, float *d_Y, void *stream);

/* One training step, loss mean_b sum_o (MLP(x_b) - t_b)^2: gradients into
 * d_grad (same layout as the parameters); if lr > 0 an Adam step (number
 * `step`, 1-based, moments d_adam_m / d_adam_v) updates d_params.  *loss_out
 * (host, may be NULL) receives the loss before the update. */
int egg_mlp_train_step_device(float *d_params, float *d_adam_m, float *d_adam_v, int32_t in,
                              int32_t hidden, int32_t out, const float *d_X, const float *d_T,
                              int64_t batch, float lr, int32_t step, float *d_work,
                              int64_t work_floats, float *d_grad, double *loss_out,
                              void *stream);

/* Grid geometry the solver derives from (n, d, eps) (EGG_SynC.cu:1249-1260):
 * out[0]=w, out[1]=od, out[2]=O, out[3]=R, out[4]=key bits; *cs_out cell side. */
int egg_sync_geometry(int64_t n, int32_t d, float eps, int64_t *out5, float *cs_out);

/* Frees the cached device workspace of the current device. */
int egg_sync_release(void);

/* Static description of an error code. */
const char *egg_sync_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif /* EGG_SYNC_H */
