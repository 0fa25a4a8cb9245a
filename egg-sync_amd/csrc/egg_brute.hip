// Brute-force SynC on the GPU: GPU_DynamicalClustering (simple_GPU_SynC.cu:105-237).
//
// Every point p scans all n points q in index order (kernel_itr,
// simple_GPU_SynC.cu:29-67): q is a neighbour if the FP32 distance
// sqrtf(sum_l (x_p,l - x_q,l)^2) is <= eps; then x'_p,l = x_p,l +
// (sum_q sinf(x_q,l - x_p,l)) / N and r_p = (sum_q expf(-dist)) / N, all
// FP32 sums in ascending q as the reference adds them.  r = mean r_p stops
// the loop once r >= lam or after max_iter passes (the reference's cap is 100,
// :126).  Labels (GPU_synCluster + rename_cluster, :70-102): a point with at
// least one other point within eps gets the smallest index among its
// eps-neighbours (itself included), and those representatives are numbered
// 0..k-1 in index order by an exclusive scan.
//
// MI355X layout: positions stay row-major [n][d] (the caller's layout); a
// workgroup of 128 targets streams the sources through LDS in tiles of 128
// points, every lane reading each source as an LDS broadcast.  r is summed in
// FP64 by one fixed-order reduction (the reference's float atomicAdd sums in
// an order that changes from run to run), so runs are bitwise repeatable.
// The loop state lives on the device; the host enqueues several passes and
// reads the state back once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <vector>

#include "../../include/egg_sync.h"
#include "egg_host.h"
#include "egg_internal.h"

namespace egg {
namespace {

constexpr int kBruteBlock = 128;   // targets per workgroup (two wavefronts)
constexpr int kReduceBlock = 1024;
constexpr int kSimpleCap = 100;    // simple_GPU_SynC.cu:126
constexpr int kPassesPerCheck = 4;

struct BruteState {
    int done;      // 1: r >= lam or capped; later passes are no-ops
    int iter;      // passes completed
    int capped;
    int max_iter;
    float lam;
    float r;       // r of the last pass, FP32 (the reference's r_local)
    double r64;    // the same mean in FP64
};

template <int D>
__global__ __launch_bounds__(kBruteBlock) void k_brute_pass(const float *__restrict__ X, int n,
                                                            float thr, float *__restrict__ Xn,
                                                            float *__restrict__ rp,
                                                            const BruteState *st) {
    if (st->done) return;
    __shared__ float s_x[kBruteBlock * D];
    const int p = blockIdx.x * kBruteBlock + threadIdx.x;
    const bool act = p < n;
    float x[D], sum[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
        x[l] = act ? X[(long long)p * D + l] : 0.f;
        sum[l] = 0.f;
    }
    float r = 0.f;
    int N = 0;
    for (int base = 0; base < n; base += kBruteBlock) {
        const int m = min(kBruteBlock, n - base);
        __syncthreads();
        for (int k = threadIdx.x; k < m * D; k += kBruteBlock)
            s_x[k] = X[(long long)base * D + k];
        __syncthreads();
        if (!act) continue;
        for (int j = 0; j < m; j++) {
            const float *q = s_x + j * D;
            float s = 0.f;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float df = x[l] - q[l];
                s += df * df;
            }
            if (s <= thr) {  // sqrtf(s) <= eps
                N += 1;
#pragma unroll
                for (int l = 0; l < D; l++) sum[l] += sinf(q[l] - x[l]);
                r += expf(-sqrtf(s));
            }
        }
    }
    if (act) {
#pragma unroll
        for (int l = 0; l < D; l++) Xn[(long long)p * D + l] = x[l] + sum[l] / (float)N;
        rp[p] = r / (float)N;
    }
}

// r = mean of r_p (FP64, fixed order), stop test, trace.
__global__ __launch_bounds__(kReduceBlock) void k_brute_reduce(const float *__restrict__ rp, int n,
                                                               BruteState *st,
                                                               float *__restrict__ trace,
                                                               int trace_cap) {
    if (st->done) return;
    __shared__ double s[kReduceBlock];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kReduceBlock) v += (double)rp[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kReduceBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r64 = s[0] / (double)n;
        const float r = (float)r64;
        const int it = st->iter + 1;
        st->iter = it;
        st->r = r;
        st->r64 = r64;
        if (trace && it <= trace_cap) trace[it - 1] = r;
        if (!(r < st->lam)) {
            st->done = 1;
        } else if (it >= st->max_iter) {
            st->done = 1;
            st->capped = 1;
        }
    }
}

// GPU_synCluster (simple_GPU_SynC.cu:70-95): the first eps-neighbour in index
// order is the smallest; a lane stops once it has seen two.
template <int D>
__global__ __launch_bounds__(kBruteBlock) void k_brute_labels(const float *__restrict__ X, int n,
                                                              float thr, int *__restrict__ C,
                                                              int *__restrict__ incl) {
    __shared__ float s_x[kBruteBlock * D];
    const int p = blockIdx.x * kBruteBlock + threadIdx.x;
    const bool act = p < n;
    float x[D];
#pragma unroll
    for (int l = 0; l < D; l++) x[l] = act ? X[(long long)p * D + l] : 0.f;
    int mn = -1, cnt = 0;
    bool need = act;
    for (int base = 0; __syncthreads_or(need) && base < n; base += kBruteBlock) {
        const int m = min(kBruteBlock, n - base);
        for (int k = threadIdx.x; k < m * D; k += kBruteBlock)
            s_x[k] = X[(long long)base * D + k];
        __syncthreads();
        if (need) {
            for (int j = 0; j < m; j++) {
                const float *q = s_x + j * D;
                float s = 0.f;
#pragma unroll
                for (int l = 0; l < D; l++) {
                    const float df = x[l] - q[l];
                    s += df * df;
                }
                if (s <= thr) {
                    if (mn < 0) mn = base + j;
                    if (++cnt >= 2) {
                        need = false;
                        break;
                    }
                }
            }
        }
    }
    if (act) {
        if (cnt >= 2) {
            C[p] = mn;
            incl[mn] = 1;
        } else {
            C[p] = -1;
        }
    }
}

// rename_cluster (simple_GPU_SynC.cu:97-102) with the exclusive scan of incl.
__global__ __launch_bounds__(kBlock) void k_brute_rename(int *__restrict__ C,
                                                         const int *__restrict__ pos, int n) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        if (C[p] >= 0) C[p] = pos[C[p]];
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

#define HIP_TRY(x)                       \
    do {                                 \
        int rc_ = hip_err((x));          \
        if (rc_) return rc_;             \
    } while (0)

int check_args(int64_t n, int32_t d, float eps) {
    if (n < 0 || n > (int64_t)(INT_MAX / 2) || d < 1) return EGG_SYNC_E_ARG;
    if (d > kMaxD) return EGG_SYNC_E_DIM;
    if (!(eps > 0.0f) || !std::isfinite(eps)) return EGG_SYNC_E_ARG;
    if ((int64_t)n * d > (int64_t)INT_MAX) return EGG_SYNC_E_ARG;
    return EGG_SYNC_OK;
}

void launch_pass(int d, const float *X, int n, float thr, float *Xn, float *rp,
                 const BruteState *st, hipStream_t s) {
    const int blocks = (n + kBruteBlock - 1) / kBruteBlock;
#define PASS_CASE(DD) \
    case DD: k_brute_pass<DD><<<blocks, kBruteBlock, 0, s>>>(X, n, thr, Xn, rp, st); break;
    switch (d) {
        EGG_EACH_D(PASS_CASE)
    }
#undef PASS_CASE
}

// Labels of positions X into C (device), representatives numbered by a scan;
// *k (device) receives their number.  incl, pos: [n] ints; tmp: scan_tmp_ints(n).
int launch_labels(int d, const float *X, int n, float thr, int *C, int *incl, int *pos,
                  int *tmp, int *k, hipStream_t s) {
    HIP_TRY(hipMemsetAsync(incl, 0, sizeof(int) * (size_t)n, s));
    const int blocks = (n + kBruteBlock - 1) / kBruteBlock;
#define LAB_CASE(DD) \
    case DD: k_brute_labels<DD><<<blocks, kBruteBlock, 0, s>>>(X, n, thr, C, incl); break;
    switch (d) {
        EGG_EACH_D(LAB_CASE)
    }
#undef LAB_CASE
    scan_exclusive(incl, pos, n, nullptr, tmp, k, nullptr, s);
    k_brute_rename<<<std::min((n + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(C, pos, n);
    return hip_err(hipGetLastError());
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace
}  // namespace egg

using namespace egg;

extern "C" {

int egg_sync_simple_solve_device(const float *d_D, int64_t n64, int32_t d, float eps, float lam,
                                 int32_t max_iter, int32_t *d_labels, float *d_final,
                                 egg_sync_stats *stats, void *stream) {
    const auto t_call = std::chrono::steady_clock::now();
    int rc = check_args(n64, d, eps);
    if (rc) return rc;
    if (n64 > 0 && (!d_D || !d_labels)) return EGG_SYNC_E_ARG;
    if (max_iter <= 0) max_iter = kSimpleCap;
    const int n = (int)n64;
    hipStream_t s = (hipStream_t)stream;
    const float thr = sqrt_threshold(eps);
    if (stats) {
        stats->iterations = 0;
        stats->status = 0;
        stats->n_clusters = 0;
        stats->n_cells = 0;
        stats->rc = NAN;
        for (double &v : stats->stage_ms) v = 0.0;
        stats->update_kernel_ms = 0.0;
        stats->update_launches = 0;
        stats->workspace_bytes = 0;
    }
    if (n == 0) {
        if (stats) stats->total_ms = ms_since(t_call);
        return EGG_SYNC_OK;
    }
    // Allocating (stage 0): two position buffers, r_p, label scratch, state
    auto t = std::chrono::steady_clock::now();
    const size_t nd = (size_t)n * d;
    const size_t tmp_ints = scan_tmp_ints(n);
    const int trace_cap = std::min(max_iter, 65536);  // r of the first passes, for stats
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bytes = 2 * al(sizeof(float) * nd) + al(sizeof(float) * n) +
                         2 * al(sizeof(int) * n) + al(sizeof(int) * tmp_ints) +
                         al(sizeof(int)) + al(sizeof(BruteState)) +
                         al(sizeof(float) * (size_t)trace_cap);
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, bytes));
    char *cp = (char *)buf.p;
    auto take = [&](size_t b) {
        char *r = cp;
        cp += al(b);
        return (void *)r;
    };
    float *A = (float *)take(sizeof(float) * nd);
    float *B = (float *)take(sizeof(float) * nd);
    float *rp = (float *)take(sizeof(float) * n);
    int *incl = (int *)take(sizeof(int) * n);
    int *pos = (int *)take(sizeof(int) * n);
    int *tmp = (int *)take(sizeof(int) * tmp_ints);
    int *k_dev = (int *)take(sizeof(int));
    BruteState *st = (BruteState *)take(sizeof(BruteState));
    float *trace = (float *)take(sizeof(float) * (size_t)trace_cap);
    BruteState hs{};
    hs.max_iter = max_iter;
    hs.lam = lam;
    HIP_TRY(hipMemcpyAsync(st, &hs, sizeof hs, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(A, d_D, sizeof(float) * nd, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (stats) stats->stage_ms[0] = ms_since(t);

    // Update (stage 2): passes A -> B -> A ..., checked every few passes;
    // with per-iteration times requested, every pass (as the reference does)
    t = std::chrono::steady_clock::now();
    const bool per_iter = stats && stats->itr_ms && stats->itr_ms_cap > 0;
    const int per_check = per_iter ? 1 : kPassesPerCheck;
    int enqueued = 0;
    auto t_it = std::chrono::steady_clock::now();
    while (enqueued < max_iter) {
        const int batch = std::min(per_check, max_iter - enqueued);
        for (int b = 0; b < batch; b++, enqueued++) {
            const float *src = (enqueued & 1) ? B : A;
            float *dst = (enqueued & 1) ? A : B;
            launch_pass(d, src, n, thr, dst, rp, st, s);
            k_brute_reduce<<<1, kReduceBlock, 0, s>>>(rp, n, st, trace, trace_cap);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&hs, st, sizeof hs, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (per_iter && hs.iter >= 1 && hs.iter <= stats->itr_ms_cap) {
            stats->itr_ms[hs.iter - 1] = (float)ms_since(t_it);
            t_it = std::chrono::steady_clock::now();
        }
        if (hs.done) break;
    }
    const float *fin = (hs.iter & 1) ? B : A;
    if (stats) stats->stage_ms[2] = ms_since(t);

    // Clustering (stage 4)
    t = std::chrono::steady_clock::now();
    rc = launch_labels(d, fin, n, thr, d_labels, incl, pos, tmp, k_dev, s);
    if (rc) return rc;
    if (d_final)
        HIP_TRY(hipMemcpyAsync(d_final, fin, sizeof(float) * nd, hipMemcpyDeviceToDevice, s));
    int k = 0;
    HIP_TRY(hipMemcpyAsync(&k, k_dev, sizeof k, hipMemcpyDeviceToHost, s));
    std::vector<float> htrace;
    if (stats && stats->rc_trace && stats->trace_cap > 0) {
        htrace.resize((size_t)std::min(hs.iter, trace_cap));
        if (!htrace.empty())
            HIP_TRY(hipMemcpyAsync(htrace.data(), trace, sizeof(float) * htrace.size(),
                                   hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (stats) stats->stage_ms[4] = ms_since(t);

    // Free Memory (stage 5)
    t = std::chrono::steady_clock::now();
    void *p = buf.p;
    buf.p = nullptr;
    HIP_TRY(hipFree(p));
    if (stats) {
        stats->stage_ms[5] = ms_since(t);
        stats->iterations = hs.iter;
        stats->status = hs.capped ? 1 : 0;
        stats->n_clusters = k;
        stats->rc = hs.r64;
        stats->workspace_bytes = (int64_t)bytes;
        for (size_t i = 0; i < htrace.size() && (int)i < stats->trace_cap; i++)
            stats->rc_trace[i] = (double)htrace[i];
        stats->total_ms = ms_since(t_call);
    }
    return hs.capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK;
}

int egg_sync_simple_solve(const float *h_D, int64_t n, int32_t d, float eps, float lam,
                          int32_t max_iter, int32_t *labels_out, float *final_out,
                          egg_sync_stats *stats_out) {
    int rc = check_args(n, d, eps);
    if (rc) return rc;
    if (n > 0 && (!h_D || !labels_out)) return EGG_SYNC_E_ARG;
    const size_t nd = (size_t)n * d;
    DevBuf dD, dL, dF;
    if (n > 0) {
        HIP_TRY(hipMalloc(&dD.p, sizeof(float) * nd));
        HIP_TRY(hipMalloc(&dL.p, sizeof(int) * (size_t)n));
        if (final_out) HIP_TRY(hipMalloc(&dF.p, sizeof(float) * nd));
        HIP_TRY(hipMemcpy(dD.p, h_D, sizeof(float) * nd, hipMemcpyHostToDevice));
    }
    rc = egg_sync_simple_solve_device((const float *)dD.p, n, d, eps, lam, max_iter,
                                      (int32_t *)dL.p, (float *)dF.p, stats_out, nullptr);
    if (rc < 0) return rc;
    if (n > 0) {
        HIP_TRY(hipMemcpy(labels_out, dL.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
        if (final_out)
            HIP_TRY(hipMemcpy(final_out, dF.p, sizeof(float) * nd, hipMemcpyDeviceToHost));
    }
    return rc;
}

int egg_sync_simple_labels_device(const float *d_D, int64_t n64, int32_t d, float eps,
                                  int32_t *d_labels, int32_t *n_clusters, void *stream) {
    int rc = check_args(n64, d, eps);
    if (rc) return rc;
    if (n64 > 0 && (!d_D || !d_labels)) return EGG_SYNC_E_ARG;
    const int n = (int)n64;
    if (n == 0) {
        if (n_clusters) *n_clusters = 0;
        return EGG_SYNC_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t tmp_ints = scan_tmp_ints(n);
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, sizeof(int) * (2 * (size_t)n + tmp_ints + 1)));
    int *incl = (int *)buf.p, *pos = incl + n, *tmp = pos + n, *k_dev = tmp + tmp_ints;
    rc = launch_labels(d, d_D, n, sqrt_threshold(eps), d_labels, incl, pos, tmp, k_dev, s);
    if (rc) return rc;
    int k = 0;
    HIP_TRY(hipMemcpyAsync(&k, k_dev, sizeof k, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (n_clusters) *n_clusters = k;
    return EGG_SYNC_OK;
}

}  // extern "C"
