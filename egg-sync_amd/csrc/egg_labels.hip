// Cluster extraction (EGG_SynC.cu:1205-1227; SynC.cpp:88-129) and input
// validation, gfx950.
#include <climits>

#include "egg_internal.h"

namespace egg {

static inline int grid_for(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

__global__ void k_label_flags(const int *__restrict__ cell_start, const DevState *st,
                              int *__restrict__ multi) {
    const int C = st->C;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x)
        multi[c] = (cell_start[c + 1] - cell_start[c]) > 1 ? 1 : 0;
}

// GPU_synCluster_float_grid: a point gets its cell's cluster id when the cell
// holds more than one point, else -1.  Ids are compacted to 0..k-1 in cell
// order (the reference leaves gaps; only the partition is meaningful).
__global__ void k_labels_egg(const int *__restrict__ point_cell, const int *__restrict__ multi,
                             const int *__restrict__ lab_off, const int *__restrict__ perm_s,
                             int n, int *__restrict__ labels) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = point_cell[i];
        labels[perm_s[i]] = multi[c] ? lab_off[c] : -1;
    }
}

void launch_labels_egg(const Geo &g, GridBufs &b, int *labels, int *multi, int *lab_off,
                       DevState *st, hipStream_t s) {
    const int gb = grid_for(g.n);
    k_label_flags<<<gb, kBlock, 0, s>>>(b.cell_start, st, multi);
    scan_exclusive(multi, lab_off, g.n, &st->C, b.scan_tmp, &st->n_clusters, nullptr, s, true);
    k_labels_egg<<<gb, kBlock, 0, s>>>(b.point_cell, multi, lab_off, b.perm_s, g.n, labels);
}

__global__ void k_invert(const int *__restrict__ perm_s, int n, int *__restrict__ inv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        inv[perm_s[i]] = i;
}

// synCluster_v2 (SynC.cpp:88-129) in one workgroup: points are visited in id
// order; an unlabelled point i gives label cl to every other point within FP32
// distance eps (overwriting earlier labels), and takes cl itself if it found
// any; the search for the next unlabelled id and the neighbour scan are
// parallel over the workgroup.  Candidates: Chebyshev radius R+1.
constexpr int kLabBlock = 1024;

template <int D>
__global__ __launch_bounds__(kLabBlock) void k_labels_sync(
    const float *__restrict__ Xs, const int *__restrict__ perm_s, const int *__restrict__ invperm,
    const int *__restrict__ point_cell, const int *__restrict__ cell_start,
    const int *__restrict__ cell_g, const int *__restrict__ ob, const int *__restrict__ oe,
    Geo g, int *labels, DevState *st) {
    __shared__ int s_min, s_any;
    const int n = g.n, od = g.od, R = g.R + 1, t = threadIdx.x;
    for (int i = t; i < n; i += kLabBlock) labels[i] = -1;
    __syncthreads();
    int span = 1;
    for (int l = 0; l < od; l++) span *= 2 * R + 1;
    int cl = 0, cursor = 0;
    while (cursor < n) {
        if (t == 0) s_min = INT_MAX;
        __syncthreads();
        int found = INT_MAX;
        for (int base = cursor; base < n; base += kLabBlock) {
            const int i = base + t;
            if (i < n && labels[i] == -1) atomicMin(&s_min, i);
            __syncthreads();
            found = s_min;
            __syncthreads();
            if (found != INT_MAX) break;
        }
        if (found == INT_MAX) break;
        const int si = invperm[found];
        float x[D];
        int g0[D];
        const int c0 = point_cell[si];
        for (int l = 0; l < D; l++) {
            x[l] = Xs[(long long)l * n + si];
            g0[l] = cell_g[(long long)l * n + c0];
        }
        if (t == 0) s_any = 0;
        __syncthreads();
        for (int k = t; k < span; k += kLabBlock) {
            int rem = k, o = 0;
            bool ok = true;
            for (int l = 0; l < od; l++) {
                const int v = g0[l] + (rem % (2 * R + 1)) - R;
                rem /= 2 * R + 1;
                if (v < 0 || v >= g.w) ok = false;
                o += v * g.wpow[l];
            }
            if (!ok) continue;
            for (int c = ob[o]; c < oe[o]; c++) {
                bool skip = false;
                for (int l = od; l < D; l++) {
                    const int dg = cell_g[(long long)l * n + c] - g0[l];
                    if (dg > R || dg < -R) skip = true;
                }
                if (skip) continue;
                for (int j = cell_start[c]; j < cell_start[c + 1]; j++) {
                    if (j == si) continue;
                    float s = 0.f;
                    for (int l = 0; l < D; l++) {
                        const float df = x[l] - Xs[(long long)l * n + j];
                        s += df * df;
                    }
                    if (s <= g.thr_eps) {
                        labels[perm_s[j]] = cl;
                        s_any = 1;
                    }
                }
            }
        }
        __syncthreads();
        if (s_any) {
            if (t == 0) labels[found] = cl;
            cl++;
        }
        cursor = found + 1;
        __syncthreads();
    }
    if (t == 0) st->n_clusters = cl;
}

void launch_labels_sync(const Geo &g, GridBufs &b, int *labels, int *invperm, DevState *st,
                        hipStream_t s) {
    k_invert<<<grid_for(g.n), kBlock, 0, s>>>(b.perm_s, g.n, invperm);
#define LS_CASE(DD)                                                                           \
    case DD:                                                                                  \
        k_labels_sync<DD><<<1, kLabBlock, 0, s>>>(b.Xs, b.perm_s, invperm, b.point_cell,      \
                                                  b.cell_start, b.cell_g, b.ob, b.oe, g,      \
                                                  labels, st);                                \
        break;
    switch (g.d) {
        EGG_EACH_D(LS_CASE)
    }
#undef LS_CASE
}

__global__ void k_validate(const float *__restrict__ aos, long long count, int *err) {
    bool bad = false;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x) {
        const float v = aos[i];
        if (!(v >= 0.0f && v <= 1.0f)) bad = true;
    }
    if (__any(bad) && __lane_id() == 0) atomicOr(err, 1);
}

void launch_validate(const float *aos, long long count, int *err, hipStream_t s) {
    k_validate<<<grid_for(count), kBlock, 0, s>>>(aos, count, err);
}

}  // namespace egg
