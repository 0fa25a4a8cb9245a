// Neighbour update, local-sync flag, extra check and FP64 order parameter
// (EGG_SynC.cu:551-763, 767-1203; SynC.cpp:30-56), gfx950.
//
// The EGG kernels work on runs -- maximal stretches of bitwise-equal
// consecutive positions in the cell-sorted order: every point of a run has
// the same inputs and hence the same result, and every scan over a cell's
// points can visit its runs once, counting each with its length.  One
// wavefront handles a tile of <= 64 runs of one cell (tiles come from a work
// queue), so the walk over candidate cells and every cell load are
// wave-uniform while each lane runs the reference's per-point tests on its
// run's position.  A scatter pass copies each run's result to its points.
// Candidate cells come from the outer-cell box of radius R (as the
// reference), pruned in the remaining dimensions to |dg| <= R+1 by a binary
// search over the column's cells (sorted by g_od): cells farther than that
// fail the reference's box test by more than a cell side, so results do not
// change.  Up to d = 3 the masked sums over streamed runs run on the matrix
// cores (stream_chunk_mfma), in the same order of FP32 additions.
#include <algorithm>
#include <cmath>

#include "egg_internal.h"

namespace egg {

// Odometer over the outer-cell offsets [-R, R]^od, dim 0 fastest (the
// enumeration order of EGG_SynC.cu:407-451), for per-thread walks.
template <int D>
struct Odo {
    int idx[D];
    bool finished;
    __device__ __forceinline__ void init(int od, int R) {
#pragma unroll
        for (int l = 0; l < D; l++) idx[l] = (l < od) ? -R : 0;
        finished = false;
    }
    __device__ __forceinline__ void next(int od, int R) {
        bool carry = true;
#pragma unroll
        for (int l = 0; l < D; l++) {
            if (l < od && carry) {
                idx[l]++;
                if (idx[l] > R) {
                    if (l == od - 1) finished = true;
                    else idx[l] = -R;
                } else {
                    carry = false;
                }
            }
        }
    }
    // outer cell id of the current offset, or -1 if outside the grid
    __device__ __forceinline__ int outer(const int *g0, const Geo &g) const {
        int o = 0;
        bool ok = true;
#pragma unroll
        for (int l = 0; l < D; l++) {
            if (l < g.od) {
                const int v = g0[l] + idx[l];
                if (v < 0 || v >= g.w) ok = false;
                o += v * g.wpow[l];
            }
        }
        return ok ? o : -1;
    }
};

// Walks the non-empty outer cells of the box [g0 - R, g0 + R]^od in the same
// order as Odo, cooperatively: the 2R+1 outer cells of a row along dim 0 have
// adjacent ids, so lane k loads the cell range of offset k and the wavefront
// visits the non-empty ones in lane order.  f(cb, ce) receives each cell
// range.  Called by a whole wavefront with wave-uniform g0 and R (2R+1 <= 64).
template <int D, typename Fn>
__device__ __forceinline__ void walk_columns(const int *g0, int R, const Geo &g,
                                             const int *__restrict__ ob,
                                             const int *__restrict__ oe, Fn &&f) {
    const int od = g.od, lane = __lane_id();
    int idx[D];
#pragma unroll
    for (int l = 0; l < D; l++) idx[l] = -R;
    for (;;) {
        bool ok = true;
        int base = 0;
#pragma unroll
        for (int l = 1; l < D; l++) {
            if (l < od) {
                const int v = g0[l] + idx[l];
                if (v < 0 || v >= g.w) ok = false;
                base += v * g.wpow[l];
            }
        }
        if (ok) {
            int cb = 0, ce = 0;
            const int v0 = g0[0] - R + lane;
            if (lane <= 2 * R && v0 >= 0 && v0 < g.w) {
                cb = ob[base + v0];
                ce = oe[base + v0];
            }
            unsigned long long m = __ballot(ce > cb);
            while (m) {
                const int k = __ffsll((long long)m) - 1;
                m &= m - 1;
                f(__builtin_amdgcn_readlane(cb, k), __builtin_amdgcn_readlane(ce, k));
            }
        }
        bool carry = true;
#pragma unroll
        for (int l = 1; l < D; l++) {
            if (l < od && carry) {
                if (++idx[l] > R) idx[l] = -R;
                else carry = false;
            }
        }
        if (carry) break;
    }
}

// First cell of [cb, ce) whose g_od is >= lo_v (cells of a column are sorted
// by g_od); the caller stops at the first cell above its range.
__device__ __forceinline__ int column_lower(const int *__restrict__ cell_g_od, int cb, int ce,
                                            int lo_v) {
    int lo = cb, hi = ce;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cell_g_od[mid] < lo_v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int D>
__device__ __forceinline__ int select_dim(const int *v, int k) {
    int r = 0;
#pragma unroll
    for (int l = 0; l < D; l++)
        if (l == k) r = v[l];
    return r;
}

template <int D>
__device__ __forceinline__ float dist2_to(const float *x, const float *__restrict__ Xs, int n,
                                          int j, float *xq) {
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < D; l++) {
        xq[l] = Xs[(long long)l * n + j];
        const float df = x[l] - xq[l];
        s += df * df;
    }
    return s;
}

// Loads the D coordinates of cell c, returns false when the cell lies outside
// the pruning window of the remaining dimensions (*stop: past the window in
// g_od, so later cells of the column are outside too).
template <int D>
__device__ __forceinline__ bool cell_in_window(const int *__restrict__ cell_g, int n, int c,
                                               const int *g0, int od, int rad, int g0od,
                                               int *gc, bool *stop) {
#pragma unroll
    for (int l = 0; l < D; l++) gc[l] = cell_g[(long long)l * n + c];
    *stop = od < D && select_dim<D>(gc, od) > g0od + rad;
    bool in = true;
#pragma unroll
    for (int l = 0; l < D; l++)
        if (l > od && (gc[l] - g0[l] > rad || g0[l] - gc[l] > rad)) in = false;
    return in && !*stop;
}

// A source run of a cell staged in LDS: its position and, for its offset v
// from the tile's anchor, sin v and cos v - 1 (in cv).
template <int D>
struct SrcPt {
    static constexpr int P = (D + 3) / 4 * 4;
    float x[P], sv[P], cv[P];
};

// Makes the LDS writes of every lane visible to the reads of the others
// (wavefront scope; the LDS serves one wavefront's accesses in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Group of tile t: the largest k with tile_off[k] <= t.
__device__ __forceinline__ int tile_group(const int *__restrict__ tile_off, int G, int t) {
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A tile seen by one lane: the tile's group (cells [gc0, gc1), all in one
// outer cell, equal in every inner coordinate but the last), the lane's run
// r (active if it belongs to the tile), its first sorted position i and its
// cell cl.
struct TileLane {
    int gc0, gc1, r, i, cl;
    bool active;
};

__device__ __forceinline__ TileLane tile_lane(int t, const int *__restrict__ tile_off, int G,
                                              const int *__restrict__ grp_cell,
                                              const int *__restrict__ cell_run,
                                              const int *__restrict__ run_start,
                                              const int *__restrict__ point_cell) {
    TileLane L;
    const int k = tile_group(tile_off, G, t);
    L.gc0 = grp_cell[k];
    L.gc1 = grp_cell[k + 1];
    L.r = cell_run[L.gc0] + (t - tile_off[k]) * kTile + __lane_id();
    L.active = L.r < cell_run[L.gc1];
    L.i = L.active ? run_start[L.r] : 0;
    L.cl = L.active ? point_cell[L.i] : L.gc0;
    return L;
}

// cell_in_window for a group: the window of the last dimension spans
// [g0[d-1] - rad, ghi + rad] (ghi = last coordinate of the group's last cell).
template <int D>
__device__ __forceinline__ bool cell_in_group_window(const int *__restrict__ cell_g, int n, int c,
                                                     const int *g0, int od, int rad, int ghi,
                                                     int *gc, bool *stop) {
#pragma unroll
    for (int l = 0; l < D; l++) gc[l] = cell_g[(long long)l * n + c];
    const int hi_od = od == D - 1 ? ghi : select_dim<D>(g0, od);
    *stop = od < D && select_dim<D>(gc, od) > hi_od + rad;
    bool in = true;
#pragma unroll
    for (int l = 0; l < D; l++) {
        const int hi = l == D - 1 ? ghi : g0[l];
        if (l > od && (gc[l] - hi > rad || g0[l] - gc[l] > rad)) in = false;
    }
    return in && !*stop;
}

__device__ __forceinline__ int next_tile(int *counter) {
    int t = 0;
    if (__lane_id() == 0) t = atomicAdd(counter, 1);
    return __builtin_amdgcn_readfirstlane(t);
}

// Work claims of the queue kernels: wavefront w of the grid's W takes item w
// first, without an atomic, and later items from the counter, offset by W.
// A pass with fewer items than wavefronts (converged iterations) takes no
// atomic at all; with more, each wavefront takes one per extra item.
__device__ __forceinline__ int wave_index() {
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
}

__device__ __forceinline__ int wave_count() { return gridDim.x * (blockDim.x / kWave); }

// Bitwise OR of v over the wavefront.
__device__ __forceinline__ unsigned wave_or(unsigned v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v |= (unsigned)__shfl_xor((int)v, o, kWave);
    return v;
}

// Classifies a sub-cell for a lane at x against threshold thr (thr_in =
// thr * (1 - 64 * 2^-24), thr_out = thr * (1 + 64 * 2^-24)): 1 if every
// point of it has FP32 squared distance <= thr, 0 if none has, 2 otherwise.
// The box is the bounding box of the sub-cell's points (smin / smax); this
// FP32 distance to the box and the reference's FP32 distance of any point in
// it are each within (d + 2) * 2^-24 relative of their exact values, so for
// d <= 20 the margin of 64 * 2^-24 covers both.
template <int D>
__device__ __forceinline__ int sub_class_box(const float *blo, const float *bhi, const float *x,
                                             float thr_in, float thr_out) {
    static_assert(2 * D + 4 < 64, "classification margin");
    float mn = 0.f, mx = 0.f;
#pragma unroll
    for (int l = 0; l < D; l++) {
        const float lo = blo[l], hi = bhi[l];
        const float ta = lo - x[l], tb = hi - x[l];
        const float a = ta * ta, b = tb * tb;
        mx += fmaxf(a, b);
        mn += x[l] < lo ? a : (x[l] > hi ? b : 0.f);
    }
    if (mx <= thr_in) return 1;
    if (mn >= thr_out) return 0;
    return 2;
}

// The boxes of sub-cells s0 .. s0 + ns - 1 (ns <= 64) into LDS, box k in
// src[k] (lower corner in x, upper in sv), so that every lane reads them as
// broadcasts instead of loading them one after another.
template <int D>
__device__ __forceinline__ void stage_boxes(const unsigned *__restrict__ smin,
                                            const unsigned *__restrict__ smax, int n, int s0,
                                            int ns, SrcPt<D> *src) {
    const int lane = __lane_id();
    if (lane < ns) {
#pragma unroll
        for (int l = 0; l < D; l++) {
            src[lane].x[l] = ord_val(smin[(long long)l * n + s0 + lane]);
            src[lane].sv[l] = ord_val(smax[(long long)l * n + s0 + lane]);
        }
    }
    wave_sync();
}

// Run of index q in a streamed range: runs [r0, r0 + total) if smask is 0,
// else the runs of the sub-cells s0 + k of the bits k of smask, in order;
// -1 past the end.
__device__ __forceinline__ int stream_run(int q, int total, int r0, int s0, unsigned smask,
                                          const int *__restrict__ sub_run) {
    if (!smask) return q < total ? r0 + q : -1;
    int before = 0, rr = -1;
    for (unsigned mm = smask; mm; mm &= mm - 1) {
        const int k = __builtin_ctz(mm);
        const int a = sub_run[s0 + k], b = sub_run[s0 + k + 1];
        if (q >= before && q < before + (b - a)) rr = a + q - before;
        before += b - a;
    }
    return rr;
}

// Runs in a streamed range.
__device__ __forceinline__ int stream_total(int r0, int r1, int s0, unsigned smask,
                                            const int *__restrict__ sub_run) {
    if (!smask) return r1 - r0;
    int total = 0;
    for (unsigned mm = smask; mm; mm &= mm - 1) {
        const int k = __builtin_ctz(mm);
        total += sub_run[s0 + k + 1] - sub_run[s0 + k];
    }
    return total;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// stream_chunk for d <= 3 with the masked sums on the matrix cores.  Each
// source q is one step: the lanes test it against eps (the same FP32
// distance as below), and two v_mfma_f32_4x4x1_16b add the row
// [w sv_0..2 | w | w cv_0..2 | 0] of q (w = run length) times each lane's
// 0/1 test result into 64 x 8 accumulators (16 blocks of 4 targets x 4
// columns per instruction: block b holds targets 4(b&7)+i, +32 in the second
// instruction, and columns 4(b>>3)+j; lane l supplies row l&3 of block l>>2's
// A and column l&3 of its B, and holds that block's rows 0-3 of column l&3 in
// its four result registers).  K = 1, so every accumulator takes one
// FP32 add per source, of the value or of 0, in source order: the sums are
// those of the scalar loop below bit for bit, the VALU is spared the six
// masked adds and the count per source.  A lane's sums enter the accumulator
// layout through LDS before the chunk and leave it after.  The count column
// is exact while a chunk's count stays below 2^24, hence n < 2^24.
template <int D>
__device__ __forceinline__ void stream_chunk_mfma(int base, int total, int r0, int s0,
                                                  unsigned smask, const int *__restrict__ sub_run,
                                                  const int *__restrict__ run_start,
                                                  const float *__restrict__ Xs, int n,
                                                  const float (&a0)[D], const float (&x)[D],
                                                  float thr, bool acc, SrcPt<D> *src,
                                                  float (&ssv)[D], float (&scm)[D], int &m_part) {
    static_assert(D <= 3 && SrcPt<D>::P == 4, "row of a source: sv[0..3] then cv[0..3]");
    const int lane = __lane_id();
    // 64 x 8 transpose buffer inside src, rows of kTr = 9 floats: an odd
    // stride puts a row-wise access of 32 lanes in 32 different banks, and
    // the accumulator-layout access too (9 * 4k + j covers 0..31 mod 32)
    constexpr int kTr = 9;
    static_assert(kWave * kTr <= kWave * (int)(sizeof(SrcPt<D>) / sizeof(float)), "T fits in src");
    float *T = reinterpret_cast<float *>(src);
    // the accumulators this lane holds: targets tc + r (and 32 + tc + r), column cc
    const int tc = lane & 28, cc = 4 * (lane >> 5) + (lane & 3);
#pragma unroll
    for (int l = 0; l < 4; l++) {
        T[lane * kTr + l] = l < D ? ssv[l] : 0.f;
        T[lane * kTr + 4 + l] = l < D ? scm[l] : 0.f;
    }
    T[lane * kTr + 3] = 0.f;  // count column: this chunk's count only
    wave_sync();
    f32x4 d1, d2;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        d1[r] = T[(tc + r) * kTr + cc];
        d2[r] = T[(32 + tc + r) * kTr + cc];
    }
    wave_sync();
    const int rr = stream_run(base + lane, total, r0, s0, smask, sub_run);
    if (rr >= 0) {
        const int j = run_start[rr];
        const float w = (float)(run_start[rr + 1] - j);  // exact: < 2^24
        float sv[D], cv[D];
#pragma unroll
        for (int l = 0; l < D; l++) {
            const float xq = Xs[(long long)l * n + j];
            src[lane].x[l] = xq;
            sincosm1_small(xq - a0[l], sv[l], cv[l]);
        }
#pragma unroll
        for (int l = 0; l < 4; l++) {
            src[lane].sv[l] = l < D ? w * sv[l] : 0.f;
            src[lane].cv[l] = l < D ? w * cv[l] : 0.f;
        }
        src[lane].sv[3] = w;
    }
    wave_sync();
    const int m = __builtin_amdgcn_readfirstlane(min(kWave, total - base));
    const float thr_l = acc ? thr : -1.f;  // lanes without acc add zeros
    constexpr int kS = sizeof(SrcPt<D>) / sizeof(float);
    const float *row = src[0].sv;  // row[q * kS + c]: column c of source q
    // per source: the lanes' 0/1 test results, lanes 0-31's to both halves of
    // the first MFMA's operand (targets 0-31) and 32-63's to the second's
    int k = 0;
    for (; k + 4 <= m; k += 4) {  // four sources per step: their LDS reads first
        float q[4][D], b[4], a[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
#pragma unroll
            for (int l = 0; l < D; l++) q[jj][l] = src[k + jj].x[l];
            b[jj] = row[(k + jj) * kS + cc];
        }
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            float sq = 0.f;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float df = x[l] - q[jj][l];
                sq += df * df;
            }
            a[jj] = sq <= thr_l ? 1.f : 0.f;
        }
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const auto pr = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[jj]),
                                                             __float_as_uint(a[jj]), false, false);
            d1 = __builtin_amdgcn_mfma_f32_4x4x1f32(__uint_as_float(pr[0]), b[jj], d1, 0, 0, 0);
            d2 = __builtin_amdgcn_mfma_f32_4x4x1f32(__uint_as_float(pr[1]), b[jj], d2, 0, 0, 0);
        }
    }
    for (; k < m; k++) {
        float sq = 0.f;
#pragma unroll
        for (int l = 0; l < D; l++) {
            const float df = x[l] - src[k].x[l];
            sq += df * df;
        }
        const float a = sq <= thr_l ? 1.f : 0.f;
        const auto pr =
            __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(a), false, false);
        const float b = row[k * kS + cc];
        d1 = __builtin_amdgcn_mfma_f32_4x4x1f32(__uint_as_float(pr[0]), b, d1, 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f32_4x4x1f32(__uint_as_float(pr[1]), b, d2, 0, 0, 0);
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        T[(tc + r) * kTr + cc] = d1[r];
        T[(32 + tc + r) * kTr + cc] = d2[r];
    }
    wave_sync();
#pragma unroll
    for (int l = 0; l < D; l++) {
        ssv[l] = T[lane * kTr + l];
        scm[l] = T[lane * kTr + 4 + l];
    }
    m_part += (int)T[lane * kTr + 3];
    wave_sync();
}

// One chunk of a streamed range: runs [base, base + 64) of it go through
// LDS (each lane loads one run and computes sin v and cos v - 1 of its
// offset v from the anchor a0); lanes with `acc` test each run against eps
// and accumulate it, weighted by its length (one rounding per run instead of
// one per point).
template <int D>
__device__ __forceinline__ void stream_chunk(int base, int total, int r0, int s0, unsigned smask,
                                             const int *__restrict__ sub_run,
                                             const int *__restrict__ run_start,
                                             const float *__restrict__ Xs, int n,
                                             const float (&a0)[D], const float (&x)[D],
                                             float thr, bool acc, bool any_acc, SrcPt<D> *src,
                                             int *len, float (&ssv)[D], float (&scm)[D],
                                             int &m_part) {
    if constexpr (D <= 3) {
        if (any_acc && n < (1 << 24)) {
            stream_chunk_mfma<D>(base, total, r0, s0, smask, sub_run, run_start, Xs, n, a0, x, thr,
                                 acc, src, ssv, scm, m_part);
            return;
        }
    }
    const int lane = __lane_id();
    const int rr = stream_run(base + lane, total, r0, s0, smask, sub_run);
    int my_len = 1;
    if (rr >= 0) {
        const int j = run_start[rr];
        my_len = run_start[rr + 1] - j;
        len[lane] = my_len;
#pragma unroll
        for (int l = 0; l < D; l++) {
            const float xq = Xs[(long long)l * n + j];
            src[lane].x[l] = xq;
            if (any_acc) sincosm1_small(xq - a0[l], src[lane].sv[l], src[lane].cv[l]);
        }
    }
    const bool multi = __any(my_len > 1);
    wave_sync();
    const int m = min(kWave, total - base);
    if (acc && !multi) {
        // single-point runs: four sources per step so their LDS reads overlap
        int k = 0;
        for (; k + 4 <= m; k += 4) {
            float sq[4];
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                float qx[D];
#pragma unroll
                for (int l = 0; l < D; l++) qx[l] = src[k + jj].x[l];
                sq[jj] = 0.f;
#pragma unroll
                for (int l = 0; l < D; l++) {
                    const float df = x[l] - qx[l];
                    sq[jj] += df * df;
                }
            }
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                if (sq[jj] <= thr) {
#pragma unroll
                    for (int l = 0; l < D; l++) {
                        ssv[l] += src[k + jj].sv[l];
                        scm[l] += src[k + jj].cv[l];
                    }
                    m_part += 1;
                }
            }
        }
        for (; k < m; k++) {
            const SrcPt<D> &qq = src[k];
            float sq = 0.f;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float df = x[l] - qq.x[l];
                sq += df * df;
            }
            if (sq <= thr) {
#pragma unroll
                for (int l = 0; l < D; l++) {
                    ssv[l] += qq.sv[l];
                    scm[l] += qq.cv[l];
                }
                m_part += 1;
            }
        }
    } else if (acc) {
        for (int k = 0; k < m; k++) {
            const SrcPt<D> &qq = src[k];
            float sq = 0.f;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float df = x[l] - qq.x[l];
                sq += df * df;
            }
            if (sq <= thr) {
                const int lk = len[k];
                const float w = (float)lk;  // exact: lk < 2^24
#pragma unroll
                for (int l = 0; l < D; l++) {
                    const float a = w * qq.sv[l];
                    const float b = w * qq.cv[l];
                    ssv[l] += a;
                    scm[l] += b;
                }
                m_part += lk;
            }
        }
    }
    wave_sync();
}

// Heavy ranges are summed in fixed point (2^-30 units, int64): integer
// additions commute, so a run's sums do not depend on which wavefront adds
// which unit, or in what order.  |sum| <= 2n <= 2^31 keeps within 2^61.
constexpr double kFix = 0x1p30;

__device__ __forceinline__ unsigned long long to_fix(float v) {
    return (unsigned long long)__double2ll_rn((double)v * kFix);
}

__device__ __forceinline__ float from_fix(long long v) { return (float)((double)v / kFix); }

template <int D>
__device__ __forceinline__ void add_fixed(long long *__restrict__ hacc, int n, int r,
                                          const float (&ps)[D], const float (&pc)[D], int pm) {
    unsigned long long *h = (unsigned long long *)hacc;
#pragma unroll
    for (int l = 0; l < D; l++) {
        atomicAdd(h + (long long)l * n + r, to_fix(ps[l]));
        atomicAdd(h + (long long)(D + l) * n + r, to_fix(pc[l]));
    }
    atomicAdd(h + (long long)(2 * D) * n + r, (unsigned long long)(long long)pm);
}

// The tile's per-lane setup: position (the anchor for lanes without a run),
// anchor a0 -- the centre of the tile's cell in the outer dimensions and of
// its block of kAnchorBlock cells in the inner ones, the same for every
// group width that divides kAnchorBlock -- and sin / cos of u = x - a0.
template <int D>
__device__ __forceinline__ void tile_setup(const TileLane &tl, const float *__restrict__ Xs,
                                           const int *__restrict__ cell_g, const Geo &g,
                                           int (&g0)[D], float (&a0)[D], float (&x)[D],
                                           float (&su)[D], float (&cu)[D]) {
    const int n = g.n;
#pragma unroll
    for (int l = 0; l < D; l++) {
        g0[l] = cell_g[(long long)l * n + tl.gc0];
        a0[l] = l < g.od ? ((float)g0[l] + 0.5f) * g.cs
                         : ((float)(g0[l] & ~(kAnchorBlock - 1)) + 0.5f * kAnchorBlock) * g.cs;
        x[l] = tl.active ? Xs[(long long)l * n + tl.i] : a0[l];
        sincos_small(x[l] - a0[l], su[l], cu[l]);
    }
}

// Final step of a run's update (EGG_SynC.cu:701-709): new position into
// rres, diagnostics; returns the local-sync test.
template <int D>
__device__ __forceinline__ bool finish_run(const float (&x)[D], const float (&su)[D],
                                           const float (&cu)[D], const float (&sum)[D],
                                           const float (&ssv)[D], const float (&scm)[D], int N,
                                           int m_part, int half, bool border, int own, int n,
                                           int r, float *__restrict__ rres,
                                           int *__restrict__ rdbg) {
    N += m_part;
    float xo[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
        const float p1 = ssv[l] * cu[l];
        const float p2 = ((float)m_part + scm[l]) * su[l];
        const float sml = p1 - p2;
        float tt = sum[l] + sml;
        tt = tt / (float)N;
        xo[l] = x[l] + tt;
    }
    const bool synced = !((!border && N != own) || (border && N != half));
#pragma unroll
    for (int l = 0; l < D; l++) rres[(long long)l * n + r] = xo[l];
    if (rdbg) {
        rdbg[4 * (long long)r + 0] = N;
        rdbg[4 * (long long)r + 1] = half;
        rdbg[4 * (long long)r + 2] = border ? 1 : 0;
        rdbg[4 * (long long)r + 3] = synced ? 1 : 0;
    }
    return synced;
}

// kernel_itr_grid_STAD5_5_1 (EGG_SynC.cu:551-763).  Per lane: full cells add
// the cell totals (EGG_SynC.cu:648-663); partial cells are streamed through
// LDS in chunks of 64 runs -- each lane loads one run and computes sin v and
// cos v - 1 of v = x_q - a0 (a0 = the tile's anchor, near the tile's cells) --
// and every lane reads each run as a broadcast and, for runs within eps,
// accumulates sin v and cos v - 1 times the run's length.  At the end of
// the tile, sum_q sin(x_q - x) = cos u * sum sin v - sin u * (m + sum (cos v - 1)),
// u = x - a0, m = number of such points.  All angles stay within a few cell
// sides of the anchor, and both sums grow like the sum of sin(x_q - x) does,
// so the result keeps the accuracy of a direct sum of sinf(x_q - x).  Border
// lanes count their eps/2-neighbours in other full cells from the same
// chunks (EGG_SynC.cu:666-681).
// Up to 3 dimensions the compiler is held to 6 wavefronts per SIMD (80
// VGPRs, a few spills: the tile's constants); more resident wavefronts hide
// the LDS and L2 latency of the streaming loop (config 1: -8%).
template <int D>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D <= 3 ? 6 : 1)))
void k_update_tile(
    const float *__restrict__ Xs, const int *__restrict__ cell_start,
    const int *__restrict__ cell_g, const float *__restrict__ Ts, const float *__restrict__ Tc,
    const int *__restrict__ ob, const int *__restrict__ oe, const int *__restrict__ run_start,
    const int *__restrict__ cell_run, const int *__restrict__ grp_cell,
    const int *__restrict__ point_cell, const int *__restrict__ tile_off,
    const int *__restrict__ cell_sub, const int *__restrict__ sub_start,
    const unsigned *__restrict__ smin, const unsigned *__restrict__ smax,
    const int *__restrict__ sub_run, const float *__restrict__ TsS,
    const float *__restrict__ TcS, float *__restrict__ rres, int *__restrict__ rdbg,
    long long *__restrict__ hacc,
    HeavyDesc *__restrict__ hdesc, int *__restrict__ dtiles, int *__restrict__ dsave, Geo g,
    DevState *st) {
    if (st->done) return;
    // heavy ranges go to k_update_heavy only up to kHeavyMaxD dimensions
    // (the deferral code costs registers)
    constexpr bool kDefer = D <= kHeavyMaxD;
    __shared__ SrcPt<D> s_src[kBlock / kWave][kWave];
    __shared__ int s_len[kBlock / kWave][kWave];
    const int n = g.n, R = g.R, od = g.od;
    const int G = st->G, T = st->TU;
    const float thr = g.thr_eps, thr_half = g.thr_half;
    // sub-cell classification bounds (sub_class); their own rounding keeps a
    // margin of at least 63 * 2^-24
    const float thr_in = thr * (1.0f - 64.0f * 0x1p-24f);
    const float thr_out = thr * (1.0f + 64.0f * 0x1p-24f);
    const float thr_half_in = thr_half * (1.0f - 64.0f * 0x1p-24f);
    const float thr_half_out = thr_half * (1.0f + 64.0f * 0x1p-24f);
    const int lane = __lane_id();
    SrcPt<D> *src = s_src[threadIdx.x / kWave];
    int *len = s_len[threadIdx.x / kWave];
    const int *cg_od = cell_g + (long long)od * n;
    bool unsynced = false;
    // tiles of the non-sparse groups only (tile_off: their scan); sparse
    // groups go to k_update_pairs
    for (int t = wave_index(); t < T; t = wave_count() + next_tile(&st->tile_next)) {
        const TileLane tl = tile_lane(t, tile_off, G, grp_cell, cell_run, run_start, point_cell);
        const int cl = tl.cl, r = tl.r;
        const bool active = tl.active;
        const int own = cell_start[cl + 1] - cell_start[cl];
        const int ghi = cell_g[(long long)(D - 1) * n + tl.gc1 - 1];
        float x[D], a0[D], sx[D], cx[D], su[D], cu[D], sum[D], ssv[D], scm[D];
        int g0[D];
        tile_setup<D>(tl, Xs, cell_g, g, g0, a0, x, su, cu);
        bool border = false;
#pragma unroll
        for (int l = 0; l < D; l++) {
            sum[l] = 0.f;
            ssv[l] = 0.f;
            scm[l] = 0.f;
            sx[l] = sinf(x[l]);
            cx[l] = cosf(x[l]);
            border |= active && on_border(x[l], g.cs);
        }
        const int g0od = select_dim<D>(g0, od);
        int N = 0, half = 0, m_part = 0;
        bool deferred = false;  // some range went to k_update_heavy
        int cur_c = 0;          // candidate cell of the range being streamed
        auto defer_tile = [&]() __attribute__((always_inline)) {
            if (!deferred) {
                deferred = true;
                if (active) {
#pragma unroll
                    for (int k = 0; k < 2 * D + 2; k++) hacc[(long long)k * n + r] = 0;
                }
                if (lane == 0) dtiles[atomicAdd(&st->ndt, 1)] = t;
            }
        };
        // queues a heavy range; false if the queue is full
        auto queue_range = [&](int r0, int s0, unsigned smask, int total,
                               int half_kind) __attribute__((always_inline)) {
            const int nunits = (total + kUnitRuns - 1) / kUnitRuns;
            unsigned long long old = 0;
            if (lane == 0)
                old = atomicAdd(&st->heavy_pack, (1ull << 32) | (unsigned long long)nunits);
            const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)old);
            const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(old >> 32));
            if (hi >= (unsigned)kHeavyCap) return false;
            if (lane == 0)
                hdesc[hi] = HeavyDesc{t, cur_c, r0, s0, smask, total, (int)lo, nunits, half_kind};
            return true;
        };
        // Streams a range of runs (stream_run) through LDS, 64 per chunk.  A
        // range of kHeavyRuns runs or more is queued for k_update_heavy
        // instead, in units any wavefront can take; the tile's runs then
        // finish in k_update_finish.
        auto scan_runs = [&](int r0, int r1, int s0, unsigned smask,
                             bool acc) __attribute__((always_inline)) {
            const bool any_acc = __any(acc);
            const int total = stream_total(r0, r1, s0, smask, sub_run);
            const bool heavy = kDefer && hacc && total >= kHeavyRuns;
            if (heavy) {
                defer_tile();
                // the float sums so far go to hacc now, so that what follows
                // does not depend on who sums the range (all sums are additive)
                if (active) add_fixed<D>(hacc, n, r, ssv, scm, m_part);
#pragma unroll
                for (int l = 0; l < D; l++) ssv[l] = scm[l] = 0.f;
                m_part = 0;
                if (queue_range(r0, s0, smask, total, 0)) return;
                // queue full: this wavefront sums the range itself, unit by
                // unit in fixed point, as k_update_heavy would
            }
            for (int base = 0; base < total; base += kWave) {
                stream_chunk<D>(base, total, r0, s0, smask, sub_run, run_start, Xs, n, a0, x,
                                thr, acc, any_acc, src, len, ssv, scm, m_part);
                if (heavy && (base + kWave == total || (base + kWave) % kUnitRuns == 0 ||
                              base + kWave > total)) {
                    if (active) add_fixed<D>(hacc, n, r, ssv, scm, m_part);
#pragma unroll
                    for (int l = 0; l < D; l++) ssv[l] = scm[l] = 0.f;
                    m_part = 0;
                }
            }
        };
        // Lanes with `hcnt` count the points of a range of runs (stream_run)
        // within eps/2.
        auto half_scan = [&](int r0, int r1, int s0, unsigned smask,
                             bool hcnt) __attribute__((always_inline)) {
            const int total = stream_total(r0, r1, s0, smask, sub_run);
            if (kDefer && hacc && total >= kHeavyRuns) {
                // a count: exact in any order, so k_update_heavy may take it
                defer_tile();
                if (queue_range(r0, s0, smask, total, 1)) return;
            }
            for (int base = 0; base < total; base += kWave) {
                const int rr = stream_run(base + lane, total, r0, s0, smask, sub_run);
                if (rr >= 0) {
                    const int j = run_start[rr];
                    len[lane] = run_start[rr + 1] - j;
#pragma unroll
                    for (int l = 0; l < D; l++) src[lane].x[l] = Xs[(long long)l * n + j];
                }
                wave_sync();
                const int m = min(kWave, total - base);
                if (hcnt) {
                    for (int k = 0; k < m; k++) {
                        float sq = 0.f;
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            const float df = x[l] - src[k].x[l];
                            sq += df * df;
                        }
                        if (sq <= thr_half) half += len[k];
                    }
                }
                wave_sync();
            }
        };
        walk_columns<D>(g0, R, g, ob, oe, [&](int cb, const int ce) {
            if (od < D) cb = column_lower(cg_od, cb, ce, g0od - R - 1);
            for (int c = cb; c < ce; c++) {
                int gc[D];
                bool stop;
                const bool in =
                    cell_in_group_window<D>(cell_g, n, c, g0, od, R + 1, ghi, gc, &stop);
                if (stop) break;
                if (!in) continue;
                float sm, lg;
                box_dist<D>(gc, x, g.cs, sm, lg);
                const bool full = active && lg <= thr;
                const bool part = active && !full && sm <= thr;
                if (full) {
#pragma unroll
                    for (int l = 0; l < D; l++) {
                        const float a = Ts[(long long)l * n + c] * cx[l];
                        const float b = Tc[(long long)l * n + c] * sx[l];
                        const float tt = a - b;
                        sum[l] += tt;
                    }
                    const int cnt = cell_start[c + 1] - cell_start[c];
                    N += cnt;
                    if (border && c == cl) half += cnt;
                }
                const bool need_half = full && border && c != cl;
                const bool any_part = __any(part), any_half = __any(need_half);
                if (!any_part && !any_half) continue;
                if (any_half) {
                    cur_c = c;
                    // eps/2-neighbours in another full cell: sub-cells wholly
                    // within eps/2 of a lane count whole, those straddling
                    // eps/2 for some lane are streamed for the wavefront
                    const int hs0 = cell_sub[c], hns = cell_sub[c + 1] - hs0;
                    if (g.sub_bits && hns <= 32) {
                        unsigned hin = 0, hst = 0;
                        stage_boxes<D>(smin, smax, n, hs0, hns, src);
                        if (need_half) {
                            for (int k = 0; k < hns; k++) {
                                const int cls = sub_class_box<D>(src[k].x, src[k].sv, x,
                                                                 thr_half_in, thr_half_out);
                                if (cls == 1) hin |= 1u << k;
                                else if (cls == 2) hst |= 1u << k;
                            }
                        }
                        wave_sync();
                        const unsigned hu = wave_or(hst);
                        const unsigned ht = hin & ~hu;
                        for (unsigned mm = wave_or(ht); mm; mm &= mm - 1) {
                            const int k = __builtin_ctz(mm);
                            if ((ht >> k) & 1u) half += sub_start[hs0 + k + 1] - sub_start[hs0 + k];
                        }
                        if (hu) half_scan(0, 0, hs0, hu, need_half);
                    } else {
                        half_scan(cell_run[c], cell_run[c + 1], 0, 0u, need_half);
                    }
                }
                if (!any_part) continue;
                const int s0 = cell_sub[c], ns = cell_sub[c + 1] - s0;
                unsigned umask = 0;
                if (g.sub_bits && ns <= 32) {
                    // sub-cells of a partial cell: each lane classifies them
                    // (wholly within eps / wholly beyond / straddling); the
                    // sub-cells some lane straddles are streamed once for the
                    // wavefront and tested point by point by every lane with
                    // a partial cell; other sub-cells wholly within eps add
                    // their totals
                    unsigned imask = 0, smask = 0;
                    stage_boxes<D>(smin, smax, n, s0, ns, src);
                    if (part) {
                        for (int k = 0; k < ns; k++) {
                            const int cls =
                                sub_class_box<D>(src[k].x, src[k].sv, x, thr_in, thr_out);
                            if (cls == 1) imask |= 1u << k;
                            else if (cls == 2) smask |= 1u << k;
                        }
                    }
                    wave_sync();
                    umask = wave_or(smask);
                    const unsigned tmask = imask & ~umask;
                    for (unsigned mm = wave_or(tmask); mm; mm &= mm - 1) {
                        const int k = __builtin_ctz(mm);
                        const int sc = s0 + k;
                        if ((tmask >> k) & 1u) {
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                const float a = TsS[(long long)l * n + sc] * cx[l];
                                const float b = TcS[(long long)l * n + sc] * sx[l];
                                const float tt = a - b;
                                sum[l] += tt;
                            }
                            N += sub_start[sc + 1] - sub_start[sc];
                        }
                    }
                    if (!umask) continue;
                }
                cur_c = c;
                scan_runs(cell_run[c], cell_run[c + 1], s0, umask, part);
            }
        });
        if (deferred) {
            // the sums so far wait for the heavy ranges: full-cell totals in
            // rres, N / half / border in dsave, streamed sums into hacc
            if (active) {
#pragma unroll
                for (int l = 0; l < D; l++) rres[(long long)l * n + r] = sum[l];
                dsave[4 * (long long)r + 0] = N;
                dsave[4 * (long long)r + 1] = half;
                dsave[4 * (long long)r + 2] = border ? 1 : 0;
                add_fixed<D>(hacc, n, r, ssv, scm, m_part);
            }
            continue;
        }
        if (active) {
            if (!finish_run<D>(x, su, cu, sum, ssv, scm, N, m_part, half, border, own, n, r, rres,
                               rdbg))
                unsynced = true;
        }
    }
    if (__any(unsynced) && lane == 0) atomicAnd(&st->flag, 0);
}

// Work units of the heavy ranges (kUnitRuns runs each), from a queue over
// all of them: every wavefront rebuilds the tile's lanes, tests the unit's
// runs as the tile kernel would and adds its sums in fixed point.
template <int D>
__global__ __launch_bounds__(kBlock) void k_update_heavy(
    const float *__restrict__ Xs, const int *__restrict__ cell_g,
    const int *__restrict__ run_start, const int *__restrict__ cell_run,
    const int *__restrict__ grp_cell, const int *__restrict__ point_cell,
    const int *__restrict__ tile_off, const int *__restrict__ sub_run,
    const HeavyDesc *__restrict__ hdesc, long long *__restrict__ hacc, Geo g, DevState *st) {
    if (st->done) return;
    const int nd = (int)min((unsigned long long)kHeavyCap, st->heavy_pack >> 32);
    if (nd == 0) return;
    __shared__ SrcPt<D> s_src[kBlock / kWave][kWave];
    __shared__ int s_len[kBlock / kWave][kWave];
    SrcPt<D> *src = s_src[threadIdx.x / kWave];
    int *len = s_len[threadIdx.x / kWave];
    const int n = g.n, G = st->G;
    const float thr = g.thr_eps;
    // descriptors k < nd hold consecutive unit ranges in the order of k
    const int end = hdesc[nd - 1].ubase + hdesc[nd - 1].nunits;
    for (int u = wave_index(); u < end; u = wave_count() + next_tile(&st->unit_next)) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (hdesc[mid].ubase <= u) lo = mid;
            else hi = mid - 1;
        }
        const HeavyDesc hd = hdesc[lo];
        const TileLane tl = tile_lane(hd.t, tile_off, G, grp_cell, cell_run, run_start,
                                      point_cell);
        float x[D], a0[D], su[D], cu[D];
        int g0[D], gc[D];
        tile_setup<D>(tl, Xs, cell_g, g, g0, a0, x, su, cu);
#pragma unroll
        for (int l = 0; l < D; l++) gc[l] = cell_g[(long long)l * n + hd.c];
        float sm, lg;
        box_dist<D>(gc, x, g.cs, sm, lg);
        const int k = u - hd.ubase;
        const int b1 = min(hd.total, (k + 1) * kUnitRuns);
        if (hd.half) {
            // border lanes of the tile, the range's cell full for them and
            // not their own: points within eps/2
            bool border = false;
#pragma unroll
            for (int l = 0; l < D; l++) border |= on_border(x[l], g.cs);
            const bool hcnt = tl.active && lg <= thr && border && hd.c != tl.cl;
            int cnt = 0;
            for (int base = k * kUnitRuns; base < b1; base += kWave) {
                const int rr = stream_run(base + __lane_id(), b1, hd.r0, hd.s0, hd.smask,
                                          sub_run);
                if (rr >= 0) {
                    const int j = run_start[rr];
                    len[__lane_id()] = run_start[rr + 1] - j;
#pragma unroll
                    for (int l = 0; l < D; l++)
                        src[__lane_id()].x[l] = Xs[(long long)l * n + j];
                }
                wave_sync();
                const int m = min(kWave, b1 - base);
                if (hcnt) {
                    for (int q = 0; q < m; q++) {
                        float sq = 0.f;
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            const float df = x[l] - src[q].x[l];
                            sq += df * df;
                        }
                        if (sq <= g.thr_half) cnt += len[q];
                    }
                }
                wave_sync();
            }
            if (hcnt)
                atomicAdd((unsigned long long *)hacc + (long long)(2 * D + 1) * n + tl.r,
                          (unsigned long long)cnt);
            continue;
        }
        const bool acc = tl.active && !(lg <= thr) && sm <= thr;
        const bool any_acc = __any(acc);
        float ps[D], pc[D];
        int pm = 0;
#pragma unroll
        for (int l = 0; l < D; l++) ps[l] = pc[l] = 0.f;
        for (int base = k * kUnitRuns; base < b1; base += kWave)
            stream_chunk<D>(base, b1, hd.r0, hd.s0, hd.smask, sub_run, run_start, Xs, n, a0, x,
                            thr, acc, any_acc, src, len, ps, pc, pm);
        if (acc) add_fixed<D>(hacc, n, tl.r, ps, pc, pm);
    }
}

// The runs of tiles with heavy ranges: their sums from hacc, then the
// tile kernel's final step.
template <int D>
__global__ __launch_bounds__(kBlock) void k_update_finish(
    const float *__restrict__ Xs, const int *__restrict__ cell_start,
    const int *__restrict__ cell_g, const int *__restrict__ run_start,
    const int *__restrict__ cell_run, const int *__restrict__ grp_cell,
    const int *__restrict__ point_cell, const int *__restrict__ tile_off,
    const int *__restrict__ dtiles, const int *__restrict__ dsave,
    const long long *__restrict__ hacc, float *__restrict__ rres, int *__restrict__ rdbg, Geo g,
    DevState *st) {
    if (st->done) return;
    const int ndt = st->ndt, n = g.n, G = st->G;
    const int waves = gridDim.x * (blockDim.x / kWave);
    bool unsynced = false;
    for (int w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; w < ndt; w += waves) {
        const TileLane tl = tile_lane(dtiles[w], tile_off, G, grp_cell, cell_run, run_start,
                                      point_cell);
        if (!tl.active) continue;
        const int r = tl.r;
        float x[D], a0[D], su[D], cu[D], sum[D], ssv[D], scm[D];
        int g0[D];
        tile_setup<D>(tl, Xs, cell_g, g, g0, a0, x, su, cu);
#pragma unroll
        for (int l = 0; l < D; l++) {
            sum[l] = rres[(long long)l * n + r];
            ssv[l] = from_fix(hacc[(long long)l * n + r]);
            scm[l] = from_fix(hacc[(long long)(D + l) * n + r]);
        }
        const int m_part = (int)hacc[(long long)(2 * D) * n + r];
        const int own = cell_start[tl.cl + 1] - cell_start[tl.cl];
        const int half = dsave[4 * (long long)r + 1] + (int)hacc[(long long)(2 * D + 1) * n + r];
        if (!finish_run<D>(x, su, cu, sum, ssv, scm, dsave[4 * (long long)r], m_part, half,
                           dsave[4 * (long long)r + 2] != 0, own, n, r, rres, rdbg))
            unsynced = true;
    }
    if (__any(unsynced) && __lane_id() == 0) atomicAnd(&st->flag, 0);
}

// ---------------------------------------------------------- pair update --
//
// The update of the runs of the cell groups sparse_group selects (the runs
// of sparse groups, or every run), point pair by point pair instead of cell
// by cell.  `list` holds the target runs (nullptr: runs 0 .. count-1).  In a sparse grid (high-dimensional
// data: 250k points in 248k cells at d = 8) the neighbour cells of a point are
// the inner cells of (2R+1)^od outer cells, almost all holding a point or
// none, and walking them costs far more than testing the points they hold.
// Here every run of the grid is a source: sources come in blocks of kBlkRuns
// consecutive runs (sorted order, so spatially compact) with bounding boxes,
// and blocks in superblocks of kSupBlks.  One wavefront takes a tile of 64
// target runs, prunes superblocks and blocks whose box lies farther from the
// tile's box than any counted point can, and streams the remaining blocks
// through LDS: the loading lane stages each source run's position, the
// bounds of its cell's box (g cs and (g+1) cs, rounded as the reference rounds
// them) and sin v, cos v - 1 of its offset v from the tile's anchor; every
// lane then tests every staged run against its target.
//
// A pair within eps, or with a clamped source (such a point lies outside its
// cell's box), gets the reference's cell decision for the source's cell
// (EGG_SynC.cu:612-636): a full cell adds its totals (T_s cos x - T_c sin x,
// EGG_SynC.cu:648-663) and its count once, when the lane meets the cell's
// first run -- every run of a full cell is within the filter radius of the
// target (it lies in the box, whose farthest corner is within eps, and FP32
// rounding is monotone), so that run is always met; a partial cell adds
// sin(x_q - x) and the length of each run within eps (EGG_SynC.cu:683-696);
// neither -> nothing.  Only cells whose outer coordinates are within R of the
// target's are candidates, as in the reference's neighbour list
// (EGG_SynC.cu:368-536).  Blocks are visited in increasing order, which is
// the reference's order of outer cells, cells and points, so the sums are
// added in the oracle's order (runs weighted by their length); N and half are
// exactly the reference's.

// Per run: its cell, flags (bit 0: the first run of its cell, bit 1: a
// coordinate clamped into cell w-1), position, cell coordinates and length,
// indexed by run so that the pair update stages a block with one coalesced
// load per field; per block of kBlkRuns runs: the bounding box of its
// positions, whether it holds a clamped run, and its cell if all its runs
// share one (-1 otherwise).  One wavefront per block.
template <int D>
__global__ __launch_bounds__(kBlock) void k_run_boxes(
    const float *__restrict__ Xs, const int *__restrict__ run_start,
    const int *__restrict__ point_cell, const int *__restrict__ cell_run,
    const int *__restrict__ cell_g, Geo g, DevState *st, int *__restrict__ rcell,
    int *__restrict__ rflag, float *__restrict__ Xr, int *__restrict__ gr,
    int *__restrict__ rlen, float *__restrict__ bmin, float *__restrict__ bmax,
    int *__restrict__ bclamp, int *__restrict__ bcell) {
    if (st->done || st->sparse_count == 0) return;
    const int n = g.n, runs = st->runs;
    const int nblk = (runs + kBlkRuns - 1) / kBlkRuns, nb = n / kBlkRuns + 1;
    const float wf = (float)g.w;
    for (int b = wave_index(); b < nblk; b += wave_count()) {
        const int r = b * kBlkRuns + __lane_id();
        const bool on = r < runs;
        float lo[D], hi[D];
        int f = 0, c = -1;
        if (on) {
            const int i = run_start[r];
            c = point_cell[i];
            bool clamped = false;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float x = Xs[(long long)l * n + i];
                lo[l] = hi[l] = x;
                clamped |= x / g.cs >= wf;  // cell_coord's clamp of w to w-1
                Xr[(long long)l * n + r] = x;
                gr[(long long)l * n + r] = cell_g[(long long)l * n + c];
            }
            f = (cell_run[c] == r ? 1 : 0) | (clamped ? 2 : 0);
            rcell[r] = c;
            rflag[r] = f;
            rlen[r] = run_start[r + 1] - i;
        } else {
#pragma unroll
            for (int l = 0; l < D; l++) {
                lo[l] = __builtin_inff();
                hi[l] = -__builtin_inff();
            }
        }
        const bool any_clamp = __any((f & 2) != 0);
        const int c_first = __builtin_amdgcn_readfirstlane(c);
        const bool one_cell = !__any(on && c != c_first);
#pragma unroll
        for (int l = 0; l < D; l++) {
            for (int o = kWave / 2; o > 0; o >>= 1) {
                lo[l] = fminf(lo[l], __shfl_xor(lo[l], o, kWave));
                hi[l] = fmaxf(hi[l], __shfl_xor(hi[l], o, kWave));
            }
            if (__lane_id() == 0) {
                bmin[(long long)l * nb + b] = lo[l];
                bmax[(long long)l * nb + b] = hi[l];
            }
        }
        if (__lane_id() == 0) {
            bclamp[b] = any_clamp ? 1 : 0;
            bcell[b] = one_cell ? c_first : -1;
        }
    }
}

// Bounding box of every superblock of kSupBlks blocks, and whether it holds a
// clamped run.
template <int D>
__global__ __launch_bounds__(kBlock) void k_super_boxes(
    const float *__restrict__ bmin, const float *__restrict__ bmax,
    const int *__restrict__ bclamp, Geo g, DevState *st, float *__restrict__ sbmin,
    float *__restrict__ sbmax, int *__restrict__ sbclamp) {
    if (st->done || st->sparse_count == 0) return;
    const int n = g.n, runs = st->runs;
    const int nblk = (runs + kBlkRuns - 1) / kBlkRuns, nb = n / kBlkRuns + 1;
    const int nsup = (nblk + kSupBlks - 1) / kSupBlks, ns = n / (kBlkRuns * kSupBlks) + 1;
    for (int sb = wave_index(); sb < nsup; sb += wave_count()) {
        const int b = sb * kSupBlks + __lane_id();
        const bool on = b < nblk;
        const bool any_clamp = __any(on && bclamp[b] != 0);
#pragma unroll
        for (int l = 0; l < D; l++) {
            float lo = on ? bmin[(long long)l * nb + b] : __builtin_inff();
            float hi = on ? bmax[(long long)l * nb + b] : -__builtin_inff();
            for (int o = kWave / 2; o > 0; o >>= 1) {
                lo = fminf(lo, __shfl_xor(lo, o, kWave));
                hi = fmaxf(hi, __shfl_xor(hi, o, kWave));
            }
            if (__lane_id() == 0) {
                sbmin[(long long)l * ns + sb] = lo;
                sbmax[(long long)l * ns + sb] = hi;
            }
        }
        if (__lane_id() == 0) sbclamp[sb] = any_clamp ? 1 : 0;
    }
}

// Whether box k of a (lo, hi) table ([D][stride]) lies within squared
// distance thr of the box [alo, ahi] (thr_clamp if the box holds a clamped
// run).  Box distances round like the point distances they bound; the
// caller's thresholds carry a relative margin of 2^-10 for that.
template <int D>
__device__ __forceinline__ bool box_near(const float (&alo)[D], const float (&ahi)[D],
                                         const float *__restrict__ blo,
                                         const float *__restrict__ bhi, long long stride, int k,
                                         float thr) {
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < D; l++) {
        const float gap = fmaxf(0.f, fmaxf(blo[l * stride + k] - ahi[l], alo[l] - bhi[l * stride + k]));
        s += gap * gap;
    }
    return s <= thr;
}

// Source runs of a block staged in LDS for the pair update.  Positions are
// kept twice: row-major for the distance tests, where every lane reads the
// same run (broadcast, 16-byte loads), and with the cell box bounds, cell
// coordinates, length, flags and cell in columns for the accepted pairs,
// where each lane reads a run of its own (consecutive runs in consecutive
// banks).
template <int D>
struct PairStage {
    static constexpr int P = (D + 3) / 4 * 4;
    float xr[kBlkRuns][P];
    float x[D][kBlkRuns], lo[D][kBlkRuns], hi[D][kBlkRuns];
    int g[D][kBlkRuns];
    int len[kBlkRuns], flag[kBlkRuns], cell[kBlkRuns];
};

// sin v: the polynomial of sin_small for |v| <= 1 when the caller knows that
// bound holds (SMALL), else sinf.
template <bool SMALL>
__device__ __forceinline__ float sin_pair(float v) {
    if (!SMALL) return sinf(v);
    const float v2 = v * v;
    float ps = -2.50521083854417e-08f;
    ps = fmaf(ps, v2, 2.75573192239859e-06f);
    ps = fmaf(ps, v2, -1.98412698412698e-04f);
    ps = fmaf(ps, v2, 8.33333333333333e-03f);
    ps = fmaf(ps, v2, -1.66666666666667e-01f);
    return fmaf(v * v2, ps, v);
}

// SMALL: every pair the filter passes is closer than 1 in each coordinate
// (filter radius below 1), so sin(x_q - x) takes the polynomial.
template <int D, bool SMALL>
__global__ __launch_bounds__(kBlock) void k_update_pairs(
    const float *__restrict__ Xs, const int *__restrict__ cell_start,
    const int *__restrict__ cell_g, const float *__restrict__ Ts, const float *__restrict__ Tc,
    const int *__restrict__ run_start, const int *__restrict__ point_cell,
    const int *__restrict__ cell_run, const int *__restrict__ list,
    const int *__restrict__ rcell, const int *__restrict__ rflag,
    const float *__restrict__ Xr, const int *__restrict__ gr, const int *__restrict__ rlen,
    const float *__restrict__ bmin, const float *__restrict__ bmax,
    const int *__restrict__ bclamp, const int *__restrict__ bcell,
    const float *__restrict__ sbmin,
    const float *__restrict__ sbmax, const int *__restrict__ sbclamp, float thr_filter,
    float thr_clamp, float *__restrict__ rres, int *__restrict__ rdbg, Geo g, DevState *st) {
    if (st->done) return;
    const int count = st->sparse_count;
    if (count == 0) return;
    __shared__ PairStage<D> s_stage[kBlock / kWave];
    PairStage<D> &S = s_stage[threadIdx.x / kWave];
    const int n = g.n, od = g.od, R = g.R, runs = st->runs, lane = __lane_id();
    const int nblk = (runs + kBlkRuns - 1) / kBlkRuns, nb = n / kBlkRuns + 1;
    const int nsup = (nblk + kSupBlks - 1) / kSupBlks, ns = n / (kBlkRuns * kSupBlks) + 1;
    const float thr = g.thr_eps, thr_half = g.thr_half;
    const float box_thr = thr_filter * (1.0f + 0x1p-10f);
    const float box_thr_clamp = thr_clamp * (1.0f + 0x1p-10f);
    const int ntiles = (count + kWave - 1) / kWave;
    bool unsynced = false;
    for (int t = wave_index(); t < ntiles; t = wave_count() + next_tile(&st->ptile_next)) {
        const int k = t * kWave + lane;
        const bool active = k < count;
        const int r = list ? list[active ? k : t * kWave] : (active ? k : t * kWave);
        const int i = run_start[r];
        const int c0 = point_cell[i];
        const int own = cell_start[c0 + 1] - cell_start[c0];
        float x[D], sum[D], sml[D], tlo[D], thi[D];
        int g0[D];
        bool border = false;
#pragma unroll
        for (int l = 0; l < D; l++) {
            x[l] = Xs[(long long)l * n + i];
            g0[l] = l < od ? cell_g[(long long)l * n + c0] : 0;
            sum[l] = 0.f;
            sml[l] = 0.f;
            border |= on_border(x[l], g.cs);
            tlo[l] = active ? x[l] : __builtin_inff();
            thi[l] = active ? x[l] : -__builtin_inff();
            for (int o = kWave / 2; o > 0; o >>= 1) {
                tlo[l] = fminf(tlo[l], __shfl_xor(tlo[l], o, kWave));
                thi[l] = fmaxf(thi[l], __shfl_xor(thi[l], o, kWave));
            }
        }
        int N = 0, half = 0;
        // the lane's decision for the last cell it classified: a cell's runs
        // are consecutive in block order, so one box test serves all of them
        int last_c = -1, last_cls = 0;  // 0 not counted, 1 full, 2 partial
        for (int sb0 = 0; sb0 < nsup; sb0 += kWave) {
            const int sbl = sb0 + lane;
            const bool sp = sbl < nsup &&
                            box_near<D>(tlo, thi, sbmin, sbmax, ns, sbl,
                                        sbclamp[sbl] ? box_thr_clamp : box_thr);
            for (unsigned long long ms = __ballot(sp); ms; ms &= ms - 1) {
                const int sb = sb0 + __ffsll((long long)ms) - 1;
                const int bl = sb * kSupBlks + lane;
                const bool bp = bl < nblk &&
                                box_near<D>(tlo, thi, bmin, bmax, nb, bl,
                                            bclamp[bl] ? box_thr_clamp : box_thr);
                for (unsigned long long mb = __ballot(bp); mb; mb &= mb - 1) {
                    const int b = sb * kSupBlks + __ffsll((long long)mb) - 1;
                    const int r0 = b * kBlkRuns;
                    // a block whose runs share one cell c (a dense cell): one
                    // box test of c per lane decides for all its runs, unless
                    // c is partial for the lane or a border lane needs its
                    // eps/2 count; if no lane needs the runs one by one, the
                    // block is not streamed, and lanes for which c is full add
                    // its totals in the block holding its first run
                    const int bc = bcell[b];
                    if (bc >= 0 && !bclamp[b]) {
                        bool need = false, add_full = false;
                        if (active) {
                            int gc[D];
                            bool win = true;
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                gc[l] = gr[(long long)l * n + r0];
                                if (l < od && (gc[l] - g0[l] > R || g0[l] - gc[l] > R)) win = false;
                            }
                            if (win) {
                                float sm, lg;
                                box_dist<D>(gc, x, g.cs, sm, lg);
                                const bool full = lg <= thr;
                                need = (!full && sm <= thr) || (full && border && bc != c0);
                                add_full = full;
                            }
                        }
                        if (!__any(need)) {
                            const int fr = cell_run[bc];
                            if (add_full && fr >= r0 && fr < r0 + kBlkRuns) {
                                const int cnt = cell_start[bc + 1] - cell_start[bc];
#pragma unroll
                                for (int l = 0; l < D; l++) {
                                    const float a = Ts[(long long)l * n + bc] * cosf(x[l]);
                                    const float bb = Tc[(long long)l * n + bc] * sinf(x[l]);
                                    const float tt = a - bb;
                                    sum[l] += tt;
                                }
                                N += cnt;
                                if (border && bc == c0) half += cnt;
                            }
                            continue;
                        }
                    }
                    const int rr = r0 + lane;
                    if (rr < runs) {
                        S.len[lane] = rlen[rr];
                        S.flag[lane] = rflag[rr];
                        S.cell[lane] = rcell[rr];
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            const int gl = gr[(long long)l * n + rr];
                            const float xq = Xr[(long long)l * n + rr];
                            S.xr[lane][l] = xq;
                            S.x[l][lane] = xq;
                            S.g[l][lane] = gl;
                            S.lo[l][lane] = (float)gl * g.cs;
                            S.hi[l][lane] = (float)(gl + 1) * g.cs;
                        }
                    }
                    wave_sync();
                    const int m = min(kBlkRuns, runs - b * kBlkRuns);
                    // the staged runs this lane must look at: within the
                    // filter radius, or clamped and within the wider one
                    unsigned long long acc = 0;
                    if (active) {
                        for (int q = 0; q < m; q++) {
                            float s = 0.f;
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                const float df = x[l] - S.xr[q][l];
                                s += df * df;
                            }
                            const bool take = s <= thr_filter ||
                                              ((S.flag[q] & 2) && s <= thr_clamp);
                            acc |= take ? (1ull << q) : 0ull;
                        }
                    }
                    // each lane walks its own accepted runs, lowest first (the
                    // block's order, hence the oracle's)
                    while (__any(acc != 0)) {
                        if (acc) {
                            const int q = __ffsll((long long)acc) - 1;
                            acc &= acc - 1;
                            const int c = S.cell[q];
                            if (c != last_c) {
                                // the box test of the source's cell, EGG_SynC.cu:612-634,
                                // for a cell of the reference's neighbour list
                                bool win = true;
#pragma unroll
                                for (int l = 0; l < D; l++) {
                                    if (l < od) {
                                        const int gq = S.g[l][q];
                                        if (gq - g0[l] > R || g0[l] - gq > R) win = false;
                                    }
                                }
                                float sm = 0.f, lg = 0.f;
#pragma unroll
                                for (int l = 0; l < D; l++) {
                                    float lo = S.lo[l][q] - x[l];
                                    float hi = S.hi[l][q] - x[l];
                                    lo *= lo;
                                    hi *= hi;
                                    if (lo < hi) {
                                        sm += lo;
                                        lg += hi;
                                    } else {
                                        sm += hi;
                                        lg += lo;
                                    }
                                }
                                last_c = c;
                                last_cls = !win ? 0 : (lg <= thr ? 1 : (sm <= thr ? 2 : 0));
                            }
                            if (last_cls == 1) {
                                // full: totals and count once, at the cell's first run
                                if (S.flag[q] & 1) {
                                    const int cnt = cell_start[c + 1] - cell_start[c];
#pragma unroll
                                    for (int l = 0; l < D; l++) {
                                        const float a = Ts[(long long)l * n + c] * cosf(x[l]);
                                        const float bb = Tc[(long long)l * n + c] * sinf(x[l]);
                                        const float tt = a - bb;
                                        sum[l] += tt;
                                    }
                                    N += cnt;
                                    if (border && c == c0) half += cnt;
                                }
                                if (border && c != c0) {
                                    float s = 0.f;
#pragma unroll
                                    for (int l = 0; l < D; l++) {
                                        const float df = x[l] - S.x[l][q];
                                        s += df * df;
                                    }
                                    if (s <= thr_half) half += S.len[q];
                                }
                            } else if (last_cls == 2) {
                                float s = 0.f;
#pragma unroll
                                for (int l = 0; l < D; l++) {
                                    const float df = x[l] - S.x[l][q];
                                    s += df * df;
                                }
                                if (s <= thr) {
                                    const int lq = S.len[q];
                                    const float w = (float)lq;  // exact: lq < 2^24
#pragma unroll
                                    for (int l = 0; l < D; l++)
                                        sml[l] += w * sin_pair<SMALL>(S.x[l][q] - x[l]);
                                    N += lq;
                                }
                            }
                        }
                    }
                    wave_sync();
                }
            }
        }
        if (active) {
            float xo[D];
#pragma unroll
            for (int l = 0; l < D; l++) {
                float tt = sum[l] + sml[l];
                tt = tt / (float)N;
                xo[l] = x[l] + tt;
            }
            const bool synced = !((!border && N != own) || (border && N != half));
#pragma unroll
            for (int l = 0; l < D; l++) rres[(long long)l * n + r] = xo[l];
            if (!synced) unsynced = true;
            if (rdbg) {
                rdbg[4 * (long long)r + 0] = N;
                rdbg[4 * (long long)r + 1] = half;
                rdbg[4 * (long long)r + 2] = border ? 1 : 0;
                rdbg[4 * (long long)r + 3] = synced ? 1 : 0;
            }
        }
    }
    if (__any(unsynced) && lane == 0) atomicAnd(&st->flag, 0);
}

// Every point takes its run's result and carries its id over to the next
// iteration's order.  The run of sorted position i is rpos[i] + rhead[i] - 1.
template <int D>
__global__ __launch_bounds__(kBlock) void k_scatter_runs(
    const int *__restrict__ rhead, const int *__restrict__ rpos, const float *__restrict__ rres,
    const int *__restrict__ rdbg, const int *__restrict__ perm_s, float *__restrict__ Xout,
    int *__restrict__ perm_out, int n, DevState *st, DebugOut dbg) {
    if (st->done) return;
    // the local-sync flag as the update left it, before the extra check
    if (blockIdx.x == 0 && threadIdx.x == 0) st->flag_local = st->flag;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        perm_out[i] = perm_s[i];
        const int r = rpos[i] + rhead[i] - 1;
#pragma unroll
        for (int l = 0; l < D; l++) Xout[(long long)l * n + i] = rres[(long long)l * n + r];
        if (rdbg) {
            if (dbg.N) dbg.N[i] = rdbg[4 * (long long)r + 0];
            if (dbg.half) dbg.half[i] = rdbg[4 * (long long)r + 1];
            if (dbg.border) dbg.border[i] = rdbg[4 * (long long)r + 2];
            if (dbg.synced) dbg.synced[i] = rdbg[4 * (long long)r + 3];
        }
    }
}

// MBR_in_eps (EGG_SynC.cu:1037-1159) for the pair (p, q), one thread: per
// dimension the coordinate, among q and q's eps/2-neighbours, closest to p;
// true if that point is within eps of p.  Used only when the pair queue of
// the extra check overflows.
template <int D>
__device__ bool mbr_in_eps(const float *xp, const float *xq, const float *__restrict__ Xs,
                           const int *__restrict__ cell_g, const int *__restrict__ ob,
                           const int *__restrict__ oe, const int *__restrict__ run_start,
                           const int *__restrict__ cell_run, const Geo &g) {
    const int n = g.n, od = g.od, R = g.R_half;
    float mbr[D];
    int g0[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
        mbr[l] = xq[l];
        g0[l] = cell_coord(xq[l], g.cs, g.w);
    }
    const int g0od = select_dim<D>(g0, od);
    const int *cg_od = cell_g + (long long)od * n;
    Odo<D> odo;
    odo.init(od, R);
    while (!odo.finished) {
        const int o = odo.outer(g0, g);
        if (o >= 0) {
            int cb = ob[o];
            const int ce = oe[o];
            if (od < D) cb = column_lower(cg_od, cb, ce, g0od - R - 1);
            for (int c = cb; c < ce; c++) {
                int gc[D];
                bool stop;
                const bool in = cell_in_window<D>(cell_g, n, c, g0, od, R + 1, g0od, gc, &stop);
                if (stop) break;
                if (!in) continue;
                float sm, lg;
                box_dist<D>(gc, xq, g.cs, sm, lg);
                if (!(sm <= g.thr_half)) continue;
                for (int rr = cell_run[c]; rr < cell_run[c + 1]; rr++) {
                    float x2[D];
                    if (dist2_to<D>(xq, Xs, n, run_start[rr], x2) <= g.thr_half) {
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            float od_ = mbr[l] - xp[l];
                            od_ *= od_;
                            float nd = x2[l] - xp[l];
                            nd *= nd;
                            if (nd < od_) mbr[l] = x2[l];
                        }
                    }
                }
            }
        }
        odo.next(od, R);
    }
    float dd = 0.f;
#pragma unroll
    for (int l = 0; l < D; l++) {
        float t = mbr[l] - xp[l];
        t *= t;
        dd += t;
    }
    return dd <= g.thr_eps;
}

// kernel_float_grid_extra_check_part1 (EGG_SynC.cu:900-1034), per tile: each
// lane looks for partners q with eps < |p-q| <= eps_extra in cells that are
// not fully within eps of it, reading the cells' points from LDS chunks, and
// queues the pairs for k_extra_mbr.  Pairs are deduplicated by position (the
// MBR test depends on x_p and x_q only): a point equal to its predecessor in
// the cell is skipped, and so is a partner equal to the lane's last queued
// one.  Should the queue overflow, the lane tests the pair itself.  Runs only
// when every point is locally synced; any hit clears the flag.
template <int D>
__global__ __launch_bounds__(kBlock) void k_extra_tile(
    const float *__restrict__ Xs, const int *__restrict__ cell_g, const int *__restrict__ ob,
    const int *__restrict__ oe, const int *__restrict__ run_start,
    const int *__restrict__ cell_run, const int *__restrict__ grp_cell,
    const int *__restrict__ point_cell, const int *__restrict__ tile_off,
    int *__restrict__ pairs, int pair_cap, Geo g, DevState *st) {
    if (st->done || st->flag == 0) return;
    __shared__ float s_x[kBlock / kWave][kWave][SrcPt<D>::P];
    __shared__ int s_pos[kBlock / kWave][kWave];
    const int n = g.n, od = g.od, R = g.R_extra;
    const int G = st->G, T = st->T;
    const int lane = __lane_id();
    float(*src)[SrcPt<D>::P] = s_x[threadIdx.x / kWave];
    int *src_pos = s_pos[threadIdx.x / kWave];
    const int *cg_od = cell_g + (long long)od * n;
    bool hit = false;
    for (int t = wave_index(); t < T; t = wave_count() + next_tile(&st->xtile_next)) {
        if (__hip_atomic_load(&st->flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        const TileLane tl = tile_lane(t, tile_off, G, grp_cell, cell_run, run_start, point_cell);
        const int c0 = tl.gc0, i = tl.i;
        const bool active = tl.active;
        const int ghi = cell_g[(long long)(D - 1) * n + tl.gc1 - 1];
        float xp[D];
        int g0[D];
#pragma unroll
        for (int l = 0; l < D; l++) {
            g0[l] = cell_g[(long long)l * n + c0];
            xp[l] = active ? Xs[(long long)l * n + i] : 0.f;
        }
        float lq[D];
        bool have_lq = false;
        const int g0od = select_dim<D>(g0, od);
        auto consider = [&](const float *xq, int j) {
            bool rep = have_lq;
#pragma unroll
            for (int l = 0; l < D; l++) rep = rep && __float_as_uint(lq[l]) == __float_as_uint(xq[l]);
            if (rep) return;
#pragma unroll
            for (int l = 0; l < D; l++) lq[l] = xq[l];
            have_lq = true;
            const int slot = atomicAdd(&st->pair_count, 1);
            if (slot < pair_cap) {
                pairs[2 * slot] = i;
                pairs[2 * slot + 1] = j;
            } else if (mbr_in_eps<D>(xp, xq, Xs, cell_g, ob, oe, run_start, cell_run, g)) {
                hit = true;
            }
        };
        walk_columns<D>(g0, R, g, ob, oe, [&](int cb, const int ce) {
            if (od < D) cb = column_lower(cg_od, cb, ce, g0od - R - 1);
            for (int c = cb; c < ce; c++) {
                int gc[D];
                bool stop;
                const bool in =
                    cell_in_group_window<D>(cell_g, n, c, g0, od, R + 1, ghi, gc, &stop);
                if (stop) break;
                if (!in) continue;
                float sm, lg;
                box_dist<D>(gc, xp, g.cs, sm, lg);
                const bool need = active && !(lg <= g.thr_eps) && sm <= g.thr_extra;
                if (!__any(need)) continue;
                const int r0 = cell_run[c], r1 = cell_run[c + 1];
                for (int base = r0; base < r1; base += kWave) {
                    const int rr = base + lane;
                    if (rr < r1) {
                        const int j = run_start[rr];
                        src_pos[lane] = j;
#pragma unroll
                        for (int l = 0; l < D; l++) src[lane][l] = Xs[(long long)l * n + j];
                    }
                    wave_sync();
                    const int m = min(kWave, r1 - base);
                    if (need) {
                        for (int k = 0; k < m; k++) {
                            float xq[D], s = 0.f;
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                xq[l] = src[k][l];
                                const float df = xp[l] - xq[l];
                                s += df * df;
                            }
                            if (!(s <= g.thr_eps) && s <= g.thr_extra) consider(xq, src_pos[k]);
                        }
                    }
                    wave_sync();
                }
            }
        });
    }
    if (__any(hit) && lane == 0) atomicAnd(&st->flag, 0);
}

// kernel_float_grid_extra_check_part2 + MBR_in_eps (EGG_SynC.cu:1037-1203),
// one wavefront per queued pair: the lanes share the scan of q's
// eps/2-neighbourhood and keep, per dimension, the smallest squared distance
// to p of a candidate coordinate (the reference keeps the coordinate with the
// smallest one, so its MBR distance is the sum of these minima); a lane-min
// reduction then gives the same value as the sequential scan.
template <int D>
__global__ __launch_bounds__(kBlock) void k_extra_mbr(
    const float *__restrict__ Xs, const int *__restrict__ cell_g, const int *__restrict__ ob,
    const int *__restrict__ oe, const int *__restrict__ run_start,
    const int *__restrict__ cell_run, const int *__restrict__ pairs, int pair_cap, Geo g,
    DevState *st) {
    if (st->done || st->flag == 0) return;
    const int n = g.n, od = g.od, R = g.R_half;
    const int npairs = min(st->pair_count, pair_cap);
    const int lane = __lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    const int *cg_od = cell_g + (long long)od * n;
    for (int k = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; k < npairs;
         k += waves) {
        if (__hip_atomic_load(&st->flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        const int ip = pairs[2 * k], iq = pairs[2 * k + 1];
        float xp[D], xq[D], best[D];
        int g0[D];
#pragma unroll
        for (int l = 0; l < D; l++) {
            xp[l] = Xs[(long long)l * n + ip];
            xq[l] = Xs[(long long)l * n + iq];
            g0[l] = cell_coord(xq[l], g.cs, g.w);
            float t = xq[l] - xp[l];  // the MBR starts at q
            t *= t;
            best[l] = t;
        }
        const int g0od = select_dim<D>(g0, od);
        walk_columns<D>(g0, R, g, ob, oe, [&](int cb, const int ce) {
            if (od < D) cb = column_lower(cg_od, cb, ce, g0od - R - 1);
            for (int c = cb; c < ce; c++) {
                int gc[D];
                bool stop;
                const bool in = cell_in_window<D>(cell_g, n, c, g0, od, R + 1, g0od, gc, &stop);
                if (stop) break;
                if (!in) continue;
                float sm, lg;
                box_dist<D>(gc, xq, g.cs, sm, lg);
                if (!(sm <= g.thr_half)) continue;
                for (int rr = cell_run[c] + lane; rr < cell_run[c + 1]; rr += kWave) {
                    float x2[D];
                    if (dist2_to<D>(xq, Xs, n, run_start[rr], x2) <= g.thr_half) {
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            float nd = x2[l] - xp[l];
                            nd *= nd;
                            best[l] = fminf(best[l], nd);
                        }
                    }
                }
            }
        });
#pragma unroll
        for (int l = 0; l < D; l++)
            for (int o = kWave / 2; o > 0; o >>= 1)
                best[l] = fminf(best[l], __shfl_xor(best[l], o, kWave));
        float dd = 0.f;
#pragma unroll
        for (int l = 0; l < D; l++) dd += best[l];
        if (dd <= g.thr_eps && lane == 0) atomicAnd(&st->flag, 0);
    }
}

// SynC pass over the grid (SynC.cpp:16-56): every point within FP32 distance
// eps (candidates: Chebyshev radius R+1 in all dims, which contains every such
// point) contributes sinf(x_j - x) to the update and expf(-|x_j - x|) to r_p.
// r_p and the block partial sums of r are FP64; the partials are combined in
// a fixed order by k_rc_finish, so r_c is deterministic.
constexpr int kRcBlocks = 1024;

template <int D, bool WRITE>
__global__ __launch_bounds__(kBlock) void k_update_sync(
    const float *__restrict__ Xs, float *__restrict__ Xout, const int *__restrict__ perm_s,
    int *__restrict__ perm_out, const int *__restrict__ point_cell,
    const int *__restrict__ cell_start, const int *__restrict__ cell_g,
    const int *__restrict__ ob, const int *__restrict__ oe, Geo g, DevState *st,
    double *__restrict__ rc_part) {
    if (WRITE && st->done) return;
    const int n = g.n, od = g.od, R = g.R + 1;
    const int *cg_od = cell_g + (long long)od * n;
    double racc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float x[D], sum[D];
        int g0[D];
        const int c0 = point_cell[i];
#pragma unroll
        for (int l = 0; l < D; l++) {
            x[l] = Xs[(long long)l * n + i];
            g0[l] = cell_g[(long long)l * n + c0];
            sum[l] = 0.f;
        }
        const int g0od = select_dim<D>(g0, od);
        int N = 0;
        double r = 0.0;
        Odo<D> odo;
        odo.init(od, R);
        while (!odo.finished) {
            const int o = odo.outer(g0, g);
            if (o >= 0) {
                int cb = ob[o];
                const int ce = oe[o];
                if (od < D) cb = column_lower(cg_od, cb, ce, g0od - R);
                for (int c = cb; c < ce; c++) {
                    int gc[D];
#pragma unroll
                    for (int l = 0; l < D; l++) gc[l] = cell_g[(long long)l * n + c];
                    if (od < D && select_dim<D>(gc, od) > g0od + R) break;
                    bool skip = false;
#pragma unroll
                    for (int l = 0; l < D; l++)
                        if (l > od && (gc[l] - g0[l] > R || g0[l] - gc[l] > R)) skip = true;
                    if (skip) continue;
                    for (int j = cell_start[c]; j < cell_start[c + 1]; j++) {
                        float xq[D];
                        const float s = dist2_to<D>(x, Xs, n, j, xq);
                        if (s <= g.thr_eps) {
                            if (WRITE) {
#pragma unroll
                                for (int l = 0; l < D; l++) sum[l] += sinf(xq[l] - x[l]);
                            }
                            N++;
                            r += (double)expf(-fabsf(sqrtf(s)));
                        }
                    }
                }
            }
            odo.next(od, R);
        }
        if (WRITE) {
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float t = sum[l] / (float)N;
                Xout[(long long)l * n + i] = x[l] + t;
            }
            perm_out[i] = perm_s[i];
        }
        racc += N > 0 ? r / (double)N : 1.0;
    }
    // deterministic block reduction
    __shared__ double s_r[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) racc += __shfl_xor(racc, o, kWave);
    if (__lane_id() == 0) s_r[threadIdx.x / kWave] = racc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) t += s_r[w];
        rc_part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_rc_finish(const double *__restrict__ rc_part,
                                                      int nparts, int n, DevState *st,
                                                      bool guard) {
    if (guard && st->done) return;
    __shared__ double s_r[kBlock];
    double t = 0.0;
    for (int k = threadIdx.x; k < nparts; k += kBlock) t += rc_part[k];
    s_r[threadIdx.x] = t;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_r[threadIdx.x] += s_r[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) st->rc = s_r[0] / (double)n;
}


__global__ void k_set_flag(DevState *st) {
    if (st->done) return;
    reset_pass_state(st);
}

__global__ void k_finish(DevState *st) {
    if (st->done) return;
    finish_iteration(st);
}

static inline int grid_for(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

// Enough workgroups to fill the chip; they drain the tile queue and exit.
constexpr int kQueueBlocks = 2048;

static inline int queue_threads(const Geo &g) {
    return g.qthreads >= kWave && g.qthreads <= kBlock && g.qthreads % kWave == 0 ? g.qthreads
                                                                                 : kBlock;
}

static inline int queue_blocks(const Geo &g) {
    const long long tiles_max = (long long)g.n / kTile + g.n;  // >= n/64 + C
    const int waves = queue_threads(g) / kWave;
    const long long cap = g.qblocks > 0 ? g.qblocks : kQueueBlocks * (kBlock / kWave) / waves;
    return (int)std::max<long long>(1, std::min<long long>(cap, (tiles_max + waves - 1) / waves));
}

// Filter radii of the pair update: pairs within eps, with a relative margin of
// 2^-12 for points that cell_coord's FP32 division puts one ulp outside their
// cell's box; for sources with a clamped coordinate (they lie up to 1 - w cs
// outside the box of cell w-1 in each such dimension), eps + sqrt(d)(1 - w cs).
static void pair_filters(const Geo &g, float &thr_filter, float &thr_clamp) {
    thr_filter = g.thr_eps * (1.0f + 0x1p-12f);
    const double gap = std::max(0.0, 1.0 - (double)g.w * (double)g.cs);
    const double r = (double)g.eps + std::sqrt((double)g.d) * gap;
    thr_clamp = std::max(thr_filter, (float)(r * r * (1.0 + 0x1p-10)));
}

void launch_update_egg(const Geo &g, GridBufs &b, DevState *st, const DebugOut *dbg,
                       hipStream_t s) {
    DebugOut d = dbg ? *dbg : DebugOut{nullptr, nullptr, nullptr, nullptr};
    const int qb = queue_blocks(g), qt = queue_threads(g);
    const long long nblk_max = g.n / kBlkRuns + 1;
    const int waves_per_block = kBlock / kWave;
    const int pb = (int)std::min<long long>(16384, (nblk_max + waves_per_block - 1) / waves_per_block);
    const int sb = (int)std::max<long long>(
        1, std::min<long long>(1024, (nblk_max / kSupBlks + waves_per_block) / waves_per_block));
    const int pq = (int)std::max<long long>(
        1, std::min<long long>(kQueueBlocks, (nblk_max + waves_per_block - 1) / waves_per_block));
    float thr_filter, thr_clamp;
    pair_filters(g, thr_filter, thr_clamp);
    // every filtered pair is closer than 1 in each coordinate
    const bool pair_small = thr_clamp < 1.0f;
#define PAIR_ARGS                                                                               \
    b.Xs, b.cell_start, b.cell_g, b.Ts, b.Tc, b.run_start, b.point_cell, b.cell_run,            \
        g.pair_mode == kPairAll ? nullptr : b.sparse_runs, b.rcell, b.rflag, b.Xr, b.gr,        \
        b.rlen, b.bmin, b.bmax, b.bclamp, b.bcell, b.sbmin, b.sbmax, b.sbclamp, thr_filter,     \
        thr_clamp, b.rres, rdbg, g, st
    int *rdbg = (d.N || d.half || d.border || d.synced) ? b.rdbg : nullptr;
#define UPD_CASE(DD)                                                                          \
    case DD:                                                                                  \
        k_update_tile<DD><<<qb, qt, 0, s>>>(b.Xs, b.cell_start, b.cell_g, b.Ts, b.Tc,     \
                                            b.ob, b.oe, b.run_start, b.cell_run,              \
                                            b.grp_cell, b.point_cell, b.utile_off,            \
                                            b.cell_sub, b.sub_start, b.smin, b.smax,          \
                                            b.sub_run, b.TsS, b.TcS, b.rres, rdbg, b.hacc,    \
                                            b.hdesc,                                          \
                                            b.dtiles, b.rdbg, g, st);                         \
        if (b.hacc) {                                                                         \
            k_update_heavy<DD><<<qb, qt, 0, s>>>(b.Xs, b.cell_g, b.run_start, b.cell_run,     \
                                                 b.grp_cell, b.point_cell, b.utile_off,       \
                                                 b.sub_run, b.hdesc, b.hacc, g, st);          \
            k_update_finish<DD><<<qb, qt, 0, s>>>(b.Xs, b.cell_start, b.cell_g, b.run_start,  \
                                                  b.cell_run, b.grp_cell, b.point_cell,       \
                                                  b.utile_off, b.dtiles, b.rdbg, b.hacc,      \
                                                  b.rres, rdbg, g, st);                       \
        }                                                                                     \
        k_run_boxes<DD><<<pb, kBlock, 0, s>>>(b.Xs, b.run_start, b.point_cell, b.cell_run,     \
                                              b.cell_g, g, st, b.rcell, b.rflag, b.Xr, b.gr,  \
                                              b.rlen, b.bmin, b.bmax, b.bclamp, b.bcell);     \
        k_super_boxes<DD><<<sb, kBlock, 0, s>>>(b.bmin, b.bmax, b.bclamp, g, st, b.sbmin,     \
                                                b.sbmax, b.sbclamp);                          \
        if (pair_small)                                                                       \
            k_update_pairs<DD, true><<<pq, kBlock, 0, s>>>(PAIR_ARGS);                        \
        else                                                                                  \
            k_update_pairs<DD, false><<<pq, kBlock, 0, s>>>(PAIR_ARGS);                       \
        k_scatter_runs<DD><<<grid_for(g.n), kBlock, 0, s>>>(b.rhead, b.rpos, b.rres, rdbg,   \
                                                            b.perm_s, b.X, b.perm, g.n, st,   \
                                                            d);                               \
        break;
    switch (g.d) {
        EGG_EACH_D(UPD_CASE)
    }
#undef UPD_CASE
#undef PAIR_ARGS
}

void launch_extra_check(const Geo &g, GridBufs &b, DevState *st, hipStream_t s) {
    const int qb = queue_blocks(g), qt = queue_threads(g);
#define EX_CASE(DD)                                                                           \
    case DD:                                                                                  \
        k_extra_tile<DD><<<qb, qt, 0, s>>>(b.Xs, b.cell_g, b.ob, b.oe, b.run_start,       \
                                           b.cell_run, b.grp_cell, b.point_cell, b.tile_off,  \
                                           b.pairs, b.pair_cap, g, st);                       \
        k_extra_mbr<DD><<<1024, kBlock, 0, s>>>(b.Xs, b.cell_g, b.ob, b.oe, b.run_start,      \
                                                b.cell_run, b.pairs, b.pair_cap, g, st);      \
        break;
    switch (g.d) {
        EGG_EACH_D(EX_CASE)
    }
#undef EX_CASE
}

int rc_partials(int n) { return kRcBlocks; }

void launch_update_sync(const Geo &g, GridBufs &b, DevState *st, double *rc_part, bool write_pos,
                        hipStream_t s) {
#define SY_CASE(DD)                                                                          \
    case DD:                                                                                 \
        if (write_pos)                                                                       \
            k_update_sync<DD, true><<<kRcBlocks, kBlock, 0, s>>>(                            \
                b.Xs, b.X, b.perm_s, b.perm, b.point_cell, b.cell_start, b.cell_g, b.ob,     \
                b.oe, g, st, rc_part);                                                       \
        else                                                                                 \
            k_update_sync<DD, false><<<kRcBlocks, kBlock, 0, s>>>(                           \
                b.Xs, nullptr, b.perm_s, nullptr, b.point_cell, b.cell_start, b.cell_g,      \
                b.ob, b.oe, g, st, rc_part);                                                 \
        break;
    switch (g.d) {
        EGG_EACH_D(SY_CASE)
    }
#undef SY_CASE
    k_rc_finish<<<1, kBlock, 0, s>>>(rc_part, kRcBlocks, g.n, st, write_pos);
}

void launch_set_flag(DevState *st, hipStream_t s) { k_set_flag<<<1, 1, 0, s>>>(st); }

void launch_finish(DevState *st, hipStream_t s) { k_finish<<<1, 1, 0, s>>>(st); }

}  // namespace egg
