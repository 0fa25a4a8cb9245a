// Shared definitions of the EGG-SynC HIP implementation (gfx950).
//
// Layout in HBM (one solve, n points, d dims, C non-empty cells, O outer cells):
//   X[d][n]            float  positions, structure-of-arrays, in the order of
//                             the last grid build (state between iterations)
//   Xs[d][n]           float  positions sorted by cell (this iteration)
//   perm[n], perm_s[n] int32  point id at each position
//   keys[KW][n]        u64    sort key (outer id, g_od, ..., g_{d-1}, split bits,
//                             position hash), KW <= kMaxKW words, low word first
//   cell_start[C+1], cell_g[d][n], cell_outer[n], Ts[d][n], Tc[d][n]
//   ob[O], oe[O]       int32  cell range of each outer cell
// All FP arithmetic that decides membership is done exactly as the
// reference's source reads (no FMA contraction: the library is built with
// -ffp-contract=off), so the integer results equal the CPU oracle's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egg {

constexpr int kMaxD = 20;         // dimensions supported by the device kernels
constexpr int kMaxKW = 4;         // 64-bit words of a sort key (keys up to 256 bits)
constexpr int kBlock = 256;       // threads per block (4 wavefronts)
constexpr int kWave = 64;
constexpr int kScanItems = 8;     // items per thread in scans
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSortItems = 8;     // items per thread in radix passes
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kSumChunk = 256;    // max points per summary work item
constexpr int kTile = 64;         // max runs per update tile (one wavefront)
constexpr int kDefaultGroup = 1;  // default cells per tile group along the last coordinate
constexpr int kAnchorBlock = 8;   // group widths are powers of two dividing this
constexpr int kSparseRuns = 8;    // cell groups with fewer runs are updated one lane per run,
constexpr int kSparseFactor = 4;  // in iterations with fewer than kSparseFactor runs per cell
constexpr int kHeavyRuns = 4096;  // a streamed range of at least this many runs is heavy:
constexpr int kUnitRuns = 1024;   // split into units of this many runs, any wavefront's work,
constexpr int kHeavyCap = 65536;  // with up to this many ranges per pass in the queue,
constexpr int kHeavyMaxD = 5;     // for d up to this (per-run fixed-point sums: (2d+2)*8 B)
constexpr int kBlkRuns = 64;      // runs per source block of the pair update
constexpr int kSupBlks = 64;      // source blocks per superblock

// Expands X(D) for every dimension count the kernels are instantiated for;
// host code dispatches on the runtime d with `switch (d) { EGG_EACH_D(X) }`.
#define EGG_EACH_D(X)                                                                       \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)  \
    X(17) X(18) X(19) X(20)

// Whether the cell groups with `runs` runs go to the pair update
// (k_update_pairs) instead of the tile kernel in this iteration.  Mode
// kPairSparse: groups of fewer than kSparseRuns runs when the grid is sparse
// overall (fewer than kSparseFactor runs per cell), where one tile per group
// would leave a wavefront with a lane or two; kPairAll: every group;
// kPairAuto: every group of a sparse grid from kPairAllMinD dimensions on
// (the tile kernel walks (2R+1)^od outer cells per tile, 2,197 at od = 3,
// while most cells hold a run or two), else as kPairSparse.
constexpr int kPairAuto = 0;
constexpr int kPairSparse = 1;
constexpr int kPairAll = 2;
constexpr int kPairAllMinD = 6;
__host__ __device__ __forceinline__ bool sparse_group(int runs, int all_runs, int cells,
                                                      int mode, int d) {
    if (mode == kPairAll) return true;
    const bool sparse_grid = (long long)all_runs < (long long)kSparseFactor * cells;
    if (mode == kPairAuto && d >= kPairAllMinD) return sparse_grid;
    return runs < kSparseRuns && sparse_grid;
}

// Grid geometry and thresholds, computed on the host (EGG_SynC.cu:1249-1260).
struct Geo {
    int n;
    int d;
    float eps;
    float cs;          // cell side
    float eps_extra;   // EGG_SynC.cu:1430
    int w;             // cells per dimension
    int od;            // outer dimensions
    int O;             // outer cells (w^od <= n)
    int R;             // ceil(eps/cs)
    int R_extra;       // ceil(eps_extra/cs)
    int R_half;        // ceil((eps/2.)/cs)
    float thr_eps;     // largest s with sqrtf(s) <= eps
    float thr_half;    // largest s with sqrtf(s) <= eps/2
    float thr_extra;   // largest s with sqrtf(s) <= eps_extra
    int bits_w;        // bits of one coordinate field in the key
    int bits_o;        // bits of the outer id field
    int key_bits;      // bits of the cell part of the key
    int sub_bits;      // bits of the sub-cell part below it (d, or 0 if keys would exceed 256 bits)
    int hash_bits;     // bits of the position hash below that (<= 16)
    int sort_bits;     // key_bits + sub_bits + hash_bits
    int kw;            // 64-bit key words (1 .. kMaxKW)
    int wpow[kMaxD];   // w^l, l < od
    int qblocks;       // workgroups of the tile-queue kernels (0 = automatic)
    int qthreads;      // threads per workgroup of those kernels (64..256, multiple of 64)
    int gw;            // update-tile groups: cells per block of the last coordinate (<= 1: one cell)
    int pair_mode;     // which groups the pair update takes (kPairAuto / kPairSparse / kPairAll)
};

// Sub-cells: every cell is split in two along each dimension at its centre
// ((g + 0.5) * cs, FP32); the sort key carries the d split bits below the
// cell fields, so the points of a sub-cell are contiguous inside their cell.
// The last cell of a dimension ends at max((w) * cs, 1) (the clamp of
// cell_coord).
__host__ __device__ __forceinline__ float cell_mid(int g, float cs) {
    return ((float)g + 0.5f) * cs;
}

// Order-preserving map of a float (not NaN) to an unsigned key and back:
// a < b  <=>  ord_key(a) < ord_key(b), so min / max of positions can be
// taken with integer atomics.
__device__ __forceinline__ unsigned ord_key(float v) {
    const unsigned u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float ord_val(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Hash of a position's FP32 bit patterns (the lowest field of the sort key):
// bitwise-equal positions get equal hashes, so inside a sub-cell they sort
// next to each other and form one run.
__host__ __device__ __forceinline__ uint32_t pos_hash_step(uint32_t h, uint32_t bits) {
    h ^= bits;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h;
}

__host__ __device__ __forceinline__ uint32_t pos_hash_final(uint32_t h) {
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Device-resident loop state (one per solve).
struct DevState {
    int done;          // 1: converged or capped; later kernels of the loop are no-ops
    int iter;          // iterations completed
    int flag;          // local-sync flag of the current iteration (1 = synced)
    int flag_local;    // flag before the extra check (diagnostics)
    int C;             // non-empty cells of the current grid
    int W;             // summary work items
    int capped;        // 1 if max_iter stopped the loop
    int max_iter;      // 0 = unlimited
    int stop_rule;     // 0: EGG local sync + extra check; 1: r_c >= lam
    int n_clusters;
    float lam;
    int pair_count;    // extra check: (p, q) pairs found this iteration
    int T;             // tiles of every cell group (<= 64 runs of one group each): extra check
    int TU;            // tiles of the groups k_update_tile updates (sparse groups excluded)
    int tile_next;     // work-queue head of the update kernel
    int xtile_next;    // work-queue head of the extra check
    int runs;          // runs of equal positions in the sorted order
    int G;             // cell groups of the update tiles
    int S;             // non-empty sub-cells
    int sparse_count;  // runs of sparse cell groups (k_update_pairs)
    int WS;            // sub-cell summary work items
    double rc;         // order parameter of the last r_c evaluation
    // heavy streamed ranges of the update (k_update_heavy): descriptors
    // created (high 32 bits) and work units allocated (low 32 bits), the
    // work-queue head over those units, tiles whose result waits for them
    unsigned long long heavy_pack;
    int unit_next;
    int ndt;
    int ptile_next;    // work-queue head of the pair update (k_update_pairs)
    // run_start[0 .. runs] describes runs of bitwise-equal positions in the
    // current order of the positions (set by k_runs; an update moves every
    // point of a run alike, so the runs survive it and the next grid build
    // sorts runs instead of points)
    int runs_valid;
    int sort_m;        // elements of the grid build's sort: runs, or n
    int sort_hist;     // 256 x the sort tiles holding them (histogram entries)
    int lens_m;        // runs whose lengths are scanned (0 without valid runs)
};

// Start of a pass: the local-sync flag and the work-queue heads.
__device__ __forceinline__ void reset_pass_state(DevState *st) {
    st->flag = 1;
    st->pair_count = 0;
    st->tile_next = 0;
    st->xtile_next = 0;
    st->heavy_pack = 0;
    st->unit_next = 0;
    st->ndt = 0;
    st->ptile_next = 0;
}

// End of an iteration (EGG_SynC.cu:1322,1425 / SynC.cpp:277): count it and
// apply the stop rule and the iteration cap.
__device__ __forceinline__ void finish_iteration(DevState *st) {
    st->iter += 1;
    if (st->stop_rule == 0) {
        if (st->flag) st->done = 1;
    } else {
        if (st->rc >= (double)st->lam) st->done = 1;
    }
    if (!st->done && st->max_iter > 0 && st->iter >= st->max_iter) {
        st->done = 1;
        st->capped = 1;
    }
}

// ---------------------------------------------------------------- helpers --

// compute_cell_id per coordinate (EGG_SynC.cu:67-74): FP32 division, truncation
// toward zero, w clamped to w-1 (other out-of-range values clamped too).
__device__ __forceinline__ int cell_coord(float x, float cs, int w) {
    float q = x / cs;
    if (!(q > -1.0f)) return 0;
    if (q >= (float)w) return w - 1;
    int t = (int)q;
    return t >= w ? w - 1 : t;
}

// On-border test (EGG_SynC.cu:590-595), in double like the reference.
__device__ __forceinline__ bool on_border(float x, float cs) {
    double xd = (double)x;
    int idm = (int)((xd - 0.00001) / (double)cs);
    int id = (int)(x / cs);
    int idp = (int)((xd + 0.00001) / (double)cs);
    return id != idm || id != idp;
}

// Box test of a point against an inner cell (EGG_SynC.cu:612-634): squared
// per-dimension distances to the two faces, the smaller summed into `small`,
// the larger into `large`.  This is the reference's formula, including its
// treatment of dimensions where x lies inside the cell's extent.
template <int D>
__device__ __forceinline__ void box_dist(const int *g, const float *x, float cs, float &small,
                                         float &large) {
    float s = 0.f, L = 0.f;
#pragma unroll
    for (int l = 0; l < D; l++) {
        float lo = (float)g[l] * cs;
        float hi = (float)(g[l] + 1) * cs;
        lo -= x[l];
        hi -= x[l];
        lo *= lo;
        hi *= hi;
        if (lo < hi) {
            s += lo;
            L += hi;
        } else {
            s += hi;
            L += lo;
        }
    }
    small = s;
    large = L;
}

// sin and cos of a small angle |v| <= 1 from their Taylor series about 0,
// Horner form with fused multiply-adds: sin to the v^11 term, cos to v^12
// (the first dropped terms, v^13/13! and v^14/14!, are below 2^-32 for
// |v| <= 1).  Max error about 1 ulp against FP64 sin / cos on [-1, 1].  Beyond
// |v| = 1 the library sinf / cosf are used.  The update kernel needs these
// only for angles within a few cell sides of a cell corner.
__host__ __device__ __forceinline__ void sincos_small(float v, float &s, float &c) {
    if (!(fabsf(v) <= 1.0f)) {
        s = sinf(v);
        c = cosf(v);
        return;
    }
    const float v2 = v * v;
    float ps = -2.50521083854417e-08f;              // -1/11!
    ps = fmaf(ps, v2, 2.75573192239859e-06f);       //  1/9!
    ps = fmaf(ps, v2, -1.98412698412698e-04f);      // -1/7!
    ps = fmaf(ps, v2, 8.33333333333333e-03f);       //  1/5!
    ps = fmaf(ps, v2, -1.66666666666667e-01f);      // -1/3!
    s = fmaf(v * v2, ps, v);
    float pc = 2.08767569878681e-09f;               //  1/12!
    pc = fmaf(pc, v2, -2.75573192239859e-07f);      // -1/10!
    pc = fmaf(pc, v2, 2.48015873015873e-05f);       //  1/8!
    pc = fmaf(pc, v2, -1.38888888888889e-03f);      // -1/6!
    pc = fmaf(pc, v2, 4.16666666666667e-02f);       //  1/4!
    pc = fmaf(pc, v2, -0.5f);                       // -1/2!
    c = fmaf(v2, pc, 1.0f);
}

// sin v alone, same polynomial as sincos_small (|v| > 1: sinf).
__host__ __device__ __forceinline__ float sin_small(float v) {
    if (!(fabsf(v) <= 1.0f)) return sinf(v);
    const float v2 = v * v;
    float ps = -2.50521083854417e-08f;
    ps = fmaf(ps, v2, 2.75573192239859e-06f);
    ps = fmaf(ps, v2, -1.98412698412698e-04f);
    ps = fmaf(ps, v2, 8.33333333333333e-03f);
    ps = fmaf(ps, v2, -1.66666666666667e-01f);
    return fmaf(v * v2, ps, v);
}

// sin v and cos v - 1 (the latter without the cancellation of cosf(v) - 1),
// same polynomial as sincos_small.
__host__ __device__ __forceinline__ void sincosm1_small(float v, float &s, float &cm1) {
    if (!(fabsf(v) <= 1.0f)) {
        s = sinf(v);
        cm1 = cosf(v) - 1.0f;
        return;
    }
    const float v2 = v * v;
    float ps = -2.50521083854417e-08f;
    ps = fmaf(ps, v2, 2.75573192239859e-06f);
    ps = fmaf(ps, v2, -1.98412698412698e-04f);
    ps = fmaf(ps, v2, 8.33333333333333e-03f);
    ps = fmaf(ps, v2, -1.66666666666667e-01f);
    s = fmaf(v * v2, ps, v);
    float pc = 2.08767569878681e-09f;
    pc = fmaf(pc, v2, -2.75573192239859e-07f);
    pc = fmaf(pc, v2, 2.48015873015873e-05f);
    pc = fmaf(pc, v2, -1.38888888888889e-03f);
    pc = fmaf(pc, v2, 4.16666666666667e-02f);
    pc = fmaf(pc, v2, -0.5f);
    cm1 = v2 * pc;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}

// Kahan-Babuska combination of two compensated sums (s, c).
__device__ __forceinline__ void two_sum_add(float &s, float &c, float s2, float c2) {
    float t = s + s2;
    float bp = t - s;
    float err = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + err;
}

}  // namespace egg
