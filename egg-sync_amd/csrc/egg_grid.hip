// Grid build and per-cell summaries (EGG_SynC.cu:248-366), gfx950.
//
// build_the_grid of the reference (two-level counting sort with atomic slot
// order and linear slot searches) becomes: per-point bit-field cell keys ->
// stable radix sort -> gather of positions into cell-sorted SoA order -> head
// flags + exclusive scan -> cell table -> outer-cell ranges.  Cells and the
// points within a cell come out in a deterministic order.
//
// kernel_inner_grid_stats (float atomics) becomes a deterministic wave-level
// Kahan-compensated reduction per cell chunk (<= 256 points), with a second
// pass combining the chunks of large cells (one wavefront per cell).
#include "egg_internal.h"

namespace egg {

__global__ void k_aos_to_soa(const float *__restrict__ aos, float *__restrict__ soa,
                             int *__restrict__ perm, int n, int d) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        for (int l = 0; l < d; l++) soa[(long long)l * n + i] = aos[i * d + l];
        if (perm) perm[i] = (int)i;
    }
}

__global__ void k_soa_to_aos_by_perm(const float *__restrict__ soa, const int *__restrict__ perm,
                                     float *__restrict__ aos, int n, int d) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long p = perm[i];
        for (int l = 0; l < d; l++) aos[p * d + l] = soa[(long long)l * n + i];
    }
}

__global__ void k_fill(int *__restrict__ p, int count, int v, const DevState *st) {
    if (st && st->done) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

static inline int grid_for(long long n, int per_block = kBlock, int cap = 65536) {
    long long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

void launch_aos_to_soa(const float *aos, float *soa, int *perm, int n, int d, hipStream_t s) {
    k_aos_to_soa<<<grid_for(n), kBlock, 0, s>>>(aos, soa, perm, n, d);
}

void launch_soa_to_aos_by_perm(const float *soa, const int *perm, float *aos, int n, int d,
                               hipStream_t s) {
    k_soa_to_aos_by_perm<<<grid_for(n), kBlock, 0, s>>>(soa, perm, aos, n, d);
}

void launch_fill(int *p, int count, int v, const DevState *st, hipStream_t s) {
    if (count > 0) k_fill<<<grid_for(count), kBlock, 0, s>>>(p, count, v, st);
}

// Key bit fields: [outer | g_od | ... | g_{d-1} | sub_0 ... sub_{d-1} | hash],
// least significant last, so numeric key order = lexicographic (outer, g_od,
// ..., g_{d-1}, sub_0, ..., sub_{d-1}, position hash); the split bits and the
// hash are present when sub_bits / hash_bits > 0.
__device__ __forceinline__ void put_bits(uint64_t (&w)[kMaxKW], uint64_t v, int shift,
                                         int width) {
    const int q = shift >> 6, o = shift & 63;
#pragma unroll
    for (int k = 0; k < kMaxKW; k++) {
        if (k == q) w[k] |= v << o;
        if (k == q + 1 && o + width > 64) w[k] |= v >> (64 - o);  // o > 0 here: width <= 32
    }
}

// Sort keys of the grid build, one per element: per run of the previous
// grid when its runs are valid (every point of a run has the run's key, and
// runs are contiguous in X's order, so a stable sort of the runs expands to
// the stable sort of the points), else per point.
template <int D>
__global__ __launch_bounds__(kBlock) void k_keys(const float *__restrict__ X, Geo g,
                                                 const int *__restrict__ run_start,
                                                 uint64_t *__restrict__ keys,
                                                 int *__restrict__ ob, int *__restrict__ oe,
                                                 DevState *st, DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int n = g.n;
    // the outer cells' ranges start empty (k_outer_ranges fills the occupied ones)
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < g.O; o += gridDim.x * blockDim.x) {
        ob[o] = 0;
        oe[o] = 0;
    }
    const bool valid = st->runs_valid != 0;
    const int m = valid ? st->runs : n;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->sort_m = m;
        st->sort_hist = 256 * max(1, (m + kSortTile - 1) / kSortTile);
        // m == n: every run is one point, run r is point r, as in a point sort
        st->lens_m = m < n ? m : 0;
    }
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const int i = valid ? run_start[e] : e;
        int c[D];
        float x[D];
#pragma unroll
        for (int l = 0; l < D; l++) {
            x[l] = X[(long long)l * n + i];
            c[l] = cell_coord(x[l], g.cs, g.w);
        }
        uint64_t kw[kMaxKW] = {};
        int shift = 0;
        if (g.hash_bits) {
            uint32_t h = 0x9E3779B9u;
#pragma unroll
            for (int l = 0; l < D; l++) h = pos_hash_step(h, __float_as_uint(x[l]));
            h = pos_hash_final(h);
            put_bits(kw, (uint64_t)(h & ((1u << g.hash_bits) - 1u)), 0, g.hash_bits);
            shift = g.hash_bits;
        }
        if (g.sub_bits) {
            // split bits, sub_{d-1} least significant
#pragma unroll
            for (int l = D - 1; l >= 0; l--) {
                put_bits(kw, x[l] >= cell_mid(c[l], g.cs) ? 1u : 0u, shift, 1);
                shift += 1;
            }
        }
#pragma unroll
        for (int l = D - 1; l >= 0; l--) {
            if (l >= g.od) {
                put_bits(kw, (uint64_t)c[l], shift, g.bits_w);
                shift += g.bits_w;
            }
        }
        uint64_t outer = 0;
#pragma unroll
        for (int l = 0; l < D; l++)
            if (l < g.od) outer += (uint64_t)c[l] * (uint64_t)g.wpow[l];
        put_bits(kw, outer, shift, g.bits_o);
#pragma unroll
        for (int k = 0; k < kMaxKW; k++)
            if (k < g.kw) keys[(long long)k * n + e] = kw[k];
    }
}

// Lengths of the sorted runs (sorted position q holds run sidx[q]), for the
// scan that places each run's points.
__global__ __launch_bounds__(kBlock) void k_run_lens(const int *__restrict__ sidx,
                                                     const int *__restrict__ run_start,
                                                     const DevState *st, int *__restrict__ lens,
                                                     const DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int m = st->lens_m;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
        const int r = sidx ? sidx[q] : q;
        lens[q] = run_start[r + 1] - run_start[r];
    }
}

// Whether sorted keys i and i-1 (stride n) agree above their low `s` bits.
__device__ __forceinline__ bool same_above(const uint64_t *__restrict__ skeys, int n, int i,
                                           int kw, int s) {
    const int q = s >> 6, o = s & 63;
    bool eq = true;
#pragma unroll
    for (int k = 0; k < kMaxKW; k++) {
        if (k < kw && k >= q) {
            uint64_t a = skeys[(long long)k * n + i], b = skeys[(long long)k * n + i - 1];
            if (k == q) {
                a >>= o;
                b >>= o;
            }
            eq = eq && a == b;
        }
    }
    return eq;
}

// Sorted positions and perm, and three head flags per sorted position: cell
// head, sub-cell head, run head (the position differs bitwise from its
// predecessor's; equal positions always share a cell and a sub-cell).  The
// sort ordered st->sort_m elements: runs of the previous grid (st->lens_m >
// 0; sorted run q holds run sidx[q], its points go to [off[q], off[q] +
// length) in their previous order) or points.
template <int D>
__global__ __launch_bounds__(kBlock) void k_gather(
    const float *__restrict__ X, const int *__restrict__ perm, const uint64_t *__restrict__ skeys,
    const int *__restrict__ sidx, const int *__restrict__ run_start,
    const int *__restrict__ off, Geo g, float *__restrict__ Xs, int *__restrict__ perm_s,
    int *__restrict__ head, int *__restrict__ shead, int *__restrict__ rhead, const DevState *st,
    DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int n = g.n, sb = g.sub_bits + g.hash_bits, hb = g.hash_bits;
    const int m = st->lens_m;  // 0: the sort ordered points
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        // sorted element q holding slot i, the slot's source j in X, and
        // whether i starts q (only then can it start a cell, sub-cell or run)
        int q = i, j, jp;
        bool first = true;
        if (m == 0) {
            j = sidx ? sidx[i] : i;
            jp = i > 0 ? (sidx ? sidx[i - 1] : i - 1) : j;
        } else {
            if (m < n) {
                int lo = 0, hi = m - 1;  // the last q with off[q] <= i
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (off[mid] <= i) lo = mid;
                    else hi = mid - 1;
                }
                q = lo;
            }
            const int r = sidx ? sidx[q] : q;
            const int o = m < n ? off[q] : q;
            first = i == o;
            j = run_start[r] + (i - o);
            jp = j;
            if (first && q > 0) jp = run_start[sidx ? sidx[q - 1] : q - 1];
        }
        bool rh = i == 0;
#pragma unroll
        for (int l = 0; l < D; l++) {
            const float x = X[(long long)l * n + j];
            Xs[(long long)l * n + i] = x;
            if (first) rh |= __float_as_uint(x) != __float_as_uint(X[(long long)l * n + jp]);
        }
        perm_s[i] = perm[j];
        int h = 1, hs = 1;
        if (!first) {
            h = hs = 0;
        } else if (q > 0) {
            // the sub-cell part: the key without its low hash_bits bits;
            // the cell part: without the split bits either
            hs = !same_above(skeys, n, q, g.kw, hb);
            h = !same_above(skeys, n, q, g.kw, sb);
        }
        head[i] = i == 0 ? 1 : h;
        if (shead) shead[i] = i == 0 ? 1 : hs;
        rhead[i] = rh ? 1 : 0;
    }
}

// Cell table per sorted position: cell of each point, first position, grid
// coordinates and outer id of each cell; with split bits (shead != nullptr)
// also the sub-cell table: first sorted position, fine coordinates 2g + split
// bit, first sub-cell of each cell, and the sub-cells' empty boxes.
template <int D>
__global__ __launch_bounds__(kBlock) void k_cells(
    const float *__restrict__ Xs, const int *__restrict__ head, const int *__restrict__ hpos,
    const int *__restrict__ shead, const int *__restrict__ spos, Geo g,
    int *__restrict__ point_cell, int *__restrict__ cell_start, int *__restrict__ cell_g,
    int *__restrict__ cell_outer, int *__restrict__ sub_start, int *__restrict__ sub_g,
    int *__restrict__ cell_sub, unsigned *__restrict__ smin, unsigned *__restrict__ smax,
    DevState *st_loop, DevState *st) {
    if (st_loop && st_loop->done) return;
    const int n = g.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int h = head[i];
        const int c = hpos[i] + h - 1;
        point_cell[i] = c;
        if (h) {
            cell_start[c] = i;
            int outer = 0;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const int gl = cell_coord(Xs[(long long)l * n + i], g.cs, g.w);
                cell_g[(long long)l * n + c] = gl;
                if (l < g.od) outer += gl * g.wpow[l];
            }
            cell_outer[c] = outer;
        }
        if (i == n - 1) {
            cell_start[c + 1] = n;
            st->C = c + 1;
        }
        if (!shead) continue;
        const int sp = spos[i] + shead[i] - 1;
        if (shead[i]) {
            sub_start[sp] = i;
#pragma unroll
            for (int l = 0; l < D; l++) {
                const float x = Xs[(long long)l * n + i];
                const int gl = cell_coord(x, g.cs, g.w);
                sub_g[(long long)l * n + sp] = 2 * gl + (x >= cell_mid(gl, g.cs) ? 1 : 0);
                // empty box, widened by the summary pass (k_summary)
                smin[(long long)l * n + sp] = 0xFFFFFFFFu;
                smax[(long long)l * n + sp] = 0u;
            }
            if (h) cell_sub[c] = sp;
        }
        if (i == n - 1) {
            sub_start[sp + 1] = n;
            cell_sub[c + 1] = sp + 1;
            st->S = sp + 1;
        }
    }
}

// Whether cell c heads a group of the update tiles.  Cells are sorted by
// (outer id, g_od, ..., g_{d-1}), so cells that share the outer cell and every
// inner coordinate but the last are adjacent; such a cell starts a new group
// when its last coordinate leaves the previous cell's block of gw cells.  With
// no inner dimension (od == d) or gw <= 1 every cell is its own group.
__device__ __forceinline__ bool group_head(const int *__restrict__ cell_g,
                                           const int *__restrict__ cell_outer, const Geo &g,
                                           int c) {
    const int n = g.n, last = g.d - 1;
    const bool grouped = g.gw > 1 && g.od < g.d;
    bool h = c == 0 || !grouped;
    if (!h) {
        h = cell_outer[c] != cell_outer[c - 1];
        for (int l = g.od; l < last && !h; l++)
            h = cell_g[(long long)l * n + c] != cell_g[(long long)l * n + c - 1];
        if (!h)
            h = cell_g[(long long)last * n + c] / g.gw != cell_g[(long long)last * n + c - 1] / g.gw;
    }
    return h;
}

// Per cell: the cell ranges of the outer cells (ob / oe), the cell's
// summary work items (nchunk, ceil(points / kSumChunk)) and whether it heads
// a group of the update tiles (gflag, group_head).
__global__ __launch_bounds__(kBlock) void k_outer_ranges(const int *__restrict__ cell_outer,
                                                         const int *__restrict__ cell_start,
                                                         const int *__restrict__ cell_g, Geo g,
                                                         const DevState *st,
                                                         int *__restrict__ ob,
                                                         int *__restrict__ oe,
                                                         int *__restrict__ nchunk,
                                                         int *__restrict__ gflag,
                                                         DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int C = st->C;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        const int o = cell_outer[c];
        if (c == 0 || cell_outer[c - 1] != o) ob[o] = c;
        if (c == C - 1 || cell_outer[c + 1] != o) oe[o] = c + 1;
        const int cnt = cell_start[c + 1] - cell_start[c];
        nchunk[c] = (cnt + kSumChunk - 1) / kSumChunk;
        gflag[c] = group_head(cell_g, cell_outer, g, c) ? 1 : 0;
    }
}

void build_grid(const Geo &g, GridBufs &b, DevState *st_loop, DevState *st, hipStream_t s,
                bool small) {
    const int n = g.n;
    const int gb = grid_for(n);
#define KEYS_CASE(DD)                                                                        \
    case DD:                                                                                 \
        k_keys<DD><<<gb, kBlock, 0, s>>>(b.X, g, b.run_start, b.keys_a, b.ob, b.oe, st, st_loop); \
        break;
    switch (g.d) {
        EGG_EACH_D(KEYS_CASE)
    }
#undef KEYS_CASE
    const int where =
        small ? small_sort(b.keys_a, b.keys_b, b.vals_b, n, g.kw, &st->sort_m, st_loop, s)
              : radix_sort(b.keys_a, b.keys_b, b.vals_a, b.vals_b, n, g.kw, g.sort_bits, b.hist,
                           b.scan_tmp, st_loop, s, true, &st->sort_m, &st->sort_hist);
    const uint64_t *skeys = where == 1 ? b.keys_b : b.keys_a;
    const int *sidx = where == 1 ? b.vals_b : (where == 0 ? b.vals_a : nullptr);
    // where the sorted runs' points go (nchunk / wi_off are free until the
    // summaries)
    k_run_lens<<<gb, kBlock, 0, s>>>(sidx, b.run_start, st, b.nchunk, st_loop);
    scan_exclusive(b.nchunk, b.wi_off, n, &st->lens_m, b.scan_tmp, nullptr, st_loop, s, true);
#define GATHER_CASE(DD)                                                                     \
    case DD:                                                                                \
        k_gather<DD><<<gb, kBlock, 0, s>>>(b.X, b.perm, skeys, sidx, b.run_start, b.wi_off, \
                                           g, b.Xs, b.perm_s, b.head,                       \
                                           g.sub_bits ? b.shead : nullptr, b.rhead, st,     \
                                           st_loop);                                        \
        break;
    switch (g.d) {
        EGG_EACH_D(GATHER_CASE)
    }
#undef GATHER_CASE
    {
        // cell, run and sub-cell heads: one scan pass over the three
        ScanSet ss{};
        ss.in[0] = b.head;
        ss.out[0] = b.hpos;
        ss.in[1] = b.rhead;
        ss.out[1] = b.rpos;
        ss.in[2] = b.shead;
        ss.out[2] = b.spos;
        scan_exclusive_multi(ss, g.sub_bits ? 3 : 2, n, nullptr, b.scan_tmp, st_loop, s, true);
    }
#define CELLS_CASE(DD)                                                                      \
    case DD:                                                                                \
        k_cells<DD><<<gb, kBlock, 0, s>>>(b.Xs, b.head, b.hpos, g.sub_bits ? b.shead : nullptr, \
                                          b.spos, g, b.point_cell, b.cell_start, b.cell_g,      \
                                          b.cell_outer, b.sub_start, b.sub_g, b.cell_sub,       \
                                          b.smin, b.smax, st_loop, st);                         \
        break;
    switch (g.d) {
        EGG_EACH_D(CELLS_CASE)
    }
#undef CELLS_CASE
    k_outer_ranges<<<gb, kBlock, 0, s>>>(b.cell_outer, b.cell_start, b.cell_g, g, st, b.ob, b.oe,
                                         b.nchunk, b.gflag, st_loop);
}

// ------------------------------------------------------------- summaries --

// Run table from the run-head scan: run_start[r], cell_run[c] (first run of
// each cell), run_start[runs] = n, cell_run[C] = runs.
__global__ __launch_bounds__(kBlock) void k_runs(const int *__restrict__ head,
                                                 const int *__restrict__ hpos,
                                                 const int *__restrict__ point_cell,
                                                 const int *__restrict__ cell_start, int n,
                                                 int *__restrict__ run_start,
                                                 int *__restrict__ cell_run, DevState *st,
                                                 DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r = hpos[i] + head[i] - 1;
        if (head[i]) run_start[r] = i;
        const int c = point_cell[i];
        if (cell_start[c] == i) cell_run[c] = r;
        if (i == n - 1) {
            run_start[r + 1] = n;
            cell_run[c + 1] = r + 1;
            st->runs = r + 1;
            st->runs_valid = 1;
        }
    }
}

// First cell of every group; grp_cell[G] = C.
__global__ __launch_bounds__(kBlock) void k_group_cells(const int *__restrict__ gflag,
                                                        const int *__restrict__ gpos,
                                                        const DevState *st,
                                                        int *__restrict__ grp_cell,
                                                        DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int C = st->C;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        if (gflag[c]) grp_cell[gpos[c]] = c;
        if (c == C - 1) grp_cell[gpos[c] + gflag[c]] = C;
    }
}

// Per group: its tiles of 64 runs (every group: the extra check's table), the
// same for the groups k_update_tile updates (0 for sparse groups), and the
// runs of sparse groups (k_update_pairs), scanned together afterwards.
__global__ __launch_bounds__(kBlock) void k_group_counts(const int *__restrict__ cell_run,
                                                         const int *__restrict__ grp_cell,
                                                         const DevState *st,
                                                         int *__restrict__ ntile,
                                                         int *__restrict__ untile,
                                                         int *__restrict__ sparse_runs,
                                                         int mode, int d, DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int G = st->G;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < G; k += gridDim.x * blockDim.x) {
        const int runs = cell_run[grp_cell[k + 1]] - cell_run[grp_cell[k]];
        const int t = (runs + kTile - 1) / kTile;
        const bool sp = sparse_group(runs, st->runs, st->C, mode, d);
        ntile[k] = t;
        untile[k] = sp ? 0 : t;
        sparse_runs[k] = sp ? runs : 0;
    }
}

// Compensated sum of two (sum, err) pairs over the wavefront by butterflies,
// combined in lane order so every lane ends with the same bits.
__device__ __forceinline__ void wave_two_sum(float &s1, float &e1, float &s2, float &e2) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const float os1 = __shfl_xor(s1, o, kWave), oe1 = __shfl_xor(e1, o, kWave);
        const float os2 = __shfl_xor(s2, o, kWave), oe2 = __shfl_xor(e2, o, kWave);
        if (lane & o) {
            float a1 = os1, b1 = oe1, a2 = os2, b2 = oe2;
            two_sum_add(a1, b1, s1, e1);
            two_sum_add(a2, b2, s2, e2);
            s1 = a1; e1 = b1; s2 = a2; e2 = b2;
        } else {
            two_sum_add(s1, e1, os1, oe1);
            two_sum_add(s2, e2, os2, oe2);
        }
    }
}

// One wavefront per work item (<= kSumChunk points of one cell).  Per dim:
// S_sin = sum(sinf(x) - sinf(g*cs)), S_cos likewise (EGG_SynC.cu:362-363),
// Kahan-compensated per lane, combined across lanes by compensated butterflies.
// Single-chunk cells get T_sin = sinf(g*cs)*cnt + S_sin directly (the bracket
// of EGG_SynC.cu:651-654); multi-chunk cells leave (sum, comp) partials.
template <int D>
__global__ __launch_bounds__(kBlock) void k_summary(
    const float *__restrict__ Xs, const int *__restrict__ cell_start,
    const int *__restrict__ cell_g, const int *__restrict__ wi_off,
    const int *__restrict__ nchunk, Geo g, const int *count, const int *nwork,
    float *__restrict__ Ts, float *__restrict__ Tc, float *__restrict__ partial,
    unsigned *__restrict__ mnk, unsigned *__restrict__ mxk, DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int n = g.n, C = *count, W = *nwork;
    const int lane = __lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int wi = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; wi < W; wi += waves) {
        // largest c with wi_off[c] <= wi.  Groups are non-empty, so each has
        // >= 1 work item: c <= wi_off[c] <= c + (W - C), hence the target lies
        // in [wi - (W - C), wi] and the search takes log2(W - C + 1) steps.
        int lo = max(0, wi - (W - C)), hi = min(wi, C - 1);
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wi_off[mid] <= wi) lo = mid;
            else hi = mid - 1;
        }
        const int c = lo;
        const int k = wi - wi_off[c];
        const int c0 = cell_start[c], c1 = cell_start[c + 1];
        const int j0 = c0 + k * kSumChunk;
        const int j1 = min(c1, j0 + kSumChunk);
        const int nch = nchunk[c];
#pragma unroll
        for (int l = 0; l < D; l++) {
            const float a = (float)cell_g[(long long)l * n + c] * g.cs;
            const float sa = sinf(a), ca = cosf(a);
            float ss = 0.f, sc_ = 0.f, cs_ = 0.f, cc = 0.f;  // (sum, comp) for sin and cos
            unsigned kmin = 0xFFFFFFFFu, kmax = 0u;           // bounding box (mnk / mxk)
            for (int j = j0 + lane; j < j1; j += kWave) {
                const float x = Xs[(long long)l * n + j];
                kmin = min(kmin, ord_key(x));
                kmax = max(kmax, ord_key(x));
                const float ts = sinf(x) - sa;
                const float tc = cosf(x) - ca;
                float y = ts - sc_;
                float t = ss + y;
                sc_ = (t - ss) - y;
                ss = t;
                y = tc - cc;
                t = cs_ + y;
                cc = (t - cs_) - y;
                cs_ = t;
            }
            // Kahan keeps sum - comp; switch to (sum, +err) form for the butterflies
            float s1 = ss, e1 = -sc_, s2 = cs_, e2 = -cc;
            wave_two_sum(s1, e1, s2, e2);
            if (mnk) {
#pragma unroll
                for (int o = kWave / 2; o > 0; o >>= 1) {
                    kmin = min(kmin, (unsigned)__shfl_xor((int)kmin, o, kWave));
                    kmax = max(kmax, (unsigned)__shfl_xor((int)kmax, o, kWave));
                }
                if (lane == 0) {
                    atomicMin(&mnk[(long long)l * n + c], kmin);
                    atomicMax(&mxk[(long long)l * n + c], kmax);
                }
            }
            if (lane == 0) {
                if (nch == 1) {
                    const float cnt = (float)(c1 - c0);
                    const float t1 = sa * cnt;
                    const float t2 = ca * cnt;
                    Ts[(long long)l * n + c] = t1 + (s1 + e1);
                    Tc[(long long)l * n + c] = t2 + (s2 + e2);
                } else {
                    float *pp = &partial[((long long)wi * D + l) * 4];
                    pp[0] = s1; pp[1] = e1; pp[2] = s2; pp[3] = e2;
                }
            }
        }
    }
}

// Cells of more than one work item: one wavefront per cell; lane k sums the
// partials k, k + 64, ... in order, then the lanes combine by butterflies.
template <int D>
__global__ __launch_bounds__(kBlock) void k_summary_combine(
    const int *__restrict__ cell_start, const int *__restrict__ cell_g,
    const int *__restrict__ wi_off, const int *__restrict__ nchunk, Geo g, const int *count,
    float *__restrict__ Ts, float *__restrict__ Tc, const float *__restrict__ partial,
    DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int n = g.n, C = *count;
    const int lane = __lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < C; c += waves) {
        const int nch = nchunk[c];
        if (nch <= 1) continue;
        const int w0 = wi_off[c];
        const float cnt = (float)(cell_start[c + 1] - cell_start[c]);
#pragma unroll
        for (int l = 0; l < D; l++) {
            float s1 = 0.f, e1 = 0.f, s2 = 0.f, e2 = 0.f;
            for (int k = lane; k < nch; k += kWave) {
                const float *pp = &partial[((long long)(w0 + k) * D + l) * 4];
                two_sum_add(s1, e1, pp[0], pp[1]);
                two_sum_add(s2, e2, pp[2], pp[3]);
            }
            wave_two_sum(s1, e1, s2, e2);
            if (lane == 0) {
                const float a = (float)cell_g[(long long)l * n + c] * g.cs;
                const float t1 = sinf(a) * cnt;
                const float t2 = cosf(a) * cnt;
                Ts[(long long)l * n + c] = t1 + (s1 + e1);
                Tc[(long long)l * n + c] = t2 + (s2 + e2);
            }
        }
    }
}

// The runs of the groups sparse_group selects, listed for k_update_pairs at
// the offsets of the scan of their counts: one thread per run (its group is
// gpos[c] + gflag[c] - 1 of its cell c).
__global__ __launch_bounds__(kBlock) void k_sparse_list(
    const int *__restrict__ run_start, const int *__restrict__ point_cell,
    const int *__restrict__ gflag, const int *__restrict__ gpos,
    const int *__restrict__ cell_run, const int *__restrict__ grp_cell,
    const int *__restrict__ off, const DevState *st, int *__restrict__ list, int mode, int d,
    DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int runs = st->runs;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < runs; r += gridDim.x * blockDim.x) {
        const int c = point_cell[run_start[r]];
        const int k = gpos[c] + gflag[c] - 1;
        const int r0 = cell_run[grp_cell[k]];
        if (sparse_group(cell_run[grp_cell[k + 1]] - r0, runs, st->C, mode, d))
            list[off[k] + r - r0] = r;
    }
}

// First run of every sub-cell (a sub-cell starts where the position changes,
// so its first sorted position heads a run; rpos holds the run-head scan).
__global__ __launch_bounds__(kBlock) void k_sub_runs(const int *__restrict__ sub_start,
                                                     const int *__restrict__ rpos,
                                                     const DevState *st,
                                                     int *__restrict__ sub_run,
                                                     int *__restrict__ snchunk,
                                                     DevState *st_loop) {
    if (st_loop && st_loop->done) return;
    const int S = st->S;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < S; k += gridDim.x * blockDim.x) {
        sub_run[k] = rpos[sub_start[k]];
        // the sub-cell's summary work items
        snchunk[k] = (sub_start[k + 1] - sub_start[k] + kSumChunk - 1) / kSumChunk;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sub_run[S] = st->runs;
}

// Kahan per-group totals (see k_summary) of the groups [starts[k], starts[k+1]),
// k < *count, with anchors coords * g.cs.
// Summaries work on "groups" of consecutive sorted points -- cells, or
// sub-cells -- given by their start offsets, coordinates and a count in
// device memory (*count groups, *nwork work items).  nchunk (work items per
// group) is filled by earlier kernels (k_outer_ranges for cells, k_sub_runs
// for sub-cells).
static void group_summaries(const Geo &g, GridBufs &b, const int *starts, const int *coords,
                            int *nchunk, int *wi_off, const int *count, int *nwork, float *Ts,
                            float *Tc, unsigned *mnk, unsigned *mxk, DevState *st_loop,
                            hipStream_t s) {
    const int n = g.n;
    scan_exclusive(nchunk, wi_off, n, count, b.scan_tmp, nwork, st_loop, s, true);
    // work items <= groups + n / kSumChunk <= n + n / kSumChunk
    const long long max_wi = (long long)n + n / kSumChunk + 1;
    const int wb = grid_for(max_wi * kWave, kBlock, 16384);
    const int cb = grid_for((long long)n * kWave, kBlock, 16384);  // one wavefront per group
#define SUM_CASE(DD)                                                                          \
    case DD:                                                                                  \
        k_summary<DD><<<wb, kBlock, 0, s>>>(b.Xs, starts, coords, wi_off, nchunk, g, count,   \
                                            nwork, Ts, Tc, b.partial, mnk, mxk, st_loop);     \
        k_summary_combine<DD><<<cb, kBlock, 0, s>>>(starts, coords, wi_off, nchunk, g, count, \
                                                    Ts, Tc, b.partial, st_loop);              \
        break;
    switch (g.d) {
        EGG_EACH_D(SUM_CASE)
    }
#undef SUM_CASE
}

void cell_summaries(const Geo &g, GridBufs &b, DevState *st_loop, DevState *st, hipStream_t s) {
    const int n = g.n;
    const int gb = grid_for(n);
    group_summaries(g, b, b.cell_start, b.cell_g, b.nchunk, b.wi_off, &st->C, &st->W, b.Ts,
                    b.Tc, nullptr, nullptr, st_loop, s);
    // runs of equal positions (heads scanned by build_grid) and the tiles over them
    k_runs<<<gb, kBlock, 0, s>>>(b.rhead, b.rpos, b.point_cell, b.cell_start, n, b.run_start,
                                 b.cell_run, st, st_loop);
    scan_exclusive(b.gflag, b.gpos, n, &st->C, b.scan_tmp, &st->G, st_loop, s, true);
    k_group_cells<<<gb, kBlock, 0, s>>>(b.gflag, b.gpos, st, b.grp_cell, st_loop);
    k_group_counts<<<gb, kBlock, 0, s>>>(b.cell_run, b.grp_cell, st, b.ntile, b.untile,
                                         b.sparse_runs, g.pair_mode, g.d, st_loop);
    {
        ScanSet ss{};
        ss.in[0] = b.ntile;
        ss.out[0] = b.tile_off;
        ss.total[0] = &st->T;
        ss.in[1] = b.untile;
        ss.out[1] = b.utile_off;
        ss.total[1] = &st->TU;
        ss.in[2] = b.sparse_runs;
        ss.out[2] = b.sparse_off;
        ss.total[2] = &st->sparse_count;
        scan_exclusive_multi(ss, 3, n, &st->G, b.scan_tmp, st_loop, s, true);
    }
    // the runs of the pair update's groups (every run, in order, in mode kPairAll)
    if (g.pair_mode != kPairAll)
        k_sparse_list<<<gb, kBlock, 0, s>>>(b.run_start, b.point_cell, b.gflag, b.gpos,
                                            b.cell_run, b.grp_cell, b.sparse_off, st,
                                            b.sparse_runs, g.pair_mode, g.d, st_loop);
    if (g.sub_bits) {
        // sub-cell runs and totals (anchors on the half-side grid)
        k_sub_runs<<<gb, kBlock, 0, s>>>(b.sub_start, b.rpos, st, b.sub_run, b.snchunk, st_loop);
        Geo gf = g;
        gf.cs = g.cs * 0.5f;
        group_summaries(gf, b, b.sub_start, b.sub_g, b.snchunk, b.swi_off, &st->S, &st->WS,
                        b.TsS, b.TcS, b.smin, b.smax, st_loop, s);
    }
}

}  // namespace egg
