// Host-side launchers shared by the translation units of the library.
#pragma once

#include "egg_common.h"

namespace egg {

// Exclusive scan of n int32 values (n_dev, if given, overrides n at run time;
// it must not exceed n).  `tmp` needs scan_tmp_ints(n) ints.  The sum of all
// values is written to *total if total != nullptr.  In-place allowed.
// With `lookback`, arrays of more than one tile are scanned in one launch
// (decoupled look-back); tmp then holds the scans' shared state and must be
// zero before the first such scan (scan_tmp_reset) -- each scan leaves it
// ready for the next on the same stream.
size_t scan_tmp_ints(int n);
void scan_exclusive(const int *in, int *out, int n, const int *n_dev, int *tmp, int *total,
                    const DevState *st, hipStream_t s, bool lookback = false);
// Zeroes the look-back state in tmp (scan_tmp_ints(n) ints); not for graph capture.
int scan_tmp_reset(int *tmp, int n, hipStream_t s);
// Nonzero if a look-back scan using tmp gave up waiting (an internal error).
int scan_tmp_error(const int *tmp, hipStream_t s, int *err);

// Exclusive scans of nch (<= kMaxScanCh) int32 arrays of the same length in
// one pass: channel c scans in[c] into out[c] (in place allowed) and writes
// its sum to *total[c] if that is not null.
constexpr int kMaxScanCh = 4;
struct ScanSet {
    const int *in[kMaxScanCh];
    int *out[kMaxScanCh];
    int *total[kMaxScanCh];
};
void scan_exclusive_multi(const ScanSet &ss, int nch, int n, const int *n_dev, int *tmp,
                          const DevState *st, hipStream_t s, bool lookback = false);

// The same stable order from one workgroup comparing every pair of the first
// *n_dev keys (for a few hundred keys, where the radix passes' launches
// dominate); the result is in the *_b buffers (returns 1).
constexpr int kSmallSortRuns = 1024;
int small_sort(const uint64_t *keys_a, uint64_t *keys_b, int *vals_b, int n, int kw,
               const int *n_dev, const DevState *st, hipStream_t s);

// Stable LSD radix sort of (key, position) pairs by key bits [0, key_bits),
// 8 bits per pass.  keys_a/keys_b: [kw][n] u64; vals_a/vals_b: [n] int32.
// On entry keys are in keys_a and vals are implicit (0..m-1); m = *n_dev
// (<= n) if n_dev is given, else n.  With n_dev, *hist_n_dev must hold
// 256 * max(1, ceil(m / kSortTile)) (the histogram entries of the tiles that
// hold elements).  Returns 0 if the sorted result is in the *_a buffers, 1 if
// in *_b, -1 if no pass was needed (identity order).
size_t sort_tmp_ints(int n);
int radix_sort(uint64_t *keys_a, uint64_t *keys_b, int *vals_a, int *vals_b, int n, int kw,
               int key_bits, int *hist, int *scan_tmp, const DevState *st, hipStream_t s,
               bool lookback = false, const int *n_dev = nullptr,
               const int *hist_n_dev = nullptr);

// A heavy streamed range of an update tile: tile t, candidate cell c, the
// range's runs (runs [r0, r0 + total) if smask is 0, else the runs of
// sub-cells s0 + k for the bits k of smask), and its work units
// [ubase, ubase + nunits) of kUnitRuns runs.
struct HeavyDesc {
    int t, c, r0, s0;
    unsigned smask;
    int total, ubase, nunits;
    int half;  // 1: an eps/2 count over a full cell (border lanes), 0: the update's sums
};

// Grid build, cell summaries (egg_grid.hip).
struct GridBufs {
    float *X;            // [d][n] current positions (state)
    int *perm;           // [n]
    float *Xs;           // [d][n] sorted positions
    int *perm_s;         // [n]
    uint64_t *keys_a, *keys_b;
    int *vals_a, *vals_b;
    int *hist;           // [256][ntiles]
    int *scan_tmp;
    int *head, *hpos;    // [n] cell-head flags of the sorted positions and their scan
    int *point_cell;     // [n]
    int *cell_start;     // [n+1]
    int *cell_g;         // [d][n]
    int *cell_outer;     // [n]
    int *ob, *oe;        // [O]
    int *nchunk, *wi_off;// [n]
    // runs: maximal stretches of bitwise-equal consecutive positions (sorted
    // order); rhead / rpos are their head flags and the flags' scan
    int *rhead, *rpos;   // [n]
    int *run_start;      // [n+1] first sorted position of each run
    int *cell_run;       // [n+1] first run of each cell
    // sub-cells (see cell_mid): shead / spos head flags and their scan,
    // first sorted position, first sub-cell of each cell, fine coordinates
    // 2g + split bit, first run, totals as Ts / Tc, summary work items
    int *shead, *spos;   // [n]
    int *sub_start;      // [n+1]
    int *cell_sub;       // [n+1]
    int *sub_g;          // [d][n]
    int *sub_run;        // [n+1]
    float *TsS, *TcS;    // [d][n]
    unsigned *smin, *smax;  // [d][n] bounding box of each sub-cell's points (ord_key)
    int *snchunk, *swi_off; // [n]
    // update tiles: cells are grouped (consecutive cells of one outer cell,
    // equal inner coordinates but the last, last one in the same block of gw
    // cells); a tile holds <= 64 runs of one group
    int *gflag, *gpos;   // [n] group-head flag of each cell and its exclusive scan
    int *grp_cell;       // [n+1] first cell of each group, grp_cell[G] = C
    int *ntile, *tile_off;// [n]  tiles per group, exclusive scan (every group: extra check)
    int *untile, *utile_off;// [n] the same for the groups of k_update_tile (not sparse)
    float *rres;         // [d][n] new position of each run
    int *sparse_runs;    // [n]   runs of sparse cell groups
    // pair update (k_update_pairs): per run its cell, flags (bit 0: first
    // run of its cell, bit 1: a coordinate clamped into cell w-1), position,
    // cell coordinates and length; bounding boxes of blocks of kBlkRuns
    // consecutive runs and of superblocks of kSupBlks blocks, whether they
    // hold a clamped run, and the cell of a block whose runs share one (-1
    // otherwise)
    int *rcell, *rflag;  // [n]
    float *Xr;           // [d][n]
    int *gr;             // [d][n]
    int *rlen;           // [n]
    float *bmin, *bmax;  // [d][n / kBlkRuns + 1]
    int *bclamp, *bcell; // [n / kBlkRuns + 1]
    float *sbmin, *sbmax;// [d][n / (kBlkRuns * kSupBlks) + 1]
    int *sbclamp;        // [n / (kBlkRuns * kSupBlks) + 1]
    int *sparse_off;     // [n]   their offsets per group (exclusive scan)
    int *rdbg;           // [4n]  N / half / border / synced of each run (diagnostics;
                         //       also the saved N / half / border of deferred runs)
    // heavy streamed ranges (k_update_heavy): their sums, per run, in
    // fixed point (2^-30): sum w sin v [d], sum w (cos v - 1) [d], count,
    // and the eps/2 count; nullptr for d > kHeavyMaxD (no deferral)
    long long *hacc;     // [2d+2][n]
    HeavyDesc *hdesc;    // [kHeavyCap]
    int *dtiles;         // [n]   tiles whose runs finish in k_update_finish
    float *Ts, *Tc;      // [d][n]
    float *partial;      // [n + n/kSumChunk + 1][4*d]
    int *pairs;          // [2 * pair_cap] extra-check (p, q) sorted positions
    int pair_cap;
};

void launch_aos_to_soa(const float *aos, float *soa, int *perm, int n, int d, hipStream_t s);
void launch_soa_to_aos_by_perm(const float *soa, const int *perm, float *aos, int n, int d,
                               hipStream_t s);
void launch_fill(int *p, int count, int v, const DevState *st, hipStream_t s);
// keys + sort + gather + cells + outer ranges.  Writes st->C.
// small: the previous grid holds at most kSmallSortRuns runs (small_sort).
void build_grid(const Geo &g, GridBufs &b, DevState *st_loop, DevState *st, hipStream_t s,
                bool small = false);
void cell_summaries(const Geo &g, GridBufs &b, DevState *st_loop, DevState *st, hipStream_t s);

// Update / extra check / r_c (egg_update.hip).
struct DebugOut {
    int *N, *half, *border, *synced;  // [n] per sorted position, nullable
};
void launch_update_egg(const Geo &g, GridBufs &b, DevState *st, const DebugOut *dbg,
                       hipStream_t s);
void launch_extra_check(const Geo &g, GridBufs &b, DevState *st, hipStream_t s);
// SynC pass: positions (write_pos) and r_c in FP64 -> st->rc.
void launch_update_sync(const Geo &g, GridBufs &b, DevState *st, double *rc_part, bool write_pos,
                        hipStream_t s);
int rc_partials(int n);
void launch_set_flag(DevState *st, hipStream_t s);
void launch_finish(DevState *st, hipStream_t s);

// Labels (egg_labels.hip).
void launch_labels_egg(const Geo &g, GridBufs &b, int *labels, int *multi, int *lab_off,
                       DevState *st, hipStream_t s);
void launch_labels_sync(const Geo &g, GridBufs &b, int *labels, int *invperm, DevState *st,
                        hipStream_t s);
void launch_validate(const float *aos, long long count, int *err, hipStream_t s);

}  // namespace egg
