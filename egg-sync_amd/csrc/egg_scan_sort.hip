// Device-wide exclusive scan and stable LSD radix sort for gfx950.
//
// Scan: 2048-element tiles (256 threads x 8 items), wave scans with lane
// shuffles; arrays of one tile or less are scanned by one workgroup in a
// single launch, longer ones by a single-pass decoupled look-back scan (the
// grid build) or reduce-then-scan in three launches (callers without
// look-back state).  Radix sort: 8-bit digits, per-tile LDS histogram,
// exclusive scan of the digit-major histogram, then a stable scatter in which
// each wavefront ranks its 64 keys by digit with 8 ballots (lanes holding the
// same digit form a 64-bit match mask).  Stability makes the grid build
// deterministic: points that share a cell keep their previous order.
#include <algorithm>

#include "egg_host.h"
#include "egg_internal.h"

namespace egg {

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a block of B threads; also
// returns the total.
template <int B = kBlock>
__device__ __forceinline__ int block_excl_scan(int v, int *s_w, int &total) {
    const int lane = __lane_id(), wid = threadIdx.x / kWave;
    int inc = wave_incl_scan(v);
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < B / kWave; w++) {
        int t = s_w[w];
        if (w < wid) woff += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return woff + inc - v;
}

// Multi-channel scans: up to kMaxScanCh int32 arrays of the same length are
// scanned by the same launches (one pass over all of them), each channel with
// its own block sums (or look-back words) and total.  Reduce-then-scan first.
template <int NC>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(ScanSet ss, int n, const int *n_dev,
                                                        int *__restrict__ bsum, int nb,
                                                        const DevState *st) {
    if (st && st->done) return;
    __shared__ int s_w[NC][kBlock / kWave];
    if (n_dev) n = min(n, *n_dev);
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        int v = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
            if (base + k < n) v += ss.in[c][base + k];
        int total;
        block_excl_scan(v, s_w[c], total);
        if (threadIdx.x == 0) bsum[c * nb + blockIdx.x] = total;
    }
}

// One workgroup of kSingleBlock threads scans whole arrays in chunks of
// kSingleTile, carrying the running total.  It scans the block sums of the
// three-kernel scan and, in its place, arrays of at most kScanTile values (one
// launch instead of three; integer sums, so the result is the same).  Larger
// arrays take the three kernels: one workgroup's thread-contiguous loads and
// stores cost ~0.85 us per thousand values on gfx950, so past a few thousand
// values the single launch is the slower one.  In place is allowed: each
// thread reads its own values before it writes them, and later chunks are
// read after earlier ones are written.
constexpr int kSingleBlock = 1024;
constexpr int kSingleTile = kSingleBlock * kScanItems;

template <int NC>
__global__ __launch_bounds__(kSingleBlock) void k_scan_single(ScanSet ss, int n, const int *n_dev,
                                                              const DevState *st) {
    if (st && st->done) return;
    __shared__ int s_w[kSingleBlock / kWave];
    if (n_dev) n = min(n, *n_dev);
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int *in = ss.in[c];
        int *out = ss.out[c];
        int carry = 0;
        for (int base = 0; base < n; base += kSingleTile) {
            const int idx0 = base + threadIdx.x * kScanItems;
            int vals[kScanItems];
            int v = 0;
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                vals[k] = (idx0 + k < n) ? in[idx0 + k] : 0;
                v += vals[k];
            }
            int tot;
            int ex = block_excl_scan<kSingleBlock>(v, s_w, tot) + carry;
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                if (idx0 + k < n) out[idx0 + k] = ex;
                ex += vals[k];
            }
            carry += tot;
        }
        if (ss.total[c] && threadIdx.x == 0) *ss.total[c] = carry;
    }
}

template <int NC>
__global__ __launch_bounds__(kBlock) void k_scan_down(ScanSet ss, int n, const int *n_dev,
                                                      const int *__restrict__ bsum, int nb,
                                                      const DevState *st) {
    if (st && st->done) return;
    __shared__ int s_w[NC][kBlock / kWave];
    if (n_dev) n = min(n, *n_dev);
    const long long tile0 = (long long)blockIdx.x * kScanTile;
    if (tile0 >= n) return;
    const long long base = tile0 + (long long)threadIdx.x * kScanItems;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        int vals[kScanItems];
        int v = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            vals[k] = (base + k < n) ? ss.in[c][base + k] : 0;
            v += vals[k];
        }
        int tot;
        int ex = block_excl_scan(v, s_w[c], tot) + bsum[c * nb + blockIdx.x];
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            if (base + k < n) ss.out[c][base + k] = ex;
            ex += vals[k];
        }
    }
}

// ------------------------------------------------- single-pass scan --
//
// Decoupled look-back (Merrill and Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", 2016).  Each workgroup takes a tile from a
// ticket counter (so a tile's predecessors belong to workgroups that started
// earlier), scans it locally and publishes its sum; wavefront 0 then walks
// back over its predecessors' published words, 64 at a time, adding sums
// until it meets an inclusive prefix, and publishes its own inclusive prefix.
// A predecessor publishes its sum before it waits on anything, so every wait
// ends; the wait is still bounded (kLbSpins polls, then an error flag and a
// wrong prefix, reported to the host by scan_tmp_error).
//
// State in tmp: [0] ticket, [1] workgroups finished, [2] scan sequence
// number, [3] error flag; from int 4 on, one 64-bit word per channel and
// tile: flag (bits 63-62: 1 sum, 2 inclusive prefix), the sequence number of
// the scan that wrote it (bits 61-32, so that words of earlier scans read as
// unpublished) and the value (bits 31-0).  The last workgroup to finish
// resets the counters and advances the sequence number.
constexpr int kLbSpins = 1 << 20;
constexpr unsigned long long kLbSum = 1ull << 62, kLbIncl = 2ull << 62;

__device__ __forceinline__ unsigned long long lb_word(unsigned long long flag, unsigned seq,
                                                      int v) {
    return flag | ((unsigned long long)(seq & 0x3fffffffu) << 32) | (unsigned)v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Exclusive prefix of tile `tile` (> 0) from its predecessors' words; one
// wavefront.
__device__ int lb_prefix(const unsigned long long *status, int tile, unsigned seq, int *err) {
    const int lane = __lane_id();
    const unsigned tag = seq & 0x3fffffffu;
    int pre = 0, spins = 0;
    for (int j = tile - 1;;) {
        const int idx = j - lane;
        int flag = 2, val = 0;  // before tile 0: an inclusive prefix of 0
        if (idx >= 0) {
            const unsigned long long w =
                __hip_atomic_load(status + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            flag = (unsigned)((w >> 32) & 0x3fffffffu) == tag ? (int)(w >> 62) : 0;
            val = (int)(unsigned)w;
        }
        const unsigned long long incl = __ballot(flag == 2);
        const unsigned long long none = __ballot(flag == 0);
        const int stop = incl ? __ffsll((long long)incl) - 1 : kWave - 1;
        const unsigned long long need = stop == kWave - 1 ? ~0ull : ((2ull << stop) - 1ull);
        if (none & need) {
            if (++spins > kLbSpins) {
                if (lane == 0) atomicOr(err, 1);
                return pre;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        pre += wave_sum(lane <= stop ? val : 0);
        if (incl) return pre;
        j -= kWave;
    }
}

template <int NC>
__global__ __launch_bounds__(kBlock) void k_scan_lookback(ScanSet ss, int n, const int *n_dev,
                                                          int *lb, const DevState *st) {
    if (st && st->done) return;
    __shared__ int s_w[NC][kBlock / kWave];
    __shared__ int s_tile, s_pre[NC];
    unsigned long long *status = (unsigned long long *)(lb + 4);
    const int nb = gridDim.x;
    if (n_dev && *n_dev <= kScanTile) {
        // one tile at run time: workgroup 0 scans it alone (no look-back
        // state is touched), the others return
        if (blockIdx.x != 0) return;
        const int m = min(n, max(0, *n_dev));
        const int base = threadIdx.x * kScanItems;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            int vals[kScanItems], v = 0;
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                vals[k] = base + k < m ? ss.in[c][base + k] : 0;
                v += vals[k];
            }
            int tot;
            int e = block_excl_scan(v, s_w[c], tot);
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                if (base + k < m) ss.out[c][base + k] = e;
                e += vals[k];
            }
            if (threadIdx.x == 0 && ss.total[c]) *ss.total[c] = tot;
        }
        return;
    }
    if (threadIdx.x == 0) s_tile = atomicAdd(&lb[0], 1);
    __syncthreads();
    const int tile = s_tile;
    const unsigned seq =
        __hip_atomic_load((const unsigned *)&lb[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n_dev) n = min(n, *n_dev);
    const long long tile0 = (long long)tile * kScanTile;
    if (n <= 0 && tile == 0 && threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++)
            if (ss.total[c]) *ss.total[c] = 0;
    }
    if (tile0 < n) {
        const long long base = tile0 + (long long)threadIdx.x * kScanItems;
        int vals[NC][kScanItems], ex[NC], tot[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            int v = 0;
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                vals[c][k] = (base + k < n) ? ss.in[c][base + k] : 0;
                v += vals[c][k];
            }
            ex[c] = block_excl_scan(v, s_w[c], tot[c]);
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < NC; c++)
                __hip_atomic_store(status + (long long)c * nb + tile,
                                   lb_word(tile == 0 ? kLbIncl : kLbSum, seq, tot[c]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x < kWave) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int pre = tile == 0 ? 0 : lb_prefix(status + (long long)c * nb, tile, seq,
                                                          &lb[3]);
                if (threadIdx.x == 0) {
                    s_pre[c] = pre;
                    if (tile > 0)
                        __hip_atomic_store(status + (long long)c * nb + tile,
                                           lb_word(kLbIncl, seq, pre + tot[c]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        __syncthreads();
        const bool last = tile0 + kScanTile >= n;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            int e = ex[c] + s_pre[c];
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                if (base + k < n) ss.out[c][base + k] = e;
                e += vals[c][k];
            }
            if (last && threadIdx.x == 0 && ss.total[c]) *ss.total[c] = s_pre[c] + tot[c];
        }
    }
    if (threadIdx.x == 0) {
        const int fin = atomicAdd(&lb[1], 1);
        if (fin == nb - 1) {
            __hip_atomic_store(&lb[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&lb[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store((unsigned *)&lb[2], seq + 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

size_t scan_tmp_ints(int n) {
    const size_t nb = (size_t)((n + kScanTile - 1) / kScanTile) + 1;
    // block sums of the three-launch scan, or the look-back state
    return std::max((size_t)kMaxScanCh * nb, 4 + 2 * (size_t)kMaxScanCh * nb);
}

int scan_tmp_reset(int *tmp, int n, hipStream_t s) {
    return hip_err(hipMemsetAsync(tmp, 0, sizeof(int) * scan_tmp_ints(n), s));
}

int scan_tmp_error(const int *tmp, hipStream_t s, int *err) {
    int rc = hip_err(hipMemcpyAsync(err, tmp + 3, sizeof(int), hipMemcpyDeviceToHost, s));
    if (!rc) rc = hip_err(hipStreamSynchronize(s));
    return rc;
}

template <int NC>
static void scan_launch(const ScanSet &ss, int n, const int *n_dev, int *tmp, const DevState *st,
                        hipStream_t s, bool lookback) {
    if (n <= kScanTile) {
        k_scan_single<NC><<<1, kSingleBlock, 0, s>>>(ss, n, n_dev, st);
        return;
    }
    const int nb = (n + kScanTile - 1) / kScanTile;
    if (lookback) {
        k_scan_lookback<NC><<<nb, kBlock, 0, s>>>(ss, n, n_dev, tmp, st);
        return;
    }
    k_scan_reduce<NC><<<nb, kBlock, 0, s>>>(ss, n, n_dev, tmp, nb, st);
    ScanSet bs{};
    for (int c = 0; c < NC; c++) {
        bs.in[c] = tmp + c * nb;
        bs.out[c] = tmp + c * nb;
        bs.total[c] = ss.total[c];
    }
    k_scan_single<NC><<<1, kSingleBlock, 0, s>>>(bs, nb, nullptr, st);
    k_scan_down<NC><<<nb, kBlock, 0, s>>>(ss, n, n_dev, tmp, nb, st);
}

void scan_exclusive_multi(const ScanSet &ss, int nch, int n, const int *n_dev, int *tmp,
                          const DevState *st, hipStream_t s, bool lookback) {
    switch (nch) {
        case 1: scan_launch<1>(ss, n, n_dev, tmp, st, s, lookback); break;
        case 2: scan_launch<2>(ss, n, n_dev, tmp, st, s, lookback); break;
        case 3: scan_launch<3>(ss, n, n_dev, tmp, st, s, lookback); break;
        default: scan_launch<4>(ss, n, n_dev, tmp, st, s, lookback); break;
    }
}

void scan_exclusive(const int *in, int *out, int n, const int *n_dev, int *tmp, int *total,
                    const DevState *st, hipStream_t s, bool lookback) {
    ScanSet ss{};
    ss.in[0] = in;
    ss.out[0] = out;
    ss.total[0] = total;
    scan_launch<1>(ss, n, n_dev, tmp, st, s, lookback);
}

// ------------------------------------------------------------ radix sort --

__device__ __forceinline__ int digit_of(const uint64_t *keys, int n, int e, int shift) {
    const uint64_t w = keys[(long long)(shift >> 6) * n + e];
    return (int)((w >> (shift & 63)) & 255ull);
}

// Keys are stored with stride n ([kw][n]); the first m = min(n, *n_dev) of
// them (all n without n_dev) are sorted.
__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint64_t *__restrict__ keys, int n,
                                                    const int *n_dev, int shift,
                                                    int *__restrict__ hist, int ntiles,
                                                    const DevState *st) {
    if (st && st->done) return;
    __shared__ int h[256];
    const int m = n_dev ? min(n, *n_dev) : n;
    // tiles holding elements: the histogram's stride (its scan covers 256
    // of them per digit)
    const int nt = n_dev ? max(1, (m + kSortTile - 1) / kSortTile) : ntiles;
    if ((int)blockIdx.x >= nt) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        long long e = base + k * kBlock + threadIdx.x;
        if (e < m) atomicAdd(&h[digit_of(keys, n, (int)e, shift)], 1);
    }
    __syncthreads();
    hist[(long long)threadIdx.x * nt + blockIdx.x] = h[threadIdx.x];
}

template <int KW>
__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint64_t *__restrict__ kin,
                                                       const int *__restrict__ vin,
                                                       uint64_t *__restrict__ kout,
                                                       int *__restrict__ vout, int n,
                                                       const int *n_dev, int shift,
                                                       const int *__restrict__ hist_scan,
                                                       int ntiles, const DevState *st) {
    if (st && st->done) return;
    const int m = n_dev ? min(n, *n_dev) : n;
    const int nt = n_dev ? max(1, (m + kSortTile - 1) / kSortTile) : ntiles;
    if ((int)blockIdx.x >= nt) return;
    __shared__ int s_off[256];
    __shared__ int s_wc[kBlock / kWave][256];
    const int t = threadIdx.x, wid = t / kWave;
    s_off[t] = hist_scan[(long long)t * nt + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) s_wc[w][t] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
    const int word = shift >> 6, sh = shift & 63;
    for (int k = 0; k < kSortItems; k++) {
        const long long e = base + k * kBlock + t;
        const bool valid = e < m;
        uint64_t kv[KW];
        int val = 0, dig = 0;
        if (valid) {
#pragma unroll
            for (int q = 0; q < KW; q++) kv[q] = kin[(long long)q * n + e];
            val = vin ? vin[e] : (int)e;
            dig = (int)((kv[word] >> sh) & 255ull);
        }
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (dig >> b) & 1;
            const unsigned long long bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const int r = __popcll(m & lanemask_lt());
        if (valid && r == 0) s_wc[wid][dig] = __popcll(m);
        __syncthreads();
        long long pos = 0;
        if (valid) {
            pos = s_off[dig] + r;
            for (int w = 0; w < wid; w++) pos += s_wc[w][dig];
        }
        __syncthreads();
        int tot = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; w++) {
            tot += s_wc[w][t];
            s_wc[w][t] = 0;
        }
        s_off[t] += tot;
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int q = 0; q < KW; q++) kout[(long long)q * n + pos] = kv[q];
            vout[pos] = val;
        }
    }
}

// Stable sort of the first m = min(n, *n_dev) keys ([kw][n] words, the most
// significant word last, as the radix sort reads them) by rank: element i goes
// to the number of keys below it plus the number of equal keys before it, the
// position a stable sort gives it.  One workgroup, m^2 comparisons: for the
// few runs of a nearly converged grid, where the radix sort's 12 launches
// cost more than the sort.  Keys in LDS up to kSmallSortLds of them.
constexpr int kSmallSortBlock = 1024;
constexpr int kSmallSortLds = 1024;

__global__ __launch_bounds__(kSmallSortBlock) void k_small_sort(
    const uint64_t *__restrict__ kin, int n, int kw, const int *n_dev,
    uint64_t *__restrict__ kout, int *__restrict__ vout, const DevState *st) {
    if (st && st->done) return;
    __shared__ uint64_t s_k[kMaxKW][kSmallSortLds];
    const int m = min(n, *n_dev);
    const bool lds = m <= kSmallSortLds;
    if (lds) {
        for (int e = threadIdx.x; e < m; e += blockDim.x)
            for (int w = 0; w < kw; w++) s_k[w][e] = kin[(long long)w * n + e];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        uint64_t ki[kMaxKW];
#pragma unroll
        for (int w = 0; w < kMaxKW; w++)
            ki[w] = w < kw ? (lds ? s_k[w][i] : kin[(long long)w * n + i]) : 0ull;
        int rank = 0;
        for (int j = 0; j < m; j++) {
            int cmp = 0;  // key j against key i, most significant word first
            for (int w = kw - 1; w >= 0 && cmp == 0; w--) {
                const uint64_t a = lds ? s_k[w][j] : kin[(long long)w * n + j];
                cmp = a < ki[w] ? -1 : (a > ki[w] ? 1 : 0);
            }
            rank += (cmp < 0 || (cmp == 0 && j < i)) ? 1 : 0;
        }
        for (int w = 0; w < kw; w++) kout[(long long)w * n + rank] = ki[w];
        vout[rank] = i;
    }
}

int small_sort(const uint64_t *keys_a, uint64_t *keys_b, int *vals_b, int n, int kw,
               const int *n_dev, const DevState *st, hipStream_t s) {
    k_small_sort<<<1, kSmallSortBlock, 0, s>>>(keys_a, n, kw, n_dev, keys_b, vals_b, st);
    return 1;
}

size_t sort_tmp_ints(int n) {
    int ntiles = (n + kSortTile - 1) / kSortTile;
    if (ntiles == 0) ntiles = 1;
    return (size_t)256 * ntiles;
}

int radix_sort(uint64_t *keys_a, uint64_t *keys_b, int *vals_a, int *vals_b, int n, int kw,
               int key_bits, int *hist, int *scan_tmp, const DevState *st, hipStream_t s,
               bool lookback, const int *n_dev, const int *hist_n_dev) {
    int ntiles = (n + kSortTile - 1) / kSortTile;
    if (ntiles == 0) ntiles = 1;
    const int passes = (key_bits + 7) / 8;
    uint64_t *kin = keys_a, *kout = keys_b;
    int *vin = nullptr, *vout = vals_b;
    int where = 0;
    for (int p = 0; p < passes; p++) {
        const int shift = p * 8;
        k_rs_hist<<<ntiles, kBlock, 0, s>>>(kin, n, n_dev, shift, hist, ntiles, st);
        scan_exclusive(hist, hist, 256 * ntiles, hist_n_dev, scan_tmp, nullptr, st, s, lookback);
#define RS_CASE(KW)                                                                          \
    case KW:                                                                                 \
        k_rs_scatter<KW><<<ntiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, n_dev, shift,    \
                                                   hist, ntiles, st);                        \
        break;
        switch (kw) {
            RS_CASE(1) RS_CASE(2) RS_CASE(3) RS_CASE(4)
        }
#undef RS_CASE
        where ^= 1;
        uint64_t *tk = kin; kin = kout; kout = tk;
        vin = vout;
        vout = (vout == vals_b) ? vals_a : vals_b;
    }
    if (passes == 0) {
        // a single cell: identity order (keys unchanged in keys_a)
        return -1;
    }
    return where;
}

}  // namespace egg
