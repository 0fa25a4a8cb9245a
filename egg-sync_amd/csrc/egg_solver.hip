// Host driver and C ABI of the EGG-SynC solver (EGG_SynC.cu:1229-1601).
//
// Differences in structure from the reference driver, same algorithm:
//   - one cached device workspace per device instead of ~25 cudaMallocs per
//     call and per-iteration mallocs in scans and the extra check;
//   - the loop state lives on the device (DevState): every kernel of an
//     iteration returns at once when `done` is set, so the host enqueues
//     `iters_per_sync` iterations and reads the state back once per batch
//     instead of 2-4 blocking copies per iteration (EGG_SynC.cu:1368,1416,1449,1483);
//   - errors come back as codes, never exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/egg_sync.h"
#include "egg_host.h"
#include "egg_internal.h"

namespace egg {
namespace {

std::mutex g_mutex;

constexpr int kTraceCap = 4096;

struct Workspace {
    void *base = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    DevState *h_state = nullptr;  // pinned readback
    hipGraphExec_t gexec = nullptr;  // captured loop body, reused while gkey matches
    std::vector<char> gkey;
};

Workspace g_ws[64];

void drop_graph(hipGraphExec_t &ge, std::vector<char> &key) {
    if (ge) (void)hipGraphExecDestroy(ge);
    ge = nullptr;
    key.clear();
}

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

int bitlen(long long v) {
    int b = 0;
    while (v > 0) {
        b++;
        v >>= 1;
    }
    return b < 1 ? 1 : b;
}

// Geometry exactly as EGG_SynC.cu:1249-1260, :563, :779, :1045, :1430.
int compute_geo(int64_t n, int d, float eps, Geo &g) {
    if (n < 1 || n > (int64_t)(INT_MAX / 2)) return EGG_SYNC_E_ARG;
    if (d < 1) return EGG_SYNC_E_ARG;
    if (d > kMaxD) return EGG_SYNC_E_DIM;
    if (!(eps > 0.0f) || !std::isfinite(eps)) return EGG_SYNC_E_ARG;
    std::memset(&g, 0, sizeof g);
    g.n = (int)n;
    g.d = d;
    g.eps = eps;
    g.cs = (float)std::sqrt(std::pow((double)eps / 2., 2.) / d);
    const float inv = 1 / g.cs;
    int w = inv >= 2147483647.0f ? 2147483647 : (int)inv;
    if (w < 1) w = 1;
    g.w = w;
    int od;
    if (w == 1 || n <= 1) {
        od = 1;
    } else {
        od = (int)(std::log((double)n) / std::log((double)w));
        od = std::max(1, std::min(d, od));
    }
    g.od = od;
    long long O = 1;
    for (int l = 0; l < od; l++) {
        g.wpow[l] = (int)O;
        O *= w;
        if (O > INT_MAX) return EGG_SYNC_E_KEY;
    }
    g.O = (int)O;
    g.R = (int)std::ceil(eps / g.cs);
    const float e2 = 2 * eps;
    const double ex = (double)e2 - (double)eps * std::sqrt(15. / 16.) + (double)eps / 2. -
                      std::sin((double)eps / 2.);
    g.eps_extra = (float)ex;
    g.R_extra = (int)std::ceil(g.eps_extra / g.cs);
    g.R_half = (int)std::ceil(((double)eps / 2.) / (double)g.cs);
    g.thr_eps = sqrt_threshold(eps);
    g.thr_half = sqrt_threshold(eps / 2);
    g.thr_extra = sqrt_threshold(g.eps_extra);
    g.bits_w = bitlen((long long)w - 1);
    g.bits_o = bitlen(O - 1);
    g.key_bits = g.bits_o + (d - od) * g.bits_w;
    if (g.key_bits > 64 * kMaxKW) return EGG_SYNC_E_KEY;
    g.sub_bits = g.key_bits + d <= 64 * kMaxKW ? d : 0;
    {
        // position hash: up to 16 bits, in the room left in the key's last word
        const int used = g.key_bits + g.sub_bits;
        g.hash_bits = std::min(16, (used + 63) / 64 * 64 - used);
    }
    g.sort_bits = g.key_bits + g.sub_bits + g.hash_bits;
    g.kw = std::max(1, (g.sort_bits + 63) / 64);
    return EGG_SYNC_OK;
}

struct Carve {
    char *p;
    size_t off = 0;
    explicit Carve(void *base) : p((char *)base) {}
    template <typename T>
    T *take(size_t count) {
        T *r = p ? (T *)(p + off) : nullptr;
        off += align_up(sizeof(T) * std::max<size_t>(count, 1));
        return r;
    }
};

struct Layout {
    GridBufs b{};
    DevState *st = nullptr;
    int *labels = nullptr;    // [n]
    float *io = nullptr;      // [n*d] input / output staging
    double *rc_part = nullptr;
    double *rc_trace = nullptr;
    int *cells_trace = nullptr;
    int *it_before = nullptr;
    int *err = nullptr;
    size_t bytes = 0;
};

Layout carve(void *base, const Geo &g) {
    Layout L;
    Carve c(base);
    const size_t n = (size_t)g.n, d = (size_t)g.d;
    GridBufs &b = L.b;
    b.X = c.take<float>(n * d);
    b.Xs = c.take<float>(n * d);
    b.perm = c.take<int>(n);
    b.perm_s = c.take<int>(n);
    b.keys_a = c.take<uint64_t>(n * g.kw);
    b.keys_b = c.take<uint64_t>(n * g.kw);
    b.vals_a = c.take<int>(n);
    b.vals_b = c.take<int>(n);
    const size_t hist = sort_tmp_ints(g.n);
    b.hist = c.take<int>(hist);
    b.scan_tmp = c.take<int>(scan_tmp_ints((int)std::max(n, hist)) + 16);
    b.head = c.take<int>(n);
    b.hpos = c.take<int>(n);
    b.point_cell = c.take<int>(n);
    b.cell_start = c.take<int>(n + 1);
    b.cell_g = c.take<int>(n * d);
    b.cell_outer = c.take<int>(n);
    b.ob = c.take<int>((size_t)g.O);
    b.oe = c.take<int>((size_t)g.O);
    b.nchunk = c.take<int>(n);
    b.wi_off = c.take<int>(n);
    b.rhead = c.take<int>(n);
    b.rpos = c.take<int>(n);
    b.run_start = c.take<int>(n + 1);
    b.cell_run = c.take<int>(n + 1);
    b.shead = c.take<int>(n);
    b.spos = c.take<int>(n);
    b.sub_start = c.take<int>(n + 1);
    b.cell_sub = c.take<int>(n + 1);
    b.sub_g = c.take<int>(n * d);
    b.sub_run = c.take<int>(n + 1);
    b.TsS = c.take<float>(n * d);
    b.TcS = c.take<float>(n * d);
    b.smin = c.take<unsigned>(n * d);
    b.smax = c.take<unsigned>(n * d);
    b.snchunk = c.take<int>(n);
    b.swi_off = c.take<int>(n);
    b.gflag = c.take<int>(n);
    b.gpos = c.take<int>(n);
    b.grp_cell = c.take<int>(n + 1);
    b.ntile = c.take<int>(n);
    b.tile_off = c.take<int>(n);
    b.untile = c.take<int>(n);
    b.utile_off = c.take<int>(n);
    b.rres = c.take<float>(n * d);
    b.sparse_runs = c.take<int>(n);
    b.rcell = c.take<int>(n);
    b.rflag = c.take<int>(n);
    b.Xr = c.take<float>(n * d);
    b.gr = c.take<int>(n * d);
    b.rlen = c.take<int>(n);
    {
        const size_t nb = n / kBlkRuns + 1, ns = n / ((size_t)kBlkRuns * kSupBlks) + 1;
        b.bmin = c.take<float>(nb * d);
        b.bmax = c.take<float>(nb * d);
        b.bclamp = c.take<int>(nb);
        b.bcell = c.take<int>(nb);
        b.sbmin = c.take<float>(ns * d);
        b.sbmax = c.take<float>(ns * d);
        b.sbclamp = c.take<int>(ns);
    }
    b.sparse_off = c.take<int>(n);
    b.rdbg = c.take<int>(4 * n);
    b.hacc = g.d <= kHeavyMaxD ? c.take<long long>((2 * d + 2) * n) : nullptr;
    b.hdesc = c.take<HeavyDesc>(kHeavyCap);
    b.dtiles = c.take<int>(n);
    b.Ts = c.take<float>(n * d);
    b.Tc = c.take<float>(n * d);
    b.partial = c.take<float>((n + n / kSumChunk + 1) * 4 * d);
    b.pair_cap = (int)std::max<size_t>(n, 65536);
    b.pairs = c.take<int>(2 * (size_t)b.pair_cap);
    L.st = c.take<DevState>(1);
    L.labels = c.take<int>(n);
    L.io = c.take<float>(n * d);
    L.rc_part = c.take<double>((size_t)rc_partials(g.n));
    L.rc_trace = c.take<double>(kTraceCap);
    L.cells_trace = c.take<int>(kTraceCap);
    L.it_before = c.take<int>(1);
    L.err = c.take<int>(1);
    L.bytes = c.off;
    return L;
}

// The look-back scans of the grid build keep their state in scan_tmp; it
// starts at zero for every solve (a workspace is reused across shapes).
static int reset_scan_state(const Geo &g, Layout &L, hipStream_t s) {
    return scan_tmp_reset(L.b.scan_tmp, (int)std::max((size_t)g.n, sort_tmp_ints(g.n)), s);
}

// EGG_SYNC_E_INTERNAL if a look-back scan on this workspace gave up waiting.
static int check_scans(const Layout &L, hipStream_t s) {
    int err = 0;
    const int rc = scan_tmp_error(L.b.scan_tmp, s, &err);
    if (rc) return rc;
    return err ? EGG_SYNC_E_INTERNAL : 0;
}

#define HIP_TRY(x)                       \
    do {                                 \
        int rc_ = hip_err((x));          \
        if (rc_) return rc_;             \
    } while (0)

// Workspace of egg_sync_solve_device_noalloc: a file-scope device array (each
// GPU the library runs on holds its own copy until the process ends).  Host
// code never uses the array's name as an address: scratch_base() takes the
// address on the current device from hipGetSymbolAddress, once per device,
// and every later use goes through that pointer.  Calls share the array, so
// two calls that overlap on one GPU race; the solve initialises everything it
// reads (loop state, error flag, every table is written before it is read).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void *g_scratch_ptr[64];

int scratch_base(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return EGG_SYNC_E_NODEV;
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return e != hipSuccess ? hip_err(e) : EGG_SYNC_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

int acquire(const Geo &g, Layout &L, Workspace *&ws) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return EGG_SYNC_E_NODEV;
    ws = &g_ws[dev];
    Layout probe = carve(nullptr, g);
    if (ws->bytes < probe.bytes) {
        drop_graph(ws->gexec, ws->gkey);
        if (ws->base) HIP_TRY(hipFree(ws->base));
        ws->base = nullptr;
        ws->bytes = 0;
        size_t want = probe.bytes + probe.bytes / 8;
        if (hipMalloc(&ws->base, want) != hipSuccess) {
            (void)hipGetLastError();
            ws->base = nullptr;
            return EGG_SYNC_E_ALLOC;
        }
        ws->bytes = want;
    }
    if (!ws->stream) HIP_TRY(hipStreamCreate(&ws->stream));
    if (!ws->h_state) HIP_TRY(hipHostMalloc((void **)&ws->h_state, sizeof(DevState), 0));
    L = carve(ws->base, g);
    return 0;
}

__global__ void k_aos_to_soa_order(const float *__restrict__ aos, const int *__restrict__ order,
                                   float *__restrict__ soa, int *__restrict__ perm, int n, int d) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int p = order ? order[i] : (int)i;
        for (int l = 0; l < d; l++) soa[(long long)l * n + i] = aos[(long long)p * d + l];
        perm[i] = p;
    }
}

__global__ void k_soa_to_rows(const int *__restrict__ soa, int n, int d, const DevState *st,
                              int *__restrict__ rows) {
    const int C = st->C;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < C;
         c += (long long)gridDim.x * blockDim.x)
        for (int l = 0; l < d; l++) rows[c * d + l] = soa[(long long)l * n + c];
}

// Learned-jump features of every point, point-id order, 4d floats each:
// position, own-cell mean sin and mean cos (the cell totals of the summary
// pass divided by the count), displacement since `prev` divided by eps.
__global__ void k_features(const float *__restrict__ Xs, const int *__restrict__ perm_s,
                           const int *__restrict__ point_cell, const int *__restrict__ cell_start,
                           const float *__restrict__ Ts, const float *__restrict__ Tc,
                           const float *__restrict__ prev, int n, int d, float inv_eps,
                           float *__restrict__ feat) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long p = perm_s[i];
        const int c = point_cell[i];
        const float inv_cnt = 1.f / (float)(cell_start[c + 1] - cell_start[c]);
        float *f = feat + p * 4 * d;
        for (int l = 0; l < d; l++) {
            const float x = Xs[(long long)l * n + i];
            f[l] = x;
            f[d + l] = Ts[(long long)l * n + c] * inv_cnt;
            f[2 * d + l] = Tc[(long long)l * n + c] * inv_cnt;
            f[3 * d + l] = prev ? (x - prev[p * d + l]) * inv_eps : 0.f;
        }
    }
}

// End of a loop iteration: finish_iteration, then record r_c and C of the
// iteration that just ran.
__global__ void k_iter_end(DevState *st, const int *it_before, double *rc_trace,
                           int *cells_trace, int cap) {
    if (!st->done) finish_iteration(st);
    const int it = st->iter;
    if (it == *it_before) return;  // the loop had already stopped
    if (it >= 1 && it <= cap) {
        rc_trace[it - 1] = st->rc;
        cells_trace[it - 1] = st->C;
    }
}

// Start of a loop iteration: the iteration count before it, then the pass
// state reset (nothing once the loop has stopped).
__global__ void k_iter_begin(DevState *st, int *it_before) {
    *it_before = st->iter;
    if (!st->done) reset_pass_state(st);
}

int grid_blocks(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    return (int)std::max(1LL, std::min(b, 65536LL));
}

void init_state(DevState &h, const egg_sync_options &opt) {
    std::memset(&h, 0, sizeof h);
    h.stop_rule = opt.stop_rule;
    h.lam = opt.lam;
    h.max_iter = opt.max_iter;
    h.flag = 1;
    h.rc = std::nan("");
}

void default_options(egg_sync_options &o) {
    std::memset(&o, 0, sizeof o);
    o.stop_rule = EGG_SYNC_STOP_EGG;
    o.label_rule = EGG_SYNC_LABELS_CELLS;
    o.lam = 1.0f - 1e-3f;
    o.max_iter = 0;
    o.compute_rc = 0;
    o.timing = 0;
    o.iters_per_sync = 4;
    o.validate = 1;
    o.use_graph = 0;
    o.queue_blocks = 0;
    o.queue_threads = 0;
    o.streams = 0;
    o.tile_group = 0;
    o.pair_groups = 0;
}

void apply_tuning(const egg_sync_options &opt, Geo &g) {
    g.qblocks = std::max(0, opt.queue_blocks);
    g.qthreads = opt.queue_threads;
    int gw = opt.tile_group > 0 ? opt.tile_group : kDefaultGroup;
    gw = std::min(gw, kAnchorBlock);
    while (gw & (gw - 1)) gw &= gw - 1;  // a power of two: groups never straddle an anchor block
    g.gw = gw;
    g.pair_mode = opt.pair_groups == kPairAll      ? kPairAll
                  : opt.pair_groups == kPairSparse ? kPairSparse
                                                   : kPairAuto;
}

template <typename T>
void key_add(std::vector<char> &k, const T &v) {
    const char *p = (const char *)&v;
    k.insert(k.end(), p, p + sizeof(T));
}


struct Events {
    std::vector<hipEvent_t> ev;
    int per_iter = 4;
    int ensure(int count) {
        while ((int)ev.size() < count) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        return 0;
    }
    ~Events() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return (double)ms;
}

// One pass of the loop body (EGG_SynC.cu:1323-1491), fully on the device.
// small: the grid holds at most kSmallSortRuns runs (the grid build sorts
// them in one launch).
void enqueue_iteration(const Geo &g, Layout &L, const egg_sync_options &opt, hipStream_t s,
                       hipEvent_t *ev, bool small = false) {
    if (ev) (void)hipEventRecord(ev[0], s);
    k_iter_begin<<<1, 1, 0, s>>>(L.st, L.it_before);
    build_grid(g, L.b, L.st, L.st, s, small);
    if (opt.stop_rule == EGG_SYNC_STOP_EGG) cell_summaries(g, L.b, L.st, L.st, s);
    if (ev) (void)hipEventRecord(ev[1], s);
    if (opt.stop_rule == EGG_SYNC_STOP_EGG) {
        launch_update_egg(g, L.b, L.st, nullptr, s);
        if (ev) (void)hipEventRecord(ev[2], s);
        launch_extra_check(g, L.b, L.st, s);
    } else {
        launch_update_sync(g, L.b, L.st, L.rc_part, true, s);
        if (ev) (void)hipEventRecord(ev[2], s);
    }
    k_iter_end<<<1, 1, 0, s>>>(L.st, L.it_before, L.rc_trace, L.cells_trace, kTraceCap);
    if (ev) (void)hipEventRecord(ev[3], s);
}

// Records `K` loop iterations enqueued by `body` on the private stream `cap`
// into an executable graph.  The loop body has no memset or host-copy nodes:
// every kernel reads the device state and returns at once after convergence,
// so replaying the graph past the last iteration changes nothing.
template <typename Body>
int capture_graph(hipStream_t cap, hipGraphExec_t &ge, Body &&body) {
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    body();
    hipGraph_t graph = nullptr;
    const hipError_t e1 = hipStreamEndCapture(cap, &graph);
    const hipError_t e0 = hipGetLastError();
    if (e1 != hipSuccess || e0 != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_err(e1 != hipSuccess ? e1 : (e0 != hipSuccess ? e0 : hipErrorUnknown));
    }
    const hipError_t e2 = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
        ge = nullptr;
        return hip_err(e2);
    }
    return 0;
}

// The solve.  With `noalloc` the workspace is the scratch array (no device
// or pinned allocation, no runtime objects: no graph, no timing events, no
// stream of its own -- the caller's stream or the null stream), and a
// workload whose workspace exceeds it is rejected with EGG_SYNC_E_TOO_BIG.
int solve_device_impl(const float *d_D, int64_t n64, int d, float eps,
                      const egg_sync_options &opt_in, int32_t *d_labels, float *d_final,
                      egg_sync_stats *so, hipStream_t s_user, const float *h_D,
                      int32_t *h_labels, float *h_final, bool noalloc = false) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    egg_sync_options opt = opt_in;
    if (noalloc) {
        opt.use_graph = 0;
        opt.timing = 0;
    }
    if (opt.stop_rule != EGG_SYNC_STOP_EGG && opt.stop_rule != EGG_SYNC_STOP_RC)
        return EGG_SYNC_E_ARG;
    if (opt.label_rule != EGG_SYNC_LABELS_CELLS && opt.label_rule != EGG_SYNC_LABELS_SYNC)
        return EGG_SYNC_E_ARG;
    if (n64 < 0 || d < 1) return EGG_SYNC_E_ARG;
    if (so) {
        so->iterations = 0;
        so->status = 0;
        so->n_clusters = 0;
        so->n_cells = 0;
        so->rc = std::nan("");
        for (int k = 0; k < 6; k++) so->stage_ms[k] = 0.0;
        so->update_kernel_ms = 0.0;
        so->update_launches = 0;
        so->workspace_bytes = 0;
    }
    if (n64 == 0) {
        if (so) so->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        return EGG_SYNC_OK;
    }
    Geo g;
    int rc = compute_geo(n64, d, eps, g);
    if (rc) return rc;
    apply_tuning(opt, g);
    const int n = g.n;
    Layout L;
    Workspace *ws = nullptr;
    DevState h_state_local;
    DevState *h_state = &h_state_local;
    hipStream_t s = s_user;
    if (noalloc) {
        void *base = nullptr;
        rc = scratch_base(&base);
        if (rc) return rc;
        if (carve(nullptr, g).bytes > kScratchBytes) return EGG_SYNC_E_TOO_BIG;
        L = carve(base, g);
    } else {
        rc = acquire(g, L, ws);
        if (rc) return rc;
        h_state = ws->h_state;
        if (!s) s = ws->stream;
    }
    if (h_D) {
        HIP_TRY(hipMemcpyAsync(L.io, h_D, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice, s));
        d_D = L.io;
    }
    if (opt.validate) {
        HIP_TRY(hipMemsetAsync(L.err, 0, sizeof(int), s));
        launch_validate(d_D, (long long)n * d, L.err, s);
        int herr = 0;
        HIP_TRY(hipMemcpyAsync(&herr, L.err, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (herr) return EGG_SYNC_E_RANGE;
    }
    launch_aos_to_soa(d_D, L.b.X, L.b.perm, n, d, s);
    DevState hs;
    init_state(hs, opt);
    HIP_TRY(hipMemcpyAsync(L.st, &hs, sizeof hs, hipMemcpyHostToDevice, s));
    rc = reset_scan_state(g, L, s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    const auto t_alloc = clk::now();

    const int K = std::max(1, opt.iters_per_sync);
    Events E;
    if (opt.timing) {
        rc = E.ensure(4 * K + 4);
        if (rc) return rc;
    }
    double st_build = 0, st_upd = 0, st_extra = 0;
    int done_iters = 0;
    // runs only merge, so once a state read shows at most kSmallSortRuns
    // valid runs, every later grid build may sort them in one launch
    bool small = false;
    auto ensure_graph = [&]() -> int {
        if (!opt.use_graph) return 0;
        std::vector<char> key;
        key_add(key, ws->base);
        key_add(key, g);
        key_add(key, K);
        key_add(key, opt.stop_rule);
        key_add(key, small);
        if (!ws->gexec || key != ws->gkey) {
            drop_graph(ws->gexec, ws->gkey);
            const int r = capture_graph(ws->stream, ws->gexec, [&] {
                for (int k = 0; k < K; k++)
                    enqueue_iteration(g, L, opt, ws->stream, nullptr, small);
            });
            if (r) return r;
            ws->gkey = key;
        }
        return 0;
    };
    rc = ensure_graph();
    if (rc) return rc;
    for (;;) {
        if (opt.use_graph) {
            if (opt.timing) (void)hipEventRecord(E.ev[0], s);
            HIP_TRY(hipGraphLaunch(ws->gexec, s));
            if (opt.timing) (void)hipEventRecord(E.ev[1], s);
        } else {
            for (int k = 0; k < K; k++)
                enqueue_iteration(g, L, opt, s, opt.timing ? &E.ev[4 * k] : nullptr, small);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(h_state, L.st, sizeof(DevState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const int real = h_state->iter - done_iters;
        if (opt.timing && opt.use_graph) {
            const double t = ms_between(E.ev[0], E.ev[1]);
            st_upd += t;
            for (int k = 0; k < real && k < K; k++)
                if (so && so->itr_ms && done_iters + k < so->itr_ms_cap)
                    so->itr_ms[done_iters + k] = (float)(t / std::max(real, 1));
        } else if (opt.timing) {
            for (int k = 0; k < real && k < K; k++) {
                hipEvent_t *ev = &E.ev[4 * k];
                const double b = ms_between(ev[0], ev[1]);
                const double u = ms_between(ev[1], ev[2]);
                const double x = ms_between(ev[2], ev[3]);
                st_build += b;
                st_upd += u;
                st_extra += x;
                if (so && so->itr_ms && done_iters + k < so->itr_ms_cap)
                    so->itr_ms[done_iters + k] = (float)(b + u + x);
            }
        }
        done_iters = h_state->iter;
        if (h_state->done) break;
        const bool want = h_state->runs_valid && h_state->runs <= kSmallSortRuns;
        if (want != small) {
            small = want;
            rc = ensure_graph();
            if (rc) return rc;
        }
    }
    const DevState loop_state = *h_state;

    // final grid on the converged positions and labels (EGG_SynC.cu:1504-1541)
    hipEvent_t *fe = opt.timing ? &E.ev[4 * K] : nullptr;
    if (fe) (void)hipEventRecord(fe[0], s);
    build_grid(g, L.b, nullptr, L.st, s,
               loop_state.runs_valid && loop_state.runs <= kSmallSortRuns);
    int *labels_dev = d_labels ? d_labels : L.labels;
    if (opt.label_rule == EGG_SYNC_LABELS_CELLS)
        launch_labels_egg(g, L.b, labels_dev, L.b.nchunk, L.b.wi_off, L.st, s);
    else
        launch_labels_sync(g, L.b, labels_dev, L.b.head, L.st, s);
    if (opt.compute_rc && opt.stop_rule == EGG_SYNC_STOP_EGG)
        launch_update_sync(g, L.b, L.st, L.rc_part, false, s);
    if (fe) (void)hipEventRecord(fe[1], s);
    float *final_dev = d_final ? d_final : (h_final ? L.io : nullptr);
    if (final_dev) launch_soa_to_aos_by_perm(L.b.Xs, L.b.perm_s, final_dev, n, d, s);
    if (h_labels)
        HIP_TRY(hipMemcpyAsync(h_labels, labels_dev, sizeof(int) * (size_t)n,
                               hipMemcpyDeviceToHost, s));
    if (h_final)
        HIP_TRY(hipMemcpyAsync(h_final, final_dev, sizeof(float) * (size_t)n * d,
                               hipMemcpyDeviceToHost, s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h_state, L.st, sizeof(DevState), hipMemcpyDeviceToHost, s));
    const int ntr = so && so->trace_cap > 0 ? std::min(std::min(so->trace_cap, kTraceCap),
                                                        loop_state.iter) : 0;
    if (ntr > 0 && so->rc_trace)
        HIP_TRY(hipMemcpyAsync(so->rc_trace, L.rc_trace, sizeof(double) * ntr,
                               hipMemcpyDeviceToHost, s));
    if (ntr > 0 && so->cells_trace)
        HIP_TRY(hipMemcpyAsync(so->cells_trace, L.cells_trace, sizeof(int) * ntr,
                               hipMemcpyDeviceToHost, s));
    if (fe) (void)hipEventRecord(fe[2], s);
    HIP_TRY(hipStreamSynchronize(s));
    const DevState fin = *h_state;
    rc = check_scans(L, s);
    if (rc) return rc;
    if (so) {
        so->iterations = loop_state.iter;
        so->status = loop_state.capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK;
        so->n_clusters = fin.n_clusters;
        so->n_cells = fin.C;
        so->rc = (opt.compute_rc || opt.stop_rule == EGG_SYNC_STOP_RC) ? fin.rc : std::nan("");
        so->stage_ms[0] = std::chrono::duration<double, std::milli>(t_alloc - t0).count();
        so->stage_ms[1] = st_build;
        so->stage_ms[2] = st_upd;
        so->stage_ms[3] = st_extra;
        if (fe) {
            so->stage_ms[4] = ms_between(fe[0], fe[1]);
            so->stage_ms[5] = ms_between(fe[1], fe[2]);
        }
        so->update_kernel_ms = st_upd;
        so->update_launches = opt.timing ? loop_state.iter : 0;
        so->workspace_bytes = (int64_t)L.bytes;
        so->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    return loop_state.capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK;
}

// Shared set-up of the single-pass entry points: positions (point-id order)
// visited in `order` into X/perm, loop state reset.
int setup_single(const float *d_D, int64_t n64, int d, float eps, const int32_t *d_order,
                 Geo &g, Layout &L, hipStream_t &s, void *stream) {
    if (!d_D) return EGG_SYNC_E_ARG;
    int rc = compute_geo(n64, d, eps, g);
    if (rc) return rc;
    Workspace *ws = nullptr;
    rc = acquire(g, L, ws);
    if (rc) return rc;
    s = stream ? (hipStream_t)stream : ws->stream;
    k_aos_to_soa_order<<<grid_blocks(g.n), kBlock, 0, s>>>(d_D, d_order, L.b.X, L.b.perm, g.n, d);
    egg_sync_options opt;
    default_options(opt);
    DevState hs;
    init_state(hs, opt);
    HIP_TRY(hipMemcpyAsync(L.st, &hs, sizeof hs, hipMemcpyHostToDevice, s));
    return reset_scan_state(g, L, s);
}

// ------------------------------------------------------------ batched solve --
//
// Independent solves share one device arena (instance k's buffers at its own
// offset, carved exactly as for a single solve) and one array of loop states.
// Instances are dealt to HIP streams, largest first onto the least-loaded
// stream; a stream runs its instances' iterations back to back, so small
// solves, which cannot fill the chip alone, overlap with each other and with
// the large ones.  Every `iters_per_sync` iterations one copy brings all loop
// states to the host, and the loop ends when every instance is done (the
// kernels of a finished instance return at once).  With use_graph the
// iterations of all streams are captured once, forked from and joined to one
// origin stream, and the graph is replayed.

// Kernel nodes of one captured loop iteration (an upper bound) and the
// largest batched graph that is captured.
constexpr int kGraphNodesPerIter = 64;
constexpr long long kMaxBatchGraphNodes = 65536;

struct BatchWs {
    void *base = nullptr;
    size_t bytes = 0;
    DevState *states = nullptr;
    DevState *h_states = nullptr;
    int states_cap = 0;
    int *err = nullptr;
    hipStream_t origin = nullptr;  // used when the caller passes no stream; capture origin
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> gkey;
};

BatchWs g_bws[64];

constexpr int kMaxStreams = 64;

int batch_resources(BatchWs &bw, size_t bytes, int n_inst, int S) {
    if (bw.bytes < bytes) {
        drop_graph(bw.gexec, bw.gkey);
        if (bw.base) HIP_TRY(hipFree(bw.base));
        bw.base = nullptr;
        bw.bytes = 0;
        const size_t want = bytes + bytes / 8;
        if (hipMalloc(&bw.base, want) != hipSuccess) {
            (void)hipGetLastError();
            return EGG_SYNC_E_ALLOC;
        }
        bw.bytes = want;
    }
    if (bw.states_cap < n_inst) {
        drop_graph(bw.gexec, bw.gkey);
        if (bw.states) HIP_TRY(hipFree(bw.states));
        if (bw.h_states) HIP_TRY(hipHostFree(bw.h_states));
        bw.states = nullptr;
        bw.h_states = nullptr;
        bw.states_cap = 0;
        if (hipMalloc((void **)&bw.states, sizeof(DevState) * n_inst) != hipSuccess) {
            (void)hipGetLastError();
            return EGG_SYNC_E_ALLOC;
        }
        HIP_TRY(hipHostMalloc((void **)&bw.h_states, sizeof(DevState) * n_inst, 0));
        bw.states_cap = n_inst;
    }
    if (!bw.err) HIP_TRY(hipMalloc((void **)&bw.err, sizeof(int)));
    if (!bw.origin) HIP_TRY(hipStreamCreate(&bw.origin));
    while ((int)bw.streams.size() < S) {
        hipStream_t t;
        HIP_TRY(hipStreamCreate(&t));
        bw.streams.push_back(t);
    }
    if (!bw.fork) HIP_TRY(hipEventCreateWithFlags(&bw.fork, hipEventDisableTiming));
    while ((int)bw.joins.size() < S) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        bw.joins.push_back(e);
    }
    return 0;
}

void release_batch(BatchWs &bw) {
    drop_graph(bw.gexec, bw.gkey);
    if (bw.base) (void)hipFree(bw.base);
    if (bw.states) (void)hipFree(bw.states);
    if (bw.h_states) (void)hipHostFree(bw.h_states);
    if (bw.err) (void)hipFree(bw.err);
    for (auto t : bw.streams) (void)hipStreamDestroy(t);
    if (bw.origin) (void)hipStreamDestroy(bw.origin);
    if (bw.fork) (void)hipEventDestroy(bw.fork);
    for (auto e : bw.joins) (void)hipEventDestroy(e);
    bw = BatchWs{};
}

int fork_streams(BatchWs &bw, hipStream_t origin, int S) {
    HIP_TRY(hipEventRecord(bw.fork, origin));
    for (int j = 0; j < S; j++) HIP_TRY(hipStreamWaitEvent(bw.streams[j], bw.fork, 0));
    return 0;
}

int join_streams(BatchWs &bw, hipStream_t origin, int S) {
    for (int j = 0; j < S; j++) {
        HIP_TRY(hipEventRecord(bw.joins[j], bw.streams[j]));
        HIP_TRY(hipStreamWaitEvent(origin, bw.joins[j], 0));
    }
    return 0;
}

int solve_batch_impl(const float *d_D, const egg_sync_instance *inst, int n_inst,
                     const egg_sync_options &opt, int32_t *d_labels, float *d_final,
                     egg_sync_instance_result *res, egg_sync_stats *so, hipStream_t s_user) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (n_inst < 0 || (n_inst > 0 && !inst)) return EGG_SYNC_E_ARG;
    if (opt.stop_rule != EGG_SYNC_STOP_EGG && opt.stop_rule != EGG_SYNC_STOP_RC)
        return EGG_SYNC_E_ARG;
    if (opt.label_rule != EGG_SYNC_LABELS_CELLS && opt.label_rule != EGG_SYNC_LABELS_SYNC)
        return EGG_SYNC_E_ARG;
    if (so) {
        so->iterations = 0;
        so->status = 0;
        so->n_clusters = 0;
        so->n_cells = 0;
        so->rc = std::nan("");
        for (int k = 0; k < 6; k++) so->stage_ms[k] = 0.0;
        so->update_kernel_ms = 0.0;
        so->update_launches = 0;
        so->workspace_bytes = 0;
    }
    if (res)
        for (int k = 0; k < n_inst; k++) res[k] = egg_sync_instance_result{0, 0, 0, 0};
    std::vector<Geo> G(n_inst);
    std::vector<size_t> off(n_inst, 0);
    std::vector<int64_t> row_off(n_inst, 0), fl_off(n_inst, 0);
    std::vector<int> active;
    size_t total = 0;
    int64_t rows = 0, fl = 0;
    for (int k = 0; k < n_inst; k++) {
        const egg_sync_instance &I = inst[k];
        if (I.n < 0 || I.data_offset < 0) return EGG_SYNC_E_ARG;
        row_off[k] = rows;
        fl_off[k] = fl;
        if (I.n == 0) continue;
        const int rc = compute_geo(I.n, I.d, I.eps, G[k]);
        if (rc) return rc;
        apply_tuning(opt, G[k]);
        off[k] = total;
        total += carve(nullptr, G[k]).bytes;
        rows += I.n;
        fl += I.n * I.d;
        active.push_back(k);
    }
    if (active.empty()) {
        if (so) so->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        return EGG_SYNC_OK;
    }
    if (!d_D || !d_labels) return EGG_SYNC_E_ARG;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return EGG_SYNC_E_NODEV;
    BatchWs &bw = g_bws[dev];
    int S = opt.streams > 0 ? std::min(opt.streams, kMaxStreams) : 4;
    S = std::max(1, std::min(S, (int)active.size()));
    int rc = batch_resources(bw, total, n_inst, S);
    if (rc) return rc;
    std::vector<Layout> Ls(n_inst);
    for (int k : active) {
        Ls[k] = carve((char *)bw.base + off[k], G[k]);
        Ls[k].st = bw.states + k;
    }
    hipStream_t s = s_user ? s_user : bw.origin;

    if (opt.validate) {
        HIP_TRY(hipMemsetAsync(bw.err, 0, sizeof(int), s));
        for (int k : active)
            launch_validate(d_D + inst[k].data_offset, inst[k].n * inst[k].d, bw.err, s);
        int herr = 0;
        HIP_TRY(hipMemcpyAsync(&herr, bw.err, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (herr) return EGG_SYNC_E_RANGE;
    }
    for (int k = 0; k < n_inst; k++) {
        init_state(bw.h_states[k], opt);
        if (inst[k].n == 0) bw.h_states[k].done = 1;
    }
    HIP_TRY(hipMemcpyAsync(bw.states, bw.h_states, sizeof(DevState) * n_inst,
                           hipMemcpyHostToDevice, s));
    for (int k : active) {
        launch_aos_to_soa(d_D + inst[k].data_offset, Ls[k].b.X, Ls[k].b.perm, G[k].n, G[k].d, s);
        rc = reset_scan_state(G[k], Ls[k], s);
        if (rc) return rc;
    }

    // largest first onto the least-loaded stream
    std::vector<int> order = active;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return inst[a].n * inst[a].d > inst[b].n * inst[b].d; });
    std::vector<std::vector<int>> per(S);
    std::vector<int64_t> load(S, 0);
    for (int k : order) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        per[j].push_back(k);
        load[j] += inst[k].n * inst[k].d;
    }
    const int K = std::max(1, opt.iters_per_sync);
    // instances still iterating; finished ones are no longer enqueued
    std::vector<char> live(n_inst, 0);
    for (int k : active) live[k] = 1;
    // instances whose grid holds at most kSmallSortRuns runs (plain launches
    // only: a captured graph keeps the radix sort)
    std::vector<char> small(n_inst, 0);
    auto enqueue_batch = [&](hipStream_t origin) -> int {
        int r = fork_streams(bw, origin, S);
        if (r) return r;
        for (int j = 0; j < S; j++)
            for (int k : per[j])
                if (live[k])
                    for (int it = 0; it < K; it++)
                        enqueue_iteration(G[k], Ls[k], opt, bw.streams[j], nullptr, small[k]);
        return join_streams(bw, origin, S);
    };
    // the graph holds the live set it was captured with; it is captured again
    // when that set has shrunk to half, and plain launches cover the rest.
    // Graphs are kept to kMaxBatchGraphNodes kernel nodes: a capture of 256
    // instances x 16 iterations (~200k nodes) ended the host process with a
    // segmentation fault, 256 x 4 (~80k) did not.
    int graph_live = 0;
    auto ensure_graph = [&](int n_live) -> int {
        if (!opt.use_graph) return 0;
        if ((long long)n_live * K * kGraphNodesPerIter > kMaxBatchGraphNodes) {
            drop_graph(bw.gexec, bw.gkey);
            graph_live = 0;
            return 0;
        }
        if (bw.gexec && graph_live > 0 && 2 * n_live > graph_live) return 0;
        std::vector<char> key;
        key_add(key, bw.base);
        key_add(key, bw.states);
        key_add(key, K);
        key_add(key, S);
        key_add(key, opt.stop_rule);
        for (int j = 0; j < S; j++)
            for (int k : per[j])
                if (live[k]) {
                    key_add(key, k);
                    key_add(key, off[k]);
                    key_add(key, G[k]);
                }
        if (!bw.gexec || key != bw.gkey) {
            drop_graph(bw.gexec, bw.gkey);
            int inner = 0;
            const int r = capture_graph(bw.origin, bw.gexec,
                                        [&] { inner = enqueue_batch(bw.origin); });
            if (r) return r;
            if (inner) {
                drop_graph(bw.gexec, bw.gkey);
                return inner;
            }
            bw.gkey = key;
        }
        graph_live = n_live;
        return 0;
    };
    int n_live = (int)active.size();
    rc = ensure_graph(n_live);
    if (rc) return rc;
    Events E;
    if (opt.timing) {
        rc = E.ensure(4);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(E.ev[0], s));
    }
    const auto t_setup = clk::now();
    int max_iter_seen = 0;
    for (;;) {
        if (opt.use_graph && graph_live == n_live) {
            HIP_TRY(hipGraphLaunch(bw.gexec, s));
        } else {
            rc = enqueue_batch(s);
            if (rc) return rc;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(bw.h_states, bw.states, sizeof(DevState) * n_inst,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        n_live = 0;
        for (int k : active) {
            const DevState &hk = bw.h_states[k];
            live[k] = !hk.done;
            n_live += live[k];
            max_iter_seen = std::max(max_iter_seen, hk.iter);
            if (!opt.use_graph) small[k] = hk.runs_valid && hk.runs <= kSmallSortRuns;
        }
        if (n_live == 0) break;
        if (opt.use_graph && n_live != graph_live) {
            rc = ensure_graph(n_live);
            if (rc) return rc;
        }
    }
    if (opt.timing) HIP_TRY(hipEventRecord(E.ev[1], s));
    // final grids and labels (EGG_SynC.cu:1504-1541), one per instance
    rc = fork_streams(bw, s, S);
    if (rc) return rc;
    for (int j = 0; j < S; j++) {
        hipStream_t sj = bw.streams[j];
        for (int k : per[j]) {
            Layout &L = Ls[k];
            const DevState &hk = bw.h_states[k];
            build_grid(G[k], L.b, nullptr, L.st, sj, hk.runs_valid && hk.runs <= kSmallSortRuns);
            int32_t *lab = d_labels + row_off[k];
            if (opt.label_rule == EGG_SYNC_LABELS_CELLS)
                launch_labels_egg(G[k], L.b, lab, L.b.nchunk, L.b.wi_off, L.st, sj);
            else
                launch_labels_sync(G[k], L.b, lab, L.b.head, L.st, sj);
            if (opt.compute_rc && opt.stop_rule == EGG_SYNC_STOP_EGG)
                launch_update_sync(G[k], L.b, L.st, L.rc_part, false, sj);
            if (d_final)
                launch_soa_to_aos_by_perm(L.b.Xs, L.b.perm_s, d_final + fl_off[k], G[k].n, G[k].d,
                                          sj);
        }
    }
    rc = join_streams(bw, s, S);
    if (rc) return rc;
    if (opt.timing) HIP_TRY(hipEventRecord(E.ev[2], s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(bw.h_states, bw.states, sizeof(DevState) * n_inst,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k : active) {
        rc = check_scans(Ls[k], s);
        if (rc) return rc;
    }
    bool capped = false;
    int64_t clusters = 0, cells = 0;
    for (int k : active) {
        const DevState &f = bw.h_states[k];
        capped = capped || f.capped;
        clusters += f.n_clusters;
        cells += f.C;
        if (res)
            res[k] = egg_sync_instance_result{f.iter, f.capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK,
                                              f.n_clusters, f.C};
    }
    if (so) {
        so->iterations = max_iter_seen;
        so->status = capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK;
        so->n_clusters = (int32_t)std::min<int64_t>(clusters, INT_MAX);
        so->n_cells = (int32_t)std::min<int64_t>(cells, INT_MAX);
        so->stage_ms[0] = std::chrono::duration<double, std::milli>(t_setup - t0).count();
        if (opt.timing) {
            so->stage_ms[2] = ms_between(E.ev[0], E.ev[1]);
            so->stage_ms[4] = ms_between(E.ev[1], E.ev[2]);
        }
        so->workspace_bytes = (int64_t)(total + sizeof(DevState) * n_inst);
        so->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    return capped ? EGG_SYNC_CAPPED : EGG_SYNC_OK;
}

}  // namespace
}  // namespace egg

using namespace egg;

extern "C" {

int egg_sync_default_options(egg_sync_options *opt) {
    if (!opt) return EGG_SYNC_E_ARG;
    default_options(*opt);
    return EGG_SYNC_OK;
}

int egg_sync_solve(const float *h_D, int64_t n, int32_t d, float eps, int32_t stop_rule,
                   int32_t *labels_out, egg_sync_stats *stats_out) {
    egg_sync_options opt;
    default_options(opt);
    opt.stop_rule = stop_rule;
    if (stop_rule == EGG_SYNC_STOP_RC) opt.label_rule = EGG_SYNC_LABELS_SYNC;
    return egg_sync_solve_ex(h_D, n, d, eps, &opt, labels_out, nullptr, stats_out);
}

int egg_sync_solve_ex(const float *h_D, int64_t n, int32_t d, float eps,
                      const egg_sync_options *opt, int32_t *labels_out, float *final_out,
                      egg_sync_stats *stats_out) {
    if ((!h_D || !labels_out) && n > 0) return EGG_SYNC_E_ARG;
    egg_sync_options o;
    if (opt) o = *opt;
    else default_options(o);
    std::lock_guard<std::mutex> lk(g_mutex);
    return solve_device_impl(nullptr, n, d, eps, o, nullptr, nullptr, stats_out, nullptr, h_D,
                             labels_out, final_out);
}

int egg_sync_solve_device(const float *d_D, int64_t n, int32_t d, float eps,
                          const egg_sync_options *opt, int32_t *d_labels, float *d_final,
                          egg_sync_stats *stats_out, void *stream) {
    if ((!d_D || !d_labels) && n > 0) return EGG_SYNC_E_ARG;
    egg_sync_options o;
    if (opt) o = *opt;
    else default_options(o);
    std::lock_guard<std::mutex> lk(g_mutex);
    return solve_device_impl(d_D, n, d, eps, o, d_labels, d_final, stats_out,
                             (hipStream_t)stream, nullptr, nullptr, nullptr);
}

int egg_sync_solve_device_noalloc(const float *d_D, int64_t n, int32_t d, float eps,
                                  const egg_sync_options *opt, int32_t *d_labels,
                                  float *d_final, egg_sync_stats *stats_out, void *stream) {
    if ((!d_D || !d_labels) && n > 0) return EGG_SYNC_E_ARG;
    egg_sync_options o;
    if (opt) o = *opt;
    else default_options(o);
    std::lock_guard<std::mutex> lk(g_mutex);
    return solve_device_impl(d_D, n, d, eps, o, d_labels, d_final, stats_out,
                             (hipStream_t)stream, nullptr, nullptr, nullptr, true);
}

int egg_sync_solve_batch_device(const float *d_D, const egg_sync_instance *inst,
                                int32_t n_inst, const egg_sync_options *opt,
                                int32_t *d_labels, float *d_final,
                                egg_sync_instance_result *results, egg_sync_stats *stats,
                                void *stream) {
    egg_sync_options o;
    if (opt) o = *opt;
    else default_options(o);
    std::lock_guard<std::mutex> lk(g_mutex);
    return solve_batch_impl(d_D, inst, n_inst, o, d_labels, d_final, results, stats,
                            (hipStream_t)stream);
}

int egg_sync_build_grid_device(const float *d_D, int64_t n, int32_t d, float eps,
                               const int32_t *d_order_in, int32_t *d_order_out,
                               int32_t *d_cell_of, int32_t *d_cell_start, int32_t *d_cell_g,
                               int32_t *d_ob, int32_t *d_oe, int32_t *C_out, void *stream) {
    std::lock_guard<std::mutex> lk(g_mutex);
    Geo g;
    Layout L;
    hipStream_t s;
    int rc = setup_single(d_D, n, d, eps, d_order_in, g, L, s, stream);
    if (rc) return rc;
    build_grid(g, L.b, nullptr, L.st, s);
    const size_t nn = (size_t)g.n;
    if (d_order_out)
        HIP_TRY(hipMemcpyAsync(d_order_out, L.b.perm_s, 4 * nn, hipMemcpyDeviceToDevice, s));
    if (d_cell_of)
        HIP_TRY(hipMemcpyAsync(d_cell_of, L.b.point_cell, 4 * nn, hipMemcpyDeviceToDevice, s));
    if (d_cell_g) k_soa_to_rows<<<grid_blocks(g.n), kBlock, 0, s>>>(L.b.cell_g, g.n, d, L.st, d_cell_g);
    if (d_ob) HIP_TRY(hipMemcpyAsync(d_ob, L.b.ob, 4 * (size_t)g.O, hipMemcpyDeviceToDevice, s));
    if (d_oe) HIP_TRY(hipMemcpyAsync(d_oe, L.b.oe, 4 * (size_t)g.O, hipMemcpyDeviceToDevice, s));
    DevState hs;
    HIP_TRY(hipMemcpyAsync(&hs, L.st, sizeof hs, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (d_cell_start)
        HIP_TRY(hipMemcpy(d_cell_start, L.b.cell_start, 4 * ((size_t)hs.C + 1),
                          hipMemcpyDeviceToDevice));
    if (C_out) *C_out = hs.C;
    HIP_TRY(hipGetLastError());
    return check_scans(L, s);
}

int egg_sync_iteration_device(const float *d_D, int64_t n, int32_t d, float eps,
                              const int32_t *d_order_in, float *d_next, int32_t *d_order_out,
                              int32_t *d_N, int32_t *d_half, int32_t *d_border,
                              int32_t *d_synced, int32_t *flag_local, int32_t *flag_final,
                              void *stream) {
    std::lock_guard<std::mutex> lk(g_mutex);
    Geo g;
    Layout L;
    hipStream_t s;
    int rc = setup_single(d_D, n, d, eps, d_order_in, g, L, s, stream);
    if (rc) return rc;
    DebugOut dbg{d_N, d_half, d_border, d_synced};
    launch_set_flag(L.st, s);
    build_grid(g, L.b, L.st, L.st, s);
    cell_summaries(g, L.b, L.st, L.st, s);
    launch_update_egg(g, L.b, L.st, &dbg, s);
    launch_extra_check(g, L.b, L.st, s);
    if (d_next) launch_soa_to_aos_by_perm(L.b.X, L.b.perm, d_next, g.n, d, s);
    if (d_order_out)
        HIP_TRY(hipMemcpyAsync(d_order_out, L.b.perm, 4 * (size_t)g.n, hipMemcpyDeviceToDevice, s));
    DevState hs;
    HIP_TRY(hipMemcpyAsync(&hs, L.st, sizeof hs, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (flag_local) *flag_local = hs.flag_local;
    if (flag_final) *flag_final = hs.flag;
    return check_scans(L, s);
}

int egg_sync_features_device(const float *d_D, const float *d_prev, int64_t n, int32_t d,
                             float eps, float *d_feat, void *stream) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!d_feat) return EGG_SYNC_E_ARG;
    Geo g;
    Layout L;
    hipStream_t s;
    int rc = setup_single(d_D, n, d, eps, nullptr, g, L, s, stream);
    if (rc) return rc;
    build_grid(g, L.b, nullptr, L.st, s);
    cell_summaries(g, L.b, nullptr, L.st, s);
    k_features<<<grid_blocks(g.n), kBlock, 0, s>>>(L.b.Xs, L.b.perm_s, L.b.point_cell,
                                                   L.b.cell_start, L.b.Ts, L.b.Tc, d_prev, g.n, d,
                                                   1.f / eps, d_feat);
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    return check_scans(L, s);
}

int egg_sync_rc_device(const float *d_D, int64_t n, int32_t d, float eps, double *rc_out,
                       void *stream) {
    std::lock_guard<std::mutex> lk(g_mutex);
    Geo g;
    Layout L;
    hipStream_t s;
    int rc = setup_single(d_D, n, d, eps, nullptr, g, L, s, stream);
    if (rc) return rc;
    build_grid(g, L.b, nullptr, L.st, s);
    launch_update_sync(g, L.b, L.st, L.rc_part, false, s);
    DevState hs;
    HIP_TRY(hipMemcpyAsync(&hs, L.st, sizeof hs, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (rc_out) *rc_out = hs.rc;
    return check_scans(L, s);
}

int egg_sync_geometry(int64_t n, int32_t d, float eps, int64_t *out5, float *cs_out) {
    Geo g;
    int rc = compute_geo(n, d, eps, g);
    if (rc) return rc;
    if (out5) {
        out5[0] = g.w;
        out5[1] = g.od;
        out5[2] = g.O;
        out5[3] = g.R;
        out5[4] = g.key_bits;
    }
    if (cs_out) *cs_out = g.cs;
    return EGG_SYNC_OK;
}

int egg_sync_release(void) {
    std::lock_guard<std::mutex> lk(g_mutex);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return EGG_SYNC_E_NODEV;
    if (dev < 0 || dev >= 64) return EGG_SYNC_E_NODEV;
    Workspace &ws = g_ws[dev];
    drop_graph(ws.gexec, ws.gkey);
    if (ws.base) (void)hipFree(ws.base);
    if (ws.stream) (void)hipStreamDestroy(ws.stream);
    if (ws.h_state) (void)hipHostFree(ws.h_state);
    ws = Workspace{};
    release_batch(g_bws[dev]);
    return EGG_SYNC_OK;
}

const char *egg_sync_error_string(int code) {
    switch (code) {
        case EGG_SYNC_OK: return "ok";
        case EGG_SYNC_CAPPED: return "iteration cap reached before convergence";
        case EGG_SYNC_E_ARG: return "invalid argument";
        case EGG_SYNC_E_RANGE: return "input values must be finite and lie in [0, 1]";
        case EGG_SYNC_E_DIM: return "dimension above EGG_SYNC_MAX_D";
        case EGG_SYNC_E_KEY: return "grid too fine: cell keys exceed 256 bits";
        case EGG_SYNC_E_ALLOC: return "device allocation failed";
        case EGG_SYNC_E_HIP: return "HIP runtime error";
        case EGG_SYNC_E_NODEV: return "no GPU device";
        case EGG_SYNC_E_INTERNAL: return "internal error: a bounded device-side wait expired";
        case EGG_SYNC_E_TOO_BIG:
            return "workload too large for the no-alloc scratch array (64 MiB of workspace)";
        default: return "unknown error";
    }
}

}  // extern "C"
