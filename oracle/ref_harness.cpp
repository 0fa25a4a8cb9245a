// Driver for the reference's own CPU SynC (src/algorithms/SynC.cpp), compiled
// from the reference sources by oracle/build_ref.sh into oracle/_ref/sync_ref.
// TEST INFRASTRUCTURE ONLY: it writes golden vectors (tests/golden/) and times
// the reference CPU path for the benchmark's CPU baseline.
//
// Input file  (little endian): int32 n, int32 d, float eps, float lam, float[n*d]
// Commands:
//   solve <in> <out>   DynamicalClustering(D, n, d, eps, lam) (SynC.cpp:268-326):
//                      labels and iteration count straight from the reference,
//                      plus a replay of its loop through the reference's
//                      ComputeNeighborhood / UpdatePoint / ComputeLocationOrder /
//                      synCluster_v2 that also records positions and r per pass.
//                      out: int32 n, d, iters_ref, iters_replay,
//                           int32 labels_ref[n], int32 labels_replay[n],
//                           float X_final[n*d], float r_trace[iters_replay]
//   step <in> <out>    one pass of ComputeNeighborhood + UpdatePoint +
//                      ComputeLocationOrder over every point (SynC.cpp:281-287).
//                      out: int32 n, d, float X_next[n*d], int32 N[n], float r_p[n]
//   time <in> <t0> <t1> <threads>
//                      the same pass for targets [t0,t1), OpenMP-parallel like
//                      DynamicalClustering_parallel (SynC.cpp:352-361); prints
//                      "seconds=<wall> targets=<t1-t0> threads=<threads>"
//   fsolve <in> <out> <B>
//                      FDynamicalClustering(D, n, d, eps, lam, B) (SynC.cpp:426-487,
//                      FSynC): labels and iteration count straight from the reference.
//                      out: int32 n, int32 iters, int32 labels[n]
//   ftime <in> <t0> <t1> <B>
//                      FSynC's pass (FDynamicalClustering, SynC.cpp:426-462): the
//                      reference's R-tree (src/structure/RTree.h, node capacity B)
//                      over all n points, then for targets [t0,t1) its range query,
//                      UpdatePoint and ComputeLocationOrder, one thread as in the
//                      reference; prints "build_seconds=<wall> seconds=<wall>
//                      targets=<t1-t0> threads=1"
#include <ATen/ATen.h>
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "algorithms/SynC.h"

// The reference's R-tree is header-only but also defines a non-inline free
// function, which SynC.cpp's copy already provides; a namespace of its own
// keeps this translation unit's copy apart.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <utility>
namespace ref_rtree {
#include "structure/RTree.h"
}

at::Tensor synCluster_v2(float *D, int n, int d, float eps);  // SynC.cpp:88 (not in SynC.h)

static bool read_input(const char *path, int &n, int &d, float &eps, float &lam,
                       std::vector<float> &D) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fread(&n, 4, 1, f) == 1 && std::fread(&d, 4, 1, f) == 1 &&
              std::fread(&eps, 4, 1, f) == 1 && std::fread(&lam, 4, 1, f) == 1;
    if (ok) {
        D.resize((size_t)n * d);
        ok = std::fread(D.data(), 4, D.size(), f) == D.size();
    }
    std::fclose(f);
    return ok;
}

template <typename T>
static void put(FILE *f, const T *p, size_t cnt) {
    if (std::fwrite(p, sizeof(T), cnt, f) != cnt) std::exit(3);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: sync_ref solve|step|time <in> ...\n");
        return 2;
    }
    std::string cmd = argv[1];
    int n, d;
    float eps, lam;
    std::vector<float> D;
    if (!read_input(argv[2], n, d, eps, lam, D)) {
        std::fprintf(stderr, "cannot read %s\n", argv[2]);
        return 2;
    }
    if (cmd == "solve" && argc >= 4) {
        std::vector<float> D0 = D;
        std::vector<at::Tensor> res = DynamicalClustering(D0.data(), n, d, eps, lam);
        at::Tensor labels_ref = res[0].contiguous();
        int iters_ref = (int)res[1].size(0);

        std::vector<float> cur = D, nxt = D;
        std::vector<float> trace;
        float r_local = 0.;
        while (r_local < lam) {
            r_local = 0.;
            for (int p = 0; p < n; p++) {
                const neighborhood N_t = ComputeNeighborhood(p, eps, cur.data(), n, d);
                UpdatePoint(p, N_t, cur.data(), nxt.data(), n, d);
                r_local += ComputeLocationOrder(p, N_t, cur.data(), n, d);
            }
            r_local /= n;
            trace.push_back(r_local);
            std::swap(cur, nxt);
        }
        at::Tensor labels_rep = synCluster_v2(cur.data(), n, d, eps).contiguous();
        int iters_rep = (int)trace.size();

        FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 2;
        put(f, &n, 1);
        put(f, &d, 1);
        put(f, &iters_ref, 1);
        put(f, &iters_rep, 1);
        put(f, labels_ref.data_ptr<int>(), (size_t)n);
        put(f, labels_rep.data_ptr<int>(), (size_t)n);
        put(f, cur.data(), cur.size());
        put(f, trace.data(), trace.size());
        std::fclose(f);
        return 0;
    }
    if (cmd == "step" && argc >= 4) {
        std::vector<float> nxt = D;
        std::vector<int> cnt(n);
        std::vector<float> rp(n);
        for (int p = 0; p < n; p++) {
            const neighborhood N_t = ComputeNeighborhood(p, eps, D.data(), n, d);
            UpdatePoint(p, N_t, D.data(), nxt.data(), n, d);
            rp[p] = ComputeLocationOrder(p, N_t, D.data(), n, d);
            cnt[p] = (int)N_t.size();
        }
        FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 2;
        put(f, &n, 1);
        put(f, &d, 1);
        put(f, nxt.data(), nxt.size());
        put(f, cnt.data(), cnt.size());
        put(f, rp.data(), rp.size());
        std::fclose(f);
        return 0;
    }
    if (cmd == "fsolve" && argc >= 5) {
        std::vector<float> D0 = D;
        const int B = std::atoi(argv[4]);
        std::vector<at::Tensor> res = FDynamicalClustering(D0.data(), n, d, eps, lam, B);
        at::Tensor labels = res[0].contiguous();
        int iters = (int)res[1].size(0);
        FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 2;
        put(f, &n, 1);
        put(f, &iters, 1);
        put(f, labels.data_ptr<int>(), (size_t)n);
        std::fclose(f);
        return 0;
    }
    if (cmd == "ftime" && argc >= 6) {
        const long t0 = std::atol(argv[3]), t1 = std::atol(argv[4]);
        const int B = std::atoi(argv[5]);
        std::vector<float> nxt = D;
        auto a = std::chrono::steady_clock::now();
        ref_rtree::RTree tree(d, B);
        for (int p = 0; p < n; p++) tree.insert(new std::pair<int, float *>(p, &D[(size_t)p * d]));
        auto b = std::chrono::steady_clock::now();
        float r_local = 0.f;
        for (long p = t0; p < t1; p++) {
            const neighborhood N_t = tree.range(eps, &D[(size_t)p * d]);
            UpdatePoint((int)p, N_t, D.data(), nxt.data(), n, d);
            r_local += ComputeLocationOrder((int)p, N_t, D.data(), n, d);
        }
        auto c = std::chrono::steady_clock::now();
        std::printf("build_seconds=%.6f seconds=%.6f targets=%ld threads=1 r=%g\n",
                    std::chrono::duration<double>(b - a).count(),
                    std::chrono::duration<double>(c - b).count(), t1 - t0, (double)r_local);
        return 0;
    }
    if (cmd == "time" && argc >= 6) {
        long t0 = std::atol(argv[3]), t1 = std::atol(argv[4]);
        int threads = std::atoi(argv[5]);
        if (threads > 0) omp_set_num_threads(threads);
        std::vector<float> nxt = D;
        std::vector<float> rs(omp_get_max_threads(), 0.f);
        auto a = std::chrono::steady_clock::now();
#pragma omp parallel for
        for (long p = t0; p < t1; p++) {
            const neighborhood N_t = ComputeNeighborhood((int)p, eps, D.data(), n, d);
            UpdatePoint((int)p, N_t, D.data(), nxt.data(), n, d);
            rs[omp_get_thread_num()] += ComputeLocationOrder((int)p, N_t, D.data(), n, d);
        }
        auto b = std::chrono::steady_clock::now();
        double s = std::chrono::duration<double>(b - a).count();
        std::printf("seconds=%.6f targets=%ld threads=%d\n", s, t1 - t0, omp_get_max_threads());
        return 0;
    }
    std::fprintf(stderr, "bad command\n");
    return 2;
}
