// MLP of the learned jump: in -> hidden -> hidden -> out, ReLU, FP32.
//
// Training is expressed as small GEMMs on a batch (activations [B x width],
// row-major) with a strided FP32 GEMM kernel: 64x64 output tiles staged
// through LDS, 4x4 outputs per thread, split over K with a deterministic
// reduction of the K-slices, so gradients are reproducible bit for bit.
// Inference for millions of points uses a fused kernel that keeps all
// weights in LDS (18.3k floats for 12-128-128-3) and one point per thread.
//
// Parameters live in one flat array: W1[h][in], b1[h], W2[h][h], b2[h],
// W3[out][h], b3[out].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/egg_sync.h"

namespace egg {
namespace mlp {

constexpr int kT = 64;      // output tile edge
constexpr int kKT = 16;     // K tile
constexpr int kThreads = 256;

struct Dims {
    int in, h, out;
    long long n_w1() const { return (long long)h * in; }
    long long off_b1() const { return n_w1(); }
    long long off_w2() const { return off_b1() + h; }
    long long off_b2() const { return off_w2() + (long long)h * h; }
    long long off_w3() const { return off_b2() + h; }
    long long off_b3() const { return off_w3() + (long long)out * h; }
    long long total() const { return off_b3() + out; }
};

// C_part[s][m][n] = sum_{k in slice s} A(m,k) * B(k,n), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn].
__global__ __launch_bounds__(kThreads) void k_gemm_part(const float *__restrict__ A,
                                                        long long sam, long long sak,
                                                        const float *__restrict__ B,
                                                        long long sbk, long long sbn,
                                                        float *__restrict__ Cpart, int M, int N,
                                                        int K, int kslice) {
    __shared__ float sA[kKT][kT + 1];
    __shared__ float sB[kKT][kT + 1];
    const int tm = blockIdx.y * kT, tn = blockIdx.x * kT, s = blockIdx.z;
    const int k0 = s * kslice, k1 = min(K, k0 + kslice);
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    float acc[4][4] = {};
    for (int kb = k0; kb < k1; kb += kKT) {
        for (int e = threadIdx.x; e < kKT * kT; e += kThreads) {
            const int kk = e / kT, mm = e % kT;
            const int k = kb + kk;
            sA[kk][mm] = (k < k1 && tm + mm < M) ? A[(tm + mm) * sam + k * sak] : 0.f;
            sB[kk][mm] = (k < k1 && tn + mm < N) ? B[k * sbk + (tn + mm) * sbn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKT; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                a[q] = sA[kk][ty * 4 + q];
                b[q] = sB[kk][tx * 4 + q];
            }
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[p][q] = fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
    float *C = Cpart + (long long)s * M * N;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = tm + ty * 4 + p, n = tn + tx * 4 + q;
            if (m < M && n < N) C[(long long)m * N + n] = acc[p][q];
        }
}

// C[m][n] (row stride ldc) = sum_s Cpart[s][m][n] (+ bias[n]); act: 0 none, 1 ReLU.
__global__ void k_gemm_reduce(const float *__restrict__ Cpart, int S, int M, int N,
                              const float *__restrict__ bias, int act, float *__restrict__ C,
                              long long ldc) {
    const long long MN = (long long)M * N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < MN;
         e += (long long)gridDim.x * blockDim.x) {
        float v = 0.f;
        for (int s = 0; s < S; s++) v += Cpart[s * MN + e];
        const int n = (int)(e % N);
        if (bias) v += bias[n];
        if (act == 1) v = fmaxf(v, 0.f);
        C[(e / N) * ldc + n] = v;
    }
}

// dZ = dH * (Z > 0) in place on dH (ReLU derivative).
__global__ void k_relu_mask(float *__restrict__ dH, const float *__restrict__ H, long long count) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
         e += (long long)gridDim.x * blockDim.x)
        if (!(H[e] > 0.f)) dH[e] = 0.f;
}

// Column sums of a [rows x cols] matrix (bias gradients), fixed order.
__global__ void k_colsum(const float *__restrict__ X, int rows, int cols, float *__restrict__ out) {
    const int c = blockIdx.x;
    __shared__ float s[kThreads];
    float v = 0.f;
    for (int r = threadIdx.x; r < rows; r += kThreads) v += X[(long long)r * cols + c];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
}

// dO = 2 (O - Y) / B, loss partials per block.
__global__ void k_mse_grad(const float *__restrict__ O, const float *__restrict__ Y, long long count,
                           float inv_b, float *__restrict__ dO, double *__restrict__ loss_part) {
    __shared__ double s[kThreads];
    double l = 0.0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
         e += (long long)gridDim.x * blockDim.x) {
        const float r = O[e] - Y[e];
        dO[e] = 2.f * r * inv_b;
        l += (double)r * (double)r;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = s[0];
}

__global__ void k_loss_finish(const double *__restrict__ part, int nparts, double inv_b,
                              double *__restrict__ loss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nparts; i++) t += part[i];
        *loss = t * inv_b;
    }
}

// Adam (Kingma & Ba 2015) on the flat parameter array.
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ Mo,
                       float *__restrict__ Ve, long long count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
         e += (long long)gridDim.x * blockDim.x) {
        const float g = G[e];
        const float m = b1 * Mo[e] + (1.f - b1) * g;
        const float v = b2 * Ve[e] + (1.f - b2) * g * g;
        Mo[e] = m;
        Ve[e] = v;
        P[e] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

// Fused inference, one point per thread, weights in LDS.  Hidden layer 2 is
// produced 8 units at a time and folded into the outputs at once, so only
// layer 1 (H units) stays in registers.
template <int IN, int H, int OUT>
__global__ __launch_bounds__(kThreads) void k_mlp_infer(const float *__restrict__ P,
                                                        const float *__restrict__ X, long long B,
                                                        float *__restrict__ Y) {
    constexpr int NP = H * IN + H + H * H + H + OUT * H + OUT;
    __shared__ float sP[NP];
    for (int e = threadIdx.x; e < NP; e += kThreads) sP[e] = P[e];
    __syncthreads();
    const float *W1 = sP, *b1 = W1 + H * IN, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H,
                *b3 = W3 + OUT * H;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < B;
         i += (long long)gridDim.x * kThreads) {
        float x[IN];
#pragma unroll
        for (int k = 0; k < IN; k++) x[k] = X[i * IN + k];
        float h1[H];
#pragma unroll
        for (int j = 0; j < H; j++) {
            float z = b1[j];
#pragma unroll
            for (int k = 0; k < IN; k++) z = fmaf(W1[j * IN + k], x[k], z);
            h1[j] = fmaxf(z, 0.f);
        }
        float o[OUT];
#pragma unroll
        for (int m = 0; m < OUT; m++) o[m] = b3[m];
        for (int j0 = 0; j0 < H; j0 += 8) {
#pragma unroll
            for (int jj = 0; jj < 8; jj++) {
                const int j = j0 + jj;
                float z = b2[j];
#pragma unroll
                for (int k = 0; k < H; k++) z = fmaf(W2[j * H + k], h1[k], z);
                z = fmaxf(z, 0.f);
#pragma unroll
                for (int m = 0; m < OUT; m++) o[m] = fmaf(W3[m * H + j], z, o[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < OUT; m++) Y[i * OUT + m] = o[m];
    }
}

inline int blocks_for(long long n, int cap = 4096) {
    return (int)std::max(1LL, std::min((long long)cap, (n + kThreads - 1) / kThreads));
}

// C[M x N] (ldc) = A B (+bias, act) via split-K partial tiles in `part`.
void gemm(const float *A, long long sam, long long sak, const float *B, long long sbk,
          long long sbn, float *C, long long ldc, int M, int N, int K, const float *bias, int act,
          float *part, long long part_cap, hipStream_t s) {
    const int tiles = ((M + kT - 1) / kT) * ((N + kT - 1) / kT);
    int S = std::max(1, std::min(256, 1024 / std::max(tiles, 1)));
    S = std::min<long long>(S, std::max<long long>(1, part_cap / ((long long)M * N)));
    S = std::min(S, std::max(1, (K + kKT - 1) / kKT));
    int kslice = (K + S - 1) / S;
    kslice = (kslice + kKT - 1) / kKT * kKT;
    S = (K + kslice - 1) / kslice;
    dim3 grid((N + kT - 1) / kT, (M + kT - 1) / kT, S);
    k_gemm_part<<<grid, kThreads, 0, s>>>(A, sam, sak, B, sbk, sbn, part, M, N, K, kslice);
    k_gemm_reduce<<<blocks_for((long long)M * N), kThreads, 0, s>>>(part, S, M, N, bias, act, C, ldc);
}

}  // namespace mlp
}  // namespace egg

using namespace egg::mlp;

extern "C" {

int egg_mlp_param_count(int32_t in, int32_t hidden, int32_t out, int64_t *count) {
    if (in < 1 || hidden < 1 || out < 1 || !count) return EGG_SYNC_E_ARG;
    *count = Dims{in, hidden, out}.total();
    return EGG_SYNC_OK;
}

size_t egg_mlp_workspace_floats(int32_t in, int32_t hidden, int32_t out, int64_t batch) {
    const long long B = batch, h = hidden;
    // Z1/H1, Z2/H2 (B x h each, ReLU outputs kept), O, dO, dH2, dH1, split-K parts
    return (size_t)(2 * B * h + 2 * B * out + 2 * B * h + 1024LL * 64 * 64 + 4096);
}

int egg_mlp_forward_device(const float *d_params, int32_t in, int32_t hidden, int32_t out,
                           const float *d_X, int64_t batch, float *d_Y, void *stream) {
    if (!d_params || !d_X || !d_Y || batch < 0) return EGG_SYNC_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (batch == 0) return EGG_SYNC_OK;
    if (in == 12 && hidden == 128 && out == 3) {
        k_mlp_infer<12, 128, 3><<<blocks_for(batch, 8192), kThreads, 0, s>>>(d_params, d_X, batch, d_Y);
    } else {
        return EGG_SYNC_E_ARG;
    }
    return hipGetLastError() == hipSuccess ? EGG_SYNC_OK : EGG_SYNC_E_HIP;
}

// One training step on a batch: loss, gradients (optional copy to d_grad) and,
// if lr > 0, an Adam update with step number `step` (1-based).
int egg_mlp_train_step_device(float *d_params, float *d_adam_m, float *d_adam_v, int32_t in,
                              int32_t hidden, int32_t out, const float *d_X, const float *d_T,
                              int64_t batch, float lr, int32_t step, float *d_work,
                              int64_t work_floats, float *d_grad, double *loss_out,
                              void *stream) {
    if (!d_params || !d_X || !d_T || !d_work || !d_grad || batch < 1) return EGG_SYNC_E_ARG;
    const long long B = batch, h = hidden;
    if ((size_t)work_floats < egg_mlp_workspace_floats(in, hidden, out, batch)) return EGG_SYNC_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const Dims D{in, hidden, out};
    float *H1 = d_work, *H2 = H1 + B * h, *O = H2 + B * h, *dO = O + B * out, *dH2 = dO + B * out,
          *dH1 = dH2 + B * h, *part = dH1 + B * h;
    const long long part_cap = 1024LL * 64 * 64;
    double *loss_part = (double *)(part + part_cap);  // 1024 doubles fit in the 4096-float tail
    float *G = d_grad;
    const float *W1 = d_params, *b1 = W1 + D.off_b1(), *W2 = W1 + D.off_w2(), *b2 = W1 + D.off_b2(),
                *W3 = W1 + D.off_w3(), *b3 = W1 + D.off_b3();
    // forward: H1 = relu(X W1^T + b1), H2 = relu(H1 W2^T + b2), O = H2 W3^T + b3
    gemm(d_X, in, 1, W1, 1, in, H1, h, (int)B, hidden, in, b1, 1, part, part_cap, s);
    gemm(H1, h, 1, W2, 1, h, H2, h, (int)B, hidden, hidden, b2, 1, part, part_cap, s);
    gemm(H2, h, 1, W3, 1, h, O, out, (int)B, out, hidden, b3, 0, part, part_cap, s);
    // loss and dO
    const int lb = 256;
    k_mse_grad<<<lb, kThreads, 0, s>>>(O, d_T, B * out, 1.f / (float)B, dO, loss_part);
    double *d_loss = loss_part + lb;
    k_loss_finish<<<1, 1, 0, s>>>(loss_part, lb, 1.0 / (double)B, d_loss);
    // backward
    float *gW1 = G, *gb1 = G + D.off_b1(), *gW2 = G + D.off_w2(), *gb2 = G + D.off_b2(),
          *gW3 = G + D.off_w3(), *gb3 = G + D.off_b3();
    gemm(dO, 1, out, H2, h, 1, gW3, h, out, hidden, (int)B, nullptr, 0, part, part_cap, s);  // dO^T H2
    k_colsum<<<out, kThreads, 0, s>>>(dO, (int)B, out, gb3);
    gemm(dO, out, 1, W3, h, 1, dH2, h, (int)B, hidden, out, nullptr, 0, part, part_cap, s);  // dO W3
    k_relu_mask<<<blocks_for(B * h), kThreads, 0, s>>>(dH2, H2, B * h);
    gemm(dH2, 1, h, H1, h, 1, gW2, h, hidden, hidden, (int)B, nullptr, 0, part, part_cap, s);
    k_colsum<<<hidden, kThreads, 0, s>>>(dH2, (int)B, hidden, gb2);
    gemm(dH2, h, 1, W2, h, 1, dH1, h, (int)B, hidden, hidden, nullptr, 0, part, part_cap, s);  // dH2 W2
    k_relu_mask<<<blocks_for(B * h), kThreads, 0, s>>>(dH1, H1, B * h);
    gemm(dH1, 1, h, d_X, in, 1, gW1, in, hidden, in, (int)B, nullptr, 0, part, part_cap, s);
    k_colsum<<<hidden, kThreads, 0, s>>>(dH1, (int)B, hidden, gb1);
    if (lr > 0.f) {
        if (!d_adam_m || !d_adam_v || step < 1) return EGG_SYNC_E_ARG;
        const float be1 = 0.9f, be2 = 0.999f;
        const float c1 = 1.f - powf(be1, (float)step), c2 = 1.f - powf(be2, (float)step);
        k_adam<<<blocks_for(D.total()), kThreads, 0, s>>>(d_params, G, d_adam_m, d_adam_v, D.total(),
                                                           lr, be1, be2, 1e-8f, c1, c2);
    }
    if (loss_out) {
        if (hipMemcpyAsync(loss_out, d_loss, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
            return EGG_SYNC_E_HIP;
        if (hipStreamSynchronize(s) != hipSuccess) return EGG_SYNC_E_HIP;
    }
    return hipGetLastError() == hipSuccess ? EGG_SYNC_OK : EGG_SYNC_E_HIP;
}

// x <- clamp(x + scale * y, 0, 1) for n x d row-major positions.
__global__ void k_jump(float *__restrict__ X, const float *__restrict__ Y, long long count,
                       float scale, float clip) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
         e += (long long)gridDim.x * blockDim.x) {
        float y = Y[e];
        if (clip > 0.f) y = fminf(fmaxf(y, -clip), clip);
        X[e] = fminf(fmaxf(X[e] + scale * y, 0.f), 1.f);
    }
}

int egg_sync_jump_device(float *d_X, const float *d_Y, int64_t n, int32_t d, float scale,
                         float clip, void *stream) {
    if ((!d_X || !d_Y) && n > 0) return EGG_SYNC_E_ARG;
    if (n == 0) return EGG_SYNC_OK;
    k_jump<<<blocks_for(n * d), kThreads, 0, (hipStream_t)stream>>>(d_X, d_Y, n * (long long)d,
                                                                     scale, clip);
    return hipGetLastError() == hipSuccess ? EGG_SYNC_OK : EGG_SYNC_E_HIP;
}

}  // extern "C"
