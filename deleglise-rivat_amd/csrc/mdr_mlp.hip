// Training and inference kernels of the quotient-range cost model
// (5 -> 32 -> 32 -> 1 FP32 MLP, ReLU; see mdr_mlp.hpp).
//
// Loss: mean squared error of log2(cycles per position) over a batch of n
// samples.  The backward pass is split in two deterministic kernels: 32 lanes
// per sample store the activations and the back-propagated errors (x, h1, h2,
// g, d2, d1) field-major -- field f of sample b at f * n + b, so the reduction
// reads consecutive samples with consecutive lanes -- then one
// workgroup per parameter sums its gradient over the batch: thread t adds
// samples t, t + 256, ... in order and a fixed LDS tree combines the 256
// partial sums.  No atomics, so a gradient is a fixed function of the batch.
// A training step applies Adam in the same workgroup that finished the
// parameter's gradient (two launches per step).
#include <hip/hip_runtime.h>
#include "mdr_mlp.hpp"
#include "mdr_kernels.hpp"

namespace mdr {

constexpr int ROW = MLP_IN + 3 * MLP_H + 1 + MLP_H;   // x, h1, h2, g, d2, d1 = 134 floats
constexpr int R_X = 0, R_H1 = MLP_IN, R_H2 = R_H1 + MLP_H, R_G = R_H2 + MLP_H, R_D2 = R_G + 1,
              R_D1 = R_D2 + MLP_H;

__global__ void k_mlp_predict(const float* w, const float* X, i64 n, float* out) {
    const i64 b = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) out[b] = mlp_forward(w, X + b * MLP_IN);
}

// Forward and backward pass of one training batch, hidden units on lanes: 32
// lanes per sample (lane i owns hidden unit i of both layers, with row i and
// column i of W2 in registers), FWD_S samples per workgroup.  The layers
// exchange h1 and d2 through LDS; the output is a 32-lane shuffle reduction.
// Writes fields x, h1, h2, g, d2, d1 of sample b at field * n + b.
constexpr int FWD_T = 256;
constexpr int FWD_S = FWD_T / MLP_H;   // 8 samples per workgroup

__global__ __launch_bounds__(FWD_T) void k_mlp_fwd(const float* w, const float* X, const float* T, const int* idx,
                                                   i64 n, float* buf, float* err) {
    __shared__ float sh1[FWD_S][MLP_H];
    __shared__ float sd2[FWD_S][MLP_H];
    const int i = threadIdx.x % MLP_H, j = threadIdx.x / MLP_H;
    const i64 b = (i64)blockIdx.x * FWD_S + j;
    const bool valid = b < n;                 // no early return: the workgroup meets at barriers
    const i64 s = valid ? (idx ? idx[b] : b) : 0;
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[s * MLP_IN + k];
    float row[MLP_H], col[MLP_H];             // W2[i][*], W2[*][i]
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) {
        row[k] = w[MLP_W2 + i * MLP_H + k];
        col[k] = w[MLP_W2 + k * MLP_H + i];
    }
    float z1 = w[MLP_B1 + i];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z1 += w[MLP_W1 + i * MLP_IN + k] * x[k];
    const float h1 = z1 > 0.f ? z1 : 0.f;
    sh1[j][i] = h1;
    __syncthreads();
    float z2 = w[MLP_B2 + i];
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) z2 += row[k] * sh1[j][k];
    const float h2 = z2 > 0.f ? z2 : 0.f;
    const float w3 = w[MLP_W3 + i];
    float o = w3 * h2;
#pragma unroll
    for (int m = MLP_H / 2; m > 0; m >>= 1) o += __shfl_xor(o, m, MLP_H);
    const float diff = (w[MLP_B3] + o) - (valid ? T[s] : 0.f);
    const float g = 2.f * diff / (float)n;
    const float d2 = z2 > 0.f ? g * w3 : 0.f;
    sd2[j][i] = d2;
    __syncthreads();
    float d1 = 0.f;
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) d1 += sd2[j][k] * col[k];
    d1 = z1 > 0.f ? d1 : 0.f;
    if (!valid) return;
    float* r = buf + b;   // field f at r[f * n]
    if (i < MLP_IN) {
        float xi = x[0];
#pragma unroll
        for (int k = 1; k < MLP_IN; ++k) xi = i == k ? x[k] : xi;
        r[(R_X + i) * n] = xi;
    }
    r[(R_H1 + i) * n] = h1;
    r[(R_H2 + i) * n] = h2;
    r[(R_D2 + i) * n] = d2;
    r[(R_D1 + i) * n] = d1;
    if (i == 0) {
        r[R_G * n] = g;
        if (err) err[b] = diff * diff;
    }
}

constexpr int REDUCE_T = 256;

struct AdamStep {
    float* m;
    float* v;
    float lr, c1, c2;   // c1 = 1 - 0.9^step, c2 = 1 - 0.999^step
};

__device__ inline void adam_update(float* w, int p, float gp, const AdamStep& o) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float mp = b1 * o.m[p] + (1.f - b1) * gp;
    const float vp = b2 * o.v[p] + (1.f - b2) * gp * gp;
    o.m[p] = mp;
    o.v[p] = vp;
    w[p] -= o.lr * (mp / o.c1) / (sqrtf(vp / o.c2) + 1e-8f);
}

// One workgroup per parameter p (grid = MLP_NPARAM).  With w non-null the
// gradient goes straight into an Adam update of w[p] (grad may be null).
__global__ __launch_bounds__(REDUCE_T) void k_mlp_reduce(const float* buf, i64 n, float* grad, float* w,
                                                         AdamStep o) {
    __shared__ float part[REDUCE_T];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    int a, c;   // gradient = sum_b field_a[b] * (c < 0 ? 1 : field_c[b])
    if (p < MLP_B1) { a = R_D1 + p / MLP_IN; c = R_X + p % MLP_IN; }
    else if (p < MLP_W2) { a = R_D1 + (p - MLP_B1); c = -1; }
    else if (p < MLP_B2) { a = R_D2 + (p - MLP_W2) / MLP_H; c = R_H1 + (p - MLP_W2) % MLP_H; }
    else if (p < MLP_W3) { a = R_D2 + (p - MLP_B2); c = -1; }
    else if (p < MLP_B3) { a = R_G; c = R_H2 + (p - MLP_W3); }
    else { a = R_G; c = -1; }
    const float* fa = buf + a * n;
    float acc = 0.f;
    if (c < 0) {
        for (i64 b = t; b < n; b += REDUCE_T) acc += fa[b];
    } else {
        const float* fc = buf + c * n;
        for (i64 b = t; b < n; b += REDUCE_T) acc += fa[b] * fc[b];
    }
    part[t] = acc;
    __syncthreads();
    for (int h = REDUCE_T / 2; h > 0; h >>= 1) {
        if (t < h) part[t] += part[t + h];
        __syncthreads();
    }
    if (t == 0) {
        if (grad) grad[p] = part[0];
        if (w) adam_update(w, p, part[0], o);
    }
}

__global__ void k_mlp_sum(const float* v, i64 n, float scale, float* out) {
    __shared__ float part[256];
    float acc = 0.f;
    for (i64 i = threadIdx.x; i < n; i += blockDim.x) acc += v[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0] * scale;
}

hipError_t mlp_predict(const float* w, const float* X, i64 n, float* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, X, n, out);
    return hipGetLastError();
}

hipError_t mlp_grad(const float* w, const float* X, const float* T, const int* idx, i64 n, float* buf, float* err,
                    float* grad, float* loss, hipStream_t st) {
    if (n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((n + FWD_S - 1) / FWD_S)), dim3(FWD_T), 0, st, w, X, T, idx, n, buf,
                       err);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(MLP_NPARAM), dim3(REDUCE_T), 0, st, buf, n, grad, (float*)nullptr,
                       AdamStep{});
    if (loss) hipLaunchKernelGGL(k_mlp_sum, dim3(1), dim3(256), 0, st, err, n, 1.f / (float)n, loss);
    return hipGetLastError();
}

hipError_t mlp_train_step(float* w, const float* X, const float* T, const int* idx, i64 n, float* buf, float* m,
                          float* v, float lr, int step, hipStream_t st) {
    if (n <= 0 || step < 1) return hipErrorInvalidValue;
    const AdamStep o{m, v, lr, 1.f - powf(0.9f, (float)step), 1.f - powf(0.999f, (float)step)};
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((n + FWD_S - 1) / FWD_S)), dim3(FWD_T), 0, st, w, X, T, idx, n, buf,
                       (float*)nullptr);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(MLP_NPARAM), dim3(REDUCE_T), 0, st, buf, n, (float*)nullptr, w, o);
    return hipGetLastError();
}

i64 mlp_row_floats() { return ROW; }

}  // namespace mdr
