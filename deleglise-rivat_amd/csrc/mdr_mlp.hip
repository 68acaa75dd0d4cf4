// Training and inference kernels of the quotient-range cost model
// (5 -> 32 -> 32 -> 1 FP32 MLP, ReLU; see mdr_mlp.hpp).
//
// Loss: mean squared error of log2(cycles per position) over a batch of n
// samples.  The backward pass is split in two deterministic kernels: one
// thread per sample stores the activations and the back-propagated errors
// (x, h1, h2, g, d2, d1), then one thread per parameter sums its gradient
// over the batch in sample order -- no atomics, so a gradient is a fixed
// function of the batch.
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

__global__ void k_mlp_fwd(const float* w, const float* X, const float* T, const int* idx, i64 n, float* buf,
                          float* err) {
    const i64 b = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const i64 s = idx ? idx[b] : b;
    float x[MLP_IN], z1[MLP_H], h1[MLP_H], z2[MLP_H];
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[s * MLP_IN + k];
    for (int i = 0; i < MLP_H; ++i) {
        float acc = w[MLP_B1 + i];
        for (int k = 0; k < MLP_IN; ++k) acc += w[MLP_W1 + i * MLP_IN + k] * x[k];
        z1[i] = acc;
        h1[i] = acc > 0.f ? acc : 0.f;
    }
    float out = w[MLP_B3];
    for (int i = 0; i < MLP_H; ++i) {
        float acc = w[MLP_B2 + i];
        for (int k = 0; k < MLP_H; ++k) acc += w[MLP_W2 + i * MLP_H + k] * h1[k];
        z2[i] = acc;
        out += w[MLP_W3 + i] * (acc > 0.f ? acc : 0.f);
    }
    const float diff = out - T[s];
    const float g = 2.f * diff / (float)n;
    float* r = buf + b * ROW;
    for (int k = 0; k < MLP_IN; ++k) r[R_X + k] = x[k];
    for (int i = 0; i < MLP_H; ++i) {
        r[R_H1 + i] = h1[i];
        r[R_H2 + i] = z2[i] > 0.f ? z2[i] : 0.f;
    }
    r[R_G] = g;
    float d2[MLP_H];
    for (int i = 0; i < MLP_H; ++i) {
        d2[i] = z2[i] > 0.f ? g * w[MLP_W3 + i] : 0.f;
        r[R_D2 + i] = d2[i];
    }
    for (int k = 0; k < MLP_H; ++k) {
        float acc = 0.f;
        for (int i = 0; i < MLP_H; ++i) acc += d2[i] * w[MLP_W2 + i * MLP_H + k];
        r[R_D1 + k] = z1[k] > 0.f ? acc : 0.f;
    }
    if (err) err[b] = diff * diff;
}

__global__ void k_mlp_reduce(const float* buf, i64 n, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    int a, c;   // gradient = sum_b row[a] * (c < 0 ? 1 : row[c])
    if (p < MLP_B1) { a = R_D1 + p / MLP_IN; c = R_X + p % MLP_IN; }
    else if (p < MLP_W2) { a = R_D1 + (p - MLP_B1); c = -1; }
    else if (p < MLP_B2) { a = R_D2 + (p - MLP_W2) / MLP_H; c = R_H1 + (p - MLP_W2) % MLP_H; }
    else if (p < MLP_W3) { a = R_D2 + (p - MLP_B2); c = -1; }
    else if (p < MLP_B3) { a = R_G; c = R_H2 + (p - MLP_W3); }
    else { a = R_G; c = -1; }
    float acc = 0.f;
    for (i64 b = 0; b < n; ++b) {
        const float* r = buf + b * ROW;
        acc += r[a] * (c < 0 ? 1.f : r[c]);
    }
    grad[p] = acc;
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

__global__ void k_mlp_adam(float* w, const float* g, float* m, float* v, float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
}

hipError_t mlp_predict(const float* w, const float* X, i64 n, float* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, X, n, out);
    return hipGetLastError();
}

hipError_t mlp_grad(const float* w, const float* X, const float* T, const int* idx, i64 n, float* buf, float* err,
                    float* grad, float* loss, hipStream_t st) {
    if (n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, X, T, idx, n, buf, err);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, buf, n, grad);
    if (loss) hipLaunchKernelGGL(k_mlp_sum, dim3(1), dim3(256), 0, st, err, n, 1.f / (float)n, loss);
    return hipGetLastError();
}

hipError_t mlp_adam(float* w, const float* g, float* m, float* v, float lr, int step, hipStream_t st) {
    const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, w, g, m, v, lr, c1, c2);
    return hipGetLastError();
}

i64 mlp_row_floats() { return ROW; }

}  // namespace mdr
