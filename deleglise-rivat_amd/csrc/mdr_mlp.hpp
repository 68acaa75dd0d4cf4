// Learned cost model of quotient ranges: a 5 -> 32 -> 32 -> 1 FP32 MLP with
// ReLU hidden layers.  It predicts log2(cycles per position) of one "cost
// piece" (a run of S1 or S2 positions whose n or j span at most a factor 4)
// from
//   f0 = log2(y)/64            y: argument of the G evaluation
//   f1 = log2(start)/64        first n (S1) or j (S2) of the piece
//   f2 = log2(end)/64          last n or j
//   f3 = kind/16               0..3 S1 {dense,compressed} x {stride 1,2}, 4..15 S2 pieces 1..12
//   f4 = log2(y/start)/64      quotient bucket: the largest quotient of the piece
// A piece's predicted cost is count * 2^output cycles; the partition cuts the
// global position space at equal predicted cost.
#ifndef MDR_MLP_HPP
#define MDR_MLP_HPP

#include "mdr_device.hpp"

namespace mdr {

constexpr int MLP_IN = 5;
constexpr int MLP_H = 32;
// parameter layout (floats)
constexpr int MLP_W1 = 0;                       // [H][IN]
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;  // [H]
constexpr int MLP_W2 = MLP_B1 + MLP_H;           // [H][H]
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;   // [H]
constexpr int MLP_W3 = MLP_B2 + MLP_H;           // [H]
constexpr int MLP_B3 = MLP_W3 + MLP_H;           // [1]
constexpr int MLP_NPARAM = MLP_B3 + 1;           // 1281

__host__ __device__ __forceinline__ float mlp_forward(const float* w, const float* x) {
    float h1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
        float s = w[MLP_B1 + i];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) s += w[MLP_W1 + i * MLP_IN + k] * x[k];
        h1[i] = s > 0.f ? s : 0.f;
    }
    float out = w[MLP_B3];
#pragma unroll 4
    for (int i = 0; i < MLP_H; ++i) {
        float s = w[MLP_B2 + i];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) s += w[MLP_W2 + i * MLP_H + k] * h1[k];
        out += w[MLP_W3 + i] * (s > 0.f ? s : 0.f);
    }
    return out;
}

__host__ __device__ __forceinline__ int seg_kind(const Seg& s) {
    if (s.type == SEG_S1_DENSE) return s.stride == 1 ? 0 : 1;
    if (s.type == SEG_S1_COMP) return s.stride == 1 ? 2 : 3;
    return 3 + s.piece;
}

__host__ __device__ __forceinline__ void piece_features(float* f, i64 y, i64 start, i64 end, int kind) {
    const float ly = log2f((float)y);
    const float ls = log2f((float)start);
    f[0] = ly * (1.f / 64);
    f[1] = ls * (1.f / 64);
    f[2] = log2f((float)end) * (1.f / 64);
    f[3] = (float)kind * (1.f / 16);
    f[4] = (ly - ls) * (1.f / 64);
}

// Value (n for S1, j for S2) at position p of a segment.
__host__ __device__ __forceinline__ i64 seg_value(const Seg& s, i64 p) {
    return s.type == SEG_S2 ? j_of_t(s.first + p) : s.first + s.stride * p;
}

// End of the cost piece that starts at position p: the first position whose
// value exceeds four times the value at p (or the segment's end).
__host__ __device__ __forceinline__ i64 piece_end(const Seg& s, i64 p) {
    const i64 v4 = 4 * seg_value(s, p);
    i64 e;
    if (s.type == SEG_S2) e = count_coprime6(v4) - s.first;
    else e = (v4 - s.first) / s.stride + 1;
    return e < s.count ? e : s.count;
}

// Predicted cycles of positions [p, pe) of a segment (one cost piece).
__device__ __forceinline__ double piece_cost(const float* w, i64 y, const Seg& s, i64 p, i64 pe) {
    float f[MLP_IN];
    piece_features(f, y, seg_value(s, p), seg_value(s, pe - 1), seg_kind(s));
    float lg = mlp_forward(w, f);
    lg = fminf(fmaxf(lg, -8.f), 24.f);   // finite (a NaN becomes -8) and bounded
    return (double)(pe - p) * (double)exp2f(lg);
}

}  // namespace mdr

#endif
