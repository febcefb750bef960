// Subtree-cost model of the learned schedule: feature transform and one-sample forward
// pass, shared by the host, the training kernels (costmodel.hip) and the scheduling
// kernel (engine.hip). No counterpart in the reference: its tasks are expanded in
// arrival order (src/host_functions.cu:173-285).
//
// Network: 7 -> 64 -> 64 -> 1, ReLU, fp32. Parameters in one flat array:
//   W1[64][7], b1[64], W2[64][64], b2[64], W3[64], b3.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace dqc_cost {

constexpr int IN = 7;
constexpr int H = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * IN;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H;
constexpr int NPARAM = OFF_B3 + 1;  // 4737

// Features of a task: log1p of members, candidates, depth, min / max out-degree and
// min / max in-degree inside the task.
__host__ __device__ inline void features(int32_t nm, int32_t tot, int32_t depth, uint32_t dmin_o, uint32_t dmax_o,
                                         uint32_t dmin_i, uint32_t dmax_i, float* x) {
    x[0] = log1pf((float)nm);
    x[1] = log1pf((float)(tot - nm));
    x[2] = log1pf((float)depth);
    x[3] = log1pf((float)dmin_o);
    x[4] = log1pf((float)dmax_o);
    x[5] = log1pf((float)dmin_i);
    x[6] = log1pf((float)dmax_i);
}

// Forward pass of one sample; the first hidden layer is kept in registers and the
// second is folded into the output sum.
__host__ __device__ inline float forward(const float* p, const float* x) {
    float h1[H];
    for (int j = 0; j < H; ++j) {
        float z = p[OFF_B1 + j];
        for (int f = 0; f < IN; ++f) z += p[OFF_W1 + j * IN + f] * x[f];
        h1[j] = z > 0.f ? z : 0.f;
    }
    float y = p[OFF_B3];
    for (int j = 0; j < H; ++j) {
        float z = p[OFF_B2 + j];
        for (int i = 0; i < H; ++i) z += p[OFF_W2 + j * H + i] * h1[i];
        if (z > 0.f) y += p[OFF_W3 + j] * z;
    }
    return y;
}

}  // namespace dqc_cost
