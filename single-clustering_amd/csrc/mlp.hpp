// mlp.hpp -- the learned active-vertex filter: an 8 -> 64 -> 64 -> 1 perceptron (ReLU hidden
// layers, logit output). Parameters are one float array in this order:
//   W1[8][64] (f * 64 + i), b1[64], W2[64][64] (i * 64 + j), b2[64], w3[64], b3.
// Inference reads BF16-rounded weights and accumulates in FP32; training keeps FP32 masters.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace scgpu {

constexpr int MLP_NF = 8;
constexpr int MLP_H = 64;
constexpr int MLP_OFF_B1 = MLP_NF * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_PARAMS = MLP_OFF_B3 + 1;  // 4801

__host__ __device__ __forceinline__ float bf16_round(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;  // round to nearest even (finite inputs)
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}

// Forward pass of one sample; w may point to LDS. Optionally keeps the hidden activations.
__device__ __forceinline__ float mlp_forward(const float *w, const float (&x)[MLP_NF], float *h1_out = nullptr,
                                             float *h2_out = nullptr) {
    float h1[MLP_H];  // fully unrolled loops keep the hidden layer in registers
#pragma unroll
    for (int i = 0; i < MLP_H; i++) {
        float a = w[MLP_OFF_B1 + i];
#pragma unroll
        for (int f = 0; f < MLP_NF; f++) a = fmaf(w[f * MLP_H + i], x[f], a);
        h1[i] = fmaxf(a, 0.0f);
        if (h1_out) h1_out[i] = h1[i];
    }
    float z = w[MLP_OFF_B3];
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_OFF_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) a = fmaf(w[MLP_OFF_W2 + i * MLP_H + j], h1[i], a);
        const float h = fmaxf(a, 0.0f);
        if (h2_out) h2_out[j] = h;
        z = fmaf(w[MLP_OFF_W3 + j], h, z);
    }
    return z;
}

}  // namespace scgpu
