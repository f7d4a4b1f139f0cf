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

// ---- wave-level BF16 MFMA inference ----------------------------------------------------------
// A wave evaluates its 64 lanes' samples together as three small GEMMs on
// v_mfma_f32_16x16x32_bf16, with the sample index as the GEMM's N dimension (four 16-sample
// tiles, one at a time) and the hidden units as M. Weights are BF16 (round to nearest even);
// every layer input is carried as an unevaluated sum of two BF16 values (hi + lo), each product
// exact in FP32, so the activations keep about 16 significant bits beyond BF16 and the result
// matches FP32 arithmetic on the BF16 weights to FP32 rounding of the sums.
//
// Operand layouts of the 16x16x32 MFMA (lane l, group g = l / 16, r = l % 16): A holds row r,
// k = 8g..8g+7; B holds column r, k = 8g..8g+7; the result holds column r, rows 4g..4g+3.
// The k order inside a 32-wide chunk is free as long as A and B agree, which lets layer 2 read
// layer 1's results in the register layout they were produced in.
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int mlp_u32x4 __attribute__((ext_vector_type(4)));

// Weights pre-arranged as per-lane MFMA A fragments (built once per weight set by
// mlp_pack_device):
//   a1[mt][l]: hidden unit j = 16 mt + r; groups 0 and 1 hold W1[0..7][j] (group 0 meets the
//              features' hi halves, group 1 their lo halves), group 2 element 0 holds b1[j]
//              (meeting a constant 1), the rest is zero;
//   a2[ot][s][l]: output unit o = 16 ot + r; element p holds W2[j][o] with
//              j = 16 (2 s + p / 4) + 4 g + p % 4 (the rows layer 1 leaves in lane group g);
//   b1, b2, w3, b3: BF16-rounded, in natural order (b2 starts layer 2's accumulators).
struct MlpPacked {
    mlp_u32x4 a1[4][64];
    mlp_u32x4 a2[4][2][64];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H], b3;
};

__host__ __device__ __forceinline__ uint32_t bf16_bits(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float bf16_bits_f32(uint32_t b) { return __uint_as_float(b << 16); }

// hi and lo BF16 halves of two values, packed as (first | second << 16)
__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi, uint32_t &lo) {
    const uint32_t h0 = bf16_bits(x0), h1 = bf16_bits(x1);
    hi = h0 | (h1 << 16);
    lo = bf16_bits(x0 - bf16_bits_f32(h0)) | (bf16_bits(x1 - bf16_bits_f32(h1)) << 16);
}

__device__ __forceinline__ mlp_f32x4 mlp_mfma(mlp_u32x4 a, mlp_u32x4 b, mlp_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16x8, a), __builtin_bit_cast(mlp_bf16x8, b),
                                                   c, 0, 0, 0);
}

// Logit of this lane's sample x. Wave-collective: every lane of the wave must call it together
// (lanes without a sample pass any finite x and ignore the result).
__device__ __forceinline__ float mlp_forward_wave(const MlpPacked *__restrict__ P, const float (&x)[MLP_NF]) {
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    uint32_t xh[4], xl[4];
#pragma unroll
    for (int q = 0; q < 4; q++) split2(x[2 * q], x[2 * q + 1], xh[q], xl[q]);
    mlp_u32x4 a1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) a1[t] = P->a1[t][lane];
    float z = 0.0f;
#pragma unroll 1
    for (int vt = 0; vt < 4; vt++) {  // one sample tile at a time keeps the live registers few
        // layer 2's weights are re-read (L1 hits) per tile rather than held in registers: the
        // empty asm hides the pointer from loop-invariant hoisting
        const MlpPacked *Q = P;
        asm volatile("" : "+s"(Q));
        // B operand of layer 1: sample 16 vt + r's features, hi halves in group 0, lo in group 1,
        // and a constant 1 in group 2 (meeting the bias b1 in a1)
        const int src = 16 * vt + r;
        mlp_u32x4 bx;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t fh = (uint32_t)__shfl((int)xh[q], src);
            const uint32_t fl = (uint32_t)__shfl((int)xl[q], src);
            bx[q] = g == 0 ? fh : (g == 1 ? fl : (g == 2 && q == 0 ? 0x3F80u : 0u));
        }
        // layer 1 (bias included) and its ReLU, split into the hi / lo B operands of layer 2
        mlp_u32x4 bh[2], bl[2];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const mlp_f32x4 acc = mlp_mfma(a1[mt], bx, mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            const float h0 = fmaxf(acc[0], 0.0f), h1 = fmaxf(acc[1], 0.0f);
            const float h2 = fmaxf(acc[2], 0.0f), h3 = fmaxf(acc[3], 0.0f);
            uint32_t hi0, lo0, hi1, lo1;
            split2(h0, h1, hi0, lo0);
            split2(h2, h3, hi1, lo1);
            const int s = mt >> 1, d = (mt & 1) * 2;
            bh[s][d] = hi0;
            bh[s][d + 1] = hi1;
            bl[s][d] = lo0;
            bl[s][d + 1] = lo1;
        }
        // layer 2, then this lane group's share of the output layer
        float zp = 0.0f;
#pragma unroll
        for (int ot = 0; ot < 4; ot++) {
            mlp_f32x4 acc = *reinterpret_cast<const mlp_f32x4 *>(Q->b2 + 16 * ot + 4 * g);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const mlp_u32x4 a2 = Q->a2[ot][s][lane];
                acc = mlp_mfma(a2, bh[s], acc);
                acc = mlp_mfma(a2, bl[s], acc);
            }
            const mlp_f32x4 w3 = *reinterpret_cast<const mlp_f32x4 *>(Q->w3 + 16 * ot + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; i++) zp = fmaf(w3[i], fmaxf(acc[i], 0.0f), zp);
        }
        zp += __shfl_xor(zp, 16);
        zp += __shfl_xor(zp, 32);
        if (vt == g) z = zp;  // this lane's own sample is 16 g + r
    }
    return P->b3 + z;
}

// Builds the packed BF16 inference weights from the FP32 parameter array (device pointers).
void mlp_pack_device(const float *w, MlpPacked *out, hipStream_t s);

}  // namespace scgpu
