// mlp.hip -- training and evaluation kernels of the learned active-vertex filter (no
// counterpart in the reference). Loss: weighted binary cross-entropy on the logit,
//   L = mean_i [ pw * y_i * softplus(-z_i) + (1 - y_i) * softplus(z_i) ],
// gradients by backpropagation; the weight gradients are reductions over the samples in a
// fixed order (one thread per parameter), so training is deterministic. Adam updates.
#include "common.hpp"
#include "mlp.hpp"

#include <cmath>
#include <random>
#include <vector>

namespace scgpu {

__device__ __forceinline__ float softplus(float z) { return z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z)); }
__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

// One thread per sample of the batch: forward, loss, and the per-sample activations and
// back-propagated signals the weight gradients need.
__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(const float *w, const float *X, const uint8_t *y,
                                                     const uint32_t *idx, uint32_t B, float pw, float *H1, float *H2,
                                                     float *D1, float *D2, float *DZ, float *loss) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint32_t s = idx ? idx[b] : b;
    float x[MLP_NF];
#pragma unroll
    for (int f = 0; f < MLP_NF; f++) x[f] = X[(uint64_t)s * MLP_NF + f];
    float h1[MLP_H], h2[MLP_H];
    const float z = mlp_forward(w, x, h1, h2);
    const float t = y[s] ? 1.0f : 0.0f;
    loss[b] = pw * t * softplus(-z) + (1.0f - t) * softplus(z);
    const float dz = (1.0f - t) * sigmoidf(z) - pw * t * sigmoidf(-z);
    DZ[b] = dz;
    for (int j = 0; j < MLP_H; j++) {
        H1[(uint64_t)b * MLP_H + j] = h1[j];
        H2[(uint64_t)b * MLP_H + j] = h2[j];
    }
    float d2[MLP_H];
    for (int j = 0; j < MLP_H; j++) {
        d2[j] = h2[j] > 0.0f ? dz * w[MLP_OFF_W3 + j] : 0.0f;
        D2[(uint64_t)b * MLP_H + j] = d2[j];
    }
    for (int i = 0; i < MLP_H; i++) {
        float a = 0.0f;
        for (int j = 0; j < MLP_H; j++) a = fmaf(w[MLP_OFF_W2 + i * MLP_H + j], d2[j], a);
        D1[(uint64_t)b * MLP_H + i] = h1[i] > 0.0f ? a : 0.0f;
    }
}

// One thread per parameter: its gradient, summed over the batch in sample order, / B.
__global__ void k_mlp_grad(const float *X, const uint32_t *idx, uint32_t B, const float *H1, const float *H2,
                           const float *D1, const float *D2, const float *DZ, float *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    double acc = 0.0;
    if (p < MLP_OFF_B1) {  // W1[f][i] = sum x_f d1_i
        const int f = p / MLP_H, i = p % MLP_H;
        for (uint32_t b = 0; b < B; b++) {
            const uint32_t s = idx ? idx[b] : b;
            acc += (double)X[(uint64_t)s * MLP_NF + f] * D1[(uint64_t)b * MLP_H + i];
        }
    } else if (p < MLP_OFF_W2) {
        const int i = p - MLP_OFF_B1;
        for (uint32_t b = 0; b < B; b++) acc += D1[(uint64_t)b * MLP_H + i];
    } else if (p < MLP_OFF_B2) {  // W2[i][j] = sum h1_i d2_j
        const int q = p - MLP_OFF_W2, i = q / MLP_H, j = q % MLP_H;
        for (uint32_t b = 0; b < B; b++) acc += (double)H1[(uint64_t)b * MLP_H + i] * D2[(uint64_t)b * MLP_H + j];
    } else if (p < MLP_OFF_W3) {
        const int j = p - MLP_OFF_B2;
        for (uint32_t b = 0; b < B; b++) acc += D2[(uint64_t)b * MLP_H + j];
    } else if (p < MLP_OFF_B3) {  // w3[j] = sum h2_j dz
        const int j = p - MLP_OFF_W3;
        for (uint32_t b = 0; b < B; b++) acc += (double)H2[(uint64_t)b * MLP_H + j] * DZ[b];
    } else {
        for (uint32_t b = 0; b < B; b++) acc += DZ[b];
    }
    grad[p] = (float)(acc / (double)B);
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    m[p] = b1 * m[p] + (1.0f - b1) * g[p];
    v[p] = b2 * v[p] + (1.0f - b2) * g[p] * g[p];
    w[p] -= lr * (m[p] / c1) / (sqrtf(v[p] / c2) + eps);
}

// FP32 weights, one sample per thread (the training arithmetic)
__global__ void k_mlp_predict(const float *w, const float *X, uint64_t n, float *out) {
    __shared__ float sw[MLP_PARAMS];
    for (int p = threadIdx.x; p < MLP_PARAMS; p += blockDim.x) sw[p] = w[p];
    __syncthreads();
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n; s += (uint64_t)gridDim.x * blockDim.x) {
        float x[MLP_NF];
#pragma unroll
        for (int f = 0; f < MLP_NF; f++) x[f] = X[s * MLP_NF + f];
        out[s] = mlp_forward(sw, x);
    }
}

// BF16 inference: the wave MFMA path the filtered sweep runs (loop bounds are block-uniform,
// so every lane of a wave reaches mlp_forward_wave together)
__global__ void k_mlp_predict_bf16(const MlpPacked *P, const float *X, uint64_t n, float *out) {
    for (uint64_t s0 = blockIdx.x * (uint64_t)blockDim.x; s0 < n; s0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = s0 + threadIdx.x;
        float x[MLP_NF];
#pragma unroll
        for (int f = 0; f < MLP_NF; f++) x[f] = s < n ? X[s * MLP_NF + f] : 0.0f;
        const float z = mlp_forward_wave(P, x);
        if (s < n) out[s] = z;
    }
}

// One thread per packed element (see MlpPacked)
__global__ void k_mlp_pack(const float *w, MlpPacked *P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 4 * 64) {  // a1[mt][lane]
        const int mt = t >> 6, lane = t & 63, g = lane >> 4, j = 16 * mt + (lane & 15);
        mlp_u32x4 v = {0u, 0u, 0u, 0u};
        if (g < 2)
            for (int q = 0; q < 4; q++)
                v[q] = bf16_bits(w[(2 * q) * MLP_H + j]) | (bf16_bits(w[(2 * q + 1) * MLP_H + j]) << 16);
        else if (g == 2)
            v[0] = bf16_bits(w[MLP_OFF_B1 + j]);
        P->a1[mt][lane] = v;
    } else if (t < 4 * 64 + 8 * 64) {  // a2[ot][s][lane]
        const int u = t - 4 * 64, ot = u >> 7, s = (u >> 6) & 1, lane = u & 63, g = lane >> 4;
        const int o = 16 * ot + (lane & 15);
        mlp_u32x4 v;
        for (int q = 0; q < 4; q++) {
            uint32_t half[2];
            for (int e = 0; e < 2; e++) {
                const int p = 2 * q + e;
                const int j = 16 * (2 * s + (p >> 2)) + 4 * g + (p & 3);
                half[e] = bf16_bits(w[MLP_OFF_W2 + j * MLP_H + o]);
            }
            v[q] = half[0] | (half[1] << 16);
        }
        P->a2[ot][s][lane] = v;
    } else if (t < 4 * 64 + 8 * 64 + MLP_H) {
        const int i = t - 12 * 64;
        P->b1[i] = bf16_round(w[MLP_OFF_B1 + i]);
        P->b2[i] = bf16_round(w[MLP_OFF_B2 + i]);
        P->w3[i] = bf16_round(w[MLP_OFF_W3 + i]);
        if (i == 0) P->b3 = bf16_round(w[MLP_OFF_B3]);
    }
}

void mlp_pack_device(const float *w, MlpPacked *out, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((12 * 64 + MLP_H + 255) / 256), dim3(256), 0, s, w, out);
    KERNEL_OK();
}

// loss and gradient of one batch (all n samples) at weights w
static double batch_grad(const float *d_w, const float *d_X, const uint8_t *d_y, const uint32_t *d_idx, uint32_t B,
                         float pw, float *d_grad, hipStream_t s) {
    DevBuf<float> H1((size_t)B * MLP_H), H2((size_t)B * MLP_H), D1((size_t)B * MLP_H), D2((size_t)B * MLP_H);
    DevBuf<float> DZ(B), L(B);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((B + 255) / 256), dim3(256), 0, s, d_w, d_X, d_y, d_idx, B, pw, H1.get(),
                       H2.get(), D1.get(), D2.get(), DZ.get(), L.get());
    KERNEL_OK();
    hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, d_X, d_idx, B,
                       (const float *)H1.get(), (const float *)H2.get(), (const float *)D1.get(),
                       (const float *)D2.get(), (const float *)DZ.get(), d_grad);
    KERNEL_OK();
    std::vector<float> h(B);
    to_host(h.data(), L.get(), B, s);
    sync_stream(s);
    double tot = 0.0;
    for (uint32_t b = 0; b < B; b++) tot += h[b];
    return tot / (double)B;
}

}  // namespace scgpu

using namespace scgpu;

extern "C" {

uint32_t sc_mlp_param_count(void) { return MLP_PARAMS; }

// Deterministic initialisation: He-uniform weights, zero biases.
void sc_mlp_init(uint64_t seed, float *w) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int p = 0; p < MLP_PARAMS; p++) w[p] = 0.0f;
    const double a1 = std::sqrt(6.0 / MLP_NF), a2 = std::sqrt(6.0 / MLP_H);
    for (int p = 0; p < MLP_OFF_B1; p++) w[p] = (float)(a1 * u(rng));
    for (int p = MLP_OFF_W2; p < MLP_OFF_B2; p++) w[p] = (float)(a2 * u(rng));
    for (int p = MLP_OFF_W3; p < MLP_OFF_B3; p++) w[p] = (float)(a2 * u(rng));
}

sc_status sc_mlp_grad(const float *w, const float *X, const uint8_t *y, uint64_t n, float pos_weight, float *grad_out,
                      double *loss_out, sc_error *err) {
    try {
        if (!w || !X || !y || !grad_out || !loss_out || n == 0) throw ScError(SC_INVALID_ARGUMENT, "bad argument");
        if (n > (1ull << 24)) throw ScError(SC_INVALID_ARGUMENT, "at most 2^24 samples per gradient");
        hipStream_t s = thread_stream();
        DevBuf<float> dw(MLP_PARAMS), dX(n * MLP_NF), dg(MLP_PARAMS);
        DevBuf<uint8_t> dy(n);
        to_device(dw.get(), w, MLP_PARAMS, s);
        to_device(dX.get(), X, n * MLP_NF, s);
        to_device(dy.get(), y, n, s);
        *loss_out = batch_grad(dw.get(), dX.get(), dy.get(), nullptr, (uint32_t)n, pos_weight, dg.get(), s);
        to_host(grad_out, dg.get(), MLP_PARAMS, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

// Minibatch Adam on (X, y): epochs over a seeded shuffle; loss_hist[e] = mean batch loss.
sc_status sc_mlp_train(const float *X, const uint8_t *y, uint64_t n, uint32_t epochs, uint32_t batch, float lr,
                       float pos_weight, uint64_t seed, float *w_inout, double *loss_hist, sc_error *err) {
    try {
        if (!X || !y || !w_inout || n == 0 || batch == 0 || epochs == 0) throw ScError(SC_INVALID_ARGUMENT, "bad argument");
        hipStream_t s = thread_stream();
        DevBuf<float> dw(MLP_PARAMS), dX(n * MLP_NF), dg(MLP_PARAMS), dm(MLP_PARAMS), dv(MLP_PARAMS);
        DevBuf<uint8_t> dy(n);
        DevBuf<uint32_t> didx(n);
        to_device(dw.get(), w_inout, MLP_PARAMS, s);
        to_device(dX.get(), X, n * MLP_NF, s);
        to_device(dy.get(), y, n, s);
        HIP_OK(hipMemsetAsync(dm.get(), 0, MLP_PARAMS * sizeof(float), s));
        HIP_OK(hipMemsetAsync(dv.get(), 0, MLP_PARAMS * sizeof(float), s));
        std::vector<uint32_t> order(n);
        for (uint64_t i = 0; i < n; i++) order[i] = (uint32_t)i;
        std::mt19937_64 rng(seed);
        uint64_t t = 0;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (uint32_t e = 0; e < epochs; e++) {
            for (uint64_t i = n - 1; i > 0; i--) std::swap(order[i], order[rng() % (i + 1)]);
            to_device(didx.get(), order.data(), n, s);
            double lsum = 0.0;
            uint64_t nb = 0;
            for (uint64_t b0 = 0; b0 < n; b0 += batch) {
                const uint32_t B = (uint32_t)std::min<uint64_t>(batch, n - b0);
                lsum += batch_grad(dw.get(), dX.get(), dy.get(), didx.get() + b0, B, pos_weight, dg.get(), s);
                nb++;
                t++;
                const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
                hipLaunchKernelGGL(k_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, dw.get(),
                                   (const float *)dg.get(), dm.get(), dv.get(), lr, b1, b2, eps, c1, c2);
                KERNEL_OK();
            }
            if (loss_hist) loss_hist[e] = lsum / (double)nb;
        }
        to_host(w_inout, dw.get(), MLP_PARAMS, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

// Logits of n samples; bf16 != 0 evaluates with BF16-rounded weights (as the filter does).
sc_status sc_mlp_predict(const float *w, const float *X, uint64_t n, int32_t bf16, float *out, sc_error *err) {
    try {
        if (!w || !X || !out) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (n == 0) return SC_OK;
        hipStream_t s = thread_stream();
        DevBuf<float> dw(MLP_PARAMS), dX(n * MLP_NF), dout(n);
        to_device(dw.get(), w, MLP_PARAMS, s);
        to_device(dX.get(), X, n * MLP_NF, s);
        if (bf16) {
            DevBuf<MlpPacked> pk(1);
            mlp_pack_device(dw.get(), pk.get(), s);
            hipLaunchKernelGGL(k_mlp_predict_bf16, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s,
                               (const MlpPacked *)pk.get(), (const float *)dX.get(), n, dout.get());
            KERNEL_OK();
            to_host(out, dout.get(), n, s);
            sync_stream(s);
            return SC_OK;
        }
        hipLaunchKernelGGL(k_mlp_predict, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const float *)dw.get(),
                           (const float *)dX.get(), n, dout.get());
        KERNEL_OK();
        to_host(out, dout.get(), n, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

}
