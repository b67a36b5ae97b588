// Kernels of the learned gear controller: a small dense network
// (16 -> 64 -> 64 -> 4) with FP32 weights and BF16-stored activations that
// predicts the output SNR of each sparsity gear from ciphertext statistics.
// It never touches the modular arithmetic; the plain noise monitor
// (rtl/m4/noise_monitor.sv, verify_m4h2_prod_calibration.py:64-82) stays the
// fallback and the arbiter of the predictions.
//
// Layouts: row-major, activations [B][features], weights [out][in].
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// BF16 storage: round-to-nearest-even from FP32, widen back exactly.
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float load_act(const void* p, size_t i, int bf16) {
    return bf16 ? bf2f(((const uint16_t*)p)[i]) : ((const float*)p)[i];
}
__device__ __forceinline__ void store_act(void* p, size_t i, float v, int bf16) {
    if (bf16)
        ((uint16_t*)p)[i] = f2bf(v);
    else
        ((float*)p)[i] = v;
}

// out[b][j] = act(sum_i in[b][i] W[j][i] + bias[j]); one thread per (b, j).
__global__ void k_dense_fwd(const void* __restrict__ in, int in_bf16, const float* __restrict__ W,
                            const float* __restrict__ bias, void* __restrict__ out, int out_bf16, int B, int I, int J,
                            int relu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * J) return;
    const int b = t / J, j = t % J;
    float acc = bias ? bias[j] : 0.0f;
    for (int i = 0; i < I; ++i) acc = fmaf(load_act(in, (size_t)b * I + i, in_bf16), W[(size_t)j * I + i], acc);
    if (relu) acc = acc > 0.0f ? acc : 0.0f;
    store_act(out, (size_t)b * J + j, acc, out_bf16);
}

// dpre[b][j] = dout[b][j] * (relu ? [out[b][j] > 0] : 1)
__global__ void k_dense_act_bwd(const float* __restrict__ dout, const void* __restrict__ out, int out_bf16,
                                float* __restrict__ dpre, int n, int relu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float g = dout[t];
    if (relu && !(load_act(out, t, out_bf16) > 0.0f)) g = 0.0f;
    dpre[t] = g;
}

// dW[j][i] = sum_b dpre[b][j] in[b][i],  db[j] = sum_b dpre[b][j]; one thread per (j, i)
// (threads with i == I compute db).
__global__ void k_dense_wgrad(const float* __restrict__ dpre, const void* __restrict__ in, int in_bf16,
                              float* __restrict__ dW, float* __restrict__ db, int B, int I, int J) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= J * (I + 1)) return;
    const int j = t / (I + 1), i = t % (I + 1);
    float acc = 0.0f;
    if (i < I) {
        for (int b = 0; b < B; ++b) acc = fmaf(dpre[(size_t)b * J + j], load_act(in, (size_t)b * I + i, in_bf16), acc);
        dW[(size_t)j * I + i] = acc;
    } else {
        for (int b = 0; b < B; ++b) acc += dpre[(size_t)b * J + j];
        db[j] = acc;
    }
}

// din[b][i] = sum_j dpre[b][j] W[j][i]
__global__ void k_dense_igrad(const float* __restrict__ dpre, const float* __restrict__ W, float* __restrict__ din,
                              int B, int I, int J) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * I) return;
    const int b = t / I, i = t % I;
    float acc = 0.0f;
    for (int j = 0; j < J; ++j) acc = fmaf(dpre[(size_t)b * J + j], W[(size_t)j * I + i], acc);
    din[t] = acc;
}

// Adam (Kingma & Ba): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// p -= lr * mhat / (sqrt(vhat) + eps), with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.0f - b1) * gi;
    const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Mean-squared error gradient: dy = 2 (y - target) / n_rows, loss partials per row.
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ target, float* __restrict__ dy,
                           float* __restrict__ row_loss, int B, int J) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float s = 0.0f;
    for (int j = 0; j < J; ++j) {
        const float d = y[(size_t)b * J + j] - target[(size_t)b * J + j];
        dy[(size_t)b * J + j] = 2.0f * d / (float)B;
        s += d * d;
    }
    row_loss[b] = s;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int ckks_gear_dense_fwd(const void* in, int in_bf16, const float* W, const float* bias, void* out, int out_bf16,
                        int B, int I, int J, int relu, void* stream) {
    if (B < 0 || I < 1 || J < 1) return -4;
    if (!B) return 0;
    if (!in || !W || !out) return -6;
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((size_t)B * J)), dim3(256), 0, (hipStream_t)stream, in, in_bf16, W,
                       bias, out, out_bf16, B, I, J, relu);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int ckks_gear_dense_bwd(const float* dout, const void* out, int out_bf16, const void* in, int in_bf16, const float* W,
                        float* dpre, float* dW, float* db, float* din, int B, int I, int J, int relu, void* stream) {
    if (B < 1 || I < 1 || J < 1) return -4;
    if (!dout || !out || !in || !W || !dpre || !dW || !db) return -6;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dense_act_bwd, dim3(blocks((size_t)B * J)), dim3(256), 0, s, dout, out, out_bf16, dpre,
                       B * J, relu);
    hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((size_t)J * (I + 1))), dim3(256), 0, s, dpre, in, in_bf16, dW, db,
                       B, I, J);
    if (din)
        hipLaunchKernelGGL(k_dense_igrad, dim3(blocks((size_t)B * I)), dim3(256), 0, s, dpre, W, din, B, I, J);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int ckks_gear_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                   int step, void* stream) {
    if (n < 0 || step < 1) return -4;
    if (!n) return 0;
    if (!p || !g || !m || !v) return -6;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, c1,
                       c2);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int ckks_gear_mse_grad(const float* y, const float* target, float* dy, float* row_loss, int B, int J, void* stream) {
    if (B < 0 || J < 1) return -4;
    if (!B) return 0;
    if (!y || !target || !dy || !row_loss) return -6;
    hipLaunchKernelGGL(k_mse_grad, dim3(blocks(B)), dim3(256), 0, (hipStream_t)stream, y, target, dy, row_loss, B, J);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

}  // extern "C"
