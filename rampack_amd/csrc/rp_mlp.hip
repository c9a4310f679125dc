// rp_mlp.hip -- the network of the learned candidate ordering (no counterpart in the reference).
//
// A multilayer perceptron 12 -> 64 -> 64 -> 64 -> 1 with GELU hidden units (the tanh form) and a sigmoid output
// predicts, from the pose of a candidate relative to a trial pose (position in the trial body frame divided by the
// total interaction range, and the relative rotation matrix), the probability that the exact overlap test finds an
// overlap.  The sweep's exact test stays the judge; the scores only order the candidates.
//
//   * training (FP32): one kernel per layer for the forward pass, the output head with the binary cross-entropy,
//     the backward pass of each layer, the weight gradients as sums over the batch in sample order, and Adam;
//   * inference (BF16 MFMA, FP32 accumulation): one wavefront per 16 pairs, v_mfma_f32_16x16x16bf16 tiles for the
//     three hidden layers, the output as a dot product of the last hidden layer with the head weights.
//
// Parameters, FP32, one array (MLP_PARAMS floats): W1 [64][12], b1 [64], W2 [64][64], b2 [64], W3 [64][64], b3 [64],
// w4 [64], b4.
#include "rp_mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

namespace rp {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ float gelu(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2 / pi)
    const float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.f + t);
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

// out[s][o] = act(b[o] + sum_i W[o][i] in[s][i]); pre keeps the sum (the backward pass needs it).
__global__ __launch_bounds__(256) void k_mlp_fwd(const float *__restrict__ in, int nin, const float *__restrict__ W,
                                                 const float *__restrict__ b, int nout, float *pre, float *act,
                                                 int B) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nout) return;
    const int s = idx / nout, o = idx - s * nout;
    float acc = b[o];
    for (int i = 0; i < nin; i++) acc += W[o * nin + i] * in[s * nin + i];
    pre[idx] = acc;
    act[idx] = gelu(acc);
}

// Output head: z = b4 + w4 . h, p = sigmoid(z), dz = (p - y) / B; the batch's binary cross-entropy summed per block
// in a fixed order into lossPart[block].
__global__ __launch_bounds__(256) void k_mlp_head(const float *__restrict__ h, const float *__restrict__ theta,
                                                  const float *__restrict__ y, int B, float *dz, float *lossPart) {
    __shared__ float sLoss[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (s < B) {
        float z = theta[MLP_OFF_B4];
        for (int j = 0; j < MLP_H; j++) z += theta[MLP_OFF_W4 + j] * h[s * MLP_H + j];
        const float p = 1.f / (1.f + expf(-z));
        dz[s] = (p - y[s]) / (float)B;
        // log(1 + exp(-|z|)) form: y log p + (1 - y) log(1 - p) without cancellation
        l = fmaxf(z, 0.f) - z * y[s] + log1pf(expf(-fabsf(z)));
    }
    sLoss[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sLoss[threadIdx.x] += sLoss[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPart[blockIdx.x] = sLoss[0];
}

// deltaIn[s][i] = (sum_o W[o][i] delta[s][o]) gelu'(pre[s][i]).
__global__ __launch_bounds__(256) void k_mlp_bwd(const float *__restrict__ delta, int nout, const float *__restrict__ W,
                                                 int nin, const float *__restrict__ pre, float *deltaIn, int B) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nin) return;
    const int s = idx / nin, i = idx - s * nin;
    float acc = 0.f;
    for (int o = 0; o < nout; o++) acc += W[o * nin + i] * delta[s * nout + o];
    deltaIn[idx] = acc * gelu_grad(pre[idx]);
}

// gW[o][i] = sum_s delta[s][o] act[s][i] (i < nin), gb[o] = sum_s delta[s][o] (i == nin), in sample order.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float *__restrict__ delta, int nout,
                                                   const float *__restrict__ act, int nin, int B, float *gW,
                                                   float *gb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nout * (nin + 1)) return;
    const int o = idx / (nin + 1), i = idx - o * (nin + 1);
    float acc = 0.f;
    if (i < nin) {
        for (int s = 0; s < B; s++) acc += delta[s * nout + o] * act[s * nin + i];
        gW[o * nin + i] = acc;
    } else {
        for (int s = 0; s < B; s++) acc += delta[s * nout + o];
        gb[o] = acc;
    }
}

// Adam (Kingma & Ba, 2015): m <- b1 m + (1 - b1) g, v <- b2 v + (1 - b2) g^2, theta -= lr m^ / (sqrt(v^) + eps).
__global__ __launch_bounds__(256) void k_adam(float *theta, const float *__restrict__ g, float *m, float *v, int n,
                                              float lr, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m[i] = b1 * m[i] + (1.f - b1) * g[i];
    v[i] = b2 * v[i] + (1.f - b2) * g[i] * g[i];
    theta[i] -= lr * (m[i] / bc1) / (sqrtf(v[i] / bc2) + eps);
}

// Rows of a dataset in the order of `rows` (a permutation) into a contiguous batch.
__global__ __launch_bounds__(256) void k_mlp_gather(const float *__restrict__ x, const float *__restrict__ y,
                                                    const int *__restrict__ rows, int B, float *bx, float *by) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * (MLP_IN + 1)) return;
    const int s = idx / (MLP_IN + 1), f = idx - s * (MLP_IN + 1);
    const int r = rows[s];
    if (f < MLP_IN) bx[s * MLP_IN + f] = x[(size_t)r * MLP_IN + f];
    else by[s] = y[r];
}

// Round to nearest even BF16 (the upper 16 bits of the FP32 pattern); finite inputs.
__device__ __forceinline__ unsigned short bf16_bits(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

// One MFMA layer for a 16-row tile: D[16][64] = A[16][K] B[K][64] with A rows in LDS (bf16, row stride lda) and the
// weights in LDS as bf16 W[n][k] (row stride ldw), so B[k][n] = W[n][k].  v_mfma_f32_16x16x16bf16: lane l supplies
// A[l % 16][4 (l / 16) .. + 3] and B[4 (l / 16) .. + 3][l % 16] and receives D[4 (l / 16) + r][l % 16], r = 0..3.
template <int K>
__device__ __forceinline__ void mfma_layer(const unsigned short *A, int lda, const unsigned short *W, int ldw,
                                           float4v acc[4]) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int row = lane & 15, kq = 4 * (lane >> 4);
    for (int t = 0; t < 4; t++) acc[t] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        short4v a;
        for (int q = 0; q < 4; q++) a[q] = (short)A[row * lda + k0 + kq + q];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            short4v bv;
            const int n = 16 * t + row;
            for (int q = 0; q < 4; q++) bv[q] = (short)W[n * ldw + k0 + kq + q];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
        }
    }
}

constexpr int INFER_WAVES = 4;

// BF16 MFMA inference: one wavefront per tile of 16 pairs (features padded from 12 to 16 with zeros), FP32
// accumulation, biases and GELU in FP32, activations rounded to BF16 between layers.
__global__ __launch_bounds__(WAVE *INFER_WAVES) void k_mlp_infer(const float *__restrict__ x, int P,
                                                                 const float *__restrict__ theta, float *score) {
    __shared__ unsigned short sW1[64 * 16], sW2[64 * 64], sW3[64 * 64];
    __shared__ float sB[3 * 64 + 64 + 1];
    __shared__ unsigned short sA[INFER_WAVES][16 * 64];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    for (int e = tid; e < 64 * 16; e += blockDim.x) {
        const int n = e / 16, k = e - n * 16;
        sW1[e] = k < MLP_IN ? bf16_bits(theta[MLP_OFF_W1 + n * MLP_IN + k]) : 0;
    }
    for (int e = tid; e < 64 * 64; e += blockDim.x) {
        sW2[e] = bf16_bits(theta[MLP_OFF_W2 + e]);
        sW3[e] = bf16_bits(theta[MLP_OFF_W3 + e]);
    }
    for (int e = tid; e < 64; e += blockDim.x) {
        sB[e] = theta[MLP_OFF_B1 + e];
        sB[64 + e] = theta[MLP_OFF_B2 + e];
        sB[128 + e] = theta[MLP_OFF_B3 + e];
        sB[192 + e] = theta[MLP_OFF_W4 + e];
    }
    if (tid == 0) sB[256] = theta[MLP_OFF_B4];
    __syncthreads();
    unsigned short *A = sA[wave];
    const int row = lane & 15, r0 = 4 * (lane >> 4);
    const int tiles = (P + 15) / 16;
    for (int tile = blockIdx.x * INFER_WAVES + wave; tile < tiles; tile += gridDim.x * INFER_WAVES) {
        // input tile: 16 rows x 16 features (12 used), bf16
        for (int e = lane; e < 16 * 16; e += WAVE) {
            const int rr = e / 16, k = e - rr * 16;
            const int p = tile * 16 + rr;
            A[rr * 64 + k] = (k < MLP_IN && p < P) ? bf16_bits(x[(size_t)p * MLP_IN + k]) : 0;
        }
        __builtin_amdgcn_wave_barrier();
        float4v acc[4];
        mfma_layer<16>(A, 64, sW1, 16, acc);
        __builtin_amdgcn_wave_barrier();  // every lane has read the inputs before they are overwritten
        for (int layer = 0; layer < 3; layer++) {
            // bias + GELU of the layer just computed, rounded into the A tile (row r0 + r, column 16 t + row)
            float h[4][4];
            for (int t = 0; t < 4; t++)
                for (int r = 0; r < 4; r++) h[t][r] = gelu(acc[t][r] + sB[64 * layer + 16 * t + row]);
            if (layer == 2) {
                // output: z = b4 + sum_j w4[j] h3[j]; each lane sums its 4 columns, then the 16 lanes of a row group
                float z[4];
                for (int r = 0; r < 4; r++) {
                    float s = 0.f;
                    for (int t = 0; t < 4; t++) s += sB[192 + 16 * t + row] * h[t][r];
                    z[r] = s;
                }
                for (int off = 8; off > 0; off >>= 1)
                    for (int r = 0; r < 4; r++) z[r] += __shfl_xor(z[r], off, 16);
                if (row == 0)
                    for (int r = 0; r < 4; r++) {
                        const int p = tile * 16 + r0 + r;
                        if (p < P) score[p] = 1.f / (1.f + expf(-(z[r] + sB[256])));
                    }
                break;
            }
            for (int t = 0; t < 4; t++)
                for (int r = 0; r < 4; r++) A[(r0 + r) * 64 + 16 * t + row] = bf16_bits(h[t][r]);
            __builtin_amdgcn_wave_barrier();
            mfma_layer<64>(A, 64, layer == 0 ? sW2 : sW3, 64, acc);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------

hipError_t mlp_forward_backward(hipStream_t s, const float *theta, int B, const float *x, const float *y,
                                MlpWork &w, float *grad) {
    const int T = 256;
    auto blocks = [&](int n) { return dim3((unsigned)((n + T - 1) / T)); };
    hipLaunchKernelGGL(k_mlp_fwd, blocks(B * MLP_H), dim3(T), 0, s, x, MLP_IN, theta + MLP_OFF_W1,
                       theta + MLP_OFF_B1, MLP_H, w.pre1, w.a1, B);
    hipLaunchKernelGGL(k_mlp_fwd, blocks(B * MLP_H), dim3(T), 0, s, w.a1, MLP_H, theta + MLP_OFF_W2,
                       theta + MLP_OFF_B2, MLP_H, w.pre2, w.a2, B);
    hipLaunchKernelGGL(k_mlp_fwd, blocks(B * MLP_H), dim3(T), 0, s, w.a2, MLP_H, theta + MLP_OFF_W3,
                       theta + MLP_OFF_B3, MLP_H, w.pre3, w.a3, B);
    hipLaunchKernelGGL(k_mlp_head, blocks(B), dim3(T), 0, s, w.a3, theta, y, B, w.dz, w.lossPart);
    // head: dw4 = sum_s dz a3, db4 = sum_s dz; delta3 = dz w4 gelu'(pre3)
    hipLaunchKernelGGL(k_mlp_wgrad, blocks(MLP_H + 1), dim3(T), 0, s, w.dz, 1, w.a3, MLP_H, B, grad + MLP_OFF_W4,
                       grad + MLP_OFF_B4);
    hipLaunchKernelGGL(k_mlp_bwd, blocks(B * MLP_H), dim3(T), 0, s, w.dz, 1, theta + MLP_OFF_W4, MLP_H, w.pre3,
                       w.d3, B);
    hipLaunchKernelGGL(k_mlp_wgrad, blocks(MLP_H * (MLP_H + 1)), dim3(T), 0, s, w.d3, MLP_H, w.a2, MLP_H, B,
                       grad + MLP_OFF_W3, grad + MLP_OFF_B3);
    hipLaunchKernelGGL(k_mlp_bwd, blocks(B * MLP_H), dim3(T), 0, s, w.d3, MLP_H, theta + MLP_OFF_W3, MLP_H, w.pre2,
                       w.d2, B);
    hipLaunchKernelGGL(k_mlp_wgrad, blocks(MLP_H * (MLP_H + 1)), dim3(T), 0, s, w.d2, MLP_H, w.a1, MLP_H, B,
                       grad + MLP_OFF_W2, grad + MLP_OFF_B2);
    hipLaunchKernelGGL(k_mlp_bwd, blocks(B * MLP_H), dim3(T), 0, s, w.d2, MLP_H, theta + MLP_OFF_W2, MLP_H, w.pre1,
                       w.d1, B);
    hipLaunchKernelGGL(k_mlp_wgrad, blocks(MLP_H * (MLP_IN + 1)), dim3(T), 0, s, w.d1, MLP_H, x, MLP_IN, B,
                       grad + MLP_OFF_W1, grad + MLP_OFF_B1);
    return hipGetLastError();
}

int mlp_loss_blocks(int B) { return (B + 255) / 256; }

hipError_t mlp_adam(hipStream_t s, float *theta, const float *grad, float *m, float *v, float lr, int step) {
    const float bc1 = 1.f - powf(0.9f, (float)step), bc2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, theta, grad, m, v, MLP_PARAMS, lr,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t mlp_gather(hipStream_t s, const float *x, const float *y, const int *rows, int B, float *bx, float *by) {
    hipLaunchKernelGGL(k_mlp_gather, dim3((unsigned)((B * (MLP_IN + 1) + 255) / 256)), dim3(256), 0, s, x, y, rows,
                       B, bx, by);
    return hipGetLastError();
}

hipError_t mlp_infer(hipStream_t s, const float *theta, int P, const float *x, float *score) {
    if (P <= 0) return hipSuccess;
    const int tiles = (P + 15) / 16;
    const unsigned blocks = (unsigned)std::min((tiles + INFER_WAVES - 1) / INFER_WAVES, 4096);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(WAVE * INFER_WAVES), 0, s, x, P, theta, score);
    return hipGetLastError();
}

}  // namespace rp
