// rp_mlp.h -- the network of the learned candidate ordering: layout of its parameters and the device launchers.
#pragma once

#include <hip/hip_runtime.h>

namespace rp {

constexpr int MLP_IN = 12;  // relative position (3, in units of the total interaction range) + relative rotation (9)
constexpr int MLP_H = 64;
constexpr int MLP_OFF_W1 = 0;                            // [64][12]
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;  // [64]
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;           // [64][64]
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;           // [64]
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H;
constexpr int MLP_PARAMS = MLP_OFF_B4 + 1;               // 9217

// Device work arrays of one training batch of B samples (B x 64 floats each, dz B floats, lossPart
// mlp_loss_blocks(B) floats).
struct MlpWork {
    float *pre1, *a1, *pre2, *a2, *pre3, *a3, *d1, *d2, *d3, *dz, *lossPart;
};

// Forward and backward pass of a batch (x: B x 12 features, y: B labels in {0, 1}): the gradient of the mean binary
// cross-entropy into grad (MLP_PARAMS floats, every entry written); the per-block loss sums into w.lossPart.
hipError_t mlp_forward_backward(hipStream_t s, const float *theta, int B, const float *x, const float *y,
                                MlpWork &w, float *grad);
int mlp_loss_blocks(int B);
// One Adam step (step counts from 1) on every parameter.
hipError_t mlp_adam(hipStream_t s, float *theta, const float *grad, float *m, float *v, float lr, int step);
// Rows `rows` (B indices) of the dataset (x: 12 features per row, y: labels) into a contiguous batch.
hipError_t mlp_gather(hipStream_t s, const float *x, const float *y, const int *rows, int B, float *bx, float *by);
// Overlap probabilities of P pairs (x: P x 12 features) with BF16 MFMA layers and FP32 accumulation.
hipError_t mlp_infer(hipStream_t s, const float *theta, int P, const float *x, float *score);

}  // namespace rp
