// The CPU build leaves out the network kernels (BF16 matrix-core code): their launchers report "not supported".
#include "rp_mlp.h"

namespace rp {

hipError_t mlp_forward_backward(hipStream_t, const float *, int, const float *, const float *, MlpWork &, float *) {
    return hipErrorNotSupported;
}
int mlp_loss_blocks(int B) { return (B + 255) / 256; }
hipError_t mlp_adam(hipStream_t, float *, const float *, float *, float *, float, int) { return hipErrorNotSupported; }
hipError_t mlp_gather(hipStream_t, const float *, const float *, const int *, int, float *, float *) {
    return hipErrorNotSupported;
}
hipError_t mlp_infer(hipStream_t, const float *, int, const float *, float *) { return hipErrorNotSupported; }

}  // namespace rp
