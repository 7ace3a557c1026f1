// unet.hpp -- host interface of the learned warm-start network (definitions in unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fftperm {
namespace unet {

struct Net;

Net* create(uint64_t seed);
void destroy(Net* n);
long num_params(Net* n);
// flat FP32 parameters: per layer the weights [cout][taps][cin] then the biases
void get_params(Net* n, float* out);
void set_params(Net* n, const float* in);
void get_grads(Net* n, float* out);
// Output channels that carry the prediction: the force densities of the three unit load
// directions, direction d's x, y, z components at channels 3d, 3d + 1, 3d + 2.
constexpr int kOutputs = 9;
// forward on the fluid mask (device, n^3, fluid = 1, n a multiple of 8); the prediction stays in
// the network
void predict(Net* n, const uint8_t* mask_dev, int edge);
// last prediction as [V][kOutputs] FP32 (device), synchronous
void pred_field(Net* n, float* out9_dev);
// direction dir of the last prediction gathered on the support voxels `active` (device, nb) into
// x0 (device, 3 nb, double, AoS), synchronous
void gather_guess(Net* n, int dir, const int64_t* active, long nb, double* x0);
// forward + masked MSE against target_dev ([V][kOutputs] FP32) + backward (+ Adam step when
// update), synchronous; returns the loss
double loss_grad(Net* n, const uint8_t* mask_dev, const float* target_dev, int edge, bool update, float lr);
hipStream_t stream(Net* n);

}  // namespace unet
}  // namespace fftperm
