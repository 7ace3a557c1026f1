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
// forward on the fluid mask (device, n^3, fluid = 1, n a multiple of 8) swapped across axis
// (0: none, 1: x <-> y, 2: x <-> z); the prediction stays in the network
void predict(Net* n, const uint8_t* mask_dev, int edge, int axis);
// last prediction as [V][3] FP32 (device), synchronous
void pred_field(Net* n, float* out3_dev);
// last prediction (made with `axis`) swapped back, components exchanged, gathered on the support
// voxels `active` (device, nb) into x0 (device, 3 nb, double), synchronous
void gather_guess(Net* n, int axis, const int64_t* active, long nb, double* x0);
// forward + masked MSE + backward (+ Adam step when update), synchronous; returns the loss
double loss_grad(Net* n, const uint8_t* mask_dev, const float* target_dev, int edge, bool update, float lr);
hipStream_t stream(Net* n);

}  // namespace unet
}  // namespace fftperm
