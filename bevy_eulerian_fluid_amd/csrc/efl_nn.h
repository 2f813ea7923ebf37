// efl_nn.h - the learned pressure preconditioner: a 4-level convolutional
// U-Net (3x3 convolutions, BF16 weights and activations, FP32 accumulation on
// MFMA), its training kernels, and the hooks the CG solver uses.
//
// No counterpart in the reference (which solves with Jacobi, Gauss-Seidel or
// one V-cycle).  Inputs per cell: the residual r / sigma, the solid fraction
// 1 - mean(area fractions) and the clamped air level set; output: an
// approximation of A^-1 r (scaled back by sigma / factor), used as the
// preconditioner z = M(r) of flexible CG.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "efl_internal.h"

namespace efl {
namespace nn {

constexpr int kLayers = 15;
constexpr int kInCh = 16;      // input tensor channels (3 features, zero padded)
constexpr int kFeatures = 3;
constexpr int kLevels = 4;     // resolutions 1, 1/2, 1/4, 1/8
constexpr float kLsScale = 0.125f;  // level-set feature = clamp(phi / 8, -1, 1)
// A dataset sample: planes r / sigma, solid fraction, clamped level set,
// mask, target / rms(target) (the five efl_nn_get_sample returns), then the
// pressure operator's stencil: diagonal and the four off-diagonal weights
// (-x, +x, -y, +y) of each unknown, so that the residual-norm loss can form
// A y for that sample.
constexpr int kSamplePlanes = 10;
constexpr int kLossMse = 0, kLossResidual = 1;  // efl_nn_set_loss

// Layer k: out = act(conv3x3(cat(up?(A), B)) + bias); A = tensor in_a[k]
// (upsampled x2 when upa), B = tensor in_b[k] (or none).
struct LayerDef {
  int ca, cb, cout, level, upa, relu, in_a, in_b;
};
// tensor ids: 0..14 = layer outputs Y_k, 15 = input X, 16..18 = pooled P0..P2
constexpr int kTX = 15, kTP0 = 16;
constexpr int kTensors = 19;
extern const LayerDef kNet[kLayers];

size_t param_count();
size_t layer_weight_offset(int k);  // [cout][cin][3][3] fp32
size_t layer_bias_offset(int k);    // [cout] fp32
double flops_per_pixel();           // forward FLOPs per level-0 pixel

// Tensor shapes for N samples of Wp x Hp (multiples of 8).
struct Shapes {
  int n, w, h;
  size_t elems(int tensor) const;  // bf16 elements
  int level(int tensor) const;
  int channels(int tensor) const;
};

// Device parameters in BF16 launch layouts.
struct Packed {
  uint16_t* wf[kLayers];  // [9][cout][cin]
  uint16_t* wb[kLayers];  // [9][cin][cout], taps flipped (input gradient)
  float* bias[kLayers];
};

// Activations of one forward pass (and their gradients for training).
struct Activations {
  Shapes s;
  uint16_t* t[kTensors];
  uint16_t* g[kTensors];   // null unless allocated for training
  uint16_t* scratch;       // concat-gradient buffer (largest level-0 concat)
  int strip[kLevels];      // convolution strip per level: 64 or 16 pixels, 0 = by size
};

// Forward pass from act.t[kTX] to act.t[14].
void forward(hipStream_t st, const Packed& p, Activations& a);

// Inference inside CG for one instance (I: its device descriptor, W x H its
// grid, done_flag: its scal[SC_DONE] in device memory): features from
// pcg_r, the network, then z (masked, scaled) with the partials of r . z and
// z . q_prev.  Every launch returns at once when the solve has stopped.
void precondition(hipStream_t st, const Packed& p, Activations& a, const DevInst* I, int W, int H,
                  const double* done_flag);

// Training state (owned by efl_nn).
struct Trainer {
  int w = 0, h = 0;        // sample size
  uint32_t capacity = 0, count = 0;
  float* data = nullptr;   // per sample: kSamplePlanes planes, fp32
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float* partial = nullptr;  // weight-gradient partials
  size_t partial_floats = 0;
  double* loss = nullptr;    // [0] sum of the last evaluation
  int* idx = nullptr;        // batch sample indices (device)
  float* target = nullptr;   // batch target plane
  float* mask = nullptr;     // batch mask plane
  double* rsums = nullptr;   // residual loss: r.r, r.Ay, Ay.Ay per batch sample
  int loss_mode = kLossMse;
  Activations act{};
  int act_n = 0;
  uint64_t adam_t = 0;
};

// Device buffers of a training batch of n samples.
int trainer_reserve_batch(Trainer& t, int n);
void pack(hipStream_t st, const float* params, Packed& p);
// Loss into t.loss and, with grads, the parameter gradients into t.grads.
// kLossMse: mean over the batch's pixels of mask (o - t)^2.  kLossResidual:
// mean over the batch's samples of 1 - (r . A y)^2 / (|r|^2 |A y|^2), y the
// output: the squared relative residual that the best multiple of y leaves
// (min over a of |r - a A y|^2 / |r|^2), which is what a flexible CG step
// with preconditioner output y achieves.
void eval_batch(hipStream_t st, Trainer& t, const Packed& p, const int* host_idx, int n, bool grads);
void adam_step(hipStream_t st, Trainer& t, float lr);
// Build one dataset sample on the device from an instance's fields and two
// device planes: r / sigma and target / trms (sigma, trms: their rms).
void make_sample(hipStream_t st, Trainer& t, uint32_t slot, const DevInst* I, int W, int H,
                 const float* dev_r, const float* dev_t, float sigma, float trms);
void read_output(hipStream_t st, const Activations& a, int n, float* dev_out);
void init_params(float* host, uint64_t seed);

}  // namespace nn
}  // namespace efl
