// mhg_mlp.hpp -- the learned limiter's network: a 128-64-64-2 perceptron
// with ReLU, BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_32x32x8_bf16) and FP32 biases.
//
// Input: a 5 x 5 stencil of five channels around a cell (density, specific
// internal energy, cell-mean x and y velocity, largest environment volume
// fraction), each channel centred on the cell's value and divided by its
// largest deviation over the stencil, channel-major (channel * 25 + (dj + 2)
// * 5 + di + 2), zero-padded to 128.  Output: the slope coefficients alpha_x,
// alpha_y of the cell, clamped to [0, 2] where they are used.
#pragma once

#include <cstring>

#include "mhg_common.hpp"

namespace mhg {

constexpr int kMlpIn = 128, kMlpHid = 64, kMlpOut = 2;
constexpr int kMlpChan = 5, kMlpSten = 25, kMlpFeat = kMlpChan * kMlpSten;  // 125 used inputs
// FP32 parameter vector: W1 [64][128], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]
constexpr int kPW1 = 0, kPB1 = kPW1 + kMlpHid * kMlpIn, kPW2 = kPB1 + kMlpHid, kPB2 = kPW2 + kMlpHid * kMlpHid,
              kPW3 = kPB2 + kMlpHid, kPB3 = kPW3 + kMlpOut * kMlpHid, kMlpParams = kPB3 + kMlpOut;
// training targets per sample: for direction d and channel k < 4, the
// normalised central slope sc and the normalised reference slope st
constexpr int kMlpTgt = 2 * 4 * 2;
// rows per block of the network kernels (one wavefront, one 32-row MFMA tile)
constexpr int kMlpRows = 32;
// largest minibatch of the allocation-free gradient (mhg_mlp_gradient_noalloc):
// its per-block partial gradients live in a fixed device array of
// kMlpNoallocRows / kMlpRows * (kMlpParams + 1) floats (6.4 MB per device)
constexpr long kMlpNoallocRows = 4096;

// Device copies of the parameters used by the kernels.
struct MlpWeights {
  const unsigned short *w1, *w2, *w3;  // bf16
  const float *b1, *b2, *b3;
};

__host__ __device__ inline unsigned short f2bf(float x) {
  // round to nearest even (finite inputs; infinities stay, NaN stays NaN);
  // written with selects only, so that device code has no branch per value
  unsigned int u;
  memcpy(&u, &x, 4);
  const unsigned int rounded = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  const unsigned int special = (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
  return (unsigned short)(((u & 0x7f800000u) == 0x7f800000u) ? special : rounded);
}
__host__ __device__ inline float bf2f(unsigned short h) {
  const unsigned int u = (unsigned int)h << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}

}  // namespace mhg
