// mhg_mlp_kernels.hpp -- launchers of the learned limiter's network kernels.
#pragma once

#include "mhg_kernels.hpp"
#include "mhg_mlp.hpp"

namespace mhg {

// State a context's inference reads (cell fields, node velocity, membership)
// and the coefficients [2][nc] it writes.
template <class R>
struct LearnedArgsT {
  const R *rho, *e, *ux, *uy;
  const int* mask;
  const R* env_f;
  R* alpha;
  double iface;  // coefficient of cells at material interfaces (< 0: classical limiter)
};

// Forward pass of rows [row0, row_end) of X ([rows][128] bf16, through perm
// if not null) into Y [rows][2]; with train, the loss and the partial
// gradients of each 32-row block into part [blocks][kMlpParams + 1]
// (inv_m = 1 / rows of the minibatch).
void launch_mlp_rows(hipStream_t st, bool train, const unsigned short* X, const float* T, const int* perm, long row0,
                     long row_end, const MlpWeights& W, float* Y, float* part, float inv_m);
void launch_mlp_reduce(hipStream_t st, const float* part, int nblk, float* grad);
void launch_mlp_adam(hipStream_t st, float* prm, float* m1, float* m2, const float* grad, float lr, float beta1,
                     float beta2, float eps, float corr1, float corr2, unsigned short* wb);
void launch_mlp_to_bf16(hipStream_t st, const float* prm, unsigned short* wb);
// Features of every cell of a 2D state and the coefficients alpha; Xout
// ([nc][128]) and Sout ([nc][5] channel scales) are written when not null.
template <class R>
void launch_learned_alpha_2d(hipStream_t st, const Params& P, const LearnedArgsT<R>& A, const MlpWeights& W,
                             float* Xout, float* Sout);

// Device address of the allocation-free gradient's partial-gradient array
// on `device` (looked up once per device; the array itself is a static
// device variable, one copy per GPU, kept until the process ends).
int mlp_noalloc_scratch(int device, float** out);

inline MlpWeights mlp_weights(const unsigned short* wb, const float* prm) {
  MlpWeights W;
  W.w1 = wb + kPW1;
  W.w2 = wb + kPW2;
  W.w3 = wb + kPW3;
  W.b1 = prm + kPB1;
  W.b2 = prm + kPB2;
  W.b3 = prm + kPB3;
  return W;
}

}  // namespace mhg
