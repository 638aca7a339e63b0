// mhg_mlp.hip -- kernels of the learned limiter's network on the matrix
// cores: feature extraction from a mesh state, forward pass (inference and
// training), backward pass to per-block partial gradients, their reduction
// and the Adam update.
//
// One block is one wavefront and owns 32 rows (samples or cells), one
// 32 x 32 MFMA tile high.  The 128-64-64 layers and every weight gradient run
// as v_mfma_f32_32x32x8_bf16 tiles (BF16 operands staged in LDS, FP32
// accumulation); the 64-2 output layer and the bias sums run on the VALU in
// FP32.  Partial gradients are written per block and summed in a fixed order,
// so training is deterministic.
#include <mutex>

#include "mhg_kernels.hpp"
#include "mhg_mlp.hpp"
#include "mhg_mlp_kernels.hpp"

namespace mhg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXs = kMlpIn + 8;   // padded LDS row of the input tile (bf16 words)
constexpr int kHs = kMlpHid + 8;  // padded LDS row of the hidden tiles

struct MlpLds {
  unsigned short x[kMlpRows][kXs];
  unsigned short h1[kMlpRows][kHs];
  unsigned short h2[kMlpRows][kHs];
  unsigned short d2[kMlpRows][kHs];  // dL/dH2, masked by the ReLU
  unsigned short d1[kMlpRows][kHs];  // dL/dH1, masked by the ReLU
  float y[kMlpRows][kMlpOut];
  float dy[kMlpRows][kMlpOut];
};

// acc += A (32 x K) * B (K x 32) with A(i, k) = a[i * ai + k * ak] and
// B(k, j) = b[k * bk + j * bj] (bf16, LDS or global).  Operand layout of
// v_mfma_f32_32x32x8_bf16: lane l holds A(l % 32, k0 + 4 (l / 32) + t) and
// B(k0 + 4 (l / 32) + t, l % 32), t = 0..3.
__device__ inline f32x16 mfma_tile(f32x16 acc, const unsigned short* a, int ai, int ak, const unsigned short* b,
                                   int bk, int bj, int K) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, q = l >> 5;
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4 av, bv;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + q * 4 + t;
      av[t] = (short)a[r * ai + k * ak];
      bv[t] = (short)b[k * bk + r * bj];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// Row of accumulator element v in this lane (its column is lane % 32).
__device__ inline int acc_row(int v) {
  const int l = threadIdx.x & 63;
  return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
}

__device__ inline f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int v = 0; v < 16; ++v) z[v] = 0.f;
  return z;
}

// Forward pass of the 32 rows staged in x (one wavefront): h1, h2 (bf16)
// and y.  h2 may alias x (x is dead once h1 is formed).  Every wavefront of
// the block calls it together (block barriers).
__device__ inline void mlp_forward(const unsigned short (*x)[kXs], unsigned short (*h1)[kHs],
                                   unsigned short (*h2)[kHs], float (*y)[kMlpOut], const MlpWeights& W) {
  const int l = threadIdx.x & 63;
  for (int nt = 0; nt < 2; ++nt) {  // H1 = relu(X W1^T + b1), B(k, j) = W1[32 nt + j][k]
    const f32x16 acc = mfma_tile(zero16(), &x[0][0], kXs, 1, W.w1 + nt * 32 * kMlpIn, 1, kMlpIn, kMlpIn);
    const int col = nt * 32 + (l & 31);
    const float b = W.b1[col];
#pragma unroll
    for (int v = 0; v < 16; ++v) h1[acc_row(v)][col] = f2bf(fmaxf(acc[v] + b, 0.f));
  }
  __syncthreads();
  for (int nt = 0; nt < 2; ++nt) {  // H2 = relu(H1 W2^T + b2)
    const f32x16 acc = mfma_tile(zero16(), &h1[0][0], kHs, 1, W.w2 + nt * 32 * kMlpHid, 1, kMlpHid, kMlpHid);
    const int col = nt * 32 + (l & 31);
    const float b = W.b2[col];
#pragma unroll
    for (int v = 0; v < 16; ++v) h2[acc_row(v)][col] = f2bf(fmaxf(acc[v] + b, 0.f));
  }
  __syncthreads();
  {  // Y = H2 W3^T + b3: lane l -> row l % 32, output l / 32
    const int row = l & 31, o = l >> 5;
    float s = 0.f;
    for (int k = 0; k < kMlpHid; ++k) s += bf2f(h2[row][k]) * bf2f(W.w3[o * kMlpHid + k]);
    y[row][o] = s + W.b3[o];
  }
  __syncthreads();
}

__device__ inline void mlp_forward_tile(MlpLds& L, const MlpWeights& W) { mlp_forward(L.x, L.h1, L.h2, L.y, W); }

// Loss of the slope regression and its gradient: for each direction d and
// channel k < 4, e = y_d sc - st; loss = sum e^2 / m; dL/dy_d = 2 sum_k e sc / m.
// Returns this lane's share of the tile's loss (rows beyond nvalid count 0).
__device__ inline float mlp_loss_tile(MlpLds& L, const float* T, const long* rows, int nvalid, float inv_m) {
  const int l = threadIdx.x & 63;
  const int row = l & 31, o = l >> 5;
  float lsum = 0.f, g = 0.f;
  if (row < nvalid) {
    const float* t = T + rows[row] * kMlpTgt + o * 8;
    const float y = L.y[row][o];
    for (int k = 0; k < 4; ++k) {
      const float sc = t[2 * k], st = t[2 * k + 1];
      const float e = y * sc - st;
      g += e * sc;
      lsum += e * e;
    }
  }
  L.dy[row][o] = 2.f * g * inv_m;
  __syncthreads();
  return lsum * inv_m;
}

// Backward pass of one tile to its partial gradient part[kMlpParams].
__device__ inline void mlp_backward_tile(MlpLds& L, const MlpWeights& W, float* part) {
  const int l = threadIdx.x & 63;
  // output layer
  for (int e = l; e < kMlpOut * kMlpHid; e += 64) {
    const int o = e / kMlpHid, k = e % kMlpHid;
    float s = 0.f;
    for (int row = 0; row < kMlpRows; ++row) s += L.dy[row][o] * bf2f(L.h2[row][k]);
    part[kPW3 + e] = s;
  }
  if (l < kMlpOut) {
    float s = 0.f;
    for (int row = 0; row < kMlpRows; ++row) s += L.dy[row][l];
    part[kPB3 + l] = s;
  }
  // dL/dH2 = dY W3, masked
  for (int e = l; e < kMlpRows * kMlpHid; e += 64) {
    const int row = e / kMlpHid, k = e % kMlpHid;
    const float s = L.dy[row][0] * bf2f(W.w3[k]) + L.dy[row][1] * bf2f(W.w3[kMlpHid + k]);
    L.d2[row][k] = f2bf(bf2f(L.h2[row][k]) > 0.f ? s : 0.f);
  }
  __syncthreads();
  {  // db2
    float s = 0.f;
    for (int row = 0; row < kMlpRows; ++row) s += bf2f(L.d2[row][l]);
    part[kPB2 + l] = s;
  }
  // dW2[n][k] = sum_rows d2[row][n] h1[row][k]: A(n, row) = d2[row][n], B(row, k) = h1[row][k]
  for (int nt = 0; nt < 2; ++nt)
    for (int kt = 0; kt < 2; ++kt) {
      const f32x16 acc = mfma_tile(zero16(), &L.d2[0][nt * 32], 1, kHs, &L.h1[0][kt * 32], kHs, 1, kMlpRows);
#pragma unroll
      for (int v = 0; v < 16; ++v) part[kPW2 + (nt * 32 + acc_row(v)) * kMlpHid + kt * 32 + (l & 31)] = acc[v];
    }
  // dL/dH1 = d2 W2, masked: A(row, n) = d2[row][n], B(n, k) = W2[n][k]
  for (int kt = 0; kt < 2; ++kt) {
    const f32x16 acc = mfma_tile(zero16(), &L.d2[0][0], kHs, 1, W.w2 + kt * 32, kMlpHid, 1, kMlpHid);
    const int col = kt * 32 + (l & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = acc_row(v);
      L.d1[row][col] = f2bf(bf2f(L.h1[row][col]) > 0.f ? acc[v] : 0.f);
    }
  }
  __syncthreads();
  {  // db1
    float s = 0.f;
    for (int row = 0; row < kMlpRows; ++row) s += bf2f(L.d1[row][l]);
    part[kPB1 + l] = s;
  }
  // dW1[n][k] = sum_rows d1[row][n] x[row][k]
  for (int nt = 0; nt < 2; ++nt)
    for (int kt = 0; kt < 4; ++kt) {
      const f32x16 acc = mfma_tile(zero16(), &L.d1[0][nt * 32], 1, kHs, &L.x[0][kt * 32], kXs, 1, kMlpRows);
#pragma unroll
      for (int v = 0; v < 16; ++v) part[kPW1 + (nt * 32 + acc_row(v)) * kMlpIn + kt * 32 + (l & 31)] = acc[v];
    }
}

// Rows [row0 + 32 blk, ...) of the sample set (through perm if given):
// forward pass; TRAIN adds the loss and the backward pass into part[blk].
template <bool TRAIN>
__global__ __launch_bounds__(64) void k_mlp_rows(const unsigned short* __restrict__ X, const float* __restrict__ T,
                                                 const int* __restrict__ perm, long row0, long row_end,
                                                 MlpWeights W, float* Y, float* part, float inv_m) {
  __shared__ MlpLds L;
  __shared__ long rows[kMlpRows];
  const int l = threadIdx.x;
  const long first = row0 + (long)blockIdx.x * kMlpRows;
  const int nvalid = (int)min((long)kMlpRows, row_end - first);
  if (l < kMlpRows) rows[l] = l < nvalid ? (perm ? (long)perm[first + l] : first + l) : 0;
  __syncthreads();
  for (int e = l; e < kMlpRows * kMlpIn; e += 64) {
    const int row = e / kMlpIn, k = e % kMlpIn;
    L.x[row][k] = row < nvalid ? X[rows[row] * kMlpIn + k] : (unsigned short)0;
  }
  __syncthreads();
  mlp_forward_tile(L, W);
  if (!TRAIN) {
    const int row = l & 31, o = l >> 5;
    if (row < nvalid) Y[rows[row] * kMlpOut + o] = L.y[row][o];
    return;
  }
  float* p = part + (long)blockIdx.x * (kMlpParams + 1);
  float loss = mlp_loss_tile(L, T, rows, nvalid, inv_m);
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
  if (l == 0) p[kMlpParams] = loss;
  mlp_backward_tile(L, W, p);
}

// grad[p] = sum over blocks of part[b][p] in block order (p <= kMlpParams:
// the last entry is the loss).
__global__ void k_mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(long)b * (kMlpParams + 1) + p];
  grad[p] = s;
}

// Adam (Kingma & Ba) on the FP32 master parameters, then their BF16 copies.
__global__ void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float beta1, float beta2, float eps,
                           float corr1, float corr2, unsigned short* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float g = grad[p];
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float w = prm[p] - lr * (a / corr1) / (sqrtf(b / corr2) + eps);
  prm[p] = w;
  wb[p] = f2bf(w);
}

// BF16 copy of the master parameters.
__global__ void k_mlp_to_bf16(const float* __restrict__ prm, unsigned short* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) wb[p] = f2bf(prm[p]);
}

// ---- features of a mesh state and inference --------------------------------
// A block of 8 wavefronts owns a tile of 32 x 8 cells, one 32-cell row per
// wavefront.  The block stages in LDS, once: the network's BF16
// weights (rows padded by 4 words so that the 32 rows of an MFMA operand
// fall on distinct banks) and biases, and the five channels (density,
// specific internal energy, mean of the 4 node velocities x and y in node
// order (i,j), (i+1,j), (i+1,j+1), (i,j+1), largest environment volume
// fraction, 1 for a pure cell) and the membership masks of the tile and a
// 2-cell halo, with indices clamped to the mesh (cells beyond the domain
// repeat the boundary cell).
//
// The layers run transposed, H^T = W X^T, on v_mfma_f32_32x32x16_bf16 (lane
// l holds A(l % 32, 8 h + t) and B(8 h + t, l % 32), h = l / 32, t = 0..7),
// so that one layer's accumulator is the next layer's B operand: lane l holds
// column l % 32 (a cell) and rows 8 (v >> 2) + 4 h + (v & 3) of each 32-row
// accumulator tile, i.e. for 16-row block m the rows 16 m + 4 h + (0..3) and
// 16 m + 8 + 4 h + (0..3) (elements v = 8 (m % 2) + t of tile m / 2).  The
// order of the inputs inside a product is free as long as the weight columns
// follow it, so W2 and W3 are staged with the middle two 4-column chunks of
// every 16 swapped (perm16).  The first B operand, the features, is built by
// each lane for its own cell and its own 64 inputs: lanes 0-31 take channels
// 0, 1 and stencil entries 0-13 of channel 2, lanes 32-63 channels 3, 4 and
// entries 14-24 of channel 2 (then the 3 padding inputs), so that for 50 of
// their 64 inputs the two halves of a wavefront read LDS words a fixed
// distance apart, and each half computes the centre and scale of 3 channels
// only (inf_feature); W1 is staged in that order.
//
// alpha = clamp(y, 0, 2).  A cell whose 3 x 3 neighbourhood does not hold one
// and the same single environment gets alpha = -1 (the classical limiter)
// unless P.learned == 2: at material interfaces independent steep slopes of
// the environments' volumes, masses and energies leave slivers of extreme
// specific energy.  EXPORT also writes the float features and the
// per-channel scales.
constexpr int kInfW = kMlpRows, kInfH = 8, kInfWaves = kInfH, kInfPW = kInfW + 4, kInfPH = kInfH + 4;
constexpr int kW1s = kMlpIn + 8, kW2s = kMlpHid + 8;  // padded LDS weight rows (bf16 words, 16-byte multiples)

struct alignas(16) InferLds {
  unsigned short w1[kMlpHid][kW1s];
  unsigned short w2[kMlpHid][kW2s];
  unsigned short w3[kMlpOut][kW2s];
  float b1[kMlpHid], b2[kMlpHid], b3[4];
  float ch[kMlpChan][kInfPH][kInfPW];
  int mask[kInfPH][kInfPW];
};

// LDS word offset of feature f (channel f / 25, stencil entry f % 25) from
// the cell's (channel 0, top-left stencil) word.
__host__ __device__ constexpr int feat_off(int f) {
  return (f / kMlpSten) * kInfPH * kInfPW + ((f % kMlpSten) / 5) * kInfPW + (f % kMlpSten) % 5;
}
constexpr int kChanWords = kInfPH * kInfPW;

// Network input (feature index, 0-127) held in slot m (0-63) of the lanes of
// half h (0: lanes 0-31, 1: lanes 32-63) of a wavefront; slot m sits at
// k = 16 (m / 8) + 8 h + m % 8 of the first product.
__host__ __device__ constexpr int inf_feature(int h, int m) {
  return m < 2 * kMlpSten ? (m / kMlpSten + 3 * h) * kMlpSten + m % kMlpSten
         : h == 0         ? m
         : m <= 60        ? m + 14
                          : kMlpFeat + (m - 61);
}
// Source column of the 4-column chunk c of a staged W2 / W3 row: inside each
// 16 columns the chunks are stored in the order 0, 2, 1, 3, the order of the
// accumulator rows a lane holds.
__host__ __device__ constexpr int perm16_chunk(int c) { return 4 * (c >> 2) + 2 * (c & 1) + ((c >> 1) & 1); }

typedef __bf16 hbf16x2 __attribute__((ext_vector_type(2)));
typedef float hf32x2 __attribute__((ext_vector_type(2)));
typedef short hs16x2 __attribute__((ext_vector_type(2)));
typedef short hs16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 hbf16x8 __attribute__((ext_vector_type(8)));

__device__ inline f32x16 mfma16(hs16x8 a, hs16x8 b, f32x16 acc) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(hbf16x8, a), __builtin_bit_cast(hbf16x8, b), acc,
                                                 0, 0, 0);
}

// Two floats to two BF16 (round to nearest even, one v_cvt_pk_bf16_f32).
__device__ inline hs16x2 bf16_pair(float a, float b) {
  const hf32x2 f = {a, b};
  return __builtin_bit_cast(hs16x2, __builtin_convertvector(f, hbf16x2));
}

// B fragment of a hidden tile for 16-row block m: relu(acc) in BF16 of
// elements 8 (m % 2) + 0..7 of tile m / 2 (the bias is in the accumulator).
// The ReLU acts on the rounded values as 16-bit integers (v_pk_max_i16 with
// 0): rounding keeps the sign and the order, so relu(bf16(x)) = bf16(relu(x)).
__device__ inline hs16x8 hidden_frag(const f32x16& acc, int half) {
  const hs16x2 z = {0, 0};
  hs16x8 h;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const hs16x2 v = __builtin_elementwise_max(bf16_pair(acc[8 * half + 2 * p], acc[8 * half + 2 * p + 1]), z);
    h[2 * p] = v[0];
    h[2 * p + 1] = v[1];
  }
  return h;
}

// Accumulator tile initialised with the bias of its rows (32 nt + acc_row).
__device__ inline f32x16 bias_tile(const float* bias, int nt) {
  const int q = (threadIdx.x & 63) >> 5;
  f32x16 acc;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const float4 b = *reinterpret_cast<const float4*>(bias + 32 * nt + 8 * jj + 4 * q);
    acc[4 * jj + 0] = b.x;
    acc[4 * jj + 1] = b.y;
    acc[4 * jj + 2] = b.z;
    acc[4 * jj + 3] = b.w;
  }
  return acc;
}

template <class R, bool EXPORT>
__global__ __launch_bounds__(64 * kInfWaves) void k_learned_alpha_2d(Params P, LearnedArgsT<R> A, MlpWeights W,
                                                                     float* Xout, float* Sout) {
  __shared__ InferLds L;
  const int t = threadIdx.x;
  const int ntx = (P.nx + kInfW - 1) / kInfW;
  const int i0 = (int)(blockIdx.x % ntx) * kInfW, j0 = (int)(blockIdx.x / ntx) * kInfH;
  // ---- stage the weights and biases
  {  // W1 with its columns in the lanes' input order: one column per thread
    static_assert(64 * kInfWaves % kMlpIn == 0, "whole columns per pass");
    const int k = t % kMlpIn, src = inf_feature((k >> 3) & 1, 8 * (k >> 4) + (k & 7));
#pragma unroll 4
    for (int r = t / kMlpIn; r < kMlpHid; r += 64 * kInfWaves / kMlpIn) L.w1[r][k] = W.w1[r * kMlpIn + src];
  }
  for (int q = t; q < kMlpHid * (kMlpHid / 4); q += 64 * kInfWaves) {
    const int r = q / (kMlpHid / 4), c4 = q % (kMlpHid / 4);
    *reinterpret_cast<uint2*>(&L.w2[r][4 * c4]) =
        *reinterpret_cast<const uint2*>(W.w2 + r * kMlpHid + 4 * perm16_chunk(c4));
  }
  if (t < kMlpOut * (kMlpHid / 4)) {
    const int r = t / (kMlpHid / 4), c4 = t % (kMlpHid / 4);
    *reinterpret_cast<uint2*>(&L.w3[r][4 * c4]) =
        *reinterpret_cast<const uint2*>(W.w3 + r * kMlpHid + 4 * perm16_chunk(c4));
  }
  if (t < kMlpHid) {
    L.b1[t] = W.b1[t];
    L.b2[t] = W.b2[t];
  }
  if (t < kMlpOut) L.b3[t] = W.b3[t];
  // ---- stage the patch
  for (int q = t; q < kInfPH * kInfPW; q += 64 * kInfWaves) {
    const int pj = q / kInfPW, pi = q % kInfPW;
    const int ci = min(max(i0 + pi - 2, 0), P.nx - 1), cj = min(max(j0 + pj - 2, 0), P.ny - 1);
    const long c = ci + (long)P.nx * cj;
    const long n0 = ci + (long)(P.nx + 1) * cj, nup = P.nx + 1;
    L.ch[0][pj][pi] = (float)A.rho[c];
    L.ch[1][pj][pi] = (float)A.e[c];
    L.ch[2][pj][pi] = (float)(R(0.25) * (((A.ux[n0] + A.ux[n0 + 1]) + A.ux[n0 + 1 + nup]) + A.ux[n0 + nup]));
    L.ch[3][pj][pi] = (float)(R(0.25) * (((A.uy[n0] + A.uy[n0 + 1]) + A.uy[n0 + 1 + nup]) + A.uy[n0 + nup]));
    const int mask = A.mask[c];
    float f = 1.f;
    if (__popc(mask) > 1) {
      f = 0.f;
      for (int e = 0; e < P.ne; ++e)
        if ((mask >> e) & 1) f = fmaxf(f, (float)A.env_f[(long)e * P.nc + c]);
    }
    L.ch[4][pj][pi] = f;
    L.mask[pj][pi] = mask;
  }
  __syncthreads();
  const int w = t >> 6, l = t & 63;
  const int cl = l & 31, q = l >> 5;
  {
    const int row = w;
    const int ci = i0 + cl, cj = j0 + row;
    const bool valid = ci < P.nx && cj < P.ny;
    const long c = valid ? ci + (long)P.nx * cj : 0;
    const float* base = &L.ch[0][row][cl];
    // ---- centre and scale of this half's channels: 3 q and 3 q + 1 (through
    // bq), and channel 2
    const float* bq = base + 3 * q * kChanWords;
    float ctr[3], S[3], inv[3];
#pragma unroll
    for (int it = 0; it < 3; ++it) {
      const float* p = it < 2 ? bq + it * kChanWords : base + 2 * kChanWords;
      const float cv = p[feat_off(12)];
      float sc = 0.f;
#pragma unroll
      for (int s = 0; s < kMlpSten; ++s) sc = fmaxf(sc, fabsf(p[feat_off(s)] - cv));
      ctr[it] = cv;
      S[it] = sc + (1e-6f * fabsf(cv) + 1e-30f);
      inv[it] = 1.f / S[it];
      __builtin_amdgcn_sched_barrier(0);  // keep one channel's 25 loads in flight, not all 75
    }
    // ---- layer 1: H1^T = W1 X^T + b1, the features of this lane's slots
    f32x16 acc1[2] = {bias_tile(L.b1, 0), bias_tile(L.b1, 1)};
#pragma unroll
    for (int j = 0; j < kMlpIn / 16; ++j) {
      float xf[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int m = 8 * j + u;
        float x;
        if (m < 2 * kMlpSten) {
          x = (bq[(m / kMlpSten) * kChanWords + feat_off(m % kMlpSten)] - ctr[m / kMlpSten]) * inv[m / kMlpSten];
        } else {  // channel 2: entry m - 50 (lanes 0-31) or m - 36 (lanes 32-63, padding past 24)
          const int s0 = m - 2 * kMlpSten, s1 = s0 + 14;
          if (s1 < kMlpSten) {
            const float v = base[2 * kChanWords + (q ? feat_off(s1) : feat_off(s0))];
            x = (v - ctr[2]) * inv[2];
          } else {
            x = q ? 0.f : (base[2 * kChanWords + feat_off(s0)] - ctr[2]) * inv[2];
          }
        }
        xf[u] = x;
        if (EXPORT && valid) Xout[c * kMlpIn + (q ? inf_feature(1, m) : inf_feature(0, m))] = x;
      }
      hs16x8 xb;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const hs16x2 v = bf16_pair(xf[2 * p], xf[2 * p + 1]);
        xb[2 * p] = v[0];
        xb[2 * p + 1] = v[1];
      }
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
        acc1[nt] = mfma16(*reinterpret_cast<const hs16x8*>(&L.w1[32 * nt + cl][16 * j + 8 * q]), xb, acc1[nt]);
      __builtin_amdgcn_sched_barrier(0);  // bounds the features held in registers to one k-step
    }
    if (EXPORT && valid) {
      Sout[c * kMlpChan + 3 * q] = S[0];
      if (q == 0) Sout[c * kMlpChan + 1] = S[1], Sout[c * kMlpChan + 2] = S[2];
      else Sout[c * kMlpChan + 4] = S[1];
    }
    // ---- layer 2: H2^T = W2 relu(H1^T) + b2
    f32x16 acc2[2] = {bias_tile(L.b2, 0), bias_tile(L.b2, 1)};
#pragma unroll
    for (int j = 0; j < kMlpHid / 16; ++j) {
      const hs16x8 hb = hidden_frag(acc1[j >> 1], j & 1);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
        acc2[nt] = mfma16(*reinterpret_cast<const hs16x8*>(&L.w2[32 * nt + cl][16 * j + 8 * q]), hb, acc2[nt]);
    }
    // ---- output layer: Y^T = W3 relu(H2^T), W3 padded to 32 rows with zeros
    f32x16 acc3 = zero16();
#pragma unroll
    for (int j = 0; j < kMlpHid / 16; ++j) {
      const hs16x8 hb = hidden_frag(acc2[j >> 1], j & 1);
      hs16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
      if (cl < kMlpOut) a = *reinterpret_cast<const hs16x8*>(&L.w3[cl][16 * j + 8 * q]);
      acc3 = mfma16(a, hb, acc3);
    }
    // lanes 0-31 hold rows 0 and 1 (y_x, y_y) of their cell in elements 0 and 1
    if (q == 0 && valid) {
      const int m0 = L.mask[row + 2][cl + 2];
      bool uni = __popc(m0) == 1;
      for (int dj = 0; dj < 3; ++dj)
        for (int di = 0; di < 3; ++di) uni = uni && L.mask[row + 1 + dj][cl + 1 + di] == m0;
      if (A.alpha) {
        const bool use = uni || P.learned == 2;
        A.alpha[c] = use ? (R)fminf(fmaxf(acc3[0] + L.b3[0], 0.f), 2.f) : (R)A.iface;
        A.alpha[P.nc + c] = use ? (R)fminf(fmaxf(acc3[1] + L.b3[1], 0.f), 2.f) : (R)A.iface;
      }
    }
  }
}

// ---- launchers ------------------------------------------------------------------
// Partial gradients of the allocation-free gradient: a fixed array, one
// copy per device, shared by every call on that device.
static __device__ __align__(256) float scratch[(kMlpNoallocRows / kMlpRows) * (kMlpParams + 1)];

int mlp_noalloc_scratch(int device, float** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static float* ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return MHG_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return MHG_E_HIP;
    if (cur != device && hipSetDevice(device) != hipSuccess) return MHG_E_HIP;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) hipSetDevice(cur);
    if (e != hipSuccess || !p) return MHG_E_HIP;
    ptr[device] = static_cast<float*>(p);
  }
  *out = ptr[device];
  return MHG_OK;
}

void launch_mlp_rows(hipStream_t st, bool train, const unsigned short* X, const float* T, const int* perm, long row0,
                     long row_end, const MlpWeights& W, float* Y, float* part, float inv_m) {
  const unsigned nblk = (unsigned)((row_end - row0 + kMlpRows - 1) / kMlpRows);
  if (nblk == 0) return;
  if (train) hipLaunchKernelGGL(k_mlp_rows<true>, dim3(nblk), dim3(64), 0, st, X, T, perm, row0, row_end, W, Y, part, inv_m);
  else hipLaunchKernelGGL(k_mlp_rows<false>, dim3(nblk), dim3(64), 0, st, X, T, perm, row0, row_end, W, Y, part, inv_m);
}

void launch_mlp_reduce(hipStream_t st, const float* part, int nblk, float* grad) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st, part, nblk, grad);
}

void launch_mlp_adam(hipStream_t st, float* prm, float* m1, float* m2, const float* grad, float lr, float beta1,
                     float beta2, float eps, float corr1, float corr2, unsigned short* wb) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, prm, m1, m2, grad, lr, beta1, beta2,
                     eps, corr1, corr2, wb);
}

void launch_mlp_to_bf16(hipStream_t st, const float* prm, unsigned short* wb) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, prm, wb);
}

template <class R>
void launch_learned_alpha_2d(hipStream_t st, const Params& P, const LearnedArgsT<R>& A, const MlpWeights& W,
                             float* Xout, float* Sout) {
  const dim3 g((unsigned)(((P.nx + kInfW - 1) / kInfW) * ((P.ny + kInfH - 1) / kInfH))), b(64 * kInfWaves);
  if (Xout) hipLaunchKernelGGL((k_learned_alpha_2d<R, true>), g, b, 0, st, P, A, W, Xout, Sout);
  else hipLaunchKernelGGL((k_learned_alpha_2d<R, false>), g, b, 0, st, P, A, W, Xout, Sout);
}

template void launch_learned_alpha_2d<double>(hipStream_t, const Params&, const LearnedArgsT<double>&,
                                              const MlpWeights&, float*, float*);
template void launch_learned_alpha_2d<float>(hipStream_t, const Params&, const LearnedArgsT<float>&,
                                             const MlpWeights&, float*, float*);

}  // namespace mhg
