// mhg_batch2d.hip -- persistent kernels of the batching form.
//
// Each kernel is launched once per phase for the whole batch with a fixed
// grid; its blocks walk a device work list of (problem, block) items and run
// the same block body as the single-mesh kernel on the problem's slice of the
// batch arrays, with the problem's Params and scalars.  A problem that has
// stopped is skipped by the bodies' stop test until the host drops it from
// the lists.  The lists and their counts live in device memory, so a
// captured graph of these launches stays valid when the host rewrites them.
#include "mhg_batch.hpp"
#include "mhg_body2d.hpp"

namespace mhg {

// ---- problem slices of the batch argument blocks ------------------------
template <class R>
__device__ inline NodeLagArgsT<R> slice(NodeLagArgsT<R> A, const BatchProblem& b) {
  const long c = b.cb, n = b.nb;
  A.p += c;
  A.q += c;
  A.m += c;
  A.xpx += n;
  A.xpy += n;
  A.ux += n;
  A.uy += n;
  A.xnx += n;
  A.xny += n;
  return A;
}

template <class R>
__device__ inline CellLagArgsT<R> slice(CellLagArgsT<R> A, const BatchProblem& b) {
  const long c = b.cb, n = b.nb, ce = (long)b.P.ne * c;
  A.xnx += n;
  A.xny += n;
  A.rho += c;
  A.e += c;
  A.p += c;
  A.cs += c;
  A.q += c;
  A.m += c;
  A.mask += c;
  A.env_f += ce;
  A.env_rho += ce;
  A.env_m += ce;
  A.env_q += ce;
  A.env_p += ce;
  A.env_e += ce;
  A.U += (long)b.P.nvu * c;
  A.Vlag += c;
  A.L += c;
  return A;
}

template <class R>
__device__ inline CellSweepArgsT<R> slice(CellSweepArgsT<R> A, const BatchProblem& b) {
  const long c = b.cb, n = b.nb;
  A.Uin += (long)b.P.nvu * c;
  A.Uout += (long)b.P.nvu * c;
  A.Vlag += c;
  A.xnx += n;
  A.xny += n;
  A.ux += n;
  A.uy += n;
  A.contrib += (long)b.P.ne * c;
  return A;
}

template <class R>
__device__ inline NodeSweepArgsT<R> slice(NodeSweepArgsT<R> A, const BatchProblem& b) {
  const long c = b.cb, n = b.nb;
  A.m += c;
  A.ux += n;
  A.uy += n;
  A.xnx += n;
  A.xny += n;
  A.contrib += (long)b.P.ne * c;
  A.Ud0 += n;
  A.Ud1 += n;
  A.Ud3 += n;
  A.dg0 += n;
  A.dg1 += n;
  A.uox += n;
  A.uoy += n;
  return A;
}

template <class R>
__device__ inline CellFinalArgsT<R> slice(CellFinalArgsT<R> A, const BatchProblem& b) {
  const long c = b.cb, n = b.nb, ce = (long)b.P.ne * c;
  A.U += (long)b.P.nvu * c;
  A.Vlag += c;
  A.L += c;
  A.ux += n;
  A.uy += n;
  A.rho += c;
  A.e += c;
  A.p += c;
  A.cs += c;
  A.q += c;
  A.m += c;
  A.mask += c;
  A.env_rho += ce;
  A.env_e += ce;
  A.env_p += ce;
  A.env_cs += ce;
  A.env_q += ce;
  A.env_m += ce;
  A.env_f += ce;
  A.env_mf += ce;
  A.block_min += b.slot;
  return A;
}

// ---- persistent kernels ----------------------------------------------------
// The loop bound is read once; items are uniform across the block.  Bodies
// that stage through LDS are followed by a barrier before the next item.
#define MHG_BATCH_ARGS \
  const BatchProblem* __restrict__ probs, DevScalars* S, const BatchItem* __restrict__ items, const int* __restrict__ count

template <class R>
__global__ __launch_bounds__(256) void kb_node_lagrange_2d(MHG_BATCH_ARGS, NodeLagArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    node_lagrange_2d_body<R, true, true, true>(b.P, S + it.p, slice(A, b), it.b);
  }
}

template <class R, bool PC, int NEM>
__global__ __launch_bounds__(256) void kb_cell_lagrange_2d(MHG_BATCH_ARGS, CellLagArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    cell_lagrange_2d_body<R, PC, NEM>(b.P, S + it.p, slice(A, b), it.b);
  }
}

template <class R, int NEM, bool FIRST, int LIM>
__global__ __launch_bounds__(256) void kb_cell_sweep_x_2d(MHG_BATCH_ARGS, CellSweepArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    cell_sweep_lane_2d_body<R, 0, NEM, FIRST, LIM>(b.P, S + it.p, slice(A, b), it.b);
  }
}

template <class R, int NEM, bool FIRST, int LIM, bool SEG>
__global__ __launch_bounds__(1024) void kb_cell_sweep_y_2d(MHG_BATCH_ARGS, CellSweepArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    cell_sweep_ytile_2d_body<R, NEM, FIRST, LIM, SEG>(b.P, S + it.p, slice(A, b), it.b);
    __syncthreads();
  }
}

template <class R, bool FIRST, int NEM, int LIM>
__global__ __launch_bounds__(256) void kb_node_sweep_x_2d(MHG_BATCH_ARGS, NodeSweepArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    node_sweep_x_2d_body<R, FIRST, NEM, LIM>(b.P, S + it.p, slice(A, b), it.b);
  }
}

template <class R, bool FIRST, int NEM, int LIM, bool SEG>
__global__ __launch_bounds__(1024) void kb_node_sweep_y_2d(MHG_BATCH_ARGS, NodeSweepArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    node_sweep_ytile_2d_body<R, FIRST, NEM, LIM, SEG>(b.P, S + it.p, slice(A, b), it.b);
    __syncthreads();
  }
}

template <class R, int NEM>
__global__ __launch_bounds__(256) void kb_cell_final_2d(MHG_BATCH_ARGS, CellFinalArgsT<R> A) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    const CellFinalArgsT<R> Ap = slice(A, b);
    cell_final_2d_body<R, 256, NEM>(b.P, S + it.p, Ap, it.b, Ap.block_min);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void kb_dt_finalize(MHG_BATCH_ARGS, const double* block_min) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchItem it = items[i];
    const BatchProblem& b = probs[it.p];
    dt_finalize_body<256, true>(b.P, S + it.p, block_min + b.slot, (int)((b.P.nc + 255) / 256));
    __syncthreads();
  }
}

#undef MHG_BATCH_ARGS

// ---- launchers ---------------------------------------------------------------
#define MHG_BNEM(ne, F) \
  do {                  \
    if ((ne) <= 1) F(1);      \
    else if ((ne) <= 2) F(2); \
    else F(3);                \
  } while (0)
#define MHG_BLIM(lim, F)       \
  do {                         \
    if ((lim) == 0) F(0);      \
    else if ((lim) == 1) F(1); \
    else if ((lim) == 2) F(2); \
    else F(kLimG);             \
  } while (0)
#define MHG_BLARGS(list) probs, S, L.items[list], L.count + (list)

template <class R>
void launch_batch_node_lagrange_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                   const NodeLagArgsT<R>& A, const BatchLaunch& L) {
  hipLaunchKernelGGL((kb_node_lagrange_2d<R>), dim3(L.grid[BL_NODE]), dim3(256), 0, st, MHG_BLARGS(BL_NODE), A);
}

template <class R>
void launch_batch_cell_lagrange_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                   const CellLagArgsT<R>& A, const BatchLaunch& L) {
  const dim3 g(L.grid[BL_CELL]), b(256);
#define CL_(N)                                                                                        \
  if (L.pc) hipLaunchKernelGGL((kb_cell_lagrange_2d<R, true, N>), g, b, 0, st, MHG_BLARGS(BL_CELL), A); \
  else hipLaunchKernelGGL((kb_cell_lagrange_2d<R, false, N>), g, b, 0, st, MHG_BLARGS(BL_CELL), A)
  MHG_BNEM(L.ne, CL_);
#undef CL_
}

template <class R, int NEM, int LM>
static void batch_cell_sweep(hipStream_t st, const BatchProblem* probs, DevScalars* S, const CellSweepArgsT<R>& A,
                             const BatchLaunch& L, int dir, bool first) {
  if (dir == 0) {
    const dim3 g(L.grid[BL_CX]), b(256);
    if (first) hipLaunchKernelGGL((kb_cell_sweep_x_2d<R, NEM, true, LM>), g, b, 0, st, MHG_BLARGS(BL_CX), A);
    else hipLaunchKernelGGL((kb_cell_sweep_x_2d<R, NEM, false, LM>), g, b, 0, st, MHG_BLARGS(BL_CX), A);
    return;
  }
  const dim3 gs(L.grid[BL_CY_SEG]), gt(L.grid[BL_CY_TALL]), b(1024);
  if (first) {
    hipLaunchKernelGGL((kb_cell_sweep_y_2d<R, NEM, true, LM, true>), gs, b, 0, st, MHG_BLARGS(BL_CY_SEG), A);
    hipLaunchKernelGGL((kb_cell_sweep_y_2d<R, NEM, true, LM, false>), gt, b, 0, st, MHG_BLARGS(BL_CY_TALL), A);
  } else {
    hipLaunchKernelGGL((kb_cell_sweep_y_2d<R, NEM, false, LM, true>), gs, b, 0, st, MHG_BLARGS(BL_CY_SEG), A);
    hipLaunchKernelGGL((kb_cell_sweep_y_2d<R, NEM, false, LM, false>), gt, b, 0, st, MHG_BLARGS(BL_CY_TALL), A);
  }
}

template <class R>
void launch_batch_cell_sweep_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const CellSweepArgsT<R>& A, const BatchLaunch& L, int dir, bool first) {
#define CS_(N, LM) batch_cell_sweep<R, N, LM>(st, probs, S, A, L, dir, first)
#define CSL_(LM)                   \
  do {                             \
    if (L.ne <= 1) CS_(1, LM);      \
    else if (L.ne <= 2) CS_(2, LM); \
    else CS_(3, LM);               \
  } while (0)
  MHG_BLIM(L.lim, CSL_);
#undef CSL_
#undef CS_
}

template <class R, int NEM, int LM>
static void batch_node_sweep(hipStream_t st, const BatchProblem* probs, DevScalars* S, const NodeSweepArgsT<R>& A,
                             const BatchLaunch& L, int dir, bool first) {
  if (dir == 0) {
    const dim3 g(L.grid[BL_NX]), b(256);
    if (first) hipLaunchKernelGGL((kb_node_sweep_x_2d<R, true, NEM, LM>), g, b, 0, st, MHG_BLARGS(BL_NX), A);
    else hipLaunchKernelGGL((kb_node_sweep_x_2d<R, false, NEM, LM>), g, b, 0, st, MHG_BLARGS(BL_NX), A);
    return;
  }
  const dim3 gs(L.grid[BL_NY_SEG]), gt(L.grid[BL_NY_TALL]), b(1024);
  if (first) {
    hipLaunchKernelGGL((kb_node_sweep_y_2d<R, true, NEM, LM, true>), gs, b, 0, st, MHG_BLARGS(BL_NY_SEG), A);
    hipLaunchKernelGGL((kb_node_sweep_y_2d<R, true, NEM, LM, false>), gt, b, 0, st, MHG_BLARGS(BL_NY_TALL), A);
  } else {
    hipLaunchKernelGGL((kb_node_sweep_y_2d<R, false, NEM, LM, true>), gs, b, 0, st, MHG_BLARGS(BL_NY_SEG), A);
    hipLaunchKernelGGL((kb_node_sweep_y_2d<R, false, NEM, LM, false>), gt, b, 0, st, MHG_BLARGS(BL_NY_TALL), A);
  }
}

template <class R>
void launch_batch_node_sweep_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const NodeSweepArgsT<R>& A, const BatchLaunch& L, int dir, bool first) {
#define NS_(N, LM) batch_node_sweep<R, N, LM>(st, probs, S, A, L, dir, first)
#define NSL_(LM)                   \
  do {                             \
    if (L.ne <= 1) NS_(1, LM);      \
    else if (L.ne <= 2) NS_(2, LM); \
    else NS_(3, LM);               \
  } while (0)
  MHG_BLIM(L.lim, NSL_);
#undef NSL_
#undef NS_
}

template <class R>
void launch_batch_cell_final_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const CellFinalArgsT<R>& A, const BatchLaunch& L) {
  const dim3 g(L.grid[BL_CELL]), b(256);
#define CF_(N) hipLaunchKernelGGL((kb_cell_final_2d<R, N>), g, b, 0, st, MHG_BLARGS(BL_CELL), A)
  MHG_BNEM(L.ne, CF_);
#undef CF_
  hipLaunchKernelGGL(kb_dt_finalize, dim3(L.grid[BL_PROB]), dim3(256), 0, st, MHG_BLARGS(BL_PROB), A.block_min);
}

#undef MHG_BLARGS
#undef MHG_BLIM
#undef MHG_BNEM

#define MHG_BATCH_INSTANTIATE(R)                                                                        \
  template void launch_batch_node_lagrange_2d<R>(hipStream_t, const BatchProblem*, DevScalars*,         \
                                                 const NodeLagArgsT<R>&, const BatchLaunch&);          \
  template void launch_batch_cell_lagrange_2d<R>(hipStream_t, const BatchProblem*, DevScalars*,         \
                                                 const CellLagArgsT<R>&, const BatchLaunch&);          \
  template void launch_batch_cell_sweep_2d<R>(hipStream_t, const BatchProblem*, DevScalars*,            \
                                              const CellSweepArgsT<R>&, const BatchLaunch&, int, bool); \
  template void launch_batch_node_sweep_2d<R>(hipStream_t, const BatchProblem*, DevScalars*,            \
                                              const NodeSweepArgsT<R>&, const BatchLaunch&, int, bool); \
  template void launch_batch_cell_final_2d<R>(hipStream_t, const BatchProblem*, DevScalars*,            \
                                              const CellFinalArgsT<R>&, const BatchLaunch&);
MHG_BATCH_INSTANTIATE(double)
MHG_BATCH_INSTANTIATE(float)
#undef MHG_BATCH_INSTANTIATE

}  // namespace mhg
