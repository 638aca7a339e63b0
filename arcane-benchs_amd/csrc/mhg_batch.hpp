// mhg_batch.hpp -- persistent kernels of the batching form (many independent
// 2D meshes advanced together).
//
// Layout: the problems' arrays are concatenated.  Problem p owns cells
// [cb, cb + nc) and nodes [nb, nb + nn) of every cell and node array, the
// slice [nvu cb, nvu (cb + nc)) of the U buffers (variable-major inside the
// slice, as for a single mesh), [ne cb, ne (cb + nc)) of the environment
// arrays and of the dual-flux scratch, and CFL minimum slots [slot, slot +
// ceil(nc / 256)).  Its Params and DevScalars are entries of device arrays.
#pragma once

#include "mhg_kernels.hpp"

namespace mhg {

struct BatchProblem {
  Params P;
  long cb, nb;  // first cell, first node
  int slot;     // first CFL block-minimum slot
  int pad;
};

// One work item of a persistent kernel: block b of problem p's launch shape.
struct BatchItem {
  int p, b;
};

// Work lists, one per launch shape:
//   NODE     node phase, 256 nodes per block
//   CELL     cell phases, 256 cells per block
//   CX, NX   x sweeps of cells and nodes: 4 row segments of 60 outputs per block
//   CY_*     y cell sweep: SEG = whole columns in segments of ny + 4 lanes,
//            TALL = 16-column x 64-row tiles
//   NY_*     y dual sweep, the same with ny + 5 lanes
//   PROB     one item per problem (dt)
enum BatchList {
  BL_NODE = 0,
  BL_CELL,
  BL_CX,
  BL_NX,
  BL_CY_SEG,
  BL_CY_TALL,
  BL_NY_SEG,
  BL_NY_TALL,
  BL_PROB,
  BL_COUNT
};

// Device work lists and the persistent grid size of each list's kernels.
struct BatchLaunch {
  const BatchItem* items[BL_COUNT];
  const int* count;  // device array of BL_COUNT item counts
  int grid[BL_COUNT];
  int ne, lim, pc;   // shared by every problem: environments, limiter, centred pseudo
};

// Blocks of each launch shape for one problem (0 if the shape does not
// apply to its mesh).  Host and device agree on these through this function.
inline int batch_blocks(const Params& P, int list) {
  const int nx = P.nx, ny = P.ny;
  switch (list) {
    case BL_NODE: return (int)((P.nn + 255) / 256);
    case BL_CELL: return (int)((P.nc + 255) / 256);
    case BL_CX: return (int)(((long)((nx + 59) / 60) * ny + 3) / 4);
    case BL_NX: return (int)(((long)((nx + 1 + 59) / 60) * (ny + 1) + 3) / 4);
    case BL_CY_SEG: {
      const int sl = ny + 4;
      if (sl > 64) return 0;
      const int tw = 16 * (64 / sl);
      return (nx + tw - 1) / tw;
    }
    case BL_CY_TALL: return ((nx + 15) / 16) * ((ny + 59) / 60);
    case BL_NY_SEG: {
      const int sl = ny + 5;
      if (sl > 64) return 0;
      const int tw = 16 * (64 / sl);
      return (nx + 1 + tw - 1) / tw;
    }
    case BL_NY_TALL: return ((nx + 1 + 15) / 16) * ((ny + 1 + 59) / 60);
    case BL_PROB: return 1;
    default: return 0;
  }
}

template <class R>
void launch_batch_node_lagrange_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                   const NodeLagArgsT<R>& A, const BatchLaunch& L);
template <class R>
void launch_batch_cell_lagrange_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                   const CellLagArgsT<R>& A, const BatchLaunch& L);
// y sweeps launch the SEG and TALL kernels back to back (either list may be empty).
template <class R>
void launch_batch_cell_sweep_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const CellSweepArgsT<R>& A, const BatchLaunch& L, int dir, bool first);
template <class R>
void launch_batch_node_sweep_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const NodeSweepArgsT<R>& A, const BatchLaunch& L, int dir, bool first);
// Final cell phase, then the dt of each problem (time += dt, iteration += 1).
template <class R>
void launch_batch_cell_final_2d(hipStream_t st, const BatchProblem* probs, DevScalars* S,
                                const CellFinalArgsT<R>& A, const BatchLaunch& L);

}  // namespace mhg
