// mhg_kernels.hpp -- kernel argument blocks and launchers.
#pragma once

#include "mhg_common.hpp"

namespace mhg {

struct NodeLagArgs {
  const double *p, *q, *m;    // cell pressure, pseudo-viscosity, mass at the start of the step
  const double *xpx, *xpy;    // node coordinates of the previous Lagrange phase
  double *ux, *uy;            // node velocity (updated in place)
  double *xnx, *xny;          // new Lagrangian coordinates
};

struct CellLagArgs {
  const double *xnx, *xny;
  const double *rho, *e, *p, *cs, *q, *m;
  const int* mask;
  const double *env_f, *env_rho, *env_m, *env_q, *env_p, *env_e;
  double *U, *Vlag, *L;
};

struct CellSweepArgs {
  const double* Uin;
  double* Uout;
  const double *Vlag, *xnx, *xny, *ux, *uy;
  double* contrib;
};

struct NodeSweepArgs {
  const double *m, *ux, *uy, *xnx, *xny, *contrib;
  double *Ud0, *Ud1, *Ud3, *dg0, *dg1, *uox, *uoy;
};

struct CellFinalArgs {
  const double* U;
  const double *Vlag, *L, *ux, *uy;
  double *rho, *e, *p, *cs, *q, *m;
  int* mask;
  double *env_rho, *env_e, *env_p, *env_cs, *env_q, *env_m, *env_f, *env_mf;
  double* block_min;  // one CFL minimum per block of k_cell_final_2d
};

int cell_final_blocks(const Params& P);
void launch_begin_iteration(hipStream_t st, DevScalars* S);
void launch_node_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeLagArgs& A,
                             bool vel, bool bc, bool pos);
void launch_cell_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const CellLagArgs& A);
void launch_cell_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgs& A,
                          int dir, bool first, int variant);
int cell_sweep_y_variants();
const char* cell_sweep_y_variant_name(int v);
void launch_node_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeSweepArgs& A,
                          int dir, bool first, int variant);
int node_sweep_variants(int dir);
const char* node_sweep_variant_name(int dir, int v);
// advance: the final dt kernel also applies time += dt, iteration += 1 of the
// iteration (fused step without a separate begin-iteration launch)
void launch_cell_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellFinalArgs& A,
                          bool advance);

}  // namespace mhg
