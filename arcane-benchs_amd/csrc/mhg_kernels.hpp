// mhg_kernels.hpp -- kernel argument blocks and launchers.
#pragma once

#include "mhg_common.hpp"

namespace mhg {

// Argument blocks of the fused step, templated on the storage type R of the
// bulk fields (double: the drop-in; float: the precision form).  Node
// coordinates, the dt scalars and the CFL block minima stay double in both.
template <class R>
struct NodeLagArgsT {
  const R *p, *q, *m;         // cell pressure, pseudo-viscosity, mass at the start of the step
  const double *xpx, *xpy;    // node coordinates of the previous Lagrange phase
  R *ux, *uy;                 // node velocity (updated in place)
  double *xnx, *xny;          // new Lagrangian coordinates
  // set: the previous Lagrange phase's coordinates are formed as x0 + dt_old
  // u from its Lagrangian velocity (2D fused steps of the Euler scheme, whose
  // coordinates are not stored); null: read from xpx, xpy
  const R *upx = nullptr, *upy = nullptr;
};

template <class R>
struct CellLagArgsT {
  const double *xnx, *xny;
  const R *rho, *e, *p, *cs, *q, *m;
  const int* mask;
  const R *env_f, *env_rho, *env_m, *env_q, *env_p, *env_e;
  R *U, *Vlag, *L;
};

template <class R>
struct CellSweepArgsT {
  const R* Uin;
  R* Uout;
  const R* Vlag;
  const double *xnx, *xny;
  const R *ux, *uy;
  R* contrib;
  const R* alpha;  // learned-limiter coefficients of the sweep direction (learned mode)
  const int* est;  // order 3: the cells' mixed / pure codes of the last remap
};

template <class R>
struct NodeSweepArgsT {
  const R* m;
  const R *ux, *uy;
  const double *xnx, *xny;
  const R* contrib;
  R *Ud0, *Ud1, *Ud3, *dg0, *dg1, *uox, *uoy;
  // total-energy fix-up (Params::etot): dual kinetic energy between the
  // sweeps, and the last sweep's per-node term of the cells' energy deposit
  R *Ud4, *dek;
};

template <class R>
struct CellFinalArgsT {
  const R* U;
  const R *Vlag, *L, *ux, *uy;
  R *rho, *e, *p, *cs, *q, *m;
  int* mask;
  R *env_rho, *env_e, *env_p, *env_cs, *env_q, *env_m, *env_f, *env_mf;
  double* block_min;  // one CFL minimum per block of k_cell_final_2d
  const R* dek;       // per-node terms of the total-energy fix-up (Params::etot)
  int* est;           // order 3: mixed / pure code of each cell (-2 mixed, else its environment, -1 none)
};

// 3D argument blocks: the 2D block (x and y components) plus the z
// components; the dual sweeps read one dual state and write the other.
template <class R>
struct NodeLag3ArgsT {
  NodeLagArgsT<R> b;
  const double* xpz;
  R* uz;
  double* xnz;
};

template <class R>
struct CellLag3ArgsT {
  CellLagArgsT<R> b;
  const double* xnz;
};

template <class R>
struct CellSweep3ArgsT {
  CellSweepArgsT<R> b;
  const double* xnz;
  const R* uz;
};

template <class R>
struct NodeSweep3ArgsT {
  const R* m;
  const R *ux, *uy, *uz;
  const double *xnx, *xny, *xnz;
  const R* contrib;
  const R *Udi0, *Udi1, *Udi2, *Udi3;  // dual state before the sweep
  R *Udo0, *Udo1, *Udo2, *Udo3;        // dual state after it
  R *dg0, *dg1, *dg2;                  // persistent nodal gradients
  R *uox, *uoy, *uoz;                  // velocity out (last sweep)
  int prevmask;                        // directions swept before this one
  // total-energy fix-up (Params::etot): dual kinetic energy before / after
  // the sweep, and the last sweep's per-node term of the cells' deposit
  const R* Udi4;
  R *Udo4, *dek;
};

template <class R>
struct CellFinal3ArgsT {
  CellFinalArgsT<R> b;
  const R* uz;
  const double* Veuler;
};

using NodeLagArgs = NodeLagArgsT<double>;
using CellLagArgs = CellLagArgsT<double>;
using CellSweepArgs = CellSweepArgsT<double>;
using NodeSweepArgs = NodeSweepArgsT<double>;
using CellFinalArgs = CellFinalArgsT<double>;

// Per-entry-point cell stages (the unfused path): each reads and writes the
// state arrays as the reference entry point of the same name does; the _n,
// tau and div u arrays exist only for this path.
struct StageArgs {
  const double *xnx, *xny;  // Lagrangian coordinates of this iteration
  double *rho, *e, *p, *cs, *q, *m, *L, *Vlag;
  const int* mask;
  const double *env_f, *env_mf, *env_m, *env_p;
  double *env_rho, *env_e, *env_q;
  double *rho_n, *e_n, *q_n, *tau, *divu;
  double *env_rho_n, *env_e_n, *env_q_n, *env_tau;
  double* U;
  double* phi;  // pente-borne: [3 ne + 1][nc] phi of the environments present (BUILD_U)
};
enum CellStage {
  STAGE_SAVE = 0,       // saveValuesAtN (cell part)
  STAGE_GEOMETRY = 1,   // computeGeometricValues: V, L
  STAGE_DENSITY = 2,    // updateDensity: rho, tau, div u
  STAGE_VISCOSITY = 3,  // computeArtificialViscosity: q
  STAGE_ENERGY = 4,     // updateEnergyAndPressure: e
  STAGE_BUILD_U = 5     // computeVariablesForRemap (cells)
};
void launch_cell_stage_2d(hipStream_t st, const Params& P, DevScalars* S, const StageArgs& A, int stage);
// Generic cell sweep of the pente-borne / -mixte options (FP64, 2D and 3D):
// k_generic_grad then k_generic_update along dir (mhg_generic.hip).
struct GenericSweepArgs {
  const double *x, *y, *z;     // Lagrangian coordinates of this iteration (z: 3D)
  const double *ux, *uy, *uz;  // node velocities at the start of the remap
  const double* Uin;
  double* Uout;
  const double* Vlag;
  double* phi;        // pente-borne: [3 ne + 1][nc], kept between iterations
  double *dar, *dav;  // pente-borne back / front face fluxes; -mixte alone: dar = gradients
  int* est;           // mixed (-2) / pure (environment, -1 none) code of each cell
  double* contrib;    // dual mass contributions of the node sweep
};
void launch_generic_sweep(hipStream_t st, const Params& P, DevScalars* S, const GenericSweepArgs& A, int dir,
                          bool first);
// computeGeometricValues of hexahedra (MahycoModule.cc:1054-1299): V and L
// from the coordinates of this iteration.
void launch_cell_geometry_3d(hipStream_t st, const Params& P, DevScalars* S, const double* xnx,
                             const double* xny, const double* xnz, double* Vlag, double* L);

// IEquationOfState entry points (per environment) and computePressionMoyenne.
struct EosArgs {
  double *rho, *e, *p, *cs;
  const int* mask;
  double *env_rho, *env_e, *env_p, *env_cs;
};
enum EosMode { EOS_APPLY = 0, EOS_INIT = 1 };
// applyEOS (p, c from rho, e) or initEOS (e, c from rho, p) of environment k
// on the cells [c0, c1) that hold it: the cell arrays for a cell holding k
// alone, the environment arrays for a cell holding several.
void launch_eos_env(hipStream_t st, const Params& P, DevScalars* S, const EosArgs& A, int k, int mode, long c0,
                    long c1);
struct PmArgs {
  double *p, *cs;
  const int* mask;
  const double *env_f, *env_p, *env_cs;
};
void launch_pression_moyenne(hipStream_t st, const Params& P, DevScalars* S, const PmArgs& A);
// CFL candidates of computeHydroDeltaT (MahycoModule.cc:2027-2071), one
// minimum per block of 256 cells into block_min (cell_final_blocks slots):
// L / (c + max over the cell's nodes of |u|), without the velocity term in
// the Lagrange-only mode.
template <class R>
struct CflArgsT {
  const R *L, *cs, *ux, *uy, *uz;
  double* block_min;
};
using CflArgs = CflArgsT<double>;
void launch_cfl(hipStream_t st, const Params& P, DevScalars* S, const CflArgs& A);
// The fused step's CFL candidates of a 2D mesh in the storage type R.
template <class R>
void launch_cfl_2d(hipStream_t st, const Params& P, DevScalars* S, const CflArgsT<R>& A);

// The last cell sweep of the remap (dir 0: x, 1: y) fused with the final
// cell phase (remapVariables, EOS, pressure average; no CFL candidates),
// for the configurations fused_final_supported accepts: classical limiters
// at order 1 or 2, no learned coefficients, no total-energy fix-up (its
// deposit comes from the dual sweep that follows), at most 4 environments.
bool fused_final_supported(const Params& P);
// The first cell sweep of the remap fused with the Lagrange cell phase
// (k_cell_lagrange_2d: V, L stored, the remap vector kept in registers), for
// the configurations of fused_final_supported without pseudo-centring.
bool fused_lag_supported(const Params& P);
template <class R>
void launch_cell_sweep_lag_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                              const CellLagArgsT<R>& LA, int dir);
template <class R>
void launch_cell_sweep_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                                const CellFinalArgsT<R>& F, int dir);
// nb: the minima in block_min (<= 0: cell_final_blocks, k_cell_final_2d's;
// the fused step's k_cfl_2d leaves cfl_blocks of them)
void launch_dt_finalize(hipStream_t st, const Params& P, DevScalars* S, const double* block_min, bool advance,
                        int nb = 0);
int cfl_blocks(const Params& P);

int cell_final_blocks(const Params& P);
void launch_begin_iteration(hipStream_t st, DevScalars* S);
// Node coordinates x0 + dt_old u of the last Lagrange phase into xx, xy.
template <class R>
void launch_lag_coords_2d(hipStream_t st, const Params& P, DevScalars* S, const R* ux, const R* uy, double* xx,
                          double* xy);
// The node phase with the previous step's dt kernel folded in: reads the
// scalars Sin and the nb block minima, writes the updated scalars to Sout.
template <class R>
void launch_node_lagrange_fold_2d(hipStream_t st, const Params& P, const DevScalars* Sin, DevScalars* Sout,
                                  const NodeLagArgsT<R>& A, bool pos, const double* block_min, int nb);
template <class R>
void launch_node_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeLagArgsT<R>& A,
                             bool vel, bool bc, bool pos);
template <class R>
void launch_cell_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const CellLagArgsT<R>& A);
// Launch variants other than the LDS tile and lane forms exist for double only;
// float contexts always run those two.
template <class R>
void launch_cell_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                          int dir, bool first, int variant);
int cell_sweep_y_variants();
const char* cell_sweep_y_variant_name(int v);
template <class R>
void launch_node_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeSweepArgsT<R>& A,
                          int dir, bool first, int variant);
int node_sweep_variants(int dir);
const char* node_sweep_variant_name(int dir, int v);
// dt_mode 0: no dt kernel (the caller computes dt later); 1: the final dt
// kernel follows; 2: it also applies time += dt, iteration += 1 of the
// iteration (fused step without a separate begin-iteration launch)
// eos = false: remapVariables only (no EOS, pressure average or CFL
// candidates; dt_mode must be 0).
template <class R>
void launch_cell_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellFinalArgsT<R>& A,
                          int dt_mode, bool eos = true);

// 3D fused step.
template <class R>
void launch_node_lagrange_3d(hipStream_t st, const Params& P, DevScalars* S, const NodeLag3ArgsT<R>& A, bool vel,
                             bool bc, bool pos);
template <class R>
void launch_cell_lagrange_3d(hipStream_t st, const Params& P, DevScalars* S, const CellLag3ArgsT<R>& A);
template <class R>
void launch_cell_sweep_3d(hipStream_t st, const Params& P, DevScalars* S, const CellSweep3ArgsT<R>& A, int dir,
                          bool first);
template <class R>
void launch_node_sweep_3d(hipStream_t st, const Params& P, DevScalars* S, const NodeSweep3ArgsT<R>& A, int dir,
                          bool first, bool last);
template <class R>
void launch_cell_final_3d(hipStream_t st, const Params& P, DevScalars* S, const CellFinal3ArgsT<R>& A, int dt_mode,
                          bool eos = true);

}  // namespace mhg
