// mhg_ctx.hpp -- host-side context of a mesh (private to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/mahyco_gpu.h"
#include "mhg_common.hpp"
#include "mhg_kernels.hpp"

namespace mhg {
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
}  // namespace mhg

struct mhg_ctx {
  mhg_options opt{};
  mhg::Params P{};
  std::string err;
  bool initialised = false;
  hipStream_t stream = nullptr;
  int device = 0;
  std::vector<mhg::DBuf> owned;

  // Bulk fields hold `bulk`-byte words: double for the drop-in, float for the
  // precision form (mhg_options.precision).  Coordinates stay double.
  int bulk = 8;
  // cells
  void *rho = nullptr, *e = nullptr, *p = nullptr, *cs = nullptr, *q = nullptr, *m = nullptr;
  void *L = nullptr, *Vlag = nullptr;
  int* mask = nullptr;
  void* U[2] = {nullptr, nullptr};
  void* contrib = nullptr;
  // environments [ne][nc]
  void *env_rho = nullptr, *env_e = nullptr, *env_p = nullptr, *env_cs = nullptr;
  void *env_q = nullptr, *env_m = nullptr, *env_f = nullptr, *env_mf = nullptr;
  // nodes
  void* ux[2] = {nullptr, nullptr};
  void* uy[2] = {nullptr, nullptr};
  double* xx[2] = {nullptr, nullptr};
  double* xy[2] = {nullptr, nullptr};
  void *Ud0 = nullptr, *Ud1 = nullptr, *Ud3 = nullptr, *dg0 = nullptr, *dg1 = nullptr;
  void *Ud4 = nullptr, *dek = nullptr;  // total-energy fix-up only
  void* UdB4 = nullptr;                 // its 3D second dual state
  int* est = nullptr;                   // order 3, generic sweeps: the cells' mixed / pure codes
  // generic sweeps (pente-borne / -mixte): phi kept between iterations
  // (pente-borne), the back / front face fluxes or the gradients [3 ne + 1][nc]
  double *phi = nullptr, *dar = nullptr, *dav = nullptr;
  // 3D: z components, a second dual state (the middle sweep reads one and
  // writes the other), the Euler volumes of the hexahedra
  double* xz[2] = {nullptr, nullptr};
  void* uz[2] = {nullptr, nullptr};
  void *Ud2 = nullptr, *dg2 = nullptr;
  void* UdB[4] = {nullptr, nullptr, nullptr, nullptr};
  double* Veuler = nullptr;
  // learned limiter: coefficients [2][nc] in the storage type (MHG_LIM_LEARNED)
  void* alpha = nullptr;
  // attached network (mhg_set_limiter_mlp): FP32 parameters and their BF16 copy
  float* mlp_prm = nullptr;
  unsigned short* mlp_wb = nullptr;
  bool use_mlp = false;
  double iface_alpha = -1.0;  // coefficient of material-interface cells (< 0: classical limiter)
  double* block_min = nullptr;
  mhg::DevScalars* S = nullptr;
  mhg::DevScalars* S_host = nullptr;  // pinned mirror
  // Second device copy of the scalars: a fused 2D step that folds the
  // previous step's dt kernel into its node phase reads S and leaves the
  // updated scalars here, and the two pointers swap.
  mhg::DevScalars* S_alt = nullptr;
  // The last enqueued fused step left its CFL minima (dt_pending_nb of them
  // in block_min) for the next step's node phase; false whenever control
  // returns to the caller.
  bool dt_pending = false;
  int dt_pending_nb = 0;

  // host mirrors of the step bookkeeping
  int xcur = 0;  // buffer with the coordinates of the last Lagrange phase
  // 2D fused steps whose kernels all form node coordinates as x0 + dt u do
  // not store them: xx[xcur], xy[xcur] are then stale, and the last Lagrange
  // phase's coordinates are x0 + dt_old u[1 - ucur] (materialize_coords
  // writes them where a caller needs the arrays)
  bool x_lazy = false;
  bool x_lazy_before = false;  // at the start of the last stepping call
  int ucur = 0;  // buffer with the current velocity
  int sens = 0;  // SensProjection
  // launch variants (mhg_autotune); the defaults are the ones it picks on
  // the triple point: the LDS tiles along y, the lane form along x
  int yvariant = 4;   // y cell sweep
  int nxvariant = 1;  // x dual sweep
  int nyvariant = 4;  // y dual sweep
  int fuse_final = 1; // fused step: the last cell sweep carries the final cell phase
  int fuse_lag = 1;   // fused step: the first cell sweep carries the Lagrange cell phase

  // per-entry-point path: the _n, tau and div u arrays (allocated on the
  // first entry-point call) and which coordinates are current between calls
  double *rho_n = nullptr, *e_n = nullptr, *q_n = nullptr, *tau = nullptr, *divu = nullptr;
  double *env_rho_n = nullptr, *env_e_n = nullptr, *env_q_n = nullptr, *env_tau = nullptr;
  bool ep_ready = false;
  int coord_phase = 0;  // 0: xx[xcur] (last Lagrange phase); 1: x0 (after
                        // saveValuesAtN); 2: xx[1 - xcur] (after updatePosition)
  bool ep_remapped = false;  // the remap of this iteration ran (cells remapped, dt not yet computed)
  int remap_stage = 0;       // 1: mhg_remap_prepare built the remap vectors of this iteration

  // captured groups of fused iterations, one per host buffer state
  struct StepGraph {
    int xcur, ucur, sens;
    bool lazy, lazy_end;  // x_lazy at the start and at the end of the group
    mhg::DevScalars *s0, *s_end;  // current scalars at the start and at the end
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  int use_graphs = 1;
  // mhg_run_async: fused iterations enqueued since the last host
  // synchronisation and the device iteration count before them; iterations
  // past the end of the run leave the device state alone, and the next call
  // that reads the host bookkeeping rolls it back to the iterations done
  int async_launched = 0;
  int async_it0 = 0;

  // kernel timing
  bool timing = false;
  std::vector<std::string> tnames;
  std::vector<double> tsum;
  std::vector<int> tcount;
  std::vector<hipEvent_t> evpool;
  std::vector<std::pair<int, int>> pending;  // (name index, event pair index)
  size_t evused = 0;

  // A problem of a batch (mhg_batch_problem): its arrays are slices of the
  // batch's arrays, its scalars an entry of the batch's scalar array; it
  // owns no memory and the batch advances it.
  mhg_batch* owner = nullptr;
  int owner_index = -1;
};


#define HIPCHK(ctx, x)                                                   \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);       \
      return MHG_E_HIP;                                                  \
    }                                                                    \
  } while (0)

namespace mhg {
int ctx_fail(mhg_ctx* c, int code, const std::string& msg);
// Device array of `count` elements of T (zeroed), owned by c.
int ctx_dalloc_bytes(mhg_ctx* c, void** out, size_t bytes);
template <class T>
int ctx_dalloc(mhg_ctx* c, T** out, size_t count) {
  void* p = nullptr;
  const int r = ctx_dalloc_bytes(c, &p, (count ? count : 1) * sizeof(T));
  *out = static_cast<T*>(p);
  return r;
}
// Validates a mesh and its options and fills the run-time parameters.
int build_params(const mhg_mesh_desc* mesh, const mhg_options* opt, Params* P);
// Brings the buffer state of a batch problem's context up to date with the
// steps the batch has run (defined with the batch).
int batch_refresh_view(mhg_ctx* c);
bool batch_running(const mhg_ctx* c);
// Drops the captured step graphs (launch arguments changed).
void ctx_drop_graphs(mhg_ctx* c);
// Features of the state and, with write_alpha, the coefficients of the
// attached network (defined with the network).
template <class R>
void enqueue_learned_alpha(mhg_ctx* c, float* Xout, float* Sout, bool write_alpha);
}  // namespace mhg

namespace mhg {

template <class R>
R* bp(void* p) {
  return static_cast<R*>(p);
}

// ---- argument blocks of the fused step for the context's buffer state ----
template <class R>
NodeLagArgsT<R> node_lag_args(mhg_ctx* c) {
  NodeLagArgsT<R> A;
  A.p = bp<R>(c->p);
  A.q = bp<R>(c->q);
  A.m = bp<R>(c->m);
  A.xpx = c->xx[c->xcur];
  A.xpy = c->xy[c->xcur];
  A.ux = bp<R>(c->ux[c->ucur]);
  A.uy = bp<R>(c->uy[c->ucur]);
  A.xnx = c->xx[1 - c->xcur];
  A.xny = c->xy[1 - c->xcur];
  return A;
}

template <class R>
CellLagArgsT<R> cell_lag_args(mhg_ctx* c) {
  CellLagArgsT<R> A;
  A.xnx = c->xx[1 - c->xcur];
  A.xny = c->xy[1 - c->xcur];
  A.rho = bp<R>(c->rho);
  A.e = bp<R>(c->e);
  A.p = bp<R>(c->p);
  A.cs = bp<R>(c->cs);
  A.q = bp<R>(c->q);
  A.m = bp<R>(c->m);
  A.mask = c->mask;
  A.env_f = bp<R>(c->env_f);
  A.env_rho = bp<R>(c->env_rho);
  A.env_m = bp<R>(c->env_m);
  A.env_q = bp<R>(c->env_q);
  A.env_p = bp<R>(c->env_p);
  A.env_e = bp<R>(c->env_e);
  A.U = bp<R>(c->U[0]);
  A.Vlag = bp<R>(c->Vlag);
  A.L = bp<R>(c->L);
  return A;
}

// Cell sweep of remap pass `pass` along idir (the passes alternate the U buffers).
template <class R>
CellSweepArgsT<R> cell_sweep_args(mhg_ctx* c, int pass, int idir) {
  CellSweepArgsT<R> A;
  A.Uin = bp<R>(c->U[pass & 1]);
  A.Uout = bp<R>(c->U[(pass + 1) & 1]);
  A.Vlag = bp<R>(c->Vlag);
  A.xnx = c->xx[1 - c->xcur];
  A.xny = c->xy[1 - c->xcur];
  A.ux = bp<R>(c->ux[c->ucur]);
  A.uy = bp<R>(c->uy[c->ucur]);
  A.contrib = bp<R>(c->contrib);
  A.alpha = c->alpha ? bp<R>(c->alpha) + (size_t)idir * c->P.nc : nullptr;
  A.est = c->est;
  return A;
}

template <class R>
NodeSweepArgsT<R> node_sweep_args(mhg_ctx* c) {
  NodeSweepArgsT<R> B{};
  B.m = bp<R>(c->m);
  B.ux = bp<R>(c->ux[c->ucur]);
  B.uy = bp<R>(c->uy[c->ucur]);
  B.xnx = c->xx[1 - c->xcur];
  B.xny = c->xy[1 - c->xcur];
  B.contrib = bp<R>(c->contrib);
  B.Ud0 = bp<R>(c->Ud0);
  B.Ud1 = bp<R>(c->Ud1);
  B.Ud3 = bp<R>(c->Ud3);
  B.dg0 = bp<R>(c->dg0);
  B.dg1 = bp<R>(c->dg1);
  B.uox = bp<R>(c->ux[1 - c->ucur]);
  B.uoy = bp<R>(c->uy[1 - c->ucur]);
  B.Ud4 = bp<R>(c->Ud4);
  B.dek = bp<R>(c->dek);
  return B;
}

template <class R>
CellFinalArgsT<R> cell_final_args(mhg_ctx* c) {
  CellFinalArgsT<R> F;
  F.U = bp<R>(c->U[c->P.dim & 1]);
  F.Vlag = bp<R>(c->Vlag);
  F.L = bp<R>(c->L);
  F.ux = bp<R>(c->ux[1 - c->ucur]);
  F.uy = bp<R>(c->uy[1 - c->ucur]);
  F.rho = bp<R>(c->rho);
  F.e = bp<R>(c->e);
  F.p = bp<R>(c->p);
  F.cs = bp<R>(c->cs);
  F.q = bp<R>(c->q);
  F.m = bp<R>(c->m);
  F.mask = c->mask;
  F.env_rho = bp<R>(c->env_rho);
  F.env_e = bp<R>(c->env_e);
  F.env_p = bp<R>(c->env_p);
  F.env_cs = bp<R>(c->env_cs);
  F.env_q = bp<R>(c->env_q);
  F.env_m = bp<R>(c->env_m);
  F.env_f = bp<R>(c->env_f);
  F.env_mf = bp<R>(c->env_mf);
  F.block_min = c->block_min;
  F.dek = bp<R>(c->dek);
  F.est = c->est;
  return F;
}

}  // namespace mhg
