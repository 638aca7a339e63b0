// mhg_host.cpp -- host side of the MI355X MaHyCo library: context, device
// memory, initial states, the C-ABI of include/mahyco_gpu.h and the launch
// sequence of one time step.
//
// The host objects mirror the reference's service structure:
//   MahycoModule entry points      -> mhg_* compute-loop calls / step()
//   IRemap (RemapADI)              -> Remap::appli() (sweep order, SensProjection)
//   IEquationOfState               -> EosParams (per environment, on the device)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mhg_common.hpp"
#include "mhg_ctx.hpp"
#include "mhg_kernels.hpp"

using namespace mhg;

namespace mhg {

int ctx_fail(mhg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int ctx_dalloc_bytes(mhg_ctx* c, void** out, size_t bytes) {
  void* p = nullptr;
  *out = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return ctx_fail(c, MHG_E_ALLOC, "hipMalloc failed");
  if (hipMemset(p, 0, bytes) != hipSuccess) {
    hipFree(p);
    return ctx_fail(c, MHG_E_HIP, "hipMemset failed");
  }
  c->owned.push_back({p, bytes});
  *out = p;
  return MHG_OK;
}

}  // namespace mhg

namespace {

int fail(mhg_ctx* c, int code, const std::string& msg) { return ctx_fail(c, code, msg); }

template <class T>
int dalloc(mhg_ctx* c, T** out, size_t count) {
  return ctx_dalloc(c, out, count);
}

// Bulk array of `count` words of the context's storage type.
int balloc(mhg_ctx* c, void** out, size_t count) {
  return ctx_dalloc_bytes(c, out, std::max<size_t>(count, 1) * (size_t)c->bulk);
}


// Host doubles to a device bulk array (rounded to float in the precision form).
hipError_t put(mhg_ctx* c, void* d, const double* h, size_t n) {
  if (!h) return hipSuccess;
  if (c->bulk == 8) return hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice);
  std::vector<float> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
  return hipMemcpy(d, t.data(), n * sizeof(float), hipMemcpyHostToDevice);
}

// A device bulk array to host doubles.
hipError_t get(mhg_ctx* c, double* h, const void* d, size_t n) {
  if (!h) return hipSuccess;
  if (c->bulk == 8) return hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost);
  std::vector<float> t(n);
  hipError_t e = hipMemcpy(t.data(), d, n * sizeof(float), hipMemcpyDeviceToHost);
  for (size_t i = 0; i < n; ++i) h[i] = t[i];
  return e;
}

double x0_of(const Params& P, int d, int i) { return P.org[d] + i * P.h[d]; }

// Node-index range selected by PrepareFaceGroup's coordinate test
// (MahycoAnnexe.cc:183-251) along one axis; returns false if not a range.
bool surface_range(const Params& P, int surf, int* axis, int* lo, int* hi, double thr) {
  const int a = surf / 2;
  const bool is_max = surf % 2;
  const int nnodes[3] = {P.nx + 1, P.ny + 1, P.nz + 1};
  if (a >= P.dim) {  // 2D: every node has z = 0, so ZMIN and ZMAX hold every face
    *axis = -1;
    *lo = 0;
    *hi = 0;
    return true;
  }
  double mx = -1.;
  for (int i = 0; i < nnodes[a]; ++i) mx = std::max(mx, x0_of(P, a, i));
  int first = -1, last = -2;
  bool gap = false;
  for (int i = 0; i < nnodes[a]; ++i) {
    const double x = x0_of(P, a, i);
    const bool in = is_max ? (std::fabs(x - mx) <= thr) : (x <= thr);
    if (in) {
      if (first < 0) first = i;
      else if (last != i - 1) gap = true;
      last = i;
    }
  }
  if (gap) return false;
  *axis = a;
  *lo = first < 0 ? 1 : first;
  *hi = first < 0 ? 0 : last;
  return true;
}

int sync(mhg_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MHG_OK;
}

int read_scalars(mhg_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->S_host, c->S, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MHG_OK;
}

int write_scalars(mhg_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->S, c->S_host, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MHG_OK;
}

// ---- kernel timing -----------------------------------------------------
int tname_index(mhg_ctx* c, const char* name) {
  for (size_t i = 0; i < c->tnames.size(); ++i)
    if (c->tnames[i] == name) return (int)i;
  c->tnames.push_back(name);
  c->tsum.push_back(0.);
  c->tcount.push_back(0);
  return (int)c->tnames.size() - 1;
}

struct TimedLaunch {
  mhg_ctx* c;
  int name = -1;
  size_t ev = 0;
  TimedLaunch(mhg_ctx* cc, const char* nm) : c(cc) {
    if (!c->timing) return;
    name = tname_index(c, nm);
    if (c->evused + 2 > c->evpool.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        hipEventCreate(&e);
        c->evpool.push_back(e);
      }
    }
    ev = c->evused;
    c->evused += 2;
    hipEventRecord(c->evpool[ev], c->stream);
  }
  ~TimedLaunch() {
    if (!c->timing) return;
    hipEventRecord(c->evpool[ev + 1], c->stream);
    c->pending.push_back({name, (int)ev});
  }
};

void collect_times(mhg_ctx* c) {
  if (c->pending.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& pr : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->evpool[pr.second], c->evpool[pr.second + 1]);
    c->tsum[pr.first] += ms;
    c->tcount[pr.first] += 1;
  }
  c->pending.clear();
  c->evused = 0;
}

// Remap (IRemap::appliRemap, RemapADIService.cc:19-60): for i < dim,
// idir = (i + SensProjection) % dim, cell sweep then dual sweep; then
// SensProjection = (SensProjection + 1) % dim and remapVariables.
// In the fused step (dt_mode 2) the last cell sweep carries the final cell
// phase (fused_final_supported), and the CFL candidates follow the last dual
// sweep, whose velocities they read.
// lag: the first cell sweep carries the Lagrange cell phase (the caller
// launched no k_cell_lagrange_2d).
// Arguments of the generic sweeps (FP64) of remap pass `pass`.
GenericSweepArgs generic_args(mhg_ctx* c, int pass) {
  GenericSweepArgs A;
  const bool d3 = c->P.dim == 3;
  A.x = c->xx[1 - c->xcur];
  A.y = c->xy[1 - c->xcur];
  A.z = d3 ? c->xz[1 - c->xcur] : nullptr;
  A.ux = bp<double>(c->ux[c->ucur]);
  A.uy = bp<double>(c->uy[c->ucur]);
  A.uz = d3 ? bp<double>(c->uz[c->ucur]) : nullptr;
  A.Uin = bp<double>(c->U[pass & 1]);
  A.Uout = bp<double>(c->U[(pass + 1) & 1]);
  A.Vlag = bp<double>(c->Vlag);
  A.phi = c->phi;
  A.dar = c->dar;
  A.dav = c->dav;
  A.est = c->est;
  A.contrib = bp<double>(c->contrib);
  return A;
}

// dt_mode 3: as 2, with the dt kernel left to the next step's node phase
// (c->dt_pending) when the step is fused.
template <class R>
void remap_2d(mhg_ctx* c, int dt_mode, bool eos = true, bool lag = false) {
  const Params& P = c->P;
  const bool fuse = (dt_mode == 2 || dt_mode == 3) && eos && c->fuse_final && fused_final_supported(P);
  if (dt_mode == 3 && !fuse) dt_mode = 2;
  for (int pass = 0; pass < P.dim; ++pass) {
    const int idir = (pass + c->sens) % P.dim;
    const bool first = (pass == 0);
    if (P.generic) {  // pente-borne / -mixte (FP64 contexts)
      TimedLaunch t(c, idir == 0 ? "cell_sweep_x" : "cell_sweep_y");
      launch_generic_sweep(c->stream, P, c->S, generic_args(c, pass), idir, first);
    } else if (lag && first) {
      TimedLaunch t(c, idir == 0 ? "cell_sweep_x_lag" : "cell_sweep_y_lag");
      launch_cell_sweep_lag_2d(c->stream, P, c->S, cell_sweep_args<R>(c, pass, idir), cell_lag_args<R>(c), idir);
    } else if (fuse && pass == P.dim - 1) {
      TimedLaunch t(c, idir == 0 ? "cell_sweep_x_final" : "cell_sweep_y_final");
      launch_cell_sweep_final_2d(c->stream, P, c->S, cell_sweep_args<R>(c, pass, idir), cell_final_args<R>(c), idir);
    } else {
      TimedLaunch t(c, idir == 0 ? "cell_sweep_x" : "cell_sweep_y");
      launch_cell_sweep_2d(c->stream, P, c->S, cell_sweep_args<R>(c, pass, idir), idir, first, c->yvariant);
    }
    {
      TimedLaunch t(c, idir == 0 ? "node_sweep_x" : "node_sweep_y");
      launch_node_sweep_2d(c->stream, P, c->S, node_sweep_args<R>(c), idir, first,
                           idir == 0 ? c->nxvariant : c->nyvariant);
    }
  }
  if (fuse) {
    TimedLaunch t(c, "cfl");
    CflArgsT<R> A;
    A.L = bp<R>(c->L);
    A.cs = bp<R>(c->cs);
    A.ux = bp<R>(c->ux[1 - c->ucur]);
    A.uy = bp<R>(c->uy[1 - c->ucur]);
    A.uz = nullptr;
    A.block_min = c->block_min;
    launch_cfl_2d(c->stream, P, c->S, A);
    if (dt_mode == 3) {
      c->dt_pending = true;
      c->dt_pending_nb = cfl_blocks(P);
    } else {
      launch_dt_finalize(c->stream, P, c->S, c->block_min, true, cfl_blocks(P));
    }
  } else if (!fuse) {
    TimedLaunch t(c, "cell_final");
    launch_cell_final_2d(c->stream, P, c->S, cell_final_args<R>(c), dt_mode, eos);
  }
  c->sens = (c->sens + 1) % P.dim;
  c->xcur = 1 - c->xcur;
  c->ucur = 1 - c->ucur;
}

// One fused iteration, enqueued on the context stream.  The time and iteration
// advance of the iteration start is applied by the final dt kernel.
// last: the step ends a sequence the caller sees, so its dt kernel runs;
// otherwise a fused step leaves it to the next step's node phase.
template <class R>
void enqueue_step_t(mhg_ctx* c, bool last) {
  if (c->use_mlp) {
    TimedLaunch t(c, "learned_alpha");
    enqueue_learned_alpha<R>(c, nullptr, nullptr, true);
  }
  const bool lag = c->fuse_lag && fused_lag_supported(c->P);
  // the node coordinates are stored only if a kernel of this step reads them:
  // the fused cell sweeps and the lane / tile dual sweeps form them from the
  // velocities, the separate cell phase and the FP64 column-walker and
  // stencil dual sweeps read them
  const bool formed = lag && c->fuse_final && fused_final_supported(c->P) &&
                      (std::is_same<R, float>::value || (c->nxvariant == 1 && c->nyvariant == 4));
  {
    TimedLaunch t(c, "node_lagrange");
    NodeLagArgsT<R> A = node_lag_args<R>(c);
    if (c->x_lazy) {  // the previous coordinates from the previous Lagrangian velocity
      A.upx = bp<R>(c->ux[1 - c->ucur]);
      A.upy = bp<R>(c->uy[1 - c->ucur]);
    }
    if (c->dt_pending) {
      launch_node_lagrange_fold_2d(c->stream, c->P, c->S, c->S_alt, A, !formed, c->block_min, c->dt_pending_nb);
      std::swap(c->S, c->S_alt);
      c->dt_pending = false;
    } else {
      launch_node_lagrange_2d(c->stream, c->P, c->S, A, true, true, !formed);
    }
  }
  if (!lag) {
    TimedLaunch t(c, "cell_lagrange");
    launch_cell_lagrange_2d(c->stream, c->P, c->S, cell_lag_args<R>(c));
  }
  // the learned coefficients' kernel, which starts the next step, reads the
  // scalars before any node phase: no deferred dt then
  remap_2d<R>(c, (last || c->use_mlp) ? 2 : 3, true, lag);
  c->x_lazy = formed;
}

// Writes the last Lagrange phase's node coordinates, x0 + dt_old u[1 - ucur],
// into xx[xcur], xy[xcur] after fused steps that did not store them (the
// values the node phase would have stored, bit for bit).
template <class R>
int materialize_coords_t(mhg_ctx* c) {
  launch_lag_coords_2d(c->stream, c->P, c->S, bp<R>(c->ux[1 - c->ucur]), bp<R>(c->uy[1 - c->ucur]),
                       c->xx[c->xcur], c->xy[c->xcur]);
  HIPCHK(c, hipGetLastError());
  c->x_lazy = false;
  return MHG_OK;
}
int materialize_coords(mhg_ctx* c) {
  if (!c->x_lazy) return MHG_OK;
  return c->bulk == 4 ? materialize_coords_t<float>(c) : materialize_coords_t<double>(c);
}

// 3D remap: three cell sweeps and three dual sweeps; the dual state goes
// velocity -> Ud -> UdB -> velocity.
// Final cell phase arguments of a 3D context (remapped velocity in the
// buffer the last dual sweep writes).
template <class R>
CellFinal3ArgsT<R> cell_final3_args(mhg_ctx* c, int unew) {
  CellFinal3ArgsT<R> F{};
  F.b.U = bp<R>(c->U[1]);
  F.b.est = c->est;
  F.b.Vlag = bp<R>(c->Vlag);
  F.b.L = bp<R>(c->L);
  F.b.ux = bp<R>(c->ux[unew]);
  F.b.uy = bp<R>(c->uy[unew]);
  F.uz = bp<R>(c->uz[unew]);
  F.b.dek = bp<R>(c->dek);
  F.b.rho = bp<R>(c->rho);
  F.b.e = bp<R>(c->e);
  F.b.p = bp<R>(c->p);
  F.b.cs = bp<R>(c->cs);
  F.b.q = bp<R>(c->q);
  F.b.m = bp<R>(c->m);
  F.b.mask = c->mask;
  F.b.env_rho = bp<R>(c->env_rho);
  F.b.env_e = bp<R>(c->env_e);
  F.b.env_p = bp<R>(c->env_p);
  F.b.env_cs = bp<R>(c->env_cs);
  F.b.env_q = bp<R>(c->env_q);
  F.b.env_m = bp<R>(c->env_m);
  F.b.env_f = bp<R>(c->env_f);
  F.b.env_mf = bp<R>(c->env_mf);
  F.b.block_min = c->block_min;
  F.Veuler = c->Veuler;
  return F;
}

template <class R>
void remap_3d(mhg_ctx* c, int dt_mode, bool eos = true) {
  const Params& P = c->P;
  const double* xn[3] = {c->xx[1 - c->xcur], c->xy[1 - c->xcur], c->xz[1 - c->xcur]};
  const R* u[3] = {bp<R>(c->ux[c->ucur]), bp<R>(c->uy[c->ucur]), bp<R>(c->uz[c->ucur])};
  void* UdA[4] = {c->Ud0, c->Ud1, c->Ud2, c->Ud3};
  int prevmask = 0;
  const char* cname[3] = {"cell_sweep_x", "cell_sweep_y", "cell_sweep_z"};
  const char* nname[3] = {"node_sweep_x", "node_sweep_y", "node_sweep_z"};
  for (int pass = 0; pass < 3; ++pass) {
    const int idir = (pass + c->sens) % 3;
    const bool first = pass == 0, last = pass == 2;
    CellSweep3ArgsT<R> A{};
    A.b.Uin = bp<R>(c->U[pass & 1]);
    A.b.est = c->est;
    A.b.Uout = bp<R>(c->U[(pass + 1) & 1]);
    A.b.Vlag = bp<R>(c->Vlag);
    A.b.xnx = xn[0];
    A.b.xny = xn[1];
    A.xnz = xn[2];
    A.b.ux = u[0];
    A.b.uy = u[1];
    A.uz = u[2];
    A.b.contrib = bp<R>(c->contrib);
    {
      TimedLaunch t(c, cname[idir]);
      if (P.generic) launch_generic_sweep(c->stream, P, c->S, generic_args(c, pass), idir, first);
      else launch_cell_sweep_3d(c->stream, P, c->S, A, idir, first);
    }
    NodeSweep3ArgsT<R> B;
    B.m = bp<R>(c->m);
    B.ux = u[0];
    B.uy = u[1];
    B.uz = u[2];
    B.xnx = xn[0];
    B.xny = xn[1];
    B.xnz = xn[2];
    B.contrib = bp<R>(c->contrib);
    void** in = (pass == 2) ? c->UdB : UdA;
    void** outb = (pass == 0) ? UdA : c->UdB;
    B.Udi0 = bp<R>(in[0]);
    B.Udi1 = bp<R>(in[1]);
    B.Udi2 = bp<R>(in[2]);
    B.Udi3 = bp<R>(in[3]);
    B.Udo0 = bp<R>(outb[0]);
    B.Udo1 = bp<R>(outb[1]);
    B.Udo2 = bp<R>(outb[2]);
    B.Udo3 = bp<R>(outb[3]);
    B.dg0 = bp<R>(c->dg0);
    B.dg1 = bp<R>(c->dg1);
    B.dg2 = bp<R>(c->dg2);
    B.uox = bp<R>(c->ux[1 - c->ucur]);
    B.uoy = bp<R>(c->uy[1 - c->ucur]);
    B.uoz = bp<R>(c->uz[1 - c->ucur]);
    B.prevmask = prevmask;
    B.Udi4 = bp<R>(pass == 2 ? c->UdB4 : c->Ud4);
    B.Udo4 = bp<R>(pass == 0 ? c->Ud4 : c->UdB4);
    B.dek = bp<R>(c->dek);
    {
      TimedLaunch t(c, nname[idir]);
      launch_node_sweep_3d(c->stream, P, c->S, B, idir, first, last);
    }
    prevmask |= 1 << idir;
  }
  {
    TimedLaunch t(c, "cell_final");
    launch_cell_final_3d(c->stream, P, c->S, cell_final3_args<R>(c, 1 - c->ucur), dt_mode, eos);
  }
  c->sens = (c->sens + 1) % 3;
  c->xcur = 1 - c->xcur;
  c->ucur = 1 - c->ucur;
}

template <class R>
void enqueue_step_3d(mhg_ctx* c) {
  NodeLag3ArgsT<R> N;
  N.b = node_lag_args<R>(c);
  N.xpz = c->xz[c->xcur];
  N.uz = bp<R>(c->uz[c->ucur]);
  N.xnz = c->xz[1 - c->xcur];
  {
    TimedLaunch t(c, "node_lagrange");
    launch_node_lagrange_3d(c->stream, c->P, c->S, N, true, true, true);
  }
  CellLag3ArgsT<R> L;
  L.b = cell_lag_args<R>(c);
  L.xnz = c->xz[1 - c->xcur];
  {
    TimedLaunch t(c, "cell_lagrange");
    launch_cell_lagrange_3d(c->stream, c->P, c->S, L);
  }
  remap_3d<R>(c, 2);
}

void enqueue_step(mhg_ctx* c, bool last) {
  if (c->P.dim == 3) {
    if (c->bulk == 4) enqueue_step_3d<float>(c);
    else enqueue_step_3d<double>(c);
    return;
  }
  if (c->bulk == 4) enqueue_step_t<float>(c, last);
  else enqueue_step_t<double>(c, last);
}

// Steps per captured graph: a multiple of the period of the host buffer state
// (xcur and ucur alternate, sens cycles through dim), so a replay leaves that
// state where it found it.
constexpr int kGraphSteps = 12;

void drop_graphs(mhg_ctx* c) {
  for (auto& g : c->graphs) hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}
}  // namespace
namespace mhg {
void ctx_drop_graphs(mhg_ctx* c) { drop_graphs(c); }
}  // namespace mhg
namespace {

// Enqueues n fused iterations.  Whole groups of kGraphSteps replay a graph
// captured once per buffer state (one launch for the group's kernels); the
// rest, and every step while kernel timing is on, launch directly.
// A sequence that ended early (a HIP error between its steps) may have left
// the last step's dt update to a node phase that never ran: run it now.
int flush_dt(mhg_ctx* c) {
  if (!c->dt_pending) return MHG_OK;
  launch_dt_finalize(c->stream, c->P, c->S, c->block_min, true, c->dt_pending_nb);
  c->dt_pending = false;
  HIPCHK(c, hipGetLastError());
  return MHG_OK;
}

int enqueue_steps(mhg_ctx* c, int n) {
  while (c->use_graphs && !c->timing && n >= kGraphSteps) {
    hipGraphExec_t exec = nullptr;
    bool lazy_end = false;
    mhg::DevScalars* s_end = nullptr;
    for (auto& g : c->graphs)
      if (g.xcur == c->xcur && g.ucur == c->ucur && g.sens == c->sens && g.lazy == c->x_lazy && g.s0 == c->S) {
        exec = g.exec;
        lazy_end = g.lazy_end;
        s_end = g.s_end;
      }
    if (!exec) {
      mhg_ctx::StepGraph g{c->xcur, c->ucur, c->sens, c->x_lazy, false, c->S, nullptr, nullptr};
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < kGraphSteps; ++s) enqueue_step(c, s == kGraphSteps - 1);
      hipGraph_t graph = nullptr;
      HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
      const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      HIPCHK(c, e);
      g.lazy_end = c->x_lazy;
      g.s_end = c->S;
      c->graphs.push_back(g);
      exec = g.exec;
      lazy_end = g.lazy_end;
      // the capture swapped the scalar buffers as a launch would: back to
      // the start, for the launch below
      if (c->S != g.s0) std::swap(c->S, c->S_alt);
      s_end = g.s_end;
    }
    HIPCHK(c, hipGraphLaunch(exec, c->stream));
    c->x_lazy = lazy_end;  // the host state the group's last step leaves
    if (c->S != s_end) std::swap(c->S, c->S_alt);
    n -= kGraphSteps;
  }
  for (int s = 0; s < n; ++s) enqueue_step(c, s == n - 1);
  HIPCHK(c, hipGetLastError());
  return MHG_OK;
}

// Undoes the host bookkeeping (coordinate and velocity buffers, sweep order)
// of `skipped` enqueued iterations that the device did not run because the
// run had ended.
void roll_back(mhg_ctx* c, int skipped, int done) {
  // no step ran: the coordinate arrays are as the call found them
  if (done == 0) c->x_lazy = c->x_lazy_before;
  for (int s = 0; s < skipped; ++s) {
    c->sens = (c->sens + c->P.dim - 1) % c->P.dim;
    c->xcur = 1 - c->xcur;
    c->ucur = 1 - c->ucur;
  }
}

// Waits for the iterations of mhg_run_async and brings the host bookkeeping
// to the iterations the device really did.
int reconcile_async(mhg_ctx* c) {
  if (c->async_launched == 0) return MHG_OK;
  const int launched = c->async_launched;
  c->async_launched = 0;
  int r = read_scalars(c);
  if (r) return r;
  collect_times(c);
  const int done = c->S_host->iteration - c->async_it0;
  roll_back(c, launched - done, done);
  return MHG_OK;
}

// The context is initialised and the host bookkeeping follows the device;
// with coords, the node coordinate arrays hold the last Lagrange phase's
// coordinates (every call but the fused steps, which form them).
int check_ready(mhg_ctx* c, bool coords = true) {
  if (!c) return MHG_E_ARG;
  if (!c->initialised) return fail(c, MHG_E_STATE, "context not initialised");
  const int r = reconcile_async(c);
  if (r || !coords) return r;
  return materialize_coords(c);
}

// A problem of a batch is advanced by mhg_batch_run only.
int check_not_batch(mhg_ctx* c) {
  if (c && c->owner) return fail(c, MHG_E_STATE, "problem of a batch: advance it with mhg_batch_run");
  return MHG_OK;
}

// The fused step starts from a state between iterations.
int check_fused(mhg_ctx* c) {
  int r = check_ready(c, false);
  if (r) return r;
  r = check_not_batch(c);
  if (r) return r;
  if (c->coord_phase != 0)
    return fail(c, MHG_E_STATE, "fused step inside an entry-point iteration: finish it with mhg_remap");
  return flush_dt(c);
}

// Pure-cell initial state (hydroStartInit, MahycoModule.cc:48-188, then
// computeCellMass 199-225): EOS init per environment, m = rho V.
// Volume of cell cc on the initial mesh by the corner-normal formula
// (the Euler volume the remap returns to).
double euler_volume(const Params& P, long cc) {
  if (P.dim == 2) {
    const int i = (int)(cc % P.nx), j = (int)(cc / P.nx);
    double x[4] = {x0_of(P, 0, i), x0_of(P, 0, i + 1), x0_of(P, 0, i + 1), x0_of(P, 0, i)};
    double y[4] = {x0_of(P, 1, j), x0_of(P, 1, j), x0_of(P, 1, j + 1), x0_of(P, 1, j + 1)};
    return quad_volume(x, y);
  }
  const int i = (int)(cc % P.nx), j = (int)((cc / P.nx) % P.ny), k = (int)(cc / ((long)P.nx * P.ny));
  V3 cn[8];
  for (int q = 0; q < 8; ++q) {
    const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
    const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
    const int dk = q >= 4 ? 1 : 0;
    cn[q] = {x0_of(P, 0, i + di), x0_of(P, 1, j + dj), x0_of(P, 2, k + dk)};
  }
  return hex_volume(cn);
}

int init_pure(mhg_ctx* c, const std::vector<int>& env, const std::vector<double>& rho,
              const std::vector<double>& pres, const double* vel) {
  if (batch_running(c)) return fail(c, MHG_E_STATE, "problem of a running batch: state fixed at the first run");
  const Params& P = c->P;
  const long NC = P.nc, NN = P.nn;
  std::vector<double> e(NC), cs(NC), m(NC), zero(NC, 0.);
  std::vector<int> mask(NC);
  for (long cc = 0; cc < NC; ++cc) {
    const int k = env[cc];
    mask[cc] = 1 << k;
    const double g = P.eos.gamma[k];
    // initEOS, PerfectGasEOSService.cc:20-31 / StiffenedGasEOSService.cc:10-22
    if (P.eos.kind[k] == MHG_EOS_STIFFENED_GAS) {
      const double t = P.eos.tension[k];
      e[cc] = (pres[cc] + (g * t)) / ((g - 1.) * rho[cc]);
      cs[cc] = std::sqrt((g / rho[cc]) * (pres[cc] + t));
    } else {
      e[cc] = pres[cc] / ((g - 1.) * rho[cc]);
      cs[cc] = std::sqrt(g * pres[cc] / rho[cc]);
    }
    const double V = euler_volume(P, cc);
    m[cc] = rho[cc] * V;
    m[cc] = 1. * m[cc];  // m_env = mass_fraction * m (199-225), mass_fraction = 1
  }
  std::vector<double> xs(NN), ys(NN), zs(NN, 0.), us(NN, 0.), vs(NN, 0.), ws(NN, 0.);
  for (long n = 0; n < NN; ++n) {
    const int i = (int)(n % (P.nx + 1)), j = (int)((n / (P.nx + 1)) % (P.ny + 1));
    const int k = (int)(n / ((long)(P.nx + 1) * (P.ny + 1)));
    xs[n] = x0_of(P, 0, i);
    ys[n] = x0_of(P, 1, j);
    if (P.dim == 3) zs[n] = x0_of(P, 2, k);
    if (vel) {
      us[n] = vel[n];
      vs[n] = vel[NN + n];
      if (P.dim == 3) ws[n] = vel[2 * NN + n];
    }
  }
  c->xcur = 0;
  c->ucur = 0;
  c->x_lazy = false;
  c->coord_phase = 0;
  c->ep_remapped = false;
  c->sens = 0;
  c->async_launched = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto up = [&](void* d, const std::vector<double>& h) { return put(c, d, h.data(), h.size()); };
  HIPCHK(c, up(c->rho, rho));
  HIPCHK(c, up(c->e, e));
  HIPCHK(c, up(c->p, pres));
  HIPCHK(c, up(c->cs, cs));
  HIPCHK(c, up(c->q, zero));
  HIPCHK(c, up(c->m, m));
  HIPCHK(c, hipMemcpy(c->mask, mask.data(), NC * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->xx[0], xs.data(), NN * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->xy[0], ys.data(), NN * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, up(c->ux[0], us));
  HIPCHK(c, up(c->uy[0], vs));
  HIPCHK(c, hipMemset(c->dg0, 0, NN * (size_t)c->bulk));
  HIPCHK(c, hipMemset(c->dg1, 0, NN * (size_t)c->bulk));
  if (P.dim == 3) {
    HIPCHK(c, hipMemcpy(c->xz[0], zs.data(), NN * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, up(c->uz[0], ws));
    HIPCHK(c, hipMemset(c->dg2, 0, NN * (size_t)c->bulk));
  }
  const size_t envb = (size_t)P.ne * NC * (size_t)c->bulk;
  void* envs[] = {c->env_rho, c->env_e, c->env_p, c->env_cs, c->env_q, c->env_m, c->env_f, c->env_mf};
  for (void* d : envs) HIPCHK(c, hipMemset(d, 0, envb));
  // the remap's cell variables as Arcane creates them: mixed / pure codes
  // zero, pente-borne's phi zero, a pure cell's fraction 1 (in its slot)
  if (c->est) HIPCHK(c, hipMemset(c->est, 0, NC * sizeof(int)));
  if (c->phi) HIPCHK(c, hipMemset(c->phi, 0, (size_t)P.nvu * NC * sizeof(double)));
  if (P.pb) {
    std::vector<double> f((size_t)P.ne * NC, 0.);
    for (long cc = 0; cc < NC; ++cc) f[(size_t)env[cc] * NC + cc] = 1.0;
    HIPCHK(c, up(c->env_f, f));
  }
  std::memset(c->S_host, 0, sizeof(DevScalars));
  c->S_host->dt = c->opt.deltat_init;
  int r = write_scalars(c);
  if (r) return r;
  c->initialised = true;
  return MHG_OK;
}

void cell_centre(const Params& P, long cc, double* xc, double* yc) {
  const int i = (int)(cc % P.nx), j = (int)(cc / P.nx);
  // 0.25 * sum of the 4 nodes in Arcane order (MahycoAnnexe.cc:107-112)
  *xc = 0.25 * (((x0_of(P, 0, i) + x0_of(P, 0, i + 1)) + x0_of(P, 0, i + 1)) + x0_of(P, 0, i));
  *yc = 0.25 * (((x0_of(P, 1, j) + x0_of(P, 1, j)) + x0_of(P, 1, j + 1)) + x0_of(P, 1, j + 1));
}

}  // namespace

namespace mhg {

// Mesh and options of mhg_create -> run-time parameters (validation included).
int build_params(const mhg_mesh_desc* mesh, const mhg_options* opt, Params* Pout) {
  if (!mesh || !opt || !Pout) return MHG_E_ARG;
  if (mesh->dim != 2 && mesh->dim != 3) return MHG_E_ARG;
  for (int d = 0; d < mesh->dim; ++d)
    if (mesh->n[d] < 1 || !(mesh->length[d] > 0.)) return MHG_E_ARG;
  // cell and node indices fit in 31 bits: the kernels split them into mesh
  // coordinates in 32-bit arithmetic
  double nodes = 1.;
  for (int d = 0; d < mesh->dim; ++d) nodes *= mesh->n[d] + 1.;
  if (nodes >= 2147483648.) return MHG_E_UNSUPPORTED;
  if (opt->nb_env < 1 || opt->nb_env > MHG_MAX_ENV || opt->nb_bc < 0 || opt->nb_bc > MHG_MAX_BC)
    return MHG_E_ARG;
  if (opt->schema_csts || opt->with_newton || opt->sans_lagrange ||
      !opt->is_euler_scheme ||
      (opt->projection_limiteur_id > MHG_LIM_ULTRABEEG && opt->projection_limiteur_id != MHG_LIM_LEARNED &&
       opt->projection_limiteur_id != MHG_LIM_LEARNED_ALL) ||
      opt->projection_limiteur_id < 0 || opt->ordre_projection < 1 || opt->ordre_projection > 3)
    return MHG_E_UNSUPPORTED;
  const bool learned = opt->projection_limiteur_id == MHG_LIM_LEARNED ||
                       opt->projection_limiteur_id == MHG_LIM_LEARNED_ALL;
  if (learned && mesh->dim != 2) return MHG_E_UNSUPPORTED;
  // order 3 (Remap-ordre-3.cc) runs with the classical or generalised
  // limiters
  if (opt->ordre_projection == 3 && learned) return MHG_E_UNSUPPORTED;
  if (opt->precision != MHG_PREC_FP64 && opt->precision != MHG_PREC_FP32_BULK) return MHG_E_UNSUPPORTED;
  // the Lagrange-only mode (with-projection = false) runs the FP64
  // entry-point kernels
  if (!opt->with_projection && (opt->precision != MHG_PREC_FP64 || learned)) return MHG_E_UNSUPPORTED;
  // pente-borne (order 2: at order 1 the reference's generic face loop moves
  // nothing, at order 3 it mixes a face value with the integrated-flux
  // formula) and -mixte: classical or generalised limiters, pure limiter
  // 0..7, debar fix 0..2; the generic sweeps run in FP64 through the
  // entry-point kernels
  const bool pb = opt->projection_pente_borne != 0, mixte = opt->projection_pente_borne_mixte != 0;
  if ((pb || mixte) && (learned || opt->projection_limiteur_pure_id < 0 ||
                        opt->projection_limiteur_pure_id > MHG_LIM_ULTRABEEG))
    return MHG_E_UNSUPPORTED;
  if (pb && (opt->ordre_projection != 2 || opt->projection_pente_borne_debar_fix < 0 ||
             opt->projection_pente_borne_debar_fix > 2))
    return MHG_E_UNSUPPORTED;
  const bool generic = opt->ordre_projection == 2 && (pb || mixte);
  if (generic && opt->precision != MHG_PREC_FP64) return MHG_E_UNSUPPORTED;
  Params P{};
  P.dim = mesh->dim;
  P.ne = opt->nb_env;
  P.nvu = 3 * P.ne + 1;
  P.nx = mesh->n[0];
  P.ny = mesh->n[1];
  P.nz = mesh->dim == 3 ? mesh->n[2] : 1;
  P.nc = (long)P.nx * P.ny * (P.dim == 3 ? P.nz : 1);
  P.nn = (long)(P.nx + 1) * (P.ny + 1) * (P.dim == 3 ? P.nz + 1 : 1);
  for (int d = 0; d < 3; ++d) {
    P.org[d] = mesh->origin[d];
    P.h[d] = d < P.dim ? mesh->length[d] / mesh->n[d] : 0.;
  }
  for (int k = 0; k < P.ne; ++k) {
    P.eos.kind[k] = opt->env[k].eos;
    P.eos.gamma[k] = opt->env[k].adiabatic_cst;
    P.eos.tension[k] = opt->env[k].limit_tension;
    P.eos.qcoef[k] = opt->visc_quad_coeff >= 0. ? opt->visc_quad_coeff
                                               : (opt->env[k].adiabatic_cst + 1) / 2.0;
  }
  P.threshold = opt->threshold;
  P.remap_threshold = opt->remap_threshold;
  P.order2 = opt->ordre_projection - 1;
  P.lim = learned ? MHG_LIM_MINMOD : opt->projection_limiteur_id;
  P.learned = !learned ? 0 : (opt->projection_limiteur_id == MHG_LIM_LEARNED_ALL ? 2 : 1);
  P.etot = opt->conservation_energie_totale ? 1 : 0;
  P.lagonly = opt->with_projection ? 0 : 1;
  P.generic = generic ? 1 : 0;
  P.pb = pb ? 1 : 0;
  P.pb_fix = opt->projection_pente_borne_debar_fix;
  P.mixte = mixte ? 1 : 0;
  P.pure_id = opt->projection_limiteur_pure_id;
  P.pseudo_centree = opt->pseudo_centree;
  P.lc = opt->longueur_caracteristique;
  P.cfl = opt->cfl;
  P.dtmin = opt->deltat_min;
  P.dtmax = opt->deltat_max;
  P.tend = opt->final_time;
  P.land = opt->land_on_final_time ? 1 : 0;
  {
    // store policy (st_stream): the remap vector, the dual fluxes and the
    // node velocities are read again by the next kernels of the same step and
    // are written with ordinary stores, which keep them in the L2 and the
    // 256 MB memory-side cache; the cell state is read again only by the
    // next step, and streams to memory once the arrays one step touches
    // exceed that cache.  Measured, ms per step (all ordinary / cell state
    // streaming / all streaming): 1792x768 triple point FP64 (~460 MB)
    // 0.261 / 0.246 / 0.255, FP32 bulk (~250 MB) 0.120 / 0.133 / 0.149;
    // 256^3 FP64 5.65 / 5.59 / 5.83, FP32 4.19 / 4.18 / 4.51.  In FP64 the
    // remap vector between the first and the last cell sweep (10 words a
    // cell on the triple point) streams as well once the fused step keeps
    // everything else in registers: cell state / cell state and remap
    // vector, triple point 0.2180 / 0.2159 ms, 256^3 5.469 / 5.429 ms; in
    // FP32 it stays cached (256^3 4.115 / 4.180 ms).
    const bool fp32 = opt->precision == MHG_PREC_FP32_BULK;
    const double wb = fp32 ? 4. : 8.;
    const double ws = (double)P.nc * wb * (6. * opt->nb_env + 20.) + (double)P.nn * 16. * P.dim;
    P.nt_store = ws > 320e6 ? (fp32 ? (1 << kStCell) : ((1 << kStCell) | (1 << kStU))) : 0;
  }
  // fused y sweeps on 8-column tiles, 512-thread blocks: more blocks per CU
  // and half the wavefronts at each barrier; 1792x768 triple point 0.2355 ->
  // 0.2313 ms per FP64 step, 0.1222 -> 0.1195 in FP32 (4 columns: 0.2498,
  // 0.1328).  The FP64 sweep with the Lagrange cell phase keeps 16 columns
  // (70.0 us against 75.7 on 8).
  P.ytw8 = opt->precision == MHG_PREC_FP32_BULK ? (kYtwLag | kYtwFinal | kYtwNode) : (kYtwFinal | kYtwNode);
  P.nbc = 0;
  for (int b = 0; b < opt->nb_bc; ++b) {
    const mhg_bc& bc = opt->bc[b];
    if (bc.surface < 0 || bc.surface > 5 || bc.type < 0 || bc.type > 2) return MHG_E_ARG;
    if (bc.type >= P.dim) {  // Vz in 2D: u_z is identically zero
      if (bc.value != 0.) return MHG_E_UNSUPPORTED;
      continue;
    }
    BcRange r;
    if (!surface_range(P, bc.surface, &r.axis, &r.lo, &r.hi, P.threshold)) return MHG_E_UNSUPPORTED;
    r.comp = bc.type;
    r.value = bc.value;
    P.bc[P.nbc++] = r;
  }
  *Pout = P;
  return MHG_OK;
}

}  // namespace mhg

extern "C" {

int mhg_abi_version(void) { return MHG_ABI_VERSION; }

void mhg_default_options(mhg_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->cfl = 0.3;
  o->threshold = 1.e-10;
  o->deltat_init = 1.e-5;
  o->deltat_min = 1.e-8;
  o->deltat_max = 0.01;
  o->final_time = 0.2;
  o->with_projection = 1;
  o->longueur_caracteristique = MHG_LC_RACINE_CUBIQUE_VOLUME;
  o->nb_bc = 6;
  const int surf[6] = {MHG_SURF_XMIN, MHG_SURF_XMAX, MHG_SURF_YMIN, MHG_SURF_YMAX, MHG_SURF_ZMIN, MHG_SURF_ZMAX};
  const int typ[6] = {MHG_BC_VX, MHG_BC_VX, MHG_BC_VY, MHG_BC_VY, MHG_BC_VZ, MHG_BC_VZ};
  for (int i = 0; i < 6; ++i) {
    o->bc[i].surface = surf[i];
    o->bc[i].type = typ[i];
    o->bc[i].value = 0.;
  }
  o->nb_env = 1;
  o->env[0].eos = MHG_EOS_PERFECT_GAS;
  o->env[0].adiabatic_cst = 1.4;
  o->env[0].limit_tension = 0.;
  o->remap_threshold = 1.e-10;
  o->is_euler_scheme = 1;
  o->ordre_projection = 2;
  o->projection_limiteur_id = MHG_LIM_MINMOD;
  o->visc_quad_coeff = -1.;
  o->precision = MHG_PREC_FP64;
}

int mhg_create(const mhg_mesh_desc* mesh, const mhg_options* opt, mhg_ctx** out) {
  if (!mesh || !opt || !out) return MHG_E_ARG;
  *out = nullptr;
  Params Pc{};
  const int pr = build_params(mesh, opt, &Pc);
  if (pr) return pr;
  mhg_ctx* c = new mhg_ctx();
  c->opt = *opt;
  c->P = Pc;
  const Params& P = c->P;
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return MHG_E_HIP;
  }
  const long NC = P.nc, NN = P.nn;
  const size_t envn = (size_t)P.ne * NC;
  int r = 0;
  c->bulk = opt->precision == MHG_PREC_FP32_BULK ? 4 : 8;
  r |= balloc(c, &c->rho, NC) | balloc(c, &c->e, NC) | balloc(c, &c->p, NC) | balloc(c, &c->cs, NC);
  r |= balloc(c, &c->q, NC) | balloc(c, &c->m, NC) | balloc(c, &c->L, NC) | balloc(c, &c->Vlag, NC);
  r |= dalloc(c, &c->mask, NC);
  r |= balloc(c, &c->U[0], (size_t)P.nvu * NC) | balloc(c, &c->U[1], (size_t)P.nvu * NC);
  r |= balloc(c, &c->contrib, envn);
  r |= balloc(c, &c->env_rho, envn) | balloc(c, &c->env_e, envn) | balloc(c, &c->env_p, envn);
  r |= balloc(c, &c->env_cs, envn) | balloc(c, &c->env_q, envn) | balloc(c, &c->env_m, envn);
  r |= balloc(c, &c->env_f, envn) | balloc(c, &c->env_mf, envn);
  for (int b = 0; b < 2; ++b) {
    r |= balloc(c, &c->ux[b], NN) | balloc(c, &c->uy[b], NN);
    r |= dalloc(c, &c->xx[b], NN) | dalloc(c, &c->xy[b], NN);
  }
  r |= balloc(c, &c->Ud0, NN) | balloc(c, &c->Ud1, NN) | balloc(c, &c->Ud3, NN);
  r |= balloc(c, &c->dg0, NN) | balloc(c, &c->dg1, NN);
  if (P.etot) r |= balloc(c, &c->Ud4, NN) | balloc(c, &c->dek, NN);
  if (P.order2 == 2 || P.generic) r |= dalloc(c, &c->est, NC);  // zero: pure in environment 0, as never written
  if (P.generic) {
    const size_t nvn = (size_t)P.nvu * NC;
    r |= dalloc(c, &c->dar, nvn);
    if (P.pb) r |= dalloc(c, &c->phi, nvn) | dalloc(c, &c->dav, nvn);
  }
  if (P.etot && P.dim == 3) r |= balloc(c, &c->UdB4, NN);
  if (P.dim == 3) {
    for (int b = 0; b < 2; ++b) r |= dalloc(c, &c->xz[b], NN) | balloc(c, &c->uz[b], NN);
    r |= balloc(c, &c->Ud2, NN) | balloc(c, &c->dg2, NN);
    for (int q = 0; q < 4; ++q) r |= balloc(c, &c->UdB[q], NN);
    r |= dalloc(c, &c->Veuler, NC);
    if (!r) {
      std::vector<double> ve(NC);
      for (long cc = 0; cc < NC; ++cc) ve[cc] = euler_volume(P, cc);
      if (hipMemcpy(c->Veuler, ve.data(), NC * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        r = MHG_E_HIP;
    }
  }
  if (P.learned) r |= balloc(c, &c->alpha, 2 * (size_t)NC);  // coefficients 0 until set
  r |= dalloc(c, &c->block_min, cell_final_blocks(P));
  r |= dalloc(c, &c->S, 1);
  r |= dalloc(c, &c->S_alt, 1);
  if (r) {
    mhg_destroy(c);
    return MHG_E_ALLOC;
  }
  if (hipHostMalloc((void**)&c->S_host, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    mhg_destroy(c);
    return MHG_E_ALLOC;
  }
  *out = c;
  return MHG_OK;
}

int mhg_destroy(mhg_ctx* c) {
  if (!c) return MHG_OK;
  if (c->owner) return fail(c, MHG_E_ARG, "problem of a batch: destroyed with the batch");
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (auto& b : c->owned) hipFree(b.p);
  for (auto e : c->evpool) hipEventDestroy(e);
  if (c->S_host) hipHostFree(c->S_host);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return MHG_OK;
}

const char* mhg_last_error(const mhg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mhg_sizes(const mhg_ctx* c, int64_t* n_cells, int64_t* n_nodes) {
  if (!c) return MHG_E_ARG;
  if (n_cells) *n_cells = c->P.nc;
  if (n_nodes) *n_nodes = c->P.nn;
  return MHG_OK;
}

// SODService (SODService.cc:3-135): cas-test 3/4/5 one environment,
// 13/14/15 two environments split at r = 0.5.
int mhg_init_sod(mhg_ctx* c, int32_t cas_test) {
  if (!c) return MHG_E_ARG;
  const Params& P = c->P;
  const bool mono = (cas_test >= 3 && cas_test <= 5), bi = (cas_test >= 13 && cas_test <= 15);
  if (!mono && !bi) return fail(c, MHG_E_ARG, "unsupported SOD cas-test");
  if (bi && P.ne < 2) return fail(c, MHG_E_ARG, "bi-material SOD needs 2 environments");
  const int axis = mono ? cas_test - 3 : cas_test - 13;
  if (P.dim != 2) return fail(c, MHG_E_UNSUPPORTED, "3D initial states: use mhg_init_cells");
  if (axis >= P.dim) return fail(c, MHG_E_ARG, "SOD axis beyond mesh dimension");
  std::vector<int> env(P.nc);
  std::vector<double> rho(P.nc), pres(P.nc);
  for (long cc = 0; cc < P.nc; ++cc) {
    double xc, yc;
    cell_centre(P, cc, &xc, &yc);
    const double r = axis == 0 ? xc : yc;
    env[cc] = mono ? 0 : (r < 0.5 ? 0 : 1);
    rho[cc] = r < 0.5 ? 1.0 : 0.125;
    pres[cc] = r < 0.5 ? 1.0 : 0.1;
  }
  return init_pure(c, env, rho, pres, nullptr);
}

int mhg_init_triple_point(mhg_ctx* c, const double rho3[3], const double p3[3]) {
  if (!c || !rho3 || !p3) return MHG_E_ARG;
  const Params& P = c->P;
  if (P.ne != 3) return fail(c, MHG_E_ARG, "triple point needs 3 environments");
  if (P.dim != 2) return fail(c, MHG_E_UNSUPPORTED, "3D initial states: use mhg_init_cells");
  std::vector<int> env(P.nc);
  std::vector<double> rho(P.nc), pres(P.nc);
  for (long cc = 0; cc < P.nc; ++cc) {
    double xc, yc;
    cell_centre(P, cc, &xc, &yc);
    const int k = (xc < 1.) ? 0 : (yc < 1.5 ? 1 : 2);
    env[cc] = k;
    rho[cc] = rho3[k];
    pres[cc] = p3[k];
  }
  return init_pure(c, env, rho, pres, nullptr);
}

int mhg_init_cells(mhg_ctx* c, const int32_t* env_of_cell, const double* density,
                   const double* pressure, const double* velocity) {
  if (!c || !env_of_cell || !density || !pressure) return MHG_E_ARG;
  const Params& P = c->P;
  std::vector<int> env(env_of_cell, env_of_cell + P.nc);
  for (int k : env)
    if (k < 0 || k >= P.ne) return fail(c, MHG_E_ARG, "bad env index");
  std::vector<double> rho(density, density + P.nc), pres(pressure, pressure + P.nc);
  return init_pure(c, env, rho, pres, velocity);
}

int mhg_upload_state(mhg_ctx* c, const mhg_state* st) {
  if (!c || !st) return MHG_E_ARG;
  if (batch_running(c)) return fail(c, MHG_E_STATE, "problem of a running batch: state fixed at the first run");
  const Params& P = c->P;
  const long NC = P.nc, NN = P.nn;
  const size_t nb = NN * sizeof(double);
  c->xcur = 0;
  c->ucur = 0;
  c->x_lazy = false;
  c->coord_phase = 0;
  c->ep_remapped = false;
  c->async_launched = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto h2d = [&](void* d, const void* h, size_t b) {
    return h ? hipMemcpy(d, h, b, hipMemcpyHostToDevice) : hipSuccess;
  };
  const size_t en = (size_t)P.ne * NC;
  HIPCHK(c, put(c, c->rho, st->density, NC));
  HIPCHK(c, put(c, c->e, st->internal_energy, NC));
  HIPCHK(c, put(c, c->p, st->pressure, NC));
  HIPCHK(c, put(c, c->cs, st->sound_speed, NC));
  HIPCHK(c, put(c, c->q, st->pseudo_viscosity, NC));
  HIPCHK(c, put(c, c->m, st->cell_mass, NC));
  HIPCHK(c, h2d(c->mask, st->env_mask, NC * sizeof(int)));
  if (c->est && st->env_mask && st->fracvol && st->fracvol_env) {
    // order 3: mixed / pure codes from the uploaded fractions (two or more
    // environments with f > 0: mixed, else the last one with f > 0)
    std::vector<int> e(NC);
    for (long cc = 0; cc < NC; ++cc) {
      const int mk = st->env_mask[cc], nbe = __builtin_popcount((unsigned)mk);
      int cnt = 0, last = -1;
      for (int k = 0; k < P.ne; ++k) {
        if (!((mk >> k) & 1)) continue;
        const double f = nbe == 1 ? st->fracvol[cc] : st->fracvol_env[(size_t)k * NC + cc];
        if (f > 0.) {
          cnt++;
          last = k;
        }
      }
      e[cc] = cnt > 1 ? -2 : last;
    }
    HIPCHK(c, h2d(c->est, e.data(), NC * sizeof(int)));
  }
  HIPCHK(c, put(c, c->env_rho, st->density_env, en));
  HIPCHK(c, put(c, c->env_e, st->internal_energy_env, en));
  HIPCHK(c, put(c, c->env_p, st->pressure_env, en));
  HIPCHK(c, put(c, c->env_cs, st->sound_speed_env, en));
  HIPCHK(c, put(c, c->env_q, st->pseudo_viscosity_env, en));
  HIPCHK(c, put(c, c->env_m, st->cell_mass_env, en));
  HIPCHK(c, put(c, c->env_f, st->fracvol_env, en));
  HIPCHK(c, put(c, c->env_mf, st->mass_fraction_env, en));
  if (P.pb && st->env_mask && st->fracvol) {
    // pente-borne: a pure cell's fraction in its environment slot; phi of
    // absent environments starts from zero
    std::vector<double> f((size_t)P.ne * NC, 0.);
    if (st->fracvol_env) std::memcpy(f.data(), st->fracvol_env, en * sizeof(double));
    for (long cc = 0; cc < NC; ++cc)
      if (__builtin_popcount((unsigned)st->env_mask[cc]) == 1)
        f[(size_t)(__builtin_ffs(st->env_mask[cc]) - 1) * NC + cc] = st->fracvol[cc];
    HIPCHK(c, put(c, c->env_f, f.data(), en));
  }
  if (c->phi) HIPCHK(c, hipMemset(c->phi, 0, (size_t)P.nvu * NC * sizeof(double)));
  if (st->velocity) {
    HIPCHK(c, put(c, c->ux[0], st->velocity, NN));
    HIPCHK(c, put(c, c->uy[0], st->velocity + NN, NN));
    if (P.dim == 3) HIPCHK(c, put(c, c->uz[0], st->velocity + 2 * NN, NN));
  }
  if (st->node_coord) {
    HIPCHK(c, h2d(c->xx[0], st->node_coord, nb));
    HIPCHK(c, h2d(c->xy[0], st->node_coord + NN, nb));
    if (P.dim == 3) HIPCHK(c, h2d(c->xz[0], st->node_coord + 2 * NN, nb));
  }
  if (st->dual_grad_phi) {
    HIPCHK(c, put(c, c->dg0, st->dual_grad_phi, NN));
    HIPCHK(c, put(c, c->dg1, st->dual_grad_phi + NN, NN));
    if (P.dim == 3) HIPCHK(c, put(c, c->dg2, st->dual_grad_phi + 2 * NN, NN));
  }
  if (st->caracteristic_length) HIPCHK(c, put(c, c->L, st->caracteristic_length, NC));
  std::memset(c->S_host, 0, sizeof(DevScalars));
  c->S_host->dt = st->global_deltat;
  c->S_host->dt_old = st->old_deltat;
  c->S_host->time = st->global_time;
  c->S_host->iteration = st->global_iteration;
  c->sens = st->sens_projection;
  int r = write_scalars(c);
  if (r) return r;
  c->initialised = true;
  return MHG_OK;
}

int mhg_download_state(mhg_ctx* c, mhg_state* st) {
  int r = check_ready(c);
  if (r) return r;
  if (c->owner) {
    r = batch_refresh_view(c);
    if (r) return r;
  }
  const Params& P = c->P;
  const long NC = P.nc, NN = P.nn;
  const size_t nb = NN * sizeof(double), en = (size_t)P.ne * NC;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto d2h = [&](void* h, const void* d, size_t b) {
    return h ? hipMemcpy(h, d, b, hipMemcpyDeviceToHost) : hipSuccess;
  };
  HIPCHK(c, get(c, st->density, c->rho, NC));
  HIPCHK(c, get(c, st->internal_energy, c->e, NC));
  HIPCHK(c, get(c, st->pressure, c->p, NC));
  HIPCHK(c, get(c, st->sound_speed, c->cs, NC));
  HIPCHK(c, get(c, st->pseudo_viscosity, c->q, NC));
  HIPCHK(c, get(c, st->cell_mass, c->m, NC));
  std::vector<int> mask(NC);
  HIPCHK(c, hipMemcpy(mask.data(), c->mask, NC * sizeof(int), hipMemcpyDeviceToHost));
  if (st->env_mask) std::memcpy(st->env_mask, mask.data(), NC * sizeof(int));
  struct EP { double* h; void* d; double* g; };
  void* gsrc[6] = {c->rho, c->e, c->p, c->cs, c->q, c->m};
  EP eps[8] = {{st->density_env, c->env_rho, st->density},
               {st->internal_energy_env, c->env_e, st->internal_energy},
               {st->pressure_env, c->env_p, st->pressure},
               {st->sound_speed_env, c->env_cs, st->sound_speed},
               {st->pseudo_viscosity_env, c->env_q, st->pseudo_viscosity},
               {st->cell_mass_env, c->env_m, st->cell_mass},
               {st->fracvol_env, c->env_f, nullptr},
               {st->mass_fraction_env, c->env_mf, nullptr}};
  // pente-borne keeps a pure cell's fraction in its environment slot
  std::vector<double> fpure;
  if (P.pb) {
    fpure.resize(en);
    HIPCHK(c, get(c, fpure.data(), c->env_f, en));
  }
  auto pure_frac = [&](long cc) {
    return P.pb ? fpure[(size_t)(__builtin_ffs(mask[cc]) - 1) * NC + cc] : 1.0;
  };
  for (int t = 0; t < 8; ++t) {
    if (!eps[t].h) continue;
    HIPCHK(c, get(c, eps[t].h, eps[t].d, en));
    // pure cells: the environment slot is the cell slot
    std::vector<double> g(NC, 1.0);
    if (t < 6) HIPCHK(c, get(c, g.data(), gsrc[t], NC));
    for (long cc = 0; cc < NC; ++cc)
      if (__builtin_popcount((unsigned)mask[cc]) == 1)
        eps[t].h[(size_t)(__builtin_ffs(mask[cc]) - 1) * NC + cc] = t == 6 ? pure_frac(cc) : g[cc];
  }
  // the cell fracvol / mass fraction of single-environment cells (1 up to
  // rounding); only pente-borne keeps and reads the fraction
  for (long cc = 0; cc < NC; ++cc) {
    const bool pure = __builtin_popcount((unsigned)mask[cc]) == 1;
    if (st->fracvol) st->fracvol[cc] = pure ? pure_frac(cc) : 0.0;
    if (st->mass_fraction) st->mass_fraction[cc] = pure ? 1.0 : 0.0;
  }
  if (st->velocity) {
    HIPCHK(c, get(c, st->velocity, c->ux[c->ucur], NN));
    HIPCHK(c, get(c, st->velocity + NN, c->uy[c->ucur], NN));
    if (P.dim == 3) HIPCHK(c, get(c, st->velocity + 2 * NN, c->uz[c->ucur], NN));
  }
  if (st->node_coord) {
    if (c->coord_phase == 1 && !P.lagonly) {  // Euler reset of saveValuesAtN: x = x0
      for (long n = 0; n < NN; ++n) {
        const int i = (int)(n % (P.nx + 1)), j = (int)((n / (P.nx + 1)) % (P.ny + 1));
        st->node_coord[n] = x0_of(P, 0, i);
        st->node_coord[NN + n] = x0_of(P, 1, j);
        if (P.dim == 3) st->node_coord[2 * NN + n] = x0_of(P, 2, (int)(n / ((long)(P.nx + 1) * (P.ny + 1))));
      }
    } else {
      const int b = c->coord_phase == 2 ? 1 - c->xcur : c->xcur;
      HIPCHK(c, d2h(st->node_coord, c->xx[b], nb));
      HIPCHK(c, d2h(st->node_coord + NN, c->xy[b], nb));
      if (P.dim == 3) HIPCHK(c, d2h(st->node_coord + 2 * NN, c->xz[b], nb));
    }
  }
  if (st->dual_grad_phi) {
    HIPCHK(c, get(c, st->dual_grad_phi, c->dg0, NN));
    HIPCHK(c, get(c, st->dual_grad_phi + NN, c->dg1, NN));
    if (P.dim == 3) HIPCHK(c, get(c, st->dual_grad_phi + 2 * NN, c->dg2, NN));
  }
  if (st->caracteristic_length) HIPCHK(c, get(c, st->caracteristic_length, c->L, NC));
  if (st->node_mass) {
    std::vector<double> m(NC);
    HIPCHK(c, get(c, m.data(), c->m, NC));
    // computeNodeMass (MahycoModule.cc:235-266): cells in ascending id
    const int nz = P.dim == 3 ? P.nz : 1;
    for (long n = 0; n < NN; ++n) {
      const int i = (int)(n % (P.nx + 1)), j = (int)((n / (P.nx + 1)) % (P.ny + 1));
      const int k = P.dim == 3 ? (int)(n / ((long)(P.nx + 1) * (P.ny + 1))) : 0;
      double s = 0.;
      for (int dk = (P.dim == 3 ? -1 : 0); dk <= 0; ++dk)
        for (int dj = -1; dj <= 0; ++dj)
          for (int di = -1; di <= 0; ++di) {
            const int ci = i + di, cj = j + dj, ck = k + dk;
            if (ci < 0 || ci >= P.nx || cj < 0 || cj >= P.ny || ck < 0 || ck >= nz) continue;
            s += m[ci + (long)P.nx * (cj + (long)P.ny * ck)];
          }
      st->node_mass[n] = (P.dim == 3 ? 0.125 : 0.25) * s;
    }
  }
  r = read_scalars(c);
  if (r) return r;
  st->global_deltat = c->S_host->dt;
  st->old_deltat = c->S_host->dt_old;
  st->global_time = c->S_host->time;
  st->global_iteration = c->S_host->iteration;
  st->sens_projection = c->sens;
  return MHG_OK;
}

// ---- per-entry-point path (Mahyco.config:30-41) ---------------------------
// Each call enqueues the kernels of one reference entry point (or service
// method) on the context stream and waits for them, so the downloaded state
// between calls is the state the reference holds between its calls.  2D and
// 3D, FP64 contexts.
namespace {

int ep_prepare(mhg_ctx* c) {
  int r = check_ready(c);
  if (r) return r;
  r = check_not_batch(c);
  if (r) return r;
  if (c->bulk != 8) return fail(c, MHG_E_UNSUPPORTED, "entry points: FP64 contexts only (use mhg_step)");
  if (c->ep_ready) return MHG_OK;
  const size_t NC = c->P.nc, envn = (size_t)c->P.ne * NC;
  r |= dalloc(c, &c->rho_n, NC) | dalloc(c, &c->e_n, NC) | dalloc(c, &c->q_n, NC);
  r |= dalloc(c, &c->tau, NC) | dalloc(c, &c->divu, NC);
  r |= dalloc(c, &c->env_rho_n, envn) | dalloc(c, &c->env_e_n, envn) | dalloc(c, &c->env_q_n, envn);
  r |= dalloc(c, &c->env_tau, envn);
  if (r) return fail(c, MHG_E_ALLOC, "entry-point buffers: hipMalloc failed");
  c->ep_ready = true;
  return MHG_OK;
}

StageArgs stage_args(mhg_ctx* c) {
  StageArgs A;
  A.xnx = c->xx[1 - c->xcur];
  A.xny = c->xy[1 - c->xcur];
  A.rho = bp<double>(c->rho);
  A.e = bp<double>(c->e);
  A.p = bp<double>(c->p);
  A.cs = bp<double>(c->cs);
  A.q = bp<double>(c->q);
  A.m = bp<double>(c->m);
  A.L = bp<double>(c->L);
  A.Vlag = bp<double>(c->Vlag);
  A.mask = c->mask;
  A.env_f = bp<double>(c->env_f);
  A.env_mf = bp<double>(c->env_mf);
  A.env_m = bp<double>(c->env_m);
  A.env_p = bp<double>(c->env_p);
  A.env_rho = bp<double>(c->env_rho);
  A.env_e = bp<double>(c->env_e);
  A.env_q = bp<double>(c->env_q);
  A.rho_n = c->rho_n;
  A.e_n = c->e_n;
  A.q_n = c->q_n;
  A.tau = c->tau;
  A.divu = c->divu;
  A.env_rho_n = c->env_rho_n;
  A.env_e_n = c->env_e_n;
  A.env_q_n = c->env_q_n;
  A.env_tau = c->env_tau;
  A.U = bp<double>(c->U[0]);
  A.phi = c->phi;
  return A;
}

EosArgs eos_args(mhg_ctx* c) {
  EosArgs A;
  A.rho = bp<double>(c->rho);
  A.e = bp<double>(c->e);
  A.p = bp<double>(c->p);
  A.cs = bp<double>(c->cs);
  A.mask = c->mask;
  A.env_rho = bp<double>(c->env_rho);
  A.env_e = bp<double>(c->env_e);
  A.env_p = bp<double>(c->env_p);
  A.env_cs = bp<double>(c->env_cs);
  return A;
}

PmArgs pm_args(mhg_ctx* c) {
  PmArgs A;
  A.p = bp<double>(c->p);
  A.cs = bp<double>(c->cs);
  A.mask = c->mask;
  A.env_f = bp<double>(c->env_f);
  A.env_p = bp<double>(c->env_p);
  A.env_cs = bp<double>(c->env_cs);
  return A;
}

int ep_finish(mhg_ctx* c) {
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MHG_OK;
}

void enqueue_stage(mhg_ctx* c, int stage) {
  if (stage == STAGE_GEOMETRY && c->P.dim == 3)
    launch_cell_geometry_3d(c->stream, c->P, c->S, c->xx[1 - c->xcur], c->xy[1 - c->xcur], c->xz[1 - c->xcur],
                            bp<double>(c->Vlag), bp<double>(c->L));
  else
    launch_cell_stage_2d(c->stream, c->P, c->S, stage_args(c), stage);  // cell-indexed beyond the geometry
}

int ep_stage(mhg_ctx* c, int stage) {
  int r = ep_prepare(c);
  if (r) return r;
  enqueue_stage(c, stage);
  return ep_finish(c);
}

int ep_node(mhg_ctx* c, bool vel, bool bc, bool pos) {
  int r = ep_prepare(c);
  if (r) return r;
  if (c->P.dim == 3) {
    NodeLag3ArgsT<double> N;
    N.b = node_lag_args<double>(c);
    N.xpz = c->xz[c->xcur];
    N.uz = bp<double>(c->uz[c->ucur]);
    N.xnz = c->xz[1 - c->xcur];
    launch_node_lagrange_3d(c->stream, c->P, c->S, N, vel, bc, pos);
  } else {
    launch_node_lagrange_2d(c->stream, c->P, c->S, node_lag_args<double>(c), vel, bc, pos);
  }
  return ep_finish(c);
}

// applyEOS of every environment, then computePressionMoyenne (the end of
// remap(), PrepareRemap.cc:552-559, and of updateEnergyAndPressure without
// projection, MahycoModule.cc:1468-1476).
void enqueue_eos_and_average(mhg_ctx* c) {
  for (int k = 0; k < c->P.ne; ++k) launch_eos_env(c->stream, c->P, c->S, eos_args(c), k, EOS_APPLY, 0, c->P.nc);
  launch_pression_moyenne(c->stream, c->P, c->S, pm_args(c));
}

// The arguments of the remap services: the dimension of the mesh, the dual
// projection (always on: sans-lagrange is refused), nb_vars_to_project =
// 3 nb_env + 5 (volumes, masses, energies, momentum, kinetic energy, pseudo;
// MahycoModule's resizeRemapVariables call) and nb_env.
int check_remap_args(mhg_ctx* c, int32_t dim, int32_t with_dual, int32_t nb_vars, int32_t nb_env) {
  if (dim != c->P.dim || with_dual != 1 || nb_env != c->P.ne || nb_vars != 3 * c->P.ne + 5)
    return fail(c, MHG_E_ARG, "remap: dimension, dual projection, variable or environment count "
                              "differs from the context's");
  if (c->P.lagonly) return fail(c, MHG_E_STATE, "remap: the context runs without projection");
  return MHG_OK;
}

int check_env(mhg_ctx* c, int32_t env) {
  if (env < 0 || env >= c->P.ne) return fail(c, MHG_E_ARG, "environment index out of range");
  return MHG_OK;
}

void enqueue_remap_fp64(mhg_ctx* c, bool eos) {
  if (c->P.dim == 3) remap_3d<double>(c, 0, eos);
  else remap_2d<double>(c, 0, eos);
}

}  // namespace

int mhg_begin_iteration(mhg_ctx* c) {
  int r = ep_prepare(c);
  if (r) return r;
  launch_begin_iteration(c->stream, c->S);
  return ep_finish(c);
}

// saveValuesAtN: rho_n, e_n, q_n <- rho, e, q; the velocity is updated in
// place and the coordinates of the last Lagrange phase stay in their buffer
// (the lagged corner normals are formed from them), so the Euler reset
// x <- x0 only changes which coordinates are current.
int mhg_save_values_at_n(mhg_ctx* c) {
  if (c && c->use_mlp && c->initialised && !c->owner) enqueue_learned_alpha<double>(c, nullptr, nullptr, true);
  int r = ep_stage(c, STAGE_SAVE);
  if (r) return r;
  c->coord_phase = 1;
  return MHG_OK;
}

int mhg_update_velocity(mhg_ctx* c) { return ep_node(c, true, false, false); }
int mhg_apply_boundary_condition(mhg_ctx* c) { return ep_node(c, false, true, false); }

int mhg_update_position(mhg_ctx* c) {
  if (c && c->coord_phase != 1) return fail(c, MHG_E_STATE, "update_position: call save_values_at_n first");
  int r = ep_node(c, false, false, true);
  if (r) return r;
  c->coord_phase = 2;
  return MHG_OK;
}

int mhg_compute_geometric_values(mhg_ctx* c) {
  if (c && c->coord_phase != 2) return fail(c, MHG_E_STATE, "compute_geometric_values: call update_position first");
  return ep_stage(c, STAGE_GEOMETRY);
}
int mhg_update_density(mhg_ctx* c) { return ep_stage(c, STAGE_DENSITY); }
int mhg_compute_artificial_viscosity(mhg_ctx* c) { return ep_stage(c, STAGE_VISCOSITY); }

// updateEnergyAndPressure (MahycoModule.cc:1439-1478): the energy update;
// without projection also the EOS of every environment and the pressure
// average.
int mhg_update_energy_and_pressure(mhg_ctx* c) {
  int r = ep_prepare(c);
  if (r) return r;
  enqueue_stage(c, STAGE_ENERGY);
  if (c->P.lagonly) enqueue_eos_and_average(c);
  return ep_finish(c);
}

// remap (PrepareRemap.cc:448-563): U from the state, the sweeps, the final
// cell phase with the EOS and the pressure average, fused; dt is left to
// mhg_compute_delta_t.  Without projection it does nothing but end the
// iteration's Lagrange phase (the mesh stays where it moved).
int mhg_remap(mhg_ctx* c) {
  if (c && c->coord_phase != 2) return fail(c, MHG_E_STATE, "remap: the Lagrange phase of this iteration has not run");
  int r = ep_prepare(c);
  if (r) return r;
  if (c->P.lagonly) {
    c->xcur = 1 - c->xcur;
  } else {
    launch_cell_stage_2d(c->stream, c->P, c->S, stage_args(c), STAGE_BUILD_U);
    enqueue_remap_fp64(c, true);
    r = ep_finish(c);
    if (r) return r;
  }
  c->coord_phase = 0;
  c->remap_stage = 0;
  c->ep_remapped = true;
  return MHG_OK;
}

// ---- the remap split as the reference's services call it ---------------
// computeVariablesForRemap + computeFaceQuantitesForRemap
// (PrepareRemap.cc:15-437): the remap vector of every cell; the face
// quantities are formed inside the sweeps from the node data.
int mhg_remap_prepare(mhg_ctx* c) {
  if (c && c->coord_phase != 2)
    return fail(c, MHG_E_STATE, "remap_prepare: the Lagrange phase of this iteration has not run");
  int r = ep_prepare(c);
  if (r) return r;
  if (c->P.lagonly) return fail(c, MHG_E_STATE, "remap_prepare: the context runs without projection");
  launch_cell_stage_2d(c->stream, c->P, c->S, stage_args(c), STAGE_BUILD_U);
  r = ep_finish(c);
  if (r) return r;
  c->remap_stage = 1;
  return MHG_OK;
}

// IRemap::appliRemap (RemapADIService.cc:19-60): the cell and dual sweeps in
// the SensProjection order, then remapVariables.  Pressures and sound speeds
// are left to the EOS and pressure-average calls.
int mhg_remap_appli(mhg_ctx* c, int32_t dimension, int32_t with_dual_projection, int32_t nb_vars_to_project,
                    int32_t nb_env) {
  int r = ep_prepare(c);
  if (r) return r;
  r = check_remap_args(c, dimension, with_dual_projection, nb_vars_to_project, nb_env);
  if (r) return r;
  if (c->remap_stage != 1) return fail(c, MHG_E_STATE, "remap_appli: call remap_prepare first");
  enqueue_remap_fp64(c, false);
  r = ep_finish(c);
  if (r) return r;
  c->coord_phase = 0;
  c->remap_stage = 0;
  c->ep_remapped = true;
  return MHG_OK;
}

// IRemap::remapVariables (RemapADIFinal.cc:19-778) from the remap vectors of
// the last appliRemap: membership, "moy", negative-value resets and the
// total-energy fix-up.  appliRemap already ends with it; a second call
// derives the same state from the same vectors.
int mhg_remap_variables(mhg_ctx* c, int32_t dimension, int32_t with_dual_projection, int32_t nb_vars_to_project,
                        int32_t nb_env) {
  int r = ep_prepare(c);
  if (r) return r;
  r = check_remap_args(c, dimension, with_dual_projection, nb_vars_to_project, nb_env);
  if (r) return r;
  if (!c->ep_remapped) return fail(c, MHG_E_STATE, "remap_variables: no remap vectors of this iteration");
  if (c->P.dim == 3) launch_cell_final_3d(c->stream, c->P, c->S, cell_final3_args<double>(c, c->ucur), 0, false);
  else launch_cell_final_2d(c->stream, c->P, c->S, cell_final_args<double>(c), 0, false);
  return ep_finish(c);
}

// IRemap::resizeRemapVariables (RemapADIService.cc:63-90): the context sized
// its remap arrays at creation; the call checks the counts.
int mhg_remap_resize_variables(mhg_ctx* c, int32_t nb_vars_to_project, int32_t nb_env) {
  if (!c) return MHG_E_ARG;
  if (nb_env != c->P.ne || nb_vars_to_project != 3 * c->P.ne + 5)
    return fail(c, MHG_E_ARG, "resize_remap_variables: counts differ from the context's");
  return MHG_OK;
}

// IRemap::synchronizeUremap (RemapADIService.cc:1050-1057): ghost-cell
// synchronisation; a context holds the whole mesh on one device, so there
// are no ghosts.
int mhg_remap_synchronize_uremap(mhg_ctx* c) {
  if (!c) return MHG_E_ARG;
  return check_ready(c);
}

// ---- IEquationOfState (mahyco/src/eos/IEquationOfState.h:45-63) --------
int mhg_eos_init(mhg_ctx* c, int32_t env) {
  int r = ep_prepare(c);
  if (r) return r;
  r = check_env(c, env);
  if (r) return r;
  launch_eos_env(c->stream, c->P, c->S, eos_args(c), env, EOS_INIT, 0, c->P.nc);
  return ep_finish(c);
}

int mhg_eos_apply(mhg_ctx* c, int32_t env) {
  int r = ep_prepare(c);
  if (r) return r;
  r = check_env(c, env);
  if (r) return r;
  launch_eos_env(c->stream, c->P, c->S, eos_args(c), env, EOS_APPLY, 0, c->P.nc);
  return ep_finish(c);
}

int mhg_eos_apply_one_cell(mhg_ctx* c, int32_t env, int64_t cell) {
  int r = ep_prepare(c);
  if (r) return r;
  r = check_env(c, env);
  if (r) return r;
  if (cell < 0 || cell >= c->P.nc) return fail(c, MHG_E_ARG, "eos_apply_one_cell: cell out of range");
  launch_eos_env(c->stream, c->P, c->S, eos_args(c), env, EOS_APPLY, cell, cell + 1);
  return ep_finish(c);
}

int mhg_eos_get_adiabatic_cst(const mhg_ctx* c, int32_t env, double* value) {
  if (!c || !value || env < 0 || env >= c->P.ne) return MHG_E_ARG;
  *value = c->opt.env[env].adiabatic_cst;
  return MHG_OK;
}

int mhg_eos_get_tension_limit_cst(const mhg_ctx* c, int32_t env, double* value) {
  if (!c || !value || env < 0 || env >= c->P.ne) return MHG_E_ARG;
  *value = c->opt.env[env].limit_tension;
  return MHG_OK;
}

// computePressionMoyenne (MahycoModule.cc:1836-1915).
int mhg_compute_pression_moyenne(mhg_ctx* c) {
  int r = ep_prepare(c);
  if (r) return r;
  launch_pression_moyenne(c->stream, c->P, c->S, pm_args(c));
  return ep_finish(c);
}

// computeDeltaT (MahycoModule.cc:2027-2191): the CFL candidates of the
// current state, then dt, the deltat-min check and the stop test.
int mhg_compute_delta_t(mhg_ctx* c, double* dt, int32_t* stop) {
  int r = ep_prepare(c);
  if (r) return r;
  if (!c->ep_remapped) return fail(c, MHG_E_STATE, "compute_delta_t: call remap first");
  CflArgs A;
  A.L = bp<double>(c->L);
  A.cs = bp<double>(c->cs);
  A.ux = bp<double>(c->ux[c->ucur]);
  A.uy = bp<double>(c->uy[c->ucur]);
  A.uz = c->P.dim == 3 ? bp<double>(c->uz[c->ucur]) : nullptr;
  A.block_min = c->block_min;
  launch_cfl(c->stream, c->P, c->S, A);
  launch_dt_finalize(c->stream, c->P, c->S, c->block_min, false, c->P.dim == 2 ? cfl_blocks(c->P) : 0);
  r = ep_finish(c);
  if (r) return r;
  c->ep_remapped = false;
  r = read_scalars(c);
  if (r) return r;
  if (dt) *dt = c->S_host->dt;
  if (stop) *stop = c->S_host->stop;
  if (c->S_host->err) return fail(c, MHG_E_DT_MIN, "dt below deltat-min");
  return MHG_OK;
}

int mhg_remap_get_ordre_projection(const mhg_ctx* c) { return c ? c->opt.ordre_projection : MHG_E_ARG; }
int mhg_remap_has_projection_pente_borne(const mhg_ctx* c) { return c ? c->opt.projection_pente_borne : MHG_E_ARG; }
int mhg_remap_has_conservation_energie_totale(const mhg_ctx* c) { return c ? c->opt.conservation_energie_totale : MHG_E_ARG; }
int mhg_remap_is_euler(const mhg_ctx* c) { return c ? c->opt.is_euler_scheme : MHG_E_ARG; }

namespace {
// One iteration through the entry-point kernels, in the order of
// Mahyco.config:30-41: the Lagrange-only mode (with-projection = false) and
// the generic sweeps (pente-borne, -mixte).
bool entry_point_steps(const Params& P) { return P.lagonly || P.generic; }
int step_entry_points(mhg_ctx* c, int32_t* stop) {
  int r = read_scalars(c);
  if (r) return r;
  if (c->S_host->stop || c->S_host->err) {  // the run has ended: nothing runs
    if (stop) *stop = c->S_host->stop;
    return c->S_host->err ? fail(c, MHG_E_DT_MIN, "dt below deltat-min") : MHG_STOP;
  }
  int (*const calls[])(mhg_ctx*) = {mhg_begin_iteration, mhg_save_values_at_n, mhg_update_velocity,
                                    mhg_apply_boundary_condition, mhg_update_position,
                                    mhg_compute_geometric_values, mhg_update_density,
                                    mhg_compute_artificial_viscosity, mhg_update_energy_and_pressure, mhg_remap};
  for (auto f : calls)
    if ((r = f(c))) return r;
  int32_t st = 0;
  r = mhg_compute_delta_t(c, nullptr, &st);
  if (stop) *stop = st;
  if (r) return r;
  return st ? MHG_STOP : MHG_OK;
}
}  // namespace

int mhg_step(mhg_ctx* c, int32_t* stop) {
  int r = check_fused(c);
  if (r) return r;
  if (entry_point_steps(c->P)) return step_entry_points(c, stop);
  r = read_scalars(c);
  if (r) return r;
  const int it0 = c->S_host->iteration;
  c->x_lazy_before = c->x_lazy;
  enqueue_step(c, true);
  HIPCHK(c, hipGetLastError());
  r = read_scalars(c);
  if (r) return r;
  roll_back(c, 1 - (c->S_host->iteration - it0), c->S_host->iteration - it0);  // a step after the end is skipped
  collect_times(c);
  if (stop) *stop = c->S_host->stop;
  if (c->S_host->err) return fail(c, MHG_E_DT_MIN, "dt below deltat-min");
  return c->S_host->stop ? MHG_STOP : MHG_OK;
}

int mhg_run_async(mhg_ctx* c, int32_t n) {
  int r = check_fused(c);
  if (r) return r;
  if (n < 0) return fail(c, MHG_E_ARG, "run_async: negative step count");
  if (entry_point_steps(c->P)) return mhg_run(c, n, nullptr) < 0 ? MHG_E_HIP : MHG_OK;  // host-synchronous
  // check_fused has reconciled earlier iterations: the device count is the
  // host's starting point for these
  r = read_scalars(c);
  if (r) return r;
  c->async_it0 = c->S_host->iteration;
  c->x_lazy_before = c->x_lazy;
  r = enqueue_steps(c, n);
  c->async_launched = n;
  return r;
}

int mhg_synchronize(mhg_ctx* c) {
  if (!c) return MHG_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_times(c);
  return reconcile_async(c);
}

int mhg_run(mhg_ctx* c, int32_t max_steps, int32_t* steps_done) {
  int r = check_fused(c);
  if (r) return r;
  if (entry_point_steps(c->P)) {
    r = read_scalars(c);
    if (r) return r;
    const int it0 = c->S_host->iteration;
    int rs = MHG_OK;
    for (int s = 0; s < max_steps && rs == MHG_OK; ++s) rs = step_entry_points(c, nullptr);
    r = read_scalars(c);
    if (r) return r;
    if (steps_done) *steps_done = c->S_host->iteration - it0;
    return rs;
  }
  r = read_scalars(c);
  if (r) return r;
  const int it0 = c->S_host->iteration;
  c->x_lazy_before = c->x_lazy;
  int launched = 0;
  while (launched < max_steps) {
    const int chunk = std::min(5 * kGraphSteps, max_steps - launched);
    r = enqueue_steps(c, chunk);
    if (r) return r;
    launched += chunk;
    r = read_scalars(c);
    if (r) return r;
    collect_times(c);
    if (c->S_host->stop || c->S_host->err) break;
  }
  // kernels of steps past the end of the run return at once; the host
  // bookkeeping (buffers, sweep order) must follow the steps really done
  const int done = c->S_host->iteration - it0;
  roll_back(c, launched - done, done);
  if (steps_done) *steps_done = done;
  if (c->S_host->err) return fail(c, MHG_E_DT_MIN, "dt below deltat-min");
  return c->S_host->stop ? MHG_STOP : MHG_OK;
}

int mhg_set_learned_alpha(mhg_ctx* c, const double* alpha) {
  if (!c || !alpha) return MHG_E_ARG;
  if (!c->P.learned) return fail(c, MHG_E_UNSUPPORTED, "set_learned_alpha: the context's limiter is not MHG_LIM_LEARNED");
  HIPCHK(c, put(c, c->alpha, alpha, 2 * (size_t)c->P.nc));
  return MHG_OK;
}

int mhg_conservation(mhg_ctx* c, double out[3]) {
  int r = check_ready(c);
  if (r) return r;
  const Params& P = c->P;
  const long NC = P.nc, NN = P.nn;
  std::vector<double> m(NC), e(NC), ux(NN), uy(NN), mn(NN);
  mhg_state st;
  std::memset(&st, 0, sizeof(st));
  std::vector<double> vel((size_t)P.dim * NN);
  st.cell_mass = m.data();
  st.internal_energy = e.data();
  st.velocity = vel.data();
  st.node_mass = mn.data();
  r = mhg_download_state(c, &st);
  if (r) return r;
  auto acc = [](double& s, double& comp, double v) {
    double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) comp += (s - t) + v;
    else comp += (v - t) + s;
    s = t;
  };
  double a = 0, ac = 0, b = 0, bc = 0, k = 0, kc = 0;
  for (long cc = 0; cc < NC; ++cc) {
    acc(a, ac, m[cc]);
    acc(b, bc, m[cc] * e[cc]);
  }
  for (long n = 0; n < NN; ++n) {
    const double u = vel[n], v = vel[NN + n], w = P.dim == 3 ? vel[2 * NN + n] : 0.;
    acc(k, kc, 0.5 * mn[n] * (u * u + v * v + w * w));
  }
  out[0] = a + ac;
  out[1] = b + bc;
  out[2] = k + kc;
  return MHG_OK;
}

// Autotuner: times every launch variant of the tunable kernels on this
// context's mesh and keeps the fastest.  The candidates write the sweep output
// buffers, the dual-flux scratch and the dual state, which hold no state
// between steps, and the boundary dual gradients, which are saved before and
// restored after; the simulation state is unchanged.
namespace {
const char* kTunable[] = {"cell_sweep_y", "node_sweep_x", "node_sweep_y", "step_graph", "final_fusion",
                          "lagrange_fusion", "store_policy", "ytile_width"};
int tunable_index(const char* kernel) {
  if (!kernel) return -1;
  for (int k = 0; k < 8; ++k)
    if (std::strcmp(kernel, kTunable[k]) == 0) return k;
  return -1;
}
int variant_count_of(int k) {
  if (k == 0) return cell_sweep_y_variants();
  if (k == 1) return node_sweep_variants(0);
  if (k == 2) return node_sweep_variants(1);
  if (k >= 3 && k <= 5) return 2;
  if (k == 6) return 16;  // one bit per kind of array (st_stream)
  if (k == 7) return 8;
  return 0;
}
const char* variant_name_of(int k, int v) {
  if (k == 3) return v ? "graph12" : "direct";
  if (k == 4 || k == 5) return v ? "fused" : "separate";
  if (k == 6) {
    static const char* names[16] = {"cached", "s1", "s2", "s3", "s4", "s5", "s6", "s7",
                                    "s8", "s9", "s10", "s11", "s12", "s13", "s14", "streaming"};
    return v >= 0 && v < 16 ? names[v] : "?";
  }
  if (k == 7) {  // 8 or 16 columns for the lag, final and dual y sweeps
    static const char* names[8] = {"l16f16n16", "l8f16n16", "l16f8n16", "l8f8n16",
                                   "l16f16n8",  "l8f16n8",  "l16f8n8",  "l8f8n8"};
    return names[v & 7];
  }
  return k == 0 ? cell_sweep_y_variant_name(v) : node_sweep_variant_name(k - 1, v);
}
int* variant_slot(mhg_ctx* c, int k) {
  if (k == 3) return &c->use_graphs;
  if (k == 4) return &c->fuse_final;
  if (k == 5) return &c->fuse_lag;
  if (k == 6) return &c->P.nt_store;
  if (k == 7) return &c->P.ytw8;
  return k == 0 ? &c->yvariant : (k == 1 ? &c->nxvariant : &c->nyvariant);
}
}  // namespace

extern "C++" {
namespace {
template <class R>
int autotune_t(mhg_ctx* c, int32_t reps, char* report, int32_t report_len) {
  const Params& P = c->P;
  CellSweepArgsT<R> A;
  A.Uin = bp<R>(c->U[0]);
  A.Uout = bp<R>(c->U[1]);
  A.Vlag = bp<R>(c->Vlag);
  A.xnx = c->xx[c->xcur];
  A.xny = c->xy[c->xcur];
  A.ux = bp<R>(c->ux[c->ucur]);
  A.uy = bp<R>(c->uy[c->ucur]);
  A.contrib = bp<R>(c->contrib);
  A.alpha = c->alpha ? bp<R>(c->alpha) + P.nc : nullptr;  // the y sweep's coefficients
  A.est = c->est;
  NodeSweepArgsT<R> B;
  B.m = bp<R>(c->m);
  B.ux = A.ux;
  B.uy = A.uy;
  B.xnx = A.xnx;
  B.xny = A.xny;
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
  const size_t gbytes = sizeof(R) * (size_t)P.nn;
  R* saved = nullptr;
  if (hipMalloc(&saved, 2 * gbytes) != hipSuccess) return fail(c, MHG_E_ALLOC, "autotune: hipMalloc failed");
  HIPCHK(c, hipMemcpyAsync(saved, c->dg0, gbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(saved + P.nn, c->dg1, gbytes, hipMemcpyDeviceToDevice, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  std::string rep;
  // the fused step forms node coordinates from velocities only with the
  // lane (x) and tile (y) dual sweeps; the column walkers and 5-node
  // stencils read stored coordinates, whose store costs the step more than
  // a dual sweep can gain, so they are not candidates there
  const bool formed = c->fuse_lag && fused_lag_supported(P) && c->fuse_final && fused_final_supported(P);
  for (int k = 0; k < 3; ++k) {
    auto launch = [&](int v) {
      if (k == 0) launch_cell_sweep_2d(c->stream, P, c->S, A, 1, true, v);
      else launch_node_sweep_2d(c->stream, P, c->S, B, k - 1, true, v);
    };
    rep += std::string(k ? " " : "") + kTunable[k] + ":";
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < variant_count_of(k); ++v) {
      if (formed && k >= 1 && v != (k == 1 ? 1 : 4)) continue;
      if (best_ms == 1e30f) best = v;
      for (int w = 0; w < 2; ++w) launch(v);
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (int n = 0; n < reps; ++n) launch(v);
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      char buf[96];
      snprintf(buf, sizeof(buf), " %s=%.4f", variant_name_of(k, v), ms);
      rep += buf;
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    *variant_slot(c, k) = best;
    drop_graphs(c);
    rep += std::string(" -> ") + variant_name_of(k, best) + ";";
  }
  HIPCHK(c, hipMemcpyAsync(c->dg0, saved, gbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->dg1, saved + P.nn, gbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  hipFree(saved);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return MHG_OK;
}
}  // namespace
}  // extern "C++"

int mhg_autotune(mhg_ctx* c, int32_t reps, char* report, int32_t report_len) {
  int r = check_ready(c);
  if (r) return r;
  r = check_not_batch(c);
  if (r) return r;
  if (reps < 1) reps = 1;
  if (c->P.dim == 3) {  // one launch shape per 3D kernel
    if (report && report_len > 0) std::strncpy(report, "3D: no launch variants", report_len - 1);
    return MHG_OK;
  }
  return c->bulk == 4 ? autotune_t<float>(c, reps, report, report_len)
                      : autotune_t<double>(c, reps, report, report_len);
}

int32_t mhg_variant_count(const char* kernel) { return variant_count_of(tunable_index(kernel)); }

int mhg_select_variant(mhg_ctx* c, const char* kernel, int32_t variant) {
  if (!c) return MHG_E_ARG;
  const int k = tunable_index(kernel);
  if (k < 0 || variant < 0 || variant >= variant_count_of(k))
    return fail(c, MHG_E_ARG, "select_variant: unknown kernel or variant");
  *variant_slot(c, k) = variant;
  drop_graphs(c);  // captured graphs hold the previous launch variants
  return MHG_OK;
}

int mhg_kernel_times(mhg_ctx* c, int32_t enable, const char** names, double* ms,
                     int32_t* launches, int32_t cap) {
  if (!c) return MHG_E_ARG;
  collect_times(c);
  const int n = (int)c->tnames.size();
  for (int i = 0; i < n && i < cap; ++i) {
    if (names) names[i] = c->tnames[i].c_str();
    if (ms) ms[i] = c->tcount[i] ? c->tsum[i] / c->tcount[i] : 0.;
    if (launches) launches[i] = c->tcount[i];
  }
  if (enable >= 0) {
    // any call with enable >= 0 resets the accumulators
    c->timing = enable != 0;
    for (int i = 0; i < n; ++i) {
      c->tsum[i] = 0.;
      c->tcount[i] = 0;
    }
  }
  return n;
}

}  // extern "C"
