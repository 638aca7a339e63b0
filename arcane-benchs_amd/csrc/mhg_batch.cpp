// mhg_batch.cpp -- host side of the batching form: many independent 2D
// meshes advanced by one set of persistent kernels.
//
// A batch holds a master context that owns the concatenated arrays, the
// stream and the buffer state (xcur, ucur, sens) shared by every running
// problem, plus one view context per problem whose pointers are the problem's
// slices; the views take the ordinary initialisation and download calls.
// Every running problem has run the same number of batch iterations, so the
// master's buffer state is theirs; a problem that stopped keeps its arrays as
// they were at its last iteration and its view derives its buffer state from
// its own iteration count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mhg_batch.hpp"
#include "mhg_ctx.hpp"

using namespace mhg;

struct mhg_batch {
  int n = 0;
  mhg_ctx* master = nullptr;    // owns every array and the stream
  std::vector<mhg_ctx*> views;  // problem contexts over slices of the arrays
  std::vector<Params> P;
  std::vector<long> cb, nb;
  std::vector<int> slot;
  long NC = 0, NN = 0;
  int nslot = 0;
  BatchProblem* dprobs = nullptr;
  DevScalars* S = nullptr;       // device, one per problem
  DevScalars* S_host = nullptr;  // pinned mirror
  BatchItem* items[BL_COUNT] = {};
  int* count = nullptr;          // device, BL_COUNT entries
  long cap[BL_COUNT] = {};
  BatchLaunch L{};
  size_t n_state = 0;            // master->owned[0, n_state): the simulation state
  // run bookkeeping
  bool started = false;
  std::vector<int> it0;          // iteration of each problem when the batch started
  std::vector<char> running;
  int sens0 = 0;
  // launch shapes (mhg_batch_autotune)
  int n_cu = 256;
  int mult = 4;                  // persistent blocks per CU of the 256-thread kernels
  int ycell_seg = 1, ynode_seg = 1;  // y sweeps in column segments where they fit
  int ymult = 2;                 // persistent blocks per CU of the 1024-thread kernels
  struct Graph {
    int xcur, ucur, sens;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  std::string err;
};

namespace {

constexpr int kBatchGraphSteps = 12;

int bfail(mhg_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIP(b, x)                                                    \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) return bfail(b, MHG_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

void drop_batch_graphs(mhg_batch* b) {
  for (auto& g : b->graphs) hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

char* off(void* p, long words, int bytes) { return static_cast<char*>(p) + words * (long)bytes; }

// Slices of the master's arrays for problem p.
void make_view(mhg_batch* b, int p, mhg_ctx* v) {
  mhg_ctx* m = b->master;
  const int w = m->bulk;
  const long c = b->cb[p], n = b->nb[p];
  const long ce = (long)b->P[p].ne * c, cu = (long)b->P[p].nvu * c;
  v->P = b->P[p];
  v->bulk = w;
  v->stream = m->stream;
  v->device = m->device;
  v->rho = off(m->rho, c, w);
  v->e = off(m->e, c, w);
  v->p = off(m->p, c, w);
  v->cs = off(m->cs, c, w);
  v->q = off(m->q, c, w);
  v->m = off(m->m, c, w);
  v->L = off(m->L, c, w);
  v->Vlag = off(m->Vlag, c, w);
  v->mask = m->mask + c;
  for (int k = 0; k < 2; ++k) {
    v->U[k] = off(m->U[k], cu, w);
    v->ux[k] = off(m->ux[k], n, w);
    v->uy[k] = off(m->uy[k], n, w);
    v->xx[k] = m->xx[k] + n;
    v->xy[k] = m->xy[k] + n;
  }
  v->contrib = off(m->contrib, ce, w);
  v->env_rho = off(m->env_rho, ce, w);
  v->env_e = off(m->env_e, ce, w);
  v->env_p = off(m->env_p, ce, w);
  v->env_cs = off(m->env_cs, ce, w);
  v->env_q = off(m->env_q, ce, w);
  v->env_m = off(m->env_m, ce, w);
  v->env_f = off(m->env_f, ce, w);
  v->env_mf = off(m->env_mf, ce, w);
  v->Ud0 = off(m->Ud0, n, w);
  v->Ud1 = off(m->Ud1, n, w);
  v->Ud3 = off(m->Ud3, n, w);
  v->dg0 = off(m->dg0, n, w);
  v->dg1 = off(m->dg1, n, w);
  v->block_min = m->block_min + b->slot[p];
  v->S = b->S + p;
  v->S_host = b->S_host + p;
  v->owner = b;
  v->owner_index = p;
}

int y_list(const mhg_batch* b, int p, bool cells) {
  const int seg = cells ? BL_CY_SEG : BL_NY_SEG;
  const bool use_seg = (cells ? b->ycell_seg : b->ynode_seg) && batch_blocks(b->P[p], seg) > 0;
  return use_seg ? seg : (cells ? BL_CY_TALL : BL_NY_TALL);
}

// Work lists of the running problems and the persistent grid sizes.
int build_lists(mhg_batch* b) {
  std::vector<BatchItem> lists[BL_COUNT];
  for (int p = 0; p < b->n; ++p) {
    if (!b->running[p]) continue;
    const int ls[7] = {BL_NODE, BL_CELL, BL_CX, BL_NX, y_list(b, p, true), y_list(b, p, false), BL_PROB};
    for (int l : ls) {
      const int nblk = batch_blocks(b->P[p], l);
      for (int k = 0; k < nblk; ++k) lists[l].push_back({p, k});
    }
  }
  int counts[BL_COUNT];
  for (int l = 0; l < BL_COUNT; ++l) {
    counts[l] = (int)lists[l].size();
    if (counts[l] > b->cap[l]) return bfail(b, MHG_E_ARG, "batch work list over capacity");
    if (counts[l]) BHIP(b, hipMemcpy(b->items[l], lists[l].data(), counts[l] * sizeof(BatchItem), hipMemcpyHostToDevice));
    b->L.items[l] = b->items[l];
    const bool big = (l >= BL_CY_SEG && l <= BL_NY_TALL);
    const long g = (long)(big ? b->ymult : b->mult) * b->n_cu;
    b->L.grid[l] = (int)std::max(1L, std::min(b->cap[l], g));
  }
  b->L.grid[BL_PROB] = std::max(1, std::min(b->n, b->n_cu));
  BHIP(b, hipMemcpy(b->count, counts, sizeof(counts), hipMemcpyHostToDevice));
  b->L.count = b->count;
  return MHG_OK;
}

template <class R>
void enqueue_batch_step_t(mhg_batch* b) {
  mhg_ctx* c = b->master;
  hipStream_t st = c->stream;
  launch_batch_node_lagrange_2d<R>(st, b->dprobs, b->S, node_lag_args<R>(c), b->L);
  launch_batch_cell_lagrange_2d<R>(st, b->dprobs, b->S, cell_lag_args<R>(c), b->L);
  for (int pass = 0; pass < 2; ++pass) {
    const int idir = (pass + c->sens) % 2;
    launch_batch_cell_sweep_2d<R>(st, b->dprobs, b->S, cell_sweep_args<R>(c, pass, idir), b->L, idir, pass == 0);
    launch_batch_node_sweep_2d<R>(st, b->dprobs, b->S, node_sweep_args<R>(c), b->L, idir, pass == 0);
  }
  launch_batch_cell_final_2d<R>(st, b->dprobs, b->S, cell_final_args<R>(c), b->L);
  c->sens = (c->sens + 1) % 2;
  c->xcur = 1 - c->xcur;
  c->ucur = 1 - c->ucur;
}

void enqueue_batch_step(mhg_batch* b) {
  if (b->master->bulk == 4) enqueue_batch_step_t<float>(b);
  else enqueue_batch_step_t<double>(b);
}

// n iterations: whole groups replay a graph captured per buffer state.
int enqueue_batch_steps(mhg_batch* b, int n) {
  mhg_ctx* c = b->master;
  while (n >= kBatchGraphSteps) {
    hipGraphExec_t exec = nullptr;
    for (auto& g : b->graphs)
      if (g.xcur == c->xcur && g.ucur == c->ucur && g.sens == c->sens) exec = g.exec;
    if (!exec) {
      mhg_batch::Graph g{c->xcur, c->ucur, c->sens, nullptr};
      BHIP(b, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < kBatchGraphSteps; ++s) enqueue_batch_step(b);
      hipGraph_t graph = nullptr;
      BHIP(b, hipStreamEndCapture(c->stream, &graph));
      const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      BHIP(b, e);
      b->graphs.push_back(g);
      exec = g.exec;
    }
    BHIP(b, hipGraphLaunch(exec, c->stream));
    n -= kBatchGraphSteps;
  }
  for (int s = 0; s < n; ++s) enqueue_batch_step(b);
  BHIP(b, hipGetLastError());
  return MHG_OK;
}

int read_all_scalars(mhg_batch* b) {
  BHIP(b, hipMemcpyAsync(b->S_host, b->S, b->n * sizeof(DevScalars), hipMemcpyDeviceToHost, b->master->stream));
  BHIP(b, hipStreamSynchronize(b->master->stream));
  return MHG_OK;
}

// First run: every problem initialised with the same sweep order; record the
// starting iterations and the running set.
int start_batch(mhg_batch* b) {
  for (int p = 0; p < b->n; ++p) {
    if (!b->views[p]->initialised) return bfail(b, MHG_E_STATE, "batch problem " + std::to_string(p) + " not initialised");
    if (b->views[p]->xcur != 0 || b->views[p]->ucur != 0 || b->views[p]->sens != b->views[0]->sens)
      return bfail(b, MHG_E_STATE, "batch problems must start from the same sweep order");
  }
  int r = read_all_scalars(b);
  if (r) return r;
  b->sens0 = b->views[0]->sens;
  mhg_ctx* c = b->master;
  c->xcur = 0;
  c->ucur = 0;
  c->sens = b->sens0;
  for (int p = 0; p < b->n; ++p) {
    b->it0[p] = b->S_host[p].iteration;
    b->running[p] = !(b->S_host[p].stop || b->S_host[p].err);
  }
  r = build_lists(b);
  if (r) return r;
  b->started = true;
  return MHG_OK;
}

long zone_cycles_now(const mhg_batch* b) {
  long z = 0;
  for (int p = 0; p < b->n; ++p) z += b->P[p].nc * (long)(b->S_host[p].iteration - b->it0[p]);
  return z;
}

}  // namespace

namespace mhg {

int batch_refresh_view(mhg_ctx* c) {
  mhg_batch* b = c->owner;
  if (!b->started) return MHG_OK;
  const int p = c->owner_index;
  HIPCHK(c, hipMemcpy(c->S_host, c->S, sizeof(DevScalars), hipMemcpyDeviceToHost));
  const int done = c->S_host->iteration - b->it0[p];
  c->xcur = done & 1;
  c->ucur = done & 1;
  c->sens = (b->sens0 + done) % 2;
  return MHG_OK;
}

bool batch_running(const mhg_ctx* c) { return c && c->owner && c->owner->started; }

}  // namespace mhg

extern "C" {

int mhg_batch_create(int32_t n, const mhg_mesh_desc* meshes, const mhg_options* opts, mhg_batch** out) {
  if (!out || n < 1 || !meshes || !opts) return MHG_E_ARG;
  *out = nullptr;
  mhg_batch* b = new mhg_batch();
  b->n = n;
  b->P.resize(n);
  for (int p = 0; p < n; ++p) {
    if (meshes[p].dim != 2) {
      delete b;
      return MHG_E_UNSUPPORTED;
    }
    const int r = build_params(&meshes[p], &opts[p], &b->P[p]);
    const mhg_options& o0 = opts[0];
    const mhg_options& o = opts[p];
    if (r || o.nb_env > 3 || o.nb_env != o0.nb_env || o.precision != o0.precision ||
        o.projection_limiteur_id == MHG_LIM_LEARNED || o.projection_limiteur_id == MHG_LIM_LEARNED_ALL ||
        o.conservation_energie_totale || o.ordre_projection == 3 || !o.with_projection || b->P[p].generic ||
        o.projection_limiteur_id != o0.projection_limiteur_id || o.ordre_projection != o0.ordre_projection ||
        o.pseudo_centree != o0.pseudo_centree) {
      delete b;
      return r ? r : MHG_E_UNSUPPORTED;
    }
  }
  // layout
  b->cb.resize(n);
  b->nb.resize(n);
  b->slot.resize(n);
  for (int p = 0; p < n; ++p) {
    b->cb[p] = b->NC;
    b->nb[p] = b->NN;
    b->slot[p] = b->nslot;
    b->NC += b->P[p].nc;
    b->NN += b->P[p].nn;
    b->nslot += batch_blocks(b->P[p], BL_CELL);
    for (int l = 0; l < BL_COUNT; ++l) b->cap[l] += batch_blocks(b->P[p], l);
  }
  {
    // one launch per phase covers every problem: the store policy follows
    // the whole batch's working set (as build_params does for one mesh)
    const double wb = opts[0].precision == MHG_PREC_FP32_BULK ? 4. : 8.;
    const double ws = (double)b->NC * wb * (6. * opts[0].nb_env + 20.) + (double)b->NN * 32.;
    for (int p = 0; p < n; ++p) b->P[p].nt_store = ws > 320e6 ? (1 << kStCell) : 0;
  }
  mhg_ctx* m = new mhg_ctx();
  b->master = m;
  m->opt = opts[0];
  m->P = b->P[0];
  m->P.dim = 2;
  m->bulk = opts[0].precision == MHG_PREC_FP32_BULK ? 4 : 8;
  if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) {
    mhg_batch_destroy(b);
    return MHG_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, m->device) == hipSuccess && prop.multiProcessorCount > 0)
    b->n_cu = prop.multiProcessorCount;
  const int ne = b->P[0].ne, nvu = b->P[0].nvu;
  const size_t NC = b->NC, NN = b->NN, w = m->bulk;
  auto ba = [&](void** d, size_t words) { return ctx_dalloc_bytes(m, d, std::max<size_t>(words, 1) * w); };
  int r = 0;
  r |= ba(&m->rho, NC) | ba(&m->e, NC) | ba(&m->p, NC) | ba(&m->cs, NC) | ba(&m->q, NC) | ba(&m->m, NC);
  r |= ba(&m->L, NC) | ba(&m->Vlag, NC) | ctx_dalloc(m, &m->mask, NC);
  r |= ba(&m->U[0], nvu * NC) | ba(&m->U[1], nvu * NC) | ba(&m->contrib, ne * NC);
  r |= ba(&m->env_rho, ne * NC) | ba(&m->env_e, ne * NC) | ba(&m->env_p, ne * NC) | ba(&m->env_cs, ne * NC);
  r |= ba(&m->env_q, ne * NC) | ba(&m->env_m, ne * NC) | ba(&m->env_f, ne * NC) | ba(&m->env_mf, ne * NC);
  for (int k = 0; k < 2; ++k) {
    r |= ba(&m->ux[k], NN) | ba(&m->uy[k], NN);
    r |= ctx_dalloc(m, &m->xx[k], NN) | ctx_dalloc(m, &m->xy[k], NN);
  }
  r |= ba(&m->Ud0, NN) | ba(&m->Ud1, NN) | ba(&m->Ud3, NN) | ba(&m->dg0, NN) | ba(&m->dg1, NN);
  r |= ctx_dalloc(m, &m->block_min, b->nslot) | ctx_dalloc(m, &b->S, n);
  b->n_state = m->owned.size();
  r |= ctx_dalloc(m, &b->dprobs, n) | ctx_dalloc(m, &b->count, BL_COUNT);
  for (int l = 0; l < BL_COUNT; ++l) r |= ctx_dalloc(m, &b->items[l], std::max(1L, b->cap[l]));
  if (r || hipHostMalloc((void**)&b->S_host, n * sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    mhg_batch_destroy(b);
    return MHG_E_ALLOC;
  }
  std::memset(b->S_host, 0, n * sizeof(DevScalars));
  std::vector<BatchProblem> hp(n);
  for (int p = 0; p < n; ++p) {
    hp[p].P = b->P[p];
    hp[p].cb = b->cb[p];
    hp[p].nb = b->nb[p];
    hp[p].slot = b->slot[p];
    hp[p].pad = 0;
  }
  if (hipMemcpy(b->dprobs, hp.data(), n * sizeof(BatchProblem), hipMemcpyHostToDevice) != hipSuccess) {
    mhg_batch_destroy(b);
    return MHG_E_HIP;
  }
  b->L.ne = ne;
  b->L.lim = b->P[0].lim;
  b->L.pc = b->P[0].pseudo_centree;
  b->views.resize(n, nullptr);
  b->it0.assign(n, 0);
  b->running.assign(n, 1);
  for (int p = 0; p < n; ++p) {
    b->views[p] = new mhg_ctx();
    b->views[p]->opt = opts[p];
    make_view(b, p, b->views[p]);
  }
  *out = b;
  return MHG_OK;
}

int mhg_batch_destroy(mhg_batch* b) {
  if (!b) return MHG_OK;
  if (b->master && b->master->stream) hipStreamSynchronize(b->master->stream);
  drop_batch_graphs(b);
  for (mhg_ctx* v : b->views) delete v;  // views own no memory
  if (b->master) {
    mhg_ctx* m = b->master;
    for (auto& d : m->owned) hipFree(d.p);
    m->owned.clear();
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
  }
  if (b->S_host) hipHostFree(b->S_host);
  delete b;
  return MHG_OK;
}

const char* mhg_batch_last_error(const mhg_batch* b) { return b ? b->err.c_str() : "null batch"; }

mhg_ctx* mhg_batch_problem(mhg_batch* b, int32_t i) {
  if (!b || i < 0 || i >= b->n) return nullptr;
  return b->views[i];
}

int mhg_batch_run(mhg_batch* b, int32_t max_steps, int32_t* n_running, int64_t* zone_cycles) {
  if (!b) return MHG_E_ARG;
  int r = MHG_OK;
  if (!b->started) {
    r = start_batch(b);
    if (r) return r;
  }
  r = read_all_scalars(b);
  if (r) return r;
  const long z0 = zone_cycles_now(b);
  int nrun = 0;
  for (int p = 0; p < b->n; ++p) nrun += b->running[p];
  int launched = 0;
  while (launched < max_steps && nrun > 0) {
    const int chunk = std::min(kBatchGraphSteps, max_steps - launched);
    r = enqueue_batch_steps(b, chunk);
    if (r) return r;
    launched += chunk;
    r = read_all_scalars(b);
    if (r) return r;
    bool changed = false;
    for (int p = 0; p < b->n; ++p)
      if (b->running[p] && (b->S_host[p].stop || b->S_host[p].err)) {
        b->running[p] = 0;
        --nrun;
        changed = true;
      }
    if (changed && nrun > 0) {
      r = build_lists(b);
      if (r) return r;
    }
  }
  // steps past a problem's end returned at once in the kernels; the master's
  // buffer state follows the iterations of the problems still running
  if (n_running) *n_running = nrun;
  if (zone_cycles) *zone_cycles = zone_cycles_now(b) - z0;
  for (int p = 0; p < b->n; ++p)
    if (b->S_host[p].err) return bfail(b, MHG_E_DT_MIN, "problem " + std::to_string(p) + ": dt below deltat-min");
  return nrun == 0 ? MHG_STOP : MHG_OK;
}

int mhg_batch_status(mhg_batch* b, int32_t i, int32_t* iterations, double* time, double* dt, int32_t* state) {
  if (!b || i < 0 || i >= b->n) return MHG_E_ARG;
  DevScalars s;
  BHIP(b, hipMemcpy(&s, b->S + i, sizeof(DevScalars), hipMemcpyDeviceToHost));
  if (iterations) *iterations = s.iteration;
  if (time) *time = s.time;
  if (dt) *dt = s.dt;
  if (state) *state = s.err ? 2 : (s.stop ? 1 : 0);
  return MHG_OK;
}

// Autotuner: the batch's arrays are copied aside, each candidate launch shape
// runs kBatchGraphSteps real iterations from that copy (timed with events),
// and the copy is put back, so the simulation state is unchanged.  The shapes
// are tuned one after the other: persistent blocks per CU of the 256-thread
// kernels, then column segments or 64-row tiles and blocks per CU of the y
// cell and dual sweeps.
int mhg_batch_autotune(mhg_batch* b, int32_t reps, char* report, int32_t report_len) {
  if (!b) return MHG_E_ARG;
  mhg_ctx* m = b->master;
  for (int p = 0; p < b->n; ++p)
    if (!b->views[p]->initialised) return bfail(b, MHG_E_STATE, "autotune: batch problem not initialised");
  if (reps < 1) reps = 1;
  const bool was_started = b->started;
  const std::vector<int> it0 = b->it0;
  const std::vector<char> running = b->running;
  if (!was_started) {
    for (int p = 0; p < b->n; ++p) b->running[p] = 1;
    m->xcur = 0;
    m->ucur = 0;
    m->sens = b->views[0]->sens;
  }
  const int xs = m->xcur, us = m->ucur, ss = m->sens;
  // snapshot the simulation state (not the work lists)
  std::vector<void*> snap(b->n_state, nullptr);
  for (size_t k = 0; k < b->n_state; ++k) {
    if (hipMalloc(&snap[k], m->owned[k].bytes) != hipSuccess) {
      for (void* q : snap)
        if (q) hipFree(q);
      return bfail(b, MHG_E_ALLOC, "autotune: snapshot allocation failed");
    }
    BHIP(b, hipMemcpyAsync(snap[k], m->owned[k].p, m->owned[k].bytes, hipMemcpyDeviceToDevice, m->stream));
  }
  auto restore = [&]() -> int {
    for (size_t k = 0; k < b->n_state; ++k)
      BHIP(b, hipMemcpyAsync(m->owned[k].p, snap[k], m->owned[k].bytes, hipMemcpyDeviceToDevice, m->stream));
    BHIP(b, hipStreamSynchronize(m->stream));
    m->xcur = xs;
    m->ucur = us;
    m->sens = ss;
    return MHG_OK;
  };
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  std::string rep;
  int rc = MHG_OK;
  auto time_shape = [&](float* ms_out) -> int {
    int r = build_lists(b);
    if (r) return r;
    drop_batch_graphs(b);
    float best = 1e30f;
    for (int k = 0; k < reps; ++k) {
      r = restore();
      if (r) return r;
      BHIP(b, hipEventRecord(e0, m->stream));
      r = enqueue_batch_steps(b, kBatchGraphSteps);
      if (r) return r;
      BHIP(b, hipEventRecord(e1, m->stream));
      BHIP(b, hipEventSynchronize(e1));
      float ms = 0.f;
      BHIP(b, hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms / kBatchGraphSteps);
    }
    *ms_out = best;
    return MHG_OK;
  };
  struct Knob {
    const char* name;
    std::vector<std::pair<int, int>> cand;  // (segments, blocks per CU)
  };
  const std::vector<int> mults = {1, 2, 4, 8};
  const std::vector<std::pair<int, int>> ycand = {{1, 1}, {1, 2}, {1, 3}, {0, 1}, {0, 2}, {0, 3}};
  char buf[128];
  // 1. blocks per CU of the 256-thread kernels
  {
    rep += "grid256:";
    int best = b->mult;
    float best_ms = 1e30f;
    for (int mu : mults) {
      b->mult = mu;
      float ms = 0.f;
      rc = time_shape(&ms);
      if (rc) break;
      snprintf(buf, sizeof(buf), " %dx=%.4f", mu, ms);
      rep += buf;
      if (ms < best_ms) {
        best_ms = ms;
        best = mu;
      }
    }
    b->mult = best;
    snprintf(buf, sizeof(buf), " -> %dx;", best);
    rep += buf;
  }
  // 2. y sweeps: segments or tiles, blocks per CU (cells and dual together)
  if (!rc) {
    rep += " ysweeps:";
    int best_seg = b->ycell_seg, best_mu = b->ymult;
    float best_ms = 1e30f;
    for (auto& cnd : ycand) {
      b->ycell_seg = b->ynode_seg = cnd.first;
      b->ymult = cnd.second;
      float ms = 0.f;
      rc = time_shape(&ms);
      if (rc) break;
      snprintf(buf, sizeof(buf), " %s%dx=%.4f", cnd.first ? "seg" : "tile64_", cnd.second, ms);
      rep += buf;
      if (ms < best_ms) {
        best_ms = ms;
        best_seg = cnd.first;
        best_mu = cnd.second;
      }
    }
    b->ycell_seg = b->ynode_seg = best_seg;
    b->ymult = best_mu;
    snprintf(buf, sizeof(buf), " -> %s%dx;", best_seg ? "seg" : "tile64_", best_mu);
    rep += buf;
  }
  const int rr = restore();
  for (void* q : snap) hipFree(q);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->it0 = it0;
  b->running = running;
  drop_batch_graphs(b);
  if (rc) return rc;
  if (rr) return rr;
  if (was_started) {
    const int r = build_lists(b);
    if (r) return r;
  }
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return MHG_OK;
}

}  // extern "C"
