// mhg_body2d.hpp -- bodies of the 2D (quad) step kernels.
//
// Each body runs one thread block's share of a kernel for the block index it
// is given: the single-mesh kernels pass their (XCD-swizzled) block index,
// the persistent batch kernels pass the local index of a work item of one
// problem.  Everything a body reads about its mesh comes from Params, its
// scalars from DevScalars and its arrays from the argument block.
#pragma once

#include <type_traits>

#include "mhg_common.hpp"
#include "mhg_kernels.hpp"
#include "mhg_kernels_impl.hpp"

namespace mhg {

__device__ inline double node_x0(const Params& P, int i) { return P.org[0] + i * P.h[0]; }
__device__ inline double node_y0(const Params& P, int j) { return P.org[1] + j * P.h[1]; }

// Lagrangian coordinate along axis A of a node with index idx along A, from
// its Lagrangian velocity component u.  In the Euler scheme every node starts
// the Lagrange phase from its initial position (saveValuesAtN resets x to x0,
// MahycoModule.cc:421, 433-441) and moves by dt u (updatePosition, 870-931):
// the node phase stores x0 + dt u with the stored velocity, so these are the
// same operations and the same double, bit for bit.  The remap sweeps form
// their coordinates this way from the velocities they read anyway instead of
// reading the stored coordinates.
template <int A, class R>
__device__ inline double lag_coord(const Params& P, int idx, double dt, R u) {
  return (P.org[A] + idx * P.h[A]) + dt * (double)u;
}

// computeDeltaT (MahycoModule.cc:2076-2191) on the minimum of the CFL
// candidates: dt = min(cfl * min_c dx_sound, 1.05 dt_old, deltat-max),
// deltat-min check, stop test with the time of this iteration.  ADVANCE
// first applies this iteration's time += dt, iteration += 1 (the work of
// k_begin_iteration, which no kernel of the step between them reads).  The
// scalars after the update; a run that has ended leaves them as they are.
template <bool ADVANCE>
__device__ __forceinline__ DevScalars dt_update(const Params& P, DevScalars s, double minimum_aux) {
  if (s.stop || s.err) return s;
  const double old = s.dt;
  const double t = ADVANCE ? s.time + old : s.time;
  if (ADVANCE) {
    s.time = t;
    s.iteration += 1;
  }
  s.dt_old = old;
  double new_dt = P.cfl * minimum_aux;
  new_dt = smin(new_dt, 1.05 * old);
  new_dt = smin(new_dt, P.dtmax);
  if (new_dt < P.dtmin) {
    s.err = 1;
    return s;
  }
  if (P.land) {  // extension: the last step ends at tend
    if (s.landing || t >= P.tend) s.stop = 1;
    else if (t + new_dt >= P.tend) {
      new_dt = P.tend - t;
      s.landing = 1;
    }
    s.dt = new_dt;
    return s;
  }
  s.dt = new_dt;
  const bool not_yet_finish = (t < P.tend);
  const bool finish = (t > P.tend);
  const bool too_much = ((t + new_dt) > P.tend);
  if ((not_yet_finish && too_much) || finish) s.stop = 1;
  return s;
}

// Arguments of the node phase with the previous step's dt kernel folded in:
// the previous CFL kernel's nb block minima and the scalars to write.
struct DtFold {
  const double* block_min;
  int nb;
  DevScalars* Sout;
};

// The folded dt kernel: the block's minimum of the CFL minima (w: this
// thread's share), computeDeltaT's update of s (the same in every block),
// written to F->Sout by block 0.
template <int BLOCK>
__device__ __forceinline__ DevScalars fold_dt(const Params& P, const DevScalars& s, double w, const DtFold* F) {
  __shared__ double red[BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) w = smin(w, __shfl_down(w, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int q = 1; q < BLOCK / 64; ++q) m = smin(m, red[q]);
  const DevScalars n = dt_update<true>(P, s, m);
  if (blockIdx.x == 0 && threadIdx.x == 0) *F->Sout = n;
  return n;
}

// ---------------------------------------------------------------------------
// K1: node phase.  updateForceAndVelocity (MahycoModule.cc:589-665) with the
// corner normals of the previous Lagrangian coordinates (saveValuesAtN copies
// m_cell_cqs computed after the previous updatePosition, 344/365),
// applyBoundaryCondition (936-1004), updatePosition (870-931) from x0 (from
// the previous coordinates with with-projection = false).
// Node mass = 1/4 sum of the cell masses (computeNodeMass 235-266): the cell
// masses do not change in the Lagrange phase, so it is recomputed here.
// ---------------------------------------------------------------------------
// FOLD (with VEL): the previous step's dt kernel folded in.  S holds the
// scalars before computeDeltaT's update; the block loads the previous CFL
// kernel's minima with its other loads, forms the coordinates with dt_old =
// S->dt (the value the update gives it), and reduces the minima before the
// velocity update, which needs the new dt (fold_dt).  Every thread of the
// block reaches that reduction.
template <class R, bool VEL, bool BC, bool POS, bool FOLD = false>
__device__ __forceinline__ void node_lagrange_2d_body(const Params& P, DevScalars* S, const NodeLagArgsT<R>& A,
                                                      long blk, const DtFold* F = nullptr) {
  bool halt = halted(S);
  const long nid0 = blk * blockDim.x + threadIdx.x;
  const bool live = nid0 < P.nn;
  if (!FOLD && !live) return;
  const long nid = live ? nid0 : 0;
  const int nnx = P.nx + 1;
  int i, j;
  divmod32(nid, nnx, j, i);
  double dt = S->dt;
  double wmin = 1.7976931348623157e308;  // FOLD: this thread's share of the minima
  if (FOLD) {
    const int b0 = threadIdx.x, b1 = threadIdx.x + 256;
    const double w0 = F->block_min[b0 < F->nb ? b0 : 0], w1 = F->block_min[b1 < F->nb ? b1 : 0];
    if (b0 < F->nb) wmin = smin(wmin, w0);
    if (b1 < F->nb) wmin = smin(wmin, w1);
    for (int b = threadIdx.x + 512; b < F->nb; b += 256) wmin = smin(wmin, F->block_min[b]);
  }
  // the velocity update in the bulk type (FP64: the same operations); the
  // corner normals stay in double
  R vx = A.ux[nid], vy = A.uy[nid];
  if (VEL) {
    double dtv = 0.5 * (S->dt_old + S->dt);
    // Every load first, from clamped indices and without branches, so that
    // a wavefront waits for memory once; out-of-mesh patch nodes read node
    // 0 and are then zeroed, out-of-mesh cells read cell 0 and are skipped.
    // 3x3 patch of previous coordinates around the node.
    double X[3][3], Y[3][3];
    bool nv[3][3];
    long pq[3][3];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        const int ii = i + di, jj = j + dj;
        const bool ok = ii >= 0 && ii < nnx && jj >= 0 && jj <= P.ny;
        pq[dj + 1][di + 1] = ok ? ii + (long)nnx * jj : 0;
        nv[dj + 1][di + 1] = ok;
      }
    // one branch, uniform over the launch, around straight-line loads
    if (A.upx != nullptr) {  // x0 + dt_old u of the previous Lagrange phase (lag_coord)
      const double dto = FOLD ? S->dt : S->dt_old;
      R ux9[3][3], uy9[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          ux9[a][b] = A.upx[pq[a][b]];
          uy9[a][b] = A.upy[pq[a][b]];
        }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          X[a][b] = lag_coord<0>(P, i + b - 1, dto, ux9[a][b]);
          Y[a][b] = lag_coord<1>(P, j + a - 1, dto, uy9[a][b]);
        }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          X[a][b] = A.xpx[pq[a][b]];
          Y[a][b] = A.xpy[pq[a][b]];
        }
    }
    // cells in ascending id: (i-1,j-1) corner 2, (i,j-1) 3, (i-1,j) 1, (i,j) 0
    const int cdi[4] = {-1, 0, -1, 0}, cdj[4] = {-1, -1, 0, 0}, cq[4] = {2, 3, 1, 0};
    R cp[4], cqv[4], cm[4];
    bool cv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ci = i + cdi[t], cj = j + cdj[t];
      cv[t] = ci >= 0 && ci < P.nx && cj >= 0 && cj < P.ny;
      const long c = cv[t] ? ci + (long)P.nx * cj : 0;
      cp[t] = A.p[c];
      cqv[t] = A.q[c];
      cm[t] = A.m[c];
    }
#pragma unroll
    for (int dj = 0; dj < 3; ++dj)
#pragma unroll
      for (int di = 0; di < 3; ++di)
        if (!nv[dj][di]) X[dj][di] = Y[dj][di] = 0.;
    R Fx = R(0), Fy = R(0), msum = R(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (!cv[t]) continue;
      // quad nodes (ci,cj),(ci+1,cj),(ci+1,cj+1),(ci,cj+1) in patch coordinates
      int bi = cdi[t] + 1, bj = cdj[t] + 1;
      double qx[4] = {X[bj][bi], X[bj][bi + 1], X[bj + 1][bi + 1], X[bj + 1][bi]};
      double qy[4] = {Y[bj][bi], Y[bj][bi + 1], Y[bj + 1][bi + 1], Y[bj + 1][bi]};
      double cx, cy;
      cqs2d(qx, qy, cq[t], &cx, &cy);
      const R pq = cp[t] + cqv[t];
      Fx += pq * (R)cx;
      Fy += pq * (R)cy;
      msum += cm[t];
    }
    if (FOLD) {
      const DevScalars sn = fold_dt<256>(P, *S, wmin, F);
      halt = sn.stop || sn.err;
      dt = sn.dt;
      dtv = 0.5 * (sn.dt_old + sn.dt);
    }
    const R mn = R(0.25) * msum;
    const R s = qdiv((R)dtv, mn);
    vx = vx + s * Fx;
    vy = vy + s * Fy;
  }
  if (BC) {
    for (int b = 0; b < P.nbc; ++b) {
      const BcRange& r = P.bc[b];
      int idx = (r.axis == 0) ? i : j;
      if (r.axis < 0 || (idx >= r.lo && idx <= r.hi)) {
        if (r.comp == 0) vx = r.value;
        else if (r.comp == 1) vy = r.value;
      }
    }
  }
  if (halt || !live) return;
  if (VEL || BC) {
    st_stream<3>(P, &A.ux[nid], (R)vx);
    st_stream<3>(P, &A.uy[nid], (R)vy);
  }
  if (POS) {
    // Euler scheme: saveValuesAtN reset x to x0; Lagrange-only: x moves on
    const double bx = P.lagonly ? A.xpx[nid] : node_x0(P, i);
    const double by = P.lagonly ? A.xpy[nid] : node_y0(P, j);
    // from the velocity as stored (the identity in double), so that the
    // sweeps' x0 + dt u from the stored velocity is the same value
    st_stream<3>(P, &A.xnx[nid], bx + dt * (double)(R)vx);
    st_stream<3>(P, &A.xny[nid], by + dt * (double)(R)vy);
  }
}

// ---------------------------------------------------------------------------
// Environment bound.  Every cell kernel is compiled for a maximum number of
// environments NEM (1, 2, 3, 4 or 8) so that per-environment values live in
// registers; the run-time count ne <= NEM guards loads and stores.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// K2: cell phase.  computeGeometricValues (1054-1299), updateDensity
// (1331-1426), computeArtificialViscosity (454-578),
// updateEnergyAndPressureforGP (1659-1770), then the remap vector
// computeVariablesForRemap_PBorn0 (PrepareRemap.cc:258-399).  Pure cells use
// the cell arrays; cells with >= 2 environments loop over their environments
// in id order.  Only the live entries of U are kept: the momentum and kinetic
// energy slots (PrepareRemap.cc:324-328) are identically zero and stay zero
// through the sweeps (phi = 0 gives a zero gradient, face value and flux).
// ---------------------------------------------------------------------------
template <class R, bool PC, int NEM>
__device__ __forceinline__ void cell_lagrange_2d_body(const Params& P, DevScalars* S, const CellLagArgsT<R>& A,
                                                      long blk) {
  if (S->stop || S->err) return;
  const long c = blk * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  const int nnx = P.nx + 1;
  int i, j;
  divmod32(c, P.nx, j, i);
  const long n0 = i + (long)nnx * j;
  const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
  const CellPre<R> in = cell_pre_load<R, PC>(A, c);
  double x[4], y[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = A.xnx[nd[q]];
    y[q] = A.xny[nd[q]];
  }
  const double Vd = quad_volume(x, y);
  const double Ld = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(x, y) : length_sqrt_2d_r<R>(Vd);
  lagrange_cell_bulk<R, PC, NEM>(P, S, A, c, Vd, Ld, in);
}

// ---------------------------------------------------------------------------
// Sweep arithmetic shared by both directions, for one variable of one cell:
// computeGradPhiFace (RemapADIService.cc:171-177), the limited cell gradient
// (497-506), the upwind face value (707-714) and the flux update (909-936).
// ---------------------------------------------------------------------------
// Lane sweep: one wavefront holds 64 consecutive cells along the sweep
// direction D (60 outputs and two halo cells on each side); stencil
// neighbours move between lanes.  D = 0: a wave is a row segment.  D = 1: a
// wave is a column segment and a block of 16 waves covers 16 adjacent
// columns, so each cache line of a row is consumed by the whole block.
// LEARN: the cell slopes of the learned limiter (learned_grad) with the
// cell's coefficient A.alpha[c]; LIM is then minmod, the fallback.
// Width of a quadrilateral cell along a sweep direction (computeGradPhiFace,
// HCELL_BY_CELLS, RemapADIService.cc:223-255): the distances from the cell
// centre to the centres of its back and front faces, a face on the mesh
// boundary left out.  a[4], t[4]: coordinates along and across the sweep of
// the nodes (i,j), (i+1,j), (i+1,j+1), (i,j+1); back face nodes b0, b1, front
// face nodes f0, f1; first / last: the cell is the first / last along the
// sweep.
__device__ __forceinline__ double cell_width_2d(const double* a, const double* t, int b0, int b1, int f0, int f1,
                                                bool first, bool last) {
  const double ca = 0.25 * (((a[0] + a[1]) + a[2]) + a[3]);
  const double ct = 0.25 * (((t[0] + t[1]) + t[2]) + t[3]);
  const double ba = (0.5 * a[b0] + 0.5 * a[b1]) - ca, bt = (0.5 * t[b0] + 0.5 * t[b1]) - ct;
  const double fa = (0.5 * a[f0] + 0.5 * a[f1]) - ca, ft = (0.5 * t[f0] + 0.5 * t[f1]) - ct;
  const double db = sqrt(ba * ba + bt * bt + 0. * 0.), df = sqrt(fa * fa + ft * ft + 0. * 0.);
  return (first ? 0. : db) + (last ? 0. : df);
}

// FINAL: the last sweep of the remap fused with the final cell phase
// (remapVariables, EOS, pressure average, k_cell_final_2d's arithmetic on
// the values it would have read back): the new remap vector stays in
// registers and the cell state goes to F.
// LAG (with FIRST): the first sweep of the remap fused with the Lagrange cell
// phase (k_cell_lagrange_2d's arithmetic on every lane, halo included): the
// remap vector comes from registers, and the output lanes store V and L
// through LA.
template <class R, int D, int NEM, bool FIRST, int LIM, bool LEARN = false, bool FINAL = false, bool LAG = false>
__device__ __forceinline__ void cell_sweep_lane_2d_body(const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                                                        int blk, const CellFinalArgsT<R>* F = nullptr,
                                                        const CellLagArgsT<R>* LA = nullptr) {
  const bool halt = halted(S);
  // halo lanes on each side: 3 for the order-3 stencil (six cells per face)
  constexpr int HL = (LIM == kLimO3) ? 3 : 2, NO = 64 - 2 * HL;
  const int lane = threadIdx.x & 63;
  const int N = (D == 0) ? P.nx : P.ny;      // cells along D
  const int T = (D == 0) ? P.ny : P.nx;      // cells across D
  const int nseg = (N + NO - 1) / NO;
  int seg, line;                             // segment along D, index across D
  if (D == 0) {
    const long wave = ((long)blk * blockDim.x + threadIdx.x) >> 6;
    divmod32(wave, nseg, line, seg);
  } else {
    const int ncolblk = (T + 15) / 16;
    line = (blk % ncolblk) * 16 + (threadIdx.x >> 6);
    seg = blk / ncolblk;
  }
  if (line >= T) return;  // uniform across the wave
  const int s = seg * NO - HL + lane;
  const bool ex = (s >= 0 && s < N);
  const bool out = ex && lane >= HL && lane <= 63 - HL;
  const int i = (D == 0) ? s : line, j = (D == 0) ? line : s;
  const int nnx = P.nx + 1;
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const long c = ex ? i + (long)P.nx * j : 0;
  const R* __restrict__ Uin = A.Uin;
  const R* __restrict__ Ud = (D == 0) ? A.ux : A.uy;          // velocity along D
  const R* __restrict__ Ut = (D == 0) ? A.uy : A.ux;          // transverse velocity
  // Every load of the cell is issued first, from clamped indices (a lane
  // out of the mesh reads cell / node 0), so that a wavefront waits for
  // memory once; the values are selected afterwards.
  const long m0 = ex ? i + (long)nnx * j : 0;
  // back face of the cell: the face at node index s along D, nodes
  // (i,j),(i,j+1) for D = 0 and (i,j),(i+1,j) for D = 1
  const bool fin = s >= 0 && s <= N;
  const long na = fin ? i + (long)nnx * j : 0, nb = (D == 0) ? na + nnx : na + 1;
  const R uda = Ud[na], udb = Ud[nb], uta = Ut[na], utb = Ut[nb];
  // the cell's two nodes off its back face: (i+1,j), (i+1,j+1) for D = 0,
  // (i+1,j+1), (i,j+1) for D = 1
  const long on1 = (D == 0) ? m0 + 1 : m0 + 1 + nnx, on2 = (D == 0) ? m0 + 1 + nnx : m0 + nnx;
  const R udon1 = Ud[on1], udon2 = Ud[on2];
  R uton1 = R(0), uton2 = R(0);
  if (LIM == kLimG || LIM == kLimO3 || LAG) {
    uton1 = Ut[on1];
    uton2 = Ut[on2];
  }
  R vlag = (FIRST && !LAG) ? A.Vlag[c] : R(1);
  const int fmask = FINAL ? F->mask[c] : 0;  // FINAL: membership mask, loaded with the rest
  // Lagrangian coordinates x0 + dt u (lag_coord); the back-face nodes (the
  // cell's first node and, along the face, its second) are the face's
  const double xfa = lag_coord<D>(P, s, dt, uda), xfb = lag_coord<D>(P, s, dt, udb);
  const double xta = lag_coord<1 - D>(P, line, dt, uta), xtb = lag_coord<1 - D>(P, line + 1, dt, utb);
  const double xon1 = lag_coord<D>(P, s + 1, dt, udon1), xon2 = lag_coord<D>(P, s + 1, dt, udon2);
  // the cell's nodes (i,j), (i+1,j), (i+1,j+1), (i,j+1): coordinate along D
  const double xa = xfa, xb = (D == 0) ? xon1 : xfb, xcc = (D == 0) ? xon2 : xon1, xdd = (D == 0) ? xfb : xon2;
  // generalised limiters, order 3, LAG: coordinates across D of the cell's nodes
  double tn[4] = {0., 0., 0., 0.};
  if (LIM == kLimG || LIM == kLimO3 || LAG) {
    const double ton1 = lag_coord<1 - D>(P, (D == 0) ? line : line + 1, dt, uton1);
    const double ton2 = lag_coord<1 - D>(P, (D == 0) ? line + 1 : line, dt, uton2);
    tn[0] = xta;
    tn[1] = (D == 0) ? ton1 : xtb;
    tn[2] = (D == 0) ? ton2 : ton1;
    tn[3] = (D == 0) ? xtb : ton2;
  }
  R pre[3][NEM];
  R preq = R(0);
  if (LAG) {  // the Lagrange cell phase of the lane's cell
    const CellPre<R> lin = cell_pre_load<R, false>(*LA, c);
    const double an[4] = {xa, xb, xcc, xdd};
    double xq[4], yq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      xq[q] = D == 0 ? an[q] : tn[q];
      yq[q] = D == 0 ? tn[q] : an[q];
    }
    const double Vd = quad_volume(xq, yq);
    const double Ld = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(xq, yq) : length_sqrt_2d_r<R>(Vd);
    R uv[NEM], um[NEM], ue[NEM], uq;
    lagrange_cell_vals<R, false, NEM>(P, S, *LA, c, (R)Vd, (R)Ld, lin, uv, um, ue, uq);
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      pre[0][k] = (ex && k < ne) ? uv[k] : R(0);
      pre[1][k] = (ex && k < ne) ? um[k] : R(0);
      pre[2][k] = (ex && k < ne) ? ue[k] : R(0);
    }
    preq = ex ? uq : R(0);
    vlag = (R)Vd;
    if (halt) return;
    if (out) {
      st_stream<0>(P, &LA->Vlag[c], (R)Vd);
      st_stream<0>(P, &LA->L[c], (R)Ld);
    }
  } else {
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        const R u = (k < ne) ? Uin[(g * ne + k) * NC + c] : R(0);
        pre[g][k] = (ex && k < ne) ? u : R(0);
      }
    const R preq_l = Uin[(3 * ne) * NC + c];
    preq = ex ? preq_l : R(0);
    if (halt) return;
  }
  // cell centre (component D, double) and phi denominator
  double xc = 0.;
  R den = R(1);
  if (ex) {
    xc = 0.25 * (((xa + xb) + xcc) + xdd);
    if (FIRST) {
      den = vlag;
    } else {
      R sv = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k < ne) sv += pre[0][k];
      den = sv;
    }
  }
  R uf = R(0);
  double lf = 0., xf = 0.;
  if (fin) {
    uf = R(0.5) * (uda + udb);
    xf = 0.5 * xfa + 0.5 * xfb;
    lf = fabs(xtb - xta);
  }
  const bool inner = ex && s >= 1;
  const double xcm = lane_up1(xc);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = lane_down1(uf);
  const double lf_f = lane_down1(lf);
  // face flux factors u_f L_f dt (flux_factor)
  const R fb = flux_factor<R>(-uf, lf, dt), ff = flux_factor<R>(uf_f, lf_f, dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  const bool o2 = P.order2 > 0;
  // learned limiter: coefficient, front-neighbour existence, face offsets
  R alpha = R(0), dbr = R(0), dfr = R(0);
  bool front_in = false;
  if (LEARN) {
    const R al = A.alpha[c];
    alpha = ex ? al : R(0);
    front_in = lane_down1((int)inner) != 0;
    dbr = (R)(xf - xc);
    dfr = (R)(lane_down1(xf) - xc);
  }
  // generalised limiters: widths of the cell and of its back / front
  // neighbours along D (the cell itself where the neighbour is missing)
  R hc = R(0), hcm = R(0), hcp = R(0);
  R h6[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};  // order 3: widths bbb .. fff of the back face
  if (LIM == kLimG || LIM == kLimO3) {
    front_in = lane_down1((int)inner) != 0;
    const double an[4] = {xa, xb, xcc, xdd};
    const double h = ex ? (D == 0 ? cell_width_2d(an, tn, 0, 3, 1, 2, s == 0, s == N - 1)
                                  : cell_width_2d(an, tn, 0, 1, 3, 2, s == 0, s == N - 1))
                        : 0.;
    const double hm = lane_up1(h), hp = lane_down1(h);
    hc = (R)h;
    hcm = (R)(inner ? hm : h);
    hcp = (R)(front_in ? hp : h);
    if (LIM == kLimO3) o3_stencil<R>((R)h, s, N, h6);
  }
  // order 3: the cell's mixed / pure code (-2 mixed, else its environment,
  // -1 none) from the previous step's final fractions in the first sweep,
  // else from phi_V > 0 (computeUremap), and the face velocity times dt
  int est = 0;
  R vdt = R(0);
  if (LIM == kLimO3) {
    if (FIRST) {
      const int el = A.est[c];
      est = ex ? el : 0;
    } else {
      int cnt = 0, last = -1;
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k < ne && qdiv(pre[0][k], den) > R(0)) {
          cnt++;
          last = k;
        }
      est = cnt > 1 ? -2 : last;
    }
    vdt = vel_dt<R>(uf, dt);
  }
  int estb = 0;
  if (LIM == kLimO3) estb = lane_up1(est);
  const bool vpure = estb != -2 && est != -2 && estb == est;
  R o3V[NEM], o3Vb[NEM], o3Vf[NEM], o3M[NEM], o3Mb[NEM], o3Mf[NEM];
  int o3up[NEM];
  EnvSlopes<R, NEM> es(P.remap_threshold);  // learned mode: consistent environment slopes
  // role: 0 volume, 1 mass, 2 energy of environment k, 3 the other variable
  auto sweep_var = [&](R u0, R& unew, R& dual, int role, int k) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = lane_up1(phi);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = lane_down1(gf);
    R phif;
    if (LIM == kLimO3) {
      R y6[6];
      o3_stencil<R>(phi, s, N, y6);
      phif = inner ? o3_flux<R>(y6, h6, vdt) : R(0);
      o3_mixed_faces<R, NEM>(vpure, role, k, vdt, phim, phi, &phif, o3V, o3Vb, o3Vf, o3M, o3Mb, o3Mf, o3up);
    } else {
      R g = R(0);
      if (LIM == kLimG && o2) {  // computeAndLimitGradPhi, UtilesRemap.cc:329-338
        const R phip = lane_down1(phi);
        g = limiter_g<R>(P.lim, gfp, gf, front_in ? phip : phi, inner ? phim : phi, hc, hcp, hcm);
      } else if (o2) {
        if (LEARN) {
          bool used;
          R gcl;
          const R gl = learned_grad(es.trace(role, phi) ? R(-1) : alpha, phim, phi, lane_down1(phi), inner,
                                    front_in, gf, gfp, dbr, dfr, &used, &gcl);
          g = es.slope(role, k, phi, gl, used, gcl);
        } else {
          g = limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp);
        }
      }
      const R gm = lane_up1(g);
      phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    }
    const R phifp = lane_down1(phif);
    const R flux_b = fb * phif;
    const R flux_f = ff * phifp;
    if (D == 0) {  // cell faces in order: bottom, right (front), top, left (back)
      unew = u0 - (flux_f + flux_b);
      dual = R(0.5) * flux_f * R(1.0) + R(0.5) * flux_b * R(-1.0);
    } else {       // bottom (back), right, top (front), left
      unew = u0 - (flux_b + flux_f);
      dual = R(0.5) * flux_b * R(-1.0) + R(0.5) * flux_f * R(1.0);
    }
  };
  R* __restrict__ Uout = A.Uout;
  R fV[NEM], fM[NEM], fE[NEM];  // FINAL: the new remap vector of the cell
#pragma unroll
  for (int k = 0; k < NEM; ++k) fV[k] = fM[k] = fE[k] = R(0);
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[0][k], un, du, 0, k);
    if (FINAL) fV[k] = un;
    else if (out) st_stream<1>(P, &Uout[k * NC + c], un);
  }
  if (LEARN) es.close_volumes(ne, alpha);
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[1][k], un, du, 1, k);
    if (un < R(0)) un = R(0);  // RemapADIService.cc:953-955
    if (FINAL) fM[k] = un;
    if (out) {
      if (!FINAL) st_stream<1>(P, &Uout[(ne + k) * NC + c], un);
      st_stream<2>(P, &A.contrib[k * NC + c], R(0.5) * du);  // fcontrib (DualPhaseRemap.cc:168-185)
    }
  }
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[2][k], un, du, 2, k);
    if (un < R(0)) un = R(0);  // RemapADIService.cc:956-958
    if (FINAL) fE[k] = un;
    else if (out) st_stream<1>(P, &Uout[(2 * ne + k) * NC + c], un);
  }
  R fQ = R(0);
  {
    R un, du;
    sweep_var(preq, un, du, 3, 0);
    if (FINAL) fQ = un;
    else if (out) st_stream<1>(P, &Uout[(3 * ne) * NC + c], un);
  }
  if (FINAL && out) {
    // Euler volume: the corner-normal formula on the initial coordinates
    double x0[4] = {node_x0(P, i), node_x0(P, i + 1), node_x0(P, i + 1), node_x0(P, i)};
    double y0[4] = {node_y0(P, j), node_y0(P, j), node_y0(P, j + 1), node_y0(P, j + 1)};
    const double ve = kCartesianEuler<R> ? P.h[0] * P.h[1] : quad_volume(x0, y0);
    final_cell_vals<R, NEM>(P, *F, c, ve, fV, fM, fE, fQ, R(0), fmask);
  }
}

// y sweep on LDS tiles: a block of 16 wavefronts owns 16 columns (one
// 128-byte line of a row) and 64 rows (60 outputs, two halo rows on each
// side).  Global loads and stores are row-major and coalesced; the tile is
// transposed through LDS so that each wavefront holds one column with one row
// per lane and runs the lane-exchange arithmetic of the x sweep.
//
// SEG (meshes with ny + 4 <= 64): a column with its halos is a segment of
// SL = ny + 4 lanes and a wavefront holds K = 64 / SL whole columns, so the
// block's tile is 16 K columns by SL rows and covers the column height at
// once.  A lane's stencil never leaves its segment: outputs sit in lanes 2 ..
// SL - 3 and read at most two lanes either side.
template <bool SEG>
struct YTileShape {
  int SL, K, TW, TP;  // segment rows, segments per wavefront, tile columns, padded row
  __device__ YTileShape(int nrows) {
    SL = SEG ? nrows + 4 : kTileR;
    K = SEG ? kTileR / SL : 1;
    TW = kTileW * K;
    TP = TW + 1;
  }
};

// FINAL: as in the lane form, the last sweep fused with the final cell
// phase; the thread that would store a cell's remap vector row-major keeps it
// and runs the final phase of that cell.
// LAG (with FIRST): as in the lane form, the Lagrange cell phase runs on
// each row-major cell of the tile (halo rows included) and its remap vector
// feeds the staging tiles from registers.
template <class R, int NEM, bool FIRST, int LIM, bool SEG, bool LEARN = false, bool FINAL = false,
          bool LAG = false, int TWC = kTileW>
__device__ __forceinline__ void cell_sweep_ytile_2d_body(const Params& P, DevScalars* S,
                                                         const CellSweepArgsT<R>& A, int L,
                                                         const CellFinalArgsT<R>* F = nullptr,
                                                         const CellLagArgsT<R>* LA = nullptr) {
  // geometry tiles [0..4]; once read, tiles [0..5] are reused as the staging
  // tiles (in, out, dual flux) of a pair of variables
  // order 3: a seventh tile for the cells' mixed / pure codes
  // TWC columns (one wavefront each) of kTileR rows, rows padded by one word
  __shared__ double lds[LIM == kLimO3 ? 7 : 6][kTileR * (TWC + 1)];
  R* const stg = reinterpret_cast<R*>(&lds[0][0]);
  constexpr int TS = kTileR * (TWC + 1);  // words of one staging tile
  constexpr int HL = (LIM == kLimO3) ? 3 : 2, NO = 64 - 2 * HL;  // halo rows, output rows
  const bool halt = halted(S);
  const int nx = P.nx, N = P.ny, nnx = nx + 1;
  const YTileShape<SEG> sh(N);
  const int TW = SEG ? sh.TW : TWC, TP = SEG ? sh.TP : TWC + 1;
  const int ncolblk = (nx + TW - 1) / TW;
  const int i0 = (L % ncolblk) * TW;
  const int s0 = SEG ? -2 : (L / ncolblk) * NO - HL;  // cell row of tile row 0
  const int t = threadIdx.x;
  // row-major position (loads, stores)
  const int tr = SEG ? t / TW : t / TWC, tc = SEG ? t % TW : t % TWC;
  const bool trow = SEG ? tr < sh.SL : true;
  // transposed position: column col, row r (lane of its segment)
  const int w = t >> 6, lane = t & 63;
  const int col = SEG ? w * sh.K + lane / sh.SL : w, r = SEG ? lane % sh.SL : lane;
  const bool tlane = SEG ? lane < sh.K * sh.SL : true;
  const int ta = r * TP + col;   // transposed LDS index
  const int ra = tr * TP + tc;   // row-major LDS index
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const bool o2 = P.order2 > 0;
  const R* __restrict__ Uin = A.Uin;
  // row-major cell of this thread
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = trow && gi < nx && gs >= 0 && gs < N;
  const bool gout = SEG ? gex : gex && tr >= HL && tr <= 63 - HL;
  const long gc = gex ? gi + (long)nx * gs : 0;
  // Every load of the geometry phase, and the first pair of variables, is
  // issued first from clamped indices (a thread out of the mesh reads cell /
  // node 0), so that a wavefront waits for memory once before the barrier.
  const bool fvalid = trow && gi < nx && gs >= 0 && gs <= N;  // back face: nodes (i, s), (i+1, s)
  const long m0 = gex ? gi + (long)nnx * gs : 0;
  const long na = fvalid ? gi + (long)nnx * gs : 0, nb = na + 1;
  const R uya = A.uy[na], uyb = A.uy[nb], uxa = A.ux[na], uxb = A.ux[nb];
  // the nodes (i+1, s+1), (i, s+1) of the cell off its back face
  const R uyc = A.uy[m0 + 1 + nnx], uyd = A.uy[m0 + nnx];
  R uxc = R(0), uxd = R(0);
  if (LIM == kLimG || LIM == kLimO3 || LAG) {
    uxc = A.ux[m0 + 1 + nnx];
    uxd = A.ux[m0 + nnx];
  }
  R vlag = (FIRST && !LAG) ? A.Vlag[gc] : R(1);
  // FINAL: the cell's membership mask, loaded with the geometry
  const int fmask = FINAL ? F->mask[gc] : 0;
  // Lagrangian coordinates x0 + dt u (lag_coord); the cell's first two
  // nodes are its back face's
  const double yfa = lag_coord<1>(P, gs, dt, uya), yfb = lag_coord<1>(P, gs, dt, uyb);
  const double xfa = lag_coord<0>(P, gi, dt, uxa), xfb = lag_coord<0>(P, gi + 1, dt, uxb);
  const double ya = yfa, yb = yfb, yc = lag_coord<1>(P, gs + 1, dt, uyc), yd = lag_coord<1>(P, gs + 1, dt, uyd);
  // generalised limiters, order 3, LAG: x coordinates of the cell's four nodes
  double tn[4] = {0., 0., 0., 0.};
  if (LIM == kLimG || LIM == kLimO3 || LAG) {
    tn[0] = xfa;
    tn[1] = xfb;
    tn[2] = lag_coord<0>(P, gi + 1, dt, uxc);
    tn[3] = lag_coord<0>(P, gi, dt, uxd);
  }
  const int nvu = P.nvu;
  R uvol[NEM];
  R lU[3 * NEM + 1];  // LAG: the remap vector of the row-major cell
#pragma unroll
  for (int v = 0; v < 3 * NEM + 1; ++v) lU[v] = R(0);
  // variables per staging round (input, output and dual flux tiles of each
  // in the six double tiles); four per round in float measured no faster
  // (0.1214 -> 0.1226 ms per triple-point step)
  constexpr int VPR = 2;
  R nx0[VPR];
  if (LAG) {  // the Lagrange cell phase of the row-major cell
    const CellPre<R> lin = cell_pre_load<R, false>(*LA, gc);
    const double yq[4] = {ya, yb, yc, yd};
    const double Vd = quad_volume(tn, yq);
    const double Ld = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(tn, yq) : length_sqrt_2d_r<R>(Vd);
    R uv[NEM], um[NEM], ue[NEM], uq;
    lagrange_cell_vals<R, false, NEM>(P, S, *LA, gc, (R)Vd, (R)Ld, lin, uv, um, ue, uq);
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      uvol[k] = (k < ne) ? uv[k] : R(0);
      if (k < ne) {
        lU[k] = uv[k];
        lU[ne + k] = um[k];
        lU[2 * ne + k] = ue[k];
      }
    }
    lU[3 * ne] = uq;
    vlag = (R)Vd;
#pragma unroll
    for (int q = 0; q < VPR; ++q) nx0[q] = lU[q];
    if (halt) return;
    if (gout) {
      st_stream<0>(P, &LA->Vlag[gc], (R)Vd);
      st_stream<0>(P, &LA->L[gc], (R)Ld);
    }
  } else {
#pragma unroll
    for (int k = 0; k < NEM; ++k) uvol[k] = (k < ne) ? Uin[k * NC + gc] : R(0);
#pragma unroll
    for (int q = 0; q < VPR; ++q) nx0[q] = q < nvu ? Uin[q * NC + gc] : R(0);
    if (halt) return;
  }
  {  // geometry of the tile, row-major (stored in the double tiles exactly)
    double xc = 0., xf = 0., lf = 0.;
    R den = R(1), uf = R(0);
    if (gex) {
      xc = 0.25 * (((ya + yb) + yc) + yd);
      if (FIRST) {
        den = vlag;
      } else {
        R sv = R(0);
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if (k < ne) sv += uvol[k];
        den = sv;
      }
    }
    if (fvalid) {
      uf = R(0.5) * (uya + uyb);
      xf = 0.5 * yfa + 0.5 * yfb;
      lf = fabs(xfb - xfa);
    }
    if (trow) {
      lds[0][ra] = xc;
      lds[1][ra] = den;
      lds[2][ra] = uf;
      lds[3][ra] = xf;
      lds[4][ra] = lf;
    }
    if ((LIM == kLimG || LIM == kLimO3) && trow) {  // width along y (faces at node rows gs and gs + 1)
      const double an[4] = {ya, yb, yc, yd};
      lds[5][ra] = gex ? cell_width_2d(an, tn, 0, 1, 3, 2, gs == 0, gs == N - 1) : 0.;
    }
    if (LIM == kLimO3 && trow) {  // the cell's mixed / pure code (see the x sweep)
      int e = 0;
      if (FIRST) {
        const int el = A.est[gc];
        e = gex ? el : 0;
      } else {
        int cnt = 0, last = -1;
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if (k < ne && qdiv(uvol[k], den) > R(0)) {
            cnt++;
            last = k;
          }
        e = cnt > 1 ? -2 : last;
      }
      lds[LIM == kLimO3 ? 6 : 5][ra] = (double)e;
    }
  }
  __syncthreads();
  const int s = s0 + r;
  const bool ex = tlane && (i0 + col < nx) && s >= 0 && s < N;
  const bool inner = ex && s >= 1;
  const double xc = tlane ? lds[0][ta] : 0., xf = tlane ? lds[3][ta] : 0., lf = tlane ? lds[4][ta] : 0.;
  const R den = tlane ? (R)lds[1][ta] : R(1), uf = tlane ? (R)lds[2][ta] : R(0);
  const double xcm = lane_up1(xc);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = lane_down1(uf);
  const double lf_f = lane_down1(lf);
  const R fb = flux_factor<R>(-uf, lf, dt), ff = flux_factor<R>(uf_f, lf_f, dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  // learned limiter: coefficient, front-neighbour existence, face offsets
  R alpha = R(0), dbr = R(0), dfr = R(0);
  bool front_in = false;
  if (LEARN) {
    alpha = ex ? A.alpha[(i0 + col) + (long)nx * s] : R(0);
    front_in = lane_down1((int)inner) != 0;
    dbr = (R)(xf - xc);
    dfr = (R)(lane_down1(xf) - xc);
  }
  // generalised limiters: widths of the cell and of its back / front
  // neighbours along y (the cell itself where the neighbour is missing)
  R hc = R(0), hcm = R(0), hcp = R(0);
  R h6[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};  // order 3: widths bbb .. fff of the back face
  if (LIM == kLimG || LIM == kLimO3) {
    front_in = lane_down1((int)inner) != 0;
    const double h = tlane ? lds[5][ta] : 0.;
    const double hm = lane_up1(h), hp = lane_down1(h);
    hc = (R)h;
    hcm = (R)(inner ? hm : h);
    hcp = (R)(front_in ? hp : h);
    if (LIM == kLimO3) o3_stencil<R>((R)h, s, N, h6);
  }
  int est = 0, estb = 0;
  R vdt = R(0);
  if (LIM == kLimO3) {
    est = tlane ? (int)lds[LIM == kLimO3 ? 6 : 5][ta] : 0;
    estb = lane_up1(est);
    vdt = vel_dt<R>(uf, dt);
  }
  const bool vpure = estb != -2 && est != -2 && estb == est;
  R o3V[NEM], o3Vb[NEM], o3Vf[NEM], o3M[NEM], o3Mb[NEM], o3Mf[NEM];
  int o3up[NEM];
  EnvSlopes<R, NEM> es(P.remap_threshold);  // learned mode: consistent environment slopes
  // one variable v on the transposed layout: new U and half the dual flux
  auto sweep_var = [&](R u0, R& un, R& dual, int v) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = lane_up1(phi);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = lane_down1(gf);
    R phif;
    if (LIM == kLimO3) {
      const int role = v < ne ? 0 : v < 2 * ne ? 1 : v < 3 * ne ? 2 : 3;
      const int k = role == 0 ? v : role == 1 ? v - ne : role == 2 ? v - 2 * ne : 0;
      R y6[6];
      o3_stencil<R>(phi, s, N, y6);
      phif = inner ? o3_flux<R>(y6, h6, vdt) : R(0);
      o3_mixed_faces<R, NEM>(vpure, role, k, vdt, phim, phi, &phif, o3V, o3Vb, o3Vf, o3M, o3Mb, o3Mf, o3up);
    } else {
      R g = R(0);
      if (LIM == kLimG && o2) {  // computeAndLimitGradPhi, UtilesRemap.cc:329-338
        const R phip = lane_down1(phi);
        g = limiter_g<R>(P.lim, gfp, gf, front_in ? phip : phi, inner ? phim : phi, hc, hcp, hcm);
      } else if (o2) {
        if (LEARN) {
          const int role = v < ne ? 0 : v < 2 * ne ? 1 : v < 3 * ne ? 2 : 3;
          const int k = role == 0 ? v : role == 1 ? v - ne : role == 2 ? v - 2 * ne : 0;
          if (role == 1 && v == ne) es.close_volumes(ne, alpha);  // every volume is done
          bool used;
          R gcl;
          const R gl = learned_grad(es.trace(role, phi) ? R(-1) : alpha, phim, phi, lane_down1(phi), inner,
                                    front_in, gf, gfp, dbr, dfr, &used, &gcl);
          g = es.slope(role, k, phi, gl, used, gcl);
        } else {
          g = limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp);
        }
      }
      const R gm = lane_up1(g);
      phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    }
    const R phifp = lane_down1(phif);
    const R flux_b = fb * phif;
    const R flux_f = ff * phifp;
    // cell faces in order: bottom (back), right, top (front), left
    un = u0 - (flux_b + flux_f);
    dual = R(0.5) * flux_b * R(-1.0) + R(0.5) * flux_f * R(1.0);
  };
  __syncthreads();  // geometry reads done: the tiles become staging tiles
  // variables VPR at a time, one barrier pair per round
  R nxq[VPR];
#pragma unroll
  for (int q = 0; q < VPR; ++q) nxq[q] = (gex && q < nvu) ? nx0[q] : R(0);
  R fU[3 * NEM + 1];  // FINAL: the new remap vector of the row-major cell
#pragma unroll
  for (int v = 0; v < 3 * NEM + 1; ++v) fU[v] = R(0);
  auto round = [&](const int va) {
    if (trow) {
#pragma unroll
      for (int q = 0; q < VPR; ++q) stg[q * TS + ra] = nxq[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int vn = va + VPR + q;  // the next round's variables, loaded ahead
      if (vn < nvu) {
        if (LAG) nxq[q] = gex ? lU[vn < 3 * NEM + 1 ? vn : 0] : R(0);
        else nxq[q] = gex ? Uin[vn * NC + gc] : R(0);
      }
    }
    R un[VPR], du[VPR];
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int v = va + q;
      un[q] = du[q] = R(0);
      if (v < nvu) {  // uniform
        const R u = tlane ? stg[q * TS + ta] : R(0);
        sweep_var(u, un[q], du[q], v);
        if (v >= ne && v < 3 * ne && un[q] < R(0)) un[q] = R(0);  // RemapADIService.cc:953-958
      }
    }
    if (tlane) {
#pragma unroll
      for (int q = 0; q < VPR; ++q) {
        const int v = va + q;
        if (v < nvu) stg[(VPR + q) * TS + ta] = un[q];
        if (v >= ne && v < 2 * ne) stg[(2 * VPR + q) * TS + ta] = R(0.5) * du[q];  // fcontrib (DualPhaseRemap.cc:168-185)
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int v = va + q;
      if (v >= nvu) continue;  // uniform
      if (FINAL) fU[v < 3 * NEM + 1 ? v : 0] = trow ? stg[(VPR + q) * TS + ra] : R(0);
      if (gout) {
        if (!FINAL) st_stream<1>(P, &A.Uout[v * NC + gc], stg[(VPR + q) * TS + ra]);
        if (v >= ne && v < 2 * ne) st_stream<2>(P, &A.contrib[(v - ne) * NC + gc], stg[(2 * VPR + q) * TS + ra]);
      }
    }
  };
  if (LEARN || LIM == kLimO3 || FINAL || LAG) {  // compile-time variable indices: per-environment values stay in registers
#pragma unroll
    for (int va = 0; va < 3 * NEM + 1; va += VPR)
      if (va < nvu) round(va);
  } else {
    for (int va = 0; va < nvu; va += VPR) round(va);
  }
  if (FINAL && gout) {
    R fV[NEM], fM[NEM], fE[NEM];
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      fV[k] = (k < ne) ? fU[k] : R(0);
      fM[k] = (k < ne) ? fU[ne + k] : R(0);
      fE[k] = (k < ne) ? fU[2 * ne + k] : R(0);
    }
    // the Euler volume from the initial node positions, formed here: the
    // opaque copies of the indices keep the compiler from carrying the
    // geometry phase's equal x0 / y0 values through the rounds (with 64
    // VGPRs that spilled 36 bytes per lane)
    int fi = gi, fs = gs;
    asm volatile("" : "+v"(fi), "+v"(fs));
    double x0[4] = {node_x0(P, fi), node_x0(P, fi + 1), node_x0(P, fi + 1), node_x0(P, fi)};
    double y0[4] = {node_y0(P, fs), node_y0(P, fs), node_y0(P, fs + 1), node_y0(P, fs + 1)};
    const double ve = kCartesianEuler<R> ? P.h[0] * P.h[1] : quad_volume(x0, y0);
    final_cell_vals<R, NEM>(P, *F, gc, ve, fV, fM, fE, fU[3 * ne], R(0), fmask);
  }
}

// ---------------------------------------------------------------------------
// K5: final cell phase.  remapVariables cell status + membership change
// (RemapADIFinal.cc:98-226), "moy" (446-643), applyEOS per environment
// (PerfectGasEOSService.cc:50-137), computePressionMoyenne (MahycoModule.cc:
// 1836-1915) and the CFL candidate of computeHydroDeltaT (2047-2061) reduced
// to one minimum per block; k_dt_finalize finishes computeDeltaT (2076-2191).
// ---------------------------------------------------------------------------
template <class R, int BLOCK, int NEM, bool EOS = true>
__device__ __forceinline__ void cell_final_2d_body(const Params& P, DevScalars* S, const CellFinalArgsT<R>& A,
                                                   long blk, double* block_min) {
  if (S->stop || S->err) return;
  if (!EOS) {  // remapVariables only: no EOS, no CFL candidate
    const long c = blk * BLOCK + threadIdx.x;
    if (c >= P.nc) return;
    int i, j;
    divmod32(c, P.nx, j, i);
    R delta_ec = R(0);
    if (P.etot) {
      const int nnx = P.nx + 1;
      const long n0 = i + (long)nnx * j;
      const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
      R sum = R(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) sum += A.dek[nd[q]];
      delta_ec = smax(R(0), sum);
    }
    double x0[4] = {node_x0(P, i), node_x0(P, i + 1), node_x0(P, i + 1), node_x0(P, i)};
    double y0[4] = {node_y0(P, j), node_y0(P, j), node_y0(P, j + 1), node_y0(P, j + 1)};
    final_cell<R, NEM, false>(P, A, c, kCartesianEuler<R> ? P.h[0] * P.h[1] : quad_volume(x0, y0), delta_ec);
    return;
  }
  const long c = blk * BLOCK + threadIdx.x;
  double dx_sound = 1.7976931348623157e308;
  if (c < P.nc) {
    const int nnx = P.nx + 1;
    int i, j;
    divmod32(c, P.nx, j, i);
    // node velocities of the CFL candidate, loaded before the cell phase
    // writes its fields so that they travel with its loads
    const long n0 = i + (long)nnx * j;
    const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
    R vx[4], vy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vx[q] = A.ux[nd[q]];
      vy[q] = A.uy[nd[q]];
    }
    const R Lc = A.L[c];
    // total-energy fix-up: the deposit, the nodes' terms in cell node order
    R delta_ec = R(0);
    if (P.etot) {
      R dk[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) dk[q] = A.dek[nd[q]];
      R sum = R(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) sum += dk[q];
      delta_ec = smax(R(0), sum);
    }
    // Euler volume: the corner-normal formula on the initial coordinates.
    double x0[4] = {node_x0(P, i), node_x0(P, i + 1), node_x0(P, i + 1), node_x0(P, i)};
    double y0[4] = {node_y0(P, j), node_y0(P, j), node_y0(P, j + 1), node_y0(P, j + 1)};
    const R c_g = final_cell<R, NEM>(P, A, c, kCartesianEuler<R> ? P.h[0] * P.h[1] : quad_volume(x0, y0), delta_ec);
    // CFL candidate: L / (c + max over the cell nodes of |u|)
    // (the root of the largest squared norm: sqrt is correctly rounded, hence
    // monotone, so it equals the largest root)
    R n2max = R(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const R ux = vx[q], uy = vy[q];
      const R n2 = ux * ux + uy * uy + R(0) * R(0);
      n2max = (n2 < n2max) ? n2max : n2;
    }
    const R vmax = sqrt(n2max);
    // std::min(minimum_aux, dx_sound) fold: a NaN candidate is skipped
    dx_sound = smin(dx_sound, (double)qdiv(Lc, c_g + vmax));
  }
  block_min_store<BLOCK>(dx_sound, block_min + blk);
}

// Minimum of nb block minima over a block of BLOCK threads, every thread
// receiving it (the minimum does not depend on the order).
template <int BLOCK>
__device__ __forceinline__ double block_minima_min(const double* block_min, int nb) {
  __shared__ double red[BLOCK / 64];
  // eight loads in flight per thread
  double w[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) w[u] = 1.7976931348623157e308;
  for (int b = threadIdx.x; b < nb; b += 8 * BLOCK) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = b + u * BLOCK;
      const double x = block_min[idx < nb ? idx : 0];
      if (idx < nb) w[u] = smin(w[u], x);
    }
  }
  double v = w[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) v = smin(v, w[u]);
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int q = 1; q < BLOCK / 64; ++q) m = smin(m, red[q]);
  return m;
}

// The dt kernel's body: one block reduces the block minima of the CFL
// kernel and updates the scalars in place.
template <int BLOCK, bool ADVANCE>
__device__ __forceinline__ void dt_finalize_body(const Params& P, DevScalars* S, const double* block_min, int nb) {
  const DevScalars s = *S;
  const double m = block_minima_min<BLOCK>(block_min, nb);
  if (s.stop || s.err) return;
  if (threadIdx.x == 0) *S = dt_update<ADVANCE>(P, s, m);
}

// Node spacings along a dual sweep, front and back.  FP64 (and the
// generalised limiters, which also need the nodes two places away) forms
// them from the Lagrangian coordinates x0 + dt u as the reference does; the
// float form as h + dt (u' - u), in float, from the velocities along the
// sweep (a difference of velocities: nothing cancels).
template <class R, int LIM>
constexpr bool kFloatSpacing = std::is_same<R, float>::value && LIM != kLimG;
template <class R, int LIM>
__device__ __forceinline__ void dual_spacing(double h, double dt, double xd, R ud, double* sf, double* sb,
                                             double* xdm, double* xdp) {
  if constexpr (kFloatSpacing<R, LIM>) {
    const R um = lane_up1(ud), up = lane_down1(ud);
    const R hr = (R)h, dtr = (R)dt;
    *sf = (double)(hr + dtr * (up - ud));
    *sb = (double)(hr + dtr * (ud - um));
    *xdm = *xdp = 0.;
  } else {
    *xdm = lane_up1(xd);
    *xdp = lane_down1(xd);
    *sf = *xdp - xd;
    *sb = xd - *xdm;
  }
}

// ---------------------------------------------------------------------------
// K4: dual (node) sweep in direction D.  computeDualGradPhi_LimC
// (DualPhaseRemap.cc:700-779), node mass fluxes (ncontrib 215-260), upwind
// momentum update (352-491) and the phi_dual refresh (494-523).  FIRST builds
// the dual state from (u, m_node) as computeVariablesForRemap_PBorn0 does
// (PrepareRemap.cc:411-433); the last sweep ends with the nodal velocity
// u = U_dual / m_dual of remapVariables (RemapADIFinal.cc:660-672).
// ---------------------------------------------------------------------------
// The same sum with the four loads issued unconditionally (clamped cell
// indices; a missing cell adds +0 to a sum that starts at +0).
template <class R>
__device__ inline R node_mass_from_cells_nb(const Params& P, const R* m, int i, int j) {
  const int cdi[4] = {-1, 0, -1, 0}, cdj[4] = {-1, -1, 0, 0};
  R v[4];
  bool ok[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ci = i + cdi[t], cj = j + cdj[t];
    ok[t] = ci >= 0 && ci < P.nx && cj >= 0 && cj < P.ny;
    v[t] = m[ok[t] ? ci + (long)P.nx * cj : 0];
  }
  R msum = R(0);
#pragma unroll
  for (int t = 0; t < 4; ++t) msum += ok[t] ? v[t] : R(0);
  return R(0.25) * msum;
}

template <class R>
__device__ inline R node_mass_from_cells(const Params& P, const R* m, int i, int j) {
  R msum = R(0);
  const int cdi[4] = {-1, 0, -1, 0}, cdj[4] = {-1, -1, 0, 0};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int ci = i + cdi[t], cj = j + cdj[t];
    if (ci < 0 || ci >= P.nx || cj < 0 || cj >= P.ny) continue;
    msum += m[ci + (long)P.nx * cj];
  }
  return R(0.25) * msum;
}

// Total-energy fix-up (RemapADIFinal.cc:751-776).  Start of the dual
// kinetic energy (PrepareRemap.cc:232, 243): U_4 = 1/2 m |u| and phi_4 =
// 1/2 |u| (the norm, not its square, as the reference has it).  U4 may be
// null (phi only).
template <class R>
__device__ __forceinline__ void dual_ecin_start(R m, R ux, R uy, R* U4, R* p4) {
  const R nrm = sqrt(ux * ux + uy * uy + R(0) * R(0));
  *p4 = R(0.5) * nrm;
  if (U4) *U4 = R(0.5) * m * nrm;
}
// A node's term of the cells' deposit after the last dual sweep: a quarter
// of U_4 / U_3 minus 1/2 |u|^2 of the remapped velocity U / U_3; +0 for a node
// without mass (skipped by the reference; a sum that starts at +0 is the
// same either way).
template <class R>
__device__ __forceinline__ R dual_ecin_term(R U0, R U1, R U3, R U4) {
  if (U3 == R(0)) return R(0);
  const R vx = qdiv(U0, U3), vy = qdiv(U1, U3);
  return R(0.25) * (qdiv(U4, U3) - R(0.5) * (vx * vx + vy * vy + R(0) * R(0)));
}

// Dual sweep along x with lane exchange: a wavefront holds 64 consecutive
// nodes of a row (60 outputs, two halo nodes on each side).  Every node forms
// its phi_dual and its limited gradient once and reads its neighbours' values
// across lanes.  front(i) sums the cells (i, j-1), (i, j) in the same order as
// back(i+1) does, so back is the front of the previous lane.
template <class R, bool FIRST, int NEM, int LIM>
__device__ __forceinline__ void node_sweep_x_2d_body(const Params& P, DevScalars* S, const NodeSweepArgsT<R>& A,
                                                     int blk) {
  const bool halt = halted(S);
  const int lane = threadIdx.x & 63;
  const int N = P.nx;
  const int nseg = (N + 1 + 59) / 60;
  const long wave = ((long)blk * blockDim.x + threadIdx.x) >> 6;
  int j, seg;
  divmod32(wave, nseg, j, seg);
  if (j > P.ny) return;  // uniform across the wave
  const int i = seg * 60 - 2 + lane;
  const bool ex = (i >= 0 && i <= N);
  const bool out = ex && lane >= 2 && lane <= 61;
  const long nid = ex ? i + (long)(N + 1) * j : 0;
  const bool inner_t = (j >= 1 && j <= P.ny - 1);
  const bool inner_d = (i >= 1 && i <= N - 1);
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  // Every load of the node is issued first, from clamped indices and
  // without branches, so that a wavefront waits for memory once: a lane out
  // of the mesh reads node 0, and a missing cell adds +0 (front starts at +0
  // and a sum of doubles never becomes -0, so the result is the guarded
  // sum's bit for bit).
  // the node's Lagrangian x, x0 + dt u (lag_coord)
  const R ud_l = A.ux[nid];
  const double xd_l = lag_coord<0>(P, i, dt, ud_l);
  // generalised limiters: x of the nodes two places back and front (beyond
  // the lanes a wavefront exchanges)
  double xmm = 0., xpp = 0.;
  if (LIM == kLimG) {
    xmm = lag_coord<0>(P, i - 2, dt, A.ux[(ex && i >= 2) ? nid - 2 : 0]);
    xpp = lag_coord<0>(P, i + 2, dt, A.ux[(ex && i + 2 <= N) ? nid + 2 : 0]);
  }
  R cf[2][NEM];
  bool cv[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int cj = j - 1 + q;
    cv[q] = ex && i < P.nx && cj >= 0 && cj < P.ny;
    const long cc = cv[q] ? i + (long)P.nx * cj : 0;
#pragma unroll
    for (int k = 0; k < NEM; ++k) cf[q][k] = (k < ne) ? A.contrib[k * NC + cc] : R(0);
  }
  R U0 = R(0), U1 = R(0), U3 = R(0), p0 = R(0), p1 = R(0);
  R U4 = R(0), p4 = R(0);  // dual kinetic energy (total-energy fix-up only)
  double xd = 0.;
  if (FIRST) {
    const R mn = node_mass_from_cells_nb(P, A.m, i, j);
    const R ux = A.ux[nid], uy = A.uy[nid];
    if (ex) {
      p0 = ux;
      p1 = uy;
      U0 = mn * p0;
      U1 = mn * p1;
      U3 = mn;
      if (P.etot) dual_ecin_start(mn, p0, p1, &U4, &p4);
    }
  } else {
    const R u0 = A.Ud0[nid], u1 = A.Ud1[nid], u3 = A.Ud3[nid];
    const R u4 = P.etot ? A.Ud4[nid] : R(0);
    if (ex) {
      U0 = u0;
      U1 = u1;
      U3 = u3;
      U4 = u4;
      // nodes inner in the first direction carry phi = U/m (DualPhaseRemap.cc:494-523)
      if (inner_t) {
        p0 = qdiv(U0, U3);
        p1 = qdiv(U1, U3);
        p4 = qdiv(U4, U3);
      } else {
        p0 = A.ux[nid];
        p1 = A.uy[nid];
        if (P.etot) dual_ecin_start<R>(R(0), p0, p1, nullptr, &p4);
      }
    }
  }
  if (halt) return;
  if (ex) xd = xd_l;
  // front dual mass flux: cells (i, j-1), (i, j), environments in order
  R front = R(0);
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int k = 0; k < NEM; ++k)
      if (k < ne) front += cv[q] ? cf[q][k] : R(0);
  double xdm, xdp, spf, spb;  // node spacings front and back (dual_spacing)
  dual_spacing<R, LIM>(P.h[0], dt, xd, ex ? ud_l : R(0), &spf, &spb, &xdm, &xdp);
  const R p0m = lane_up1(p0), p0p = lane_down1(p0);
  const R p1m = lane_up1(p1), p1p = lane_down1(p1);
  const R back = lane_up1(front);
  // limited nodal gradient (computeDualGradPhi_LimC, DualPhaseRemap.cc:700-779)
  R g0 = R(0), g1 = R(0);
  if (P.order2 > 0 && ex) {
    if (inner_d) {
      const R dxf = (R)spf, dxb = (R)spb;
      const R gf0 = qdiv(p0p - p0, dxf), gb0 = qdiv(p0 - p0m, dxb);
      const R gf1 = qdiv(p1p - p1, dxf), gb1 = qdiv(p1 - p1m, dxb);
      if (LIM == kLimG) {  // computeDualGradPhi, UtilesRemap.cc:15-68, 350-382
        const R h0 = (R)(0.5 * (xdp - xdm));
        const R hm = (R)(i >= 2 ? 0.5 * (xd - xmm) : 0.), hp = (R)(i <= N - 2 ? 0.5 * (xpp - xd) : 0.);
        g0 = limiter_g<R>(P.lim, gf0, gb0, p0p, p0m, h0, hp, hm);
        g1 = limiter_g<R>(P.lim, gf1, gb1, p1p, p1m, h0, hp, hm);
      } else {
        g0 = limited_grad<LIM, false, R, kBranchFree<R>>(gb0, gf0);
        g1 = limited_grad<LIM, false, R, kBranchFree<R>>(gb1, gf1);
      }
      if (!inner_t && out) {  // kept for the sweep in the other direction
        A.dg0[nid] = g0;
        A.dg1[nid] = g1;
      }
    } else {  // boundary node along x: gradient left by an earlier sweep
      g0 = A.dg0[nid];
      g1 = A.dg1[nid];
    }
  }
  const R g0m = lane_up1(g0), g0p = lane_down1(g0);
  const R g1m = lane_up1(g1), g1p = lane_down1(g1);
  R p4m = R(0), p4p = R(0);
  if (P.etot) {
    p4m = lane_up1(p4);
    p4p = lane_down1(p4);
  }
  if (!out) return;
  if (inner_d) {
    U3 += back - front;
    // the kinetic energy is carried with the upwind value (its gradient is 0)
    if (P.etot) U4 += back * (back > R(0) ? p4m : p4) - front * (front < R(0) ? p4p : p4);
    const R order2 = (R)P.order2;
    const R ufront = R(0.5) * (p0 + p0);  // DualPhaseRemap.cc:391
    const int signfront = (ufront > R(0)) ? 1 : -1;
    const double xfr = spf;
    const bool sf = (front < R(0));
    const R tf = upwind_offset<R>(signfront, xfr, dt, ufront);
    const R FV0 = (sf ? p0p : p0) + order2 * (R(0.5) * (sf ? g0p : g0) * tf);
    const R FV1 = (sf ? p1p : p1) + order2 * (R(0.5) * (sf ? g1p : g1) * tf);
    const R uback = R(0.5) * (p0m + p0);
    const int signback = (uback > R(0)) ? 1 : -1;
    const double xbk = spb;
    const bool sb = (back > R(0));
    const R tb = upwind_offset<R>(signback, xbk, dt, uback);
    const R BV0 = (sb ? p0m : p0) + order2 * (R(0.5) * (sb ? g0m : g0) * tb);
    const R BV1 = (sb ? p1m : p1) + order2 * (R(0.5) * (sb ? g1m : g1) * tb);
    U0 += back * BV0 - front * FV0;
    U1 += back * BV1 - front * FV1;
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    if (fabs(U0) < thresold) U0 = R(0);
    if (fabs(U1) < thresold) U1 = R(0);
  }
  if (FIRST) {
    st_stream<2>(P, &A.Ud0[nid], U0);
    st_stream<2>(P, &A.Ud1[nid], U1);
    st_stream<2>(P, &A.Ud3[nid], U3);
    if (P.etot) st_stream<2>(P, &A.Ud4[nid], U4);
  } else if (U3 != R(0)) {
    st_stream<3>(P, &A.uox[nid], qdiv(U0, U3));
    st_stream<3>(P, &A.uoy[nid], qdiv(U1, U3));
  } else {
    st_stream<3>(P, &A.uox[nid], A.ux[nid]);
    st_stream<3>(P, &A.uoy[nid], A.uy[nid]);
  }
  if (!FIRST && P.etot) st_stream<2>(P, &A.dek[nid], dual_ecin_term(U0, U1, U3, U4));
}

// Dual sweep along y on LDS tiles: a block of 16 wavefronts owns 16 node
// columns and 64 node rows (60 outputs, two halo rows on each side).  The
// row-major pass loads the nodes (coalesced) and forms phi_dual and the front
// flux; the tile is transposed through LDS, each wavefront runs the lane
// arithmetic of the x dual sweep on one column, and the increments of the
// dual state travel back through LDS to be added and stored row-major.
// SEG: whole node columns of SL = ny + 5 lanes (ny + 1 nodes and two halos
// on each side), K = 64 / SL of them per wavefront, as in the cell sweep.
template <class R, bool FIRST, int NEM, int LIM, bool SEG, int TWC = kTileW>
__device__ __forceinline__ void node_sweep_ytile_2d_body(const Params& P, DevScalars* S,
                                                         const NodeSweepArgsT<R>& A, int L) {
  __shared__ double ldx[kTileR * (TWC + 1)];   // coordinate along y
  __shared__ R lds[4][kTileR * (TWC + 1)];     // phi_dual x, y, front flux, kinetic energy; then the increments
  const bool halt = halted(S);
  const int nx = P.nx, N = P.ny, nnx = nx + 1;
  const YTileShape<SEG> sh(N + 1);
  const int TW = SEG ? sh.TW : TWC, TP = SEG ? sh.TP : TWC + 1;
  const int ncolblk = (nnx + TW - 1) / TW;
  const int i0 = (L % ncolblk) * TW;
  const int s0 = SEG ? -2 : (L / ncolblk) * 60 - 2;  // node row of tile row 0
  const int t = threadIdx.x;
  const int tr = SEG ? t / TW : t / TWC, tc = SEG ? t % TW : t % TWC;
  const bool trow = SEG ? tr < sh.SL : true;
  const int w = t >> 6, lane = t & 63;
  const int col = SEG ? w * sh.K + lane / sh.SL : w, r = SEG ? lane % sh.SL : lane;
  const bool tlane = SEG ? lane < sh.K * sh.SL : true;
  const int ta = r * TP + col;   // transposed LDS index
  const int ra = tr * TP + tc;   // row-major LDS index
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  // ---- row-major pass
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = trow && gi <= nx && gs >= 0 && gs <= N;
  const bool gout = SEG ? gex : gex && tr >= 2 && tr <= 61;
  const long gq = gex ? gi + (long)nnx * gs : 0;
  R U0 = R(0), U1 = R(0), U3 = R(0), U4 = R(0);
  {
    // every load first, from clamped indices and without branches (one wait
    // for memory; a missing cell adds +0 to a sum that starts at +0)
    const R ud_l = A.uy[gq];
    const double xd_l = lag_coord<1>(P, gs, dt, ud_l);  // the node's Lagrangian y, x0 + dt u
    R cf[2][NEM];
    bool cv[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ci = gi - 1 + q;
      cv[q] = gex && gs < N && ci >= 0 && ci < nx;
      const long cc = cv[q] ? ci + (long)nx * gs : 0;
#pragma unroll
      for (int k = 0; k < NEM; ++k) cf[q][k] = (k < ne) ? A.contrib[k * NC + cc] : R(0);
    }
    double xd = 0.;
    R p0 = R(0), p1 = R(0), f = R(0), p4 = R(0);
    if (FIRST) {
      const R mn = node_mass_from_cells_nb(P, A.m, gi, gs);
      const R ux = A.ux[gq], uy = A.uy[gq];
      if (gex) {
        p0 = ux;
        p1 = uy;
        U0 = mn * p0;
        U1 = mn * p1;
        U3 = mn;
        if (P.etot) dual_ecin_start(mn, p0, p1, &U4, &p4);
      }
    } else {
      const R u0 = A.Ud0[gq], u1 = A.Ud1[gq], u3 = A.Ud3[gq];
      const R u4 = P.etot ? A.Ud4[gq] : R(0);
      if (gex) {
        const bool inner_t = (gi >= 1 && gi <= nx - 1);
        U0 = u0;
        U1 = u1;
        U3 = u3;
        U4 = u4;
        if (inner_t) {  // DualPhaseRemap.cc:494-523
          p0 = qdiv(U0, U3);
          p1 = qdiv(U1, U3);
          p4 = qdiv(U4, U3);
        } else {
          p0 = A.ux[gq];
          p1 = A.uy[gq];
          if (P.etot) dual_ecin_start<R>(R(0), p0, p1, nullptr, &p4);
        }
      }
    }
    if (halt) return;
    if (gex) xd = xd_l;
    // front flux: cells (i-1, s), (i, s), environments in order
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k < ne) f += cv[q] ? cf[q][k] : R(0);
    if (trow) {
      ldx[ra] = kFloatSpacing<R, LIM> ? (double)(gex ? ud_l : R(0)) : xd;  // float spacing: the velocity
      lds[0][ra] = p0;
      lds[1][ra] = p1;
      lds[2][ra] = f;
      lds[3][ra] = p4;
    }
  }
  __syncthreads();
  // ---- transposed pass: column i, node row s
  const int i = i0 + col, s = s0 + r;
  const bool ex = tlane && i <= nx && s >= 0 && s <= N;
  const bool out = SEG ? ex : ex && lane >= 2 && lane <= 61;
  const bool inner_t = (i >= 1 && i <= nx - 1);
  const bool inner_d = (s >= 1 && s <= N - 1);
  const double xd = tlane ? ldx[ta] : 0.;
  const R p0 = tlane ? lds[0][ta] : R(0), p1 = tlane ? lds[1][ta] : R(0);
  const R front = tlane ? lds[2][ta] : R(0);
  double xdm, xdp, spf, spb;  // node spacings front and back (dual_spacing)
  dual_spacing<R, LIM>(P.h[1], dt, xd, (R)xd, &spf, &spb, &xdm, &xdp);
  const R p0m = lane_up1(p0), p0p = lane_down1(p0);
  const R p1m = lane_up1(p1), p1p = lane_down1(p1);
  const R back = lane_up1(front);
  // generalised limiters: y of the nodes two rows back and front (beyond
  // the tile's halo rows)
  double xmm = 0., xpp = 0.;
  if (LIM == kLimG) {
    xmm = lag_coord<1>(P, s - 2, dt, A.uy[(ex && s >= 2) ? i + (long)nnx * (s - 2) : 0]);
    xpp = lag_coord<1>(P, s + 2, dt, A.uy[(ex && s + 2 <= N) ? i + (long)nnx * (s + 2) : 0]);
  }
  R g0 = R(0), g1 = R(0);
  if (P.order2 > 0 && ex) {
    const long q = i + (long)nnx * s;
    if (inner_d) {
      const R dxf = (R)spf, dxb = (R)spb;
      const R gf0 = qdiv(p0p - p0, dxf), gb0 = qdiv(p0 - p0m, dxb);
      const R gf1 = qdiv(p1p - p1, dxf), gb1 = qdiv(p1 - p1m, dxb);
      if (LIM == kLimG) {  // computeDualGradPhi, UtilesRemap.cc:15-68, 350-382
        const R h0 = (R)(0.5 * (xdp - xdm));
        const R hm = (R)(s >= 2 ? 0.5 * (xd - xmm) : 0.), hp = (R)(s <= N - 2 ? 0.5 * (xpp - xd) : 0.);
        g0 = limiter_g<R>(P.lim, gf0, gb0, p0p, p0m, h0, hp, hm);
        g1 = limiter_g<R>(P.lim, gf1, gb1, p1p, p1m, h0, hp, hm);
      } else {
        g0 = limited_grad<LIM, false, R, kBranchFree<R>>(gb0, gf0);
        g1 = limited_grad<LIM, false, R, kBranchFree<R>>(gb1, gf1);
      }
      if (!inner_t && out) {  // kept for the sweep in the other direction
        A.dg0[q] = g0;
        A.dg1[q] = g1;
      }
    } else {  // boundary node along y: gradient left by an earlier sweep
      g0 = A.dg0[q];
      g1 = A.dg1[q];
    }
  }
  const R g0m = lane_up1(g0), g0p = lane_down1(g0);
  const R g1m = lane_up1(g1), g1p = lane_down1(g1);
  R p4 = R(0), p4m = R(0), p4p = R(0);
  if (P.etot) {
    p4 = tlane ? lds[3][ta] : R(0);
    p4m = lane_up1(p4);
    p4p = lane_down1(p4);
  }
  R d0 = R(0), d1 = R(0), d3 = R(0), d4 = R(0);
  if (out && inner_d) {
    const R order2 = (R)P.order2;
    d3 = back - front;
    // the kinetic energy is carried with the upwind value (its gradient is 0)
    if (P.etot) d4 = back * (back > R(0) ? p4m : p4) - front * (front < R(0) ? p4p : p4);
    const R ufront = R(0.5) * (p1 + p1);  // DualPhaseRemap.cc:391
    const int signfront = (ufront > R(0)) ? 1 : -1;
    const double xfr = spf;
    const bool sf = (front < R(0));
    const R tf = upwind_offset<R>(signfront, xfr, dt, ufront);
    const R FV0 = (sf ? p0p : p0) + order2 * (R(0.5) * (sf ? g0p : g0) * tf);
    const R FV1 = (sf ? p1p : p1) + order2 * (R(0.5) * (sf ? g1p : g1) * tf);
    const R uback = R(0.5) * (p1m + p1);
    const int signback = (uback > R(0)) ? 1 : -1;
    const double xbk = spb;
    const bool sb = (back > R(0));
    const R tb = upwind_offset<R>(signback, xbk, dt, uback);
    const R BV0 = (sb ? p0m : p0) + order2 * (R(0.5) * (sb ? g0m : g0) * tb);
    const R BV1 = (sb ? p1m : p1) + order2 * (R(0.5) * (sb ? g1m : g1) * tb);
    d0 = back * BV0 - front * FV0;
    d1 = back * BV1 - front * FV1;
  }
  __syncthreads();  // transposed reads done before the tiles are reused
  if (tlane) {
    lds[0][ta] = d0;
    lds[1][ta] = d1;
    lds[2][ta] = d3;
    lds[3][ta] = d4;
  }
  __syncthreads();
  // ---- row-major: U += increments, store
  if (!gout) return;
  if (gs >= 1 && gs <= N - 1) {
    U3 += lds[2][ra];
    U0 += lds[0][ra];
    U1 += lds[1][ra];
    if (P.etot) U4 += lds[3][ra];
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    if (fabs(U0) < thresold) U0 = R(0);
    if (fabs(U1) < thresold) U1 = R(0);
  }
  if (FIRST) {
    st_stream<2>(P, &A.Ud0[gq], U0);
    st_stream<2>(P, &A.Ud1[gq], U1);
    st_stream<2>(P, &A.Ud3[gq], U3);
    if (P.etot) st_stream<2>(P, &A.Ud4[gq], U4);
  } else if (U3 != R(0)) {
    st_stream<3>(P, &A.uox[gq], qdiv(U0, U3));
    st_stream<3>(P, &A.uoy[gq], qdiv(U1, U3));
  } else {
    st_stream<3>(P, &A.uox[gq], A.ux[gq]);
    st_stream<3>(P, &A.uoy[gq], A.uy[gq]);
  }
  if (!FIRST && P.etot) st_stream<2>(P, &A.dek[gq], dual_ecin_term(U0, U1, U3, U4));
}

}  // namespace mhg
