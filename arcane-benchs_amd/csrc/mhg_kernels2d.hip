// mhg_kernels2d.hip -- 2D (quad) kernels of the MaHyCo time step for gfx950.
//
// Data layout in HBM (structure of arrays, FP64):
//   cells  i + nx*j : rho, e, p, c, q, m, mask(int32), L, Vlag
//            U[v][c] (v < 3*ne+1: volumes, masses, energies per environment,
//            then q*V) in two ping-pong buffers, contrib[k][c] (half the dual
//            mass flux of each environment)
//   envs   [k][c] : rho, e, p, c, q, m, f, mf  (read/written for cells that
//            hold >= 2 environments only)
//   nodes  i + (nx+1)*j : ux, uy, two ping-pong coordinate buffers (x, y),
//            Ud[3] (dual momentum x, y, mass), dgrad[2] (only the ring of
//            nodes on the domain boundary is ever read back)
// The face normals and the outer face normals of the reference are exactly
// +/- e_dir on a Cartesian mesh (InitGeometricValues, MahycoModule.cc:1008-1050
// evaluates to (1,0,0)-type vectors bit for bit), so dot products with them
// reduce to one component and are not stored.
#include <cstdio>
#include <type_traits>

#include "mhg_common.hpp"
#include "mhg_kernels.hpp"
#include "mhg_kernels_impl.hpp"

namespace mhg {

__device__ inline double node_x0(const Params& P, int i) { return P.org[0] + i * P.h[0]; }
__device__ inline double node_y0(const Params& P, int j) { return P.org[1] + j * P.h[1]; }

// ---------------------------------------------------------------------------
// Start of iteration: time += dt, iteration += 1 (Arcane time-loop manager).
// ---------------------------------------------------------------------------
__global__ void k_begin_iteration(DevScalars* S) {
  if (S->stop || S->err) return;
  S->time += S->dt;
  S->iteration += 1;
}

// ---------------------------------------------------------------------------
// K1: node phase.  updateForceAndVelocity (MahycoModule.cc:589-665) with the
// corner normals of the previous Lagrangian coordinates (saveValuesAtN copies
// m_cell_cqs computed after the previous updatePosition, 344/365),
// applyBoundaryCondition (936-1004), updatePosition (870-931) from x0.
// Node mass = 1/4 sum of the cell masses (computeNodeMass 235-266): the cell
// masses do not change in the Lagrange phase, so it is recomputed here.
// ---------------------------------------------------------------------------
template <class R, bool VEL, bool BC, bool POS>
__global__ void k_node_lagrange_2d(Params P, DevScalars* S, NodeLagArgsT<R> A) {
  if (S->stop || S->err) return;
  const long nid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nid >= P.nn) return;
  const int nnx = P.nx + 1;
  const int i = (int)(nid % nnx), j = (int)(nid / nnx);
  const double dt = S->dt;
  double vx = A.ux[nid], vy = A.uy[nid];
  if (VEL) {
    const double dtv = 0.5 * (S->dt_old + S->dt);
    // 3x3 patch of previous coordinates around the node.
    double X[3][3], Y[3][3];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        int ii = i + di, jj = j + dj;
        double xv = 0., yv = 0.;
        if (ii >= 0 && ii < nnx && jj >= 0 && jj <= P.ny) {
          long q = ii + (long)nnx * jj;
          xv = A.xpx[q];
          yv = A.xpy[q];
        }
        X[dj + 1][di + 1] = xv;
        Y[dj + 1][di + 1] = yv;
      }
    double Fx = 0., Fy = 0., msum = 0.;
    // cells in ascending id: (i-1,j-1) corner 2, (i,j-1) 3, (i-1,j) 1, (i,j) 0
    const int cdi[4] = {-1, 0, -1, 0}, cdj[4] = {-1, -1, 0, 0}, cq[4] = {2, 3, 1, 0};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int ci = i + cdi[t], cj = j + cdj[t];
      if (ci < 0 || ci >= P.nx || cj < 0 || cj >= P.ny) continue;
      long c = ci + (long)P.nx * cj;
      // quad nodes (ci,cj),(ci+1,cj),(ci+1,cj+1),(ci,cj+1) in patch coordinates
      int bi = cdi[t] + 1, bj = cdj[t] + 1;
      double qx[4] = {X[bj][bi], X[bj][bi + 1], X[bj + 1][bi + 1], X[bj + 1][bi]};
      double qy[4] = {Y[bj][bi], Y[bj][bi + 1], Y[bj + 1][bi + 1], Y[bj + 1][bi]};
      double cx, cy;
      cqs2d(qx, qy, cq[t], &cx, &cy);
      double pq = (double)A.p[c] + (double)A.q[c];
      Fx += pq * cx;
      Fy += pq * cy;
      msum += (double)A.m[c];
    }
    double mn = 0.25 * msum;
    double s = dtv / mn;
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
  if (VEL || BC) {
    st_stream(&A.ux[nid], (R)vx);
    st_stream(&A.uy[nid], (R)vy);
  }
  if (POS) {
    st_stream(&A.xnx[nid], node_x0(P, i) + dt * vx);
    st_stream(&A.xny[nid], node_y0(P, j) + dt * vy);
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
__global__ __launch_bounds__(256) void k_cell_lagrange_2d(Params P, DevScalars* S, CellLagArgsT<R> A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  const int nnx = P.nx + 1;
  const int i = (int)(c % P.nx), j = (int)(c / P.nx);
  const long n0 = i + (long)nnx * j;
  const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
  double x[4], y[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = A.xnx[nd[q]];
    y[q] = A.xny[nd[q]];
  }
  const double Vd = quad_volume(x, y);
  const double Ld = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(x, y) : pow(Vd, 0.5);
  lagrange_cell_bulk<R, PC, NEM>(P, S, A, c, Vd, Ld);
}

// ---------------------------------------------------------------------------
// Per-entry-point cell stages.  The same formulas as k_cell_lagrange_2d, with
// every intermediate kept in the state arrays between calls:
//   SAVE       saveValuesAtN cell part (MahycoModule.cc:308-420): rho_n, e_n,
//              q_n <- rho, e, q (global and mixed environments)
//   GEOMETRY   computeGeometricValues (1054-1299): V, L from the coordinates
//   DENSITY    updateDensity (1331-1426): rho = m/V, tau, div u; mixed
//              rho_k = m_k/(f_k V), tau_k
//   VISCOSITY  computeArtificialViscosity (454-578): q, q_k, q = sum f_k q_k
//   ENERGY     updateEnergyAndPressureforGP (1659-1770): e, e_k,
//              e = sum mf_k e_k (no EOS: with-projection)
//   BUILD_U    computeVariablesForRemap_PBorn0 (PrepareRemap.cc:258-399),
//              live entries of U
// ---------------------------------------------------------------------------
template <int STAGE, bool PC, int NEM>
__global__ __launch_bounds__(256) void k_cell_stage_2d(Params P, DevScalars* S, StageArgs A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  const int ne = P.ne;
  const long NC = P.nc;
  const int mask = A.mask[c];
  const int nbe = __popc(mask);
  const int kk = __ffs(mask) - 1;
  if (STAGE == STAGE_SAVE) {
    A.rho_n[c] = A.rho[c];
    A.e_n[c] = A.e[c];
    A.q_n[c] = A.q[c];
    if (nbe >= 2) {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        const long ek = k * NC + c;
        A.env_rho_n[ek] = A.env_rho[ek];
        A.env_e_n[ek] = A.env_e[ek];
        A.env_q_n[ek] = A.env_q[ek];
      }
    }
  } else if (STAGE == STAGE_GEOMETRY) {
    const int nnx = P.nx + 1;
    const int i = (int)(c % P.nx), j = (int)(c / P.nx);
    const long n0 = i + (long)nnx * j;
    const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
    double x[4], y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      x[q] = A.xnx[nd[q]];
      y[q] = A.xny[nd[q]];
    }
    const double V = quad_volume(x, y);
    A.Vlag[c] = V;
    A.L[c] = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(x, y) : pow(V, 0.5);
  } else if (STAGE == STAGE_DENSITY) {
    const double inv_dt = 1.0 / S->dt;
    const double V = A.Vlag[c];
    const double rho_n = A.rho_n[c];
    const double rho = A.m[c] / V;
    const double tau = 0.5 * (1.0 / rho_n + 1.0 / rho);
    A.rho[c] = rho;
    A.tau[c] = tau;
    A.divu[c] = inv_dt * (1.0 / rho - 1.0 / rho_n) / tau;
    if (nbe >= 2) {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        const long ek = k * NC + c;
        const double rk = A.env_m[ek] / (A.env_f[ek] * V);
        A.env_rho[ek] = rk;
        A.env_tau[ek] = 0.5 * (1.0 / A.env_rho_n[ek] + 1.0 / rk);
      }
    }
  } else if (STAGE == STAGE_VISCOSITY) {
    const double divu = A.divu[c], L = A.L[c], cs = A.cs[c];
    double q = 0.;
    if (nbe == 1 && divu < 0.0) q = viscosity(P.eos.qcoef[kk], A.tau[c], L, cs, divu);
    if (nbe >= 2) {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        const long ek = k * NC + c;
        double qk = 0.;
        if (divu < 0.0) qk = viscosity(P.eos.qcoef[k], A.env_tau[ek], L, cs, divu);
        A.env_q[ek] = qk;
        q += qk * A.env_f[ek];
      }
    }
    A.q[c] = q;
  } else if (STAGE == STAGE_ENERGY) {
    double e = 0.;
    if (nbe == 1)
      e = compute_eint(PC, P.eos.gamma[kk], A.q_n[c], A.q[c], A.rho_n[c], A.rho[c], A.p[c], A.e_n[c]);
    if (nbe >= 2) {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        const long ek = k * NC + c;
        const double ev = compute_eint(PC, P.eos.gamma[k], A.env_q_n[ek], A.env_q[ek], A.env_rho_n[ek],
                                       A.env_rho[ek], A.env_p[ek], A.env_e_n[ek]);
        A.env_e[ek] = ev;
        e += A.env_mf[ek] * ev;
      }
    }
    A.e[c] = e;
  } else if (STAGE == STAGE_BUILD_U) {
    const double V = A.Vlag[c];
    double uv[NEM], um[NEM], ue[NEM];
    double uq = 0.;
#pragma unroll
    for (int k = 0; k < NEM; ++k) uv[k] = um[k] = ue[k] = 0.;
    if (nbe == 1) {
      const double rho = A.rho[c];
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k == kk) {
          uv[k] = V;
          um[k] = V * rho;
          ue[k] = V * rho * A.e[c];
        }
      uq = V * A.q[c];
    } else if (nbe >= 2) {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        const long ek = k * NC + c;
        const double Vk = A.env_f[ek] * V;
        const double rk = A.env_rho[ek];
        uv[k] = Vk;
        um[k] = Vk * rk;
        ue[k] = Vk * rk * A.env_e[ek];
        uq = Vk * A.env_q[ek];  // PrepareRemap.cc:372 assigns: the last environment wins
      }
    }
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      if (k >= ne) break;
      A.U[k * NC + c] = uv[k];
      A.U[(ne + k) * NC + c] = um[k];
      A.U[(2 * ne + k) * NC + c] = ue[k];
    }
    A.U[(3 * ne) * NC + c] = uq;
  }
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
template <class R, int D, int NEM, bool FIRST, int LIM>
__global__ __launch_bounds__(D == 0 ? 256 : 1024) void k_cell_sweep_lane_2d(Params P, DevScalars* S,
                                                                          CellSweepArgsT<R> A) {
  if (S->stop || S->err) return;
  const int lane = threadIdx.x & 63;
  const int N = (D == 0) ? P.nx : P.ny;      // cells along D
  const int T = (D == 0) ? P.ny : P.nx;      // cells across D
  const int nseg = (N + 59) / 60;
  int seg, line;                             // segment along D, index across D
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  if (D == 0) {
    const long wave = ((long)blk * blockDim.x + threadIdx.x) >> 6;
    line = (int)(wave / nseg);
    seg = (int)(wave % nseg);
  } else {
    const int ncolblk = (T + 15) / 16;
    line = (blk % ncolblk) * 16 + (threadIdx.x >> 6);
    seg = blk / ncolblk;
  }
  if (line >= T) return;  // uniform across the wave
  const int s = seg * 60 - 2 + lane;
  const bool ex = (s >= 0 && s < N);
  const bool out = ex && lane >= 2 && lane <= 61;
  const int i = (D == 0) ? s : line, j = (D == 0) ? line : s;
  const int nnx = P.nx + 1;
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const long c = ex ? i + (long)P.nx * j : 0;
  const R* __restrict__ Uin = A.Uin;
  const double* __restrict__ Xd = (D == 0) ? A.xnx : A.xny;   // coordinate along D
  const double* __restrict__ Xt = (D == 0) ? A.xny : A.xnx;   // transverse coordinate
  const R* __restrict__ Ud = (D == 0) ? A.ux : A.uy;          // velocity along D
  // cell centre (component D, double) and phi denominator
  double xc = 0.;
  R den = R(1);
  if (ex) {
    const long m0 = i + (long)nnx * j;
    const double a = Xd[m0], b = Xd[m0 + 1], cc = Xd[m0 + 1 + nnx], dd = Xd[m0 + nnx];
    xc = 0.25 * (((a + b) + cc) + dd);
    if (FIRST) {
      den = A.Vlag[c];
    } else {
      R sv = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k < ne) sv += Uin[k * NC + c];
      den = sv;
    }
  }
  // back face of the cell: the face at node index s along D, nodes
  // (i,j),(i,j+1) for D = 0 and (i,j),(i+1,j) for D = 1
  R uf = R(0);
  double lf = 0., xf = 0.;
  if (s >= 0 && s <= N) {
    const long na = i + (long)nnx * j, nb = (D == 0) ? na + nnx : na + 1;
    uf = R(0.5) * (Ud[na] + Ud[nb]);
    xf = 0.5 * Xd[na] + 0.5 * Xd[nb];
    lf = fabs(Xt[nb] - Xt[na]);
  }
  const bool inner = ex && s >= 1;
  const double xcm = __shfl_up(xc, 1, 64);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = __shfl_down(uf, 1, 64);
  const double lf_f = __shfl_down(lf, 1, 64);
  // face flux factors u_f L_f dt (double products, then the storage type)
  const R fb = (R)(-1.0 * uf * lf * dt), ff = (R)(1.0 * uf_f * lf_f * dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  const bool o2 = P.order2 > 0;
  R pre[3][NEM];
  const R preq = ex ? Uin[(3 * ne) * NC + c] : R(0);
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int k = 0; k < NEM; ++k) pre[g][k] = (ex && k < ne) ? Uin[(g * ne + k) * NC + c] : R(0);
  auto sweep_var = [&](R u0, R& unew, R& dual) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = __shfl_up(phi, 1, 64);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = __shfl_down(gf, 1, 64);
    const R g = o2 ? limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp) : R(0);
    const R gm = __shfl_up(g, 1, 64);
    const R phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    const R phifp = __shfl_down(phif, 1, 64);
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
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[0][k], un, du);
    if (out) st_stream(&Uout[k * NC + c], un);
  }
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[1][k], un, du);
    if (un < R(0)) un = R(0);  // RemapADIService.cc:953-955
    if (out) {
      st_stream(&Uout[(ne + k) * NC + c], un);
      st_stream(&A.contrib[k * NC + c], R(0.5) * du);  // fcontrib (DualPhaseRemap.cc:168-185)
    }
  }
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    R un, du;
    sweep_var(pre[2][k], un, du);
    if (un < R(0)) un = R(0);  // RemapADIService.cc:956-958
    if (out) st_stream(&Uout[(2 * ne + k) * NC + c], un);
  }
  {
    R un, du;
    sweep_var(preq, un, du);
    if (out) st_stream(&Uout[(3 * ne) * NC + c], un);
  }
}

// y sweep: each lane owns one column and walks a strip of rows, keeping a
// two-cell window of phi, face gradients, limited gradients and face values
// in registers.  Row r loads cell r, forms the gradient of face r, the
// limited gradient of cell r-1, the face value of face r-1 and updates cell
// r-2.  Face r lies between cells r-1 and r (node row r).  The variables of U
// are independent in a sweep (only the denominator sum_k U_vol,k couples
// them), so a thread carries VG of them and the variable group is a grid
// dimension.  Rows are processed in chunks of R: all loads of a chunk are
// issued before its arithmetic so that R rows of memory traffic are in flight.
// y sweep on LDS tiles: a block of 16 wavefronts owns 16 columns (one
// 128-byte line of a row) and 64 rows (60 outputs, two halo rows on each
// side).  Global loads and stores are row-major and coalesced; the tile is
// transposed through LDS so that each wavefront holds one column with one row
// per lane and runs the lane-exchange arithmetic of the x sweep.


template <class R, int NEM, bool FIRST, int LIM>
__global__ __launch_bounds__(1024) void k_cell_sweep_ytile_2d(Params P, DevScalars* S, CellSweepArgsT<R> A) {
  // geometry tiles [0..4]; once read, tiles [0..5] are reused as the staging
  // tiles (in, out, dual flux) of a pair of variables
  __shared__ double lds[6][kTileR][kTileP];
  typedef R Tile[kTileR][kTileP];
  Tile* stg = reinterpret_cast<Tile*>(&lds[0][0][0]);
  if (S->stop || S->err) return;  // uniform over the grid
  const int nx = P.nx, N = P.ny, nnx = nx + 1;
  const int ncolblk = (nx + kTileW - 1) / kTileW;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (L % ncolblk) * kTileW;
  const int s0 = (L / ncolblk) * 60 - 2;  // cell row of tile row 0
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;     // row-major position (loads, stores)
  const int w = t >> 6, lane = t & 63;    // transposed position: column w, row lane
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const bool o2 = P.order2 > 0;
  const R* __restrict__ Uin = A.Uin;
  // row-major cell of this thread
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = gi < nx && gs >= 0 && gs < N;
  const bool gout = gex && tr >= 2 && tr <= 61;
  const long gc = gex ? gi + (long)nx * gs : 0;
  {  // geometry of the tile, row-major (stored in the double tiles exactly)
    double xc = 0., xf = 0., lf = 0.;
    R den = R(1), uf = R(0);
    if (gex) {
      const long m0 = gi + (long)nnx * gs;
      const double a = A.xny[m0], b = A.xny[m0 + 1], cc = A.xny[m0 + 1 + nnx], dd = A.xny[m0 + nnx];
      xc = 0.25 * (((a + b) + cc) + dd);
      if (FIRST) {
        den = A.Vlag[gc];
      } else {
        R sv = R(0);
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if (k < ne) sv += Uin[k * NC + gc];
        den = sv;
      }
    }
    if (gi < nx && gs >= 0 && gs <= N) {  // back face: nodes (i, s), (i+1, s)
      const long na = gi + (long)nnx * gs, nb = na + 1;
      uf = R(0.5) * (A.uy[na] + A.uy[nb]);
      xf = 0.5 * A.xny[na] + 0.5 * A.xny[nb];
      lf = fabs(A.xnx[nb] - A.xnx[na]);
    }
    lds[0][tr][tc] = xc;
    lds[1][tr][tc] = den;
    lds[2][tr][tc] = uf;
    lds[3][tr][tc] = xf;
    lds[4][tr][tc] = lf;
  }
  __syncthreads();
  const int s = s0 + lane;
  const bool ex = (i0 + w < nx) && s >= 0 && s < N;
  const bool inner = ex && s >= 1;
  const double xc = lds[0][lane][w], xf = lds[3][lane][w], lf = lds[4][lane][w];
  const R den = (R)lds[1][lane][w], uf = (R)lds[2][lane][w];
  const double xcm = __shfl_up(xc, 1, 64);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = __shfl_down(uf, 1, 64);
  const double lf_f = __shfl_down(lf, 1, 64);
  const R fb = (R)(-1.0 * uf * lf * dt), ff = (R)(1.0 * uf_f * lf_f * dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  // one variable on the transposed layout: new U and half the dual flux
  auto sweep_var = [&](R u0, R& un, R& dual) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = __shfl_up(phi, 1, 64);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = __shfl_down(gf, 1, 64);
    const R g = o2 ? limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp) : R(0);
    const R gm = __shfl_up(g, 1, 64);
    const R phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    const R phifp = __shfl_down(phif, 1, 64);
    const R flux_b = fb * phif;
    const R flux_f = ff * phifp;
    // cell faces in order: bottom (back), right, top (front), left
    un = u0 - (flux_b + flux_f);
    dual = R(0.5) * flux_b * R(-1.0) + R(0.5) * flux_f * R(1.0);
  };
  __syncthreads();  // geometry reads done: the tiles become staging tiles
  const int nvu = P.nvu;
  // variables two at a time (a, b = a + 1), one barrier pair per round
  R nxa = gex ? Uin[gc] : R(0);
  R nxb = (gex && nvu > 1) ? Uin[NC + gc] : R(0);
  for (int va = 0; va < nvu; va += 2) {
    const int vb = va + 1;
    const bool hb = vb < nvu;  // uniform
    stg[0][tr][tc] = nxa;
    stg[1][tr][tc] = nxb;
    __syncthreads();
    if (va + 2 < nvu) nxa = gex ? Uin[(va + 2) * NC + gc] : R(0);
    if (va + 3 < nvu) nxb = gex ? Uin[(va + 3) * NC + gc] : R(0);
    const R ua = stg[0][lane][w], ub = stg[1][lane][w];
    R una, dua, unb, dub;
    sweep_var(ua, una, dua);
    sweep_var(ub, unb, dub);
    if (va >= ne && va < 3 * ne && una < R(0)) una = R(0);  // RemapADIService.cc:953-958
    if (vb >= ne && vb < 3 * ne && unb < R(0)) unb = R(0);
    const bool ma = (va >= ne && va < 2 * ne), mb = hb && (vb >= ne && vb < 2 * ne);
    stg[2][lane][w] = una;
    stg[3][lane][w] = unb;
    if (ma) stg[4][lane][w] = R(0.5) * dua;  // fcontrib (DualPhaseRemap.cc:168-185)
    if (mb) stg[5][lane][w] = R(0.5) * dub;
    __syncthreads();
    if (gout) {
      st_stream(&A.Uout[va * NC + gc], stg[2][tr][tc]);
      if (hb) st_stream(&A.Uout[vb * NC + gc], stg[3][tr][tc]);
      if (ma) st_stream(&A.contrib[(va - ne) * NC + gc], stg[4][tr][tc]);
      if (mb) st_stream(&A.contrib[(vb - ne) * NC + gc], stg[5][tr][tc]);
    }
  }
}

template <bool FIRST, int STRIP, int VG, int R, int LIM>
__global__ __launch_bounds__(64) void k_cell_sweep_y_2d(Params P, DevScalars* S, CellSweepArgs A) {
  if (S->stop || S->err) return;
  // logical tile: variable group fastest, then column block, then strip
  const int ngroups = (P.nvu + VG - 1) / VG;
  const int ncb = (P.nx + 63) / 64;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int vg = L % ngroups;
  const int t = L / ngroups;
  const int i = (t % ncb) * 64 + threadIdx.x;
  const int strip = t / ncb;
  if (i >= P.nx) return;
  const int j0 = strip * STRIP;
  const int j1 = min(j0 + STRIP, P.ny);
  const int v0 = vg * VG;
  const int nnx = P.nx + 1;
  const long NC = P.nc;
  const int ne = P.ne;
  const int nvu = P.nvu;
  const double dt = S->dt;
  const double order2 = (double)P.order2;
  const bool o2 = P.order2 > 0;
  const double* __restrict__ Uin = A.Uin;
  double* __restrict__ Uout = A.Uout;
  double ph1[VG], ph2[VG], gf1[VG], g2[VG], pf2[VG], U1[VG], U2[VG];
#pragma unroll
  for (int s = 0; s < VG; ++s) ph1[s] = ph2[s] = gf1[s] = g2[s] = pf2[s] = U1[s] = U2[s] = 0.;
  double xc1 = 0., xc2 = 0.;                      // centres of cells r-1, r-2
  double uf1 = 0., lf1 = 0., xf1 = 0., dx1 = 1.;  // face r-1
  double uf2 = 0., lf2 = 0.;                      // face r-2
  bool in1 = false;                               // face r-1 inner
  for (int r0 = j0 - 2; r0 <= j1 + 1; r0 += R) {
    // ---- loads of the chunk: node rows r0..r0+R, cell rows r0..r0+R-1
    double ya[R + 1], yb[R + 1], xa[R + 1], xb[R + 1], va[R + 1], vb[R + 1];
#pragma unroll
    for (int t = 0; t <= R; ++t) {
      const int r = r0 + t;
      ya[t] = yb[t] = xa[t] = xb[t] = va[t] = vb[t] = 0.;
      if (r >= 0 && r <= P.ny) {
        const long na = i + (long)nnx * r;
        ya[t] = A.xny[na];
        yb[t] = A.xny[na + 1];
        xa[t] = A.xnx[na];
        xb[t] = A.xnx[na + 1];
        va[t] = A.uy[na];
        vb[t] = A.uy[na + 1];
      }
    }
    double den[R], Ur[R][VG];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int r = r0 + t;
      const bool rv = (r >= 0 && r < P.ny);
      const long c = rv ? i + (long)P.nx * r : 0;
      den[t] = 1.;
      if (rv) {
        if (FIRST) {
          den[t] = A.Vlag[c];
        } else {
          double sv = 0.;
          for (int k = 0; k < ne; ++k) sv += Uin[k * NC + c];
          den[t] = sv;
        }
      }
#pragma unroll
      for (int s = 0; s < VG; ++s) Ur[t][s] = (rv && v0 + s < nvu) ? Uin[(v0 + s) * NC + c] : 0.;
    }
    // ---- arithmetic of the chunk, row by row
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int r = r0 + t;
      if (r > j1 + 1) break;
      const bool rv = (r >= 0 && r < P.ny);
      // cell centre: nodes (i,r),(i+1,r),(i+1,r+1),(i,r+1) in quad order
      const double xc = rv ? 0.25 * (((ya[t] + yb[t]) + yb[t + 1]) + ya[t + 1]) : 0.;
      // face r: nodes (i,r), (i+1,r)
      const bool fv = (r >= 0 && r <= P.ny);
      const double uf = fv ? 0.5 * (va[t] + vb[t]) : 0.;
      const double xf = fv ? 0.5 * ya[t] + 0.5 * yb[t] : 0.;
      const double lf = fv ? fabs(xb[t] - xa[t]) : 0.;
      const bool inner = rv && r >= 1;
      const double dx = inner ? xc - xc1 : 1.;
      const bool upm1 = in1 && (uf1 * dx1 > 0.0);
      const double dot1 = in1 ? (upm1 ? xf1 - xc2 : xf1 - xc1) : 0.;
      const int out_row = r - 2;
      const bool out = (out_row >= j0 && out_row < j1);
      const long co = out ? i + (long)P.nx * out_row : 0;
#pragma unroll
      for (int s = 0; s < VG; ++s) {
        const int v = v0 + s;
        if (v >= nvu) break;
        const double ur = Ur[t][s];
        const double phi = rv ? ur / den[t] : 0.;
        const double gf = face_grad(inner, phi, ph1[s], dx);
        const double g1 = o2 ? limited_grad<LIM, true>(gf1[s], gf) : 0.;  // cell r-1
        const double pf1 = in1 ? (upm1 ? ph2[s] + order2 * dot1 * g2[s] : ph1[s] + order2 * dot1 * g1) : 0.;
        if (out) {
          const double flux_b = -1.0 * uf2 * lf2 * dt * pf2[s];
          const double flux_f = 1.0 * uf1 * lf1 * dt * pf1;
          // cell faces in order: bottom (back), right, top (front), left
          double un = U2[s] - (flux_b + flux_f);
          if (v >= ne && v < 3 * ne && un < 0.) un = 0.;
          st_stream(&Uout[v * NC + co], un);
          if (v >= ne && v < 2 * ne) {
            const double dual = 0.5 * flux_b * -1.0 + 0.5 * flux_f * 1.0;
            st_stream(&A.contrib[(v - ne) * NC + co], 0.5 * dual);
          }
        }
        ph2[s] = ph1[s];
        ph1[s] = phi;
        gf1[s] = gf;
        g2[s] = g1;
        pf2[s] = pf1;
        U2[s] = U1[s];
        U1[s] = ur;
      }
      xc2 = xc1;
      xc1 = xc;
      uf2 = uf1;
      lf2 = lf1;
      uf1 = uf;
      lf1 = lf;
      xf1 = xf;
      dx1 = dx;
      in1 = inner;
    }
  }
}

// ---------------------------------------------------------------------------
// K5: final cell phase.  remapVariables cell status + membership change
// (RemapADIFinal.cc:98-226), "moy" (446-643), applyEOS per environment
// (PerfectGasEOSService.cc:50-137), computePressionMoyenne (MahycoModule.cc:
// 1836-1915) and the CFL candidate of computeHydroDeltaT (2047-2061) reduced
// to one minimum per block; k_dt_finalize finishes computeDeltaT (2076-2191).
// ---------------------------------------------------------------------------
template <class R, int BLOCK, int NEM>
__global__ __launch_bounds__(BLOCK) void k_cell_final_2d(Params P, DevScalars* S, CellFinalArgsT<R> A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double dx_sound = 1.7976931348623157e308;
  if (c < P.nc) {
    const int nnx = P.nx + 1;
    const int i = (int)(c % P.nx), j = (int)(c / P.nx);
    // Euler volume: the corner-normal formula on the initial coordinates.
    double x0[4] = {node_x0(P, i), node_x0(P, i + 1), node_x0(P, i + 1), node_x0(P, i)};
    double y0[4] = {node_y0(P, j), node_y0(P, j), node_y0(P, j + 1), node_y0(P, j + 1)};
    const R c_g = final_cell<R, NEM>(P, A, c, quad_volume(x0, y0));
    // CFL candidate: L / (c + max over the cell nodes of |u|)
    const long n0 = i + (long)nnx * j;
    const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
    R vmax = R(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const R ux = A.ux[nd[q]], uy = A.uy[nd[q]];
      const R un = sqrt(ux * ux + uy * uy + R(0) * R(0));
      vmax = (un < vmax) ? vmax : un;
    }
    // std::min(minimum_aux, dx_sound) fold: a NaN candidate is skipped
    dx_sound = smin(dx_sound, (double)qdiv(A.L[c], c_g + vmax));
  }
  block_min_store<BLOCK>(dx_sound, A.block_min);
}

// computeDeltaT (MahycoModule.cc:2076-2191) on the block minima of
// k_cell_final_2d: dt = min(cfl * min_c dx_sound, 1.05 dt_old, deltat-max),
// deltat-min check, stop test with the time of this iteration.  ADVANCE
// first applies this iteration's time += dt, iteration += 1 (the work of
// k_begin_iteration, which no kernel of the step between them reads).
template <int BLOCK, bool ADVANCE>
__global__ __launch_bounds__(BLOCK) void k_dt_finalize(Params P, DevScalars* S, const double* block_min, int nb) {
  if (S->stop || S->err) return;
  __shared__ double red[BLOCK / 64];
  double v = 1.7976931348623157e308;
  for (int b = threadIdx.x; b < nb; b += BLOCK) v = smin(v, block_min[b]);
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double minimum_aux = red[0];
  for (int w = 1; w < BLOCK / 64; ++w) minimum_aux = smin(minimum_aux, red[w]);
  const double old = S->dt;
  if (ADVANCE) {
    S->time += old;
    S->iteration += 1;
  }
  S->dt_old = old;
  double new_dt = P.cfl * minimum_aux;
  new_dt = smin(new_dt, 1.05 * old);
  new_dt = smin(new_dt, P.dtmax);
  if (new_dt < P.dtmin) {
    S->err = 1;
    return;
  }
  S->dt = new_dt;
  const double t = S->time;
  const bool not_yet_finish = (t < P.tend);
  const bool finish = (t > P.tend);
  const bool too_much = ((t + new_dt) > P.tend);
  if ((not_yet_finish && too_much) || finish) S->stop = 1;
}

// ---------------------------------------------------------------------------
// K4: dual (node) sweep in direction D.  computeDualGradPhi_LimC
// (DualPhaseRemap.cc:700-779), node mass fluxes (ncontrib 215-260), upwind
// momentum update (352-491) and the phi_dual refresh (494-523).  FIRST builds
// the dual state from (u, m_node) as computeVariablesForRemap_PBorn0 does
// (PrepareRemap.cc:411-433); the last sweep ends with the nodal velocity
// u = U_dual / m_dual of remapVariables (RemapADIFinal.cc:660-672).
// ---------------------------------------------------------------------------
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

template <int D, bool FIRST, int LIM>
__global__ void k_node_sweep_2d(Params P, DevScalars* S, NodeSweepArgs A) {
  if (S->stop || S->err) return;
  const long nid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nid >= P.nn) return;
  const int nnx = P.nx + 1;
  const int i = (int)(nid % nnx), j = (int)(nid / nnx);
  const int s = (D == 0) ? i : j;
  const int N = (D == 0) ? P.nx : P.ny;           // nodes 0..N along D
  const int T = (D == 0) ? P.ny : P.nx;           // nodes 0..T along the other direction
  const int st = (D == 0) ? j : i;
  const long nstride = (D == 0) ? 1 : nnx;
  const bool inner_d = (s >= 1 && s <= N - 1);
  const bool inner_t = (st >= 1 && st <= T - 1);  // inner in the other direction
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;

  // Dual state of this node before the sweep.
  double U0, U1, U3;
  if (FIRST) {
    const double mn = node_mass_from_cells(P, A.m, i, j);
    U0 = mn * A.ux[nid];
    U1 = mn * A.uy[nid];
    U3 = mn;
  } else {
    U0 = A.Ud0[nid];
    U1 = A.Ud1[nid];
    U3 = A.Ud3[nid];
  }
  if (inner_d) {
    // phi_dual (u_x, u_y) and coordinate along D on the 5-node stencil.
    bool ex[5];
    double ph0[5], ph1[5], xd[5];
    const double* __restrict__ Xd = (D == 0) ? A.xnx : A.xny;
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
      const int so = s + o;
      ex[o + 2] = (so >= 0 && so <= N);
      ph0[o + 2] = ph1[o + 2] = 0.;
      xd[o + 2] = 0.;
      if (!ex[o + 2]) continue;
      const long q = nid + o * nstride;
      xd[o + 2] = Xd[q];
      // In the second sweep the nodes that were inner in the first direction
      // carry phi = U/m (494-523); the others still hold the initial phi = u.
      if (FIRST || !inner_t) {
        ph0[o + 2] = A.ux[q];
        ph1[o + 2] = A.uy[q];
      } else {
        const double m3 = A.Ud3[q];
        ph0[o + 2] = A.Ud0[q] / m3;
        ph1[o + 2] = A.Ud1[q] / m3;
      }
    }
    // limited nodal gradients of the stencil nodes -1, 0, +1
    double g0[3] = {0., 0., 0.}, g1[3] = {0., 0., 0.};
    if (P.order2 > 0) {
#pragma unroll
      for (int o = -1; o <= 1; ++o) {
        const int so = s + o;
        if (so >= 1 && so <= N - 1) {
          const double dxf = xd[o + 3] - xd[o + 2], dxb = xd[o + 2] - xd[o + 1];
          const double gf0 = (ph0[o + 3] - ph0[o + 2]) / dxf, gb0 = (ph0[o + 2] - ph0[o + 1]) / dxb;
          const double gf1 = (ph1[o + 3] - ph1[o + 2]) / dxf, gb1 = (ph1[o + 2] - ph1[o + 1]) / dxb;
          g0[o + 1] = limited_grad<LIM, false>(gb0, gf0);
          g1[o + 1] = limited_grad<LIM, false>(gb1, gf1);
        } else {  // boundary node along D: value left by an earlier sweep
          const long q = nid + o * nstride;
          g0[o + 1] = A.dg0[q];
          g1[o + 1] = A.dg1[q];
        }
      }
      // keep the gradient of nodes that lie on the boundary of the other direction
      if (!inner_t) {
        A.dg0[nid] = g0[1];
        A.dg1[nid] = g1[1];
      }
    }
    // node mass fluxes (back cells then front cells, ascending shift)
    double back = 0., front = 0.;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
      for (int sh = -1; sh <= 0; ++sh) {
        int ci, cj;
        if (D == 0) {
          ci = i - (pass == 0 ? 1 : 0);
          cj = j + sh;
        } else {
          ci = i + sh;
          cj = j - (pass == 0 ? 1 : 0);
        }
        if (ci < 0 || ci >= P.nx || cj < 0 || cj >= P.ny) continue;
        const long cc = ci + (long)P.nx * cj;
        for (int k = 0; k < ne; ++k) {
          const double w = A.contrib[k * NC + cc];
          if (pass == 0) back += w;
          else front += w;
        }
      }
    }
    U3 += back - front;
    const double order2 = (double)P.order2;
    const double* phd = (D == 0) ? ph0 : ph1;  // component along D
    const double ufront = 0.5 * (phd[2] + phd[2]);  // DualPhaseRemap.cc:391
    const int signfront = (ufront > 0) ? 1 : -1;
    const double xfr = xd[3] - xd[2];
    const int sf = (front < 0.) ? 1 : 0;
    const double tf = signfront * xfr - dt * ufront;
    const double FV0 = ph0[sf + 2] + order2 * (0.5 * g0[sf + 1] * tf);
    const double FV1 = ph1[sf + 2] + order2 * (0.5 * g1[sf + 1] * tf);
    const double uback = 0.5 * (phd[1] + phd[2]);
    const int signback = (uback > 0) ? 1 : -1;
    const double xbk = xd[2] - xd[1];
    const int sb = (back > 0.) ? -1 : 0;
    const double tb = signback * xbk - dt * uback;
    const double BV0 = ph0[sb + 2] + order2 * (0.5 * g0[sb + 1] * tb);
    const double BV1 = ph1[sb + 2] + order2 * (0.5 * g1[sb + 1] * tb);
    U0 += back * BV0 - front * FV0;
    U1 += back * BV1 - front * FV1;
    const double thresold = 1.e-300;
    if (fabs(U0) < thresold) U0 = 0.;
    if (fabs(U1) < thresold) U1 = 0.;
  }
  if (FIRST) {
    st_stream(&A.Ud0[nid], U0);
    st_stream(&A.Ud1[nid], U1);
    st_stream(&A.Ud3[nid], U3);
  } else {
    if (U3 != 0.) {
      st_stream(&A.uox[nid], U0 / U3);
      st_stream(&A.uoy[nid], U1 / U3);
    } else {
      st_stream(&A.uox[nid], A.ux[nid]);
      st_stream(&A.uoy[nid], A.uy[nid]);
    }
  }
}

// Dual sweep along x with lane exchange: a wavefront holds 64 consecutive
// nodes of a row (60 outputs, two halo nodes on each side).  Every node forms
// its phi_dual and its limited gradient once and reads its neighbours' values
// across lanes.  front(i) sums the cells (i, j-1), (i, j) in the same order as
// back(i+1) does, so back is the front of the previous lane.
template <class R, bool FIRST, int NEM, int LIM>
__global__ __launch_bounds__(256) void k_node_sweep_x_2d(Params P, DevScalars* S, NodeSweepArgsT<R> A) {
  if (S->stop || S->err) return;
  const int lane = threadIdx.x & 63;
  const int N = P.nx;
  const int nseg = (N + 1 + 59) / 60;
  const long wave = ((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
  const int j = (int)(wave / nseg);
  if (j > P.ny) return;  // uniform across the wave
  const int seg = (int)(wave % nseg);
  const int i = seg * 60 - 2 + lane;
  const bool ex = (i >= 0 && i <= N);
  const bool out = ex && lane >= 2 && lane <= 61;
  const long nid = ex ? i + (long)(N + 1) * j : 0;
  const bool inner_t = (j >= 1 && j <= P.ny - 1);
  const bool inner_d = (i >= 1 && i <= N - 1);
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;
  R U0 = R(0), U1 = R(0), U3 = R(0), p0 = R(0), p1 = R(0);
  double xd = 0.;
  if (ex) {
    xd = A.xnx[nid];
    if (FIRST) {
      const R mn = node_mass_from_cells(P, A.m, i, j);
      p0 = A.ux[nid];
      p1 = A.uy[nid];
      U0 = mn * p0;
      U1 = mn * p1;
      U3 = mn;
    } else {
      U0 = A.Ud0[nid];
      U1 = A.Ud1[nid];
      U3 = A.Ud3[nid];
      // nodes inner in the first direction carry phi = U/m (DualPhaseRemap.cc:494-523)
      if (inner_t) {
        p0 = qdiv(U0, U3);
        p1 = qdiv(U1, U3);
      } else {
        p0 = A.ux[nid];
        p1 = A.uy[nid];
      }
    }
  }
  // front dual mass flux: cells (i, j-1), (i, j), environments in order
  R front = R(0);
  if (ex && i < P.nx) {
#pragma unroll
    for (int sh = -1; sh <= 0; ++sh) {
      const int cj = j + sh;
      if (cj < 0 || cj >= P.ny) continue;
      const long cc = i + (long)P.nx * cj;
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (k >= ne) break;
        front += A.contrib[k * NC + cc];
      }
    }
  }
  const double xdm = __shfl_up(xd, 1, 64), xdp = __shfl_down(xd, 1, 64);
  const R p0m = __shfl_up(p0, 1, 64), p0p = __shfl_down(p0, 1, 64);
  const R p1m = __shfl_up(p1, 1, 64), p1p = __shfl_down(p1, 1, 64);
  const R back = __shfl_up(front, 1, 64);
  // limited nodal gradient (computeDualGradPhi_LimC, DualPhaseRemap.cc:700-779)
  R g0 = R(0), g1 = R(0);
  if (P.order2 > 0 && ex) {
    if (inner_d) {
      const R dxf = (R)(xdp - xd), dxb = (R)(xd - xdm);
      const R gf0 = qdiv(p0p - p0, dxf), gb0 = qdiv(p0 - p0m, dxb);
      const R gf1 = qdiv(p1p - p1, dxf), gb1 = qdiv(p1 - p1m, dxb);
      g0 = limited_grad<LIM, false, R, kBranchFree<R>>(gb0, gf0);
      g1 = limited_grad<LIM, false, R, kBranchFree<R>>(gb1, gf1);
      if (!inner_t && out) {  // kept for the sweep in the other direction
        A.dg0[nid] = g0;
        A.dg1[nid] = g1;
      }
    } else {  // boundary node along x: gradient left by an earlier sweep
      g0 = A.dg0[nid];
      g1 = A.dg1[nid];
    }
  }
  const R g0m = __shfl_up(g0, 1, 64), g0p = __shfl_down(g0, 1, 64);
  const R g1m = __shfl_up(g1, 1, 64), g1p = __shfl_down(g1, 1, 64);
  if (!out) return;
  if (inner_d) {
    U3 += back - front;
    const R order2 = (R)P.order2;
    const R ufront = R(0.5) * (p0 + p0);  // DualPhaseRemap.cc:391
    const int signfront = (ufront > R(0)) ? 1 : -1;
    const double xfr = xdp - xd;
    const bool sf = (front < R(0));
    const R tf = (R)(signfront * xfr - dt * ufront);
    const R FV0 = (sf ? p0p : p0) + order2 * (R(0.5) * (sf ? g0p : g0) * tf);
    const R FV1 = (sf ? p1p : p1) + order2 * (R(0.5) * (sf ? g1p : g1) * tf);
    const R uback = R(0.5) * (p0m + p0);
    const int signback = (uback > R(0)) ? 1 : -1;
    const double xbk = xd - xdm;
    const bool sb = (back > R(0));
    const R tb = (R)(signback * xbk - dt * uback);
    const R BV0 = (sb ? p0m : p0) + order2 * (R(0.5) * (sb ? g0m : g0) * tb);
    const R BV1 = (sb ? p1m : p1) + order2 * (R(0.5) * (sb ? g1m : g1) * tb);
    U0 += back * BV0 - front * FV0;
    U1 += back * BV1 - front * FV1;
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    if (fabs(U0) < thresold) U0 = R(0);
    if (fabs(U1) < thresold) U1 = R(0);
  }
  if (FIRST) {
    st_stream(&A.Ud0[nid], U0);
    st_stream(&A.Ud1[nid], U1);
    st_stream(&A.Ud3[nid], U3);
  } else if (U3 != R(0)) {
    st_stream(&A.uox[nid], qdiv(U0, U3));
    st_stream(&A.uoy[nid], qdiv(U1, U3));
  } else {
    st_stream(&A.uox[nid], A.ux[nid]);
    st_stream(&A.uoy[nid], A.uy[nid]);
  }
}

// Dual sweep along y on LDS tiles: a block of 16 wavefronts owns 16 node
// columns and 64 node rows (60 outputs, two halo rows on each side).  The
// row-major pass loads the nodes (coalesced) and forms phi_dual and the front
// flux; the tile is transposed through LDS, each wavefront runs the lane
// arithmetic of the x dual sweep on one column, and the increments of the
// dual state travel back through LDS to be added and stored row-major.
template <class R, bool FIRST, int NEM, int LIM>
__global__ __launch_bounds__(1024) void k_node_sweep_ytile_2d(Params P, DevScalars* S, NodeSweepArgsT<R> A) {
  __shared__ double ldx[kTileR][kTileP];   // coordinate along y
  __shared__ R lds[3][kTileR][kTileP];     // phi_dual x, y and front flux; then the increments
  if (S->stop || S->err) return;  // uniform over the grid
  const int nx = P.nx, N = P.ny, nnx = nx + 1;
  const int ncolblk = (nnx + kTileW - 1) / kTileW;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (L % ncolblk) * kTileW;
  const int s0 = (L / ncolblk) * 60 - 2;  // node row of tile row 0
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;
  const int w = t >> 6, lane = t & 63;
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;
  // ---- row-major pass
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = gi <= nx && gs >= 0 && gs <= N;
  const bool gout = gex && tr >= 2 && tr <= 61;
  const long gq = gex ? gi + (long)nnx * gs : 0;
  R U0 = R(0), U1 = R(0), U3 = R(0);
  {
    double xd = 0.;
    R p0 = R(0), p1 = R(0), f = R(0);
    if (gex) {
      const bool inner_t = (gi >= 1 && gi <= nx - 1);
      xd = A.xny[gq];
      if (FIRST) {
        const R mn = node_mass_from_cells(P, A.m, gi, gs);
        p0 = A.ux[gq];
        p1 = A.uy[gq];
        U0 = mn * p0;
        U1 = mn * p1;
        U3 = mn;
      } else {
        U0 = A.Ud0[gq];
        U1 = A.Ud1[gq];
        U3 = A.Ud3[gq];
        if (inner_t) {  // DualPhaseRemap.cc:494-523
          p0 = qdiv(U0, U3);
          p1 = qdiv(U1, U3);
        } else {
          p0 = A.ux[gq];
          p1 = A.uy[gq];
        }
      }
      if (gs < N) {  // front flux: cells (i-1, s), (i, s), environments in order
#pragma unroll
        for (int sh = -1; sh <= 0; ++sh) {
          const int ci = gi + sh;
          if (ci < 0 || ci >= nx) continue;
          const long cc = ci + (long)nx * gs;
#pragma unroll
          for (int k = 0; k < NEM; ++k) {
            if (k >= ne) break;
            f += A.contrib[k * NC + cc];
          }
        }
      }
    }
    ldx[tr][tc] = xd;
    lds[0][tr][tc] = p0;
    lds[1][tr][tc] = p1;
    lds[2][tr][tc] = f;
  }
  __syncthreads();
  // ---- transposed pass: column i, node row s
  const int i = i0 + w, s = s0 + lane;
  const bool ex = i <= nx && s >= 0 && s <= N;
  const bool out = ex && lane >= 2 && lane <= 61;
  const bool inner_t = (i >= 1 && i <= nx - 1);
  const bool inner_d = (s >= 1 && s <= N - 1);
  const double xd = ldx[lane][w];
  const R p0 = lds[0][lane][w], p1 = lds[1][lane][w];
  const R front = lds[2][lane][w];
  const double xdm = __shfl_up(xd, 1, 64), xdp = __shfl_down(xd, 1, 64);
  const R p0m = __shfl_up(p0, 1, 64), p0p = __shfl_down(p0, 1, 64);
  const R p1m = __shfl_up(p1, 1, 64), p1p = __shfl_down(p1, 1, 64);
  const R back = __shfl_up(front, 1, 64);
  R g0 = R(0), g1 = R(0);
  if (P.order2 > 0 && ex) {
    const long q = i + (long)nnx * s;
    if (inner_d) {
      const R dxf = (R)(xdp - xd), dxb = (R)(xd - xdm);
      const R gf0 = qdiv(p0p - p0, dxf), gb0 = qdiv(p0 - p0m, dxb);
      const R gf1 = qdiv(p1p - p1, dxf), gb1 = qdiv(p1 - p1m, dxb);
      g0 = limited_grad<LIM, false, R, kBranchFree<R>>(gb0, gf0);
      g1 = limited_grad<LIM, false, R, kBranchFree<R>>(gb1, gf1);
      if (!inner_t && out) {  // kept for the sweep in the other direction
        A.dg0[q] = g0;
        A.dg1[q] = g1;
      }
    } else {  // boundary node along y: gradient left by an earlier sweep
      g0 = A.dg0[q];
      g1 = A.dg1[q];
    }
  }
  const R g0m = __shfl_up(g0, 1, 64), g0p = __shfl_down(g0, 1, 64);
  const R g1m = __shfl_up(g1, 1, 64), g1p = __shfl_down(g1, 1, 64);
  R d0 = R(0), d1 = R(0), d3 = R(0);
  if (out && inner_d) {
    const R order2 = (R)P.order2;
    d3 = back - front;
    const R ufront = R(0.5) * (p1 + p1);  // DualPhaseRemap.cc:391
    const int signfront = (ufront > R(0)) ? 1 : -1;
    const double xfr = xdp - xd;
    const bool sf = (front < R(0));
    const R tf = (R)(signfront * xfr - dt * ufront);
    const R FV0 = (sf ? p0p : p0) + order2 * (R(0.5) * (sf ? g0p : g0) * tf);
    const R FV1 = (sf ? p1p : p1) + order2 * (R(0.5) * (sf ? g1p : g1) * tf);
    const R uback = R(0.5) * (p1m + p1);
    const int signback = (uback > R(0)) ? 1 : -1;
    const double xbk = xd - xdm;
    const bool sb = (back > R(0));
    const R tb = (R)(signback * xbk - dt * uback);
    const R BV0 = (sb ? p0m : p0) + order2 * (R(0.5) * (sb ? g0m : g0) * tb);
    const R BV1 = (sb ? p1m : p1) + order2 * (R(0.5) * (sb ? g1m : g1) * tb);
    d0 = back * BV0 - front * FV0;
    d1 = back * BV1 - front * FV1;
  }
  __syncthreads();  // transposed reads done before the tiles are reused
  lds[0][lane][w] = d0;
  lds[1][lane][w] = d1;
  lds[2][lane][w] = d3;
  __syncthreads();
  // ---- row-major: U += increments, store
  if (!gout) return;
  if (gs >= 1 && gs <= N - 1) {
    U3 += lds[2][tr][tc];
    U0 += lds[0][tr][tc];
    U1 += lds[1][tr][tc];
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    if (fabs(U0) < thresold) U0 = R(0);
    if (fabs(U1) < thresold) U1 = R(0);
  }
  if (FIRST) {
    st_stream(&A.Ud0[gq], U0);
    st_stream(&A.Ud1[gq], U1);
    st_stream(&A.Ud3[gq], U3);
  } else if (U3 != R(0)) {
    st_stream(&A.uox[gq], qdiv(U0, U3));
    st_stream(&A.uoy[gq], qdiv(U1, U3));
  } else {
    st_stream(&A.uox[gq], A.ux[gq]);
    st_stream(&A.uoy[gq], A.uy[gq]);
  }
}

// Dual sweep along y: each lane owns one node column and walks a strip of
// STRIP node rows (two halo rows above and below).  Row r's phi_dual,
// coordinate and front flux enter a four-row register window, its gradient
// g(r-1) is formed once, and node r-2 is finished.  front(j) sums the cells
// (i-1, j), (i, j) in the order back(j+1) does.
template <bool FIRST, int NEM, int LIM, int STRIP>
__global__ __launch_bounds__(64) void k_node_sweep_y_2d(Params P, DevScalars* S, NodeSweepArgs A) {
  if (S->stop || S->err) return;
  const int nnx = P.nx + 1;
  const int ncb = (nnx + 63) / 64;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int i = (L % ncb) * 64 + threadIdx.x;
  if (i > P.nx) return;
  const int N = P.ny;
  const int j0 = (L / ncb) * STRIP;
  const int j1 = min(j0 + STRIP, N + 1);  // output node rows [j0, j1)
  const bool inner_t = (i >= 1 && i <= P.nx - 1);
  const long NC = P.nc;
  const int ne = P.ne;
  const double dt = S->dt;
  const bool o2 = P.order2 > 0;
  const double order2 = (double)P.order2;
  // windows: [0] row r-3, [1] r-2, [2] r-1, [3] r
  double xw[4] = {0., 0., 0., 0.}, p0w[4] = {0., 0., 0., 0.}, p1w[4] = {0., 0., 0., 0.};
  double fw[4] = {0., 0., 0., 0.};
  double U0w[3] = {0., 0., 0.}, U1w[3] = {0., 0., 0.}, U3w[3] = {0., 0., 0.};  // rows r-2, r-1, r
  double g0w[3] = {0., 0., 0.}, g1w[3] = {0., 0., 0.};                         // g(r-3), g(r-2), g(r-1)
  for (int r = j0 - 2; r <= j1 + 1; ++r) {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      xw[t] = xw[t + 1];
      p0w[t] = p0w[t + 1];
      p1w[t] = p1w[t + 1];
      fw[t] = fw[t + 1];
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      U0w[t] = U0w[t + 1];
      U1w[t] = U1w[t + 1];
      U3w[t] = U3w[t + 1];
      g0w[t] = g0w[t + 1];
      g1w[t] = g1w[t + 1];
    }
    double xd = 0., p0 = 0., p1 = 0., U0 = 0., U1 = 0., U3 = 0., f = 0.;
    if (r >= 0 && r <= N) {
      const long q = i + (long)nnx * r;
      xd = A.xny[q];
      if (FIRST) {
        p0 = A.ux[q];
        p1 = A.uy[q];
      } else {
        U0 = A.Ud0[q];
        U1 = A.Ud1[q];
        U3 = A.Ud3[q];
        if (inner_t) {
          p0 = qdiv(U0, U3);
          p1 = qdiv(U1, U3);
        } else {
          p0 = A.ux[q];
          p1 = A.uy[q];
        }
      }
      if (r < N) {
#pragma unroll
        for (int sh = -1; sh <= 0; ++sh) {
          const int ci = i + sh;
          if (ci < 0 || ci >= P.nx) continue;
          const long cc = ci + (long)P.nx * r;
#pragma unroll
          for (int k = 0; k < NEM; ++k) {
            if (k >= ne) break;
            f += A.contrib[k * NC + cc];
          }
        }
      }
    }
    xw[3] = xd;
    p0w[3] = p0;
    p1w[3] = p1;
    fw[3] = f;
    U0w[2] = U0;
    U1w[2] = U1;
    U3w[2] = U3;
    // gradient of node row r-1
    const int rg = r - 1;
    double g0 = 0., g1 = 0.;
    if (o2 && rg >= j0 - 1 && rg >= 0 && rg <= N) {
      if (rg >= 1 && rg <= N - 1) {
        const double dxf = xw[3] - xw[2], dxb = xw[2] - xw[1];
        const double gf0 = (p0w[3] - p0w[2]) / dxf, gb0 = (p0w[2] - p0w[1]) / dxb;
        const double gf1 = (p1w[3] - p1w[2]) / dxf, gb1 = (p1w[2] - p1w[1]) / dxb;
        g0 = limited_grad<LIM, false>(gb0, gf0);
        g1 = limited_grad<LIM, false>(gb1, gf1);
        if (!inner_t && rg >= j0 && rg < j1) {
          const long q = i + (long)nnx * rg;
          A.dg0[q] = g0;
          A.dg1[q] = g1;
        }
      } else {
        const long q = i + (long)nnx * rg;
        g0 = A.dg0[q];
        g1 = A.dg1[q];
      }
    }
    g0w[2] = g0;
    g1w[2] = g1;
    // finish node row o = r - 2
    const int o = r - 2;
    if (o < j0 || o >= j1) continue;
    const long nid = i + (long)nnx * o;
    double V0, V1, V3;
    if (FIRST) {
      const double mn = node_mass_from_cells(P, A.m, i, o);
      V0 = mn * p0w[1];
      V1 = mn * p1w[1];
      V3 = mn;
    } else {
      V0 = U0w[0];
      V1 = U1w[0];
      V3 = U3w[0];
    }
    if (o >= 1 && o <= N - 1) {
      const double back = fw[0], front = fw[1];
      V3 += back - front;
      const double ufront = 0.5 * (p1w[1] + p1w[1]);  // DualPhaseRemap.cc:391
      const int signfront = (ufront > 0) ? 1 : -1;
      const double xfr = xw[2] - xw[1];
      const bool sf = (front < 0.);
      const double tf = signfront * xfr - dt * ufront;
      const double FV0 = (sf ? p0w[2] : p0w[1]) + order2 * (0.5 * (sf ? g0w[2] : g0w[1]) * tf);
      const double FV1 = (sf ? p1w[2] : p1w[1]) + order2 * (0.5 * (sf ? g1w[2] : g1w[1]) * tf);
      const double uback = 0.5 * (p1w[0] + p1w[1]);
      const int signback = (uback > 0) ? 1 : -1;
      const double xbk = xw[1] - xw[0];
      const bool sb = (back > 0.);
      const double tb = signback * xbk - dt * uback;
      const double BV0 = (sb ? p0w[0] : p0w[1]) + order2 * (0.5 * (sb ? g0w[0] : g0w[1]) * tb);
      const double BV1 = (sb ? p1w[0] : p1w[1]) + order2 * (0.5 * (sb ? g1w[0] : g1w[1]) * tb);
      V0 += back * BV0 - front * FV0;
      V1 += back * BV1 - front * FV1;
      const double thresold = 1.e-300;
      if (fabs(V0) < thresold) V0 = 0.;
      if (fabs(V1) < thresold) V1 = 0.;
    }
    if (FIRST) {
      st_stream(&A.Ud0[nid], V0);
      st_stream(&A.Ud1[nid], V1);
      st_stream(&A.Ud3[nid], V3);
    } else if (V3 != 0.) {
      st_stream(&A.uox[nid], V0 / V3);
      st_stream(&A.uoy[nid], V1 / V3);
    } else {
      st_stream(&A.uox[nid], A.ux[nid]);
      st_stream(&A.uoy[nid], A.uy[nid]);
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

#define MHG_NEM_DISPATCH(ne, F) \
  do {                          \
    if ((ne) <= 1) F(1);        \
    else if ((ne) <= 2) F(2);   \
    else if ((ne) <= 3) F(3);   \
    else if ((ne) <= 4) F(4);   \
    else F(8);                  \
  } while (0)

void launch_begin_iteration(hipStream_t st, DevScalars* S) {
  hipLaunchKernelGGL(k_begin_iteration, dim3(1), dim3(1), 0, st, S);
}

template <class R>
void launch_node_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeLagArgsT<R>& A,
                             bool vel, bool bc, bool pos) {
  const int B = 256;
  dim3 g(nblocks(P.nn, B)), b(B);
  if (vel && bc && pos) hipLaunchKernelGGL((k_node_lagrange_2d<R, true, true, true>), g, b, 0, st, P, S, A);
  else if (vel && !bc && !pos) hipLaunchKernelGGL((k_node_lagrange_2d<R, true, false, false>), g, b, 0, st, P, S, A);
  else if (!vel && bc && !pos) hipLaunchKernelGGL((k_node_lagrange_2d<R, false, true, false>), g, b, 0, st, P, S, A);
  else if (!vel && !bc && pos) hipLaunchKernelGGL((k_node_lagrange_2d<R, false, false, true>), g, b, 0, st, P, S, A);
}

template <class R>
void launch_cell_lagrange_2d(hipStream_t st, const Params& P, DevScalars* S, const CellLagArgsT<R>& A) {
  const int B = 256;
  dim3 g(nblocks(P.nc, B)), b(B);
#define L_(N)                                                                                     \
  if (P.pseudo_centree) hipLaunchKernelGGL((k_cell_lagrange_2d<R, true, N>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_lagrange_2d<R, false, N>), g, b, 0, st, P, S, A)
  MHG_NEM_DISPATCH(P.ne, L_);
#undef L_
}

// Launch variants of the y sweep (strip length, variables per thread, rows
// per load chunk of the column walker; strip 0 = LDS tiles), chosen per mesh
// by the autotuner (mhg_autotune).
struct YVariant {
  int strip, vg, r;
};
static const YVariant kYVariants[] = {{32, 4, 2}, {8, 4, 2}, {16, 4, 4}, {8, 5, 2}, {0, 0, 0}};
int cell_sweep_y_variants() { return (int)(sizeof(kYVariants) / sizeof(kYVariants[0])); }
const char* cell_sweep_y_variant_name(int v) {
  static char buf[16][32];
  const YVariant& y = kYVariants[v];
  if (y.strip == 0) snprintf(buf[v], 32, "tile16x64");
  else snprintf(buf[v], 32, "strip%d_vg%d_r%d", y.strip, y.vg, y.r);
  return buf[v];
}

#define MHG_LIM_DISPATCH(lim, F) \
  do {                           \
    if ((lim) == 0) F(0);        \
    else if ((lim) == 1) F(1);   \
    else F(2);                   \
  } while (0)

template <int STRIP, int VG, int RR>
static void launch_y(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgs& A, bool first) {
  const long ntiles = (long)nblocks(P.nx, 64) * ((P.ny + STRIP - 1) / STRIP) * ((P.nvu + VG - 1) / VG);
  dim3 g((unsigned)ntiles), b(64);
#define YL_(LM)                                                                                    \
  if (first) hipLaunchKernelGGL((k_cell_sweep_y_2d<true, STRIP, VG, RR, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_sweep_y_2d<false, STRIP, VG, RR, LM>), g, b, 0, st, P, S, A)
  MHG_LIM_DISPATCH(P.lim, YL_);
#undef YL_
}

template <class R, int NEM>
static void launch_lane_x(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                          bool first) {
  const long waves = (long)((P.nx + 59) / 60) * P.ny;
  dim3 g(nblocks(waves * 64, 256)), b(256);
#define XL_(LM)                                                                                    \
  if (first) hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, true, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, false, LM>), g, b, 0, st, P, S, A)
  MHG_LIM_DISPATCH(P.lim, XL_);
#undef XL_
}

template <class R, int NEM>
static void launch_tile_y(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                          bool first) {
  dim3 g((unsigned)(((P.nx + kTileW - 1) / kTileW) * ((P.ny + 59) / 60))), b(1024);
#define YT_(LM)                                                                                  \
  if (first) hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, true, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, false, LM>), g, b, 0, st, P, S, A)
  MHG_LIM_DISPATCH(P.lim, YT_);
#undef YT_
}

template <class R>
void launch_cell_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                          int dir, bool first, int variant) {
  if (dir == 0) {
#define X_(N) launch_lane_x<R, N>(st, P, S, A, first)
    MHG_NEM_DISPATCH(P.ne, X_);
#undef X_
    return;
  }
  if constexpr (std::is_same<R, double>::value) {
    switch (variant) {
      case 0: launch_y<32, 4, 2>(st, P, S, A, first); return;
      case 1: launch_y<8, 4, 2>(st, P, S, A, first); return;
      case 2: launch_y<16, 4, 4>(st, P, S, A, first); return;
      case 3: launch_y<8, 5, 2>(st, P, S, A, first); return;
      default: break;
    }
  }
#define T_(N) launch_tile_y<R, N>(st, P, S, A, first)
  MHG_NEM_DISPATCH(P.ne, T_);
#undef T_
}

// Launch variants of the dual sweeps: 0 = one thread per node re-forming the
// 5-node stencil; x: 1 = lane exchange; y: 1, 2, 3 = column walkers of 8, 16
// and 32 rows, 4 = LDS tiles.
int node_sweep_variants(int dir) { return dir == 0 ? 2 : 5; }
const char* node_sweep_variant_name(int dir, int v) {
  static const char* xn[] = {"stencil5", "lanes64"};
  static const char* yn[] = {"stencil5", "walk8", "walk16", "walk32", "tile16x64"};
  if (v < 0 || v >= node_sweep_variants(dir)) return "?";
  return dir == 0 ? xn[v] : yn[v];
}

template <int STRIP, int NEM, int LM>
static void launch_node_y_walk(hipStream_t st, const Params& P, DevScalars* S, const NodeSweepArgs& A,
                               bool first) {
  const int ncb = (P.nx + 1 + 63) / 64;
  const int nst = (P.ny + 1 + STRIP - 1) / STRIP;
  dim3 g((unsigned)(ncb * nst)), b(64);
  if (first) hipLaunchKernelGGL((k_node_sweep_y_2d<true, NEM, LM, STRIP>), g, b, 0, st, P, S, A);
  else hipLaunchKernelGGL((k_node_sweep_y_2d<false, NEM, LM, STRIP>), g, b, 0, st, P, S, A);
}

static void launch_node_fp64_variant(hipStream_t st, const Params& P, DevScalars* S, const NodeSweepArgs& A,
                                     int dir, bool first, int variant) {
  if (variant == 0) {
    const int B = 256;
    dim3 g(nblocks(P.nn, B)), b(B);
#define NL_(LM)                                                                                       \
  if (dir == 0 && first) hipLaunchKernelGGL((k_node_sweep_2d<0, true, LM>), g, b, 0, st, P, S, A);      \
  else if (dir == 0) hipLaunchKernelGGL((k_node_sweep_2d<0, false, LM>), g, b, 0, st, P, S, A);        \
  else if (first) hipLaunchKernelGGL((k_node_sweep_2d<1, true, LM>), g, b, 0, st, P, S, A);            \
  else hipLaunchKernelGGL((k_node_sweep_2d<1, false, LM>), g, b, 0, st, P, S, A)
    MHG_LIM_DISPATCH(P.lim, NL_);
#undef NL_
    return;
  }
#define NY_(STR, N, LM) launch_node_y_walk<STR, N, LM>(st, P, S, A, first)
#define NYN_(STR, LM)                  \
  do {                                 \
    if (P.ne <= 1) NY_(STR, 1, LM);      \
    else if (P.ne <= 2) NY_(STR, 2, LM); \
    else if (P.ne <= 3) NY_(STR, 3, LM); \
    else if (P.ne <= 4) NY_(STR, 4, LM); \
    else NY_(STR, 8, LM);                \
  } while (0)
#define NYL_(LM)                              \
  do {                                        \
    if (variant == 1) NYN_(8, LM);            \
    else if (variant == 2) NYN_(16, LM);      \
    else NYN_(32, LM);                        \
  } while (0)
  MHG_LIM_DISPATCH(P.lim, NYL_);
#undef NYL_
#undef NYN_
#undef NY_
}

template <class R>
void launch_node_sweep_2d(hipStream_t st, const Params& P, DevScalars* S, const NodeSweepArgsT<R>& A,
                          int dir, bool first, int variant) {
  if constexpr (std::is_same<R, double>::value) {
    // the 5-node stencil (both directions) and the y column walkers
    if (variant <= 0 || variant >= node_sweep_variants(dir) || (dir == 1 && variant <= 3)) {
      launch_node_fp64_variant(st, P, S, A, dir, first, variant <= 0 || variant >= node_sweep_variants(dir) ? 0 : variant);
      return;
    }
  }
  if (dir == 0) {
    const long waves = (long)((P.nx + 1 + 59) / 60) * (P.ny + 1);
    dim3 g(nblocks(waves * 64, 256)), b(256);
#define NX_(N, LM)                                                                               \
  if (first) hipLaunchKernelGGL((k_node_sweep_x_2d<R, true, N, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_node_sweep_x_2d<R, false, N, LM>), g, b, 0, st, P, S, A)
#define NXL_(LM)                    \
  do {                              \
    if (P.ne <= 1) NX_(1, LM);      \
    else if (P.ne <= 2) NX_(2, LM); \
    else if (P.ne <= 3) NX_(3, LM); \
    else if (P.ne <= 4) NX_(4, LM); \
    else NX_(8, LM);                \
  } while (0)
    MHG_LIM_DISPATCH(P.lim, NXL_);
#undef NXL_
#undef NX_
    return;
  }
  dim3 g((unsigned)(((P.nx + 1 + kTileW - 1) / kTileW) * ((P.ny + 1 + 59) / 60))), b(1024);
#define NT_(N, LM)                                                                                 \
  if (first) hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, true, N, LM>), g, b, 0, st, P, S, A);    \
  else hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, false, N, LM>), g, b, 0, st, P, S, A)
#define NTL_(LM)                    \
  do {                              \
    if (P.ne <= 1) NT_(1, LM);      \
    else if (P.ne <= 2) NT_(2, LM); \
    else if (P.ne <= 3) NT_(3, LM); \
    else if (P.ne <= 4) NT_(4, LM); \
    else NT_(8, LM);                \
  } while (0)
  MHG_LIM_DISPATCH(P.lim, NTL_);
#undef NTL_
#undef NT_
}

int cell_final_blocks(const Params& P) { return (int)nblocks(P.nc, 256); }

void launch_cell_stage_2d(hipStream_t st, const Params& P, DevScalars* S, const StageArgs& A, int stage) {
  dim3 g(nblocks(P.nc, 256)), b(256);
#define CS_(STG, N)                                                                                   \
  if (P.pseudo_centree) hipLaunchKernelGGL((k_cell_stage_2d<STG, true, N>), g, b, 0, st, P, S, A);   \
  else hipLaunchKernelGGL((k_cell_stage_2d<STG, false, N>), g, b, 0, st, P, S, A)
#define CSN_(STG)                      \
  do {                                 \
    if (P.ne <= 1) CS_(STG, 1);        \
    else if (P.ne <= 2) CS_(STG, 2);   \
    else if (P.ne <= 3) CS_(STG, 3);   \
    else if (P.ne <= 4) CS_(STG, 4);   \
    else CS_(STG, 8);                  \
  } while (0)
  switch (stage) {
    case STAGE_SAVE: CSN_(STAGE_SAVE); break;
    case STAGE_GEOMETRY: CSN_(STAGE_GEOMETRY); break;
    case STAGE_DENSITY: CSN_(STAGE_DENSITY); break;
    case STAGE_VISCOSITY: CSN_(STAGE_VISCOSITY); break;
    case STAGE_ENERGY: CSN_(STAGE_ENERGY); break;
    default: CSN_(STAGE_BUILD_U); break;
  }
#undef CSN_
#undef CS_
}

void launch_dt_finalize(hipStream_t st, const Params& P, DevScalars* S, const double* block_min, bool advance) {
  const int nb = cell_final_blocks(P);
  if (advance) hipLaunchKernelGGL((k_dt_finalize<1024, true>), dim3(1), dim3(1024), 0, st, P, S, block_min, nb);
  else hipLaunchKernelGGL((k_dt_finalize<1024, false>), dim3(1), dim3(1024), 0, st, P, S, block_min, nb);
}

template <class R>
void launch_cell_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellFinalArgsT<R>& A,
                          int dt_mode) {
  constexpr int B = 256;
  dim3 g(nblocks(P.nc, B)), b(B);
#define F_(N) hipLaunchKernelGGL((k_cell_final_2d<R, B, N>), g, b, 0, st, P, S, A)
  MHG_NEM_DISPATCH(P.ne, F_);
#undef F_
  if (dt_mode > 0) launch_dt_finalize(st, P, S, A.block_min, dt_mode == 2);
}

// The fused step for both storage types.
#define MHG_INSTANTIATE(R)                                                                           \
  template void launch_node_lagrange_2d<R>(hipStream_t, const Params&, DevScalars*,                  \
                                           const NodeLagArgsT<R>&, bool, bool, bool);                \
  template void launch_cell_lagrange_2d<R>(hipStream_t, const Params&, DevScalars*,                  \
                                           const CellLagArgsT<R>&);                                  \
  template void launch_cell_sweep_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const CellSweepArgsT<R>&, int, bool, int);                   \
  template void launch_node_sweep_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const NodeSweepArgsT<R>&, int, bool, int);                   \
  template void launch_cell_final_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const CellFinalArgsT<R>&, int);
MHG_INSTANTIATE(double)
MHG_INSTANTIATE(float)
#undef MHG_INSTANTIATE

}  // namespace mhg
