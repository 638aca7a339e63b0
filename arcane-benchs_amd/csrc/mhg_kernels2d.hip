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
#include "mhg_body2d.hpp"

namespace mhg {

// ---------------------------------------------------------------------------
// Start of iteration: time += dt, iteration += 1 (Arcane time-loop manager).
// ---------------------------------------------------------------------------
__global__ void k_begin_iteration(DevScalars* S) {
  if (S->stop || S->err) return;
  S->time += S->dt;
  S->iteration += 1;
}

// ---------------------------------------------------------------------------
// Single-mesh step kernels: one block of the grid is one block of the body.
// The sweeps, the node phase and the final cell phase map the block index so
// that each XCD runs a contiguous range of body blocks (the rows a block
// shares with its neighbours then stay in that XCD's L2); the cell Lagrange
// phase measured no gain from it and keeps the hardware order.
// ---------------------------------------------------------------------------
template <class R, bool VEL, bool BC, bool POS>
__global__ void k_node_lagrange_2d(Params P, DevScalars* S, NodeLagArgsT<R> A) {
  node_lagrange_2d_body<R, VEL, BC, POS>(P, S, A, xcd_block(blockIdx.x, gridDim.x));
}

// The node phase of a fused step with the previous step's dt kernel folded
// in (node_lagrange_2d_body's FOLD): Sin is only read, so a block that starts
// late sees what the first saw; block 0 writes the updated scalars to
// F.Sout, which the rest of the step reads.  A run that has ended copies Sin
// to F.Sout unchanged.
template <class R, bool VEL, bool BC, bool POS>
__global__ __launch_bounds__(256) void k_node_lagrange_fold_2d(Params P, DevScalars* Sin, DtFold F,
                                                               NodeLagArgsT<R> A) {
  node_lagrange_2d_body<R, VEL, BC, POS, true>(P, Sin, A, xcd_block(blockIdx.x, gridDim.x), &F);
}

template <class R, bool PC, int NEM>
__global__ __launch_bounds__(256) void k_cell_lagrange_2d(Params P, DevScalars* S, CellLagArgsT<R> A) {
  cell_lagrange_2d_body<R, PC, NEM>(P, S, A, blockIdx.x);
}

template <class R, int D, int NEM, bool FIRST, int LIM, bool LEARN = false>
__global__ __launch_bounds__(D == 0 ? 256 : 1024) void k_cell_sweep_lane_2d(Params P, DevScalars* S,
                                                                          CellSweepArgsT<R> A) {
  cell_sweep_lane_2d_body<R, D, NEM, FIRST, LIM, LEARN>(P, S, A, xcd_block(blockIdx.x, gridDim.x));
}

template <class R, int NEM, bool FIRST, int LIM, bool LEARN = false>
__global__ __launch_bounds__(1024) void k_cell_sweep_ytile_2d(Params P, DevScalars* S, CellSweepArgsT<R> A) {
  cell_sweep_ytile_2d_body<R, NEM, FIRST, LIM, false, LEARN>(P, S, A, xcd_block(blockIdx.x, gridDim.x));
}

// The last cell sweep of the remap fused with the final cell phase.
template <class R, int NEM, int LIM>
__global__ __launch_bounds__(256) void k_cell_sweep_x_final_2d(Params P, DevScalars* S, CellSweepArgsT<R> A,
                                                               CellFinalArgsT<R> F) {
  cell_sweep_lane_2d_body<R, 0, NEM, false, LIM, false, true>(P, S, A, xcd_block(blockIdx.x, gridDim.x), &F);
}

// The first cell sweep of the remap fused with the Lagrange cell phase.
template <class R, int NEM, int LIM>
__global__ __launch_bounds__(256) void k_cell_sweep_x_lag_2d(Params P, DevScalars* S, CellSweepArgsT<R> A,
                                                             CellLagArgsT<R> LA) {
  cell_sweep_lane_2d_body<R, 0, NEM, true, LIM, false, false, true>(P, S, A, xcd_block(blockIdx.x, gridDim.x),
                                                                     nullptr, &LA);
}

template <class R, int NEM, int LIM, int TWC = kTileW>
__global__ __launch_bounds__(64 * TWC) void k_cell_sweep_y_lag_2d(Params P, DevScalars* S, CellSweepArgsT<R> A,
                                                                     CellLagArgsT<R> LA) {
  cell_sweep_ytile_2d_body<R, NEM, true, LIM, false, false, false, true, TWC>(
      P, S, A, xcd_block(blockIdx.x, gridDim.x), nullptr, &LA);
}

// A bound on waves per SIMD keeps the scalar registers of both argument
// blocks within budget (the unbounded kernel took 102 and ran one block per
// CU): 8 in float; 7 in FP64, whose 8-wave budget of 64 VGPRs spilled 12
// bytes per lane (0.2137 -> 0.2128-0.2134 ms per triple-point step without;
// float at 7 ran 0.1103 -> 0.1114).
template <class R, int NEM, int LIM, int TWC = kTileW>
__global__ __launch_bounds__(64 * TWC, sizeof(R) == 8 ? 7 : 8) void k_cell_sweep_y_final_2d(Params P, DevScalars* S,
                                                                       CellSweepArgsT<R> A, CellFinalArgsT<R> F) {
  cell_sweep_ytile_2d_body<R, NEM, false, LIM, false, false, true, false, TWC>(
      P, S, A, xcd_block(blockIdx.x, gridDim.x), &F);
}

template <class R, int BLOCK, int NEM, bool EOS = true>
__global__ __launch_bounds__(BLOCK) void k_cell_final_2d(Params P, DevScalars* S, CellFinalArgsT<R> A) {
  cell_final_2d_body<R, BLOCK, NEM, EOS>(P, S, A, xcd_block(blockIdx.x, gridDim.x), A.block_min);
}

template <int BLOCK, bool ADVANCE>
__global__ __launch_bounds__(BLOCK) void k_dt_finalize(Params P, DevScalars* S, const double* block_min, int nb) {
  dt_finalize_body<BLOCK, ADVANCE>(P, S, block_min, nb);
}

template <class R, bool FIRST, int NEM, int LIM>
__global__ __launch_bounds__(256) void k_node_sweep_x_2d(Params P, DevScalars* S, NodeSweepArgsT<R> A) {
  node_sweep_x_2d_body<R, FIRST, NEM, LIM>(P, S, A, xcd_block(blockIdx.x, gridDim.x));
}

template <class R, bool FIRST, int NEM, int LIM, int TWC = kTileW>
__global__ __launch_bounds__(64 * TWC) void k_node_sweep_ytile_2d(Params P, DevScalars* S, NodeSweepArgsT<R> A) {
  node_sweep_ytile_2d_body<R, FIRST, NEM, LIM, false, TWC>(P, S, A, xcd_block(blockIdx.x, gridDim.x));
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
  const int ne = env_count<NEM>(P);
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
    int i, j;
    divmod32(c, P.nx, j, i);
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
    A.L[c] = (P.lc == MHG_LC_FACES_OPPOSEES) ? length_faces_opposees_2d(x, y) : length_sqrt_2d(V);
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
    if (P.pb) {
      // computeVariablesForRemap with pente-borne (PrepareRemap.cc:159-215):
      // fraction, density, energy of each environment present, the last
      // one's viscosity; the slots of absent environments are left as they are
      // (a pure cell's fraction is kept in its environment slot)
      if (nbe == 1) {
        A.phi[kk * NC + c] = A.env_f[kk * NC + c];
        A.phi[(ne + kk) * NC + c] = A.rho[c];
        A.phi[(2 * ne + kk) * NC + c] = A.e[c];
        A.phi[(3 * ne) * NC + c] = A.q[c];
      } else if (nbe >= 2) {
#pragma unroll
        for (int k = 0; k < NEM; ++k) {
          if (!((mask >> k) & 1)) continue;
          const long ek = k * NC + c;
          A.phi[k * NC + c] = A.env_f[ek];
          A.phi[(ne + k) * NC + c] = A.env_rho[ek];
          A.phi[(2 * ne + k) * NC + c] = A.env_e[ek];
          A.phi[(3 * ne) * NC + c] = A.env_q[ek];
        }
      }
    }
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
          st_stream<1>(P, &Uout[v * NC + co], un);
          if (v >= ne && v < 2 * ne) {
            const double dual = 0.5 * flux_b * -1.0 + 0.5 * flux_f * 1.0;
            st_stream<2>(P, &A.contrib[(v - ne) * NC + co], 0.5 * dual);
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



template <int D, bool FIRST, int LIM>
__global__ void k_node_sweep_2d(Params P, DevScalars* S, NodeSweepArgs A) {
  if (S->stop || S->err) return;
  const long nid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nid >= P.nn) return;
  const int nnx = P.nx + 1;
  int i, j;
  divmod32(nid, nnx, j, i);
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
    st_stream<2>(P, &A.Ud0[nid], U0);
    st_stream<2>(P, &A.Ud1[nid], U1);
    st_stream<2>(P, &A.Ud3[nid], U3);
  } else {
    if (U3 != 0.) {
      st_stream<3>(P, &A.uox[nid], U0 / U3);
      st_stream<3>(P, &A.uoy[nid], U1 / U3);
    } else {
      st_stream<3>(P, &A.uox[nid], A.ux[nid]);
      st_stream<3>(P, &A.uoy[nid], A.uy[nid]);
    }
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
  const int ne = env_count<NEM>(P);
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
      st_stream<2>(P, &A.Ud0[nid], V0);
      st_stream<2>(P, &A.Ud1[nid], V1);
      st_stream<2>(P, &A.Ud3[nid], V3);
    } else if (V3 != 0.) {
      st_stream<3>(P, &A.uox[nid], V0 / V3);
      st_stream<3>(P, &A.uoy[nid], V1 / V3);
    } else {
      st_stream<3>(P, &A.uox[nid], A.ux[nid]);
      st_stream<3>(P, &A.uoy[nid], A.uy[nid]);
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
  else if (vel && bc && !pos) hipLaunchKernelGGL((k_node_lagrange_2d<R, true, true, false>), g, b, 0, st, P, S, A);
}

template <class R>
void launch_node_lagrange_fold_2d(hipStream_t st, const Params& P, const DevScalars* Sin, DevScalars* Sout,
                                  const NodeLagArgsT<R>& A, bool pos, const double* block_min, int nb) {
  const int B = 256;
  dim3 g(nblocks(P.nn, B)), b(B);
  const DtFold F{block_min, nb, Sout};
  DevScalars* in = const_cast<DevScalars*>(Sin);
  if (pos) hipLaunchKernelGGL((k_node_lagrange_fold_2d<R, true, true, true>), g, b, 0, st, P, in, F, A);
  else hipLaunchKernelGGL((k_node_lagrange_fold_2d<R, true, true, false>), g, b, 0, st, P, in, F, A);
}

// The last Lagrange phase's node coordinates x0 + dt_old u (lag_coord) into
// the coordinate arrays, after fused steps that formed them without storing
// them.  It runs whatever the run's state (an ended run keeps its arrays).
template <class R>
__global__ void k_lag_coords_2d(Params P, DevScalars* S, const R* ux, const R* uy, double* xx, double* xy) {
  const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= P.nn) return;
  const int nnx = P.nx + 1;
  int i, j;
  divmod32(n, nnx, j, i);
  const double dt = S->dt_old;
  xx[n] = lag_coord<0>(P, i, dt, ux[n]);
  xy[n] = lag_coord<1>(P, j, dt, uy[n]);
}

template <class R>
void launch_lag_coords_2d(hipStream_t st, const Params& P, DevScalars* S, const R* ux, const R* uy, double* xx,
                          double* xy) {
  const int B = 256;
  hipLaunchKernelGGL(k_lag_coords_2d<R>, dim3(nblocks(P.nn, B)), dim3(B), 0, st, P, S, ux, uy, xx, xy);
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
  const int no = P.order2 == 2 ? 58 : 60;  // output lanes per wavefront (order 3: three halo lanes a side)
  const long waves = (long)((P.nx + no - 1) / no) * P.ny;
  dim3 g(nblocks(waves * 64, 256)), b(256);
  if (P.learned) {
    if (first) hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, true, 0, true>), g, b, 0, st, P, S, A);
    else hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, false, 0, true>), g, b, 0, st, P, S, A);
    return;
  }
#define XL_(LM)                                                                                    \
  if (first) hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, true, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_sweep_lane_2d<R, 0, NEM, false, LM>), g, b, 0, st, P, S, A)
  if (P.order2 == 2) {
    XL_(kLimO3);
  } else if (P.lim >= kLimG) {
    XL_(kLimG);
  } else {
    MHG_LIM_DISPATCH(P.lim, XL_);
  }
#undef XL_
}

template <class R, int NEM>
static void launch_tile_y(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                          bool first) {
  const int no = P.order2 == 2 ? 58 : 60;  // output rows per tile (order 3: three halo rows a side)
  dim3 g((unsigned)(((P.nx + kTileW - 1) / kTileW) * ((P.ny + no - 1) / no))), b(1024);
  if (P.learned) {
    if (first) hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, true, 0, true>), g, b, 0, st, P, S, A);
    else hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, false, 0, true>), g, b, 0, st, P, S, A);
    return;
  }
#define YT_(LM)                                                                                  \
  if (first) hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, true, LM>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_sweep_ytile_2d<R, NEM, false, LM>), g, b, 0, st, P, S, A)
  if (P.order2 == 2) {
    YT_(kLimO3);
  } else if (P.lim >= kLimG) {
    YT_(kLimG);
  } else {
    MHG_LIM_DISPATCH(P.lim, YT_);
  }
#undef YT_
}

bool fused_final_supported(const Params& P) {
  return !P.learned && !P.etot && !P.lagonly && !P.generic && P.ne <= 4 && P.lim < kLimG && P.order2 != 2;
}

template <class R>
void launch_cell_sweep_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                                const CellFinalArgsT<R>& F, int dir) {
  if (dir == 0) {
    const long waves = (long)((P.nx + 59) / 60) * P.ny;
    dim3 g(nblocks(waves * 64, 256)), b(256);
#define XF_(N, LM) hipLaunchKernelGGL((k_cell_sweep_x_final_2d<R, N, LM>), g, b, 0, st, P, S, A, F)
#define XFN_(LM)                          \
  do {                                    \
    if (P.ne <= 1) XF_(1, LM);            \
    else if (P.ne <= 2) XF_(2, LM);       \
    else if (P.ne <= 3) XF_(3, LM);       \
    else XF_(4, LM);                      \
  } while (0)
    MHG_LIM_DISPATCH(P.lim, XFN_);
#undef XFN_
#undef XF_
    return;
  }
  const int tw = (P.ytw8 & kYtwFinal) ? 8 : kTileW;
  dim3 g((unsigned)(((P.nx + tw - 1) / tw) * ((P.ny + 59) / 60))), b(64 * tw);
#define YF_(N, LM)                                                                                     \
  if (tw == 8) hipLaunchKernelGGL((k_cell_sweep_y_final_2d<R, N, LM, 8>), g, b, 0, st, P, S, A, F); \
  else hipLaunchKernelGGL((k_cell_sweep_y_final_2d<R, N, LM>), g, b, 0, st, P, S, A, F)
#define YFN_(LM)                          \
  do {                                    \
    if (P.ne <= 1) YF_(1, LM);            \
    else if (P.ne <= 2) YF_(2, LM);       \
    else if (P.ne <= 3) YF_(3, LM);       \
    else YF_(4, LM);                      \
  } while (0)
  MHG_LIM_DISPATCH(P.lim, YFN_);
#undef YFN_
#undef YF_
}

bool fused_lag_supported(const Params& P) { return fused_final_supported(P) && !P.pseudo_centree; }

template <class R>
void launch_cell_sweep_lag_2d(hipStream_t st, const Params& P, DevScalars* S, const CellSweepArgsT<R>& A,
                              const CellLagArgsT<R>& LA, int dir) {
  if (dir == 0) {
    const long waves = (long)((P.nx + 59) / 60) * P.ny;
    dim3 g(nblocks(waves * 64, 256)), b(256);
#define XG_(N, LM) hipLaunchKernelGGL((k_cell_sweep_x_lag_2d<R, N, LM>), g, b, 0, st, P, S, A, LA)
#define XGN_(LM)                          \
  do {                                    \
    if (P.ne <= 1) XG_(1, LM);            \
    else if (P.ne <= 2) XG_(2, LM);       \
    else if (P.ne <= 3) XG_(3, LM);       \
    else XG_(4, LM);                      \
  } while (0)
    MHG_LIM_DISPATCH(P.lim, XGN_);
#undef XGN_
#undef XG_
    return;
  }
  const int tw = (P.ytw8 & kYtwLag) ? 8 : kTileW;
  dim3 g((unsigned)(((P.nx + tw - 1) / tw) * ((P.ny + 59) / 60))), b(64 * tw);
#define YG_(N, LM)                                                                                   \
  if (tw == 8) hipLaunchKernelGGL((k_cell_sweep_y_lag_2d<R, N, LM, 8>), g, b, 0, st, P, S, A, LA); \
  else hipLaunchKernelGGL((k_cell_sweep_y_lag_2d<R, N, LM>), g, b, 0, st, P, S, A, LA)
#define YGN_(LM)                          \
  do {                                    \
    if (P.ne <= 1) YG_(1, LM);            \
    else if (P.ne <= 2) YG_(2, LM);       \
    else if (P.ne <= 3) YG_(3, LM);       \
    else YG_(4, LM);                      \
  } while (0)
  MHG_LIM_DISPATCH(P.lim, YGN_);
#undef YGN_
#undef YG_
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
    // the column walkers have no learned or generalised-limiter form
    switch (P.learned || P.lim >= kLimG || P.order2 == 2 ? -1 : variant) {
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
    // the 5-node stencil (both directions) and the y column walkers (no
    // generalised-limiter or total-energy form)
    if (P.lim < kLimG && !P.etot &&
        (variant <= 0 || variant >= node_sweep_variants(dir) || (dir == 1 && variant <= 3))) {
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
    if (P.lim >= kLimG) {
      NXL_(kLimG);
    } else {
      MHG_LIM_DISPATCH(P.lim, NXL_);
    }
#undef NXL_
#undef NX_
    return;
  }
  const int tw = (P.ytw8 & kYtwNode) ? 8 : kTileW;
  dim3 g((unsigned)(((P.nx + 1 + tw - 1) / tw) * ((P.ny + 1 + 59) / 60))), b(64 * tw);
#define NT_(N, LM)                                                                                        \
  if (tw == 8 && first) hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, true, N, LM, 8>), g, b, 0, st, P, S, A); \
  else if (tw == 8) hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, false, N, LM, 8>), g, b, 0, st, P, S, A);   \
  else if (first) hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, true, N, LM>), g, b, 0, st, P, S, A);         \
  else hipLaunchKernelGGL((k_node_sweep_ytile_2d<R, false, N, LM>), g, b, 0, st, P, S, A)
#define NTL_(LM)                    \
  do {                              \
    if (P.ne <= 1) NT_(1, LM);      \
    else if (P.ne <= 2) NT_(2, LM); \
    else if (P.ne <= 3) NT_(3, LM); \
    else if (P.ne <= 4) NT_(4, LM); \
    else NT_(8, LM);                \
  } while (0)
  if (P.lim >= kLimG) {
    NTL_(kLimG);
  } else {
    MHG_LIM_DISPATCH(P.lim, NTL_);
  }
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

// ---------------------------------------------------------------------------
// Entry-point kernels of the EOS services and of computePressionMoyenne.
// ---------------------------------------------------------------------------
// applyEOS / initEOS of environment k (PerfectGasEOSService.cc:20-137,
// StiffenedGasEOSService.cc:10-59) on the cells [c0, c1) holding it.
template <int MODE>
__global__ __launch_bounds__(256) void k_eos_env(Params P, DevScalars* S, EosArgs A, int k, long c0, long c1) {
  if (S->stop || S->err) return;
  const long c = c0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  const int mask = A.mask[c];
  if (!((mask >> k) & 1)) return;
  const bool pure = __popc(mask) == 1;
  const long ek = (long)k * P.nc + c;
  double* rho = pure ? A.rho + c : A.env_rho + ek;
  double* e = pure ? A.e + c : A.env_e + ek;
  double* p = pure ? A.p + c : A.env_p + ek;
  double* cs = pure ? A.cs + c : A.env_cs + ek;
  if (MODE == EOS_APPLY) {
    double pv, cv;
    eos_apply(P.eos, k, *rho, *e, &pv, &cv);
    *p = pv;
    *cs = cv;
  } else {
    const double g = P.eos.gamma[k], r = *rho, pr = *p;
    if (P.eos.kind[k] == MHG_EOS_STIFFENED_GAS) {
      const double t = P.eos.tension[k];
      *e = (pr + (g * t)) / ((g - 1.) * r);
      *cs = sqrt((g / r) * (pr + t));
    } else {
      *e = pr / ((g - 1.) * r);
      *cs = sqrt(g * pr / r);
    }
  }
}

// computePressionMoyenne (MahycoModule.cc:1836-1915): a cell holding
// several environments takes p = sum_k f_k p_k and c = max(1e-20, max_k c_k)
// in ascending k; a cell holding none p = 0, c = 1e-20; pure cells keep
// their values.
template <int NEM>
__global__ __launch_bounds__(256) void k_pression_moyenne(Params P, DevScalars* S, PmArgs A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  const int mask = A.mask[c];
  if (__popc(mask) == 1) return;
  double p = 0., cs = 1.e-20;
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (!((mask >> k) & 1)) continue;
    const long ek = (long)k * P.nc + c;
    p += A.env_f[ek] * A.env_p[ek];
    cs = smax(A.env_cs[ek], cs);
  }
  A.p[c] = p;
  A.cs[c] = cs;
}

// CFL candidates of quads: the same arithmetic as k_cell_final_2d's, in the
// storage type R of the bulk fields.  A block of kCflThreads owns a
// contiguous run of cells (XCD-contiguous runs) and leaves one minimum: with
// at most kCflBlocks blocks the kernel that reduces them (the dt kernel, or
// every block of the next step's node phase) reads few minima.
constexpr int kCflBlocks = 512, kCflThreads = 1024;
int cfl_blocks(const Params& P) {
  const int nb = (int)nblocks(P.nc, kCflThreads);
  return nb < kCflBlocks ? nb : kCflBlocks;
}

template <class R>
__global__ __launch_bounds__(kCflThreads) void k_cfl_2d(Params P, DevScalars* S, CflArgsT<R> A) {
  const bool halt = halted(S);
  const int nb = (int)gridDim.x;
  const int lb = xcd_block(blockIdx.x, nb);
  const long per = ((P.nc + nb - 1) / nb + kCflThreads - 1) / kCflThreads * kCflThreads;
  const long c0 = (long)lb * per, c1 = (c0 + per < P.nc) ? c0 + per : P.nc;
  double dx_sound = 1.7976931348623157e308;
  const int nnx = P.nx + 1;
  // the thread's first cell (i, j); each later one is kCflThreads further on,
  // so (i, j) advance without a division and the node loads do not wait
  // for one
  long c = c0 + threadIdx.x;
  int i, j;
  divmod32(c, P.nx, j, i);
#pragma unroll 2
  for (; c < c1; c += kCflThreads) {
    R vmax = R(0);
    const R Lc = A.L[c], cc = A.cs[c];
    if (!P.lagonly) {
      const long n0 = i + (long)nnx * j;
      const long nd[4] = {n0, n0 + 1, n0 + 1 + nnx, n0 + nnx};
      R vx[4], vy[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vx[q] = A.ux[nd[q]];
        vy[q] = A.uy[nd[q]];
      }
      // the root of the largest squared norm: sqrt is correctly rounded, hence
      // monotone, so it equals the largest root (one root per cell, not per node)
      R n2max = R(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const R ux = vx[q], uy = vy[q];
        const R n2 = ux * ux + uy * uy + R(0) * R(0);
        n2max = (n2 < n2max) ? n2max : n2;
      }
      vmax = sqrt(n2max);
    }
    dx_sound = smin(dx_sound, (double)qdiv(Lc, cc + vmax));
    i += kCflThreads;
    while (i >= P.nx) {
      i -= P.nx;
      ++j;
    }
  }
  if (halt) return;
  block_min_store<kCflThreads>(dx_sound, A.block_min + lb);
}

void launch_eos_env(hipStream_t st, const Params& P, DevScalars* S, const EosArgs& A, int k, int mode, long c0,
                    long c1) {
  if (c1 <= c0) return;
  dim3 g(nblocks(c1 - c0, 256)), b(256);
  if (mode == EOS_INIT) hipLaunchKernelGGL((k_eos_env<EOS_INIT>), g, b, 0, st, P, S, A, k, c0, c1);
  else hipLaunchKernelGGL((k_eos_env<EOS_APPLY>), g, b, 0, st, P, S, A, k, c0, c1);
}

void launch_pression_moyenne(hipStream_t st, const Params& P, DevScalars* S, const PmArgs& A) {
  dim3 g(nblocks(P.nc, 256)), b(256);
#define PM_(N) hipLaunchKernelGGL((k_pression_moyenne<N>), g, b, 0, st, P, S, A)
  MHG_NEM_DISPATCH(P.ne, PM_);
#undef PM_
}

void launch_cfl_3d(hipStream_t st, const Params& P, DevScalars* S, const CflArgs& A);

void launch_cfl(hipStream_t st, const Params& P, DevScalars* S, const CflArgs& A) {
  if (P.dim == 3) {
    launch_cfl_3d(st, P, S, A);
    return;
  }
  hipLaunchKernelGGL(k_cfl_2d<double>, dim3(cfl_blocks(P)), dim3(kCflThreads), 0, st, P, S, A);
}

template <class R>
void launch_cfl_2d(hipStream_t st, const Params& P, DevScalars* S, const CflArgsT<R>& A) {
  hipLaunchKernelGGL(k_cfl_2d<R>, dim3(cfl_blocks(P)), dim3(kCflThreads), 0, st, P, S, A);
}

void launch_dt_finalize(hipStream_t st, const Params& P, DevScalars* S, const double* block_min, bool advance,
                        int nb) {
  if (nb <= 0) nb = cell_final_blocks(P);
  if (advance) hipLaunchKernelGGL((k_dt_finalize<1024, true>), dim3(1), dim3(1024), 0, st, P, S, block_min, nb);
  else hipLaunchKernelGGL((k_dt_finalize<1024, false>), dim3(1), dim3(1024), 0, st, P, S, block_min, nb);
}

template <class R>
void launch_cell_final_2d(hipStream_t st, const Params& P, DevScalars* S, const CellFinalArgsT<R>& A,
                          int dt_mode, bool eos) {
  constexpr int B = 256;
  dim3 g(nblocks(P.nc, B)), b(B);
#define F_(N) hipLaunchKernelGGL((k_cell_final_2d<R, B, N>), g, b, 0, st, P, S, A)
#define FN_(N) hipLaunchKernelGGL((k_cell_final_2d<R, B, N, false>), g, b, 0, st, P, S, A)
  if (eos) MHG_NEM_DISPATCH(P.ne, F_);
  else MHG_NEM_DISPATCH(P.ne, FN_);
#undef FN_
#undef F_
  if (dt_mode > 0) launch_dt_finalize(st, P, S, A.block_min, dt_mode == 2);
}

// The fused step for both storage types.
#define MHG_INSTANTIATE(R)                                                                           \
  template void launch_node_lagrange_2d<R>(hipStream_t, const Params&, DevScalars*,                  \
                                           const NodeLagArgsT<R>&, bool, bool, bool);                \
  template void launch_node_lagrange_fold_2d<R>(hipStream_t, const Params&, const DevScalars*,       \
                                                DevScalars*, const NodeLagArgsT<R>&, bool,           \
                                                const double*, int);                                 \
  template void launch_cell_lagrange_2d<R>(hipStream_t, const Params&, DevScalars*,                  \
                                           const CellLagArgsT<R>&);                                  \
  template void launch_lag_coords_2d<R>(hipStream_t, const Params&, DevScalars*, const R*, const R*,  \
                                        double*, double*);                                           \
  template void launch_cell_sweep_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const CellSweepArgsT<R>&, int, bool, int);                   \
  template void launch_node_sweep_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const NodeSweepArgsT<R>&, int, bool, int);                   \
  template void launch_cell_final_2d<R>(hipStream_t, const Params&, DevScalars*,                     \
                                        const CellFinalArgsT<R>&, int, bool);                        \
  template void launch_cell_sweep_final_2d<R>(hipStream_t, const Params&, DevScalars*,               \
                                              const CellSweepArgsT<R>&, const CellFinalArgsT<R>&, int); \
  template void launch_cfl_2d<R>(hipStream_t, const Params&, DevScalars*, const CflArgsT<R>&);              \
  template void launch_cell_sweep_lag_2d<R>(hipStream_t, const Params&, DevScalars*, const CellSweepArgsT<R>&, \
                                            const CellLagArgsT<R>&, int);
MHG_INSTANTIATE(double)
MHG_INSTANTIATE(float)
#undef MHG_INSTANTIATE

}  // namespace mhg
