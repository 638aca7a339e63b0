// mhg_kernels3d.hip -- 3D (hexa8) kernels of the MaHyCo time step for gfx950.
//
// Same structure as the 2D path: node phase, cell phase, one cell sweep and
// one dual sweep per direction (order (i + SensProjection) % 3), final cell
// phase, dt.  Layout: cells i + nx*(j + ny*k), nodes i + (nx+1)*(j + (ny+1)*k),
// bulk fields in the storage type R, coordinates and Euler volumes in double.
// Every Cartesian face normal is +e_dir (InitGeometricValues,
// MahycoModule.cc:1016-1025, with face nodes ordered from the smallest id), so
// dot products with face normals reduce to one component.
#include <algorithm>
#include <type_traits>

#include "mhg_common.hpp"
#include "mhg_kernels.hpp"
#include "mhg_kernels_impl.hpp"

namespace mhg {

__device__ inline long nid3(const Params& P, int i, int j, int k) {
  return i + (long)(P.nx + 1) * (j + (long)(P.ny + 1) * k);
}
__device__ inline long cid3(const Params& P, int i, int j, int k) {
  return i + (long)P.nx * (j + (long)P.ny * k);
}
__device__ inline double x0_3d(const Params& P, int d, int i) { return P.org[d] + i * P.h[d]; }

// The 8 nodes of cell (ci, cj, ck) in hexa8 order.
__device__ inline void hex_nodes(const Params& P, const double* X, const double* Y, const double* Z, int ci,
                                 int cj, int ck, V3* c) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
    const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
    const int dk = q >= 4 ? 1 : 0;
    const long n = nid3(P, ci + di, cj + dj, ck + dk);
    c[q] = {X[n], Y[n], Z[n]};
  }
}

// ---------------------------------------------------------------------------
// Node phase (MahycoModule.cc:589-665, 936-1004, 870-931): force from the 8
// cells around the node in ascending id with the corner normals of the
// previous Lagrangian coordinates, velocity, boundary conditions, position.
// ---------------------------------------------------------------------------
template <class R, bool VEL, bool BC, bool POS>
__global__ __launch_bounds__(256) void k_node_lagrange_3d(Params P, DevScalars* S, NodeLag3ArgsT<R> A) {
  if (S->stop || S->err) return;
  const long nid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nid >= P.nn) return;
  const int nnx = P.nx + 1, nny = P.ny + 1;
  int i, j, k, t_;
  divmod32(nid, nnx, t_, i);
  divmod32(t_, nny, k, j);
  const double dt = S->dt;
  double vx = A.b.ux[nid], vy = A.b.uy[nid], vz = A.uz[nid];
  if (VEL) {
    const double dtv = 0.5 * (S->dt_old + S->dt);
    V3 F = {0., 0., 0.};
    double msum = 0.;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int di = (t & 1) - 1, dj = ((t >> 1) & 1) - 1, dk = (t >> 2) - 1;
      const int ci = i + di, cj = j + dj, ck = k + dk;
      if (ci < 0 || ci >= P.nx || cj < 0 || cj >= P.ny || ck < 0 || ck >= P.nz) continue;
      V3 cn[8], fc[6];
      hex_nodes(P, A.b.xpx, A.b.xpy, A.xpz, ci, cj, ck, cn);
#pragma unroll
      for (int f = 0; f < 6; ++f) fc[f] = hex_face_centre(cn, f);
      // the node is corner 6 - t of the cell (nodeCells order, hexa8 numbering)
      V3 C;
      switch (t) {
        case 0: C = hex_corner_normal<6>(cn, fc); break;
        case 1: C = hex_corner_normal<7>(cn, fc); break;
        case 2: C = hex_corner_normal<5>(cn, fc); break;
        case 3: C = hex_corner_normal<4>(cn, fc); break;
        case 4: C = hex_corner_normal<2>(cn, fc); break;
        case 5: C = hex_corner_normal<3>(cn, fc); break;
        case 6: C = hex_corner_normal<1>(cn, fc); break;
        default: C = hex_corner_normal<0>(cn, fc); break;
      }
      const long c = cid3(P, ci, cj, ck);
      const double pq = (double)A.b.p[c] + (double)A.b.q[c];
      F = v3add(F, v3mul(pq, C));
      msum += (double)A.b.m[c];
    }
    const double mn = 0.125 * msum;
    const double s = dtv / mn;
    vx = vx + s * F.x;
    vy = vy + s * F.y;
    vz = vz + s * F.z;
  }
  if (BC) {
    for (int b = 0; b < P.nbc; ++b) {
      const BcRange& r = P.bc[b];
      const int idx = (r.axis == 0) ? i : (r.axis == 1 ? j : k);
      if (r.axis < 0 || (idx >= r.lo && idx <= r.hi)) {
        if (r.comp == 0) vx = r.value;
        else if (r.comp == 1) vy = r.value;
        else vz = r.value;
      }
    }
  }
  if (VEL || BC) {
    st_stream<3>(P, &A.b.ux[nid], (R)vx);
    st_stream<3>(P, &A.b.uy[nid], (R)vy);
    st_stream<3>(P, &A.uz[nid], (R)vz);
  }
  if (POS) {
    // Euler scheme: from x0; Lagrange-only (with-projection = false): from
    // the previous coordinates
    const double bx = P.lagonly ? A.b.xpx[nid] : x0_3d(P, 0, i);
    const double by = P.lagonly ? A.b.xpy[nid] : x0_3d(P, 1, j);
    const double bz = P.lagonly ? A.xpz[nid] : x0_3d(P, 2, k);
    st_stream<3>(P, &A.b.xnx[nid], bx + dt * vx);
    st_stream<3>(P, &A.b.xny[nid], by + dt * vy);
    st_stream<3>(P, &A.xnz[nid], bz + dt * vz);
  }
}

// The same three phases for every node, on tiles: a block owns 15 x 15 nodes
// of each node plane of a range of planes along z and marches up the range.
// For each cell plane it forms once per cell (16 x 16 cells, the tile's
// nodes and their low-side neighbours) the 24 half cross products of
// computeCQs and the 8 corner forces (p + q) C_q into LDS; each node then adds
// its 8 cells' corner forces in ascending cell id, the 4 cells below from
// the previous plane (kept in registers) and the 4 above from this one, the
// order and arithmetic of k_node_lagrange_3d, so the two agree bit for bit.
// A block starting above the first plane forms the cell plane below its
// range first.
constexpr int kNl3Tile = 15;                  // nodes per tile side
constexpr int kNl3Cells = kNl3Tile + 1;       // cells per tile side (low halo)
template <class R>
__global__ __launch_bounds__(256) void k_node_lagrange_3d_tiled(Params P, DevScalars* S, NodeLag3ArgsT<R> A, int ntx,
                                                              int nty, int kchunk) {
  if (S->stop || S->err) return;
  constexpr int TC = kNl3Cells;
  __shared__ double sF[3][8][TC * TC];  // corner forces: component, corner, cell slot
  __shared__ double sM[TC * TC];        // cell masses
  const int tid = threadIdx.x;
  const int ntile = ntx * nty;
  const int tile = blockIdx.x % ntile, chunk = blockIdx.x / ntile;
  const int i0 = (tile % ntx) * kNl3Tile, j0 = (tile / ntx) * kNl3Tile;
  const int nnx = P.nx + 1, nny = P.ny + 1, nnz = P.nz + 1;
  const int k0 = chunk * kchunk, k1 = min(k0 + kchunk, nnz);
  const int ti = tid % TC, tj = tid / TC;
  const int ni = i0 + ti, nj = j0 + tj;  // node of the thread in the node phase
  const bool node_ok = ti < kNl3Tile && tj < kNl3Tile && ni < nnx && nj < nny;
  const int ci = i0 - 1 + ti, cj = j0 - 1 + tj;  // cell of the thread in the cell phase
  const bool cell_ok = ci >= 0 && ci < P.nx && cj >= 0 && cj < P.ny;
  const double dt = S->dt, dtv = 0.5 * (S->dt_old + S->dt);
  auto cell_phase = [&](int kc) {
    if (!cell_ok) return;
    V3 cn[8], fc[6], n[24];
    hex_nodes(P, A.b.xpx, A.b.xpy, A.xpz, ci, cj, kc, cn);
#pragma unroll
    for (int f = 0; f < 6; ++f) fc[f] = hex_face_centre(cn, f);
#pragma unroll
    for (int t = 0; t < 24; ++t) n[t] = hex_term_value(cn, fc, t);
    const long c = cid3(P, ci, cj, kc);
    const double pq = (double)A.b.p[c] + (double)A.b.q[c];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const V3 f = v3mul(pq, hex_corner_sum(n, q));
      sF[0][q][tid] = f.x;
      sF[1][q][tid] = f.y;
      sF[2][q][tid] = f.z;
    }
    sM[tid] = (double)A.b.m[c];
  };
  // the node's cells in a cell plane: t = 0..3 (di, dj) = (-1,-1), (0,-1),
  // (-1,0), (0,0); corner of the node in the cell: `below` for the plane
  // under the node, else the plane above
  auto add_plane = [&](V3& F, double& msum, bool below) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int di = (t & 1) - 1, dj = (t >> 1) - 1;
      const int gi = ni + di, gj = nj + dj;
      if (gi < 0 || gi >= P.nx || gj < 0 || gj >= P.ny) continue;
      const int slot = (ti + di + 1) + TC * (tj + dj + 1);
      const int q = below ? (t == 0 ? 6 : (t == 1 ? 7 : (t == 2 ? 5 : 4))) : (t == 0 ? 2 : (t == 1 ? 3 : (t == 2 ? 1 : 0)));
      F = v3add(F, V3{sF[0][q][slot], sF[1][q][slot], sF[2][q][slot]});
      msum += sM[slot];
    }
  };
  V3 Fn = {0., 0., 0.};
  double mn = 0.;
  if (k0 > 0) {
    cell_phase(k0 - 1);
    __syncthreads();
    if (node_ok) add_plane(Fn, mn, true);
    __syncthreads();
  }
  for (int kc = k0; kc < k1; ++kc) {
    if (kc < P.nz) cell_phase(kc);
    __syncthreads();
    if (node_ok) {
      V3 F = Fn;
      double msum = mn;
      if (kc < P.nz) add_plane(F, msum, false);
      const long nid = nid3(P, ni, nj, kc);
      double vx = A.b.ux[nid], vy = A.b.uy[nid], vz = A.uz[nid];
      const double mnode = 0.125 * msum;
      const double s = dtv / mnode;
      vx = vx + s * F.x;
      vy = vy + s * F.y;
      vz = vz + s * F.z;
      for (int b = 0; b < P.nbc; ++b) {
        const BcRange& r = P.bc[b];
        const int idx = (r.axis == 0) ? ni : (r.axis == 1 ? nj : kc);
        if (r.axis < 0 || (idx >= r.lo && idx <= r.hi)) {
          if (r.comp == 0) vx = r.value;
          else if (r.comp == 1) vy = r.value;
          else vz = r.value;
        }
      }
      st_stream<3>(P, &A.b.ux[nid], (R)vx);
      st_stream<3>(P, &A.b.uy[nid], (R)vy);
      st_stream<3>(P, &A.uz[nid], (R)vz);
      const double bx = P.lagonly ? A.b.xpx[nid] : x0_3d(P, 0, ni);
      const double by = P.lagonly ? A.b.xpy[nid] : x0_3d(P, 1, nj);
      const double bz = P.lagonly ? A.xpz[nid] : x0_3d(P, 2, kc);
      st_stream<3>(P, &A.b.xnx[nid], bx + dt * vx);
      st_stream<3>(P, &A.b.xny[nid], by + dt * vy);
      st_stream<3>(P, &A.xnz[nid], bz + dt * vz);
      Fn = V3{0., 0., 0.};
      mn = 0.;
      if (kc < P.nz) add_plane(Fn, mn, true);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Cell phase: volume and length of the hexahedron (MahycoModule.cc:1054-1267),
// then the shared density / viscosity / energy / remap-vector arithmetic.
// ---------------------------------------------------------------------------
template <class R, bool PC, int NEM>
__global__ __launch_bounds__(256) void k_cell_lagrange_3d(Params P, DevScalars* S, CellLag3ArgsT<R> A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  int i, j, k, t_;
  divmod32(c, P.nx, t_, i);
  divmod32(t_, P.ny, k, j);
  const CellPre<R> in = cell_pre_load<R, PC>(A.b, c);
  V3 cn[8];
  hex_nodes(P, A.b.xnx, A.b.xny, A.xnz, i, j, k, cn);
  const double Vd = hex_volume(cn);
  const double Ld = (P.lc == MHG_LC_FACES_OPPOSEES) ? hex_length_faces_opposees(cn) : pow(Vd, 1. / 3.);
  lagrange_cell_bulk<R, PC, NEM>(P, S, A.b, c, Vd, Ld, in);
}

// ---------------------------------------------------------------------------
// Cell sweeps.  Geometry of cell c along D: centre 0.125 * (sum of its 8
// nodes) (cellCoordFromNodes), back face along D (node index s along D):
// centre sum of 0.25 * x over its 4 nodes, velocity 0.25 * (sum of the 4
// node velocities) . e_D, length |0.5 (x2 - x0) x (x3 - x1)|_D
// (PrepareRemap.cc:15-123).  Fluxes of a cell: back face then front face
// (hexa8 face order), contrib = 0.25 * dual flux (DualPhaseRemap.cc:147-188).
// ---------------------------------------------------------------------------
// Face nodes of the D-face with node index (i, j, k) in Arcane order.
__device__ inline void face_nodes3(const Params& P, int D, int i, int j, int k, long* n) {
  if (D == 0) {
    n[0] = nid3(P, i, j, k); n[1] = nid3(P, i, j + 1, k);
    n[2] = nid3(P, i, j + 1, k + 1); n[3] = nid3(P, i, j, k + 1);
  } else if (D == 1) {
    n[0] = nid3(P, i, j, k); n[1] = nid3(P, i + 1, j, k);
    n[2] = nid3(P, i + 1, j, k + 1); n[3] = nid3(P, i, j, k + 1);
  } else {
    n[0] = nid3(P, i, j, k); n[1] = nid3(P, i + 1, j, k);
    n[2] = nid3(P, i + 1, j + 1, k); n[3] = nid3(P, i, j + 1, k);
  }
}

// Geometry of one cell of a sweep along D: centre component, and its back
// face's centre component, velocity and length.
template <class R, int D>
__device__ inline void sweep_geometry3(const Params& P, const CellSweep3ArgsT<R>& A, int i, int j, int k,
                                       bool cell_ok, bool face_ok, double* xc, double* xf, R* uf,
                                       double* lf) {
  const double* X[3] = {A.b.xnx, A.b.xny, A.xnz};
  const R* U3[3] = {A.b.ux, A.b.uy, A.uz};
  const double* Xd = X[D];
  // every load first, from clamped indices (cell / face (0, 0, 0) when out of
  // the mesh), so that a wavefront waits for memory once
  const int ci = cell_ok ? i : 0, cj = cell_ok ? j : 0, ck = cell_ok ? k : 0;
  double cx[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
    const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
    const int dk = q >= 4 ? 1 : 0;
    cx[q] = Xd[nid3(P, ci + di, cj + dj, ck + dk)];
  }
  long n[4];
  face_nodes3(P, D, face_ok ? i : 0, face_ok ? j : 0, face_ok ? k : 0, n);
  double fx[4];
  R fu[4];
  V3 xq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    fx[q] = Xd[n[q]];
    fu[q] = U3[D][n[q]];
    xq[q] = V3{A.b.xnx[n[q]], A.b.xny[n[q]], A.xnz[n[q]]};
  }
  *xc = 0.;
  if (cell_ok) {
    double somme = 0.;
#pragma unroll
    for (int q = 0; q < 8; ++q) somme += cx[q];
    *xc = 0.125 * somme;
  }
  *xf = 0.;
  *uf = R(0);
  *lf = 0.;
  if (face_ok) {
    double fc = 0., vm = 0.;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      fc += 0.25 * fx[q];
      vm += (double)fu[q];
    }
    *xf = fc;
    *uf = (R)(0.25 * vm);
    const V3 t = v3mul(0.5, v3cross(v3sub(xq[2], xq[0]), v3sub(xq[3], xq[1])));
    *lf = fabs(D == 0 ? t.x : (D == 1 ? t.y : t.z));
  }
}

// Width of a hexahedron along D (computeGradPhiFace, HCELL_BY_CELLS,
// RemapADIService.cc:223-255): the distances from the cell centre (1/8 of
// the sum of its nodes) to the centres of its back and front faces (1/4 of
// each node, in the face's node order), the face on the mesh boundary left
// out.  s: the cell's index along D, N: cells along D.
template <class R, int D>
__device__ inline double cell_width_3d(const Params& P, const CellSweep3ArgsT<R>& A, int i, int j, int k, int s,
                                       int N) {
  const double* X[3] = {A.b.xnx, A.b.xny, A.xnz};
  double c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double somme = 0.;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
      const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
      const int dk = q >= 4 ? 1 : 0;
      somme += X[d][nid3(P, i + di, j + dj, k + dk)];
    }
    c[d] = 0.125 * somme;
  }
  auto face_dist = [&](int fi, int fj, int fk) {
    long n[4];
    face_nodes3(P, D, fi, fj, fk, n);
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double fc = 0.;
#pragma unroll
      for (int q = 0; q < 4; ++q) fc += 0.25 * X[d][n[q]];
      v[d] = fc - c[d];
    }
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  };
  const double db = face_dist(i, j, k);
  const double df = face_dist(i + (D == 0), j + (D == 1), k + (D == 2));
  return (s == 0 ? 0. : db) + (s == N - 1 ? 0. : df);
}

// x sweep: a wavefront holds 64 consecutive cells of a row (j, k).
template <class R, int NEM, bool FIRST, int LIM>
__global__ __launch_bounds__(256) void k_cell_sweep_x_3d(Params P, DevScalars* S, CellSweep3ArgsT<R> A) {
  if (S->stop || S->err) return;
  const int lane = threadIdx.x & 63;
  const int N = P.nx;
  constexpr int HL = (LIM == kLimO3) ? 3 : 2, NO = 64 - 2 * HL;  // halo lanes (order 3: 3), outputs
  const int nseg = (N + NO - 1) / NO;
  const long wave = ((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
  int line_, seg;
  divmod32(wave, nseg, line_, seg);
  const long line = line_;
  if (line >= (long)P.ny * P.nz) return;  // uniform across the wave
  int j, k;
  divmod32(line, P.ny, k, j);
  const int i = seg * NO - HL + lane;
  const bool ex = (i >= 0 && i < N);
  const bool out = ex && lane >= HL && lane <= 63 - HL;
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const long c = ex ? cid3(P, i, j, k) : 0;
  const R* __restrict__ Uin = A.b.Uin;
  double xc, xf, lf;
  R uf;
  sweep_geometry3<R, 0>(P, A, i, j, k, ex, i >= 0 && i <= N, &xc, &xf, &uf, &lf);
  R den = R(1);
  R pre[3][NEM];
  const R preq_l = Uin[(3 * ne) * NC + c];  // c is 0 for a lane out of the mesh
  const R preq = ex ? preq_l : R(0);
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int q = 0; q < NEM; ++q) {
      const R u = (q < ne) ? Uin[(g * ne + q) * NC + c] : R(0);
      pre[g][q] = (ex && q < ne) ? u : R(0);
    }
  const R vlag = FIRST ? A.b.Vlag[c] : R(1);
  if (ex) {
    if (FIRST) {
      den = vlag;
    } else {
      R sv = R(0);
#pragma unroll
      for (int q = 0; q < NEM; ++q)
        if (q < ne) sv += pre[0][q];
      den = sv;
    }
  }
  const bool inner = ex && i >= 1;
  const double xcm = lane_up1(xc);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = lane_down1(uf);
  const double lf_f = lane_down1(lf);
  const R fb = flux_factor<R>(-uf, lf, dt), ff = flux_factor<R>(uf_f, lf_f, dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  const bool o2 = P.order2 > 0;
  const int sd = i;  // index along the sweep
  int est_own = 0;   // order 3: the cell's code (first sweep: the last remap's; then phi_V > 0)
  if (LIM == kLimO3) {
    if (FIRST) {
      const int el = A.b.est[c];
      est_own = ex ? el : 0;
    } else {
      int cnt = 0, last = -1;
#pragma unroll
      for (int q = 0; q < NEM; ++q)
        if (q < ne && qdiv(pre[0][q], den) > R(0)) {
          cnt++;
          last = q;
        }
      est_own = cnt > 1 ? -2 : last;
    }
  }
  // generalised limiters: widths of the cell and of its back / front
  // neighbours along D (the cell itself where the neighbour is missing)
  bool front_in = false;
  R hc = R(0), hcm = R(0), hcp = R(0);
  R h6[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};  // order 3: widths bbb .. fff of the back face
  if (LIM == kLimG || LIM == kLimO3) {
    front_in = lane_down1((int)inner) != 0;
    const double h = ex ? cell_width_3d<R, 0>(P, A, i, j, k, i, N) : 0.;
    const double hm = lane_up1(h), hp = lane_down1(h);
    hc = (R)h;
    hcm = (R)(inner ? hm : h);
    hcp = (R)(front_in ? hp : h);
    if (LIM == kLimO3) o3_stencil<R>((R)h, sd, N, h6);
  }
  // order 3: mixed / pure codes of the cell and its back neighbour
  int est = 0, estb = 0;
  R vdt = R(0);
  if (LIM == kLimO3) {
    est = est_own;
    estb = lane_up1(est);
    vdt = vel_dt<R>(uf, dt);
  }
  const bool vpure = estb != -2 && est != -2 && estb == est;
  R o3V[NEM], o3Vb[NEM], o3Vf[NEM], o3M[NEM], o3Mb[NEM], o3Mf[NEM];
  int o3up[NEM];
  auto sweep_var = [&](R u0, R& un, R& dual, int role, int kq) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = lane_up1(phi);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = lane_down1(gf);
    R phif;
    if (LIM == kLimO3) {  // RemapADIService.cc:583-640
      R y6[6];
      o3_stencil<R>(phi, sd, N, y6);
      phif = inner ? o3_flux<R>(y6, h6, vdt) : R(0);
      o3_mixed_faces<R, NEM>(vpure, role, kq, vdt, phim, phi, &phif, o3V, o3Vb, o3Vf, o3M, o3Mb, o3Mf, o3up);
    } else {
      R g = R(0);
      if (LIM == kLimG && o2) {  // computeAndLimitGradPhi, UtilesRemap.cc:329-338
        const R phip = lane_down1(phi);
        g = limiter_g<R>(P.lim, gfp, gf, front_in ? phip : phi, inner ? phim : phi, hc, hcp, hcm);
      } else if (o2) {
        g = limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp);
      }
      const R gm = lane_up1(g);
      phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    }
    const R phifp = lane_down1(phif);
    const R flux_b = fb * phif, flux_f = ff * phifp;
    R fsum = R(0);  // faces in hexa8 order: back (x-) before front (x+)
    fsum = fsum + flux_b;
    fsum = fsum + flux_f;
    un = u0 - fsum;
    R d = R(0);
    d = d + R(0.5) * flux_b * R(-1.0);
    d = d + R(0.5) * flux_f * R(1.0);
    dual = d;
  };
  R* __restrict__ Uout = A.b.Uout;
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int q = 0; q < NEM; ++q) {
      if (q >= ne) break;
      R un, du;
      sweep_var(pre[g][q], un, du, g, q);
      if (g > 0 && un < R(0)) un = R(0);  // RemapADIService.cc:953-958
      if (out) {
        st_stream<1>(P, &Uout[(g * ne + q) * NC + c], un);
        if (g == 1) st_stream<2>(P, &A.b.contrib[q * NC + c], R(0.25) * du);
      }
    }
  {
    R un, du;
    sweep_var(preq, un, du, 3, 0);
    if (out) st_stream<1>(P, &Uout[(3 * ne) * NC + c], un);
  }
}

// y and z sweeps on LDS tiles: a block of 16 wavefronts owns 16 columns (i)
// and 64 rows along D (60 outputs) at one index of the third direction; the
// tile is transposed through LDS so that a wavefront holds one column.
template <class R, int D, int NEM, bool FIRST, int LIM>
__global__ __launch_bounds__(1024) void k_cell_sweep_tile_3d(Params P, DevScalars* S, CellSweep3ArgsT<R> A) {
  __shared__ double lds[LIM == kLimO3 ? 7 : 6][kTileR][kTileP];  // order 3: a tile for the cells' codes
  typedef R Tile[kTileR][kTileP];
  Tile* stg = reinterpret_cast<Tile*>(&lds[0][0][0]);
  if (S->stop || S->err) return;  // uniform over the grid
  const int nx = P.nx;
  const int N = (D == 1) ? P.ny : P.nz;      // cells along D
  const int T = (D == 1) ? P.nz : P.ny;      // the third direction
  const int ncolblk = (nx + kTileW - 1) / kTileW;
  constexpr int HL = (LIM == kLimO3) ? 3 : 2, NO = 64 - 2 * HL;  // halo rows (order 3: 3), outputs
  const int nseg = (N + NO - 1) / NO;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (L % ncolblk) * kTileW;
  const int s0 = ((L / ncolblk) % nseg) * NO - HL;
  const int o = L / (ncolblk * nseg);        // index along the third direction
  if (o >= T) return;                        // uniform over the block
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;
  const int w = t >> 6, lane = t & 63;
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  const double dt = S->dt;
  const R order2 = (R)P.order2;
  const bool o2 = P.order2 > 0;
  const R* __restrict__ Uin = A.b.Uin;
  auto cell_ijk = [&](int ii, int s, int* ci, int* cj, int* ck) {
    *ci = ii;
    *cj = (D == 1) ? s : o;
    *ck = (D == 1) ? o : s;
  };
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = gi < nx && gs >= 0 && gs < N;
  const bool gout = gex && tr >= HL && tr <= 63 - HL;
  int ci, cj, ck;
  cell_ijk(gi, gs, &ci, &cj, &ck);
  const long gc = gex ? cid3(P, ci, cj, ck) : 0;
  {
    double xc, xf, lf;
    R uf, den = R(1);
    sweep_geometry3<R, D>(P, A, ci, cj, ck, gex, gi < nx && gs >= 0 && gs <= N, &xc, &xf, &uf, &lf);
    if (gex) {
      if (FIRST) {
        den = A.b.Vlag[gc];
      } else {
        R sv = R(0);
#pragma unroll
        for (int q = 0; q < NEM; ++q)
          if (q < ne) sv += Uin[q * NC + gc];
        den = sv;
      }
    }
    lds[0][tr][tc] = xc;
    lds[1][tr][tc] = (double)den;
    lds[2][tr][tc] = (double)uf;
    lds[3][tr][tc] = xf;
    lds[4][tr][tc] = lf;
    if (LIM == kLimG || LIM == kLimO3) lds[5][tr][tc] = gex ? cell_width_3d<R, D>(P, A, ci, cj, ck, gs, N) : 0.;
    if (LIM == kLimO3) {  // the cell's code: the last remap's in the first sweep, then phi_V > 0
      int e = 0;
      if (FIRST) {
        const int el = A.b.est[gc];
        e = gex ? el : 0;
      } else {
        int cnt = 0, last = -1;
#pragma unroll
        for (int q = 0; q < NEM; ++q)
          if (q < ne && qdiv(Uin[q * NC + gc], den) > R(0)) {
            cnt++;
            last = q;
          }
        e = cnt > 1 ? -2 : last;
      }
      lds[LIM == kLimO3 ? 6 : 5][tr][tc] = (double)e;
    }
  }
  __syncthreads();
  const int s = s0 + lane;
  const bool ex = (i0 + w < nx) && s >= 0 && s < N;
  const bool inner = ex && s >= 1;
  const double xc = lds[0][lane][w], xf = lds[3][lane][w], lf = lds[4][lane][w];
  const R den = (R)lds[1][lane][w], uf = (R)lds[2][lane][w];
  const double xcm = lane_up1(xc);
  const double dx = inner ? xc - xcm : 1.;
  const bool upm = inner && ((double)uf * dx > 0.0);
  const double dotxf = inner ? (upm ? xf - xcm : xf - xc) : 0.;
  const R uf_f = lane_down1(uf);
  const double lf_f = lane_down1(lf);
  const R fb = flux_factor<R>(-uf, lf, dt), ff = flux_factor<R>(uf_f, lf_f, dt);
  const R dxr = (R)dx, dotr = (R)dotxf;
  const int sd = s;  // index along the sweep
  const int est_own = (LIM == kLimO3) ? (int)lds[LIM == kLimO3 ? 6 : 5][lane][w] : 0;
  // generalised limiters: widths of the cell and of its back / front
  // neighbours along D (the cell itself where the neighbour is missing)
  bool front_in = false;
  R hc = R(0), hcm = R(0), hcp = R(0);
  R h6[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};  // order 3: widths bbb .. fff of the back face
  if (LIM == kLimG || LIM == kLimO3) {
    front_in = lane_down1((int)inner) != 0;
    const double h = lds[5][lane][w];
    const double hm = lane_up1(h), hp = lane_down1(h);
    hc = (R)h;
    hcm = (R)(inner ? hm : h);
    hcp = (R)(front_in ? hp : h);
    if (LIM == kLimO3) o3_stencil<R>((R)h, sd, N, h6);
  }
  // order 3: mixed / pure codes of the cell and its back neighbour
  int est = 0, estb = 0;
  R vdt = R(0);
  if (LIM == kLimO3) {
    est = est_own;
    estb = lane_up1(est);
    vdt = vel_dt<R>(uf, dt);
  }
  const bool vpure = estb != -2 && est != -2 && estb == est;
  R o3V[NEM], o3Vb[NEM], o3Vf[NEM], o3M[NEM], o3Mb[NEM], o3Mf[NEM];
  int o3up[NEM];
  auto sweep_var = [&](R u0, R& un, R& dual, int role, int kq) {
    const R phi = ex ? qdiv(u0, den) : R(0);
    const R phim = lane_up1(phi);
    const R gf = face_grad(inner, phi, phim, dxr);
    const R gfp = lane_down1(gf);
    R phif;
    if (LIM == kLimO3) {  // RemapADIService.cc:583-640
      R y6[6];
      o3_stencil<R>(phi, sd, N, y6);
      phif = inner ? o3_flux<R>(y6, h6, vdt) : R(0);
      o3_mixed_faces<R, NEM>(vpure, role, kq, vdt, phim, phi, &phif, o3V, o3Vb, o3Vf, o3M, o3Mb, o3Mf, o3up);
    } else {
      R g = R(0);
      if (LIM == kLimG && o2) {  // computeAndLimitGradPhi, UtilesRemap.cc:329-338
        const R phip = lane_down1(phi);
        g = limiter_g<R>(P.lim, gfp, gf, front_in ? phip : phi, inner ? phim : phi, hc, hcp, hcm);
      } else if (o2) {
        g = limited_grad<LIM, true, R, kBranchFree<R>>(gf, gfp);
      }
      const R gm = lane_up1(g);
      phif = inner ? (upm ? phim + order2 * dotr * gm : phi + order2 * dotr * g) : R(0);
    }
    const R phifp = lane_down1(phif);
    const R flux_b = fb * phif, flux_f = ff * phifp;
    R fsum = R(0);  // faces in hexa8 order: back before front
    fsum = fsum + flux_b;
    fsum = fsum + flux_f;
    un = u0 - fsum;
    R d = R(0);
    d = d + R(0.5) * flux_b * R(-1.0);
    d = d + R(0.5) * flux_f * R(1.0);
    dual = d;
  };
  __syncthreads();  // geometry reads done: the tiles become staging tiles
  const int nvu = P.nvu;
  R nxa = gex ? Uin[gc] : R(0);
  R nxb = (gex && nvu > 1) ? Uin[NC + gc] : R(0);
  auto role_of = [&](int v) { return v < ne ? 0 : v < 2 * ne ? 1 : v < 3 * ne ? 2 : 3; };
  auto env_of = [&](int v) { return v < ne ? v : v < 2 * ne ? v - ne : v < 3 * ne ? v - 2 * ne : 0; };
  auto round = [&](const int va) {
    const int vb = va + 1;
    const bool hb = vb < nvu;  // uniform
    stg[0][tr][tc] = nxa;
    stg[1][tr][tc] = nxb;
    __syncthreads();
    if (va + 2 < nvu) nxa = gex ? Uin[(va + 2) * NC + gc] : R(0);
    if (va + 3 < nvu) nxb = gex ? Uin[(va + 3) * NC + gc] : R(0);
    const R ua = stg[0][lane][w], ub = stg[1][lane][w];
    R una, dua, unb, dub;
    sweep_var(ua, una, dua, role_of(va), env_of(va));
    sweep_var(ub, unb, dub, role_of(vb), env_of(vb));
    if (va >= ne && va < 3 * ne && una < R(0)) una = R(0);  // RemapADIService.cc:953-958
    if (vb >= ne && vb < 3 * ne && unb < R(0)) unb = R(0);
    const bool ma = (va >= ne && va < 2 * ne), mb = hb && (vb >= ne && vb < 2 * ne);
    stg[2][lane][w] = una;
    stg[3][lane][w] = unb;
    if (ma) stg[4][lane][w] = R(0.25) * dua;
    if (mb) stg[5][lane][w] = R(0.25) * dub;
    __syncthreads();
    if (gout) {
      st_stream<1>(P, &A.b.Uout[va * NC + gc], stg[2][tr][tc]);
      if (hb) st_stream<1>(P, &A.b.Uout[vb * NC + gc], stg[3][tr][tc]);
      if (ma) st_stream<2>(P, &A.b.contrib[(va - ne) * NC + gc], stg[4][tr][tc]);
      if (mb) st_stream<2>(P, &A.b.contrib[(vb - ne) * NC + gc], stg[5][tr][tc]);
    }
  };
  if (LIM == kLimO3) {  // compile-time variable indices: per-environment values stay in registers
#pragma unroll
    for (int va = 0; va < 3 * NEM + 1; va += 2)
      if (va < nvu) round(va);
  } else {
    for (int va = 0; va < nvu; va += 2) round(va);
  }
}

// ---------------------------------------------------------------------------
// Dual sweeps (DualPhaseRemap.cc:22-845).  phi_dual of a node is U/m if the
// node was inner in a direction already swept in this remap (494-523), else
// its velocity; the back flux of a node is the front flux of its neighbour
// behind it; front(node) sums 0.25 * dual flux of the 4 cells ahead of it
// (dir1 = D+1 outer, dir2 = D+2 inner, environments innermost).
// ---------------------------------------------------------------------------
template <class R>
struct DualNode {
  R U[4];    // dual state before the sweep
  R p[3];    // phi_dual
  R U4, p4;  // dual kinetic energy and its phi (total-energy fix-up only)
  double xd; // coordinate along D
  R front;   // front mass flux
};

template <class R, int D, bool FIRST, int NEM>
__device__ inline DualNode<R> dual_node_load(const Params& P, const NodeSweep3ArgsT<R>& A, int i, int j, int k,
                                             bool ok) {
  // Every load is issued first, from clamped indices and without branches,
  // so that a wavefront waits for memory once: a node out of the mesh reads
  // node 0 and is zeroed below, a missing cell reads cell 0 and adds +0 to a
  // sum that starts at +0 (so the sums are the guarded sums bit for bit).
  const long nid = ok ? nid3(P, i, j, k) : 0;
  const double* X[3] = {A.xnx, A.xny, A.xnz};
  const double xd_l = X[D][nid];
  const R u[3] = {A.ux[nid], A.uy[nid], A.uz[nid]};
  R cm[8];
  bool cmv[8];
  R Ul[4];
  if (FIRST) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {  // node mass 0.125 * sum of the 8 cells (computeNodeMass)
      const int ci = i + (t & 1) - 1, cj = j + ((t >> 1) & 1) - 1, ck = k + (t >> 2) - 1;
      cmv[t] = ci >= 0 && ci < P.nx && cj >= 0 && cj < P.ny && ck >= 0 && ck < P.nz;
      cm[t] = A.m[cmv[t] ? cid3(P, ci, cj, ck) : 0];
    }
  } else {
    Ul[0] = A.Udi0[nid];
    Ul[1] = A.Udi1[nid];
    Ul[2] = A.Udi2[nid];
    Ul[3] = A.Udi3[nid];
  }
  const R u4l = (!FIRST && P.etot) ? A.Udi4[nid] : R(0);
  // front cells: index D at the node, dir1 and dir2 shifted by -1 and 0
  const int d1 = (D + 1) % 3, d2 = (D + 2) % 3;
  const int nn3[3] = {P.nx, P.ny, P.nz};
  const int base[3] = {i, j, k};
  R cf[4][NEM];
  bool cfv[4];
#pragma unroll
  for (int s1 = -1; s1 <= 0; ++s1)
#pragma unroll
    for (int s2 = -1; s2 <= 0; ++s2) {
      const int q = 2 * (s1 + 1) + (s2 + 1);
      int cc[3] = {i, j, k};
      cc[d1] += s1;
      cc[d2] += s2;
      cfv[q] = ok && base[D] < nn3[D] && cc[d1] >= 0 && cc[d1] < nn3[d1] && cc[d2] >= 0 && cc[d2] < nn3[d2];
      const long cell = cfv[q] ? cid3(P, cc[0], cc[1], cc[2]) : 0;
#pragma unroll
      for (int e = 0; e < NEM; ++e) cf[q][e] = (e < env_count<NEM>(P)) ? A.contrib[e * P.nc + cell] : R(0);
    }
  DualNode<R> d;
  for (int q = 0; q < 4; ++q) d.U[q] = R(0);
  for (int q = 0; q < 3; ++q) d.p[q] = R(0);
  d.U4 = d.p4 = R(0);
  d.xd = 0.;
  d.front = R(0);
  if (!ok) return d;
  d.xd = xd_l;
  if (FIRST) {
    R msum = R(0);
#pragma unroll
    for (int t = 0; t < 8; ++t) msum += cmv[t] ? cm[t] : R(0);
    const R mn = R(0.125) * msum;
    for (int q = 0; q < 3; ++q) {
      d.p[q] = u[q];
      d.U[q] = mn * u[q];
    }
    d.U[3] = mn;
    if (P.etot) {  // 1/2 m |u| and 1/2 |u| (PrepareRemap.cc:232, 243)
      const R nrm = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      d.p4 = R(0.5) * nrm;
      d.U4 = R(0.5) * mn * nrm;
    }
  } else {
    for (int q = 0; q < 4; ++q) d.U[q] = Ul[q];
    const int idx[3] = {i, j, k};
    const int nn[3] = {P.nx, P.ny, P.nz};
    bool inner_prev = false;
    for (int e = 0; e < 3; ++e)
      if (((A.prevmask >> e) & 1) && idx[e] >= 1 && idx[e] <= nn[e] - 1) inner_prev = true;
    for (int q = 0; q < 3; ++q) d.p[q] = inner_prev ? qdiv(d.U[q], d.U[3]) : u[q];
    if (P.etot) {
      d.U4 = u4l;
      d.p4 = inner_prev ? qdiv(d.U4, d.U[3]) : R(0.5) * sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    }
  }
  if (base[D] < nn3[D]) {
    R f = R(0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < NEM; ++e)
        if (e < env_count<NEM>(P)) f += cfv[q] ? cf[q][e] : R(0);
    d.front = f;
  }
  return d;
}

// Increments of the dual state of an inner node along D (352-491), given the
// node and its two neighbours along D (phi, coordinate, gradient).
template <class R, int D>
__device__ inline void dual_increments(const Params& P, double dt, const R* pm, const R* p0, const R* pp, double xdm,
                                       double xd, double xdp, const R* gm, const R* g0, const R* gp, R back, R front,
                                       R* dU) {
  const R order2 = (R)P.order2;
  const R ufront = R(0.5) * (p0[D] + p0[D]);  // DualPhaseRemap.cc:391
  const int signfront = (ufront > R(0)) ? 1 : -1;
  const double xfr = xdp - xd;
  const bool sf = (front < R(0));
  const R tf = upwind_offset<R>(signfront, xfr, dt, ufront);
  const R uback = R(0.5) * (pm[D] + p0[D]);
  const int signback = (uback > R(0)) ? 1 : -1;
  const double xbk = xd - xdm;
  const bool sb = (back > R(0));
  const R tb = upwind_offset<R>(signback, xbk, dt, uback);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const R FV = (sf ? pp[q] : p0[q]) + order2 * (R(0.5) * (sf ? gp[q] : g0[q]) * tf);
    const R BV = (sb ? pm[q] : p0[q]) + order2 * (R(0.5) * (sb ? gm[q] : g0[q]) * tb);
    dU[q] = back * BV - front * FV;
  }
  dU[3] = back - front;
}

template <class R, bool FIRST, bool LAST>
__device__ inline void dual_store(const Params& P, const NodeSweep3ArgsT<R>& A, long nid, const R* U, R U4) {
  if (P.etot) {
    if (!LAST) {
      st_stream<2>(P, &A.Udo4[nid], U4);
    } else {  // the node's term of the total-energy deposit (RemapADIFinal.cc:757-764)
      R t = R(0);
      if (U[3] != R(0)) {
        const R vx = qdiv(U[0], U[3]), vy = qdiv(U[1], U[3]), vz = qdiv(U[2], U[3]);
        t = R(0.25) * (qdiv(U4, U[3]) - R(0.5) * (vx * vx + vy * vy + vz * vz));
      }
      st_stream<2>(P, &A.dek[nid], t);
    }
  }
  if (!LAST) {
    st_stream<2>(P, &A.Udo0[nid], U[0]);
    st_stream<2>(P, &A.Udo1[nid], U[1]);
    st_stream<2>(P, &A.Udo2[nid], U[2]);
    st_stream<2>(P, &A.Udo3[nid], U[3]);
  } else if (U[3] != R(0)) {  // RemapADIFinal.cc:660-672
    st_stream<3>(P, &A.uox[nid], qdiv(U[0], U[3]));
    st_stream<3>(P, &A.uoy[nid], qdiv(U[1], U[3]));
    st_stream<3>(P, &A.uoz[nid], qdiv(U[2], U[3]));
  } else {
    st_stream<3>(P, &A.uox[nid], A.ux[nid]);
    st_stream<3>(P, &A.uoy[nid], A.uy[nid]);
    st_stream<3>(P, &A.uoz[nid], A.uz[nid]);
  }
}

// True if the node lies on the domain boundary of a direction other than D:
// its gradient along D is read back as a stale value by a later sweep.
template <int D>
__device__ inline bool on_other_boundary(const Params& P, int i, int j, int k) {
  const int idx[3] = {i, j, k};
  const int nn[3] = {P.nx, P.ny, P.nz};
  bool b = false;
#pragma unroll
  for (int e = 0; e < 3; ++e)
    if (e != D && (idx[e] == 0 || idx[e] == nn[e])) b = true;
  return b;
}

template <class R, int D, bool FIRST, bool LAST, int NEM, int LIM>
__global__ __launch_bounds__(256) void k_node_sweep_x_3d(Params P, DevScalars* S, NodeSweep3ArgsT<R> A) {
  if (S->stop || S->err) return;
  const int lane = threadIdx.x & 63;
  const int N = P.nx;
  const int nseg = (N + 1 + 59) / 60;
  const long wave = ((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
  int line_, seg;
  divmod32(wave, nseg, line_, seg);
  const long line = line_;
  if (line >= (long)(P.ny + 1) * (P.nz + 1)) return;  // uniform across the wave
  int j, k;
  divmod32(line, P.ny + 1, k, j);
  const int i = seg * 60 - 2 + lane;
  const bool ex = (i >= 0 && i <= N);
  const bool out = ex && lane >= 2 && lane <= 61;
  const bool inner_d = (i >= 1 && i <= N - 1);
  const double dt = S->dt;
  const DualNode<R> nd = dual_node_load<R, 0, FIRST, NEM>(P, A, i, j, k, ex);
  const long nid = ex ? nid3(P, i, j, k) : 0;
  R pm[3], pp[3], g0[3] = {R(0), R(0), R(0)}, gm[3], gp[3];
  const double xdm = lane_up1(nd.xd), xdp = lane_down1(nd.xd);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    pm[q] = lane_up1(nd.p[q]);
    pp[q] = lane_down1(nd.p[q]);
  }
  const R back = lane_up1(nd.front);
  R* dg[3] = {A.dg0, A.dg1, A.dg2};
  // generalised limiters: dual widths (UtilesRemap.cc:46-57) from the nodes
  // one and two places back and front along x
  R h0 = R(0), hm = R(0), hp = R(0);
  if (LIM == kLimG) {
    const double xmm = A.xnx[(ex && i >= 2) ? nid - 2 : 0], xpp = A.xnx[(ex && i + 2 <= N) ? nid + 2 : 0];
    h0 = (R)(0.5 * (xdp - xdm));
    hm = (R)(i >= 2 ? 0.5 * (nd.xd - xmm) : 0.);
    hp = (R)(i <= N - 2 ? 0.5 * (xpp - nd.xd) : 0.);
  }
  if (P.order2 > 0 && ex) {
    if (inner_d) {
      const R dxf = (R)(xdp - nd.xd), dxb = (R)(nd.xd - xdm);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const R gf = qdiv(pp[q] - nd.p[q], dxf), gb = qdiv(nd.p[q] - pm[q], dxb);
        g0[q] = (LIM == kLimG) ? limiter_g<R>(P.lim, gf, gb, pp[q], pm[q], h0, hp, hm)
                               : limited_grad<LIM, false, R, kBranchFree<R>>(gb, gf);
      }
      if (out && on_other_boundary<0>(P, i, j, k))
        for (int q = 0; q < 3; ++q) dg[q][nid] = g0[q];
    } else {
      for (int q = 0; q < 3; ++q) g0[q] = dg[q][nid];
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    gm[q] = lane_up1(g0[q]);
    gp[q] = lane_down1(g0[q]);
  }
  R p4m = R(0), p4p = R(0);
  if (P.etot) {
    p4m = lane_up1(nd.p4);
    p4p = lane_down1(nd.p4);
  }
  if (!out) return;
  R U[4] = {nd.U[0], nd.U[1], nd.U[2], nd.U[3]};
  R U4 = nd.U4;
  if (inner_d) {
    R dU[4];
    dual_increments<R, 0>(P, dt, pm, nd.p, pp, xdm, nd.xd, xdp, gm, g0, gp, back, nd.front, dU);
    // the kinetic energy is carried with the upwind value (its gradient is 0)
    if (P.etot) U4 += back * (back > R(0) ? p4m : nd.p4) - nd.front * (nd.front < R(0) ? p4p : nd.p4);
    U[3] += dU[3];
    for (int q = 0; q < 3; ++q) U[q] += dU[q];
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    for (int q = 0; q < 3; ++q)
      if (fabs(U[q]) < thresold) U[q] = R(0);
  }
  dual_store<R, FIRST, LAST>(P, A, nid, U, U4);
}

// y and z dual sweeps on LDS tiles (columns i, rows along D).
template <class R, int D, bool FIRST, bool LAST, int NEM, int LIM>
__global__ __launch_bounds__(1024) void k_node_sweep_tile_3d(Params P, DevScalars* S, NodeSweep3ArgsT<R> A) {
  __shared__ double ldx[kTileR][kTileP];
  __shared__ R lds[5][kTileR][kTileP];  // phi x, y, z, front flux, kinetic energy; then the increments
  if (S->stop || S->err) return;
  const int nx = P.nx, nnx = nx + 1;
  const int N = (D == 1) ? P.ny : P.nz;  // nodes 0..N along D
  const int T = (D == 1) ? P.nz : P.ny;  // nodes 0..T along the third direction
  const int ncolblk = (nnx + kTileW - 1) / kTileW;
  const int nseg = (N + 1 + 59) / 60;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (L % ncolblk) * kTileW;
  const int s0 = ((L / ncolblk) % nseg) * 60 - 2;
  const int o = L / (ncolblk * nseg);
  if (o > T) return;  // uniform over the block
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;
  const int w = t >> 6, lane = t & 63;
  const double dt = S->dt;
  auto node_ijk = [&](int ii, int s, int* ni, int* nj, int* nk) {
    *ni = ii;
    *nj = (D == 1) ? s : o;
    *nk = (D == 1) ? o : s;
  };
  // ---- row-major: load the nodes
  const int gi = i0 + tc, gs = s0 + tr;
  const bool gex = gi <= nx && gs >= 0 && gs <= N;
  const bool gout = gex && tr >= 2 && tr <= 61;
  int ni, nj, nk;
  node_ijk(gi, gs, &ni, &nj, &nk);
  const DualNode<R> nd = dual_node_load<R, D, FIRST, NEM>(P, A, ni, nj, nk, gex);
  const long gq = gex ? nid3(P, ni, nj, nk) : 0;
  ldx[tr][tc] = nd.xd;
  lds[0][tr][tc] = nd.p[0];
  lds[1][tr][tc] = nd.p[1];
  lds[2][tr][tc] = nd.p[2];
  lds[3][tr][tc] = nd.front;
  lds[4][tr][tc] = nd.p4;
  __syncthreads();
  // ---- transposed: column i, row s
  const int i = i0 + w, s = s0 + lane;
  const bool ex = i <= nx && s >= 0 && s <= N;
  const bool out = ex && lane >= 2 && lane <= 61;
  const bool inner_d = (s >= 1 && s <= N - 1);
  int ti, tj, tk;
  node_ijk(i, s, &ti, &tj, &tk);
  const double xd = ldx[lane][w];
  const R front = lds[3][lane][w];
  const double xdm = lane_up1(xd), xdp = lane_down1(xd);
  const R back = lane_up1(front);
  R* dg[3] = {A.dg0, A.dg1, A.dg2};
  // generalised limiters: dual widths (UtilesRemap.cc:46-57) from the nodes
  // one and two places back and front along D
  R h0 = R(0), hm = R(0), hp = R(0);
  if (LIM == kLimG) {
    const double* XD = (D == 1) ? A.xny : A.xnz;
    int ai, aj, ak, bi, bj, bk;
    node_ijk(i, s - 2, &ai, &aj, &ak);
    node_ijk(i, s + 2, &bi, &bj, &bk);
    const double xmm = XD[(ex && s >= 2) ? nid3(P, ai, aj, ak) : 0];
    const double xpp = XD[(ex && s + 2 <= N) ? nid3(P, bi, bj, bk) : 0];
    h0 = (R)(0.5 * (xdp - xdm));
    hm = (R)(s >= 2 ? 0.5 * (xd - xmm) : 0.);
    hp = (R)(s <= N - 2 ? 0.5 * (xpp - xd) : 0.);
  }
  // the increments of the inner nodes (352-491), one velocity component at a
  // time (dual_increments' arithmetic, fewer values live at once): upwind
  // terms from the velocity along D, then per component the phi of the node
  // and its neighbours, the limited gradients and the face values
  const R order2 = (R)P.order2;
  const R pD = lds[D][lane][w], pDm = lane_up1(pD);
  const R ufront = R(0.5) * (pD + pD);  // DualPhaseRemap.cc:391
  const int signfront = (ufront > R(0)) ? 1 : -1;
  const bool sf = (front < R(0));
  const R tf = (R)(signfront * (xdp - xd) - dt * ufront);
  const R uback = R(0.5) * (pDm + pD);
  const int signback = (uback > R(0)) ? 1 : -1;
  const bool sb = (back > R(0));
  const R tb = (R)(signback * (xd - xdm) - dt * uback);
  const bool o2ex = P.order2 > 0 && ex;
  const long q0 = o2ex ? nid3(P, ti, tj, tk) : 0;
  const bool keep_dg = out && on_other_boundary<D>(P, ti, tj, tk);
  R dU[4] = {R(0), R(0), R(0), R(0)}, dU4 = R(0);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const R p0 = lds[q][lane][w];
    const R pm = lane_up1(p0), pp = lane_down1(p0);
    R g0 = R(0);
    if (o2ex) {
      if (inner_d) {
        const R dxf = (R)(xdp - xd), dxb = (R)(xd - xdm);
        const R gf = qdiv(pp - p0, dxf), gb = qdiv(p0 - pm, dxb);
        g0 = (LIM == kLimG) ? limiter_g<R>(P.lim, gf, gb, pp, pm, h0, hp, hm)
                            : limited_grad<LIM, false, R, kBranchFree<R>>(gb, gf);
        if (keep_dg) dg[q][q0] = g0;
      } else {
        g0 = dg[q][q0];
      }
    }
    const R gm = lane_up1(g0), gp = lane_down1(g0);
    if (out && inner_d) {
      const R FV = (sf ? pp : p0) + order2 * (R(0.5) * (sf ? gp : g0) * tf);
      const R BV = (sb ? pm : p0) + order2 * (R(0.5) * (sb ? gm : g0) * tb);
      dU[q] = back * BV - front * FV;
    }
  }
  if (out && inner_d) dU[3] = back - front;
  if (P.etot) {
    // the kinetic energy is carried with the upwind value (its gradient is 0)
    const R p4 = lds[4][lane][w];
    const R p4m = lane_up1(p4), p4p = lane_down1(p4);
    if (out && inner_d) dU4 = back * (back > R(0) ? p4m : p4) - front * (front < R(0) ? p4p : p4);
  }
  __syncthreads();  // transposed reads done before the tiles are reused
#pragma unroll
  for (int q = 0; q < 4; ++q) lds[q][lane][w] = dU[q];
  lds[4][lane][w] = dU4;
  __syncthreads();
  // ---- row-major: U += increments, store
  if (!gout) return;
  R U[4] = {nd.U[0], nd.U[1], nd.U[2], nd.U[3]};
  R U4 = nd.U4;
  if (gs >= 1 && gs <= N - 1) {
    if (P.etot) U4 += lds[4][tr][tc];
    U[3] += lds[3][tr][tc];
    for (int q = 0; q < 3; ++q) U[q] += lds[q][tr][tc];
    const R thresold = (R)1.e-300;  // 0 in float: the filter only acts in double
    for (int q = 0; q < 3; ++q)
      if (fabs(U[q]) < thresold) U[q] = R(0);
  }
  dual_store<R, FIRST, LAST>(P, A, gq, U, U4);
}

// ---------------------------------------------------------------------------
// Final cell phase with the stored Euler volume; CFL over the 8 nodes.
// ---------------------------------------------------------------------------
template <class R, int BLOCK, int NEM, bool EOS = true>
__global__ __launch_bounds__(BLOCK) void k_cell_final_3d(Params P, DevScalars* S, CellFinal3ArgsT<R> A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double dx_sound = 1.7976931348623157e308;
  if (!EOS && c >= P.nc) return;  // remapVariables only: no CFL candidate
  if (c < P.nc) {
  int i, j, k, t_;
  divmod32(c, P.nx, t_, i);
  divmod32(t_, P.ny, k, j);
    // total-energy fix-up: the deposit, the 8 nodes' terms in cell node order
    R delta_ec = R(0);
    if (P.etot) {
      R sum = R(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
        const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
        sum += A.b.dek[nid3(P, i + di, j + dj, k + (q >= 4 ? 1 : 0))];
      }
      delta_ec = smax(R(0), sum);
    }
    const double ve = kCartesianEuler<R> ? P.h[0] * P.h[1] * P.h[2] : A.Veuler[c];
    const R c_g = final_cell<R, NEM, EOS>(P, A.b, c, ve, delta_ec);
    if (!EOS) return;
    // the root of the largest squared norm: sqrt is correctly rounded, hence
    // monotone, so it equals the largest root (one root per cell, not per node)
    R n2max = R(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
      const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
      const int dk = q >= 4 ? 1 : 0;
      const long n = nid3(P, i + di, j + dj, k + dk);
      const R ux = A.b.ux[n], uy = A.b.uy[n], uz = A.uz[n];
      const R n2 = ux * ux + uy * uy + uz * uz;
      n2max = (n2 < n2max) ? n2max : n2;
    }
    const R vmax = sqrt(n2max);
    dx_sound = smin(dx_sound, (double)qdiv(A.b.L[c], c_g + vmax));
  }
  block_min_store<BLOCK>(dx_sound, A.b.block_min + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Entry-point kernels of hexahedra: computeGeometricValues (V, L) and the
// CFL candidates, the same arithmetic as k_cell_lagrange_3d's geometry and
// k_cell_final_3d's candidates.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cell_geometry_3d(Params P, DevScalars* S, const double* xnx,
                                                          const double* xny, const double* xnz, double* Vlag,
                                                          double* L) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  int i, j, k, t_;
  divmod32(c, P.nx, t_, i);
  divmod32(t_, P.ny, k, j);
  V3 cn[8];
  hex_nodes(P, xnx, xny, xnz, i, j, k, cn);
  const double Vd = hex_volume(cn);
  Vlag[c] = Vd;
  L[c] = (P.lc == MHG_LC_FACES_OPPOSEES) ? hex_length_faces_opposees(cn) : pow(Vd, 1. / 3.);
}

__global__ __launch_bounds__(256) void k_cfl_3d(Params P, DevScalars* S, CflArgs A) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  double dx_sound = 1.7976931348623157e308;
  if (c < P.nc) {
    double vmax = 0.;
    if (!P.lagonly) {
    int i, j, k, t_;
  divmod32(c, P.nx, t_, i);
  divmod32(t_, P.ny, k, j);
      // the root of the largest squared norm (sqrt is monotone)
      double n2max = 0.;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
        const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
        const int dk = q >= 4 ? 1 : 0;
        const long n = nid3(P, i + di, j + dj, k + dk);
        const double ux = A.ux[n], uy = A.uy[n], uz = A.uz[n];
        const double n2 = ux * ux + uy * uy + uz * uz;
        n2max = (n2 < n2max) ? n2max : n2;
      }
      vmax = sqrt(n2max);
    }
    dx_sound = smin(dx_sound, A.L[c] / (A.cs[c] + vmax));
  }
  block_min_store<256>(dx_sound, A.block_min + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks3(long n, int b) { return (unsigned)((n + b - 1) / b); }

#define MHG3_NEM(ne, F)         \
  do {                          \
    if ((ne) <= 1) F(1);        \
    else if ((ne) <= 2) F(2);   \
    else if ((ne) <= 3) F(3);   \
    else if ((ne) <= 4) F(4);   \
    else F(8);                  \
  } while (0)
#define MHG3_LIM(lim, F)       \
  do {                         \
    if ((lim) == 0) F(0);      \
    else if ((lim) == 1) F(1); \
    else if ((lim) == 2) F(2); \
    else F(kLimG);             \
  } while (0)

template <class R>
void launch_node_lagrange_3d(hipStream_t st, const Params& P, DevScalars* S, const NodeLag3ArgsT<R>& A, bool vel,
                             bool bc, bool pos) {
  dim3 g(nblocks3(P.nn, 256)), b(256);
  if (vel && bc && pos) {
    // tiles of 15 x 15 nodes, node planes split into ranges for >= ~2048 blocks
    const int ntx = (P.nx + 1 + kNl3Tile - 1) / kNl3Tile, nty = (P.ny + 1 + kNl3Tile - 1) / kNl3Tile;
    const int nnz = P.nz + 1;
    const int nchunk = std::max(1, std::min(nnz, (2048 + ntx * nty - 1) / (ntx * nty)));
    const int kchunk = (nnz + nchunk - 1) / nchunk;
    const int nblk = ntx * nty * ((nnz + kchunk - 1) / kchunk);
    hipLaunchKernelGGL((k_node_lagrange_3d_tiled<R>), dim3(nblk), dim3(256), 0, st, P, S, A, ntx, nty, kchunk);
  }
  else if (vel && !bc && !pos) hipLaunchKernelGGL((k_node_lagrange_3d<R, true, false, false>), g, b, 0, st, P, S, A);
  else if (!vel && bc && !pos) hipLaunchKernelGGL((k_node_lagrange_3d<R, false, true, false>), g, b, 0, st, P, S, A);
  else if (!vel && !bc && pos) hipLaunchKernelGGL((k_node_lagrange_3d<R, false, false, true>), g, b, 0, st, P, S, A);
}

template <class R>
void launch_cell_lagrange_3d(hipStream_t st, const Params& P, DevScalars* S, const CellLag3ArgsT<R>& A) {
  dim3 g(nblocks3(P.nc, 256)), b(256);
#define L_(N)                                                                                     \
  if (P.pseudo_centree) hipLaunchKernelGGL((k_cell_lagrange_3d<R, true, N>), g, b, 0, st, P, S, A); \
  else hipLaunchKernelGGL((k_cell_lagrange_3d<R, false, N>), g, b, 0, st, P, S, A)
  MHG3_NEM(P.ne, L_);
#undef L_
}

template <class R, int D, int NEM, int LM>
static void launch_sweep3(hipStream_t st, const Params& P, DevScalars* S, const CellSweep3ArgsT<R>& A, bool first) {
  if (D == 0) {
    const int no = LM == kLimO3 ? 58 : 60;  // output lanes per wavefront
    const long waves = (long)((P.nx + no - 1) / no) * P.ny * P.nz;
    dim3 g(nblocks3(waves * 64, 256)), b(256);
    if (first) hipLaunchKernelGGL((k_cell_sweep_x_3d<R, NEM, true, LM>), g, b, 0, st, P, S, A);
    else hipLaunchKernelGGL((k_cell_sweep_x_3d<R, NEM, false, LM>), g, b, 0, st, P, S, A);
    return;
  }
  constexpr int DD = D == 0 ? 1 : D;
  const int N = (DD == 1) ? P.ny : P.nz, T = (DD == 1) ? P.nz : P.ny;
  const int no = LM == kLimO3 ? 58 : 60;  // output rows per tile
  dim3 g((unsigned)(((P.nx + kTileW - 1) / kTileW) * ((N + no - 1) / no) * T)), b(1024);
  if (first) hipLaunchKernelGGL((k_cell_sweep_tile_3d<R, DD, NEM, true, LM>), g, b, 0, st, P, S, A);
  else hipLaunchKernelGGL((k_cell_sweep_tile_3d<R, DD, NEM, false, LM>), g, b, 0, st, P, S, A);
}

template <class R>
void launch_cell_sweep_3d(hipStream_t st, const Params& P, DevScalars* S, const CellSweep3ArgsT<R>& A, int dir,
                          bool first) {
#define SW_(N, LM)                                                  \
  do {                                                              \
    if (dir == 0) launch_sweep3<R, 0, N, LM>(st, P, S, A, first);   \
    else if (dir == 1) launch_sweep3<R, 1, N, LM>(st, P, S, A, first); \
    else launch_sweep3<R, 2, N, LM>(st, P, S, A, first);            \
  } while (0)
#define SWL_(LM)                     \
  do {                               \
    if (P.ne <= 1) SW_(1, LM);       \
    else if (P.ne <= 2) SW_(2, LM);  \
    else if (P.ne <= 3) SW_(3, LM);  \
    else if (P.ne <= 4) SW_(4, LM);  \
    else SW_(8, LM);                 \
  } while (0)
  if (P.order2 == 2) {
    SWL_(kLimO3);
  } else {
    MHG3_LIM(P.lim, SWL_);
  }
#undef SWL_
#undef SW_
}

template <class R, int D, bool FIRST, bool LAST, int NEM, int LM>
static void launch_dual3(hipStream_t st, const Params& P, DevScalars* S, const NodeSweep3ArgsT<R>& A) {
  if (D == 0) {
    const long waves = (long)((P.nx + 1 + 59) / 60) * (P.ny + 1) * (P.nz + 1);
    dim3 g(nblocks3(waves * 64, 256)), b(256);
    hipLaunchKernelGGL((k_node_sweep_x_3d<R, 0, FIRST, LAST, NEM, LM>), g, b, 0, st, P, S, A);
    return;
  }
  constexpr int DD = D == 0 ? 1 : D;
  const int N = (DD == 1) ? P.ny : P.nz, T = (DD == 1) ? P.nz : P.ny;
  dim3 g((unsigned)(((P.nx + 1 + kTileW - 1) / kTileW) * ((N + 1 + 59) / 60) * (T + 1))), b(1024);
  hipLaunchKernelGGL((k_node_sweep_tile_3d<R, DD, FIRST, LAST, NEM, LM>), g, b, 0, st, P, S, A);
}

template <class R>
void launch_node_sweep_3d(hipStream_t st, const Params& P, DevScalars* S, const NodeSweep3ArgsT<R>& A, int dir,
                          bool first, bool last) {
#define DS_(D, N, LM)                                                                  \
  do {                                                                                 \
    if (first) launch_dual3<R, D, true, false, N, LM>(st, P, S, A);                    \
    else if (last) launch_dual3<R, D, false, true, N, LM>(st, P, S, A);                \
    else launch_dual3<R, D, false, false, N, LM>(st, P, S, A);                         \
  } while (0)
#define DSD_(N, LM)                   \
  do {                                \
    if (dir == 0) DS_(0, N, LM);      \
    else if (dir == 1) DS_(1, N, LM); \
    else DS_(2, N, LM);               \
  } while (0)
#define DSL_(LM)                      \
  do {                                \
    if (P.ne <= 1) DSD_(1, LM);       \
    else if (P.ne <= 2) DSD_(2, LM);  \
    else if (P.ne <= 3) DSD_(3, LM);  \
    else if (P.ne <= 4) DSD_(4, LM);  \
    else DSD_(8, LM);                 \
  } while (0)
  MHG3_LIM(P.lim, DSL_);
#undef DSL_
#undef DSD_
#undef DS_
}

template <class R>
void launch_cell_final_3d(hipStream_t st, const Params& P, DevScalars* S, const CellFinal3ArgsT<R>& A, int dt_mode,
                          bool eos) {
  constexpr int B = 256;
  dim3 g(nblocks3(P.nc, B)), b(B);
#define F_(N) hipLaunchKernelGGL((k_cell_final_3d<R, B, N>), g, b, 0, st, P, S, A)
#define FN_(N) hipLaunchKernelGGL((k_cell_final_3d<R, B, N, false>), g, b, 0, st, P, S, A)
  if (eos) MHG3_NEM(P.ne, F_);
  else MHG3_NEM(P.ne, FN_);
#undef FN_
#undef F_
  if (dt_mode > 0) launch_dt_finalize(st, P, S, A.b.block_min, dt_mode == 2);
}

void launch_cell_geometry_3d(hipStream_t st, const Params& P, DevScalars* S, const double* xnx,
                             const double* xny, const double* xnz, double* Vlag, double* L) {
  hipLaunchKernelGGL(k_cell_geometry_3d, dim3(nblocks3(P.nc, 256)), dim3(256), 0, st, P, S, xnx, xny, xnz, Vlag, L);
}

void launch_cfl_3d(hipStream_t st, const Params& P, DevScalars* S, const CflArgs& A) {
  hipLaunchKernelGGL(k_cfl_3d, dim3(cell_final_blocks(P)), dim3(256), 0, st, P, S, A);
}

#define MHG3_INSTANTIATE(R)                                                                                   \
  template void launch_node_lagrange_3d<R>(hipStream_t, const Params&, DevScalars*, const NodeLag3ArgsT<R>&, \
                                           bool, bool, bool);                                                \
  template void launch_cell_lagrange_3d<R>(hipStream_t, const Params&, DevScalars*, const CellLag3ArgsT<R>&); \
  template void launch_cell_sweep_3d<R>(hipStream_t, const Params&, DevScalars*, const CellSweep3ArgsT<R>&,   \
                                        int, bool);                                                          \
  template void launch_node_sweep_3d<R>(hipStream_t, const Params&, DevScalars*, const NodeSweep3ArgsT<R>&,   \
                                        int, bool, bool);                                                    \
  template void launch_cell_final_3d<R>(hipStream_t, const Params&, DevScalars*, const CellFinal3ArgsT<R>&,   \
                                        int, bool);
MHG3_INSTANTIATE(double)
MHG3_INSTANTIATE(float)
#undef MHG3_INSTANTIATE

}  // namespace mhg
