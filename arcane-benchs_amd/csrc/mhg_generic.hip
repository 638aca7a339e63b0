// mhg_generic.hip -- the generic cell sweep of the remap for gfx950 (FP64):
// the pente-borne fluxes (projection-pente-borne) and the pure-neighbourhood
// limiter switch (projection-pente-borne-mixte), RemapADIService.cc:319-431
// (computeGradPhiCell's generic loop), 549-581 (face quantities), 733-872
// (computeUremap), UtilesRemap.cc:73-814 (INTY, computeY0, computexgxd,
// computeygyd, computeFluxPP, computeFluxPPPure).  2D and 3D, one thread per
// cell, two kernels per sweep:
//   k_generic_grad    limited gradient of each variable; with pente-borne,
//                     the integrated flux leaving the cell through its back
//                     face (dar) and through its front face (dav);
//                     without, the gradient itself (into dar)
//   k_generic_update  face fluxes (pente-borne: dav of the back cell minus dar
//                     of the front cell; else the upwind face value), the new
//                     remap vector, non-negative masses and energies, phi
//                     (pente-borne form), the mixed / pure code and the dual
//                     mass contributions of the node sweep.
// These options are off by default and host-only in the reference; the
// kernels favour a direct reading of its loops over speed.
#include "mhg_common.hpp"
#include "mhg_kernels.hpp"

namespace mhg {
namespace {

// Arcane's math::min / math::max.
__device__ inline double amin(double a, double b) { return (a < b) ? a : b; }
__device__ inline double amax(double a, double b) { return (a < b) ? b : a; }

template <int DIM>
__device__ inline long gnode(const Params& P, int i, int j, int k) {
  return DIM == 2 ? i + (long)(P.nx + 1) * j : i + (long)(P.nx + 1) * (j + (long)(P.ny + 1) * k);
}

// Geometry of one cell for a sweep along D: its centre, the centres, normal
// velocities and lengths along D of its back and front faces
// (computeFaceQuantitesForRemap, PrepareRemap.cc:15-123; cellCoordFromNodes).
struct CellGeo {
  double cc[3], fb[3], ff[3];
  double ub, uf, lb, lf;
};

template <int DIM, int D>
__device__ inline void face_nodes(const Params& P, int i, int j, int k, long* n) {
  if (DIM == 2) {
    n[0] = gnode<2>(P, i, j, 0);
    n[1] = D == 0 ? gnode<2>(P, i, j + 1, 0) : gnode<2>(P, i + 1, j, 0);
  } else if (D == 0) {
    n[0] = gnode<3>(P, i, j, k); n[1] = gnode<3>(P, i, j + 1, k);
    n[2] = gnode<3>(P, i, j + 1, k + 1); n[3] = gnode<3>(P, i, j, k + 1);
  } else if (D == 1) {
    n[0] = gnode<3>(P, i, j, k); n[1] = gnode<3>(P, i + 1, j, k);
    n[2] = gnode<3>(P, i + 1, j, k + 1); n[3] = gnode<3>(P, i, j, k + 1);
  } else {
    n[0] = gnode<3>(P, i, j, k); n[1] = gnode<3>(P, i + 1, j, k);
    n[2] = gnode<3>(P, i + 1, j + 1, k); n[3] = gnode<3>(P, i, j + 1, k);
  }
}

template <int DIM, int D>
__device__ inline void face_geo(const Params& P, const GenericSweepArgs& A, int i, int j, int k, double* fc,
                                double* u, double* len) {
  const double* X[3] = {A.x, A.y, A.z};
  const double* U[3] = {A.ux, A.uy, A.uz};
  constexpr int NB = DIM == 2 ? 2 : 4;
  const double w = DIM == 2 ? 0.5 : 0.25;
  long n[4];
  face_nodes<DIM, D>(P, i, j, k, n);
  for (int d = 0; d < 3; ++d) {
    double s = 0.;
    if (d < DIM)
      for (int q = 0; q < NB; ++q) s += w * X[d][n[q]];
    fc[d] = s;
  }
  double vm = 0.;
  for (int q = 0; q < NB; ++q) vm += U[D][n[q]];
  *u = w * vm;
  if (DIM == 2) {
    const double* T = X[1 - D];  // the coordinate across the face
    *len = fabs(T[n[1]] - T[n[0]]);
  } else {
    V3 x[4];
    for (int q = 0; q < 4; ++q) x[q] = V3{X[0][n[q]], X[1][n[q]], X[2][n[q]]};
    const V3 t = v3mul(0.5, v3cross(v3sub(x[2], x[0]), v3sub(x[3], x[1])));
    *len = fabs(D == 0 ? t.x : (D == 1 ? t.y : t.z));
  }
}

template <int DIM, int D>
__device__ inline void cell_geo(const Params& P, const GenericSweepArgs& A, int i, int j, int k, CellGeo* g) {
  const double* X[3] = {A.x, A.y, A.z};
  constexpr int NB = DIM == 2 ? 4 : 8;
  const double w = DIM == 2 ? 0.25 : 0.125;
  long nd[8];
  for (int q = 0; q < NB; ++q) {
    const int di = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
    const int dj = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
    const int dk = q >= 4 ? 1 : 0;
    nd[q] = gnode<DIM>(P, i + di, j + dj, k + dk);
  }
  for (int d = 0; d < 3; ++d) {
    double s = 0.;
    if (d < DIM)
      for (int q = 0; q < NB; ++q) s += X[d][nd[q]];
    g->cc[d] = w * s;
  }
  face_geo<DIM, D>(P, A, i, j, k, g->fb, &g->ub, &g->lb);
  face_geo<DIM, D>(P, A, i + (D == 0), j + (D == 1), k + (D == 2), g->ff, &g->uf, &g->lf);
}

// Width along D (computeGradPhiFace, HCELL_BY_CELLS, RemapADIService.cc:
// 223-255): distances from the centre to the back and front face centres,
// a face on the mesh boundary left out.
__device__ inline double cell_width(const CellGeo& g, int s, int N) {
  auto dist = [&](const double* f) {
    const double a = f[0] - g.cc[0], b = f[1] - g.cc[1], c = f[2] - g.cc[2];
    return sqrt(a * a + b * b + c * c);
  };
  double h = 0.;
  if (s > 0) h += dist(g.fb);
  if (s < N - 1) h += dist(g.ff);
  return h;
}

// ---- pente-borne (UtilesRemap.cc:73-603) -----------------------------------
__device__ inline double pb_inty(double X, double x0, double y0, double x1, double y1, double thr) {
  double flux = 0.;
  if (fabs(x1 - x0) > thr) {
    const double eta = amin(amax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// computeY0: the monotonicity thresholds of limiter lim.
__device__ inline void pb_y0(int lim, double yp, double ym, double h0, double hp, double hm, double* y0plus,
                             double* y0moins) {
  double a = 0., b = 0.;
  if (lim == MHG_LIM_MINMODG || lim == MHG_LIM_MINMOD) {
    a = yp;
    b = ym;
  } else if (lim == MHG_LIM_SUPERBEEG || lim == MHG_LIM_SUPERBEE) {
    a = ((h0 + hm) * yp + h0 * ym) / (2 * h0 + hm);
    b = ((h0 + hp) * ym + h0 * yp) / (2 * h0 + hp);
  } else if (lim == MHG_LIM_VANLEERG || lim == MHG_LIM_VANLEER) {
    const double lo = amin(yp, ym), hi = amax(yp, ym);
    double xp = (h0 * h0 + 3 * h0 * hm + 2 * hm * hm) * yp;
    double xm = (h0 * h0 + 3 * h0 * hp + 2 * hp * hp) * ym;
    xp += (h0 * h0 - h0 * hp - 2 * hp * hp + 2 * h0 * hm) * ym;
    xm += (h0 * h0 - h0 * hm - 2 * hm * hm + 2 * h0 * hp) * yp;
    xp /= (2 * h0 * h0 + 5 * h0 * hm + 2 * hm * hm - h0 * hp - 2 * hp * hp);
    xm /= (2 * h0 * h0 + 5 * h0 * hp + 2 * hp * hp - h0 * hm - 2 * hm * hm);
    a = amin(amax(xp, lo), hi);
    b = amin(amax(xm, lo), hi);
  } else if (lim == MHG_LIM_ULTRABEEG) {
    a = (yp + ym) / 2.;
    b = (yp + ym) / 2.;
  } else if (lim == MHG_LIM_ARITHMETICG) {
    a = ((h0 + hm + hp) * yp + h0 * ym) / (2 * h0 + hm + hp);
    b = ((h0 + hm + hp) * ym + h0 * yp) / (2 * h0 + hm + hp);
  }
  *y0plus = a;
  *y0moins = b;
}

// One variable's integrated flux over ppv next to the back face (type 0)
// or the front face (type 1): computexgxd, computeygyd, the three INTY
// pieces, the zero floor and formula 16.
__device__ inline double pb_var_flux(int type, int lim, double ppv, double h0, double hp, double hm, double y0,
                                     double yp, double ym, double grady, double thr) {
  double y0plus, y0moins;
  pb_y0(lim, yp, ym, h0, hp, hm, &y0plus, &y0moins);
  double xplus = 1.;
  if (fabs(y0plus - yp) > thr) xplus = (y0 - yp) / (y0plus - yp) - 1. / 2.;
  double xmoins = 1.;
  if (fabs(y0moins - ym) > thr) xmoins = (y0 - ym) / (y0moins - ym) - 1. / 2.;
  const double xd = +h0 * amin(amax(xplus, -1. / 2.), 1. / 2.);
  const double xg = -h0 * amin(amax(xmoins, -1. / 2.), 1. / 2.);
  const double xtd = y0 + h0 / 2 * grady;
  const double xtg = y0 - h0 / 2 * grady;
  const double yd = amin(amax(xtd, amin(yp, 2. * y0moins - ym)), amax(yp, 2. * y0moins - ym));
  const double yg = amin(amax(xtg, amin(ym, 2. * y0plus - yp)), amax(ym, 2. * y0plus - yp));
  double f1, f1m, f2, f2m, f3, f3m;
  if (type == 0) {
    f1m = pb_inty(-h0 / 2. + ppv, -h0 / 2., ym, xg, yg, thr);
    f1 = pb_inty(-h0 / 2., -h0 / 2., ym, xg, yg, thr);
    f2m = pb_inty(-h0 / 2. + ppv, xg, yg, xd, yd, thr);
    f2 = pb_inty(-h0 / 2., xg, yg, xd, yd, thr);
    f3m = pb_inty(-h0 / 2. + ppv, xd, yd, h0 / 2., yp, thr);
    f3 = pb_inty(-h0 / 2., xd, yd, h0 / 2., yp, thr);
  } else {
    f1 = pb_inty(h0 / 2. - ppv, -h0 / 2., ym, xg, yg, thr);
    f1m = pb_inty(h0 / 2., -h0 / 2., ym, xg, yg, thr);
    f2 = pb_inty(h0 / 2. - ppv, xg, yg, xd, yd, thr);
    f2m = pb_inty(h0 / 2., xg, yg, xd, yd, thr);
    f3 = pb_inty(h0 / 2. - ppv, xd, yd, h0 / 2., yp, thr);
    f3m = pb_inty(h0 / 2., xd, yd, h0 / 2., yp, thr);
  }
  double F = smax(((f1m - f1) + (f2m - f2)) + (f3m - f3), 0.);
  if (((yp - y0) * (ym - y0)) >= 0.) F = y0 * ppv;
  return F;
}

constexpr int kNV = 3 * kMaxEnv + 1;  // live variables: V_k, m_k, E_k, q V

// phi of variable v of cell x: pente-borne keeps it; otherwise U / den, den
// the Lagrangian volume in the first sweep, the sum of the environment
// volumes after (computeUremap_PBorn0).
__device__ inline double phi_of(const Params& P, const GenericSweepArgs& A, bool first, long x, int v, double den) {
  return P.pb ? A.phi[v * P.nc + x] : A.Uin[v * P.nc + x] / den;
}
__device__ inline double den_of(const Params& P, const GenericSweepArgs& A, bool first, long x) {
  if (P.pb) return 1.;
  if (first) return A.Vlag[x];
  double s = 0.;
  for (int k = 0; k < P.ne; ++k) s += A.Uin[k * P.nc + x];
  return s;
}

template <int DIM, int D>
__device__ inline void cell_ijk(const Params& P, long c, int* i, int* j, int* k, int* s, int* N, long* stride) {
  *i = (int)(c % P.nx);
  *j = (int)((c / P.nx) % P.ny);
  *k = DIM == 3 ? (int)(c / ((long)P.nx * P.ny)) : 0;
  *s = D == 0 ? *i : (D == 1 ? *j : *k);
  *N = D == 0 ? P.nx : (D == 1 ? P.ny : P.nz);
  *stride = D == 0 ? 1 : (D == 1 ? (long)P.nx : (long)P.nx * P.ny);
}

template <int DIM, int D>
__global__ __launch_bounds__(256) void k_generic_grad(Params P, DevScalars* S, GenericSweepArgs A, int first) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  int i, j, k, s, N;
  long stride;
  cell_ijk<DIM, D>(P, c, &i, &j, &k, &s, &N, &stride);
  const bool hb = s > 0, hf = s < N - 1;
  const long b = hb ? c - stride : c, f = hf ? c + stride : c;
  CellGeo gc, gb, gf;
  cell_geo<DIM, D>(P, A, i, j, k, &gc);
  cell_geo<DIM, D>(P, A, i - (D == 0 && hb), j - (D == 1 && hb), k - (D == 2 && hb), &gb);
  cell_geo<DIM, D>(P, A, i + (D == 0 && hf), j + (D == 1 && hf), k + (D == 2 && hf), &gf);
  const double h0 = cell_width(gc, s, N);
  const double hm = hb ? cell_width(gb, s - 1, N) : h0;
  const double hp = hf ? cell_width(gf, s + 1, N) : h0;
  const double dxb = gc.cc[D] - gb.cc[D], dxf = gf.cc[D] - gc.cc[D];
  // -mixte: a cell whose neighbourhood along the sweep is pure in one
  // environment uses the pure limiter when projection-limiteur-pure-id is 1
  bool vp = false;
  if (P.mixte) {
    const int ec = A.est[c], eb = A.est[b], ef = A.est[f];
    vp = ec != -2 && eb != -2 && ef != -2 && ec == ef && ec == eb;
  }
  const int lim = (P.pure_id == 1 && vp) ? P.pure_id : P.lim;
  const bool first_b = first != 0;
  const double denc = den_of(P, A, first_b, c), denb = den_of(P, A, first_b, b), denf = den_of(P, A, first_b, f);
  const int nv = P.nvu;
  const long NC = P.nc;
  double g[kNV], y0[kNV], yp[kNV], ym[kNV];
  for (int v = 0; v < nv; ++v) {
    y0[v] = phi_of(P, A, first_b, c, v, denc);
    ym[v] = phi_of(P, A, first_b, b, v, denb);
    yp[v] = phi_of(P, A, first_b, f, v, denf);
    const double gbv = hb ? (y0[v] - ym[v]) / dxb : 0.;
    const double gfv = hf ? (yp[v] - y0[v]) / dxf : 0.;
    g[v] = lim < MHG_LIM_MINMODG ? limited_cell_grad(lim, gbv, gfv)
                                 : limiter_g<double>(lim, gfv, gbv, yp[v], ym[v], h0, hp, hm);
  }
  if (!P.pb) {
    for (int v = 0; v < nv; ++v) A.dar[v * NC + c] = g[v];
    return;
  }
  // computeFluxPP / computeFluxPPPure for the back face (outgoing velocity
  // -u_back, type 0) and the front face (u_front, type 1)
  const double thr = P.remap_threshold, dt = S->dt;
  const int ylim = vp ? P.pure_id : P.lim;
  const int ne = P.ne;
  for (int t = 0; t < 2; ++t) {
    const double vout = t == 0 ? -1.0 * gc.ub : 1.0 * gc.uf;
    const double ppv = 0.5 * (vout + fabs(vout)) * dt;
    double F[kNV];
    for (int v = 0; v < nv; ++v) F[v] = pb_var_flux(t, ylim, ppv, h0, hp, hm, y0[v], yp[v], ym[v], g[v], thr);
    if (P.pb_fix == 1) {
      double sfv = 0.;
      for (int q = 0; q < ne; ++q) sfv += F[q];
      if (fabs(sfv) > thr) {
        for (int q = 0; q < ne; ++q) {
          if (!vp) F[ne + q] = (F[ne + q] / sfv) * F[q];
          F[2 * ne + q] = (F[2 * ne + q] / sfv) * F[ne + q];
        }
      } else {
        for (int v = 0; v < nv; ++v) F[v] = 0.;
      }
    } else if (P.pb_fix == 2) {
      for (int q = 0; q < ne; ++q) {
        if (!vp) F[ne + q] = y0[ne + q] * F[q];
        F[2 * ne + q] = y0[2 * ne + q] * F[ne + q];
      }
    }
    if (ppv == 0.)
      for (int v = 0; v < nv; ++v) F[v] = 0.;
    double* out = t == 0 ? A.dar : A.dav;
    for (int v = 0; v < nv; ++v) out[v * NC + c] = F[v];
  }
}

template <int DIM, int D>
__global__ __launch_bounds__(256) void k_generic_update(Params P, DevScalars* S, GenericSweepArgs A, int first) {
  if (S->stop || S->err) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.nc) return;
  int i, j, k, s, N;
  long stride;
  cell_ijk<DIM, D>(P, c, &i, &j, &k, &s, &N, &stride);
  const bool hb = s > 0, hf = s < N - 1;
  const long b = hb ? c - stride : c, f = hf ? c + stride : c;
  CellGeo gc;
  cell_geo<DIM, D>(P, A, i, j, k, &gc);
  const int nv = P.nvu, ne = P.ne;
  const long NC = P.nc;
  const double dt = S->dt;
  double pfb[kNV], pff[kNV];
  double fb_fac, ff_fac;  // face flux factor: n.e_D L (pente-borne), n.e_D u L dt (else)
  if (P.pb) {
    for (int v = 0; v < nv; ++v) {
      pfb[v] = hb ? A.dav[v * NC + b] - A.dar[v * NC + c] : 0.;
      pff[v] = hf ? A.dav[v * NC + c] - A.dar[v * NC + f] : 0.;
    }
    fb_fac = -1.0 * gc.lb;
    ff_fac = 1.0 * gc.lf;
  } else {
    // upwind face values phi_up + order2 ((x_f - x_up) . e_D) grad_up
    // (computeUpwindFaceQuantitiesForProjection_PBorn0_O2, 658-720)
    CellGeo gb, gf;
    cell_geo<DIM, D>(P, A, i - (D == 0 && hb), j - (D == 1 && hb), k - (D == 2 && hb), &gb);
    cell_geo<DIM, D>(P, A, i + (D == 0 && hf), j + (D == 1 && hf), k + (D == 2 && hf), &gf);
    const bool first_b = first != 0;
    const double denc = den_of(P, A, first_b, c), denb = den_of(P, A, first_b, b), denf = den_of(P, A, first_b, f);
    const double order2 = (double)P.order2;
    const bool upb = hb && gc.ub * (gc.cc[D] - gb.cc[D]) > 0.0;  // back face: upwind is the back cell
    const bool upf = hf && gc.uf * (gf.cc[D] - gc.cc[D]) > 0.0;  // front face: upwind is this cell
    const double dotb = upb ? gc.fb[D] - gb.cc[D] : gc.fb[D] - gc.cc[D];
    const double dotf = upf ? gc.ff[D] - gc.cc[D] : gc.ff[D] - gf.cc[D];
    for (int v = 0; v < nv; ++v) {
      const double pc = phi_of(P, A, first_b, c, v, denc), gcv = A.dar[v * NC + c];
      pfb[v] = !hb ? 0. : (upb ? phi_of(P, A, first_b, b, v, denb) + order2 * dotb * A.dar[v * NC + b]
                              : pc + order2 * dotb * gcv);
      pff[v] = !hf ? 0. : (upf ? pc + order2 * dotf * gcv
                              : phi_of(P, A, first_b, f, v, denf) + order2 * dotf * A.dar[v * NC + f]);
    }
    fb_fac = -1.0 * gc.ub * gc.lb * dt;
    ff_fac = 1.0 * gc.uf * gc.lf * dt;
  }
  double Un[kNV], dual[kNV];
  for (int v = 0; v < nv; ++v) {
    const double flux_b = fb_fac * pfb[v], flux_f = ff_fac * pff[v];
    Un[v] = A.Uin[v * NC + c] - (flux_b + flux_f);
    dual[v] = 0.5 * flux_b * -1.0 + 0.5 * flux_f * 1.0;
  }
  for (int q = 0; q < ne; ++q) {
    if (Un[ne + q] < 0.) Un[ne + q] = 0.;
    if (Un[2 * ne + q] < 0.) Un[2 * ne + q] = 0.;
  }
  double sv = 0.;
  for (int q = 0; q < ne; ++q) sv += Un[q];
  int cnt = 0, last = -1;
  for (int q = 0; q < ne; ++q) {
    const double fq = Un[q] / sv;
    if (P.pb) {
      A.phi[q * NC + c] = fq;
      A.phi[(ne + q) * NC + c] = (Un[q] != 0.) ? Un[ne + q] / Un[q] : 0.;
      A.phi[(2 * ne + q) * NC + c] = (Un[ne + q] != 0.) ? Un[2 * ne + q] / Un[ne + q] : 0.;
    }
    if (fq > 0.) {
      cnt++;
      last = q;
    }
  }
  A.est[c] = cnt > 1 ? -2 : last;
  for (int v = 0; v < nv; ++v) A.Uout[v * NC + c] = Un[v];
  const double onc = DIM == 2 ? 0.5 : 0.25;  // DualPhaseRemap.cc:147-188
  for (int q = 0; q < ne; ++q) A.contrib[q * NC + c] = onc * dual[ne + q];
}

template <int DIM, int D>
void launch_dir(hipStream_t st, const Params& P, DevScalars* S, const GenericSweepArgs& A, bool first) {
  const unsigned nb = (unsigned)((P.nc + 255) / 256);
  k_generic_grad<DIM, D><<<nb, 256, 0, st>>>(P, S, A, first ? 1 : 0);
  k_generic_update<DIM, D><<<nb, 256, 0, st>>>(P, S, A, first ? 1 : 0);
}

}  // namespace

void launch_generic_sweep(hipStream_t st, const Params& P, DevScalars* S, const GenericSweepArgs& A, int dir,
                          bool first) {
  if (P.dim == 2) {
    if (dir == 0) launch_dir<2, 0>(st, P, S, A, first);
    else launch_dir<2, 1>(st, P, S, A, first);
  } else {
    if (dir == 0) launch_dir<3, 0>(st, P, S, A, first);
    else if (dir == 1) launch_dir<3, 1>(st, P, S, A, first);
    else launch_dir<3, 2>(st, P, S, A, first);
  }
}

}  // namespace mhg
