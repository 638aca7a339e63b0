// mahyco_oracle.cpp -- CPU FP64 oracle of MaHyCo's Lagrange + Euler-ADI
// remap time step.
//
// TEST INFRASTRUCTURE ONLY: this is the checker the GPU path is compared
// against (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).
// It is never linked into or called by the product library.
//
// It restates, function by function, the arithmetic of the reference files
// (paths relative to the reference root):
//   mahyco/src/MahycoModule.cc, MahycoAnnexe.cc, PrepareRemap.cc,
//   mahyco/src/Remap/RemapADIService.{h,cc}, DualPhaseRemap.cc, RemapADIFinal.cc,
//   mahyco/src/eos/perfectgas/PerfectGasEOSService.cc,
//   mahyco/src/eos/stiffenedgas/StiffenedGasEOSService.cc,
//   mahyco/src/casTest/SOD/SODService.cc,
//   mahyco/src/accenv/AccEnvDefaultService.cc:240-349, MultiEnvUtils.h:303-440.
// Storage follows the reference's model: one global array per material
// variable plus per-environment values for cells holding >= 2 environments;
// a cell holding one environment has no separate environment value (its
// environment value *is* the global value, MultiEnvUtils.h:303-440).
//
// Conventions the reference leaves to Arcane (not pinned by the reference,
// fixed here and mirrored by the GPU path):
//   * node coordinates x_i = origin + i * (length / n);
//   * node->cell order is ascending cell id; cell->node order is Arcane's
//     quad4/hexa8 order; face->node order is ascending node id in 2D and
//     chosen so the face normal is +e_dir in 3D;
//   * never-written arrays start at zero;
//   * the time-loop manager increments time by dt before the compute-loop
//     entry points of each iteration (mho_begin_iteration here).
// The whole file is compiled with -ffp-contract=off: every a*b+c is a
// rounded multiply followed by a rounded add, as written.

#include "mahyco_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct R3 {
  double x, y, z;
  R3() : x(0.), y(0.), z(0.) {}
  R3(double a, double b, double c) : x(a), y(b), z(c) {}
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline R3 operator+(const R3& a, const R3& b) { return R3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline R3 operator-(const R3& a, const R3& b) { return R3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline R3 operator*(double s, const R3& a) { return R3(s * a.x, s * a.y, s * a.z); }
inline R3& operator+=(R3& a, const R3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline double dot(const R3& a, const R3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline R3 cross(const R3& a, const R3& b) {
  return R3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double normL2(const R3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline R3 absolute(const R3& a) { return R3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }

// std::max/std::min semantics (used by the reference's limiters and clamps).
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double smin(double a, double b) { return (b < a) ? b : a; }

// Limiters, RemapADIService.h:260-286.
inline double limiter(int id, double r) {
  switch (id) {
    case 0: return smax(0.0, smin(1.0, r));                                  // MinMod
    case 1: return smax(0.0, smax(smin(2.0 * r, 1.0), smin(r, 2.0)));       // SuperBee
    case 2: return (r <= 0.0 ? 0.0 : 2.0 * r / (1.0 + r));                  // VanLeer
    default: return 0.0;                                                    // DefaultO1
  }
}

// Generalised limiters, fluxLimiterG (UtilesRemap.cc:127-173): the slope of
// a cell (or dual cell) of width h0 between neighbours of widths hp (front)
// and hm (back), from the front and back gradients gp, gm and the front and
// back values yp, ym.  Zero at a slope break (gp gm < 0); otherwise the
// limiter's candidate slope, bounded by the simple-slope limit
// min(|gp (h0+hp)/h0|, |gm (h0+hm)/h0|), with the sign of yp - ym.
inline double flux_limiter_g(int id, double gp, double gm, double yp, double ym, double h0, double hp,
                             double hm) {
  if (gp * gm < 0.0) return 0.;
  const bool up = (yp - ym) > 0.;
  double grady = 0.;
  if (id == MHG_LIM_MINMODG) {         // formula 9c
    grady = up ? smin(std::fabs(gp), std::fabs(gm)) : -smin(std::fabs(gp), std::fabs(gm));
  } else if (id == MHG_LIM_SUPERBEEG) {  // formula 9g
    grady = up ? smax(std::fabs(gp), std::fabs(gm)) : -smax(std::fabs(gp), std::fabs(gm));
  } else if (id == MHG_LIM_VANLEERG) {   // formula 9e
    const double lp = (h0 / 2. + hp) / (h0 + hp + hm);
    const double lm = (h0 / 2. + hm) / (h0 + hp + hm);
    grady = (lp * gp + lm * gm) != 0. ? gp * gm / (lp * gp + lm * gm) : 0.;
  } else if (id == MHG_LIM_ULTRABEEG) {
    grady = (yp - ym) / h0;
  } else if (id == MHG_LIM_ARITHMETICG) {
    const double lp = (h0 / 2. + hp) / (h0 + hp + hm);
    const double lm = (h0 / 2. + hm) / (h0 + hp + hm);
    grady = lm * gp + lp * gm;
  }
  // simple-slope limit (formula 10)
  const double gMp = gp * (h0 + hp) / h0;
  const double gMm = gm * (h0 + hm) / h0;
  const double gM = smin(std::fabs(gMp), std::fabs(gMm));
  return up ? smin(std::fabs(gM), std::fabs(grady)) : -smin(std::fabs(gM), std::fabs(grady));
}

// Learned slope of one cell and variable (extension, MHG_LIM_LEARNED): the
// central slope alpha (g_b + g_f) / 2, kept if the face values it gives,
// phi + g db at the back face and phi + g df at the front face, lie within
// [min, max] of the cell's phi and its existing neighbours' along the sweep;
// otherwise, and for alpha < 0, the minmod slope (RemapADIService.cc:497-506
// with Lim = minmod).
// *used tells whether a non-zero learned slope was kept.
inline double learned_grad(double alpha, double phim, double phi, double phip, bool back_in, bool front_in,
                           double gb, double gf, double db, double df, bool* used = nullptr) {
  double gcl = 0.;
  if (gb != 0.) gcl += 0.5 * (limiter(0, gf / gb) * gb);
  if (gf != 0.) gcl += 0.5 * (limiter(0, gb / gf) * gf);
  if (used) *used = false;
  if (alpha < 0.) return gcl;
  const double gc = alpha * (0.5 * (gb + gf));
  double mn = phi, mx = phi;
  if (back_in) {
    mn = smin(mn, phim);
    mx = smax(mx, phim);
  }
  if (front_in) {
    mn = smin(mn, phip);
    mx = smax(mx, phip);
  }
  const double vb = phi + gc * db, vf = phi + gc * df;
  const bool ok = vb >= mn && vb <= mx && vf >= mn && vf <= mx;
  if (used) *used = ok && gc != 0.;
  return ok ? gc : gcl;
}

// Consistent environment slopes of the learned mode (extension): in a mixed
// cell (two or more environments with phi_V > 0) with alpha >= 0, the mass
// and energy slopes of environment k follow its volume slope, scaled by the
// environment's mass and energy per unit of its volume in the cell:
//   g_m = (phi_m / phi_V) g_V,  g_E = (phi_E / phi_V) g_V,
// when a non-zero learned volume slope was kept and phi_V exceeds the remap
// threshold (so the ratios stay finite); otherwise they take the classical
// (minmod) slope, so that alpha = 1e30 (every non-zero learned slope
// rejected) is still minmod.  The volume slope of a trace environment (0 <
// phi_V <= remap threshold) is classical too, so that its three slopes are
// the classical method's (a learned volume slope beside classical mass and
// energy slopes moves volume without mass, or mass without volume).  A face value of an
// environment's mass or energy is then its volume face value times the
// upwind cell's density or density times specific energy, so no face carries
// mass or energy without volume (the source of sliver cells of extreme
// specific energy when the three slopes are chosen independently).
// phi, g: the cell's nv values and slopes; gcl: the classical slopes.
inline void consistent_env_slopes(int ne, const double* phi, double* g, const double* gcl, const bool* used_v,
                                  double threshold) {
  for (int k = 0; k < ne; ++k) {
    const int vm = ne + k, ve = 2 * ne + k;
    if (used_v[k] && phi[k] > threshold) {
      g[vm] = (phi[vm] / phi[k]) * g[k];
      g[ve] = (phi[ve] / phi[k]) * g[k];
    } else {
      g[vm] = gcl[vm];
      g[ve] = gcl[ve];
    }
  }
}

// Limiter of the classical branches: the options' limiter, minmod in the
// learned mode (its fallback and its dual sweeps).
inline bool is_learned(const mhg_options& o) {
  return o.projection_limiteur_id == MHG_LIM_LEARNED || o.projection_limiteur_id == MHG_LIM_LEARNED_ALL;
}
inline int classical_limiter(const mhg_options& o) {
  return is_learned(o) ? MHG_LIM_MINMOD : o.projection_limiteur_id;
}
inline bool is_generalised(const mhg_options& o) {
  return o.projection_limiteur_id >= MHG_LIM_MINMODG && o.projection_limiteur_id <= MHG_LIM_ULTRABEEG;
}
// projection-pente-borne / -mixte: the generic gradient loop of
// computeGradPhiCell (RemapADIService.cc:319-431).
inline bool is_pb(const mhg_options& o) { return o.projection_pente_borne != 0; }
inline bool generic_grad(const mhg_options& o) {
  return o.ordre_projection > 1 && (o.projection_pente_borne || o.projection_pente_borne_mixte);
}
inline bool has_limited_gradients(const mhg_options& o) {
  return o.projection_limiteur_id < MHG_LIM_MINMODG || is_learned(o) || is_generalised(o);
}

// compute_eint, MahycoModule.cc:1630-1654.
inline double compute_eint(bool pseudo_centree, double adiabatic_cst, double pseudo_viscosity_n,
                           double pseudo_viscosity, double density_n, double density,
                           double pressure, double internal_energy_n) {
  double pseudo = 0.;
  if (pseudo_centree &&
      ((pseudo_viscosity_n + pseudo_viscosity) * (1.0 / density - 1.0 / density_n) < 0.))
    pseudo = 0.5 * (pseudo_viscosity + pseudo_viscosity_n);
  if (!pseudo_centree && (pseudo_viscosity * (1.0 / density - 1.0 / density_n) < 0.))
    pseudo = pseudo_viscosity;
  double denom = 1 + 0.5 * (adiabatic_cst - 1.0) * density * (1.0 / density - 1.0 / density_n);
  double numer = internal_energy_n - (0.5 * pressure + pseudo) * (1.0 / density - 1.0 / density_n);
  return numer / denom;
}

// Pseudo-viscosity, MahycoModule.cc:517-528 (qcoef = (gamma+1)/2 in the reference).
inline double viscosity(double qcoef, double tau, double L, double c, double divu) {
  return 1. / tau * (-0.5 * L * c * divu + qcoef * L * L * divu * divu);
}

// 3D corner normals, MahycoModule.cc:2195-2258.
void computeCQs(const R3 nc[8], const R3 fc[6], R3 out[8]) {
  const R3 c0 = fc[0], c1 = fc[1], c2 = fc[2], c3 = fc[3], c4 = fc[4], c5 = fc[5];
  const R3 n1a04 = 0.5 * cross(nc[0] - c0, nc[3] - c0);
  const R3 n1a03 = 0.5 * cross(nc[3] - c0, nc[2] - c0);
  const R3 n1a02 = 0.5 * cross(nc[2] - c0, nc[1] - c0);
  const R3 n1a01 = 0.5 * cross(nc[1] - c0, nc[0] - c0);
  const R3 n2a05 = 0.5 * cross(nc[0] - c1, nc[4] - c1);
  const R3 n2a12 = 0.5 * cross(nc[4] - c1, nc[7] - c1);
  const R3 n2a08 = 0.5 * cross(nc[7] - c1, nc[3] - c1);
  const R3 n2a04 = 0.5 * cross(nc[3] - c1, nc[0] - c1);
  const R3 n3a01 = 0.5 * cross(nc[0] - c2, nc[1] - c2);
  const R3 n3a06 = 0.5 * cross(nc[1] - c2, nc[5] - c2);
  const R3 n3a09 = 0.5 * cross(nc[5] - c2, nc[4] - c2);
  const R3 n3a05 = 0.5 * cross(nc[4] - c2, nc[0] - c2);
  const R3 n4a09 = 0.5 * cross(nc[4] - c3, nc[5] - c3);
  const R3 n4a10 = 0.5 * cross(nc[5] - c3, nc[6] - c3);
  const R3 n4a11 = 0.5 * cross(nc[6] - c3, nc[7] - c3);
  const R3 n4a12 = 0.5 * cross(nc[7] - c3, nc[4] - c3);
  const R3 n5a02 = 0.5 * cross(nc[1] - c4, nc[2] - c4);
  const R3 n5a07 = 0.5 * cross(nc[2] - c4, nc[6] - c4);
  const R3 n5a10 = 0.5 * cross(nc[6] - c4, nc[5] - c4);
  const R3 n5a06 = 0.5 * cross(nc[5] - c4, nc[1] - c4);
  const R3 n6a03 = 0.5 * cross(nc[2] - c5, nc[3] - c5);
  const R3 n6a08 = 0.5 * cross(nc[3] - c5, nc[7] - c5);
  const R3 n6a11 = 0.5 * cross(nc[7] - c5, nc[6] - c5);
  const R3 n6a07 = 0.5 * cross(nc[6] - c5, nc[2] - c5);
  const double w = 1. / 12.;
  out[0] = w * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) +
                (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09));
  out[1] = w * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) +
                (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07));
  out[2] = w * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) +
                (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08));
  out[3] = w * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) +
                (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11));
  out[4] = w * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) +
                (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11));
  out[5] = w * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) +
                (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02));
  out[6] = w * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) +
                (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08));
  out[7] = w * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) +
                (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03));
}
// NOTE: the reference writes (5*(...) + (...)) * (1./12.); the product of a
// Real3 by a scalar is commutative component-wise, so w*(...) is identical.

// 2D corner normals, MahycoModule.cc:1093-1112 (host loop in the reference).
void computeCQs2D(const R3 coord_in[4], R3 out[4]) {
  R3 coord[5];
  for (int i = 0; i < 4; ++i) coord[i] = coord_in[i];
  coord[4] = coord[0];
  R3 npc[5];
  for (int i = 0; i < 4; ++i) {
    npc[i + 1].x = 0.5 * (coord[i + 1].y - coord[i].y);
    npc[i + 1].y = 0.5 * (coord[i].x - coord[i + 1].x);
  }
  npc[0] = npc[4];
  for (int ii = 0; ii < 4; ++ii) out[ii] = npc[ii + 1] + npc[ii];
}

// Face centres of a hexahedron (MahycoModule.cc:1080-1087).
void hexFaceCoords(const R3 c[8], R3 fc[6]) {
  fc[0] = 0.25 * (c[0] + c[3] + c[2] + c[1]);
  fc[1] = 0.25 * (c[0] + c[4] + c[7] + c[3]);
  fc[2] = 0.25 * (c[0] + c[1] + c[5] + c[4]);
  fc[3] = 0.25 * (c[4] + c[5] + c[6] + c[7]);
  fc[4] = 0.25 * (c[1] + c[2] + c[6] + c[5]);
  fc[5] = 0.25 * (c[2] + c[3] + c[7] + c[6]);
}

// Material variable: global values + per-environment values for cells that
// hold >= 2 environments.
struct MatVar {
  std::vector<double> g;
  std::vector<double> e;  // [nenv][nc]
  long nc = 0;
  void alloc(long ncells, int nenv) {
    nc = ncells;
    g.assign(ncells, 0.);
    e.assign((size_t)nenv * ncells, 0.);
  }
  double& env(int k, long c) { return e[(size_t)k * nc + c]; }
};

}  // namespace

struct mho_ctx {
  mhg_options opt;
  int dim = 2;
  int n[3] = {1, 1, 1};   // cells per direction (n[2] = 1 in 2D)
  int nn[3] = {1, 1, 1};  // nodes per direction (nn[2] = 1 in 2D)
  long NC = 0, NN = 0, NF = 0;
  long nfd[3] = {0, 0, 0}, foff[3] = {0, 0, 0};
  int fdim[3][3];
  int nbnode = 4, nbface = 4;
  double org[3] = {0, 0, 0}, len[3] = {1, 1, 1};
  std::string err;
  bool initialised = false;

  int nenv = 1, nvar = 8;
  std::vector<double> gamma_env, ptens_env;
  std::vector<int> eos_env;

  // cell material variables (Mahyco.axl:15-75)
  MatVar density, density_n, pressure, pressure_n, pseudo_viscosity, pseudo_viscosity_n,
      pseudo_viscosity_nmoins1, mass_fraction, cell_mass, internal_energy, internal_energy_n,
      dpde, fracvol, cell_volume, cell_volume_n, sound_speed, tau_density;
  // cell plain variables
  std::vector<double> euler_volume, div_u, caracteristic_length;
  std::vector<int> env_id, env_mask, est_mixte, est_pure, cell_status;
  std::vector<R3> cell_coord, cell_cqs, cell_cqs_n, outer_face_normal;  // [nc][nbnode], [nc][2*dim]
  std::vector<double> u_lagrange, phi_lagrange, grad_phi, dual_phi_flux;  // [nc][nvar]
  std::vector<double> back_flux_contrib_env, front_flux_contrib_env;     // [nc][nenv]
  // node variables
  std::vector<R3> velocity, velocity_n, node_coord, node_coord_0, force;
  std::vector<double> node_mass;
  std::vector<double> u_dual_lagrange, phi_dual_lagrange, dual_grad_phi;  // [nn][nvar]
  // learned-limiter coefficients [dim][nc] (MHG_LIM_LEARNED)
  std::vector<double> learned_alpha;
  bool landing = false;  // land_on_final_time: the step under way ends at final_time
  std::vector<double> back_flux_mass, front_flux_mass, back_flux_mass_env, front_flux_mass_env;
  // face variables
  std::vector<R3> face_coord, face_normal, face_length_lagrange;
  std::vector<double> face_normal_velocity, deltax_lagrange;
  std::vector<double> grad_phi_face, phi_face;  // [nf][nvar]
  std::vector<double> h_cell;                   // [nc] cell width along the sweep
  // pente-borne: integrated back / front face fluxes of each cell [nc][nvar]
  std::vector<double> delta_phi_face_ar, delta_phi_face_av;
  // boundary conditions: node lists per bc
  std::vector<std::vector<long>> bc_nodes;
  // scalars
  double global_deltat = 0., global_old_deltat = 0., old_deltat = 0., global_time = 0.;
  int global_iteration = 0, sens_projection = 0;

  // ---- topology -------------------------------------------------------
  long cellId(int i, int j, int k) const { return i + (long)n[0] * (j + (long)n[1] * k); }
  long nodeId(int i, int j, int k) const { return i + (long)nn[0] * (j + (long)nn[1] * k); }
  void cellIdx(long c, int* idx) const {
    idx[0] = (int)(c % n[0]);
    long r = c / n[0];
    idx[1] = (int)(r % n[1]);
    idx[2] = (int)(r / n[1]);
  }
  void nodeIdx(long nd, int* idx) const {
    idx[0] = (int)(nd % nn[0]);
    long r = nd / nn[0];
    idx[1] = (int)(r % nn[1]);
    idx[2] = (int)(r / nn[1]);
  }
  long faceId(int d, int i, int j, int k) const {
    return foff[d] + i + (long)fdim[d][0] * (j + (long)fdim[d][1] * k);
  }
  void faceIdx(long f, int* d, int* idx) const {
    int dd = 0;
    while (dd + 1 < dim && f >= foff[dd + 1]) ++dd;
    long r = f - foff[dd];
    idx[0] = (int)(r % fdim[dd][0]);
    r /= fdim[dd][0];
    idx[1] = (int)(r % fdim[dd][1]);
    idx[2] = (int)(r / fdim[dd][1]);
    *d = dd;
  }
  // Arcane quad4 / hexa8 node order.
  int cellNodes(long c, long* out) const {
    int id[3];
    cellIdx(c, id);
    int i = id[0], j = id[1], k = id[2];
    out[0] = nodeId(i, j, k);
    out[1] = nodeId(i + 1, j, k);
    out[2] = nodeId(i + 1, j + 1, k);
    out[3] = nodeId(i, j + 1, k);
    if (dim == 3) {
      out[4] = nodeId(i, j, k + 1);
      out[5] = nodeId(i + 1, j, k + 1);
      out[6] = nodeId(i + 1, j + 1, k + 1);
      out[7] = nodeId(i, j + 1, k + 1);
      return 8;
    }
    return 4;
  }
  // Arcane quad4 faces (0,1),(1,2),(2,3),(3,0); hexa8 faces
  // (0,3,2,1),(0,4,7,3),(0,1,5,4),(4,5,6,7),(1,2,6,5),(2,3,7,6).
  int cellFaces(long c, long* out) const {
    int id[3];
    cellIdx(c, id);
    int i = id[0], j = id[1], k = id[2];
    if (dim == 2) {
      out[0] = faceId(1, i, j, 0);
      out[1] = faceId(0, i + 1, j, 0);
      out[2] = faceId(1, i, j + 1, 0);
      out[3] = faceId(0, i, j, 0);
      return 4;
    }
    out[0] = faceId(2, i, j, k);
    out[1] = faceId(0, i, j, k);
    out[2] = faceId(1, i, j, k);
    out[3] = faceId(2, i, j, k + 1);
    out[4] = faceId(0, i + 1, j, k);
    out[5] = faceId(1, i, j + 1, k);
    return 6;
  }
  int faceNodes(long f, long* out) const {
    int d, id[3];
    faceIdx(f, &d, id);
    int i = id[0], j = id[1], k = id[2];
    if (dim == 2) {
      out[0] = nodeId(i, j, 0);
      out[1] = (d == 0) ? nodeId(i, j + 1, 0) : nodeId(i + 1, j, 0);
      return 2;
    }
    if (d == 0) {
      out[0] = nodeId(i, j, k); out[1] = nodeId(i, j + 1, k);
      out[2] = nodeId(i, j + 1, k + 1); out[3] = nodeId(i, j, k + 1);
    } else if (d == 1) {
      // Arcane orders the nodes of a face from its smallest id towards the
      // smaller neighbour: (i,j,k), (i+1,j,k), (i+1,j,k+1), (i,j,k+1); with
      // InitGeometricValues' formula (MahycoModule.cc:1020-1023) every
      // Cartesian face normal is then +e_dir
      out[0] = nodeId(i, j, k); out[1] = nodeId(i + 1, j, k);
      out[2] = nodeId(i + 1, j, k + 1); out[3] = nodeId(i, j, k + 1);
    } else {
      out[0] = nodeId(i, j, k); out[1] = nodeId(i + 1, j, k);
      out[2] = nodeId(i + 1, j + 1, k); out[3] = nodeId(i, j + 1, k);
    }
    return 4;
  }
  // previous / next cell of a face in its direction (-1 if none).
  void faceCells(long f, long* prev, long* next) const {
    int d, id[3];
    faceIdx(f, &d, id);
    *prev = -1;
    *next = -1;
    if (id[d] > 0) {
      int q[3] = {id[0], id[1], id[2]};
      q[d] -= 1;
      *prev = cellId(q[0], q[1], q[2]);
    }
    if (id[d] < n[d]) *next = cellId(id[0], id[1], id[2]);
  }
  // Cells of a node in ascending id order, with the node's index in each cell.
  int nodeCells(long nd, long* cells, int* corner) const {
    int id[3];
    nodeIdx(nd, id);
    int cnt = 0;
    int kmin = (dim == 3) ? -1 : 0;
    for (int dk = kmin; dk <= 0; ++dk)
      for (int dj = -1; dj <= 0; ++dj)
        for (int di = -1; di <= 0; ++di) {
          int ci = id[0] + di, cj = id[1] + dj, ck = id[2] + dk;
          if (ci < 0 || ci >= n[0] || cj < 0 || cj >= n[1] || ck < 0 || ck >= n[2]) continue;
          int ox = -di, oy = -dj, oz = -dk;
          int q = (oy == 0) ? (ox == 0 ? 0 : 1) : (ox == 0 ? 3 : 2);
          cells[cnt] = cellId(ci, cj, ck);
          corner[cnt] = q + 4 * oz;
          ++cnt;
        }
    return cnt;
  }
  int nbEnvOf(long c) const { return __builtin_popcount((unsigned)env_mask[c]); }
  // Environments of a cell in ascending id order.
  int envsOf(long c, int* out) const {
    int cnt = 0;
    for (int k = 0; k < nenv; ++k)
      if (env_mask[c] & (1 << k)) out[cnt++] = k;
    return cnt;
  }
  // Environment value with Arcane's pure-cell aliasing.
  double& EV(MatVar& v, int k, long c) { return (nbEnvOf(c) == 1) ? v.g[c] : v.env(k, c); }
  bool isImpure(long c, int k) const {
    return (env_mask[c] & (1 << k)) && __builtin_popcount((unsigned)env_mask[c]) >= 2;
  }
  bool isPure(long c, int k) const { return env_mask[c] == (1 << k); }
  double gammaOf(int k) const { return gamma_env[k]; }
  double qcoef(int k) const {
    return opt.visc_quad_coeff >= 0. ? opt.visc_quad_coeff : (gamma_env[k] + 1) / 2.0;
  }
};

namespace {

int fail(mho_ctx* ctx, int code, const char* msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// env_id, computeMultiEnvGlobalCellId (AccEnvDefaultService.cc:308-349).
void computeEnvId(mho_ctx& S) {
  for (long c = 0; c < S.NC; ++c) {
    int ne = S.nbEnvOf(c);
    if (ne != 1) {
      S.env_id[c] = -ne - 1;
    } else {
      int k = 0;
      while (!(S.env_mask[c] & (1 << k))) ++k;
      S.env_id[c] = k;
    }
  }
}

// Volume (and L) from the current cqs and coordinates (MahycoModule.cc:1116-1267).
void computeVolumeAndLength(mho_ctx& S) {
  const double inv_dim = 1. / S.dim;
  const bool faces_opp = S.opt.longueur_caracteristique == MHG_LC_FACES_OPPOSEES;
  const double racine = S.dim == 2 ? .5 : 1. / 3.;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long nodes[8];
    int nb = S.cellNodes(c, nodes);
    double volume = 0;
    R3 coord[8];
    for (int i = 0; i < nb; ++i) {
      volume += dot(S.node_coord[nodes[i]], S.cell_cqs[c * S.nbnode + i]);
      coord[i] = S.node_coord[nodes[i]];
    }
    volume *= inv_dim;
    S.cell_volume.g[c] = volume;
    if (faces_opp) {
      R3 fc[6];
      fc[0] = 0.25 * (coord[0] + coord[3] + coord[2] + coord[1]);
      fc[1] = 0.25 * (coord[0] + coord[4] + coord[7] + coord[3]);
      fc[2] = 0.25 * (coord[0] + coord[1] + coord[5] + coord[4]);
      fc[3] = 0.25 * (coord[4] + coord[5] + coord[6] + coord[7]);
      fc[4] = 0.25 * (coord[1] + coord[2] + coord[6] + coord[5]);
      fc[5] = 0.25 * (coord[2] + coord[3] + coord[7] + coord[6]);
      R3 m1 = fc[0] - fc[3], m2 = fc[2] - fc[5], m3 = fc[1] - fc[4];
      double d1 = normL2(m1), d2 = normL2(m2), d3 = normL2(m3);
      double num = d1 * d2 * d3;
      double den = d1 * d2 + d1 * d3 + d2 * d3;
      S.caracteristic_length[c] = num / den;
    } else {
      S.caracteristic_length[c] = std::pow(volume, racine);
    }
  }
}

// computeGeometricValues, MahycoModule.cc:1054-1299.
void computeGeometricValues(mho_ctx& S) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long nodes[8];
    int nb = S.cellNodes(c, nodes);
    R3 coord[8];
    for (int i = 0; i < nb; ++i) coord[i] = S.node_coord[nodes[i]];
    R3* out = &S.cell_cqs[c * S.nbnode];
    if (S.dim == 3) {
      R3 fc[6];
      hexFaceCoords(coord, fc);
      computeCQs(coord, fc, out);
    } else {
      computeCQs2D(coord, out);
    }
  }
  computeVolumeAndLength(S);
  // mixed V_k = f_k V (1274-1295)
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) S.cell_volume.env(k, c) = S.fracvol.env(k, c) * S.cell_volume.g[c];
}

// EOS (PerfectGasEOSService.cc:50-137, StiffenedGasEOSService.cc:27-44).
void eosCell(const mho_ctx& S, int k, double rho, double e, double* p, double* cs, double* dpde) {
  const double g = S.gamma_env[k];
  if (S.eos_env[k] == MHG_EOS_STIFFENED_GAS) {
    const double t = S.ptens_env[k];
    double pressure = ((g - 1.) * rho * e) - (g * t);
    *p = pressure;
    *cs = std::sqrt((g / rho) * (pressure + t));
    *dpde = (g - 1.) * rho;
  } else {
    double pressure = (g - 1.) * rho * e;
    *p = pressure;
    *cs = std::sqrt(g * pressure / rho);
    *dpde = (g - 1.) * rho;
  }
}

void applyEOS(mho_ctx& S, int k) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    if (S.isPure(c, k)) {
      eosCell(S, k, S.density.g[c], S.internal_energy.g[c], &S.pressure.g[c],
              &S.sound_speed.g[c], &S.dpde.g[c]);
    } else if (S.isImpure(c, k)) {
      eosCell(S, k, S.density.env(k, c), S.internal_energy.env(k, c), &S.pressure.env(k, c),
              &S.sound_speed.env(k, c), &S.dpde.env(k, c));
    }
  }
}

// computePressionMoyenne, MahycoModule.cc:1836-1915.
void computePressionMoyenne(mho_ctx& S) {
  for (long c = 0; c < S.NC; ++c) {
    if (S.env_id[c] < 0) {
      S.pressure.g[c] = 0.;
      S.sound_speed.g[c] = 1.e-20;
    }
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        S.pressure.g[c] += S.fracvol.env(k, c) * S.pressure.env(k, c);
        S.sound_speed.g[c] = smax(S.sound_speed.env(k, c), S.sound_speed.g[c]);
      }
}

// initEOS of one cell (PerfectGasEOSService.cc:20-31, StiffenedGasEOSService.cc:10-22):
// internal energy and sound speed from density and pressure.
void eosInitCell(const mho_ctx& S, int k, double rho, double p, double* e, double* cs) {
  const double g = S.gamma_env[k];
  if (S.eos_env[k] == MHG_EOS_STIFFENED_GAS) {
    const double t = S.ptens_env[k];
    *e = (p + (g * t)) / ((g - 1.) * rho);
    *cs = std::sqrt((g / rho) * (p + t));
  } else {
    *e = p / ((g - 1.) * rho);
    *cs = std::sqrt(g * p / rho);
  }
}

void initEOS(mho_ctx& S, int k) {
  for (long c = 0; c < S.NC; ++c) {
    if (!(S.env_mask[c] & (1 << k))) continue;
    eosInitCell(S, k, S.EV(S.density, k, c), S.EV(S.pressure, k, c), &S.EV(S.internal_energy, k, c),
                &S.EV(S.sound_speed, k, c));
  }
}

// PrepareFaceGroup (MahycoAnnexe.cc:183-251) + _initBoundaryConditionsForAcc (257-281).
void prepareBoundaryConditions(mho_ctx& S) {
  R3 maxCoor(-1., -1., -1.);
  for (long nd = 0; nd < S.NN; ++nd) {
    maxCoor.x = std::max(maxCoor.x, S.node_coord[nd].x);
    maxCoor.y = std::max(maxCoor.y, S.node_coord[nd].y);
    maxCoor.z = std::max(maxCoor.z, S.node_coord[nd].z);
  }
  const double thr = S.opt.threshold;
  std::vector<std::vector<char>> in_group(6, std::vector<char>(S.NN, 0));
  for (long f = 0; f < S.NF; ++f) {
    long fn[4];
    int nb = S.faceNodes(f, fn);
    bool fl[6] = {true, true, true, true, true, true};
    for (int i = 0; i < nb; ++i) {
      const R3& x = S.node_coord[fn[i]];
      if (x.x > thr) fl[0] = false;
      if (x.y > thr) fl[2] = false;
      if (x.z > thr) fl[4] = false;
      if (std::fabs(x.x - maxCoor.x) > thr) fl[1] = false;
      if (std::fabs(x.y - maxCoor.y) > thr) fl[3] = false;
      if (std::fabs(x.z - maxCoor.z) > thr) fl[5] = false;
    }
    // group ids: 0 XMIN 1 XMAX 2 YMIN 3 YMAX 4 ZMIN 5 ZMAX
    for (int gI = 0; gI < 6; ++gI)
      if (fl[gI])
        for (int i = 0; i < nb; ++i) in_group[gI][fn[i]] = 1;
  }
  S.bc_nodes.clear();
  for (int b = 0; b < S.opt.nb_bc; ++b) {
    std::vector<long> lst;
    int sidx = S.opt.bc[b].surface;  // MHG_SURF_* uses the same numbering
    for (long nd = 0; nd < S.NN; ++nd)
      if (in_group[sidx][nd]) lst.push_back(nd);
    S.bc_nodes.push_back(lst);
  }
}

// InitGeometricValues, MahycoModule.cc:1008-1050.
void initGeometricValues(mho_ctx& S) {
  for (long nd = 0; nd < S.NN; ++nd) S.node_coord_0[nd] = S.node_coord[nd];
  const double one_over_nbnode = S.dim == 2 ? .5 : .25;
  for (long f = 0; f < S.NF; ++f) {
    long fn[4];
    S.faceNodes(f, fn);
    R3 nrm;
    if (S.dim == 3) {
      R3 v1 = S.node_coord[fn[2]] - S.node_coord[fn[0]];
      R3 v2 = S.node_coord[fn[3]] - S.node_coord[fn[1]];
      nrm.x = v1.y * v2.z - v1.z * v2.y;
      nrm.y = -(v2.x * v1.z - v2.z * v1.x);
      nrm.z = v1.x * v2.y - v1.y * v2.x;
    } else {
      nrm.x = S.node_coord[fn[1]].y - S.node_coord[fn[0]].y;
      nrm.y = S.node_coord[fn[1]].x - S.node_coord[fn[0]].x;
    }
    double nl = normL2(nrm);
    nrm.x /= nl;
    nrm.y /= nl;
    nrm.z /= nl;
    S.face_normal[f] = nrm;
  }
  for (long f = 0; f < S.NF; ++f) {
    long fn[4];
    int nb = S.faceNodes(f, fn);
    R3 fc;
    for (int i = 0; i < nb; ++i) fc += one_over_nbnode * S.node_coord[fn[i]];
    S.face_coord[f] = fc;
  }
  const int nfc = 2 * S.dim;
  for (long c = 0; c < S.NC; ++c) {
    long cf[6];
    S.cellFaces(c, cf);
    for (int i = 0; i < nfc; ++i) {
      R3 v = S.face_coord[cf[i]] - S.cell_coord[c];
      double nl = normL2(v);
      S.outer_face_normal[c * nfc + i] = R3(v.x / nl, v.y / nl, v.z / nl);
    }
  }
}

void cellCoordFromNodes(mho_ctx& S) {
  const double one_over_nbnode = S.dim == 2 ? .25 : .125;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long nodes[8];
    int nb = S.cellNodes(c, nodes);
    R3 somme(0., 0., 0.);
    for (int i = 0; i < nb; ++i) somme += S.node_coord[nodes[i]];
    S.cell_coord[c] = one_over_nbnode * somme;
  }
}

// computeNodeMass (MahycoModule.cc:235-266) / "nodem" (RemapADIFinal.cc:712-718).
void computeNodeMass(mho_ctx& S) {
  const double one_over_nbnode = S.dim == 2 ? .25 : .125;
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    long cells[8];
    int corner[8];
    int nb = S.nodeCells(nd, cells, corner);
    double sum_mass = 0.;
    for (int i = 0; i < nb; ++i) sum_mass += S.cell_mass.g[cells[i]];
    S.node_mass[nd] = one_over_nbnode * sum_mass;
  }
}

// Common part of the init: hydroStartInit (MahycoModule.cc:48-188) after the
// case has set env membership and the initial ρ, p, u; then computeCellMass
// (199-225) and computeNodeMass (235-266).
void finishInit(mho_ctx& S) {
  computeEnvId(S);
  // computeGeometricValues at init (96): cqs, V, L, mixed V_k = f_k V (f_k unset -> 0 here).
  computeGeometricValues(S);
  prepareBoundaryConditions(S);
  for (int k = 0; k < S.nenv; ++k) initEOS(S, k);
  // mixed-cell averages (113-127)
  for (long c = 0; c < S.NC; ++c) {
    if (S.nbEnvOf(c) != 1) {
      S.internal_energy.g[c] = 0.;
      int ev[MHG_MAX_ENV];
      int ne = S.envsOf(c, ev);
      for (int q = 0; q < ne; ++q) {
        int k = ev[q];
        S.internal_energy.g[c] += S.internal_energy.env(k, c) * S.mass_fraction.env(k, c);
        S.sound_speed.g[c] = std::max(S.sound_speed.env(k, c), S.sound_speed.g[c]);
      }
    }
  }
  for (long c = 0; c < S.NC; ++c) S.euler_volume[c] = S.cell_volume.g[c];
  initGeometricValues(S);
  // volume recomputed from cqs (154-182)
  {
    const double inv_dim = 1. / S.dim;
    for (long c = 0; c < S.NC; ++c) {
      long nodes[8];
      int nb = S.cellNodes(c, nodes);
      double volume = 0;
      for (int i = 0; i < nb; ++i)
        volume += dot(S.node_coord[nodes[i]], S.cell_cqs[c * S.nbnode + i]);
      volume *= inv_dim;
      S.cell_volume.g[c] = volume;
    }
  }
  // computeCellMass (199-225)
  for (long c = 0; c < S.NC; ++c) S.cell_mass.g[c] = S.density.g[c] * S.cell_volume.g[c];
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.env_mask[c] & (1 << k))
        S.EV(S.cell_mass, k, c) = S.EV(S.mass_fraction, k, c) * S.cell_mass.g[c];
  computeNodeMass(S);
  S.initialised = true;
}

// ---------------------------------------------------------------------------
// Compute-loop entry points
// ---------------------------------------------------------------------------

void beginIteration(mho_ctx& S) {
  S.global_time += S.global_deltat;
  S.global_iteration += 1;
}

// saveValuesAtN, MahycoModule.cc:308-450.
void saveValuesAtN(mho_ctx& S) {
  S.global_old_deltat = S.old_deltat;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    S.pseudo_viscosity_nmoins1.g[c] = S.pseudo_viscosity_n.g[c];
    S.pseudo_viscosity_n.g[c] = S.pseudo_viscosity.g[c];
    S.pressure_n.g[c] = S.pressure.g[c];
    S.cell_volume_n.g[c] = S.cell_volume.g[c];
    S.density_n.g[c] = S.density.g[c];
    S.internal_energy_n.g[c] = S.internal_energy.g[c];
    for (int i = 0; i < S.nbnode; ++i) S.cell_cqs_n[c * S.nbnode + i] = S.cell_cqs[c * S.nbnode + i];
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        S.pseudo_viscosity_nmoins1.env(k, c) = S.pseudo_viscosity_n.env(k, c);
        S.pseudo_viscosity_n.env(k, c) = S.pseudo_viscosity.env(k, c);
        S.pressure_n.env(k, c) = S.pressure.env(k, c);
        S.cell_volume_n.env(k, c) = S.cell_volume.env(k, c);
        S.density_n.env(k, c) = S.density.env(k, c);
        S.internal_energy_n.env(k, c) = S.internal_energy.env(k, c);
      }
  const bool copy_velocity = !S.opt.sans_lagrange;
  const bool copy_node_coord = S.opt.with_projection && S.opt.is_euler_scheme;
  for (long nd = 0; nd < S.NN; ++nd) {
    if (copy_velocity) S.velocity_n[nd] = S.velocity[nd];
    if (copy_node_coord) S.node_coord[nd] = S.node_coord_0[nd];
  }
}

// updateForceAndVelocity, MahycoModule.cc:589-665 (called from updateVelocity 678-731).
void updateVelocity(mho_ctx& S) {
  const double dt = 0.5 * (S.global_old_deltat + S.global_deltat);
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    long cells[8];
    int corner[8];
    int nb = S.nodeCells(nd, cells, corner);
    R3 node_force(0., 0., 0.);
    for (int i = 0; i < nb; ++i) {
      long c = cells[i];
      node_force += (S.pressure_n.g[c] + S.pseudo_viscosity_n.g[c]) *
                    S.cell_cqs_n[c * S.nbnode + corner[i]];
    }
    S.force[nd] = node_force;
    S.velocity[nd] = S.velocity_n[nd] + (dt / S.node_mass[nd]) * node_force;
  }
}

// applyBoundaryCondition, MahycoModule.cc:936-1004.
void applyBoundaryCondition(mho_ctx& S) {
  for (int b = 0; b < S.opt.nb_bc; ++b) {
    const double value = S.opt.bc[b].value;
    const int type = S.opt.bc[b].type;
    for (long nd : S.bc_nodes[b]) {
      R3 v = S.velocity[nd];
      if (type == MHG_BC_VX) v.x = value;
      else if (type == MHG_BC_VY) v.y = value;
      else if (type == MHG_BC_VZ) v.z = value;
      S.velocity[nd] = v;
    }
  }
}

// updatePosition, MahycoModule.cc:870-931.
void updatePosition(mho_ctx& S) {
  const double deltat = S.global_deltat;
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) S.node_coord[nd] += deltat * S.velocity[nd];
  cellCoordFromNodes(S);
}

// updateDensity, MahycoModule.cc:1307-1426.
void updateDensity(mho_ctx& S) {
  const double inv_deltat = 1.0 / S.global_deltat;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    double density_n = S.density_n.g[c];
    double new_density = S.cell_mass.g[c] / S.cell_volume.g[c];
    double tau = 0.5 * (1.0 / density_n + 1.0 / new_density);
    S.density.g[c] = new_density;
    S.tau_density.g[c] = tau;
    S.div_u[c] = inv_deltat * (1.0 / S.density.g[c] - 1.0 / density_n) / S.tau_density.g[c];
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        double rho = S.cell_mass.env(k, c) / S.cell_volume.env(k, c);
        S.density.env(k, c) = rho;
        S.tau_density.env(k, c) = 0.5 * (1.0 / S.density_n.env(k, c) + 1.0 / rho);
      }
}

// computeArtificialViscosity, MahycoModule.cc:454-578.
void computeArtificialViscosity(mho_ctx& S) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    S.pseudo_viscosity.g[c] = 0.;
    int env = S.env_id[c];
    if (env >= 0 && S.div_u[c] < 0.0) {
      S.pseudo_viscosity.g[c] = viscosity(S.qcoef(env), S.tau_density.g[c],
                                          S.caracteristic_length[c], S.sound_speed.g[c],
                                          S.div_u[c]);
    }
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        double qk = 0.;
        if (S.div_u[c] < 0.0)
          qk = viscosity(S.qcoef(k), S.tau_density.env(k, c), S.caracteristic_length[c],
                         S.sound_speed.g[c], S.div_u[c]);
        S.pseudo_viscosity.env(k, c) = qk;
        S.pseudo_viscosity.g[c] += qk * S.fracvol.env(k, c);
      }
}

// updateEnergyAndPressureforGP, MahycoModule.cc:1659-1770 (non-CSTS path),
// called from updateEnergyAndPressure 1439-1478.
void updateEnergyAndPressure(mho_ctx& S) {
  const bool pc = S.opt.pseudo_centree != 0;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    S.internal_energy.g[c] = 0.;
    int env = S.env_id[c];
    if (env >= 0)
      S.internal_energy.g[c] = compute_eint(
          pc, S.gammaOf(env), S.pseudo_viscosity_n.g[c], S.pseudo_viscosity.g[c],
          S.density_n.g[c], S.density.g[c], S.pressure.g[c], S.internal_energy_n.g[c]);
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        double ek = compute_eint(pc, S.gammaOf(k), S.pseudo_viscosity_n.env(k, c),
                                 S.pseudo_viscosity.env(k, c), S.density_n.env(k, c),
                                 S.density.env(k, c), S.pressure.env(k, c),
                                 S.internal_energy_n.env(k, c));
        S.internal_energy.env(k, c) = ek;
        S.internal_energy.g[c] += S.mass_fraction.env(k, c) * ek;
      }
  if (!S.opt.with_projection) {
    for (int k = 0; k < S.nenv; ++k) applyEOS(S, k);
    computePressionMoyenne(S);
  }
}

// computeVariablesForRemap_PBorn0, PrepareRemap.cc:258-437.
void computeVariablesForRemap(mho_ctx& S) {
  const int ne = S.nenv, nv = S.nvar;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    double* U = &S.u_lagrange[c * nv];
    for (int v = 0; v < nv; ++v) U[v] = 0;
    int env = S.env_id[c];
    if (env >= 0) {
      double V = S.cell_volume.g[c];
      U[env] = V;
      U[ne + env] = V * S.density.g[c];
      U[2 * ne + env] = V * S.density.g[c] * S.internal_energy.g[c];
      U[3 * ne + 0] = 0.;
      U[3 * ne + 1] = 0.;
      U[3 * ne + 2] = 0.;
      U[3 * ne + 3] = 0.;
      U[3 * ne + 4] = V * S.pseudo_viscosity.g[c];
    }
  }
  for (int k = 0; k < ne; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        double* U = &S.u_lagrange[c * nv];
        double Vk = S.cell_volume.env(k, c);
        U[k] = Vk;
        U[ne + k] = Vk * S.density.env(k, c);
        U[2 * ne + k] = Vk * S.density.env(k, c) * S.internal_energy.env(k, c);
        U[3 * ne + 0] = 0.;
        U[3 * ne + 1] = 0.;
        U[3 * ne + 2] = 0.;
        U[3 * ne + 3] = 0.;
        U[3 * ne + 4] = Vk * S.pseudo_viscosity.env(k, c);
      }
  if (is_pb(S.opt)) {
    // computeVariablesForRemap, PrepareRemap.cc:159-215: each environment of
    // the cell, in ascending order, writes its fraction, density, energy and
    // viscosity (the last one's viscosity stays); an environment absent from
    // the cell leaves its slots as the previous remap left them
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c)
      for (int k = 0; k < ne; ++k) {
        if (!(S.env_mask[c] & (1 << k))) continue;
        double* phi = &S.phi_lagrange[c * nv];
        phi[k] = S.EV(S.fracvol, k, c);
        phi[ne + k] = S.EV(S.density, k, c);
        phi[2 * ne + k] = S.EV(S.internal_energy, k, c);
        phi[3 * ne + 0] = 0.;
        phi[3 * ne + 1] = 0.;
        phi[3 * ne + 2] = 0.;
        phi[3 * ne + 3] = 0.;
        phi[3 * ne + 4] = S.EV(S.pseudo_viscosity, k, c);
      }
  } else {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c)
      for (int v = 0; v < nv; ++v)
        S.phi_lagrange[c * nv + v] = S.u_lagrange[c * nv + v] / S.cell_volume.g[c];
  }
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    const R3& u = S.velocity[nd];
    double m = S.node_mass[nd];
    double* Ud = &S.u_dual_lagrange[nd * nv];
    double* Pd = &S.phi_dual_lagrange[nd * nv];
    Ud[0] = m * u.x;
    Ud[1] = m * u.y;
    Ud[2] = m * u.z;
    Ud[3] = m;
    Ud[4] = 0.5 * m * normL2(u);
    Pd[0] = u.x;
    Pd[1] = u.y;
    Pd[2] = u.z;
    Pd[3] = m;
    Pd[4] = 0.5 * normL2(u);
  }
}

// computeFaceQuantitesForRemap, PrepareRemap.cc:15-123 (non-CSTS path).
void computeFaceQuantitesForRemap(mho_ctx& S) {
  const double one_over_nbnode = S.dim == 2 ? .5 : .25;
#pragma omp parallel for schedule(static)
  for (long f = 0; f < S.NF; ++f) {
    long fn[4];
    int nb = S.faceNodes(f, fn);
    if (S.dim == 3) {
      R3 coord[4];
      for (int i = 0; i < 4; ++i) coord[i] = S.node_coord[fn[i]];
      R3 v1 = coord[2] - coord[0], v2 = coord[3] - coord[1];
      R3 tmp = 0.5 * cross(v1, v2);
      S.face_length_lagrange[f] = absolute(tmp);
    } else {
      R3 t = S.node_coord[fn[1]] - S.node_coord[fn[0]];
      S.face_length_lagrange[f] = R3(std::fabs(t.y), std::fabs(t.x), 0.);
    }
    R3 vitesse_moy(0., 0., 0.);
    R3 fc(0., 0., 0.);
    for (int i = 0; i < nb; ++i) {
      fc += one_over_nbnode * S.node_coord[fn[i]];
      vitesse_moy += S.velocity[fn[i]];
    }
    S.face_coord[f] = fc;
    S.face_normal_velocity[f] = dot(one_over_nbnode * vitesse_moy, S.face_normal[f]);
  }
}

// Faces of direction d: [foff[d], foff[d] + nfd[d]).
bool isInnerFaceOfDir(const mho_ctx& S, long f, int d, long* prev, long* next) {
  S.faceCells(f, prev, next);
  (void)d;
  return *prev >= 0 && *next >= 0;
}

// computeGradPhiFace, RemapADIService.cc:91-271 (gphi kernel 163-178).
void computeGradPhiFace(mho_ctx& S, int idir) {
  const int nv = S.nvar;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < S.nfd[idir]; ++q) {
    long f = S.foff[idir] + q, pc, nc;
    if (!isInnerFaceOfDir(S, f, idir, &pc, &nc)) continue;
    S.deltax_lagrange[f] = dot(S.cell_coord[nc] - S.cell_coord[pc], S.face_normal[f]);
    for (int v = 0; v < nv; ++v)
      S.grad_phi_face[f * nv + v] =
          (S.phi_lagrange[nc * nv + v] - S.phi_lagrange[pc * nv + v]) / S.deltax_lagrange[f];
  }
}

// back / front faces of a cell in direction d.
void cellDirFaces(const mho_ctx& S, long c, int d, long* back, long* front) {
  int id[3];
  S.cellIdx(c, id);
  *back = S.faceId(d, id[0], id[1], id[2]);
  int q[3] = {id[0], id[1], id[2]};
  q[d] += 1;
  *front = S.faceId(d, q[0], q[1], q[2]);
}

// Cell widths along idir (computeGradPhiFace, HCELL_BY_CELLS,
// RemapADIService.cc:223-255): the distances from the cell centre to the
// centres of its back and front faces, the face on the mesh boundary
// excluded.
void computeCellWidths(mho_ctx& S, int idir) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    int id[3];
    S.cellIdx(c, id);
    long bf, ff;
    cellDirFaces(S, c, idir, &bf, &ff);
    double h = 0.;
    if (id[idir] > 0) h += normL2(S.face_coord[bf] - S.cell_coord[c]);
    if (id[idir] < S.n[idir] - 1) h += normL2(S.face_coord[ff] - S.cell_coord[c]);
    S.h_cell[c] = h;
  }
}

// computeGradPhiCell, RemapADIService.cc:283-317 -> _PBorn0_LimC 447-517.
// Generalised limiters: the generic loop 349-396 with computeAndLimitGradPhi
// (UtilesRemap.cc:306-339), whose back / front cell is the cell itself on the
// mesh boundary.
void computeGradPhiCellGeneric(mho_ctx& S, int idir);
void computeGradPhiCell(mho_ctx& S, int idir) {
  const int nv = S.nvar;
  std::fill(S.grad_phi.begin(), S.grad_phi.end(), 0.);
  if (generic_grad(S.opt)) {
    computeGradPhiCellGeneric(S, idir);
    return;
  }
  if (S.opt.ordre_projection > 1 && S.opt.projection_pente_borne == 0 &&
      is_learned(S.opt)) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c) {
      long bf, ff, pb, nb, pf, nf;
      cellDirFaces(S, c, idir, &bf, &ff);
      const bool back_in = isInnerFaceOfDir(S, bf, idir, &pb, &nb);
      const bool front_in = isInnerFaceOfDir(S, ff, idir, &pf, &nf);
      const double db = dot(S.face_coord[bf] - S.cell_coord[c], S.face_normal[bf]);
      const double df = dot(S.face_coord[ff] - S.cell_coord[c], S.face_normal[ff]);
      const double alpha = S.learned_alpha.empty() ? 0. : S.learned_alpha[(size_t)idir * S.NC + c];
      const double* ph = &S.phi_lagrange[c * nv];
      double* g = &S.grad_phi[c * nv];
      int present = 0;
      for (int k = 0; k < S.nenv; ++k) present += ph[k] > 0. ? 1 : 0;
      const bool cons = alpha >= 0. && present >= 2;
      double gcl[3 * MHG_MAX_ENV + 8];
      bool used[MHG_MAX_ENV] = {};
      for (int v = 0; v < nv; ++v) {
        const double phi = ph[v];
        const double phim = back_in ? S.phi_lagrange[pb * nv + v] : phi;
        const double phip = front_in ? S.phi_lagrange[nf * nv + v] : phi;
        const double gb = S.grad_phi_face[bf * nv + v], gf = S.grad_phi_face[ff * nv + v];
        bool u = false;
        // a trace environment (0 < phi_V <= remap threshold) keeps the classical slope
        const bool trace = v < S.nenv && phi > 0. && phi <= S.opt.remap_threshold;
        g[v] = learned_grad(trace ? -1. : alpha, phim, phi, phip, back_in, front_in, gb, gf, db, df, &u);
        if (v < S.nenv) used[v] = u;
        if (cons) gcl[v] = learned_grad(-1., phim, phi, phip, back_in, front_in, gb, gf, db, df);
      }
      if (cons) consistent_env_slopes(S.nenv, ph, g, gcl, used, S.opt.remap_threshold);
    }
  } else if (S.opt.ordre_projection > 1 && S.opt.projection_pente_borne == 0 && is_generalised(S.opt)) {
    computeCellWidths(S, idir);
    const int lim = S.opt.projection_limiteur_id;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c) {
      long bf, ff, pb, nb, pf, nf;
      cellDirFaces(S, c, idir, &bf, &ff);
      const long bc = isInnerFaceOfDir(S, bf, idir, &pb, &nb) ? pb : c;
      const long fc = isInnerFaceOfDir(S, ff, idir, &pf, &nf) ? nf : c;
      for (int v = 0; v < nv; ++v)
        S.grad_phi[c * nv + v] =
            flux_limiter_g(lim, S.grad_phi_face[ff * nv + v], S.grad_phi_face[bf * nv + v],
                           S.phi_lagrange[fc * nv + v], S.phi_lagrange[bc * nv + v], S.h_cell[c],
                           S.h_cell[fc], S.h_cell[bc]);
    }
  } else if (S.opt.ordre_projection > 1 && S.opt.projection_pente_borne == 0 &&
             S.opt.projection_limiteur_id < MHG_LIM_MINMODG) {
    const int lim = S.opt.projection_limiteur_id;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c) {
      long bf, ff;
      cellDirFaces(S, c, idir, &bf, &ff);
      for (int v = 0; v < nv; ++v) {
        double g = 0.;
        double gb = S.grad_phi_face[bf * nv + v];
        double gf = S.grad_phi_face[ff * nv + v];
        if (gb != 0.) g += 0.5 * (limiter(lim, gf / gb) * gb);
        if (gf != 0.) g += 0.5 * (limiter(lim, gb / gf) * gf);
        S.grad_phi[c * nv + v] = g;
      }
    }
  }
}

// Order-3 face value (Remap-ordre-3.cc): Arcane's math::min / math::max
// ((a < b) ? a : b and (a < b) ? b : a).
inline double amin(double a, double b) { return (a < b) ? a : b; }
inline double amax(double a, double b) { return (a < b) ? b : a; }
// evaluate_grad (7-14)
inline double o3_grad(double hm, double h0, double hp, double ym, double y0, double yp) {
  return h0 / (hm + h0 + hp) * ((2. * hm + h0) / (h0 + hp) * (yp - y0) + (h0 + 2. * hp) / (hm + h0) * (y0 - ym));
}
// evaluate_ystar (17-30)
inline double o3_ystar(double hmm, double hm, double hp, double hpp, double ymm, double ym, double yp, double ypp,
                       double gradm, double gradp) {
  (void)ymm;
  (void)ypp;
  const double tmp1 = (2. * hp * hm) / (hm + hp) *
                      ((hmm + hm) / (2. * hm + hp) - (hpp + hp) / (2. * hp + hm)) * (yp - ym);
  const double tmp2 = -hm * (hmm + hm) / (2. * hm + hp) * gradp + hp * (hp + hpp) / (hm + 2. * hp) * gradm;
  return ym + hm / (hm + hp) * (yp - ym) + 1. / (hmm + hm + hp + hpp) * (tmp1 + tmp2);
}
// ComputeFluxOrdre3 (69-131): the mean over the swept region of the
// parabolic reconstruction of the upwind cell, bounded by the TVD interval.
inline double o3_flux(double ymmm, double ymm, double ym, double yp, double ypp, double yppp, double hmmm,
                      double hmm, double hm, double hp, double hpp, double hppp, double vdt) {
  const double gradmm = o3_grad(hmmm, hmm, hm, ymmm, ymm, ym);
  const double gradm = o3_grad(hmm, hm, hp, ymm, ym, yp);
  const double gradp = o3_grad(hm, hp, hpp, ym, yp, ypp);
  const double gradpp = o3_grad(hp, hpp, hppp, yp, ypp, yppp);
  const double ystarm = o3_ystar(hmmm, hmm, hm, hp, ymmm, ymm, ym, yp, gradmm, gradm);
  const double ystar = o3_ystar(hmm, hm, hp, hpp, ymm, ym, yp, ypp, gradm, gradp);
  const double ystarp = o3_ystar(hm, hp, hpp, hppp, ym, yp, ypp, yppp, gradp, gradpp);
  const double ym_m = ystarm, ym_p = ystar, yp_m = ystar, yp_p = ystarp;
  const double grad_m = ym_p - ym_m, grad_p = yp_p - yp_m;
  const double ym6 = 6. * (ym - 0.5 * (ym_m + ym_p));
  const double yp6 = 6. * (yp - 0.5 * (yp_m + yp_p));
  double flux;
  if (vdt > 0.) {  // evaluate_fm (33-38)
    flux = ym_p - 0.5 * vdt / hm * (grad_m - (1. - 2. / 3. * vdt / hm) * ym6);
  } else if (vdt < 0.) {  // evaluate_fp (41-46)
    flux = yp_m + 0.5 * (-vdt) / hp * (grad_p - (1. - 2. / 3. * (-vdt) / hp) * yp6);
  } else {
    return 0.;
  }
  const double num = vdt / hm, nup = vdt / hp;
  const double ym_ym = ym + (1. - num) / num * (ym - ymm);
  const double yp_ym = yp - (1. + nup) / nup * (yp - ypp);
  const double i1a = amin(ym, yp), i1b = amax(ym, yp);
  const double i2a = vdt >= 0. ? amin(ym, ym_ym) : amin(yp, yp_ym);
  const double i2b = vdt >= 0. ? amax(ym, ym_ym) : amax(yp, yp_ym);
  double la = 0., lb = 0.;
  if (!((i1b < i2a) || (i2b < i1a))) {
    la = amax(i1a, i2a);
    lb = amin(i1b, i2b);
  }
  if (flux < la) flux = la;
  if (flux > lb) flux = lb;
  return flux;
}

// ---- pente-borne (UtilesRemap.cc:73-814) -------------------------------
// Arcane's math::min / math::max (amin / amax) and std::max (smax), as the
// reference calls them; abs of a Real is the floating-point absolute value.

// INTY (UtilesRemap.cc:83-94): integral from -infinity to X of the linear
// function through (x0, y0) and (x1, y1) restricted to [x0, x1]; zero for a
// degenerate interval (|x1 - x0| <= threshold).
inline double pb_inty(double X, double x0, double y0, double x1, double y1, double thr) {
  double flux = 0.;
  if (std::fabs(x1 - x0) > thr) {
    const double eta = amin(amax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// computeY0 (UtilesRemap.cc:184-240): the monotonicity thresholds y0plus /
// y0moins of the single-slope reconstruction for limiter lim (zero for an
// identifier without a rule).
inline void pb_y0(int lim, double yp, double ym, double h0, double hp, double hm, double* y0plus,
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

// computexgxd and computeygyd (UtilesRemap.cc:252-305): the abscissae xg, xd
// (cell centre at 0, width h0) and ordinates yg, yd of the inner support
// points of the three-piece reconstruction through (-h0/2, ym) and (h0/2, yp).
inline void pb_support(double y0, double yp, double ym, double h0, double y0plus, double y0moins, double grady,
                       double thr, double* xg, double* yg, double* xd, double* yd) {
  double xplus = 1.;
  if (std::fabs(y0plus - yp) > thr) xplus = (y0 - yp) / (y0plus - yp) - 1. / 2.;
  double xmoins = 1.;
  if (std::fabs(y0moins - ym) > thr) xmoins = (y0 - ym) / (y0moins - ym) - 1. / 2.;
  *xd = +h0 * amin(amax(xplus, -1. / 2.), 1. / 2.);
  *xg = -h0 * amin(amax(xmoins, -1. / 2.), 1. / 2.);
  const double xtd = y0 + h0 / 2 * grady;
  const double xtg = y0 - h0 / 2 * grady;
  const double ad = amin(yp, 2. * y0moins - ym), bd = amax(yp, 2. * y0moins - ym);
  const double ag = amin(ym, 2. * y0plus - yp), bg = amax(ym, 2. * y0plus - yp);
  *yd = amin(amax(xtd, ad), bd);
  *yg = amin(amax(xtg, ag), bg);
}

// One variable's pente-borne flux of computeFluxPP / computeFluxPPPure
// (UtilesRemap.cc:420-530): the integral of the reconstruction over the
// length ppv (the outgoing volume per unit face length) next to the back
// face (type 0) or the front face (type 1), floored at zero; the first-order
// value y0 ppv at a local extremum (formula 16).
inline double pb_var_flux(int type, double ppv, double h0, double y0, double yp, double ym, double xg, double yg,
                          double xd, double yd, double thr) {
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

// computeFluxPP (pure = false, UtilesRemap.cc:406-603) and computeFluxPPPure
// (pure = true, 628-814) for one face of cell c: F[nv] from the cell's and
// its back / front neighbours' phi, widths and the cell's limited gradient,
// then the debar fix (1: environment mass and energy fluxes scaled by the
// volume flux, 2: taken from the cell's density and energy; the pure form
// fixes energies only) and zero for a face without outgoing volume.  The
// dual flux the reference also forms here is overwritten by computeUremap
// before any use, so it is not computed.
void pb_face_flux(mho_ctx& S, long c, long fc, long bc, double v_out, int type, bool pure, double* F) {
  const int nv = S.nvar, ne = S.nenv;
  const double thr = S.opt.remap_threshold, dt = S.global_deltat;
  const int lim = pure ? S.opt.projection_limiteur_pure_id : S.opt.projection_limiteur_id;
  const double ppv = 0.5 * (v_out + std::fabs(v_out)) * dt;
  const double* P = S.phi_lagrange.data();
  const double h0 = S.h_cell[c], hp = S.h_cell[fc], hm = S.h_cell[bc];
  for (int v = 0; v < nv; ++v) {
    const double y0 = P[c * nv + v], yp = P[fc * nv + v], ym = P[bc * nv + v];
    double y0plus, y0moins, xg, yg, xd, yd;
    pb_y0(lim, yp, ym, h0, hp, hm, &y0plus, &y0moins);
    pb_support(y0, yp, ym, h0, y0plus, y0moins, S.grad_phi[c * nv + v], thr, &xg, &yg, &xd, &yd);
    F[v] = pb_var_flux(type, ppv, h0, y0, yp, ym, xg, yg, xd, yd, thr);
  }
  const int fix = S.opt.projection_pente_borne_debar_fix;
  if (fix == 1) {
    double sfm = 0., sfv = 0.;
    for (int k = 0; k < ne; ++k) sfv += F[k];
    if (std::fabs(sfv) > thr) {
      for (int k = 0; k < ne; ++k) {
        if (!pure) F[ne + k] = (F[ne + k] / sfv) * F[k];
        F[2 * ne + k] = (F[2 * ne + k] / sfv) * F[ne + k];
        sfm += F[ne + k];
      }
      F[3 * ne] = (F[3 * ne] / sfv) * sfm;
      F[3 * ne + 1] = (F[3 * ne + 1] / sfv) * sfm;
      F[3 * ne + 2] = (F[3 * ne + 2] / sfv) * sfm;
      F[3 * ne + 3] = P[c * nv + 3 * ne + 3] * sfv;
    } else {
      for (int v = 0; v < nv; ++v) F[v] = 0.;
    }
  } else if (fix == 2) {
    double sfm = 0., sfv = 0.;
    for (int k = 0; k < ne; ++k) {
      if (!pure) F[ne + k] = P[c * nv + ne + k] * F[k];
      F[2 * ne + k] = P[c * nv + 2 * ne + k] * F[ne + k];
      sfm += F[ne + k];
      sfv += F[k];
    }
    F[3 * ne] = P[c * nv + 3 * ne] * sfm;
    F[3 * ne + 1] = P[c * nv + 3 * ne + 1] * sfm;
    F[3 * ne + 2] = P[c * nv + 3 * ne + 2] * sfm;
    F[3 * ne + 3] = P[c * nv + 3 * ne + 3] * sfv;
  }
  if (ppv == 0.)
    for (int v = 0; v < nv; ++v) F[v] = 0.;
}

// computeGradPhiCell's generic loop (RemapADIService.cc:319-431) for
// projection-pente-borne or -mixte: per cell, its back / front neighbours
// along idir (the cell itself on the mesh boundary), the limiter (the pure
// one where projection-limiteur-pure-id is 1 and the neighbourhood is pure
// under -mixte), the limited gradient (computeAndLimitGradPhi,
// UtilesRemap.cc:306-339), and with pente-borne the integrated fluxes of the
// back face (outgoing velocity -u_back) and of the front face (u_front).
void computeGradPhiCellGeneric(mho_ctx& S, int idir) {
  const int nv = S.nvar, nfc = 2 * S.dim;
  computeCellWidths(S, idir);
  std::fill(S.delta_phi_face_ar.begin(), S.delta_phi_face_ar.end(), 0.);
  std::fill(S.delta_phi_face_av.begin(), S.delta_phi_face_av.end(), 0.);
  const R3 dirproj(0.5 * (1 - idir) * (2 - idir), 1.0 * idir * (2 - idir), -0.5 * idir * (1 - idir));
  const bool mixte = S.opt.projection_pente_borne_mixte != 0;
  const int pure_id = S.opt.projection_limiteur_pure_id;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long bf, ff, pb, nb, pf, nf, cf[6];
    cellDirFaces(S, c, idir, &bf, &ff);
    const long bc = isInnerFaceOfDir(S, bf, idir, &pb, &nb) ? pb : c;
    const long fc = isInnerFaceOfDir(S, ff, idir, &pf, &nf) ? nf : c;
    S.cellFaces(c, cf);
    int ib = 0, iff = 0;
    for (int q = 0; q < nfc; ++q) {
      if (cf[q] == bf) ib = q;
      if (cf[q] == ff) iff = q;
    }
    const bool vp = mixte && S.est_mixte[c] == 0 && S.est_mixte[fc] == 0 && S.est_mixte[bc] == 0 &&
                    S.est_pure[c] == S.est_pure[fc] && S.est_pure[c] == S.est_pure[bc];
    int lim = S.opt.projection_limiteur_id;
    if (pure_id == 1 && vp) lim = pure_id;
    for (int v = 0; v < nv; ++v) {
      const double gb = S.grad_phi_face[bf * nv + v], gf = S.grad_phi_face[ff * nv + v];
      double g = 0.;
      if (lim < MHG_LIM_MINMODG) {
        if (gb != 0.) g += 0.5 * (limiter(lim, gf / gb) * gb);
        if (gf != 0.) g += 0.5 * (limiter(lim, gb / gf) * gf);
      } else {
        g = flux_limiter_g(lim, gf, gb, S.phi_lagrange[fc * nv + v], S.phi_lagrange[bc * nv + v], S.h_cell[c],
                           S.h_cell[fc], S.h_cell[bc]);
      }
      S.grad_phi[c * nv + v] = g;
    }
    if (is_pb(S.opt)) {
      const double v_ar = dot(S.outer_face_normal[c * nfc + ib], dirproj) * S.face_normal_velocity[bf];
      const double v_av = dot(S.outer_face_normal[c * nfc + iff], dirproj) * S.face_normal_velocity[ff];
      pb_face_flux(S, c, fc, bc, v_ar, 0, vp, &S.delta_phi_face_ar[c * nv]);
      pb_face_flux(S, c, fc, bc, v_av, 1, vp, &S.delta_phi_face_av[c * nv]);
    }
  }
}

// computeUpwindFaceQuantitiesForProjection with pente-borne
// (RemapADIService.cc:549-581): the net integrated flux through each inner
// face, the back cell's front-face flux minus the front cell's back-face flux.
void computeUpwindFaceQuantitiesPB(mho_ctx& S, int idir) {
  const int nv = S.nvar;
  std::fill(S.phi_face.begin(), S.phi_face.end(), 0.);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < S.nfd[idir]; ++q) {
    long f = S.foff[idir] + q, b, nf;
    if (!isInnerFaceOfDir(S, f, idir, &b, &nf)) continue;
    for (int v = 0; v < nv; ++v)
      S.phi_face[f * nv + v] = S.delta_phi_face_av[b * nv + v] - S.delta_phi_face_ar[nf * nv + v];
  }
}

// computeUremap with pente-borne (RemapADIService.cc:733-872): face flux
// n.e_dir L phi_face (phi_face is already the integral), dual flux as half
// the signed sum, non-negative masses and energies, then phi in the
// pente-borne form (fraction, density, velocity from the cell momentum,
// specific energy; the kinetic-energy slot with the total-energy fix-up;
// the viscosity slot kept) and the mixed / pure codes.
void computeUremapPB(mho_ctx& S, int idir) {
  const int nv = S.nvar, ne = S.nenv, nfc = 2 * S.dim;
  const R3 dirproj(0.5 * (1 - idir) * (2 - idir), 1.0 * idir * (2 - idir), -0.5 * idir * (1 - idir));
  std::fill(S.dual_phi_flux.begin(), S.dual_phi_flux.end(), 0.);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long cf[6];
    S.cellFaces(c, cf);
    double flux_face[3 * MHG_MAX_ENV + 8];
    for (int v = 0; v < nv; ++v) flux_face[v] = 0.;
    for (int index = 0; index < nfc; ++index) {
      const long f = cf[index];
      if (std::fabs(dot(S.face_normal[f], dirproj)) >= 1.0E-10) {
        const double face_length = S.face_length_lagrange[f][idir];
        const R3& ofn = S.outer_face_normal[c * nfc + index];
        const double ofn_dir = dot(ofn, dirproj);
        for (int v = 0; v < nv; ++v) {
          const double flux = ofn_dir * face_length * S.phi_face[f * nv + v];
          flux_face[v] += flux;
          S.dual_phi_flux[c * nv + v] += 0.5 * flux * ofn[idir];
        }
      }
    }
    double* U = &S.u_lagrange[c * nv];
    for (int v = 0; v < nv; ++v) U[v] -= flux_face[v];
    for (int k = 0; k < ne; ++k) {
      if (U[ne + k] < 0.) U[ne + k] = 0.;
      if (U[2 * ne + k] < 0.) U[2 * ne + k] = 0.;
    }
    double somme_volume = 0.;
    for (int k = 0; k < ne; ++k) somme_volume += U[k];
    double* phi = &S.phi_lagrange[c * nv];
    double somme_masse = 0.;
    for (int k = 0; k < ne; ++k) {
      phi[k] = U[k] / somme_volume;
      phi[ne + k] = (U[k] != 0.) ? U[ne + k] / U[k] : 0.;
      somme_masse += U[ne + k];
    }
    if (somme_masse != 0) {
      phi[3 * ne] = U[3 * ne] / somme_masse;
      phi[3 * ne + 1] = U[3 * ne + 1] / somme_masse;
    }
    for (int k = 0; k < ne; ++k) phi[2 * ne + k] = (U[ne + k] != 0.) ? U[2 * ne + k] / U[ne + k] : 0.;
    if (S.opt.conservation_energie_totale) phi[3 * ne + 2] = U[3 * ne + 2] / somme_masse;
    int imatcell = 0, imatpure = -1;
    for (int k = 0; k < ne; ++k)
      if (phi[k] > 0.) {
        imatcell++;
        imatpure = k;
      }
    S.est_mixte[c] = imatcell > 1 ? 1 : 0;
    S.est_pure[c] = imatcell > 1 ? -1 : imatpure;
  }
}

// Order-3 face values (RemapADIService.cc:583-640): six cells along idir
// around each inner face, the last existing cell standing in for the missing
// ones; at a face whose neighbourhood is not pure, an environment's mass and
// energy face values follow its volume face value through the upwind cell's
// ratios.
void computeUpwindFaceQuantities3(mho_ctx& S, int idir) {
  const int nv = S.nvar, ne = S.nenv;
  std::fill(S.phi_face.begin(), S.phi_face.end(), 0.);
  computeCellWidths(S, idir);
  const double deltat = S.global_deltat;
  const long delta = (idir == 0) ? 1 : (idir == 1 ? S.n[0] : (long)S.n[0] * S.n[1]);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < S.nfd[idir]; ++q) {
    long f = S.foff[idir] + q, b, fr;
    if (!isInnerFaceOfDir(S, f, idir, &b, &fr)) continue;
    int ib[3], iff[3];
    S.cellIdx(b, ib);
    S.cellIdx(fr, iff);
    const long bb = ib[idir] >= 1 ? b - delta : b;
    const long bbb = ib[idir] >= 2 ? b - 2 * delta : bb;
    const long ff = iff[idir] <= S.n[idir] - 2 ? fr + delta : fr;
    const long fff = iff[idir] <= S.n[idir] - 3 ? fr + 2 * delta : ff;
    const double vdt = S.face_normal_velocity[f] * deltat;
    const double* P = S.phi_lagrange.data();
    const double* h = S.h_cell.data();
    double* F = &S.phi_face[f * nv];
    for (int v = 0; v < nv; ++v)
      F[v] = o3_flux(P[bbb * nv + v], P[bb * nv + v], P[b * nv + v], P[fr * nv + v], P[ff * nv + v],
                     P[fff * nv + v], h[bbb], h[bb], h[b], h[fr], h[ff], h[fff], vdt);
    const bool voisinage_pure = S.est_mixte[b] == 0 && S.est_mixte[fr] == 0 && S.est_pure[b] == S.est_pure[fr];
    if (!voisinage_pure) {
      for (int k = 0; k < ne; ++k) {
        long up = -1;
        if (vdt > 0. && P[b * nv + k] != 0. && P[b * nv + ne + k] != 0.) up = b;
        else if (vdt < 0. && P[fr * nv + k] != 0. && P[fr * nv + ne + k] != 0.) up = fr;
        if (up < 0) continue;
        F[ne + k] = (P[up * nv + ne + k] / P[up * nv + k]) * F[k];
        F[2 * ne + k] = (P[up * nv + 2 * ne + k] / P[up * nv + ne + k]) * F[ne + k];
      }
    }
  }
}

// computeUpwindFaceQuantitiesForProjection_PBorn0_O2, RemapADIService.cc:658-720.
void computeUpwindFaceQuantities(mho_ctx& S, int idir) {
  const int nv = S.nvar;
  std::fill(S.phi_face.begin(), S.phi_face.end(), 0.);
  const int order2 = S.opt.ordre_projection - 1;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < S.nfd[idir]; ++q) {
    long f = S.foff[idir] + q, pc, nc;
    if (!isInnerFaceOfDir(S, f, idir, &pc, &nc)) continue;
    long up = (S.face_normal_velocity[f] * S.deltax_lagrange[f] > 0.0 ? pc : nc);
    double dot_xf = dot(S.face_coord[f] - S.cell_coord[up], S.face_normal[f]);
    for (int v = 0; v < nv; ++v)
      S.phi_face[f * nv + v] = S.phi_lagrange[up * nv + v] + order2 * dot_xf * S.grad_phi[up * nv + v];
  }
}

// computeUremap_PBorn0, RemapADIService.cc:883-1040.
void computeUremap(mho_ctx& S, int idir) {
  const int nv = S.nvar, ne = S.nenv;
  const double deltat = S.global_deltat;
  const int nfc = 2 * S.dim;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    long cf[6];
    S.cellFaces(c, cf);
    for (int v = 0; v < nv; ++v) {
      S.dual_phi_flux[c * nv + v] = 0.;
      double flux_face = 0.;
      for (int index = 0; index < nfc; ++index) {
        long f = cf[index];
        double nfi = S.face_normal[f][idir];
        if (std::fabs(nfi) >= 1.0E-10) {
          double face_normal_velocity = S.face_normal_velocity[f];
          double face_length = S.face_length_lagrange[f][idir];
          const R3& ofn = S.outer_face_normal[c * nfc + index];
          double ofn_dir = ofn[idir];
          double flux = ofn_dir * face_normal_velocity * face_length * deltat * S.phi_face[f * nv + v];
          flux_face += flux;
          S.dual_phi_flux[c * nv + v] += 0.5 * flux * ofn[idir];
        }
      }
      S.u_lagrange[c * nv + v] = S.u_lagrange[c * nv + v] - flux_face;
    }
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    double* U = &S.u_lagrange[c * nv];
    for (int k = 0; k < ne; ++k) {
      if (U[ne + k] < 0.) U[ne + k] = 0.;
      if (U[2 * ne + k] < 0.) U[2 * ne + k] = 0.;
    }
    double somme_volume = 0.;
    for (int k = 0; k < ne; ++k) somme_volume += U[k];
    for (int v = 0; v < nv; ++v) S.phi_lagrange[c * nv + v] = U[v] / somme_volume;
    int imatcell = 0, imatpure = -1;
    for (int k = 0; k < ne; ++k) {
      if (S.phi_lagrange[c * nv + k] > 0.) {
        imatcell++;
        imatpure = k;
      }
      if (imatcell > 1) {
        S.est_mixte[c] = 1;
        S.est_pure[c] = -1;
      } else {
        S.est_mixte[c] = 0;
        S.est_pure[c] = imatpure;
      }
    }
  }
}

// computeDualGradPhi_LimC, DualPhaseRemap.cc:700-845.
void computeDualGradPhi(mho_ctx& S, int idir) {
  const int nv = S.nvar;
  const int lim = classical_limiter(S.opt);
  long delta = (idir == 0) ? 1 : (idir == 1 ? S.nn[0] : (long)S.nn[0] * S.nn[1]);
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    int id[3];
    S.nodeIdx(nd, id);
    if (id[idir] < 1 || id[idir] > S.nn[idir] - 2) continue;  // inner nodes of idir
    long bn = nd - delta, fn = nd + delta;
    double gf[3], gb[3];
    for (int v = 0; v < 3; ++v) {
      gf[v] = (S.phi_dual_lagrange[fn * nv + v] - S.phi_dual_lagrange[nd * nv + v]) /
              (S.node_coord[fn][idir] - S.node_coord[nd][idir]);
      gb[v] = (S.phi_dual_lagrange[nd * nv + v] - S.phi_dual_lagrange[bn * nv + v]) /
              (S.node_coord[nd][idir] - S.node_coord[bn][idir]);
    }
    if (is_generalised(S.opt)) {
      // computeDualGradPhi / computeAndLimitGradPhiDual, UtilesRemap.cc:15-68,
      // 350-382: dual widths from the node coordinates along idir, the width
      // beyond a boundary node taken as 0.  (On 3 nodes along idir the
      // reference reads a missing node here; its width is taken as 0 too.)
      auto X = [&](long n) { return S.node_coord[n][idir]; };
      const bool bbe = id[idir] >= 2, ffe = id[idir] <= S.nn[idir] - 3;
      const double h0 = 0.5 * (X(fn) - X(bn));
      double hm, hp;
      if (!bbe) {
        hm = 0.;
        hp = ffe ? 0.5 * (X(fn + delta) - X(nd)) : 0.;
      } else if (!ffe) {
        hp = 0.;
        hm = 0.5 * (X(nd) - X(bn - delta));
      } else {
        hm = 0.5 * (X(nd) - X(bn - delta));
        hp = 0.5 * (X(fn + delta) - X(nd));
      }
      for (int v = 0; v < 3; ++v)
        S.dual_grad_phi[nd * nv + v] =
            flux_limiter_g(S.opt.projection_limiteur_id, gf[v], gb[v], S.phi_dual_lagrange[fn * nv + v],
                           S.phi_dual_lagrange[bn * nv + v], h0, hp, hm);
      continue;
    }
    for (int v = 0; v < 3; ++v) {
      double& g = S.dual_grad_phi[nd * nv + v];
      g = 0.;
      if (gb[v] != 0.) g += 0.5 * limiter(lim, gf[v] / gb[v]) * gb[v];
      if (gf[v] != 0.) g += 0.5 * limiter(lim, gb[v] / gf[v]) * gf[v];
    }
  }
}

// computeDualUremap, DualPhaseRemap.cc:22-676 (accelerated branch 127-523).
void computeDualUremap(mho_ctx& S, int idir) {
  const int nv = S.nvar, ne = S.nenv;
  const double deltat = S.global_deltat;
  if (S.opt.ordre_projection > 1 && has_limited_gradients(S.opt)) computeDualGradPhi(S, idir);
  const double oneovernbcell = (S.dim == 2 ? 0.5 : 0.25);
  // fcontrib (147-188): every cell is the back cell of its front face and the
  // front cell of its back face, so both contributions are set for every cell.
#pragma omp parallel for schedule(static)
  for (long q = 0; q < S.nfd[idir]; ++q) {
    long f = S.foff[idir] + q, pc, nc;
    S.faceCells(f, &pc, &nc);
    if (pc >= 0)
      for (int k = 0; k < ne; ++k)
        S.back_flux_contrib_env[pc * ne + k] = oneovernbcell * S.dual_phi_flux[pc * nv + ne + k];
    if (nc >= 0)
      for (int k = 0; k < ne; ++k)
        S.front_flux_contrib_env[nc * ne + k] = oneovernbcell * S.dual_phi_flux[nc * nv + ne + k];
  }
  // ncontrib (215-341)
  const int dir1 = (idir + 1) % S.dim;
  const int dir2 = (idir + 2) % S.dim;
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    int id[3];
    S.nodeIdx(nd, id);
    for (int k = 0; k < ne; ++k) {
      S.back_flux_mass_env[nd * ne + k] = 0.;
      S.front_flux_mass_env[nd * ne + k] = 0.;
    }
    S.back_flux_mass[nd] = 0.;
    S.front_flux_mass[nd] = 0.;
    const int s2lo = (S.dim == 3) ? -1 : 0;
    for (int pass = 0; pass < 2; ++pass) {  // 0 = back cells, 1 = front cells
      for (int sht1 = -1; sht1 <= 0; ++sht1)
        for (int sht2 = s2lo; sht2 <= 0; ++sht2) {
          int ci[3] = {id[0], id[1], id[2]};
          if (pass == 0) ci[idir] -= 1;
          ci[dir1] += sht1;
          if (S.dim == 3) ci[dir2] += sht2;
          bool ok = ci[0] >= 0 && ci[0] < S.n[0] && ci[1] >= 0 && ci[1] < S.n[1] &&
                    ci[2] >= 0 && ci[2] < S.n[2];
          if (!ok) continue;
          long c = S.cellId(ci[0], ci[1], ci[2]);
          for (int k = 0; k < ne; ++k) {
            if (pass == 0) {
              S.back_flux_mass_env[nd * ne + k] += S.back_flux_contrib_env[c * ne + k];
              S.back_flux_mass[nd] += S.back_flux_contrib_env[c * ne + k];
            } else {
              S.front_flux_mass_env[nd * ne + k] += S.front_flux_contrib_env[c * ne + k];
              S.front_flux_mass[nd] += S.front_flux_contrib_env[c * ne + k];
            }
          }
        }
    }
  }
  // upwind dual update (352-491), inner nodes of idir
  const int order2 = S.opt.ordre_projection - 1;
  long delta = (idir == 0) ? 1 : (idir == 1 ? S.nn[0] : (long)S.nn[0] * S.nn[1]);
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    int id[3];
    S.nodeIdx(nd, id);
    if (id[idir] < 1 || id[idir] > S.nn[idir] - 2) continue;
    long bn = nd - delta, fn = nd + delta;
    const double* Pd = S.phi_dual_lagrange.data();
    const double* G = S.dual_grad_phi.data();
    double* Ud = &S.u_dual_lagrange[nd * nv];
    double bfm = S.back_flux_mass[nd], ffm = S.front_flux_mass[nd];
    Ud[3] += bfm - ffm;
    double FV[3] = {0., 0., 0.}, BV[3] = {0., 0., 0.};
    double FE = 0., BE = 0.;
    int signfront;
    double ufront = 0.5 * (Pd[nd * nv + idir] + Pd[nd * nv + idir]);
    if (ufront > 0) signfront = 1;
    else signfront = -1;
    double xf = S.node_coord[fn][idir] - S.node_coord[nd][idir];
    long src = (ffm < 0.) ? fn : nd;
    for (int v = 0; v < 3; ++v)
      FV[v] = Pd[src * nv + v] + order2 * (0.5 * G[src * nv + v] * (signfront * xf - deltat * ufront));
    FE = Pd[src * nv + 4] + order2 * (0.5 * G[src * nv + 4] * (signfront * xf - deltat * ufront));
    int signback;
    double uback = 0.5 * (Pd[bn * nv + idir] + Pd[nd * nv + idir]);
    if (uback > 0) signback = 1;
    else signback = -1;
    double xb = S.node_coord[nd][idir] - S.node_coord[bn][idir];
    long srcb = (bfm > 0.) ? bn : nd;
    for (int v = 0; v < 3; ++v)
      BV[v] = Pd[srcb * nv + v] + order2 * (0.5 * G[srcb * nv + v] * (signback * xb - deltat * uback));
    BE = Pd[srcb * nv + 4] + order2 * (0.5 * G[srcb * nv + 4] * (signback * xb - deltat * uback));
    for (int v = 0; v < 3; ++v) Ud[v] += bfm * BV[v] - ffm * FV[v];
    Ud[4] += bfm * BE - ffm * FE;
  }
  // finalisation (494-523)
  const double thresold = 1.e-300;
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    int id[3];
    S.nodeIdx(nd, id);
    if (id[idir] < 1 || id[idir] > S.nn[idir] - 2) continue;
    double* Ud = &S.u_dual_lagrange[nd * nv];
    double* Pd = &S.phi_dual_lagrange[nd * nv];
    if (std::fabs(Ud[0]) < thresold) Ud[0] = 0.;
    if (std::fabs(Ud[1]) < thresold) Ud[1] = 0.;
    if (std::fabs(Ud[2]) < thresold) Ud[2] = 0.;
    Pd[0] = Ud[0] / Ud[3];
    Pd[1] = Ud[1] / Ud[3];
    Pd[2] = Ud[2] / Ud[3];
    Pd[3] = Ud[3];
    Pd[4] = Ud[4] / Ud[3];
  }
}

// remapVariables, RemapADIFinal.cc:19-778.
void remapVariables(mho_ctx& S) {
  const int nv = S.nvar, ne = S.nenv;
  const double threshold = S.opt.remap_threshold;
  // cell status per environment (98-143) and membership change (145-226)
  std::vector<int> new_mask(S.env_mask);
  for (int k = 0; k < ne; ++k) {
    for (long c = 0; c < S.NC; ++c) {
      const double* U = &S.u_lagrange[c * nv];
      double fvol = U[k] / S.euler_volume[c];
      bool cell_in_env = (S.env_mask[c] & (1 << k)) != 0;
      int cell_status = 0;
      if (fvol < threshold && cell_in_env) {
        cell_status = -1;
      } else if (fvol > threshold && !cell_in_env) {
        double volt = 0.;
        for (int kk = 0; kk < ne; ++kk) volt += U[kk];
        double fvol_norm = U[k] / volt;
        if (fvol_norm > threshold && U[ne + k] != 0.) cell_status = +1;
      }
      S.cell_status[c] = cell_status;
      if (cell_status < 0) new_mask[c] &= ~(1 << k);
      if (cell_status > 0) new_mask[c] |= (1 << k);
    }
  }
  S.env_mask.swap(new_mask);
  computeEnvId(S);
  // "moy" (446-643)
#pragma omp parallel for schedule(static)
  for (long c = 0; c < S.NC; ++c) {
    const double* U = &S.u_lagrange[c * nv];
    int ev[MHG_MAX_ENV];
    int nbe = S.envsOf(c, ev);
    double vol = S.euler_volume[c];
    S.cell_volume.g[c] = vol;
    double volt = 0., masset = 0.;
    for (int k = 0; k < ne; ++k) {
      volt += U[k];
      masset += U[ne + k];
    }
    double unsurvolt = 1. / volt;
    double somme_frac = 0.;
    double unsurvol = 1. / vol;
    for (int q = 0; q < nbe; ++q) {
      int k = ev[q];
      double fvol = U[k] * unsurvolt;
      if (fvol < threshold) fvol = 0.;
      S.EV(S.fracvol, k, c) = fvol;
      somme_frac += fvol;
    }
    int matcell = 0, imatpure = -1;
    double unsursomme_frac = 1. / somme_frac;
    // learned mode (extension): an environment whose fraction ends at or
    // below the threshold leaves the cell, so that no member has a zero
    // fraction and density (the EOS would give c = sqrt(0/0) and the next
    // Lagrange phase rho_k = 0/0); the reference keeps it (membership tested
    // on U_V / V_euler, the fraction on U_V / sum U_V)
    const bool prune = is_learned(S.opt);
    bool pruned = false;
    for (int q = 0; q < nbe; ++q) {
      int k = ev[q];
      double fvol = S.EV(S.fracvol, k, c) * unsursomme_frac;
      if (prune && fvol <= threshold) {
        fvol = 0.;
        S.env_mask[c] &= ~(1 << k);
        pruned = true;
      }
      if (fvol > 0.) {
        matcell++;
        imatpure = k;
      }
      S.EV(S.fracvol, k, c) = fvol;
    }
    if (pruned) nbe = S.envsOf(c, ev);
    if (matcell > 1) {
      S.est_mixte[c] = 1;
      S.est_pure[c] = -1;
    } else {
      S.est_mixte[c] = 0;
      S.est_pure[c] = imatpure;
    }
    if (masset != 0.) {
      double unsurmasset = 1. / masset;
      double fmasset = 0.;
      for (int q = 0; q < nbe; ++q) {
        int k = ev[q];
        double mf = U[ne + k] * unsurmasset;
        if (S.EV(S.fracvol, k, c) < threshold) mf = 0.;
        fmasset += mf;
        S.EV(S.mass_fraction, k, c) = mf;
      }
      if (fmasset != 0.) {
        double unsurfmasset = 1. / fmasset;
        for (int q = 0; q < nbe; ++q) {
          int k = ev[q];
          S.EV(S.mass_fraction, k, c) = S.EV(S.mass_fraction, k, c) * unsurfmasset;
        }
      }
    }
    double density_nplus1 = 0.;
    for (int q = 0; q < nbe; ++q) {
      int k = ev[q];
      double dk = 0.;
      double fvol = S.EV(S.fracvol, k, c);
      if (fvol > threshold) {
        double vol_norm = U[k] * vol * unsurvolt;
        dk = U[ne + k] / vol_norm;
      }
      S.EV(S.density, k, c) = dk;
      density_nplus1 += fvol * dk;
    }
    if (S.est_mixte[c]) S.density.g[c] = density_nplus1;
    S.cell_mass.g[c] = S.euler_volume[c] * density_nplus1;
    double energie_nplus1 = 0., pseudo_nplus1 = 0.;
    for (int q = 0; q < nbe; ++q) {
      int k = ev[q];
      double fvol = S.EV(S.fracvol, k, c);
      double mf = S.EV(S.mass_fraction, k, c);
      double ek = 0.;
      if (fvol > threshold && U[ne + k] != 0.) ek = U[2 * ne + k] / U[ne + k];
      S.EV(S.internal_energy, k, c) = ek;
      energie_nplus1 += mf * ek;
      S.EV(S.cell_mass, k, c) = mf * S.cell_mass.g[c];
      double pv = U[3 * ne + 4] * unsurvol;
      S.EV(S.pseudo_viscosity, k, c) = pv;
      pseudo_nplus1 += fvol * pv;
    }
    S.internal_energy.g[c] = energie_nplus1;
    S.pseudo_viscosity.g[c] = pseudo_nplus1;
    for (int q = 0; q < nbe; ++q) {
      int k = ev[q];
      if (S.EV(S.density, k, c) < 0. || S.EV(S.internal_energy, k, c) < 0.) {
        S.EV(S.internal_energy, k, c) = 0.;
        S.EV(S.density, k, c) = 0.;
        S.EV(S.fracvol, k, c) = 0.;
        S.EV(S.mass_fraction, k, c) = 0.;
      }
    }
  }
  if (is_learned(S.opt)) computeEnvId(S);  // after the learned mode's pruning
  // velocity from the dual (660-672)
#pragma omp parallel for schedule(static)
  for (long nd = 0; nd < S.NN; ++nd) {
    const double* Ud = &S.u_dual_lagrange[nd * nv];
    double mp = Ud[3];
    if (mp != 0.) {
      S.velocity[nd].x = Ud[0] / mp;
      S.velocity[nd].y = Ud[1] / mp;
      S.velocity[nd].z = Ud[2] / mp;
    }
  }
  computeNodeMass(S);
  // mixed V_k = f_k V (722-744)
  for (int k = 0; k < ne; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) S.cell_volume.env(k, c) = S.fracvol.env(k, c) * S.cell_volume.g[c];
  // total-energy fix-up (751-776): per node of the cell, a quarter of the
  // kinetic energy the dual remap carried (U_4 / U_3, from 1/2 m |u| at the
  // start of the remap) minus 1/2 |u|^2 of the remapped velocity; the sum,
  // where positive, goes into the internal energies (by mass fraction) and,
  // for a cell with other than one environment, into the cell's average.
  if (S.opt.conservation_energie_totale) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < S.NC; ++c) {
      long nodes[8];
      const int nb = S.cellNodes(c, nodes);
      double delta_ec = 0.;
      for (int i = 0; i < nb; ++i) {
        const double* Ud = &S.u_dual_lagrange[nodes[i] * nv];
        if (Ud[3] != 0.) {
          const R3& u = S.velocity[nodes[i]];
          const double ec_proj = Ud[4] / Ud[3];
          const double ec_reconst = 0.5 * (u.x * u.x + u.y * u.y + u.z * u.z);
          delta_ec += 0.25 * (ec_proj - ec_reconst);
        }
      }
      delta_ec = smax(0., delta_ec);
      int ev[MHG_MAX_ENV];
      const int nbe = S.envsOf(c, ev);
      for (int q = 0; q < nbe; ++q)
        S.EV(S.internal_energy, ev[q], c) += S.EV(S.mass_fraction, ev[q], c) * delta_ec;
      if (nbe != 1) S.internal_energy.g[c] += delta_ec;
    }
  }
}

// appliRemap, RemapADIService.cc:19-60.
void appliRemap(mho_ctx& S) {
  for (int i = 0; i < S.dim; ++i) {
    int idir = (i + S.sens_projection) % S.dim;
    computeGradPhiFace(S, idir);
    computeGradPhiCell(S, idir);
    if (S.opt.ordre_projection == 3) computeUpwindFaceQuantities3(S, idir);
    else if (is_pb(S.opt)) computeUpwindFaceQuantitiesPB(S, idir);
    else computeUpwindFaceQuantities(S, idir);
    if (is_pb(S.opt)) computeUremapPB(S, idir);
    else computeUremap(S, idir);
    computeDualUremap(S, idir);
  }
  S.sens_projection = (S.sens_projection + 1) % S.dim;
  remapVariables(S);
}

// remap, PrepareRemap.cc:448-563.
void remap(mho_ctx& S) {
  if (!S.opt.with_projection) return;
  computeVariablesForRemap(S);
  computeFaceQuantitesForRemap(S);
  appliRemap(S);
  // reset of the _n fields (485-550)
  for (long c = 0; c < S.NC; ++c) {
    S.pseudo_viscosity_n.g[c] = 0.;
    S.pressure_n.g[c] = 0.;
    S.cell_volume_n.g[c] = 0.;
    S.density_n.g[c] = 0.;
    S.internal_energy_n.g[c] = 0.;
    S.tau_density.g[c] = 0.;
  }
  for (int k = 0; k < S.nenv; ++k)
    for (long c = 0; c < S.NC; ++c)
      if (S.isImpure(c, k)) {
        S.pseudo_viscosity_n.env(k, c) = 0.;
        S.pressure_n.env(k, c) = 0.;
        S.cell_volume_n.env(k, c) = 0.;
        S.density_n.env(k, c) = 0.;
        S.internal_energy_n.env(k, c) = 0.;
        S.tau_density.env(k, c) = 0.;
      }
  for (int k = 0; k < S.nenv; ++k) applyEOS(S, k);
  computePressionMoyenne(S);
}

// computeHydroDeltaT (MahycoModule.cc:2027-2071) + computeDeltaT (2076-2191).
int computeDeltaT(mho_ctx& S, double* out_dt, int* stop) {
  S.global_old_deltat = S.global_deltat;
  S.old_deltat = S.global_old_deltat;
  const bool with_projection = S.opt.with_projection != 0;
  double minimum_aux = std::numeric_limits<double>::max();
#pragma omp parallel for schedule(static) reduction(min : minimum_aux)
  for (long c = 0; c < S.NC; ++c) {
    double cell_dx = S.caracteristic_length[c];
    double sound_speed = S.sound_speed.g[c];
    double vmax = 0.;
    if (with_projection) {
      long nodes[8];
      int nb = S.cellNodes(c, nodes);
      for (int i = 0; i < nb; ++i) {
        double un = normL2(S.velocity[nodes[i]]);
        vmax = (un < vmax) ? vmax : un;  // math::max(|u|, vmax)
      }
    }
    double dx_sound = cell_dx / (sound_speed + vmax);
    minimum_aux = std::min(minimum_aux, dx_sound);
  }
  double new_dt = S.opt.cfl * minimum_aux;
  new_dt = std::min(new_dt, 1.05 * S.global_old_deltat);
  new_dt = std::min(new_dt, S.opt.deltat_max);
  *stop = 0;
  if (new_dt < S.opt.deltat_min) {
    *out_dt = new_dt;
    return MHG_E_DT_MIN;
  }
  const double stop_time = S.opt.final_time;
  if (S.opt.land_on_final_time) {  // extension: the last step ends at final_time
    if (S.landing || S.global_time >= stop_time) *stop = 1;
    else if (S.global_time + new_dt >= stop_time) {
      new_dt = stop_time - S.global_time;
      S.landing = true;
    }
    S.global_deltat = new_dt;
    *out_dt = new_dt;
    return MHG_OK;
  }
  S.global_deltat = new_dt;
  bool not_yet_finish = (S.global_time < stop_time);
  bool finish = (S.global_time > stop_time);
  bool too_much = ((S.global_time + new_dt) > stop_time);
  if (not_yet_finish && too_much) *stop = 1;
  if (finish) *stop = 1;
  *out_dt = new_dt;
  return MHG_OK;
}

}  // namespace

// ===========================================================================
// C interface
// ===========================================================================
namespace {

void allocAll(mho_ctx& S) {
  const long NC = S.NC, NN = S.NN, NF = S.NF;
  const int ne = S.nenv, nv = S.nvar;
  MatVar* mv[] = {&S.density, &S.density_n, &S.pressure, &S.pressure_n, &S.pseudo_viscosity,
                  &S.pseudo_viscosity_n, &S.pseudo_viscosity_nmoins1, &S.mass_fraction,
                  &S.cell_mass, &S.internal_energy, &S.internal_energy_n, &S.dpde, &S.fracvol,
                  &S.cell_volume, &S.cell_volume_n, &S.sound_speed, &S.tau_density};
  for (MatVar* m : mv) m->alloc(NC, ne);
  S.euler_volume.assign(NC, 0.);
  S.div_u.assign(NC, 0.);
  S.caracteristic_length.assign(NC, 0.);
  S.env_id.assign(NC, 0);
  S.env_mask.assign(NC, 0);
  S.est_mixte.assign(NC, 0);
  S.est_pure.assign(NC, 0);
  S.cell_status.assign(NC, 0);
  S.cell_coord.assign(NC, R3());
  S.cell_cqs.assign((size_t)NC * S.nbnode, R3());
  S.cell_cqs_n.assign((size_t)NC * S.nbnode, R3());
  S.outer_face_normal.assign((size_t)NC * 2 * S.dim, R3());
  S.u_lagrange.assign((size_t)NC * nv, 0.);
  S.phi_lagrange.assign((size_t)NC * nv, 0.);
  S.grad_phi.assign((size_t)NC * nv, 0.);
  S.dual_phi_flux.assign((size_t)NC * nv, 0.);
  S.back_flux_contrib_env.assign((size_t)NC * ne, 0.);
  S.front_flux_contrib_env.assign((size_t)NC * ne, 0.);
  S.velocity.assign(NN, R3());
  S.velocity_n.assign(NN, R3());
  S.node_coord.assign(NN, R3());
  S.node_coord_0.assign(NN, R3());
  S.force.assign(NN, R3());
  S.node_mass.assign(NN, 0.);
  S.u_dual_lagrange.assign((size_t)NN * nv, 0.);
  S.phi_dual_lagrange.assign((size_t)NN * nv, 0.);
  S.dual_grad_phi.assign((size_t)NN * nv, 0.);
  S.back_flux_mass.assign(NN, 0.);
  S.front_flux_mass.assign(NN, 0.);
  S.back_flux_mass_env.assign((size_t)NN * ne, 0.);
  S.front_flux_mass_env.assign((size_t)NN * ne, 0.);
  S.face_coord.assign(NF, R3());
  S.face_normal.assign(NF, R3());
  S.face_length_lagrange.assign(NF, R3());
  S.face_normal_velocity.assign(NF, 0.);
  S.deltax_lagrange.assign(NF, 0.);
  S.grad_phi_face.assign((size_t)NF * nv, 0.);
  S.phi_face.assign((size_t)NF * nv, 0.);
  S.h_cell.assign(NC, 0.);
  S.delta_phi_face_ar.assign((size_t)NC * nv, 0.);
  S.delta_phi_face_av.assign((size_t)NC * nv, 0.);
}

// Mesh generation and the geometry-only part of hydroStartInit.
void buildMesh(mho_ctx& S) {
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = (d < S.dim) ? S.len[d] / S.n[d] : 0.;
  for (int k = 0; k < S.nn[2]; ++k)
    for (int j = 0; j < S.nn[1]; ++j)
      for (int i = 0; i < S.nn[0]; ++i) {
        R3 x(S.org[0] + i * h[0], S.org[1] + j * h[1], S.dim == 3 ? S.org[2] + k * h[2] : 0.);
        S.node_coord[S.nodeId(i, j, k)] = x;
      }
  cellCoordFromNodes(S);
  computeGeometricValues(S);
  for (long c = 0; c < S.NC; ++c) S.euler_volume[c] = S.cell_volume.g[c];
  initGeometricValues(S);
  // the boundary face groups of the initial mesh (PrepareFaceGroup runs at
  // init on it), so that a context given its state by mho_upload_state
  // without an init call has them too
  prepareBoundaryConditions(S);
  S.global_deltat = S.opt.deltat_init;
}

void resetPhysics(mho_ctx& S) {
  MatVar* mv[] = {&S.density, &S.density_n, &S.pressure, &S.pressure_n, &S.pseudo_viscosity,
                  &S.pseudo_viscosity_n, &S.pseudo_viscosity_nmoins1, &S.mass_fraction,
                  &S.cell_mass, &S.internal_energy, &S.internal_energy_n, &S.dpde, &S.fracvol,
                  &S.cell_volume_n, &S.sound_speed, &S.tau_density};
  for (MatVar* m : mv) {
    std::fill(m->g.begin(), m->g.end(), 0.);
    std::fill(m->e.begin(), m->e.end(), 0.);
  }
  for (long nd = 0; nd < S.NN; ++nd) {
    S.node_coord[nd] = S.node_coord_0[nd];
    S.velocity[nd] = R3();
    S.velocity_n[nd] = R3();
  }
  std::fill(S.dual_grad_phi.begin(), S.dual_grad_phi.end(), 0.);
  std::fill(S.grad_phi_face.begin(), S.grad_phi_face.end(), 0.);
  std::fill(S.env_mask.begin(), S.env_mask.end(), 0);
  // the remap's cell variables as Arcane creates them (zero): pente-borne
  // reads phi of environments absent from a cell, -mixte the mixed / pure codes
  std::fill(S.phi_lagrange.begin(), S.phi_lagrange.end(), 0.);
  std::fill(S.est_mixte.begin(), S.est_mixte.end(), 0);
  std::fill(S.est_pure.begin(), S.est_pure.end(), 0);
  S.global_deltat = S.opt.deltat_init;
  S.global_old_deltat = S.old_deltat = S.global_time = 0.;
  S.global_iteration = 0;
  S.landing = false;
  S.sens_projection = 0;
  cellCoordFromNodes(S);
}

int checkReady(mho_ctx* ctx) {
  if (!ctx) return MHG_E_ARG;
  if (!ctx->initialised) return fail(ctx, MHG_E_STATE, "context not initialised");
  return MHG_OK;
}

}  // namespace

extern "C" {

int mho_abi_version(void) { return MHG_ABI_VERSION; }

void mho_default_options(mhg_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->cfl = 0.3;
  o->threshold = 1.e-10;
  o->deltat_init = 1.e-5;   // Data.1.arc:68
  o->deltat_min = 1.e-8;    // Data.1.arc:69
  o->deltat_max = 0.01;     // Data.1.arc:70
  o->final_time = 0.2;      // Data.1.arc:73
  o->with_projection = 1;
  o->longueur_caracteristique = MHG_LC_RACINE_CUBIQUE_VOLUME;
  // Data.1.arc:75-104: reflecting walls
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

int mho_create(const mhg_mesh_desc* mesh, const mhg_options* opt, mho_ctx** out) {
  if (!mesh || !opt || !out) return MHG_E_ARG;
  *out = nullptr;
  if (mesh->dim != 2 && mesh->dim != 3) return MHG_E_ARG;
  for (int d = 0; d < mesh->dim; ++d)
    if (mesh->n[d] < 1 || !(mesh->length[d] > 0.)) return MHG_E_ARG;
  if (opt->nb_env < 1 || opt->nb_env > MHG_MAX_ENV || opt->nb_bc < 0 || opt->nb_bc > MHG_MAX_BC)
    return MHG_E_ARG;
  if (opt->schema_csts || opt->with_newton || opt->sans_lagrange || !opt->is_euler_scheme ||
      !has_limited_gradients(*opt) || opt->projection_limiteur_id < 0 || opt->ordre_projection < 1 ||
      opt->ordre_projection > 3)
    return MHG_E_UNSUPPORTED;
  // pente-borne: order 2 (at order 1 the reference's generic face loop moves
  // nothing, at order 3 it mixes a face value with the integrated-flux
  // formula), classical or generalised limiters, debar fix 0..2
  if ((opt->projection_pente_borne || opt->projection_pente_borne_mixte) &&
      (is_learned(*opt) || opt->projection_limiteur_pure_id < 0 ||
       opt->projection_limiteur_pure_id > MHG_LIM_ULTRABEEG))
    return MHG_E_UNSUPPORTED;
  if (opt->projection_pente_borne &&
      (opt->ordre_projection != 2 || opt->projection_pente_borne_debar_fix < 0 ||
       opt->projection_pente_borne_debar_fix > 2))
    return MHG_E_UNSUPPORTED;
  for (int b = 0; b < opt->nb_bc; ++b)
    if (opt->bc[b].surface < 0 || opt->bc[b].surface > 5 || opt->bc[b].type < 0 ||
        opt->bc[b].type > 2)
      return MHG_E_ARG;
  mho_ctx* S = new mho_ctx();
  S->opt = *opt;
  S->dim = mesh->dim;
  for (int d = 0; d < 3; ++d) {
    S->n[d] = (d < S->dim) ? mesh->n[d] : 1;
    S->nn[d] = (d < S->dim) ? mesh->n[d] + 1 : 1;
    S->org[d] = mesh->origin[d];
    S->len[d] = mesh->length[d];
  }
  S->NC = (long)S->n[0] * S->n[1] * S->n[2];
  S->NN = (long)S->nn[0] * S->nn[1] * S->nn[2];
  long off = 0;
  for (int d = 0; d < 3; ++d) {
    for (int e = 0; e < 3; ++e) S->fdim[d][e] = S->n[e] + ((e == d && d < S->dim) ? 1 : 0);
    S->foff[d] = off;
    S->nfd[d] = (d < S->dim) ? (long)S->fdim[d][0] * S->fdim[d][1] * S->fdim[d][2] : 0;
    off += S->nfd[d];
  }
  S->NF = off;
  S->nbnode = (S->dim == 2) ? 4 : 8;
  S->nbface = 2 * S->dim;
  S->nenv = opt->nb_env;
  S->nvar = 3 * S->nenv + 3 + 1 + 1;  // MahycoAnnexe.cc:90
  for (int k = 0; k < S->nenv; ++k) {
    S->gamma_env.push_back(opt->env[k].adiabatic_cst);
    S->ptens_env.push_back(opt->env[k].limit_tension);
    S->eos_env.push_back(opt->env[k].eos);
  }
  allocAll(*S);
  buildMesh(*S);
  *out = S;
  return MHG_OK;
}

int mho_destroy(mho_ctx* ctx) {
  delete ctx;
  return MHG_OK;
}

const char* mho_last_error(const mho_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mho_sizes(const mho_ctx* ctx, int64_t* n_cells, int64_t* n_nodes) {
  if (!ctx) return MHG_E_ARG;
  if (n_cells) *n_cells = ctx->NC;
  if (n_nodes) *n_nodes = ctx->NN;
  return MHG_OK;
}

int mho_set_threads(int32_t n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// SODService::initMat / initVar (SODService.cc:3-135), then finishInit.
int mho_init_sod(mho_ctx* ctx, int32_t cas_test) {
  if (!ctx) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  bool mono = (cas_test == 3 || cas_test == 4 || cas_test == 5);
  bool bi = (cas_test == 13 || cas_test == 14 || cas_test == 15);
  if (!mono && !bi) return fail(ctx, MHG_E_ARG, "unsupported SOD cas-test");
  if (bi && S.nenv < 2) return fail(ctx, MHG_E_ARG, "bi-material SOD needs 2 environments");
  int axis = mono ? cas_test - 3 : cas_test - 13;
  if (axis >= S.dim) return fail(ctx, MHG_E_ARG, "SOD axis beyond mesh dimension");
  resetPhysics(S);
  for (long c = 0; c < S.NC; ++c) {
    double r = S.cell_coord[c][axis];
    int env = mono ? 0 : (r < 0.5 ? 0 : 1);
    S.env_mask[c] = 1 << env;
    S.density.g[c] = (r < 0.5) ? 1.0 : 0.125;
    S.pressure.g[c] = (r < 0.5) ? 1.0 : 0.1;
    S.fracvol.g[c] = 1.;
    S.mass_fraction.g[c] = 1.;
  }
  finishInit(S);
  return MHG_OK;
}

int mho_init_triple_point(mho_ctx* ctx, const double rho[3], const double p[3]) {
  if (!ctx || !rho || !p) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  if (S.nenv != 3) return fail(ctx, MHG_E_ARG, "triple point needs 3 environments");
  resetPhysics(S);
  for (long c = 0; c < S.NC; ++c) {
    const R3& x = S.cell_coord[c];
    int env = (x.x < 1.) ? 0 : (x.y < 1.5 ? 1 : 2);
    S.env_mask[c] = 1 << env;
    S.density.g[c] = rho[env];
    S.pressure.g[c] = p[env];
    S.fracvol.g[c] = 1.;
    S.mass_fraction.g[c] = 1.;
  }
  finishInit(S);
  return MHG_OK;
}

int mho_init_cells(mho_ctx* ctx, const int32_t* env_of_cell, const double* density,
                   const double* pressure, const double* velocity) {
  if (!ctx || !env_of_cell || !density || !pressure) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  for (long c = 0; c < S.NC; ++c)
    if (env_of_cell[c] < 0 || env_of_cell[c] >= S.nenv) return fail(ctx, MHG_E_ARG, "bad env index");
  resetPhysics(S);
  for (long c = 0; c < S.NC; ++c) {
    S.env_mask[c] = 1 << env_of_cell[c];
    S.density.g[c] = density[c];
    S.pressure.g[c] = pressure[c];
    S.fracvol.g[c] = 1.;
    S.mass_fraction.g[c] = 1.;
  }
  if (velocity)
    for (long nd = 0; nd < S.NN; ++nd)
      for (int d = 0; d < S.dim; ++d) S.velocity[nd][d] = velocity[(size_t)d * S.NN + nd];
  finishInit(S);
  return MHG_OK;
}

int mho_upload_state(mho_ctx* ctx, const mhg_state* st) {
  if (!ctx || !st) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  const long NC = S.NC, NN = S.NN;
  if (st->env_mask)
    for (long c = 0; c < NC; ++c) S.env_mask[c] = st->env_mask[c];
  computeEnvId(S);
  // pente-borne's phi of absent environments is not part of the state: an
  // upload starts it from zero
  std::fill(S.phi_lagrange.begin(), S.phi_lagrange.end(), 0.);
  struct P { double* src; double* srce; MatVar* mv; };
  P ps[] = {{st->density, st->density_env, &S.density},
            {st->internal_energy, st->internal_energy_env, &S.internal_energy},
            {st->pressure, st->pressure_env, &S.pressure},
            {st->sound_speed, st->sound_speed_env, &S.sound_speed},
            {st->pseudo_viscosity, st->pseudo_viscosity_env, &S.pseudo_viscosity},
            {st->cell_mass, st->cell_mass_env, &S.cell_mass},
            {st->fracvol, st->fracvol_env, &S.fracvol},
            {st->mass_fraction, st->mass_fraction_env, &S.mass_fraction}};
  for (auto& q : ps) {
    if (q.src) std::memcpy(q.mv->g.data(), q.src, sizeof(double) * NC);
    if (q.srce) std::memcpy(q.mv->e.data(), q.srce, sizeof(double) * NC * S.nenv);
  }
  for (long nd = 0; nd < NN; ++nd)
    for (int d = 0; d < S.dim; ++d) {
      if (st->velocity) S.velocity[nd][d] = st->velocity[(size_t)d * NN + nd];
      if (st->node_coord) S.node_coord[nd][d] = st->node_coord[(size_t)d * NN + nd];
      if (st->dual_grad_phi) S.dual_grad_phi[nd * S.nvar + d] = st->dual_grad_phi[(size_t)d * NN + nd];
    }
  if (st->node_mass) std::memcpy(S.node_mass.data(), st->node_mass, sizeof(double) * NN);
  if (st->caracteristic_length)
    std::memcpy(S.caracteristic_length.data(), st->caracteristic_length, sizeof(double) * NC);
  // corner normals at the uploaded coordinates (they are what cqs_n will hold)
#pragma omp parallel for schedule(static)
  for (long c = 0; c < NC; ++c) {
    long nodes[8];
    int nb = S.cellNodes(c, nodes);
    R3 coord[8];
    for (int i = 0; i < nb; ++i) coord[i] = S.node_coord[nodes[i]];
    if (S.dim == 3) {
      R3 fc[6];
      hexFaceCoords(coord, fc);
      computeCQs(coord, fc, &S.cell_cqs[c * S.nbnode]);
    } else {
      computeCQs2D(coord, &S.cell_cqs[c * S.nbnode]);
    }
  }
  // mixed / pure flags (read by the order-3 face values) from the uploaded
  // fractions: mixed if two or more environments have f > 0, else the last
  // environment with f > 0 (-1: none), as remapVariables sets them
  // (RemapADIFinal.cc:270-293)
  for (long c = 0; c < NC; ++c) {
    int ev[MHG_MAX_ENV];
    const int nbe = S.envsOf(c, ev);
    int matcell = 0, imatpure = -1;
    for (int q = 0; q < nbe; ++q)
      if (S.EV(S.fracvol, ev[q], c) > 0.) {
        matcell++;
        imatpure = ev[q];
      }
    S.est_mixte[c] = matcell > 1 ? 1 : 0;
    S.est_pure[c] = matcell > 1 ? -1 : imatpure;
  }
  S.global_deltat = st->global_deltat;
  S.old_deltat = st->old_deltat;
  S.global_time = st->global_time;
  S.landing = false;
  S.sens_projection = st->sens_projection;
  S.global_iteration = st->global_iteration;
  S.initialised = true;
  return MHG_OK;
}

int mho_download_state(mho_ctx* ctx, mhg_state* st) {
  if (!ctx || !st) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  const long NC = S.NC, NN = S.NN;
  if (st->env_mask)
    for (long c = 0; c < NC; ++c) st->env_mask[c] = S.env_mask[c];
  struct P { double* dst; double* dste; MatVar* mv; };
  P ps[] = {{st->density, st->density_env, &S.density},
            {st->internal_energy, st->internal_energy_env, &S.internal_energy},
            {st->pressure, st->pressure_env, &S.pressure},
            {st->sound_speed, st->sound_speed_env, &S.sound_speed},
            {st->pseudo_viscosity, st->pseudo_viscosity_env, &S.pseudo_viscosity},
            {st->cell_mass, st->cell_mass_env, &S.cell_mass},
            {st->fracvol, st->fracvol_env, &S.fracvol},
            {st->mass_fraction, st->mass_fraction_env, &S.mass_fraction}};
  for (auto& q : ps) {
    if (q.dst) std::memcpy(q.dst, q.mv->g.data(), sizeof(double) * NC);
    if (q.dste) {
      std::memcpy(q.dste, q.mv->e.data(), sizeof(double) * NC * S.nenv);
      // pure cells: the environment value is the cell value
      for (long c = 0; c < NC; ++c)
        if (S.nbEnvOf(c) == 1) q.dste[(size_t)S.env_id[c] * NC + c] = q.mv->g[c];
    }
  }
  for (long nd = 0; nd < NN; ++nd)
    for (int d = 0; d < S.dim; ++d) {
      if (st->velocity) st->velocity[(size_t)d * NN + nd] = S.velocity[nd][d];
      if (st->node_coord) st->node_coord[(size_t)d * NN + nd] = S.node_coord[nd][d];
      if (st->dual_grad_phi) st->dual_grad_phi[(size_t)d * NN + nd] = S.dual_grad_phi[nd * S.nvar + d];
    }
  if (st->node_mass) std::memcpy(st->node_mass, S.node_mass.data(), sizeof(double) * NN);
  if (st->caracteristic_length)
    std::memcpy(st->caracteristic_length, S.caracteristic_length.data(), sizeof(double) * NC);
  st->global_deltat = S.global_deltat;
  st->old_deltat = S.old_deltat;
  st->global_time = S.global_time;
  st->sens_projection = S.sens_projection;
  st->global_iteration = S.global_iteration;
  return MHG_OK;
}

int mho_begin_iteration(mho_ctx* ctx) {
  int r = checkReady(ctx);
  if (r) return r;
  beginIteration(*ctx);
  return MHG_OK;
}
int mho_save_values_at_n(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; saveValuesAtN(*ctx); return MHG_OK; }
int mho_update_velocity(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; updateVelocity(*ctx); return MHG_OK; }
int mho_apply_boundary_condition(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; applyBoundaryCondition(*ctx); return MHG_OK; }
int mho_update_position(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; updatePosition(*ctx); return MHG_OK; }
int mho_compute_geometric_values(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; computeGeometricValues(*ctx); return MHG_OK; }
int mho_update_density(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; updateDensity(*ctx); return MHG_OK; }
int mho_compute_artificial_viscosity(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; computeArtificialViscosity(*ctx); return MHG_OK; }
int mho_update_energy_and_pressure(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; updateEnergyAndPressure(*ctx); return MHG_OK; }
int mho_remap(mho_ctx* ctx) { int r = checkReady(ctx); if (r) return r; remap(*ctx); return MHG_OK; }
int mho_compute_delta_t(mho_ctx* ctx, double* new_dt, int32_t* stop) {
  int r = checkReady(ctx);
  if (r) return r;
  double dt = 0.;
  int st = 0;
  r = computeDeltaT(*ctx, &dt, &st);
  if (new_dt) *new_dt = dt;
  if (stop) *stop = st;
  if (r == MHG_E_DT_MIN) ctx->err = "dt below deltat-min";
  return r;
}
// The remap split as the reference's services call it (PrepareRemap.cc:448-563).
static int checkRemapArgs(mho_ctx* ctx, int32_t dim, int32_t wd, int32_t nvar, int32_t ne) {
  if (dim != ctx->dim || wd != 1 || ne != ctx->nenv || nvar != ctx->nvar)
    return fail(ctx, MHG_E_ARG, "remap: counts differ from the context's");
  if (!ctx->opt.with_projection) return fail(ctx, MHG_E_STATE, "remap: no projection");
  return MHG_OK;
}
int mho_remap_prepare(mho_ctx* ctx) {
  int r = checkReady(ctx);
  if (r) return r;
  if (!ctx->opt.with_projection) return fail(ctx, MHG_E_STATE, "remap_prepare: no projection");
  computeVariablesForRemap(*ctx);
  computeFaceQuantitesForRemap(*ctx);
  return MHG_OK;
}
int mho_remap_appli(mho_ctx* ctx, int32_t dim, int32_t wd, int32_t nvar, int32_t ne) {
  int r = checkReady(ctx);
  if (r) return r;
  if ((r = checkRemapArgs(ctx, dim, wd, nvar, ne))) return r;
  appliRemap(*ctx);
  return MHG_OK;
}
int mho_remap_variables(mho_ctx* ctx, int32_t dim, int32_t wd, int32_t nvar, int32_t ne) {
  int r = checkReady(ctx);
  if (r) return r;
  if ((r = checkRemapArgs(ctx, dim, wd, nvar, ne))) return r;
  remapVariables(*ctx);
  return MHG_OK;
}
int mho_remap_resize_variables(mho_ctx* ctx, int32_t nvar, int32_t ne) {
  if (!ctx) return MHG_E_ARG;
  return (ne == ctx->nenv && nvar == ctx->nvar) ? MHG_OK : fail(ctx, MHG_E_ARG, "counts differ");
}
int mho_remap_synchronize_uremap(mho_ctx* ctx) { return checkReady(ctx); }
int mho_eos_init(mho_ctx* ctx, int32_t env) {
  int r = checkReady(ctx);
  if (r) return r;
  if (env < 0 || env >= ctx->nenv) return fail(ctx, MHG_E_ARG, "environment index");
  initEOS(*ctx, env);
  return MHG_OK;
}
int mho_eos_apply(mho_ctx* ctx, int32_t env) {
  int r = checkReady(ctx);
  if (r) return r;
  if (env < 0 || env >= ctx->nenv) return fail(ctx, MHG_E_ARG, "environment index");
  applyEOS(*ctx, env);
  return MHG_OK;
}
// applyOneCellEOS (PerfectGasEOSService.cc / StiffenedGasEOSService.cc:49-60)
int mho_eos_apply_one_cell(mho_ctx* ctx, int32_t env, int64_t c) {
  int r = checkReady(ctx);
  if (r) return r;
  mho_ctx& S = *ctx;
  if (env < 0 || env >= S.nenv || c < 0 || c >= S.NC) return fail(ctx, MHG_E_ARG, "environment or cell index");
  if (S.isPure(c, env))
    eosCell(S, env, S.density.g[c], S.internal_energy.g[c], &S.pressure.g[c], &S.sound_speed.g[c], &S.dpde.g[c]);
  else if (S.isImpure(c, env))
    eosCell(S, env, S.density.env(env, c), S.internal_energy.env(env, c), &S.pressure.env(env, c),
            &S.sound_speed.env(env, c), &S.dpde.env(env, c));
  return MHG_OK;
}
int mho_eos_get_adiabatic_cst(const mho_ctx* ctx, int32_t env, double* v) {
  if (!ctx || !v || env < 0 || env >= ctx->nenv) return MHG_E_ARG;
  *v = ctx->gamma_env[env];
  return MHG_OK;
}
int mho_eos_get_tension_limit_cst(const mho_ctx* ctx, int32_t env, double* v) {
  if (!ctx || !v || env < 0 || env >= ctx->nenv) return MHG_E_ARG;
  *v = ctx->ptens_env[env];
  return MHG_OK;
}
int mho_compute_pression_moyenne(mho_ctx* ctx) {
  int r = checkReady(ctx);
  if (r) return r;
  computePressionMoyenne(*ctx);
  return MHG_OK;
}

int mho_remap_get_ordre_projection(const mho_ctx* ctx) { return ctx ? ctx->opt.ordre_projection : MHG_E_ARG; }
int mho_remap_has_projection_pente_borne(const mho_ctx* ctx) { return ctx ? ctx->opt.projection_pente_borne : MHG_E_ARG; }
int mho_remap_has_conservation_energie_totale(const mho_ctx* ctx) { return ctx ? ctx->opt.conservation_energie_totale : MHG_E_ARG; }
int mho_remap_is_euler(const mho_ctx* ctx) { return ctx ? ctx->opt.is_euler_scheme : MHG_E_ARG; }

int mho_step(mho_ctx* ctx, int32_t* stop) {
  int r = checkReady(ctx);
  if (r) return r;
  mho_ctx& S = *ctx;
  beginIteration(S);
  saveValuesAtN(S);
  updateVelocity(S);
  applyBoundaryCondition(S);
  updatePosition(S);
  computeGeometricValues(S);
  updateDensity(S);
  computeArtificialViscosity(S);
  updateEnergyAndPressure(S);
  remap(S);
  double dt;
  int st = 0;
  r = computeDeltaT(S, &dt, &st);
  if (stop) *stop = st;
  if (r) {
    S.err = "dt below deltat-min";
    return r;
  }
  return st ? MHG_STOP : MHG_OK;
}

int mho_run(mho_ctx* ctx, int32_t max_steps, int32_t* steps_done) {
  int done = 0, r = MHG_OK;
  for (; done < max_steps;) {
    int32_t st = 0;
    r = mho_step(ctx, &st);
    if (r < 0) break;
    ++done;
    if (st) break;
  }
  if (steps_done) *steps_done = done;
  return r;
}

// Neumaier-compensated sums of mass, internal energy and kinetic energy.
int mho_conservation(mho_ctx* ctx, double out[3]) {
  if (!ctx || !out) return MHG_E_ARG;
  mho_ctx& S = *ctx;
  auto acc = [](double& s, double& comp, double v) {
    double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) comp += (s - t) + v;
    else comp += (v - t) + s;
    s = t;
  };
  double m = 0, mc = 0, e = 0, ec = 0, k = 0, kc = 0;
  for (long c = 0; c < S.NC; ++c) {
    acc(m, mc, S.cell_mass.g[c]);
    acc(e, ec, S.cell_mass.g[c] * S.internal_energy.g[c]);
  }
  for (long nd = 0; nd < S.NN; ++nd) {
    const R3& u = S.velocity[nd];
    acc(k, kc, 0.5 * S.node_mass[nd] * dot(u, u));
  }
  out[0] = m + mc;
  out[1] = e + ec;
  out[2] = k + kc;
  return MHG_OK;
}

// ---- probes -----------------------------------------------------------
double mho_probe_cqs(int32_t dim, const double* coord, double* out_cqs) {
  int nb = (dim == 2) ? 4 : 8;
  R3 c[8], q[8];
  for (int i = 0; i < nb; ++i) c[i] = R3(coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]);
  if (dim == 3) {
    R3 fc[6];
    hexFaceCoords(c, fc);
    computeCQs(c, fc, q);
  } else {
    computeCQs2D(c, q);
  }
  double volume = 0;
  for (int i = 0; i < nb; ++i) {
    volume += dot(c[i], q[i]);
    out_cqs[3 * i] = q[i].x;
    out_cqs[3 * i + 1] = q[i].y;
    out_cqs[3 * i + 2] = q[i].z;
  }
  return volume * (1. / dim);
}
double mho_probe_limiter(int32_t id, double r) { return limiter(id, r); }
double mho_probe_pb_inty(double X, double x0, double y0, double x1, double y1, double thr) {
  return pb_inty(X, x0, y0, x1, y1, thr);
}
double mho_probe_pb_flux(int32_t type, int32_t lim, double ppv, double h0, double hp, double hm, double y0,
                         double yp, double ym, double grady, double thr) {
  double y0plus, y0moins, xg, yg, xd, yd;
  pb_y0(lim, yp, ym, h0, hp, hm, &y0plus, &y0moins);
  pb_support(y0, yp, ym, h0, y0plus, y0moins, grady, thr, &xg, &yg, &xd, &yd);
  return pb_var_flux(type, ppv, h0, y0, yp, ym, xg, yg, xd, yd, thr);
}

int mho_probe_cell_widths(mho_ctx* c, int32_t idir, double* out) {
  if (!c || !out || idir < 0 || idir >= c->dim) return MHG_E_ARG;
  computeFaceQuantitesForRemap(*c);
  computeCellWidths(*c, idir);
  std::copy(c->h_cell.begin(), c->h_cell.end(), out);
  return MHG_OK;
}

double mho_probe_flux3(const double* y6, const double* h6, double vdt) {
  return o3_flux(y6[0], y6[1], y6[2], y6[3], y6[4], y6[5], h6[0], h6[1], h6[2], h6[3], h6[4], h6[5], vdt);
}

double mho_probe_limiter_g(int32_t id, double gp, double gm, double yp, double ym, double h0, double hp,
                           double hm) {
  return flux_limiter_g(id, gp, gm, yp, ym, h0, hp, hm);
}

double mho_probe_learned_grad(double alpha, double phim, double phi, double phip, int32_t back_in,
                              int32_t front_in, double gb, double gf, double db, double df) {
  return learned_grad(alpha, phim, phi, phip, back_in != 0, front_in != 0, gb, gf, db, df);
}

void mho_probe_consistent_slopes(int32_t ne, const double* phi, double* g, const double* gcl, const int32_t* used) {
  bool u[MHG_MAX_ENV] = {};
  for (int k = 0; k < ne && k < MHG_MAX_ENV; ++k) u[k] = used[k] != 0;
  consistent_env_slopes(ne, phi, g, gcl, u, 1e-10);
}

int mho_set_learned_alpha(mho_ctx* c, const double* alpha) {
  if (!c || !alpha) return MHG_E_ARG;
  c->learned_alpha.assign(alpha, alpha + (size_t)c->dim * c->NC);
  return MHG_OK;
}
double mho_probe_eint(int32_t pc, double gamma, double q_n, double q, double rho_n, double rho,
                      double p, double e_n) {
  return compute_eint(pc != 0, gamma, q_n, q, rho_n, rho, p, e_n);
}
double mho_probe_viscosity(double gamma, double tau, double L, double c, double divu) {
  return viscosity((gamma + 1) / 2.0, tau, L, c, divu);
}
void mho_probe_eos_pg(double gamma, double rho, double e, double* out3) {
  mho_probe_eos(MHG_EOS_PERFECT_GAS, gamma, 0., rho, e, out3);
}
void mho_probe_eos(int32_t eos, double gamma, double limit_tension, double rho, double e, double* out3) {
  mho_ctx S;
  S.gamma_env.push_back(gamma);
  S.ptens_env.push_back(limit_tension);
  S.eos_env.push_back(eos);
  eosCell(S, 0, rho, e, &out3[0], &out3[1], &out3[2]);
}
void mho_probe_eos_init(int32_t eos, double gamma, double limit_tension, double rho, double p, double* out2) {
  mho_ctx S;
  S.gamma_env.push_back(gamma);
  S.ptens_env.push_back(limit_tension);
  S.eos_env.push_back(eos);
  eosInitCell(S, 0, rho, p, &out2[0], &out2[1]);
}

}  // extern "C"
