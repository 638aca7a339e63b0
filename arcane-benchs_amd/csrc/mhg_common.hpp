// mhg_common.hpp -- device-side building blocks shared by the MaHyCo kernels.
//
// Every formula keeps the operation order of the reference expression it
// cites; the library is compiled with -ffp-contract=off so a*b+c is a rounded
// product followed by a rounded sum, as in the reference source.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/mahyco_gpu.h"

namespace mhg {

constexpr int kMaxEnv = MHG_MAX_ENV;
constexpr int kMaxBc = MHG_MAX_BC;

// std::max / std::min semantics ((a<b)?b:a and (b<a)?b:a).
template <class T>
__host__ __device__ inline T smax(T a, T b) { return (a < b) ? b : a; }
template <class T>
__host__ __device__ inline T smin(T a, T b) { return (b < a) ? b : a; }

// Quotient in the arithmetic type: IEEE division in double (the drop-in keeps
// the reference's rounding bit for bit), the hardware-reciprocal quotient
// (within 2 ulp) in the float precision form, whose contract is an error
// bound, not bit equality.
__host__ __device__ inline double qdiv(double a, double b) { return a / b; }
__device__ inline float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Characteristic length of a quadrilateral, math::pow(volume, 1/2)
// (MahycoModule.cc:1259, racine-cubique-volume in 2D): the exact value is the
// square root, so the correctly rounded sqrt (about 20 instructions, where
// the device pow takes about 200).  glibc's pow, which the CPU restatement
// calls, differs from the correctly rounded root by one ulp in about 0.09% of
// inputs.
__host__ __device__ inline double length_sqrt_2d(double V) { return sqrt(V); }
// The same in the storage type R of the bulk fields: the float form stores L
// in float and takes the hardware square root of the rounded volume (one
// instruction, within 1 ulp of the rounded root); L is a length scale of
// the viscosity and the CFL bound, where it cancels nothing.
template <class R>
__device__ inline double length_sqrt_2d_r(double V) {
  if constexpr (std::is_same<R, float>::value) return (double)__builtin_amdgcn_sqrtf((float)V);
  else return sqrt(V);
}

// Face flux factor u_f L_f dt of a cell sweep, and the upwind offset
// s x - dt u of a dual sweep.  FP64 forms them in double and rounds once, as
// the reference's double arithmetic does; the float form multiplies in float
// (float-sized factors, no cancellation: |dt u| stays below the cell size).
template <class R>
__device__ inline R flux_factor(R u, double l, double dt) {
  if constexpr (std::is_same<R, float>::value) return u * (float)l * (float)dt;
  else return (R)(1.0 * u * l * dt);
}
template <class R>
__device__ inline R upwind_offset(int sign, double dx, double dt, R u) {
  if constexpr (std::is_same<R, float>::value) return (float)(sign * dx) - (float)dt * u;
  else return (R)(sign * dx - dt * u);
}
template <class R>
__device__ inline R vel_dt(R u, double dt) {
  if constexpr (std::is_same<R, float>::value) return u * (float)dt;
  else return (R)((double)u * dt);
}

// Limiters, RemapADIService.h:260-286.
template <int LIM, class T = double>
__host__ __device__ inline T limiter(T r) {
  if (LIM == 0) return smax(T(0), smin(T(1), r));                            // MinMod
  if (LIM == 1) return smax(T(0), smax(smin(T(2) * r, T(1)), smin(r, T(2))));  // SuperBee
  if (LIM == 2) return (r <= T(0) ? T(0) : qdiv(T(2) * r, T(1) + r));       // VanLeer
  return T(0);                                                               // DefaultO1
}
__host__ __device__ inline double limiter_rt(int lim, double r) {
  switch (lim) {
    case 0: return limiter<0>(r);
    case 1: return limiter<1>(r);
    case 2: return limiter<2>(r);
    default: return 0.0;
  }
}

// Limited cell gradient, computeGradPhiCell_PBorn0_LimC (RemapADIService.cc:497-506).
__host__ __device__ inline double limited_cell_grad(int lim, double gb, double gf) {
  double g = 0.;
  if (gb != 0.) g += 0.5 * (limiter_rt(lim, gf / gb) * gb);
  if (gf != 0.) g += 0.5 * (limiter_rt(lim, gb / gf) * gf);
  return g;
}
// Limited nodal gradient, computeDualGradPhi_LimC (DualPhaseRemap.cc:771-777):
// note the different association (0.5 * L) * g.
__host__ __device__ inline double limited_node_grad(int lim, double gb, double gf) {
  double g = 0.;
  if (gb != 0.) g += 0.5 * limiter_rt(lim, gf / gb) * gb;
  if (gf != 0.) g += 0.5 * limiter_rt(lim, gb / gf) * gf;
  return g;
}

// The minmod limited gradient 0.5 lim(gf/gb) gb + 0.5 lim(gb/gf) gf (each
// term only where its denominator is non-zero) with one quotient.  With gb
// and gf of the same sign the term of the larger one has lim = 1 exactly and
// the other term's quotient is small / large, so the quotient is taken once,
// of the operands selected by |gf| >= |gb|; with opposite signs or a zero
// operand both terms vanish and the sum stays +0.  Bit for bit the guarded
// two-term sum for finite operands; a NaN operand gives NaN.
// In float (the precision form, an error bound rather than bit equality)
// the sum is its exact value sign(gb) min(|gb|, |gf|), without a quotient.
template <bool CELL, class T>
__host__ __device__ inline T minmod_grad(T gb, T gf) {
  const bool same = (gf > T(0) && gb > T(0)) || (gf < T(0) && gb < T(0));
  const bool unord = !(gb == gb) || !(gf == gf);
  if (std::is_same<T, float>::value) {
    const T m = smin(fabs(gb), fabs(gf));
    const T g = same ? (gb > T(0) ? m : -m) : T(0);
    return unord ? gb + gf : g;
  }
  // the quotient only where some lane has a non-zero slope: a wavefront over
  // uniform data (zero gradients, absent environments) skips it
  T g = T(0);
  if (same) {
    const bool big_f = fabs(gf) >= fabs(gb);
    const T l = smax(T(0), smin(T(1), qdiv(big_f ? gb : gf, big_f ? gf : gb)));
    const T lb = big_f ? T(1) : l, lf = big_f ? l : T(1);
    const T t1 = CELL ? T(0.5) * (lb * gb) : T(0.5) * lb * gb;
    const T t2 = CELL ? T(0.5) * (lf * gf) : T(0.5) * lf * gf;
    g = (T(0) + t1) + t2;
  }
  return unord ? gb + gf : g;
}

// Limited gradients with the limiter fixed at compile time.  BF (branch-free)
// evaluates the reference formula for both terms and selects: each term is
// taken only where its denominator is non-zero, exactly as the guarded sums
// of the reference (a skipped term adds +0 to a sum that starts at +0).
// Minmod always takes minmod_grad; the other limiters evaluate the quotient
// as the reference.
template <int LIM, bool CELL, class T, bool BF = false>
__host__ __device__ inline T limited_grad(T gb, T gf) {
  if (LIM == 0) return minmod_grad<CELL>(gb, gf);
  if (BF) {
    const T l1 = limiter<LIM>(qdiv(gf, gb)), l2 = limiter<LIM>(qdiv(gb, gf));
    const T t1 = CELL ? T(0.5) * (l1 * gb) : T(0.5) * l1 * gb;
    const T t2 = CELL ? T(0.5) * (l2 * gf) : T(0.5) * l2 * gf;
    T g = T(0);
    g = g + ((gb != T(0)) ? t1 : T(0));
    g = g + ((gf != T(0)) ? t2 : T(0));
    return g;
  }
  T g = T(0);
  if (gb != T(0)) g += CELL ? T(0.5) * (limiter<LIM>(qdiv(gf, gb)) * gb) : T(0.5) * limiter<LIM>(qdiv(gf, gb)) * gb;
  if (gf != T(0)) g += CELL ? T(0.5) * (limiter<LIM>(qdiv(gb, gf)) * gf) : T(0.5) * limiter<LIM>(qdiv(gb, gf)) * gf;
  return g;
}

// Template value of the sweeps' LIM for the generalised limiters; the
// limiter id (MHG_LIM_MINMODG .. MHG_LIM_ULTRABEEG) is then Params::lim.
constexpr int kLimG = 3;

// Generalised limiters, fluxLimiterG (UtilesRemap.cc:127-173): slope of a
// cell (or dual cell) of width h0 between a front neighbour of width hp and a
// back neighbour of width hm, from the front / back gradients gp, gm and the
// front / back values yp, ym.  Zero at a slope break; otherwise the
// limiter's candidate, bounded by min(|gp (h0+hp)/h0|, |gm (h0+hm)/h0|), with
// the sign of yp - ym.
template <class T>
__host__ __device__ inline T limiter_g(int id, T gp, T gm, T yp, T ym, T h0, T hp, T hm) {
  if (gp * gm < T(0)) return T(0);
  const bool up = (yp - ym) > T(0);
  const T agp = fabs(gp), agm = fabs(gm);
  T grady = T(0);
  if (id == MHG_LIM_MINMODG) {  // formula 9c
    grady = up ? smin(agp, agm) : -smin(agp, agm);
  } else if (id == MHG_LIM_SUPERBEEG) {  // formula 9g
    grady = up ? smax(agp, agm) : -smax(agp, agm);
  } else if (id == MHG_LIM_VANLEERG || id == MHG_LIM_ARITHMETICG) {  // formulas 9e and the mean
    const T lp = qdiv(h0 / T(2) + hp, h0 + hp + hm);
    const T lm = qdiv(h0 / T(2) + hm, h0 + hp + hm);
    if (id == MHG_LIM_VANLEERG) {
      const T den = lp * gp + lm * gm;
      grady = den != T(0) ? qdiv(gp * gm, den) : T(0);
    } else {
      grady = lm * gp + lp * gm;
    }
  } else if (id == MHG_LIM_ULTRABEEG) {
    grady = qdiv(yp - ym, h0);
  }
  // simple-slope limit (formula 10)
  const T gM = smin(fabs(qdiv(gp * (h0 + hp), h0)), fabs(qdiv(gm * (h0 + hm), h0)));
  const T g = smin(fabs(gM), fabs(grady));
  return up ? g : -g;
}

// Template value of the cell sweeps' LIM for the order-3 face values
// (ordre-projection 3); the dual sweeps keep Params::lim.
constexpr int kLimO3 = 4;

// Order-3 face value, ComputeFluxOrdre3 (Remap-ordre-3.cc:69-131): the mean
// over the swept region of the parabolic reconstruction of the upwind cell,
// bounded by the TVD interval.  y / h: values and widths of the cells bbb,
// bb, b | f, ff, fff around the face; vdt: face velocity times dt.  min and
// max follow Arcane's math::min / math::max.
template <class T>
__host__ __device__ inline T amin3(T a, T b) { return (a < b) ? a : b; }
template <class T>
__host__ __device__ inline T amax3(T a, T b) { return (a < b) ? b : a; }
template <class T>
__host__ __device__ inline T o3_grad(T hm, T h0, T hp, T ym, T y0, T yp) {  // evaluate_grad (7-14)
  return qdiv(h0, hm + h0 + hp) *
         (qdiv(T(2) * hm + h0, h0 + hp) * (yp - y0) + qdiv(h0 + T(2) * hp, hm + h0) * (y0 - ym));
}
template <class T>
__host__ __device__ inline T o3_ystar(T hmm, T hm, T hp, T hpp, T ym, T yp, T gradm, T gradp) {  // (17-30)
  const T tmp1 = qdiv(T(2) * hp * hm, hm + hp) *
                 (qdiv(hmm + hm, T(2) * hm + hp) - qdiv(hpp + hp, T(2) * hp + hm)) * (yp - ym);
  const T tmp2 = qdiv(-hm * (hmm + hm), T(2) * hm + hp) * gradp + qdiv(hp * (hp + hpp), hm + T(2) * hp) * gradm;
  return ym + qdiv(hm, hm + hp) * (yp - ym) + qdiv(T(1), hmm + hm + hp + hpp) * (tmp1 + tmp2);
}
template <class T>
__host__ __device__ inline T o3_flux(const T* y, const T* h, T vdt) {
  const T gradmm = o3_grad(h[0], h[1], h[2], y[0], y[1], y[2]);
  const T gradm = o3_grad(h[1], h[2], h[3], y[1], y[2], y[3]);
  const T gradp = o3_grad(h[2], h[3], h[4], y[2], y[3], y[4]);
  const T gradpp = o3_grad(h[3], h[4], h[5], y[3], y[4], y[5]);
  const T ystarm = o3_ystar(h[0], h[1], h[2], h[3], y[1], y[2], gradmm, gradm);
  const T ystar = o3_ystar(h[1], h[2], h[3], h[4], y[2], y[3], gradm, gradp);
  const T ystarp = o3_ystar(h[2], h[3], h[4], h[5], y[3], y[4], gradp, gradpp);
  const T ym = y[2], yp = y[3], hm = h[2], hp = h[3];
  const T grad_m = ystar - ystarm, grad_p = ystarp - ystar;
  const T ym6 = T(6) * (ym - T(0.5) * (ystarm + ystar));
  const T yp6 = T(6) * (yp - T(0.5) * (ystar + ystarp));
  T flux;
  if (vdt > T(0)) {  // evaluate_fm (33-38)
    flux = ystar - qdiv(T(0.5) * vdt, hm) * (grad_m - (T(1) - qdiv(T(2.0 / 3.0) * vdt, hm)) * ym6);
  } else if (vdt < T(0)) {  // evaluate_fp (41-46), with the reference's sign of the curvature term
    const T x = -vdt;
    flux = ystar + qdiv(T(0.5) * x, hp) * (grad_p - (T(1) - qdiv(T(2.0 / 3.0) * x, hp)) * yp6);
  } else {
    return T(0);
  }
  const T num = qdiv(vdt, hm), nup = qdiv(vdt, hp);
  const T ym_ym = ym + qdiv(T(1) - num, num) * (ym - y[1]);
  const T yp_ym = yp - qdiv(T(1) + nup, nup) * (yp - y[4]);
  const T i1a = amin3(ym, yp), i1b = amax3(ym, yp);
  const T i2a = vdt >= T(0) ? amin3(ym, ym_ym) : amin3(yp, yp_ym);
  const T i2b = vdt >= T(0) ? amax3(ym, ym_ym) : amax3(yp, yp_ym);
  T la = T(0), lb = T(0);
  if (!((i1b < i2a) || (i2b < i1a))) {
    la = amax3(i1a, i2a);
    lb = amin3(i1b, i2b);
  }
  if (flux < la) flux = la;
  if (flux > lb) flux = lb;
  return flux;
}

// compute_eint, MahycoModule.cc:1630-1654.
template <class T>
__host__ __device__ inline T compute_eint(bool pseudo_centree, T adiabatic_cst, T q_n, T q, T rho_n,
                                          T rho, T p, T e_n) {
  const T dv = qdiv(T(1), rho) - qdiv(T(1), rho_n);  // 1/rho - 1/rho_n
  T pseudo = T(0);
  if (pseudo_centree && ((q_n + q) * dv < T(0))) pseudo = T(0.5) * (q + q_n);
  if (!pseudo_centree && (q * dv < T(0))) pseudo = q;
  T denom = T(1) + T(0.5) * (adiabatic_cst - T(1)) * rho * dv;
  T numer = e_n - (T(0.5) * p + pseudo) * dv;
  return qdiv(numer, denom);
}

// Pseudo-viscosity, MahycoModule.cc:523-526.
template <class T>
__host__ __device__ inline T viscosity(T qcoef, T tau, T L, T c, T divu) {
  return qdiv(T(1), tau) * (T(-0.5) * L * c * divu + qcoef * L * L * divu * divu);
}

// Equation of state of one environment.
struct EosParams {
  int kind[kMaxEnv];
  double gamma[kMaxEnv];
  double tension[kMaxEnv];
  double qcoef[kMaxEnv];
};

// PerfectGas compute_pressure_sndspd_PG (PerfectGasEOSService.cc:38-45) and
// StiffenedGas applyEOS (StiffenedGasEOSService.cc:35-42).
template <class T>
__host__ __device__ inline void eos_apply(const EosParams& E, int k, T rho, T e, T* p, T* c) {
  const T g = (T)E.gamma[k];
  if (E.kind[k] == MHG_EOS_STIFFENED_GAS) {
    const T t = (T)E.tension[k];
    T pressure = ((g - T(1)) * rho * e) - (g * t);
    *p = pressure;
    *c = sqrt(qdiv(g, rho) * (pressure + t));
  } else {
    T pressure = (g - T(1)) * rho * e;
    *p = pressure;
    *c = sqrt(qdiv(g * pressure, rho));
  }
}

// 2D corner normal of corner q of a quad whose nodes are (x[i], y[i]) in
// Arcane quad4 order (MahycoModule.cc:1102-1111):
//   npc[e+1] = 0.5*(y[e+1]-y[e], x[e]-x[e+1]), npc[0] = npc[4],
//   cqs[q] = npc[q+1] + npc[q].
__host__ __device__ inline void cqs2d(const double* x, const double* y, int q, double* cx,
                                      double* cy) {
  const int a = q, b = (q + 1) & 3, m = (q + 3) & 3;
  // npc[q+1]: edge a -> b ; npc[q]: edge m -> a
  double e1x = 0.5 * (y[b] - y[a]);
  double e1y = 0.5 * (x[a] - x[b]);
  double e0x = 0.5 * (y[a] - y[m]);
  double e0y = 0.5 * (x[m] - x[a]);
  *cx = e1x + e0x;
  *cy = e1y + e0y;
}

// Volume (1/2) sum_i x_i . C_i of a quad (MahycoModule.cc:1244-1256).
__host__ __device__ inline double quad_volume(const double* x, const double* y) {
  double volume = 0;
  for (int q = 0; q < 4; ++q) {
    double cx, cy;
    cqs2d(x, y, q, &cx, &cy);
    volume += x[q] * cx + y[q] * cy;
  }
  volume *= 0.5;
  return volume;
}

// Characteristic length "faces-opposees" (MahycoModule.cc:1199-1219); in 2D
// the reference reads coord[4..7] which are zero.
__host__ __device__ inline double length_faces_opposees_2d(const double* x, const double* y) {
  double cx[8] = {x[0], x[1], x[2], x[3], 0., 0., 0., 0.};
  double cy[8] = {y[0], y[1], y[2], y[3], 0., 0., 0., 0.};
  auto fc = [&](const double* c, int a, int b, int cc, int d) {
    return 0.25 * (c[a] + c[b] + c[cc] + c[d]);
  };
  double f0x = fc(cx, 0, 3, 2, 1), f0y = fc(cy, 0, 3, 2, 1);
  double f1x = fc(cx, 0, 4, 7, 3), f1y = fc(cy, 0, 4, 7, 3);
  double f2x = fc(cx, 0, 1, 5, 4), f2y = fc(cy, 0, 1, 5, 4);
  double f3x = fc(cx, 4, 5, 6, 7), f3y = fc(cy, 4, 5, 6, 7);
  double f4x = fc(cx, 1, 2, 6, 5), f4y = fc(cy, 1, 2, 6, 5);
  double f5x = fc(cx, 2, 3, 7, 6), f5y = fc(cy, 2, 3, 7, 6);
  double m1x = f0x - f3x, m1y = f0y - f3y;
  double m2x = f2x - f5x, m2y = f2y - f5y;
  double m3x = f1x - f4x, m3y = f1y - f4y;
  double d1 = sqrt(m1x * m1x + m1y * m1y + 0. * 0.);
  double d2 = sqrt(m2x * m2x + m2y * m2y + 0. * 0.);
  double d3 = sqrt(m3x * m3x + m3y * m3y + 0. * 0.);
  double num = d1 * d2 * d3;
  double den = d1 * d2 + d1 * d3 + d2 * d3;
  return num / den;
}

// ---- 3D geometry (hexa8, Arcane node order) --------------------------------
struct V3 {
  double x, y, z;
};
__host__ __device__ inline V3 v3add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 v3sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 v3mul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline V3 v3cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Face centres of a hexahedron, 0.25 * (sum of its 4 nodes in face order)
// (MahycoModule.cc:1080-1087): faces (0,3,2,1), (0,4,7,3), (0,1,5,4),
// (4,5,6,7), (1,2,6,5), (2,3,7,6).
constexpr int kHexFace[6][4] = {{0, 3, 2, 1}, {0, 4, 7, 3}, {0, 1, 5, 4}, {4, 5, 6, 7}, {1, 2, 6, 5}, {2, 3, 7, 6}};
__host__ __device__ inline V3 hex_face_centre(const V3* c, int f) {
  return v3mul(0.25, v3add(v3add(v3add(c[kHexFace[f][0]], c[kHexFace[f][1]]), c[kHexFace[f][2]]), c[kHexFace[f][3]]));
}

// The 24 half cross products of computeCQs (MahycoModule.cc:2195-2258),
// term t = 0.5 * cross(c[a] - fc[f], c[b] - fc[f]) with (f, a, b) below, and
// the 12 terms of each corner in the reference's summation order: corner q
// is (1/12) * (5 * (sum of the first six) + (sum of the last six)).
constexpr int kHexTerm[24][3] = {{0, 0, 3}, {0, 3, 2}, {0, 2, 1}, {0, 1, 0},   // n1a04 n1a03 n1a02 n1a01
                                  {1, 0, 4}, {1, 4, 7}, {1, 7, 3}, {1, 3, 0},   // n2a05 n2a12 n2a08 n2a04
                                  {2, 0, 1}, {2, 1, 5}, {2, 5, 4}, {2, 4, 0},   // n3a01 n3a06 n3a09 n3a05
                                  {3, 4, 5}, {3, 5, 6}, {3, 6, 7}, {3, 7, 4},   // n4a09 n4a10 n4a11 n4a12
                                  {4, 1, 2}, {4, 2, 6}, {4, 6, 5}, {4, 5, 1},   // n5a02 n5a07 n5a10 n5a06
                                  {5, 2, 3}, {5, 3, 7}, {5, 7, 6}, {5, 6, 2}};  // n6a03 n6a08 n6a11 n6a07
constexpr int kHexCorner[8][12] = {{3, 0, 7, 4, 11, 8, 2, 1, 6, 5, 9, 10},
                                   {3, 2, 8, 9, 19, 16, 0, 1, 10, 11, 18, 17},
                                   {2, 1, 17, 16, 23, 20, 3, 0, 19, 18, 22, 21},
                                   {1, 0, 6, 7, 21, 20, 3, 2, 4, 5, 23, 22},
                                   {4, 5, 11, 10, 12, 15, 6, 7, 8, 9, 13, 14},
                                   {9, 10, 12, 13, 18, 19, 8, 11, 15, 14, 17, 16},
                                   {14, 13, 18, 17, 23, 22, 15, 12, 19, 16, 20, 21},
                                   {6, 5, 15, 14, 22, 21, 7, 4, 12, 13, 23, 20}};
__host__ __device__ inline V3 hex_term_value(const V3* c, const V3* fc, int t) {
  const int f = kHexTerm[t][0], a = kHexTerm[t][1], b = kHexTerm[t][2];
  return v3mul(0.5, v3cross(v3sub(c[a], fc[f]), v3sub(c[b], fc[f])));
}
// Corner q from the 24 terms n[] (or from the coordinates when n is null).
__host__ __device__ inline V3 hex_corner_sum(const V3* n, int q) {
  V3 a = n[kHexCorner[q][0]];
#pragma unroll
  for (int s = 1; s < 6; ++s) a = v3add(a, n[kHexCorner[q][s]]);
  V3 b = n[kHexCorner[q][6]];
#pragma unroll
  for (int s = 7; s < 12; ++s) b = v3add(b, n[kHexCorner[q][s]]);
  const double w = 1. / 12.;
  return v3mul(w, v3add(v3mul(5., a), b));
}
// Corner q of the cell whose nodes are c[0..7] (fc: its 6 face centres),
// forming only the 12 terms it needs.
template <int Q>
__host__ __device__ inline V3 hex_corner_normal(const V3* c, const V3* fc) {
  V3 a = hex_term_value(c, fc, kHexCorner[Q][0]);
#pragma unroll
  for (int s = 1; s < 6; ++s) a = v3add(a, hex_term_value(c, fc, kHexCorner[Q][s]));
  V3 b = hex_term_value(c, fc, kHexCorner[Q][6]);
#pragma unroll
  for (int s = 7; s < 12; ++s) b = v3add(b, hex_term_value(c, fc, kHexCorner[Q][s]));
  const double w = 1. / 12.;
  return v3mul(w, v3add(v3mul(5., a), b));
}
// Volume (1/3) sum_i x_i . C_i (MahycoModule.cc:1244-1256) of a hexahedron:
// the 24 terms once, then the 8 corners (computeCQs, MahycoModule.cc:2195-2258).
__host__ __device__ inline double hex_volume(const V3* c) {
  V3 fc[6], n[24];
#pragma unroll
  for (int f = 0; f < 6; ++f) fc[f] = hex_face_centre(c, f);
#pragma unroll
  for (int t = 0; t < 24; ++t) n[t] = hex_term_value(c, fc, t);
  double volume = 0.;
#pragma unroll
  for (int q = 0; q < 8; ++q) volume += v3dot(c[q], hex_corner_sum(n, q));
  const double inv_dim = 1. / 3.;
  return volume * inv_dim;
}
// Characteristic length "faces-opposees" (MahycoModule.cc:1199-1219).
__host__ __device__ inline double hex_length_faces_opposees(const V3* c) {
  V3 fc[6];
  for (int f = 0; f < 6; ++f) fc[f] = hex_face_centre(c, f);
  const V3 m1 = v3sub(fc[0], fc[3]), m2 = v3sub(fc[2], fc[5]), m3 = v3sub(fc[1], fc[4]);
  const double d1 = sqrt(m1.x * m1.x + m1.y * m1.y + m1.z * m1.z);
  const double d2 = sqrt(m2.x * m2.x + m2.y * m2.y + m2.z * m2.z);
  const double d3 = sqrt(m3.x * m3.x + m3.y * m3.y + m3.z * m3.z);
  return (d1 * d2 * d3) / (d1 * d2 + d1 * d3 + d2 * d3);
}

// Boundary condition on an index range of one axis (the node set that
// PrepareFaceGroup's coordinate test selects on a Cartesian mesh).
struct BcRange {
  int axis;   // -1: every node
  int lo, hi; // inclusive node-index range along axis (lo > hi: empty)
  int comp;   // velocity component
  double value;
};


// Neighbour-lane reads within a 64-lane wavefront through the DPP whole-wave
// shifts of gfx9-family VALU instructions (wave_shr:1 = 0x138, wave_shl:1 =
// 0x130) instead of an LDS permute.  lane_up1: lane i receives lane i - 1;
// lane_down1: lane i receives lane i + 1; a lane whose source lies outside
// the wavefront (lane 0 up, lane 63 down) receives 0 (bound_ctrl).  Those
// lanes are halo lanes of every sweep, whose results no output lane reads,
// and a shift that keeps no old value is one instruction instead of a copy
// and a shift (or none, where the compiler folds it into its consumer).
// 64-bit values move as two 32-bit halves.
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;

template <int CTRL>
__device__ inline int dpp_shift(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ inline float dpp_shift(float v) {
  return __int_as_float(dpp_shift<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ inline double dpp_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shift<CTRL>((int)(b & 0xffffffffLL)), hi = dpp_shift<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T>
__device__ inline T lane_up1(T v) {
  return dpp_shift<kDppWaveShr1>(v);
}
template <class T>
__device__ inline T lane_down1(T v) {
  return dpp_shift<kDppWaveShl1>(v);
}

// XCD-aware block order.  The hardware hands block b to XCD b % 8; this maps
// it to logical block (b % 8) * (n / 8) + b / 8 so that every XCD walks a
// contiguous band of logical blocks and neighbouring tiles share its L2.
__device__ inline int xcd_block(int b, int n) {
  const int full = n & ~7;
  if (b >= full) return b;
  const int per = full >> 3;
  return (b & 7) * per + (b >> 3);
}

// Scalars kept on the device so steps can run back to back.
struct DevScalars {
  double dt;        // m_global_deltat
  double dt_old;    // m_old_deltat
  double time;      // m_global_time
  int iteration;
  int stop;         // end of run reached
  int err;          // MHG_E_DT_MIN raised
  int landing;      // the step under way ends at the final time (land_on_final_time)
};

// The run has ended (stop) or failed (err): the kernels of a replayed step
// then leave memory as it is.  The step's main kernels read the flags with
// their first loads and test them before their first global store, so that
// the flags' latency overlaps the loads' instead of preceding them (1.6 % of
// an FP64 triple-point step, 2.2 % in float).  The flags are uniform over the
// launch, so the early return is uniform over every block.
__device__ __forceinline__ bool halted(const DevScalars* S) { return (S->stop | S->err) != 0; }

// (n / d, n % d) in 32-bit unsigned arithmetic: mesh indices are below 2^31
// (mhg_create refuses larger meshes), and a 64-bit division expands to a
// branch and twice the instructions.
__device__ __forceinline__ void divmod32(long n, int d, int& q, int& r) {
  const unsigned un = (unsigned)n, ud = (unsigned)d, uq = un / ud;
  q = (int)uq;
  r = (int)(un - uq * ud);
}

// Params::ytw8 bits: the y sweeps that run on 8-column tiles (the others on
// 16 columns): the first cell sweep with the Lagrange cell phase, the last
// with the final cell phase, the dual sweeps.
constexpr int kYtwLag = 1, kYtwFinal = 2, kYtwNode = 4;

// Euler volume (area in 2D) of a cell of the Cartesian generator's uniform
// mesh.  FP64 forms it as the reference does, with the corner-normal formula
// on the initial node coordinates; the float form takes the exact cell size
// h_x h_y (h_z), which that formula reproduces to about 1e-13 relative, and
// saves the formula's FP64 instructions (3D: the stored volume's load).
template <class R>
constexpr bool kCartesianEuler = std::is_same<R, float>::value;

// Run-time parameters of a context.
struct Params {
  int dim, ne, nvu;  // nvu = 3*ne+1 live cell variables
  int nx, ny, nz;
  long nc, nn;
  double org[3], h[3];
  EosParams eos;
  double threshold, remap_threshold;
  int order2, lim, pseudo_centree, lc;
  int learned;  // cell sweeps use the learned-limiter slope (lim is then minmod); 2: the
                // network's coefficients also at material interfaces
  int land;     // shorten the last step to end at tend
  int etot;     // total-energy fix-up of the remap (conservation-energie-totale)
  int lagonly;  // with-projection = false: no remap, the mesh keeps moving
                // (updatePosition from the previous coordinates), the CFL
                // has no velocity term
  // generic cell sweeps (order 2 with projection-pente-borne or -mixte; FP64,
  // entry-point kernels): pb the pente-borne fluxes with debar fix pb_fix,
  // mixte the pure-neighbourhood switch to limiter pure_id
  int generic, pb, pb_fix, mixte, pure_id;
  // store policy of the step kernels' outputs, one bit per kind of array
  // (st_stream): set, non-temporal stores (the lines stream to memory);
  // clear, ordinary stores (they stay in the L2 / memory-side cache for the
  // next kernel of the step); chosen per context from the working set,
  // settable by the caller
  int nt_store;
  int ytw8;  // y sweeps on LDS tiles of 8 columns (512-thread blocks) instead of 16: kYtw* bits
  double cfl, dtmin, dtmax, tend;
  int nbc;
  BcRange bc[kMaxBc];
};

// Store of a value that this kernel writes once and a later kernel reads.
// With Params::nt_store (a working set beyond the memory-side cache) it is
// non-temporal, so that it streams out instead of staying dirty in the L2 of
// the writing XCD (each kernel ends with a write-back of that L2); otherwise
// an ordinary store, which the next kernels of the step find in the caches.
// The flag is uniform over the launch (a scalar branch).
// (The ordinary store is written as a relaxed wavefront-scope atomic store,
// which is the same plain global store: as `*p = v` the compiler merges the
// two branches into one store and drops the non-temporal hint.)
// CLS names the kind of array (bit CLS of Params::nt_store selects streaming
// for it): kStCell the cell state and geometry the next step reads, kStU the
// remap vector between the cell sweeps, kStFlux the dual mass fluxes and dual
// state between a cell sweep and the dual sweeps, kStNode node velocities
// and coordinates.
constexpr int kStCell = 0, kStU = 1, kStFlux = 2, kStNode = 3;
template <int CLS, class T>
__device__ inline void st_stream(const Params& P, T* p, T v) {
  if ((P.nt_store >> CLS) & 1) __builtin_nontemporal_store(v, p);
  else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Environment count inside a kernel compiled for at most NEM environments.
// The launchers choose NEM from {1, 2, 3, 4, 8} as the smallest bound >= ne,
// so NEM = ne whenever NEM <= 4: the count is then a compile-time constant
// and the per-environment guards (k < ne) fold away.
template <int NEM>
__host__ __device__ inline int env_count(const Params& P) {
  return NEM <= 4 ? NEM : P.ne;
}

}  // namespace mhg
