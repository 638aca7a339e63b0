// mhg_kernels_impl.hpp -- device functions shared by the 2D and 3D kernels.
#pragma once

#include <type_traits>

#include "mhg_common.hpp"
#include "mhg_kernels.hpp"

namespace mhg {

// LDS tiles of the y (and z) sweeps: 16 columns, 64 rows, padded rows.
constexpr int kTileW = 16, kTileR = 64, kTileP = kTileW + 1;

// Limited gradients: branch-free in float (two hardware-reciprocal quotients
// cost less than the divergent branches) for the limiters other than minmod,
// which takes minmod_grad in both types.
template <class R>
constexpr bool kBranchFree = std::is_same<R, float>::value;

template <class T>
__device__ inline T face_grad(bool inner, T phi, T phim, T dx) {
  return inner ? qdiv(phi - phim, dx) : T(0);
}

// Learned-limiter slope of a cell (MHG_LIM_LEARNED): alpha (g_b + g_f) / 2,
// kept if the face values phi + g db (back) and phi + g df (front) stay
// within [min, max] of phi and the existing neighbours along the sweep,
// otherwise the minmod slope; alpha < 0 selects the minmod slope outright.
template <class R>
__device__ inline R learned_grad(R alpha, R phim, R phi, R phip, bool back_in, bool front_in, R gb, R gf, R db,
                                 R df, bool* used = nullptr, R* classical = nullptr) {
  const R gcl = limited_grad<0, true, R, kBranchFree<R>>(gb, gf);
  if (classical) *classical = gcl;
  if (used) *used = false;
  if (alpha < R(0)) return gcl;
  const R gc = alpha * (R(0.5) * (gb + gf));
  R mn = phi, mx = phi;
  if (back_in) {
    mn = smin(mn, phim);
    mx = smax(mx, phim);
  }
  if (front_in) {
    mn = smin(mn, phip);
    mx = smax(mx, phip);
  }
  const R vb = phi + gc * db, vf = phi + gc * df;
  const bool ok = vb >= mn && vb <= mx && vf >= mn && vf <= mx;
  if (used) *used = ok && gc != R(0);
  return ok ? gc : gcl;
}

// Consistent environment slopes of the learned mode, per lane.  In a mixed
// cell (two or more environments with phi_V > 0) with alpha >= 0, the mass
// and energy slopes of environment k follow its volume slope, g_m = (phi_m /
// phi_V) g_V and g_E = (phi_E / phi_V) g_V, when a non-zero learned volume
// slope was kept and phi_V exceeds the remap threshold; otherwise they take
// the classical slope.  An environment's mass and energy face values are
// then its volume face value times the upwind cell's density and energy
// density, so no face carries mass or energy without volume.  The volume
// slope of a trace environment (0 < phi_V <= remap threshold) is classical,
// so that its three slopes are the classical method's.
template <class R, int NEM>
struct EnvSlopes {
  R gV[NEM], pV[NEM];
  bool uV[NEM];
  bool cons = false;
  R thr;
  __device__ explicit EnvSlopes(double threshold) : thr((R)threshold) {}
  // a trace environment's volume (slope forced classical)
  __device__ __forceinline__ bool trace(int role, R phi) const { return role == 0 && phi > R(0) && phi <= thr; }
  // slope of variable v (role: 0 volume, 1 mass, 2 energy, 3 other; k its
  // environment) from its learned slope gl (kept: used) and classical gcl
  __device__ __forceinline__ R slope(int role, int k, R phi, R gl, bool used, R gcl) {
    R g = gl;
    if (role == 0) {
      gV[k] = gl;
      pV[k] = phi;
      uV[k] = used;
    } else if ((role == 1 || role == 2) && cons) {
      g = (uV[k] && pV[k] > thr) ? qdiv(phi, pV[k]) * gV[k] : gcl;
    }
    return g;
  }
  // after the volumes: a mixed cell with an active learned coefficient
  __device__ __forceinline__ void close_volumes(int ne, R alpha) {
    int present = 0;
#pragma unroll
    for (int k = 0; k < NEM; ++k)
      if (k < ne && pV[k] > R(0)) ++present;
    cons = alpha >= R(0) && present >= 2;
  }
};

// Mass-conservative arithmetic of the float precision form.  The reference
// forms a cell's remapped mass as V sum_k f_k rho_k with rho_k = U_m,k /
// (f_k V) (RemapADIFinal.cc:446-560) and the mass entering the remap as
// V (m / V) (PrepareRemap.cc:300-372): both are the masses themselves, and the
// drop-in path keeps those operations bit for bit.  In float the quotients
// (hardware reciprocal, within 2 ulp) are biased, and the mass of the 1792x768
// triple point drifted by 3e-7 over 500 steps (FP64: 5e-12).  The float form
// therefore takes the stored mass as the remapped one, U_m = m, and the new
// cell mass as sum_k U_m,k, so that the remap only moves mass between cells.
template <class R>
constexpr bool kMassCons = std::is_same<R, float>::value;

// Cell fields the Lagrange cell phase reads, loaded by the caller before the
// node coordinates so that a wavefront issues all of them together.
template <class R>
struct CellPre {
  R rho_n, m, cs, p, e, q;
  int mask;
};
template <class R, bool PC>
__device__ __forceinline__ CellPre<R> cell_pre_load(const CellLagArgsT<R>& A, long c) {
  CellPre<R> r;
  r.rho_n = A.rho[c];
  r.m = A.m[c];
  r.cs = A.cs[c];
  r.p = A.p[c];
  r.e = A.e[c];
  r.q = PC ? A.q[c] : R(0);
  r.mask = A.mask[c];
  return r;
}

// Density, viscosity, energy and the remap vector of cell c from its
// Lagrangian volume V and characteristic length L (both dimensions), into
// uv, um, ue (per environment) and uq.
template <class R, bool PC, int NEM>
__device__ inline void lagrange_cell_vals(const Params& P, const DevScalars* S, const CellLagArgsT<R>& A, long c,
                                          R V, R L, const CellPre<R>& in, R (&uv)[NEM], R (&um)[NEM],
                                          R (&ue)[NEM], R& uq) {
  const R inv_dt = qdiv(R(1.0), (R)S->dt);
  const R rho_n = in.rho_n;
  const R rho = qdiv(in.m, V);
  const R tau = R(0.5) * (qdiv(R(1.0), rho_n) + qdiv(R(1.0), rho));
  const R divu = qdiv(inv_dt * (qdiv(R(1.0), rho) - qdiv(R(1.0), rho_n)), tau);
  const R cs = in.cs;
  const int mask = in.mask;
  const long NC = P.nc;
  uq = R(0);
#pragma unroll
  for (int k = 0; k < NEM; ++k) uv[k] = um[k] = ue[k] = R(0);
  if (__popc(mask) == 1) {
    const int kk = __ffs(mask) - 1;
    R qv = R(0);
    if (divu < R(0)) qv = viscosity((R)P.eos.qcoef[kk], tau, L, cs, divu);
    const R qn = in.q;
    const R e = compute_eint(PC, (R)P.eos.gamma[kk], qn, qv, rho_n, rho, in.p, in.e);
#pragma unroll
    for (int k = 0; k < NEM; ++k)
      if (k == kk) {
        uv[k] = V;
        // float precision form: the stored mass itself, the quantity the
        // remap conserves (kMassCons), where the reference forms V (m / V)
        um[k] = kMassCons<R> ? in.m : V * rho;
        ue[k] = kMassCons<R> ? um[k] * e : V * rho * e;
      }
    uq = V * qv;
  } else if (mask != 0) {
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      if (!(mask & (1 << k))) continue;
      const long ek = k * NC + c;
      const R f = A.env_f[ek];
      const R Vk = f * V;
      const R rho_nk = A.env_rho[ek];
      const R rk = qdiv(A.env_m[ek], Vk);
      const R tk = R(0.5) * (qdiv(R(1.0), rho_nk) + qdiv(R(1.0), rk));
      R qk = R(0);
      if (divu < R(0)) qk = viscosity((R)P.eos.qcoef[k], tk, L, cs, divu);
      const R qnk = PC ? A.env_q[ek] : R(0);
      const R ekv = compute_eint(PC, (R)P.eos.gamma[k], qnk, qk, rho_nk, rk, A.env_p[ek], A.env_e[ek]);
      uv[k] = Vk;
      um[k] = kMassCons<R> ? A.env_m[ek] : Vk * rk;
      ue[k] = kMassCons<R> ? um[k] * ekv : Vk * rk * ekv;
      uq = Vk * qk;  // PrepareRemap.cc:372 assigns: the last environment wins
    }
  }
}

// The same, stored: the remap vector, the Lagrangian volume and L.
template <class R, bool PC, int NEM>
__device__ inline void lagrange_cell_bulk(const Params& P, const DevScalars* S, const CellLagArgsT<R>& A,
                                          long c, double Vd, double Ld, const CellPre<R>& in) {
  // geometry computed in double, the rest in the storage type
  const R V = (R)Vd, L = (R)Ld;
  const int ne = env_count<NEM>(P);
  const long NC = P.nc;
  R uv[NEM], um[NEM], ue[NEM], uq;
  lagrange_cell_vals<R, PC, NEM>(P, S, A, c, V, L, in, uv, um, ue, uq);
  R* U = A.U;
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    if (k >= ne) break;
    st_stream<1>(P, &U[k * NC + c], uv[k]);
    st_stream<1>(P, &U[(ne + k) * NC + c], um[k]);
    st_stream<1>(P, &U[(2 * ne + k) * NC + c], ue[k]);
  }
  st_stream<1>(P, &U[(3 * ne) * NC + c], uq);
  st_stream<0>(P, &A.Vlag[c], V);
  st_stream<0>(P, &A.L[c], L);
}

// Final cell phase of cell c with Euler volume vold: membership, "moy", EOS,
// pressure average; returns the cell's sound speed for the CFL candidate.
// The remap vector comes from the caller (volumes, masses and energies of
// the environments, zero beyond ne, and q V), so that the last cell sweep
// can run this phase on the values it has just formed.  EOS = false stops
// after remapVariables (RemapADIFinal.cc:19-778): the pressures and sound
// speeds are left to the EOS and pressure-average entry points
// (PrepareRemap.cc:552-559), and the return value is 0.  old_mask_in: the
// cell's membership mask when the caller loaded it ahead (-1: load it here).
template <class R, int NEM, bool EOS = true>
__device__ inline R final_cell_vals(const Params& P, const CellFinalArgsT<R>& A, long c, double vold,
                                    const R (&Uvol)[NEM], const R (&Umas)[NEM], const R (&Uene)[NEM], R Uq,
                                    R delta_ec = R(0), int old_mask_in = -1) {
  {
    const long NC = P.nc;
    const int ne = env_count<NEM>(P);
    const R thr = (R)P.remap_threshold;
    const R vol = (R)vold;
    const int old_mask = old_mask_in >= 0 ? old_mask_in : A.mask[c];
    // cell status per environment (110-143)
    R volt = R(0);
#pragma unroll
    for (int k = 0; k < NEM; ++k)
      if (k < ne) volt += Uvol[k];
    int mask = old_mask;
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      if (k >= ne) break;
      const R fvol = qdiv(Uvol[k], vol);
      const bool in_env = (old_mask >> k) & 1;
      if (fvol < thr && in_env) {
        mask &= ~(1 << k);
      } else if (fvol > thr && !in_env) {
        const R fn = qdiv(Uvol[k], volt);
        if (fn > thr && Umas[k] != R(0)) mask |= (1 << k);
      }
    }
    // "moy"
    R masset = R(0);
    volt = R(0);
#pragma unroll
    for (int k = 0; k < NEM; ++k)
      if (k < ne) {
        volt += Uvol[k];
        masset += Umas[k];
      }
    const R unsurvolt = qdiv(R(1), volt);
    const R unsurvol = qdiv(R(1), vol);
    R fv[NEM], mf[NEM], dk[NEM], ek[NEM], pk[NEM], ck[NEM];
    R somme_frac = R(0);
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      fv[k] = R(0);
      pk[k] = ck[k] = R(0);
      if (!((mask >> k) & 1)) continue;
      R f = Uvol[k] * unsurvolt;
      if (f < thr) f = R(0);
      fv[k] = f;
      somme_frac += f;
    }
    int matcell = 0, imatpure = -1;
    const R unsursomme_frac = qdiv(R(1), somme_frac);
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      if (!((mask >> k) & 1)) continue;
      R f = fv[k] * unsursomme_frac;
      // learned mode and the float precision form: an environment whose
      // fraction ends at or below the threshold leaves the cell (no member
      // with zero fraction and density, whose EOS gives a 0/0 sound speed).
      // The presence test above divides by the Euler volume and the
      // fraction by the sum of the environment volumes; in float the two
      // differ by ~1e-7 relative, so a volume near the threshold can pass
      // the first and be zeroed by the second.
      if ((P.learned || std::is_same<R, float>::value) && f <= thr) {
        f = R(0);
        mask &= ~(1 << k);
      }
      if (f > R(0)) {
        matcell++;
        imatpure = k;
      }
      fv[k] = f;
    }
    // order 3, generic sweeps: the cell's mixed / pure code for the next
    // remap (RemapADIFinal.cc:286-293)
    if ((P.order2 == 2 || P.generic) && A.est) A.est[c] = matcell > 1 ? -2 : imatpure;
    const int nbe = __popc(mask);
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      // a zero total mass leaves the mass fractions as they were (513-538)
      mf[k] = R(0);
      if (nbe == 1) mf[k] = R(1.0);
      else if (k < ne && ((old_mask >> k) & 1) && ((mask >> k) & 1)) mf[k] = A.env_mf[k * NC + c];
    }
    if (masset != R(0)) {
      const R unsurmasset = qdiv(R(1), masset);
      R fmasset = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        R m = Umas[k] * unsurmasset;
        if (fv[k] < thr) m = R(0);
        fmasset += m;
        mf[k] = m;
      }
      if (fmasset != R(0)) {
        const R unsurfmasset = qdiv(R(1), fmasset);
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if ((mask >> k) & 1) mf[k] = mf[k] * unsurfmasset;
      }
    }
    R density_nplus1 = R(0);
    int kpure = 0;
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      dk[k] = R(0);
      if (!((mask >> k) & 1)) continue;
      kpure = k;
      if (fv[k] > thr) {
        const R vol_norm = Uvol[k] * vol * unsurvolt;
        dk[k] = qdiv(Umas[k], vol_norm);
      }
      density_nplus1 += fv[k] * dk[k];
    }
    R rho_g;
    if (matcell > 1) rho_g = density_nplus1;
    else if (nbe == 1) {
      rho_g = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k == kpure) rho_g = dk[k];
    } else {
      rho_g = qdiv(A.m[c], A.Vlag[c]);  // left as updateDensity set it (m/V)
    }
    R mass_g = vol * density_nplus1;
    if (kMassCons<R>) {  // the sum of the remapped masses it equals (see kMassCons)
      mass_g = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (((mask >> k) & 1) && fv[k] > thr) mass_g += Umas[k];
    }
    R energie = R(0), pseudo = R(0);
    const R pv = Uq * unsurvol;
    R mk[NEM];
#pragma unroll
    for (int k = 0; k < NEM; ++k) {
      ek[k] = R(0);
      mk[k] = R(0);
      if (!((mask >> k) & 1)) continue;
      if (fv[k] > thr && Umas[k] != R(0)) ek[k] = qdiv(Uene[k], Umas[k]);
      energie += mf[k] * ek[k];
      mk[k] = mf[k] * mass_g;
      pseudo += fv[k] * pv;
    }
    R e_g = energie, q_g = pseudo;
    R p_g = R(0), c_g = R(1.e-20);
    if (nbe == 1) {
      // pure cell: the environment slot is the cell slot
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k == kpure) mass_g = mk[k];
      R mfp = R(0);
#pragma unroll
      for (int k = 0; k < NEM; ++k)
        if (k == kpure) mfp = mf[k];
      bool reset = false;
      if (rho_g < R(0) || e_g < R(0)) {
        e_g = R(0);
        rho_g = R(0);
        mfp = R(0);
        reset = true;
      }
      // pente-borne reads a pure cell's fraction (1 up to rounding, 0 after
      // a reset): kept in its environment slot
      if (P.pb) {
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if (k == kpure) A.env_f[k * NC + c] = reset ? R(0) : fv[k];
      }
      // total-energy fix-up (RemapADIFinal.cc:751-776), before the EOS
      if (P.etot) e_g += mfp * delta_ec;
      if (EOS) {
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if (k == kpure) eos_apply(P.eos, k, rho_g, e_g, &p_g, &c_g);
      }
    } else {
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        if (dk[k] < R(0) || ek[k] < R(0)) {
          ek[k] = R(0);
          dk[k] = R(0);
          fv[k] = R(0);
          mf[k] = R(0);
        }
      }
      // total-energy fix-up (RemapADIFinal.cc:751-776), before the EOS: the
      // environments by mass fraction, the cell's average in full
      if (P.etot) {
#pragma unroll
        for (int k = 0; k < NEM; ++k)
          if ((mask >> k) & 1) ek[k] += mf[k] * delta_ec;
        e_g += delta_ec;
      }
      // EOS per environment, then computePressionMoyenne; each environment's
      // values are stored as soon as they are final (shorter live ranges)
#pragma unroll
      for (int k = 0; k < NEM; ++k) {
        if (!((mask >> k) & 1)) continue;
        if (EOS) {
          eos_apply(P.eos, k, dk[k], ek[k], &pk[k], &ck[k]);
          p_g += fv[k] * pk[k];
          c_g = smax(ck[k], c_g);
        }
        if (nbe >= 2) {
          const long e = k * NC + c;
          A.env_rho[e] = dk[k];
          A.env_e[e] = ek[k];
          if (EOS) {
            A.env_p[e] = pk[k];
            A.env_cs[e] = ck[k];
          }
          A.env_q[e] = pv;
          A.env_m[e] = mk[k];
          A.env_f[e] = fv[k];
          A.env_mf[e] = mf[k];
        }
      }
    }
    st_stream<0>(P, &A.rho[c], rho_g);
    st_stream<0>(P, &A.e[c], e_g);
    if (EOS) {
      st_stream<0>(P, &A.p[c], p_g);
      st_stream<0>(P, &A.cs[c], c_g);
    }
    st_stream<0>(P, &A.q[c], q_g);
    st_stream<0>(P, &A.m[c], mass_g);
    st_stream<0>(P, &A.mask[c], mask);
    return EOS ? c_g : R(0);
  }
}

// The final cell phase on the remap vector in memory.
template <class R, int NEM, bool EOS = true>
__device__ inline R final_cell(const Params& P, const CellFinalArgsT<R>& A, long c, double vold,
                               R delta_ec = R(0)) {
  const long NC = P.nc;
  const int ne = env_count<NEM>(P);
  R Uvol[NEM], Umas[NEM], Uene[NEM];
#pragma unroll
  for (int k = 0; k < NEM; ++k) {
    Uvol[k] = Umas[k] = Uene[k] = R(0);
    if (k < ne) {
      Uvol[k] = A.U[k * NC + c];
      Umas[k] = A.U[(ne + k) * NC + c];
      Uene[k] = A.U[(2 * ne + k) * NC + c];
    }
  }
  const R Uq = A.U[(3 * ne) * NC + c];
  return final_cell_vals<R, NEM, EOS>(P, A, c, vold, Uvol, Umas, Uene, Uq, delta_ec);
}

// Reduction of the CFL candidates of a block to one minimum, stored at slot.
template <int BLOCK>
__device__ inline void block_min_store(double v, double* slot) {
  __shared__ double red[BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < BLOCK / 64; ++w) b = smin(b, red[w]);
    *slot = b;
  }
}

// Order 3 (RemapADIService.cc:583-640): the six values bbb, bb, b | f, ff,
// fff around the back face of the cell in this lane (s: its index along the
// sweep, N: cells along it), from the lanes three behind to two ahead; a
// cell beyond the mesh is replaced by the last existing one, as the
// reference's previous() / next() walk does.  Every lane runs the shifts.
template <class R>
__device__ __forceinline__ void o3_stencil(R v, int s, int N, R* y) {
  const R m1 = lane_up1(v), m2 = lane_up1(m1), m3 = lane_up1(m2);
  const R p1 = lane_down1(v), p2 = lane_down1(p1);
  y[2] = m1;
  y[1] = s >= 2 ? m2 : m1;
  y[0] = s >= 3 ? m3 : y[1];
  y[3] = v;
  y[4] = s + 1 <= N - 1 ? p1 : v;
  y[5] = s + 2 <= N - 1 ? p2 : y[4];
}

// Order 3 at a face whose neighbourhood is not pure (RemapADIService.cc:
// 623-638): an environment's mass face value is its volume face value times
// the upwind cell's mass per volume, and the energy face value the new mass
// face value times its energy per mass, when the upwind cell (b for vdt >
// 0, f for vdt < 0) holds the environment's volume and mass.  Roles come in
// the order volumes, masses, energies; the per-environment registers carry
// the volume and mass values between them.  pb, pf: the variable's value in
// cells b and f.
template <class R, int NEM>
__device__ __forceinline__ void o3_mixed_faces(bool vpure, int role, int k, R vdt, R pb, R pf, R* phif, R* V,
                                               R* Vb, R* Vf, R* M, R* Mb, R* Mf, int* up) {
#pragma unroll
  for (int q = 0; q < NEM; ++q) {
    if (q != k) continue;
    if (role == 0) {
      V[q] = *phif;
      Vb[q] = pb;
      Vf[q] = pf;
    } else if (role == 1) {
      int u = 0;
      if (!vpure) {
        if (vdt > R(0) && Vb[q] != R(0) && pb != R(0)) u = 1;
        else if (vdt < R(0) && Vf[q] != R(0) && pf != R(0)) u = 2;
      }
      if (u == 1) *phif = qdiv(pb, Vb[q]) * V[q];
      if (u == 2) *phif = qdiv(pf, Vf[q]) * V[q];
      up[q] = u;
      M[q] = *phif;
      Mb[q] = pb;
      Mf[q] = pf;
    } else if (role == 2) {
      if (up[q] == 1) *phif = qdiv(pb, Mb[q]) * M[q];
      if (up[q] == 2) *phif = qdiv(pf, Mf[q]) * M[q];
    }
  }
}

}  // namespace mhg
