// Directional flux sweeps: AUSM+ inviscid flux with MUSCL/NND-2/1st-order
// reconstruction (src/InviscidScheme.cuh:21-149, src/Reconstruction.cu:5-46,
// src/Limiter.cuh:8-41) and the 2nd-order central viscous flux
// (src/ViscousScheme.cuh:22-615), fused per direction: every face flux
// G = F_inviscid - F_viscous is computed once into LDS and each cell takes
// dq(c) -= G(c+1/2) - G(c-1/2), one read-modify-write of dq per direction.
//
// Work layout (one 256-thread workgroup per tile):
//  * i-sweep: tile = TW rows (j) x L cells (i); faces are enumerated row-major
//    so consecutive lanes read consecutive i (coalesced).
//  * j/k-sweeps: tile = 64 cells (i) x L cells along the sweep; faces are
//    enumerated i-fastest so every wavefront reads one contiguous i-row.
// The tile's first face is shared with the previous tile (recomputed, 1/L).
#include <cstdlib>

#include "ostc_kernels.h"
#include "ostc_thermo.h"

namespace ostc {

// minmod with sgn(0)=+1 (src/Limiter.cuh:10-13, src/gxl_lib/Math.hpp:9-15)
template <typename R>
__device__ __forceinline__ R minmod(R a, R b) {
  const R sa = a < (R)0 ? (R)-1 : (R)1, sb = b < (R)0 ? (R)-1 : (R)1;
  return (R)0.5 * (sa + sb) * dmin(fabs(a), fabs(b));
}

// AUSM+ flux through the face between the points at idl and idl+st
// (src/InviscidScheme.cuh:70-149); pv of points idl-st .. idl+2st.
template <typename R, int NS, int NT>
__device__ __forceinline__ void inviscid_face(const ZoneView<R> &z, const ostc_param *__restrict__ p, int dir,
                                              int64_t idl, int64_t st, R *__restrict__ fc) {
  constexpr int NV = 5 + NS + NT;
  R pv[4][NV];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t id = idl + (q - 1) * st;
#pragma unroll
    for (int l = 0; l < 5; ++l) pv[q][l] = z.bv[id + l * z.plane];
#pragma unroll
    for (int l = 0; l < NS + NT; ++l) pv[q][5 + l] = z.sv[id + l * z.plane];
  }
  R pl[NV], pr[NV];
  const int rec = p->reconstruction;
  if (rec == 2) {
    const R kappa = (R)(1.0 / 3.0);
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R d_i = pv[1][l] - pv[0][l], d_i1 = pv[2][l] - pv[1][l], d_i2 = pv[3][l] - pv[2][l];
      const R dnl = minmod(d_i, d_i1), dpl = minmod(d_i1, d_i);
      const R dnr = minmod(d_i1, d_i2), dpr = minmod(d_i2, d_i1);
      pl[l] = pv[1][l] + (R)0.25 * (((R)1 - kappa) * dnl + ((R)1 + kappa) * dpl);
      pr[l] = pv[2][l] - (R)0.25 * (((R)1 - kappa) * dpr + ((R)1 + kappa) * dnr);
    }
  } else if (rec == 3) {
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R d_i = pv[1][l] - pv[0][l], d_i1 = pv[2][l] - pv[1][l], d_i2 = pv[3][l] - pv[2][l];
      pl[l] = pv[1][l] + (R)0.5 * minmod(d_i, d_i1);
      pr[l] = pv[2][l] - (R)0.5 * minmod(d_i1, d_i2);
    }
  } else {
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      pl[l] = pv[1][l];
      pr[l] = pv[2][l];
    }
  }
  R el = (R)0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  R er = (R)0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  R El, Er, gam_l = (R)kGammaAir, gam_r = (R)kGammaAir;
  if constexpr (NS > 0) {
    // mixture E and gamma from the reconstructed T (src/InviscidScheme.cuh:33-61)
    R mwl = 0, mwr = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const R inv = inv_mw<R>(p, l);
      mwl += pl[5 + l] * inv;
      mwr += pr[5 + l] * inv;
    }
    const R tl = fdiv<R>(pl[4], pl[0] * (R)kRu * mwl), tr = fdiv<R>(pr[4], pr[0] * (R)kRu * mwr);
    R hl[NS], hr[NS], cpl[NS], cpr[NS];
    enthalpy_cp<R>(tl, hl, cpl, p, NS);
    enthalpy_cp<R>(tr, hr, cpr, p, NS);
    R cpL = 0, cpR = 0, cvL = 0, cvR = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const R rw = r_over_mw<R>(p, l);
      cpL += cpl[l] * pl[5 + l];
      cpR += cpr[l] * pr[5 + l];
      cvL += pl[5 + l] * (cpl[l] - rw);
      cvR += pr[5 + l] * (cpr[l] - rw);
      el += hl[l] * pl[5 + l];
      er += hr[l] * pr[5 + l];
    }
    El = pl[0] * el - pl[4];
    Er = pr[0] * er - pr[4];
    gam_l = fdiv<R>(cpL, cvL);
    gam_r = fdiv<R>(cpR, cvR);
  } else {
    El = el * pl[0] + fdiv<R>(pl[4], (R)(kGammaAir - 1));
    Er = er * pr[0] + fdiv<R>(pr[4], (R)(kGammaAir - 1));
  }
  const int64_t idr = idl + st;
  const R jl = z.jac[idl], jr = z.jac[idr];
  const int64_t mo = (int64_t)(dir * 3) * z.plane;
  const R k1 = (R)0.5 * (jl * z.metric[mo + idl] + jr * z.metric[mo + idr]);
  const R k2 = (R)0.5 * (jl * z.metric[mo + z.plane + idl] + jr * z.metric[mo + z.plane + idr]);
  const R k3 = (R)0.5 * (jl * z.metric[mo + 2 * z.plane + idl] + jr * z.metric[mo + 2 * z.plane + idr]);
  const R gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const R igk = fdiv<R>((R)1, gk);
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) * igk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) * igk;
  const R c = (R)0.5 * (sqrt(fdiv<R>(gam_l * pl[4], pl[0])) + sqrt(fdiv<R>(gam_r * pr[4], pr[0])));
  const R ic = fdiv<R>((R)1, c);
  const R mach_l = ul * ic, mach_r = ur * ic;
  R mlp, mrn, plp, prn;
  const R alpha = (R)(3 / 16.0);
  if (fabs(mach_l) > (R)1) {
    mlp = (R)0.5 * (mach_l + fabs(mach_l));
    plp = fdiv<R>(mlp, mach_l);
  } else {
    const R a = (mach_l + (R)1) * (mach_l + (R)1) * (R)0.25;
    const R b = (mach_l * mach_l - (R)1) * (mach_l * mach_l - (R)1);
    mlp = a + (R)0.125 * b;
    plp = a * ((R)2 - mach_l) + alpha * mach_l * b;
  }
  if (fabs(mach_r) > (R)1) {
    mrn = (R)0.5 * (mach_r - fabs(mach_r));
    prn = fdiv<R>(mrn, mach_r);
  } else {
    const R a = (mach_r - (R)1) * (mach_r - (R)1) * (R)0.25;
    const R b = (mach_r * mach_r - (R)1) * (mach_r * mach_r - (R)1);
    mrn = -a - (R)0.125 * b;
    prn = a * ((R)2 + mach_r) - alpha * mach_r * b;
  }
  const R p_coeff = plp * pl[4] + prn * pr[4];
  const R m_half = mlp + mrn;
  const R mach_pos = (R)0.5 * (m_half + fabs(m_half));
  const R mach_neg = (R)0.5 * (m_half - fabs(m_half));
  const R mass_flux = c * (pl[0] * mach_pos + pr[0] * mach_neg);
  const R coeff = mass_flux * gk;
  const bool up = mass_flux >= (R)0;
  fc[0] = coeff;
  fc[1] = coeff * (up ? pl[1] : pr[1]) + p_coeff * k1;
  fc[2] = coeff * (up ? pl[2] : pr[2]) + p_coeff * k2;
  fc[3] = coeff * (up ? pl[3] : pr[3]) + p_coeff * k3;
  fc[4] = up ? fdiv<R>(coeff * (El + pl[4]), pl[0]) : fdiv<R>(coeff * (Er + pr[4]), pr[0]);
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] = coeff * (up ? pl[l] : pr[l]);
}

// Derivatives of plane f in (xi, eta, zeta) at the face between id0 and
// id0+s[DIR]: two-point along DIR, four-point average across
// (src/ViscousScheme.cuh:40-55 and their gv/hv counterparts).
template <typename R, int DIR>
__device__ __forceinline__ void face_der(const R *__restrict__ f, int64_t id0, const int64_t s[3], R d[3]) {
  const int64_t id1 = id0 + s[DIR];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q == DIR)
      d[q] = f[id1] - f[id0];
    else
      d[q] = (R)0.25 * (f[id0 + s[q]] - f[id0 - s[q]] + f[id1 + s[q]] - f[id1 - s[q]]);
  }
}

// Species part of the viscous face flux of a mixture (src/ViscousScheme.cuh:
// 158-215): the diffusion fluxes with the correction velocity and the
// enthalpy they carry.  fv[0] receives the energy term, fv[1 + l] species l.
template <typename R, int NS, int NT, int DIR>
__device__ __forceinline__ void viscous_face_species(const ZoneView<R> &z, const ostc_param *__restrict__ p,
                                                     int64_t id0, const int64_t s[3], R *__restrict__ fv) {
  const int64_t id1 = id0 + s[DIR];
  const int64_t P = z.plane;
  R mm[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) mm[q] = (R)0.5 * (z.metric[q * P + id0] + z.metric[q * P + id1]);
  const R xi_x = mm[0], xi_y = mm[1], xi_z = mm[2];
  const R eta_x = mm[3], eta_y = mm[4], eta_z = mm[5];
  const R zeta_x = mm[6], zeta_y = mm[7], zeta_z = mm[8];
  const int64_t mo = (int64_t)(DIR * 3) * P;
  const R ja = z.jac[id0], jb = z.jac[id1];
  const R kx = (R)0.5 * (z.metric[mo + id0] * ja + z.metric[mo + id1] * jb);
  const R ky = (R)0.5 * (z.metric[mo + P + id0] * ja + z.metric[mo + P + id1] * jb);
  const R kz = (R)0.5 * (z.metric[mo + 2 * P + id0] * ja + z.metric[mo + 2 * P + id1] * jb);
  const R *svp = z.sv;
  R turb_d = 0;
  if constexpr (NT > 0) turb_d = fdiv<R>((R)0.5 * (z.mut[id0] + z.mut[id1]), (R)p->Sct);
  R h[NS], dif[NS], yx[NS], yy[NS], yz[NS];
  const R tm = (R)0.5 * (z.bv[5 * P + id0] + z.bv[5 * P + id1]);
  enthalpy<R>(tm, h, p, NS);
  R rho_uc = 0, rho_vc = 0, rho_wc = 0, e = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    dif[l] = (R)0.5 * (z.rho_d[l * P + id0] + z.rho_d[l * P + id1]) + turb_d;
    R dy[3];
    face_der<R, DIR>(svp + l * P, id0, s, dy);
    yx[l] = dy[0] * xi_x + dy[1] * eta_x + dy[2] * zeta_x;
    yy[l] = dy[0] * xi_y + dy[1] * eta_y + dy[2] * zeta_y;
    yz[l] = dy[0] * xi_z + dy[1] * eta_z + dy[2] * zeta_z;
    const R dh = dif[l] * h[l];
    e += kx * dh * yx[l];
    e += ky * dh * yy[l];
    e += kz * dh * yz[l];
    rho_uc += dif[l] * yx[l];
    rho_vc += dif[l] * yy[l];
    rho_wc += dif[l] * yz[l];
  }
  fv[0] = e;
#pragma unroll
  for (int l = 0; l < NS; ++l)
    fv[1 + l] = dif[l] * (kx * yx[l] + ky * yy[l] + kz * yz[l]) -
                (R)0.5 * (svp[l * P + id0] + svp[l * P + id1]) * (kx * rho_uc + ky * rho_vc + kz * rho_wc);
}

// Viscous face flux (src/ViscousScheme.cuh:22-219 and the gv/hv copies).
// SPECIES = false leaves the mixture's species terms to viscous_face_species
// (fv[5 .. 5+NS) = 0 and no enthalpy diffusion in fv[4]).
template <typename R, int NS, int NT, int DIR, bool SPECIES = true>
__device__ __forceinline__ void viscous_face(const ZoneView<R> &z, const ostc_param *__restrict__ p, int64_t id0,
                                             const int64_t s[3], R *__restrict__ fv) {
  constexpr int NV = 5 + NS + NT;
  const int64_t id1 = id0 + s[DIR];
  const int64_t P = z.plane;
  R mm[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) mm[q] = (R)0.5 * (z.metric[q * P + id0] + z.metric[q * P + id1]);
  const R xi_x = mm[0], xi_y = mm[1], xi_z = mm[2];
  const R eta_x = mm[3], eta_y = mm[4], eta_z = mm[5];
  const R zeta_x = mm[6], zeta_y = mm[7], zeta_z = mm[8];
  R du[3], dv[3], dw[3], dt[3];
  face_der<R, DIR>(z.bv + 1 * P, id0, s, du);
  face_der<R, DIR>(z.bv + 2 * P, id0, s, dv);
  face_der<R, DIR>(z.bv + 3 * P, id0, s, dw);
  face_der<R, DIR>(z.bv + 5 * P, id0, s, dt);
  const R u_x = du[0] * xi_x + du[1] * eta_x + du[2] * zeta_x;
  const R u_y = du[0] * xi_y + du[1] * eta_y + du[2] * zeta_y;
  const R u_z = du[0] * xi_z + du[1] * eta_z + du[2] * zeta_z;
  const R v_x = dv[0] * xi_x + dv[1] * eta_x + dv[2] * zeta_x;
  const R v_y = dv[0] * xi_y + dv[1] * eta_y + dv[2] * zeta_y;
  const R v_z = dv[0] * xi_z + dv[1] * eta_z + dv[2] * zeta_z;
  const R w_x = dw[0] * xi_x + dw[1] * eta_x + dw[2] * zeta_x;
  const R w_y = dw[0] * xi_y + dw[1] * eta_y + dw[2] * zeta_y;
  const R w_z = dw[0] * xi_z + dw[1] * eta_z + dw[2] * zeta_z;
  const R mul = (R)0.5 * (z.mul[id0] + z.mul[id1]);
  R mut = 0;
  if constexpr (NT > 0) mut = (R)0.5 * (z.mut[id0] + z.mut[id1]);
  const R vis = mul + mut;
  R tau_xx = fdiv<R>(vis * ((R)4 * u_x - (R)2 * v_y - (R)2 * w_z), (R)3);
  R tau_yy = fdiv<R>(vis * ((R)4 * v_y - (R)2 * u_x - (R)2 * w_z), (R)3);
  R tau_zz = fdiv<R>(vis * ((R)4 * w_z - (R)2 * u_x - (R)2 * v_y), (R)3);
  const R tau_xy = vis * (u_y + v_x), tau_xz = vis * (u_z + w_x), tau_yz = vis * (v_z + w_y);
  const R *svp = z.sv;
  if constexpr (NT > 0) {
    const R tt = (R)(-2.0 / 3) * (R)0.5 *
                 (z.bv[id0] * svp[NS * P + id0] + z.bv[id1] * svp[NS * P + id1]);
    tau_xx += tt;
    tau_yy += tt;
    tau_zz += tt;
  }
  const int64_t mo = (int64_t)(DIR * 3) * P;
  const R ja = z.jac[id0], jb = z.jac[id1];
  const R kx = (R)0.5 * (z.metric[mo + id0] * ja + z.metric[mo + id1] * jb);
  const R ky = (R)0.5 * (z.metric[mo + P + id0] * ja + z.metric[mo + P + id1] * jb);
  const R kz = (R)0.5 * (z.metric[mo + 2 * P + id0] * ja + z.metric[mo + 2 * P + id1] * jb);
  fv[0] = 0;
  fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  const R um = (R)0.5 * (z.bv[P + id0] + z.bv[P + id1]);
  const R vm = (R)0.5 * (z.bv[2 * P + id0] + z.bv[2 * P + id1]);
  const R wm = (R)0.5 * (z.bv[3 * P + id0] + z.bv[3 * P + id1]);
  const R t_x = dt[0] * xi_x + dt[1] * eta_x + dt[2] * zeta_x;
  const R t_y = dt[0] * xi_y + dt[1] * eta_y + dt[2] * zeta_y;
  const R t_z = dt[0] * xi_z + dt[1] * eta_z + dt[2] * zeta_z;
  R cond = (R)0.5 * (z.lambda[id0] + z.lambda[id1]);
  if constexpr (NT > 0) cond += (R)0.5 * (z.lambda_t[id0] + z.lambda_t[id1]);
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + cond * (kx * t_x + ky * t_y + kz * t_z);
  if constexpr (NS > 0 && !SPECIES) {
#pragma unroll
    for (int l = 0; l < NS; ++l) fv[5 + l] = 0;
  } else if constexpr (NS > 0) {
    R turb_d = 0;
    if constexpr (NT > 0) turb_d = fdiv<R>(mut, (R)p->Sct);
    R h[NS], dif[NS], yx[NS], yy[NS], yz[NS];
    const R tm = (R)0.5 * (z.bv[5 * P + id0] + z.bv[5 * P + id1]);
    enthalpy<R>(tm, h, p, NS);
    R rho_uc = 0, rho_vc = 0, rho_wc = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      dif[l] = (R)0.5 * (z.rho_d[l * P + id0] + z.rho_d[l * P + id1]) + turb_d;
      R dy[3];
      face_der<R, DIR>(svp + l * P, id0, s, dy);
      yx[l] = dy[0] * xi_x + dy[1] * eta_x + dy[2] * zeta_x;
      yy[l] = dy[0] * xi_y + dy[1] * eta_y + dy[2] * zeta_y;
      yz[l] = dy[0] * xi_z + dy[1] * eta_z + dy[2] * zeta_z;
      const R dh = dif[l] * h[l];
      fv[4] += kx * dh * yx[l];
      fv[4] += ky * dh * yy[l];
      fv[4] += kz * dh * yz[l];
      rho_uc += dif[l] * yx[l];
      rho_vc += dif[l] * yy[l];
      rho_wc += dif[l] * yz[l];
    }
#pragma unroll
    for (int l = 0; l < NS; ++l)
      fv[5 + l] = dif[l] * (kx * yx[l] + ky * yy[l] + kz * yz[l]) -
                  (R)0.5 * (svp[l * P + id0] + svp[l * P + id1]) * (kx * rho_uc + ky * rho_vc + kz * rho_wc);
  }
  if constexpr (NT > 0) {
    R dk[3], dom[3];
    face_der<R, DIR>(svp + NS * P, id0, s, dk);
    face_der<R, DIR>(svp + (NS + 1) * P, id0, s, dom);
    const R k_x = dk[0] * xi_x + dk[1] * eta_x + dk[2] * zeta_x;
    const R k_y = dk[0] * xi_y + dk[1] * eta_y + dk[2] * zeta_y;
    const R k_z = dk[0] * xi_z + dk[1] * eta_z + dk[2] * zeta_z;
    const R o_x = dom[0] * xi_x + dom[1] * eta_x + dom[2] * zeta_x;
    const R o_y = dom[0] * xi_y + dom[1] * eta_y + dom[2] * zeta_y;
    const R o_z = dom[0] * xi_z + dom[1] * eta_z + dom[2] * zeta_z;
    const R wd = (R)0.5 * (z.wall_dist[id0] + z.wall_dist[id1]);
    R f1 = 1;
    if (wd > (R)1e-25) {
      const R km = (R)0.5 * (svp[NS * P + id0] + svp[NS * P + id1]);
      const R om = (R)0.5 * (svp[(NS + 1) * P + id0] + svp[(NS + 1) * P + id1]);
      const R param1 = fdiv<R>(sqrt(km), (R)0.09 * om * wd);
      const R rhom = (R)0.5 * (z.bv[id0] + z.bv[id1]);
      const R d2 = wd * wd;
      const R param2 = fdiv<R>((R)500 * mul, rhom * d2 * om);
      const R cd = dmax((R)1e-20, fdiv<R>((R)2 * rhom * (R)sst::sigma_omega2, om) * (k_x * o_x + k_y * o_y + k_z * o_z));
      const R param3 = fdiv<R>((R)4 * rhom * (R)sst::sigma_omega2 * km, cd * d2);
      const R arg1 = dmin(dmax(param1, param2), param3);
      f1 = tanh(arg1 * arg1 * arg1 * arg1);
    }
    const R sigma_k = (R)sst::sigma_k2 + (R)sst::delta_sigma_k * f1;
    const R sigma_o = (R)sst::sigma_omega2 + (R)sst::delta_sigma_omega * f1;
    fv[5 + NS] = (mul + mut * sigma_k) * (kx * k_x + ky * k_y + kz * k_z);
    fv[6 + NS] = (mul + mut * sigma_o) * (kx * o_x + ky * o_y + kz * o_z);
  }
}

// ACC: 0 the first writer of dq, 1 accumulating into it, 2 the first writer
// of the flow and species rows, accumulating into the turbulence rows (which
// the SST source wrote before)
template <typename R, int NS, int NT, int DIR, bool INV, bool VISC, int ACC>
__global__ __launch_bounds__(256) void sweep_kernel(ZoneView<R> z, const ostc_param *__restrict__ p, int L,
                                                    int TW) {
  constexpr int NV = 5 + NS + NT;
  extern __shared__ unsigned char smem_raw[];
  R *fc = reinterpret_cast<R *>(smem_raw);
  const int64_t s[3] = {1, (int64_t)z.nx, (int64_t)z.nx * z.ny};
  const int ext[3] = {z.mx, z.my, z.mz};
  // tile origin (cells)
  int o[3];
  const uint3 bk = xcd_block();
  if (DIR == 0) {
    o[0] = bk.x * L;
    o[1] = bk.y * TW;
    o[2] = bk.z;
  } else if (DIR == 1) {
    o[0] = bk.x * TW;
    o[1] = bk.y * L;
    o[2] = bk.z;
  } else {
    o[0] = bk.x * TW;
    o[1] = bk.y;
    o[2] = bk.z * L;
  }
  const int A = DIR == 0 ? 1 : 0;  // across index direction inside the tile
  const int nfaces = TW * (L + 1);
  for (int q = threadIdx.x; q < nfaces; q += blockDim.x) {
    int a, f;
    if (DIR == 0) {
      a = q / (L + 1);
      f = q - a * (L + 1) - 1;
    } else {
      f = q / TW - 1;
      a = q - (f + 1) * TW;
    }
    int c[3] = {o[0], o[1], o[2]};
    c[A] += a;
    c[DIR] += f;
    if (c[A] >= ext[A] || c[DIR] >= ext[DIR]) continue;
    const int64_t id0 = z.idx(c[0], c[1], c[2]);
    R g[NV];
    if (INV) {
      inviscid_face<R, NS, NT>(z, p, DIR, id0, s[DIR], g);
    } else {
#pragma unroll
      for (int l = 0; l < NV; ++l) g[l] = 0;
    }
    if (VISC) {
      R v[NV];
      // mixtures: the species terms in a second pass over the faces below,
      // so that the two halves of the viscous flux never hold registers at
      // the same time
      viscous_face<R, NS, NT, DIR, (NS == 0)>(z, p, id0, s, v);
#pragma unroll
      for (int l = 0; l < NV; ++l) g[l] -= v[l];
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) fc[l * nfaces + q] = g[l];
  }
  if constexpr (VISC && NS > 0) {
    // same face -> thread mapping as above: each thread updates its own faces
    for (int q = threadIdx.x; q < nfaces; q += blockDim.x) {
      int a, f;
      if (DIR == 0) {
        a = q / (L + 1);
        f = q - a * (L + 1) - 1;
      } else {
        f = q / TW - 1;
        a = q - (f + 1) * TW;
      }
      int c[3] = {o[0], o[1], o[2]};
      c[A] += a;
      c[DIR] += f;
      if (c[A] >= ext[A] || c[DIR] >= ext[DIR]) continue;
      const int64_t id0 = z.idx(c[0], c[1], c[2]);
      R v[1 + NS];
      viscous_face_species<R, NS, NT, DIR>(z, p, id0, s, v);
      fc[4 * nfaces + q] -= v[0];
#pragma unroll
      for (int l = 0; l < NS; ++l) fc[(5 + l) * nfaces + q] -= v[1 + l];
    }
  }
  __syncthreads();
  const int ncells = TW * L;
  for (int cidx = threadIdx.x; cidx < ncells; cidx += blockDim.x) {
    int a, f;
    if (DIR == 0) {
      a = cidx / L;
      f = cidx - a * L;
    } else {
      f = cidx / TW;
      a = cidx - f * TW;
    }
    int c[3] = {o[0], o[1], o[2]};
    c[A] += a;
    c[DIR] += f;
    if (c[A] >= ext[A] || c[DIR] >= ext[DIR]) continue;
    const int qr = DIR == 0 ? a * (L + 1) + f + 1 : (f + 1) * TW + a;
    const int ql = DIR == 0 ? qr - 1 : qr - TW;
    const int64_t id = z.idx(c[0], c[1], c[2]);
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R d = fc[l * nfaces + qr] - fc[l * nfaces + ql];
      if (ACC == 1 || (ACC == 2 && l >= 5 + NS))
        z.dq[id + l * z.plane] -= d;
      else
        z.dq[id + l * z.plane] = -d;
    }
  }
}

// ---------------------------------------------------------------------------
// Marching flux kernels for single-species air (NS = 0).  The per-face sweeps
// above load every stencil value of every face from global memory (~110 load
// instructions per j/k face); the vector-memory address unit, not HBM, bounds
// them.  Here each wave owns a 64-lane segment of an i-row (cells
// i0-2 .. i0+61, of which lanes 2..61 are written) and marches along j or k:
// every cell is loaded once, i-neighbours come from DPP lane shifts, the
// marched direction from a register window of rows, and only the central
// differences across the third direction are loaded (two cells per value).
//  * MARCH = 1: i-faces and j-faces, dq = (dq - d_i) - d_j
//  * MARCH = 2: k-faces, dq -= d_k
// Same face arithmetic as inviscid_face / viscous_face, with the cross
// derivatives 0.25*(f(L+e) - f(L-e) + f(R+e) - f(R-e)) formed as
// 0.25*(D_e f(L) + D_e f(R)), D_e f = f(+e) - f(-e).

// value of lane l-1 / l+1 (DPP wave shift; the first / last lane keeps x)
template <typename R>
__device__ __forceinline__ R lane_prev(R x) {
  if constexpr (sizeof(R) == 4) {
    const int v = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(R, __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false));
  } else {
    const long long v = __builtin_bit_cast(long long, x);
    const int lo = (int)v, hi = (int)(v >> 32);
    const int a = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
    const int b = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(R, (long long)(((unsigned long long)(unsigned)b << 32) | (unsigned)a));
  }
}
template <typename R>
__device__ __forceinline__ R lane_next(R x) {
  if constexpr (sizeof(R) == 4) {
    const int v = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(R, __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false));
  } else {
    const long long v = __builtin_bit_cast(long long, x);
    const int lo = (int)v, hi = (int)(v >> 32);
    const int a = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
    const int b = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(R, (long long)(((unsigned long long)(unsigned)b << 32) | (unsigned)a));
  }
}

// what a face needs of each of its two cells
template <typename R, int NT>
struct FaceCell {
  R pv[5 + NT];  // rho u v w p [k omega]
  R t;
  R m[9];        // metric (xi_x xi_y xi_z eta_x ... zeta_z)
  R jac, mul, mut, lam, lamt, wd;
};

template <typename R, int NT>
__device__ __forceinline__ void face_cell_load(const ZoneView<R> &z, int64_t q, FaceCell<R, NT> &c, bool visc) {
  const int64_t P = z.plane;
#pragma unroll
  for (int l = 0; l < 5; ++l) c.pv[l] = z.bv[l * P + q];
#pragma unroll
  for (int l = 0; l < NT; ++l) c.pv[5 + l] = z.sv[l * P + q];
#pragma unroll
  for (int r = 0; r < 9; ++r) c.m[r] = z.metric[r * P + q];
  c.jac = z.jac[q];
  c.t = c.mul = c.mut = c.lam = c.lamt = c.wd = 0;
  if (visc) {
    c.t = z.bv[5 * P + q];
    c.mul = z.mul[q];
    c.lam = z.lambda[q];
    if constexpr (NT > 0) {
      c.mut = z.mut[q];
      c.lamt = z.lambda_t[q];
      c.wd = z.wall_dist[q];
    }
  }
}

// gradient variables: u v w T [k omega]
template <typename R, int NT>
__device__ __forceinline__ R grad_var(const FaceCell<R, NT> &c, int g) {
  return g < 3 ? c.pv[1 + g] : (g == 3 ? c.t : c.pv[5 + (g - 4)]);
}

template <typename R, int NT>
__device__ __forceinline__ void face_cell_shift(const FaceCell<R, NT> &a, FaceCell<R, NT> &b, bool next) {
#pragma unroll
  for (int l = 0; l < 5 + NT; ++l) b.pv[l] = next ? lane_next(a.pv[l]) : lane_prev(a.pv[l]);
#pragma unroll
  for (int r = 0; r < 9; ++r) b.m[r] = next ? lane_next(a.m[r]) : lane_prev(a.m[r]);
  b.t = next ? lane_next(a.t) : lane_prev(a.t);
  b.jac = next ? lane_next(a.jac) : lane_prev(a.jac);
  b.mul = next ? lane_next(a.mul) : lane_prev(a.mul);
  b.mut = next ? lane_next(a.mut) : lane_prev(a.mut);
  b.lam = next ? lane_next(a.lam) : lane_prev(a.lam);
  b.lamt = next ? lane_next(a.lamt) : lane_prev(a.lamt);
  b.wd = next ? lane_next(a.wd) : lane_prev(a.wd);
}

// minmod of minmod() above in select form (same values bit for bit:
// sgn(0) = +1, equal signs -> +-min(|a|,|b|), else +0)
template <typename R>
__device__ __forceinline__ R minmod_sel(R a, R b) {
  const bool na = a < (R)0, nb = b < (R)0;
  const R m = dmin(fabs(a), fabs(b));
  return na == nb ? (na ? -m : m) : (R)0;
}

// AUSM+ (src/InviscidScheme.cuh:70-149) on values: pv = points L-1, L, R, R+1.
// minmod is symmetric, so MUSCL's minmod(d_i, d_i1) and minmod(d_i1, d_i)
// are one value (formed once).
template <typename R, int NT, int DIR>
__device__ __forceinline__ void inviscid_face_v(const ostc_param *__restrict__ p, const R (&pv)[4][5 + NT],
                                                const FaceCell<R, NT> &cl, const FaceCell<R, NT> &cr,
                                                R *__restrict__ fc) {
  constexpr int NV = 5 + NT;
  R pl[NV], pr[NV];
  const int rec = p->reconstruction;
  if (rec == 2) {
    const R kappa = (R)(1.0 / 3.0);
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R d_i = pv[1][l] - pv[0][l], d_i1 = pv[2][l] - pv[1][l], d_i2 = pv[3][l] - pv[2][l];
      const R ml = minmod_sel(d_i, d_i1), mr = minmod_sel(d_i1, d_i2);
      pl[l] = pv[1][l] + (R)0.25 * (((R)1 - kappa) * ml + ((R)1 + kappa) * ml);
      pr[l] = pv[2][l] - (R)0.25 * (((R)1 - kappa) * mr + ((R)1 + kappa) * mr);
    }
  } else if (rec == 3) {
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R d_i = pv[1][l] - pv[0][l], d_i1 = pv[2][l] - pv[1][l], d_i2 = pv[3][l] - pv[2][l];
      pl[l] = pv[1][l] + (R)0.5 * minmod_sel(d_i, d_i1);
      pr[l] = pv[2][l] - (R)0.5 * minmod_sel(d_i1, d_i2);
    }
  } else {
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      pl[l] = pv[1][l];
      pr[l] = pv[2][l];
    }
  }
  const R el = (R)0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  const R er = (R)0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  const R El = el * pl[0] + fdiv<R>(pl[4], (R)(kGammaAir - 1));
  const R Er = er * pr[0] + fdiv<R>(pr[4], (R)(kGammaAir - 1));
  const R gam_l = (R)kGammaAir, gam_r = (R)kGammaAir;
  const R jl = cl.jac, jr = cr.jac;
  const R k1 = (R)0.5 * (jl * cl.m[DIR * 3] + jr * cr.m[DIR * 3]);
  const R k2 = (R)0.5 * (jl * cl.m[DIR * 3 + 1] + jr * cr.m[DIR * 3 + 1]);
  const R k3 = (R)0.5 * (jl * cl.m[DIR * 3 + 2] + jr * cr.m[DIR * 3 + 2]);
  const R gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const R igk = fdiv<R>((R)1, gk);
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) * igk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) * igk;
  const R c = (R)0.5 * (sqrt(fdiv<R>(gam_l * pl[4], pl[0])) + sqrt(fdiv<R>(gam_r * pr[4], pr[0])));
  const R ic = fdiv<R>((R)1, c);
  const R mach_l = ul * ic, mach_r = ur * ic;
  R mlp, mrn, plp, prn;
  const R alpha = (R)(3 / 16.0);
  if (fabs(mach_l) > (R)1) {
    mlp = (R)0.5 * (mach_l + fabs(mach_l));
    plp = fdiv<R>(mlp, mach_l);
  } else {
    const R a = (mach_l + (R)1) * (mach_l + (R)1) * (R)0.25;
    const R b = (mach_l * mach_l - (R)1) * (mach_l * mach_l - (R)1);
    mlp = a + (R)0.125 * b;
    plp = a * ((R)2 - mach_l) + alpha * mach_l * b;
  }
  if (fabs(mach_r) > (R)1) {
    mrn = (R)0.5 * (mach_r - fabs(mach_r));
    prn = fdiv<R>(mrn, mach_r);
  } else {
    const R a = (mach_r - (R)1) * (mach_r - (R)1) * (R)0.25;
    const R b = (mach_r * mach_r - (R)1) * (mach_r * mach_r - (R)1);
    mrn = -a - (R)0.125 * b;
    prn = a * ((R)2 + mach_r) - alpha * mach_r * b;
  }
  const R p_coeff = plp * pl[4] + prn * pr[4];
  const R m_half = mlp + mrn;
  const R mach_pos = (R)0.5 * (m_half + fabs(m_half));
  const R mach_neg = (R)0.5 * (m_half - fabs(m_half));
  const R mass_flux = c * (pl[0] * mach_pos + pr[0] * mach_neg);
  const R coeff = mass_flux * gk;
  const bool up = mass_flux >= (R)0;
  fc[0] = coeff;
  fc[1] = coeff * (up ? pl[1] : pr[1]) + p_coeff * k1;
  fc[2] = coeff * (up ? pl[2] : pr[2]) + p_coeff * k2;
  fc[3] = coeff * (up ? pl[3] : pr[3]) + p_coeff * k3;
  fc[4] = up ? fdiv<R>(coeff * (El + pl[4]), pl[0]) : fdiv<R>(coeff * (Er + pr[4]), pr[0]);
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] = coeff * (up ? pl[l] : pr[l]);
}

// viscous face flux (src/ViscousScheme.cuh:22-219) on values; dl/dr[e][g] =
// D_e of gradient variable g at L / R (row e == DIR unused)
template <typename R, int NT, int DIR>
__device__ __forceinline__ void viscous_face_v(const ostc_param *__restrict__ p, const FaceCell<R, NT> &cl,
                                               const FaceCell<R, NT> &cr, const R (&dl)[3][4 + NT],
                                               const R (&dr)[3][4 + NT], R *__restrict__ fv) {
  constexpr int NG = 4 + NT;
  R mm[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) mm[q] = (R)0.5 * (cl.m[q] + cr.m[q]);
  const R xi_x = mm[0], xi_y = mm[1], xi_z = mm[2];
  const R eta_x = mm[3], eta_y = mm[4], eta_z = mm[5];
  const R zeta_x = mm[6], zeta_y = mm[7], zeta_z = mm[8];
  R d[NG][3];
#pragma unroll
  for (int gq = 0; gq < NG; ++gq)
#pragma unroll
    for (int e = 0; e < 3; ++e)
      d[gq][e] = e == DIR ? grad_var(cr, gq) - grad_var(cl, gq) : (R)0.25 * (dl[e][gq] + dr[e][gq]);
  const R *du = d[0], *dv = d[1], *dw = d[2], *dt = d[3];
  const R u_x = du[0] * xi_x + du[1] * eta_x + du[2] * zeta_x;
  const R u_y = du[0] * xi_y + du[1] * eta_y + du[2] * zeta_y;
  const R u_z = du[0] * xi_z + du[1] * eta_z + du[2] * zeta_z;
  const R v_x = dv[0] * xi_x + dv[1] * eta_x + dv[2] * zeta_x;
  const R v_y = dv[0] * xi_y + dv[1] * eta_y + dv[2] * zeta_y;
  const R v_z = dv[0] * xi_z + dv[1] * eta_z + dv[2] * zeta_z;
  const R w_x = dw[0] * xi_x + dw[1] * eta_x + dw[2] * zeta_x;
  const R w_y = dw[0] * xi_y + dw[1] * eta_y + dw[2] * zeta_y;
  const R w_z = dw[0] * xi_z + dw[1] * eta_z + dw[2] * zeta_z;
  const R mul = (R)0.5 * (cl.mul + cr.mul);
  R mut = 0;
  if constexpr (NT > 0) mut = (R)0.5 * (cl.mut + cr.mut);
  const R vis = mul + mut;
  R tau_xx = fdiv<R>(vis * ((R)4 * u_x - (R)2 * v_y - (R)2 * w_z), (R)3);
  R tau_yy = fdiv<R>(vis * ((R)4 * v_y - (R)2 * u_x - (R)2 * w_z), (R)3);
  R tau_zz = fdiv<R>(vis * ((R)4 * w_z - (R)2 * u_x - (R)2 * v_y), (R)3);
  const R tau_xy = vis * (u_y + v_x), tau_xz = vis * (u_z + w_x), tau_yz = vis * (v_z + w_y);
  if constexpr (NT > 0) {
    const R tt = (R)(-2.0 / 3) * (R)0.5 * (cl.pv[0] * cl.pv[5] + cr.pv[0] * cr.pv[5]);
    tau_xx += tt;
    tau_yy += tt;
    tau_zz += tt;
  }
  const R ja = cl.jac, jb = cr.jac;
  const R kx = (R)0.5 * (cl.m[DIR * 3] * ja + cr.m[DIR * 3] * jb);
  const R ky = (R)0.5 * (cl.m[DIR * 3 + 1] * ja + cr.m[DIR * 3 + 1] * jb);
  const R kz = (R)0.5 * (cl.m[DIR * 3 + 2] * ja + cr.m[DIR * 3 + 2] * jb);
  fv[0] = 0;
  fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  const R um = (R)0.5 * (cl.pv[1] + cr.pv[1]);
  const R vm = (R)0.5 * (cl.pv[2] + cr.pv[2]);
  const R wm = (R)0.5 * (cl.pv[3] + cr.pv[3]);
  const R t_x = dt[0] * xi_x + dt[1] * eta_x + dt[2] * zeta_x;
  const R t_y = dt[0] * xi_y + dt[1] * eta_y + dt[2] * zeta_y;
  const R t_z = dt[0] * xi_z + dt[1] * eta_z + dt[2] * zeta_z;
  R cond = (R)0.5 * (cl.lam + cr.lam);
  if constexpr (NT > 0) cond += (R)0.5 * (cl.lamt + cr.lamt);
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + cond * (kx * t_x + ky * t_y + kz * t_z);
  if constexpr (NT > 0) {
    const R *dk = d[4], *dom = d[5];
    const R k_x = dk[0] * xi_x + dk[1] * eta_x + dk[2] * zeta_x;
    const R k_y = dk[0] * xi_y + dk[1] * eta_y + dk[2] * zeta_y;
    const R k_z = dk[0] * xi_z + dk[1] * eta_z + dk[2] * zeta_z;
    const R o_x = dom[0] * xi_x + dom[1] * eta_x + dom[2] * zeta_x;
    const R o_y = dom[0] * xi_y + dom[1] * eta_y + dom[2] * zeta_y;
    const R o_z = dom[0] * xi_z + dom[1] * eta_z + dom[2] * zeta_z;
    const R wd = (R)0.5 * (cl.wd + cr.wd);
    R f1 = 1;
    if (wd > (R)1e-25) {
      const R km = (R)0.5 * (cl.pv[5] + cr.pv[5]);
      const R om = (R)0.5 * (cl.pv[6] + cr.pv[6]);
      const R param1 = fdiv<R>(sqrt(km), (R)0.09 * om * wd);
      const R rhom = (R)0.5 * (cl.pv[0] + cr.pv[0]);
      const R d2 = wd * wd;
      const R param2 = fdiv<R>((R)500 * mul, rhom * d2 * om);
      const R cd = dmax((R)1e-20, fdiv<R>((R)2 * rhom * (R)sst::sigma_omega2, om) * (k_x * o_x + k_y * o_y + k_z * o_z));
      const R param3 = fdiv<R>((R)4 * rhom * (R)sst::sigma_omega2 * km, cd * d2);
      const R arg1 = dmin(dmax(param1, param2), param3);
      f1 = tanh(arg1 * arg1 * arg1 * arg1);
    }
    const R sigma_k = (R)sst::sigma_k2 + (R)sst::delta_sigma_k * f1;
    const R sigma_o = (R)sst::sigma_omega2 + (R)sst::delta_sigma_omega * f1;
    fv[5] = (mul + mut * sigma_k) * (kx * k_x + ky * k_y + kz * k_z);
    fv[6] = (mul + mut * sigma_o) * (kx * o_x + ky * o_y + kz * o_z);
  }
}

template <typename R, int NT, int DIR, bool INV, bool VISC>
__device__ __forceinline__ void face_flux_v(const ostc_param *__restrict__ p, const R (&pv)[4][5 + NT],
                                            const FaceCell<R, NT> &cl, const FaceCell<R, NT> &cr,
                                            const R (&dl)[3][4 + NT], const R (&dr)[3][4 + NT], R *__restrict__ g) {
  constexpr int NV = 5 + NT;
  if constexpr (INV) {
    inviscid_face_v<R, NT, DIR>(p, pv, cl, cr, g);
  } else {
#pragma unroll
    for (int l = 0; l < NV; ++l) g[l] = 0;
  }
  if constexpr (VISC) {
    R v[NV];
    viscous_face_v<R, NT, DIR>(p, cl, cr, dl, dr, v);
#pragma unroll
    for (int l = 0; l < NV; ++l) g[l] -= v[l];
  }
}

constexpr int kSegUseful = 60;  // written lanes 2..61 of a 64-lane segment

// ACC (MARCH = 1): 0 the first writer of dq, 1 accumulating into it, 2 the
// first writer of the flow rows, accumulating into the turbulence rows
// (which the SST source wrote before); MARCH = 2 always accumulates
template <typename R, int NT, int MARCH, bool INV, bool VISC, int ACC>
__global__ __launch_bounds__(256) void k_flux_march(ZoneView<R> z, const ostc_param *__restrict__ p, int mlen) {
  constexpr int NV = 5 + NT, NG = 4 + NT;
  constexpr int X = 3 - MARCH;  // the direction whose central differences are loaded
  const int lane = threadIdx.x, w = threadIdx.y;
  const uint3 bk = xcd_block();
  const int ext[3] = {z.mx, z.my, z.mz};
  const int i = (int)bk.x * kSegUseful - 2 + lane;
  const int ic = min(i, z.mx + 1);
  const bool useful = lane >= 2 && lane < 2 + kSegUseful && i < z.mx;
  const int cx = (int)bk.y * (int)blockDim.y + w;
  if (cx >= ext[X]) return;
  const int m0 = (int)bk.z * mlen, m1 = min(m0 + mlen, ext[MARCH]);
  const int64_t sX = X == 1 ? (int64_t)z.nx : (int64_t)z.nx * z.ny;
  const int64_t P = z.plane;
  auto cell = [&](int m) -> int64_t { return MARCH == 1 ? z.idx(ic, m, cx) : z.idx(ic, cx, m); };
  auto load_dx = [&](int64_t q, R *dx) {
#pragma unroll
    for (int gq = 0; gq < NG; ++gq) {
      const R *f = gq < 3 ? z.bv + (1 + gq) * P : (gq == 3 ? z.bv + 5 * P : z.sv + (gq - 4) * P);
      dx[gq] = f[q + sX] - f[q - sX];
    }
  };
  // window: pvA = row c-2 (pv only); cw[0..2] = rows c-1, c, c+1; dxw = their
  // D_X.  (Reading row c+2 one step ahead instead costs the k kernel a wave
  // per SIMD and measured slower.)
  R pvA[NV];
  FaceCell<R, NT> cw[3];
  R dxw[3][NG];
  R gprev[NV], diprev[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) gprev[l] = diprev[l] = 0;
  // prologue: rows m0-2, m0-1, m0
  {
    FaceCell<R, NT> c0;
    face_cell_load<R, NT>(z, cell(m0 - 2), c0, VISC);
#pragma unroll
    for (int l = 0; l < NV; ++l) pvA[l] = c0.pv[l];
  }
  face_cell_load<R, NT>(z, cell(m0 - 1), cw[0], VISC);
  face_cell_load<R, NT>(z, cell(m0), cw[1], VISC);
  if (VISC) {
    load_dx(cell(m0 - 1), dxw[0]);
    load_dx(cell(m0), dxw[1]);
  }
  for (int r = m0 + 1; r <= m1 + 1; ++r) {
    // load row r = c+1
    face_cell_load<R, NT>(z, cell(r), cw[2], VISC);
    if (VISC) load_dx(cell(r), dxw[2]);
    const int c = r - 1;
    // derivative sets of rows c-1 (L of the marched face) and c (R)
    R dL[3][NG], dR[3][NG];
#pragma unroll
    for (int gq = 0; gq < NG; ++gq) {
      const R f0 = grad_var(cw[0], gq), f1 = grad_var(cw[1], gq);
      dL[0][gq] = lane_next(f0) - lane_prev(f0);
      dR[0][gq] = lane_next(f1) - lane_prev(f1);
      dL[X][gq] = dxw[0][gq];
      dR[X][gq] = dxw[1][gq];
      dL[MARCH][gq] = 0;
      // D_MARCH of row c (cross term of its i-faces)
      dR[MARCH][gq] = grad_var(cw[2], gq) - f0;
    }
    R acc[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) acc[l] = 0;
    if constexpr (MARCH == 1) {
      // i-faces of row c: each lane the face on its left (L = lane-1, R = lane)
      FaceCell<R, NT> cm;
      face_cell_shift<R, NT>(cw[1], cm, false);
      R pvi[4][NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        pvi[2][l] = cw[1].pv[l];
        pvi[1][l] = cm.pv[l];
        pvi[0][l] = lane_prev(cm.pv[l]);
        pvi[3][l] = lane_next(cw[1].pv[l]);
      }
      R dli[3][NG], dri[3][NG];
#pragma unroll
      for (int gq = 0; gq < NG; ++gq) {
        dri[0][gq] = 0;
        dri[1][gq] = dR[1][gq];
        dri[2][gq] = dR[2][gq];
        dli[0][gq] = 0;
        dli[1][gq] = lane_prev(dR[1][gq]);
        dli[2][gq] = lane_prev(dR[2][gq]);
      }
      R gi[NV];
      face_flux_v<R, NT, 0, INV, VISC>(p, pvi, cm, cw[1], dli, dri, gi);
#pragma unroll
      for (int l = 0; l < NV; ++l) acc[l] = lane_next(gi[l]) - gi[l];  // d_i(c)
    }
    // marched face c-1/2 between rows c-1 and c
    {
      R pvm[4][NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        pvm[0][l] = pvA[l];
        pvm[1][l] = cw[0].pv[l];
        pvm[2][l] = cw[1].pv[l];
        pvm[3][l] = cw[2].pv[l];
      }
      R dRm[3][NG];
#pragma unroll
      for (int gq = 0; gq < NG; ++gq) {
        dRm[0][gq] = dR[0][gq];
        dRm[X][gq] = dR[X][gq];
        dRm[MARCH][gq] = 0;
      }
      R gm[NV];
      face_flux_v<R, NT, MARCH, INV, VISC>(p, pvm, cw[0], cw[1], dL, dRm, gm);
      // row c-1 is complete: d_M(c-1) = G(c-1/2) - G(c-3/2)
      const int cm1 = c - 1;
      if (useful && cm1 >= m0 && cm1 < m1) {
        const int64_t q = cell(cm1);
#pragma unroll
        for (int l = 0; l < NV; ++l) {
          const R dmv = gm[l] - gprev[l];
          R v;
          if constexpr (MARCH == 1) {
            const R base = (ACC == 1 || (ACC == 2 && l >= 5)) ? z.dq[q + l * P] - diprev[l] : -diprev[l];
            v = base - dmv;
          } else {
            v = z.dq[q + l * P] - dmv;
          }
          z.dq[q + l * P] = v;
        }
      }
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        gprev[l] = gm[l];
        diprev[l] = acc[l];
      }
    }
    // shift the window
#pragma unroll
    for (int l = 0; l < NV; ++l) pvA[l] = cw[0].pv[l];
    cw[0] = cw[1];
    cw[1] = cw[2];
#pragma unroll
    for (int gq = 0; gq < NG; ++gq) {
      dxw[0][gq] = dxw[1][gq];
      dxw[1][gq] = dxw[2][gq];
    }
  }
}

template <typename R, int NT>
static int launch_march(const ZoneView<R> &z, const ostc_param *dp, bool inv, bool visc, bool first_writer,
                        hipStream_t st, bool keep_turb) {
  const int segs = (z.mx + kSegUseful - 1) / kSegUseful;
  // rows marched per wave (launch tuning; 3 extra rows are loaded per chunk)
  static const int mlen = [] {
    const char *e = std::getenv("OSTC_FLUX_MLEN");
    const int v = e ? std::atoi(e) : 8;
    return v < 1 ? 1 : v;
  }();
  const dim3 block(64, 4);
#define OSTC_MARCH(M, I, V, A) \
  hipLaunchKernelGGL((k_flux_march<R, NT, M, I, V, A>), grid, block, 0, st, z, dp, mlen)
  for (int m = 1; m <= 2; ++m) {
    if (m == 2 && z.mz == 1) break;
    const int X = 3 - m;
    const int extX = X == 1 ? z.my : z.mz, extM = m == 1 ? z.my : z.mz;
    const dim3 grid((unsigned)segs, (unsigned)((extX + 3) / 4), (unsigned)((extM + mlen - 1) / mlen));
    const int acc = !(first_writer && m == 1) ? 1 : (keep_turb && NT > 0 ? 2 : 0);
    prof_mark(m == 1 ? P_FLUX_IJ : P_FLUX_K, st, true);
    if (m == 1) {
      if (acc == 2) {
        if (inv && visc) OSTC_MARCH(1, true, true, 2);
        else if (inv) OSTC_MARCH(1, true, false, 2);
        else OSTC_MARCH(1, false, true, 2);
      } else if (inv && visc) { if (acc) OSTC_MARCH(1, true, true, 1); else OSTC_MARCH(1, true, true, 0); }
      else if (inv) { if (acc) OSTC_MARCH(1, true, false, 1); else OSTC_MARCH(1, true, false, 0); }
      else { if (acc) OSTC_MARCH(1, false, true, 1); else OSTC_MARCH(1, false, true, 0); }
    } else {
      if (inv && visc) OSTC_MARCH(2, true, true, 1);
      else if (inv) OSTC_MARCH(2, true, false, 1);
      else OSTC_MARCH(2, false, true, 1);
    }
    prof_mark(m == 1 ? P_FLUX_IJ : P_FLUX_K, st, false);
    if (hipGetLastError() != hipSuccess) return OSTC_ERR_HIP;
  }
#undef OSTC_MARCH
  return OSTC_OK;
}

// LDS budget of one sweep workgroup: 160 KB per CU holds three of them, the
// residency the kernels' registers allow (a larger tile would cost a
// workgroup per CU)
constexpr int kSweepLds = 53248;

template <typename R, int NS, int NT>
static int launch_sweeps(const ZoneView<R> &z, const ostc_param *dp, bool inv, bool visc, bool first_writer,
                         hipStream_t st, bool keep_turb) {
  constexpr int NV = 5 + NS + NT;
  const int ext[3] = {z.mx, z.my, z.mz};
  for (int dir = 0; dir < 3; ++dir) {
    if (dir == 2 && z.mz == 1) break;
    int L, TW;
    dim3 grid;
    if (dir == 0) {
      L = ext[0] < 256 ? ext[0] : 256;
      TW = 1024 / (L + 1);
      const int lds_cap = kSweepLds / ((L + 1) * NV * (int)sizeof(R));
      if (TW > lds_cap) TW = lds_cap;
      if (TW < 1) TW = 1;
      if (TW > ext[1]) TW = ext[1];
      grid = dim3((ext[0] + L - 1) / L, (ext[1] + TW - 1) / TW, ext[2]);
    } else {
      // cells along the sweep per tile (launch tuning; environment override)
      static const int l_env = [] {
        const char *e = std::getenv("OSTC_SWEEP_L");
        return e ? std::atoi(e) : 0;
      }();
      TW = ext[0] < 64 ? ext[0] : 64;
      const auto lds_ok = [&](int l) { return TW * (l + 1) * NV * (int)sizeof(R) <= kSweepLds; };
      if (l_env > 0) {
        L = l_env;
      } else {
        // the longest tile (<= 16 cells) whose TW * (L + 1) faces fill whole
        // rounds of the 256 threads (a partly filled last round costs as much
        // as a full one), else the longest that fits
        L = 0;
        for (int c = 16; c >= 2 && L == 0; --c)
          if ((TW * (c + 1)) % 256 == 0 && lds_ok(c)) L = c;
        if (L == 0) L = 16;
      }
      while (L > 1 && !lds_ok(L)) L /= 2;
      if (L > ext[dir]) L = ext[dir];
      if (dir == 1)
        grid = dim3((ext[0] + TW - 1) / TW, (ext[1] + L - 1) / L, ext[2]);
      else
        grid = dim3((ext[0] + TW - 1) / TW, ext[1], (ext[2] + L - 1) / L);
    }
    const size_t lds = (size_t)TW * (L + 1) * NV * sizeof(R);
    const bool acc = !(first_writer && dir == 0);
    // the first writer of the flow and species rows, the turbulence rows
    // accumulated (the SST source wrote them)
    const bool turb_acc = !acc && keep_turb && NT > 0;
    prof_mark(P_FLUX_I + dir, st, true);
#define OSTC_SWEEP(D, I, V, A) \
  hipLaunchKernelGGL((sweep_kernel<R, NS, NT, D, I, V, A>), grid, dim3(256), lds, st, z, dp, L, TW)
#define OSTC_SWEEP_D(D)                               \
  do {                                                \
    if (inv && visc) {                                \
      if (acc) OSTC_SWEEP(D, true, true, 1);          \
      else OSTC_SWEEP(D, true, true, 0);              \
    } else if (inv) {                                 \
      if (acc) OSTC_SWEEP(D, true, false, 1);         \
      else OSTC_SWEEP(D, true, false, 0);             \
    } else {                                          \
      if (acc) OSTC_SWEEP(D, false, true, 1);         \
      else OSTC_SWEEP(D, false, true, 0);             \
    }                                                 \
  } while (0)
    if (turb_acc) {
      if (inv && visc)
        OSTC_SWEEP(0, true, true, 2);
      else if (inv)
        OSTC_SWEEP(0, true, false, 2);
      else
        OSTC_SWEEP(0, false, true, 2);
    } else if (dir == 0) {
      OSTC_SWEEP_D(0);
    } else if (dir == 1) {
      OSTC_SWEEP_D(1);
    } else {
      OSTC_SWEEP_D(2);
    }
#undef OSTC_SWEEP_D
#undef OSTC_SWEEP
    prof_mark(P_FLUX_I + dir, st, false);
    if (hipGetLastError() != hipSuccess) return OSTC_ERR_HIP;
  }
  return OSTC_OK;
}

template <typename R>
int flux_sweeps(const ZoneView<R> &z, const ostc_param *dp, int ns, int nt, bool inv, bool visc, bool first_writer,
                hipStream_t st, bool keep_turb) {
  if (ns == 0 && nt == 0) return launch_march<R, 0>(z, dp, inv, visc, first_writer, st, keep_turb);
  if (ns == 0 && nt == 2) return launch_march<R, 2>(z, dp, inv, visc, first_writer, st, keep_turb);
  if (ns == 9 && nt == 0) return launch_sweeps<R, 9, 0>(z, dp, inv, visc, first_writer, st, keep_turb);
  if (ns == 9 && nt == 2) return launch_sweeps<R, 9, 2>(z, dp, inv, visc, first_writer, st, keep_turb);
  if (ns == 20 && nt == 0) return launch_sweeps<R, 20, 0>(z, dp, inv, visc, first_writer, st, keep_turb);
  if (ns == 20 && nt == 2) return launch_sweeps<R, 20, 2>(z, dp, inv, visc, first_writer, st, keep_turb);
  return OSTC_ERR_ARG;
}

template int flux_sweeps<float>(const ZoneView<float> &, const ostc_param *, int, int, bool, bool, bool, hipStream_t,
                                bool);
template int flux_sweeps<double>(const ZoneView<double> &, const ostc_param *, int, int, bool, bool, bool,
                                 hipStream_t, bool);

}  // namespace ostc
