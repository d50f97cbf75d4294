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
      const R inv = (R)(1.0 / p->mw[l]);
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
      const R rw = (R)(kRu / p->mw[l]);
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

// Viscous face flux (src/ViscousScheme.cuh:22-219 and the gv/hv copies).
template <typename R, int NS, int NT, int DIR>
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
  if constexpr (NS > 0) {
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

template <typename R, int NS, int NT, int DIR, bool INV, bool VISC, bool ACCUM>
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
      viscous_face<R, NS, NT, DIR>(z, p, id0, s, v);
#pragma unroll
      for (int l = 0; l < NV; ++l) g[l] -= v[l];
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) fc[l * nfaces + q] = g[l];
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
      if (ACCUM)
        z.dq[id + l * z.plane] -= d;
      else
        z.dq[id + l * z.plane] = -d;
    }
  }
}

template <typename R, int NS, int NT>
static int launch_sweeps(const ZoneView<R> &z, const ostc_param *dp, bool inv, bool visc, bool first_writer,
                         hipStream_t st) {
  constexpr int NV = 5 + NS + NT;
  const int ext[3] = {z.mx, z.my, z.mz};
  for (int dir = 0; dir < 3; ++dir) {
    if (dir == 2 && z.mz == 1) break;
    int L, TW;
    dim3 grid;
    if (dir == 0) {
      L = ext[0] < 256 ? ext[0] : 256;
      TW = 1024 / (L + 1);
      const int lds_cap = 65536 / ((L + 1) * NV * (int)sizeof(R));
      if (TW > lds_cap) TW = lds_cap;
      if (TW < 1) TW = 1;
      if (TW > ext[1]) TW = ext[1];
      grid = dim3((ext[0] + L - 1) / L, (ext[1] + TW - 1) / TW, ext[2]);
    } else {
      TW = ext[0] < 64 ? ext[0] : 64;
      L = 16;
      while (L > 1 && TW * (L + 1) * NV * (int)sizeof(R) > 65536) L /= 2;
      if (L > ext[dir]) L = ext[dir];
      if (dir == 1)
        grid = dim3((ext[0] + TW - 1) / TW, (ext[1] + L - 1) / L, ext[2]);
      else
        grid = dim3((ext[0] + TW - 1) / TW, ext[1], (ext[2] + L - 1) / L);
    }
    const size_t lds = (size_t)TW * (L + 1) * NV * sizeof(R);
    const bool acc = !(first_writer && dir == 0);
    prof_mark(P_FLUX_I + dir, st, true);
#define OSTC_SWEEP(D, I, V, A) \
  hipLaunchKernelGGL((sweep_kernel<R, NS, NT, D, I, V, A>), grid, dim3(256), lds, st, z, dp, L, TW)
#define OSTC_SWEEP_D(D)                               \
  do {                                                \
    if (inv && visc) {                                \
      if (acc) OSTC_SWEEP(D, true, true, true);       \
      else OSTC_SWEEP(D, true, true, false);          \
    } else if (inv) {                                 \
      if (acc) OSTC_SWEEP(D, true, false, true);      \
      else OSTC_SWEEP(D, true, false, false);         \
    } else {                                          \
      if (acc) OSTC_SWEEP(D, false, true, true);      \
      else OSTC_SWEEP(D, false, true, false);         \
    }                                                 \
  } while (0)
    if (dir == 0) OSTC_SWEEP_D(0);
    else if (dir == 1) OSTC_SWEEP_D(1);
    else OSTC_SWEEP_D(2);
#undef OSTC_SWEEP_D
#undef OSTC_SWEEP
    prof_mark(P_FLUX_I + dir, st, false);
    if (hipGetLastError() != hipSuccess) return OSTC_ERR_HIP;
  }
  return OSTC_OK;
}

template <typename R>
int flux_sweeps(const ZoneView<R> &z, const ostc_param *dp, int ns, int nt, bool inv, bool visc, bool first_writer,
                hipStream_t st) {
  if (ns == 0 && nt == 0) return launch_sweeps<R, 0, 0>(z, dp, inv, visc, first_writer, st);
  if (ns == 0 && nt == 2) return launch_sweeps<R, 0, 2>(z, dp, inv, visc, first_writer, st);
  if (ns == 9 && nt == 0) return launch_sweeps<R, 9, 0>(z, dp, inv, visc, first_writer, st);
  if (ns == 9 && nt == 2) return launch_sweeps<R, 9, 2>(z, dp, inv, visc, first_writer, st);
  return OSTC_ERR_ARG;
}

template int flux_sweeps<float>(const ZoneView<float> &, const ostc_param *, int, int, bool, bool, bool, hipStream_t);
template int flux_sweeps<double>(const ZoneView<double> &, const ostc_param *, int, int, bool, bool, bool,
                                 hipStream_t);

}  // namespace ostc
