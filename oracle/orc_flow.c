/* Flow-field stages of the per-iteration path.  TEST INFRASTRUCTURE ONLY. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define NSMAX ORC_MAX_SPEC
#define NVMAX (7 + ORC_MAX_SPEC + 4)

/* gxl::sgn: src/gxl_lib/Math.hpp:9-15 (sgn(0) = +1) */
static inline int sgn(double a) { return a < 0 ? -1 : 1; }
/* minmod limiter: src/Limiter.cuh:10-13 (ids 0 and 1 share this body) */
static inline double minmod(double a, double b) { return 0.5 * (sgn(a) + sgn(b)) * dmin(fabs(a), fabs(b)); }

/* ---------------------------------------------------------------- setup */

/* compute_total_energy: src/FieldOperation.cuh:15-35 */
static void total_energy(orc_zone *z, const orc_param *p, int i, int j, int k) {
  const double u = F(bv, i, j, k, 1), v = F(bv, i, j, k, 2), w = F(bv, i, j, k, 3);
  double vel = u * u + v * v + w * w;
  S(vel, i, j, k) = vel;
  F(cv, i, j, k, 4) = 0.5 * F(bv, i, j, k, 0) * vel;
  if (p->mix_model != ORC_AIR) {
    double h[NSMAX];
    orc_enthalpy(F(bv, i, j, k, 5), h, p);
    for (int l = 0; l < p->n_spec; ++l) F(cv, i, j, k, 4) += h[l] * F(cv, i, j, k, 5 + l);
    F(cv, i, j, k, 4) -= F(bv, i, j, k, 4);
  } else {
    F(cv, i, j, k, 4) += F(bv, i, j, k, 4) / (ORC_GAMMA_AIR - 1);
  }
  S(vel, i, j, k) = sqrt(S(vel, i, j, k));
}

/* compute_cv_from_bv: src/FieldOperation.cuh:37-73 (all points incl. ghosts) */
void orc_compute_cv_from_bv(orc_zone *z, const orc_param *p) {
  const int g = z->ngg;
  for (int k = -g; k < z->mz + g; ++k)
    for (int j = -g; j < z->my + g; ++j)
      for (int i = -g; i < z->mx + g; ++i) {
        const double rho = F(bv, i, j, k, 0);
        F(cv, i, j, k, 0) = rho;
        F(cv, i, j, k, 1) = rho * F(bv, i, j, k, 1);
        F(cv, i, j, k, 2) = rho * F(bv, i, j, k, 2);
        F(cv, i, j, k, 3) = rho * F(bv, i, j, k, 3);
        for (int l = 0; l < p->n_scalar; ++l) F(cv, i, j, k, 5 + l) = rho * F(sv, i, j, k, l);
        total_energy(z, p, i, j, k);
      }
}

/* update_physical_properties: src/FieldOperation.cuh:75-109 (interior + 1) */
void orc_update_physical_properties(orc_zone *z, const orc_param *p) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = -1; k < z->mz + 1; ++k)
    for (int j = -1; j < z->my + 1; ++j)
      for (int i = -1; i < z->mx + 1; ++i) {
        const double t = F(bv, i, j, k, 5);
        if (p->mix_model != ORC_AIR) {
          const int ns = p->n_spec;
          double mw = 0, cp_tot = 0, cvv = 0, cp[NSMAX], y[NSMAX];
          orc_cp(t, cp, p);
          for (int l = 0; l < ns; ++l) {
            y[l] = F(sv, i, j, k, l);
            mw += y[l] / p->mw[l];
            cp_tot += y[l] * cp[l];
            cvv += y[l] * (cp[l] - ORC_RU / p->mw[l]);
          }
          mw = 1 / mw;
          S(cp, i, j, k) = cp_tot;
          S(gamma, i, j, k) = cp_tot / cvv;
          S(acoustic_speed, i, j, k) = sqrt(S(gamma, i, j, k) * ORC_RU * t / mw);
          double mu, lam, rhoD[NSMAX];
          orc_transport_point(p, t, mw, y, cp, &mu, &lam, rhoD);
          S(mul, i, j, k) = mu;
          S(thermal_conductivity, i, j, k) = lam;
          for (int l = 0; l < ns; ++l) F(rho_D, i, j, k, l) = rhoD[l];
        } else {
          const double c_temp = ORC_GAMMA_AIR * ORC_RU / ORC_MW_AIR;
          const double cp = ORC_GAMMA_AIR * ORC_RU / ORC_MW_AIR / (ORC_GAMMA_AIR - 1);
          S(acoustic_speed, i, j, k) = sqrt(c_temp * t);
          S(mul, i, j, k) = orc_sutherland(t);
          S(thermal_conductivity, i, j, k) = S(mul, i, j, k) * cp / p->Pr;
        }
        S(mach, i, j, k) = S(vel, i, j, k) / S(acoustic_speed, i, j, k);
      }
}

/* SST::compute_mut: src/SST.cu:6-58 (rho*k read from cv) */
static void sst_compute_mut(orc_zone *z, int i, int j, int k, double mul, const orc_param *p) {
  const double xi_x = M(i, j, k, 1, 1), xi_y = M(i, j, k, 1, 2), xi_z = M(i, j, k, 1, 3);
  const double eta_x = M(i, j, k, 2, 1), eta_y = M(i, j, k, 2, 2), eta_z = M(i, j, k, 2, 3);
  const double zeta_x = M(i, j, k, 3, 1), zeta_y = M(i, j, k, 3, 2), zeta_z = M(i, j, k, 3, 3);
#define D1(l) (F(bv, i + 1, j, k, l) - F(bv, i - 1, j, k, l))
#define D2(l) (F(bv, i, j + 1, k, l) - F(bv, i, j - 1, k, l))
#define D3(l) (F(bv, i, j, k + 1, l) - F(bv, i, j, k - 1, l))
  const double u_y = 0.5 * (xi_y * D1(1) + eta_y * D2(1) + zeta_y * D3(1));
  const double u_z = 0.5 * (xi_z * D1(1) + eta_z * D2(1) + zeta_z * D3(1));
  const double v_x = 0.5 * (xi_x * D1(2) + eta_x * D2(2) + zeta_x * D3(2));
  const double v_z = 0.5 * (xi_z * D1(2) + eta_z * D2(2) + zeta_z * D3(2));
  const double w_x = 0.5 * (xi_x * D1(3) + eta_x * D2(3) + zeta_x * D3(3));
  const double w_y = 0.5 * (xi_y * D1(3) + eta_y * D2(3) + zeta_y * D3(3));
#undef D1
#undef D2
#undef D3
  const int ns = p->n_spec;
  const double rhoK = F(cv, i, j, k, ns + 5);
  const double tke = F(sv, i, j, k, ns);
  const double omega = F(sv, i, j, k, ns + 1);
  const double vorticity = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
  const double density = F(bv, i, j, k, 0);
  double f2 = 1;
  const double dy = S(wall_distance, i, j, k);
  if (dy > 1e-25) {
    const double param1 = 2 * sqrt(tke) / (0.09 * omega * dy);
    const double param2 = 500 * mul / (density * dy * dy * omega);
    const double arg2 = dmax(param1, param2);
    f2 = tanh(arg2 * arg2);
  }
  double mut = 0;
  const double denominator = dmax(SST_A_1 * omega, vorticity * f2);
  if (denominator > 1e-25) mut = SST_A_1 * rhoK / denominator;
  S(mut, i, j, k) = mut;
}

/* initialize_mut: src/FieldOperation.cuh:111-137 (interior + 1) */
void orc_initialize_mut(orc_zone *z, const orc_param *p) {
  if (p->turb_method != ORC_RANS) return;
  for (int k = -1; k < z->mz + 1; ++k)
    for (int j = -1; j < z->my + 1; ++j)
      for (int i = -1; i < z->mx + 1; ++i) {
        const double t = F(bv, i, j, k, 5);
        double mul = orc_sutherland(t);
        if (p->mix_model != ORC_AIR) {
          double mw = 0, y[NSMAX];
          for (int l = 0; l < p->n_spec; ++l) {
            y[l] = F(sv, i, j, k, l);
            mw += y[l] / p->mw[l];
          }
          mw = 1 / mw;
          mul = orc_mixture_viscosity(p, t, mw, y);
        }
        sst_compute_mut(z, i, j, k, mul, p);
      }
}

/* ---------------------------------------------------------------- source */

/* SST::compute_source_and_mut: src/SST.cu:60-178 */
static void sst_source(orc_zone *z, int i, int j, int k, const orc_param *p) {
  const int ns = p->n_spec;
  const double xi_x = M(i, j, k, 1, 1), xi_y = M(i, j, k, 1, 2), xi_z = M(i, j, k, 1, 3);
  const double eta_x = M(i, j, k, 2, 1), eta_y = M(i, j, k, 2, 2), eta_z = M(i, j, k, 2, 3);
  const double zeta_x = M(i, j, k, 3, 1), zeta_y = M(i, j, k, 3, 2), zeta_z = M(i, j, k, 3, 3);
#define B1(l) (F(bv, i + 1, j, k, l) - F(bv, i - 1, j, k, l))
#define B2(l) (F(bv, i, j + 1, k, l) - F(bv, i, j - 1, k, l))
#define B3(l) (F(bv, i, j, k + 1, l) - F(bv, i, j, k - 1, l))
#define V1(l) (F(sv, i + 1, j, k, l) - F(sv, i - 1, j, k, l))
#define V2(l) (F(sv, i, j + 1, k, l) - F(sv, i, j - 1, k, l))
#define V3(l) (F(sv, i, j, k + 1, l) - F(sv, i, j, k - 1, l))
  const double u_x = 0.5 * (xi_x * B1(1) + eta_x * B2(1) + zeta_x * B3(1));
  const double u_y = 0.5 * (xi_y * B1(1) + eta_y * B2(1) + zeta_y * B3(1));
  const double u_z = 0.5 * (xi_z * B1(1) + eta_z * B2(1) + zeta_z * B3(1));
  const double v_x = 0.5 * (xi_x * B1(2) + eta_x * B2(2) + zeta_x * B3(2));
  const double v_y = 0.5 * (xi_y * B1(2) + eta_y * B2(2) + zeta_y * B3(2));
  const double v_z = 0.5 * (xi_z * B1(2) + eta_z * B2(2) + zeta_z * B3(2));
  const double w_x = 0.5 * (xi_x * B1(3) + eta_x * B2(3) + zeta_x * B3(3));
  const double w_y = 0.5 * (xi_y * B1(3) + eta_y * B2(3) + zeta_y * B3(3));
  const double w_z = 0.5 * (xi_z * B1(3) + eta_z * B2(3) + zeta_z * B3(3));
  const double density = F(bv, i, j, k, 0);
  const double omega = F(sv, i, j, k, ns + 1);
  const double k_x = 0.5 * (xi_x * V1(ns) + eta_x * V2(ns) + zeta_x * V3(ns));
  const double k_y = 0.5 * (xi_y * V1(ns) + eta_y * V2(ns) + zeta_y * V3(ns));
  const double k_z = 0.5 * (xi_z * V1(ns) + eta_z * V2(ns) + zeta_z * V3(ns));
  const double omega_x = 0.5 * (xi_x * V1(ns + 1) + eta_x * V2(ns + 1) + zeta_x * V3(ns + 1));
  const double omega_y = 0.5 * (xi_y * V1(ns + 1) + eta_y * V2(ns + 1) + zeta_y * V3(ns + 1));
  const double omega_z = 0.5 * (xi_z * V1(ns + 1) + eta_z * V2(ns + 1) + zeta_z * V3(ns + 1));
#undef B1
#undef B2
#undef B3
#undef V1
#undef V2
#undef V3
  const double inter_var = 2 * density * SST_SIGMA_OMEGA2 / omega * (k_x * omega_x + k_y * omega_y + k_z * omega_z);
  const double tke = F(sv, i, j, k, ns);
  const double rhoK = density * tke;
  const double vorticity = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
  double f1 = 1, f2 = 1;
  const double dy = S(wall_distance, i, j, k);
  if (dy > 1e-25) {
    const double param1 = sqrt(tke) / (0.09 * omega * dy);
    const double d2 = dy * dy;
    const double param2 = 500 * S(mul, i, j, k) / (density * d2 * omega);
    const double arg2 = dmax(2 * param1, param2);
    f2 = tanh(arg2 * arg2);
    const double CDkomega = dmax(1e-20, inter_var);
    const double param3 = 4 * density * SST_SIGMA_OMEGA2 * tke / (CDkomega * d2);
    const double arg1 = dmin(dmax(param1, param2), param3);
    f1 = tanh(arg1 * arg1 * arg1 * arg1);
  }
  double mut = 0;
  const double denominator = dmax(SST_A_1 * omega, vorticity * f2);
  if (denominator > 1e-25) mut = SST_A_1 * rhoK / denominator;
  S(mut, i, j, k) = mut;
  const double beta = SST_BETA_2 + SST_DELTA_BETA * f1;
  if (mut > 1e-25) {
    const double divU = u_x + v_y + w_z;
    const double divU2 = divU * divU;
    const double prod_k = mut * (2 * (u_x * u_x + v_y * v_y + w_z * w_z) - 2.0 / 3 * divU2 + (u_y + v_x) * (u_y + v_x) +
                                 (u_z + w_x) * (u_z + w_x) + (v_z + w_y) * (v_z + w_y)) -
                          2.0 / 3 * rhoK * divU;
    const double diss_k = SST_BETA_STAR * rhoK * omega;
    const double jac = S(jac, i, j, k);
    F(dq, i, j, k, ns + 5) += jac * (prod_k - diss_k);
    const double gamma = SST_GAMMA2 + SST_DELTA_GAMMA * f1;
    const double prod_omega = gamma * density / mut * prod_k + (1 - f1) * inter_var;
    const double diss_omega = beta * density * omega * omega;
    F(dq, i, j, k, ns + 6) += jac * (prod_omega - diss_omega);
  }
  if (p->turb_implicit == 1) {
    F(turb_src_jac, i, j, k, 0) = -2 * SST_BETA_STAR * omega;
    F(turb_src_jac, i, j, k, 1) = -2 * beta * omega;
  }
}

/* compute_source: src/SourceTerm.cuh:11-40 */
void orc_compute_source(orc_zone *z, const orc_param *p) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        if (p->turb_method == ORC_RANS) {
          sst_source(z, i, j, k, p);
          if (p->mix_model != ORC_AIR) {
            S(turb_therm_cond, i, j, k) = S(mut, i, j, k) * S(cp, i, j, k) / p->Prt;
          } else {
            const double cp = ORC_GAMMA_AIR * ORC_RU / ORC_MW_AIR / (ORC_GAMMA_AIR - 1);
            S(turb_therm_cond, i, j, k) = S(mut, i, j, k) * cp / p->Prt;
          }
        }
        if (p->mix_model == ORC_FR) orc_finite_rate_chemistry(z, i, j, k, p);
      }
}

/* ---------------------------------------------------------------- inviscid */

/* reconstruction: src/InviscidScheme.cuh:21-68 with MUSCL/NND2/first order
 * (src/Reconstruction.cu:5-46).  pv holds 4 consecutive points (i-1..i+2) of
 * n_var primitives; outputs pv_l/pv_r of n_var+2 (E at n_var, gamma at
 * n_var+1 for mixtures). */
static void reconstruct(const double *pv, double *pv_l, double *pv_r, const orc_param *p) {
  const int nv = p->n_var;
  const double *p0 = pv, *p1 = pv + nv, *p2 = pv + 2 * nv, *p3 = pv + 3 * nv;
  switch (p->reconstruction) {
    case 2: {
      const double kappa = 1.0 / 3.0;
      for (int l = 0; l < nv; ++l) {
        const double delta_i = p1[l] - p0[l];
        const double delta_i1 = p2[l] - p1[l];
        const double delta_i2 = p3[l] - p2[l];
        const double delta_neg_l = minmod(delta_i, delta_i1);
        const double delta_pos_l = minmod(delta_i1, delta_i);
        const double delta_neg_r = minmod(delta_i1, delta_i2);
        const double delta_pos_r = minmod(delta_i2, delta_i1);
        pv_l[l] = p1[l] + 0.25 * ((1 - kappa) * delta_neg_l + (1 + kappa) * delta_pos_l);
        pv_r[l] = p2[l] - 0.25 * ((1 - kappa) * delta_pos_r + (1 + kappa) * delta_neg_r);
      }
      break;
    }
    case 3:
      for (int l = 0; l < nv; ++l) {
        const double delta_i = p1[l] - p0[l];
        const double delta_i1 = p2[l] - p1[l];
        const double delta_i2 = p3[l] - p2[l];
        pv_l[l] = p1[l] + 0.5 * minmod(delta_i, delta_i1);
        pv_r[l] = p2[l] - 0.5 * minmod(delta_i1, delta_i2);
      }
      break;
    default:
      for (int l = 0; l < nv; ++l) {
        pv_l[l] = p1[l];
        pv_r[l] = p2[l];
      }
  }
  if (p->mix_model != ORC_AIR) {
    double el = 0.5 * (pv_l[1] * pv_l[1] + pv_l[2] * pv_l[2] + pv_l[3] * pv_l[3]);
    double er = 0.5 * (pv_r[1] * pv_r[1] + pv_r[2] * pv_r[2] + pv_r[3] * pv_r[3]);
    const int ns = p->n_spec;
    double mw_inv_l = 0, mw_inv_r = 0;
    for (int l = 0; l < ns; ++l) {
      mw_inv_l += pv_l[5 + l] / p->mw[l];
      mw_inv_r += pv_r[5 + l] / p->mw[l];
    }
    const double t_l = pv_l[4] / (pv_l[0] * ORC_RU * mw_inv_l);
    const double t_r = pv_r[4] / (pv_r[0] * ORC_RU * mw_inv_r);
    double hl[NSMAX], hr[NSMAX], cpl_i[NSMAX], cpr_i[NSMAX];
    orc_enthalpy_and_cp(t_l, hl, cpl_i, p);
    orc_enthalpy_and_cp(t_r, hr, cpr_i, p);
    double cpl = 0, cpr = 0, cvl = 0, cvr = 0;
    for (int l = 0; l < ns; ++l) {
      cpl += cpl_i[l] * pv_l[l + 5];
      cpr += cpr_i[l] * pv_r[l + 5];
      cvl += pv_l[l + 5] * (cpl_i[l] - ORC_RU / p->mw[l]);
      cvr += pv_r[l + 5] * (cpr_i[l] - ORC_RU / p->mw[l]);
      el += hl[l] * pv_l[l + 5];
      er += hr[l] * pv_r[l + 5];
    }
    pv_l[nv] = pv_l[0] * el - pv_l[4];
    pv_r[nv] = pv_r[0] * er - pv_r[4];
    pv_l[nv + 1] = cpl / cvl;
    pv_r[nv + 1] = cpr / cvr;
  } else {
    const double el = 0.5 * (pv_l[1] * pv_l[1] + pv_l[2] * pv_l[2] + pv_l[3] * pv_l[3]);
    const double er = 0.5 * (pv_r[1] * pv_r[1] + pv_r[2] * pv_r[2] + pv_r[3] * pv_r[3]);
    pv_l[nv] = el * pv_l[0] + pv_l[4] / (ORC_GAMMA_AIR - 1);
    pv_r[nv] = er * pv_r[0] + pv_r[4] / (ORC_GAMMA_AIR - 1);
  }
}

/* AUSMP_compute_inviscid_flux: src/InviscidScheme.cuh:70-149.  ml/mr are the
 * 3 metric entries of row (dir+1) at the face's left and right points. */
static void ausm_plus(const double *pv, const double *ml, const double *mr, double jac_l, double jac_r,
                      const orc_param *p, double *fc) {
  double pv_l[NVMAX], pv_r[NVMAX];
  const int nv = p->n_var;
  reconstruct(pv, pv_l, pv_r, p);
  const double k1 = 0.5 * (jac_l * ml[0] + jac_r * mr[0]);
  const double k2 = 0.5 * (jac_l * ml[1] + jac_r * mr[1]);
  const double k3 = 0.5 * (jac_l * ml[2] + jac_r * mr[2]);
  const double grad_k_div_jac = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pv_l[1] + k2 * pv_l[2] + k3 * pv_l[3]) / grad_k_div_jac;
  const double ur = (k1 * pv_r[1] + k2 * pv_r[2] + k3 * pv_r[3]) / grad_k_div_jac;
  const double pl = pv_l[4], pr = pv_r[4], rho_l = pv_l[0], rho_r = pv_r[0];
  double gam_l = ORC_GAMMA_AIR, gam_r = ORC_GAMMA_AIR;
  if (p->n_spec > 0) {
    gam_l = pv_l[nv + 1];
    gam_r = pv_r[nv + 1];
  }
  const double c = 0.5 * (sqrt(gam_l * pl / rho_l) + sqrt(gam_r * pr / rho_r));
  const double mach_l = ul / c, mach_r = ur / c;
  double mlp, mrn, plp, prn;
  const double alpha = 3 / 16.0;
  if (fabs(mach_l) > 1) {
    mlp = 0.5 * (mach_l + fabs(mach_l));
    plp = mlp / mach_l;
  } else {
    const double a = (mach_l + 1) * (mach_l + 1) * 0.25;
    const double b = (mach_l * mach_l - 1) * (mach_l * mach_l - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - mach_l) + alpha * mach_l * b;
  }
  if (fabs(mach_r) > 1) {
    mrn = 0.5 * (mach_r - fabs(mach_r));
    prn = mrn / mach_r;
  } else {
    const double a = (mach_r - 1) * (mach_r - 1) * 0.25;
    const double b = (mach_r * mach_r - 1) * (mach_r * mach_r - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mach_r) - alpha * mach_r * b;
  }
  const double p_coeff = plp * pl + prn * pr;
  const double m_half = mlp + mrn;
  const double mach_pos = 0.5 * (m_half + fabs(m_half));
  const double mach_neg = 0.5 * (m_half - fabs(m_half));
  const double mass_flux_half = c * (rho_l * mach_pos + rho_r * mach_neg);
  const double coeff = mass_flux_half * grad_k_div_jac;
  const double *s = mass_flux_half >= 0 ? pv_l : pv_r;
  fc[0] = coeff;
  fc[1] = coeff * s[1] + p_coeff * k1;
  fc[2] = coeff * s[2] + p_coeff * k2;
  fc[3] = coeff * s[3] + p_coeff * k3;
  fc[4] = coeff * (s[nv] + s[4]) / s[0];
  for (int l = 5; l < nv; ++l) fc[l] = coeff * s[l];
}

/* One AUSM+ face flux from 4 points of n_var primitives (exposed for tests). */
void orc_ausm_face(const orc_param *p, const double *pv, const double *ml, const double *mr, double jl, double jr,
                   double *fc) {
  ausm_plus(pv, ml, mr, jl, jr, p, fc);
}

/* inviscid_flux_1d for one direction (src/SchemeSelector.cuh:70-157):
 * face fluxes F(c+1/2), c = -1..m-1, then dq(c) -= F(c+1/2) - F(c-1/2). */
void orc_inviscid_flux_dir(orc_zone *z, const orc_param *p, int dir) {
  const int ext[3] = {z->mx, z->my, z->mz};
  const int nv = p->n_var;
  const int m = ext[dir];
  const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
#pragma omp parallel
  {
    double *fc = (double *)malloc(sizeof(double) * (size_t)(m + 1) * nv);
    double pv[4 * NVMAX];
#pragma omp for collapse(2) schedule(static)
    for (int b = 0; b < ext[o2]; ++b)
      for (int a = 0; a < ext[o1]; ++a) {
        for (int f = -1; f < m; ++f) {
          int idx[3];
          idx[o1] = a;
          idx[o2] = b;
          double ml[3] = {0, 0, 0}, mr[3] = {0, 0, 0}, jl = 0, jr = 0;
          for (int q = 0; q < 4; ++q) {
            idx[dir] = f - 1 + q;
            const int i = idx[0], j = idx[1], k = idx[2];
            for (int l = 0; l < 5; ++l) pv[q * nv + l] = F(bv, i, j, k, l);
            for (int l = 0; l < p->n_scalar; ++l) pv[q * nv + 5 + l] = F(sv, i, j, k, l);
            if (q == 1) {
              for (int c = 1; c < 4; ++c) ml[c - 1] = M(i, j, k, dir + 1, c);
              jl = S(jac, i, j, k);
            } else if (q == 2) {
              for (int c = 1; c < 4; ++c) mr[c - 1] = M(i, j, k, dir + 1, c);
              jr = S(jac, i, j, k);
            }
          }
          ausm_plus(pv, ml, mr, jl, jr, p, &fc[(f + 1) * nv]);
        }
        for (int c = 0; c < m; ++c) {
          int idx[3];
          idx[o1] = a;
          idx[o2] = b;
          idx[dir] = c;
          for (int l = 0; l < nv; ++l) F(dq, idx[0], idx[1], idx[2], l) -= fc[(c + 1) * nv + l] - fc[c * nv + l];
        }
      }
    free(fc);
  }
}

/* compute_inviscid_flux: src/SchemeSelector.cuh:36-68 (i, j, then k sweep) */
void orc_inviscid_flux(orc_zone *z, const orc_param *p) {
  orc_inviscid_flux_dir(z, p, 0);
  orc_inviscid_flux_dir(z, p, 1);
  if (z->mz > 1) orc_inviscid_flux_dir(z, p, 2);
}

/* ---------------------------------------------------------------- viscous */

/* compute_fv/gv/hv_2nd_order: src/ViscousScheme.cuh:22-615.  Face between
 * (i,j,k) and its +dir neighbour. */
static void viscous_face(orc_zone *z, const orc_param *p, int dir, int i, int j, int k, double *fv) {
  const int e[3] = {dir == 0, dir == 1, dir == 2};
  const int i1 = i + e[0], j1 = j + e[1], k1 = k + e[2];
  double mm[3][3];
  for (int r = 1; r <= 3; ++r)
    for (int c = 1; c <= 3; ++c) mm[r - 1][c - 1] = 0.5 * (M(i, j, k, r, c) + M(i1, j1, k1, r, c));
  const double xi_x = mm[0][0], xi_y = mm[0][1], xi_z = mm[0][2];
  const double eta_x = mm[1][0], eta_y = mm[1][1], eta_z = mm[1][2];
  const double zeta_x = mm[2][0], zeta_y = mm[2][1], zeta_z = mm[2][2];

  /* derivative of field arr, variable l, along computational direction d:
   * two-point along dir, four-point average of central differences across */
#define DER(arr, l, d, out)                                                                                       \
  do {                                                                                                            \
    if ((d) == dir) {                                                                                             \
      out = F(arr, i1, j1, k1, l) - F(arr, i, j, k, l);                                                           \
    } else {                                                                                                      \
      const int ed[3] = {(d) == 0, (d) == 1, (d) == 2};                                                           \
      out = 0.25 * (F(arr, i + ed[0], j + ed[1], k + ed[2], l) - F(arr, i - ed[0], j - ed[1], k - ed[2], l) +     \
                    F(arr, i1 + ed[0], j1 + ed[1], k1 + ed[2], l) - F(arr, i1 - ed[0], j1 - ed[1], k1 - ed[2], l)); \
    }                                                                                                             \
  } while (0)

  double u_xi, u_eta, u_zeta, v_xi, v_eta, v_zeta, w_xi, w_eta, w_zeta, t_xi, t_eta, t_zeta;
  DER(bv, 1, 0, u_xi);
  DER(bv, 1, 1, u_eta);
  DER(bv, 1, 2, u_zeta);
  DER(bv, 2, 0, v_xi);
  DER(bv, 2, 1, v_eta);
  DER(bv, 2, 2, v_zeta);
  DER(bv, 3, 0, w_xi);
  DER(bv, 3, 1, w_eta);
  DER(bv, 3, 2, w_zeta);
  DER(bv, 5, 0, t_xi);
  DER(bv, 5, 1, t_eta);
  DER(bv, 5, 2, t_zeta);

  const double u_x = u_xi * xi_x + u_eta * eta_x + u_zeta * zeta_x;
  const double u_y = u_xi * xi_y + u_eta * eta_y + u_zeta * zeta_y;
  const double u_z = u_xi * xi_z + u_eta * eta_z + u_zeta * zeta_z;
  const double v_x = v_xi * xi_x + v_eta * eta_x + v_zeta * zeta_x;
  const double v_y = v_xi * xi_y + v_eta * eta_y + v_zeta * zeta_y;
  const double v_z = v_xi * xi_z + v_eta * eta_z + v_zeta * zeta_z;
  const double w_x = w_xi * xi_x + w_eta * eta_x + w_zeta * zeta_x;
  const double w_y = w_xi * xi_y + w_eta * eta_y + w_zeta * zeta_y;
  const double w_z = w_xi * xi_z + w_eta * eta_z + w_zeta * zeta_z;

  const double mul = 0.5 * (S(mul, i, j, k) + S(mul, i1, j1, k1));
  double mut = 0;
  if (p->turb_method == ORC_RANS) mut = 0.5 * (S(mut, i, j, k) + S(mut, i1, j1, k1));
  const double viscosity = mul + mut;
  double tau_xx = viscosity * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  double tau_yy = viscosity * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  double tau_zz = viscosity * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double tau_xy = viscosity * (u_y + v_x);
  const double tau_xz = viscosity * (u_z + w_x);
  const double tau_yz = viscosity * (v_z + w_y);
  const int ns = p->n_spec;
  if (p->turb_method == ORC_RANS && p->rans_model == 2) {
    const double tt = -2.0 / 3 * 0.5 * (F(bv, i, j, k, 0) * F(sv, i, j, k, ns) + F(bv, i1, j1, k1, 0) * F(sv, i1, j1, k1, ns));
    tau_xx += tt;
    tau_yy += tt;
    tau_zz += tt;
  }
  const double kx = 0.5 * (M(i, j, k, dir + 1, 1) * S(jac, i, j, k) + M(i1, j1, k1, dir + 1, 1) * S(jac, i1, j1, k1));
  const double ky = 0.5 * (M(i, j, k, dir + 1, 2) * S(jac, i, j, k) + M(i1, j1, k1, dir + 1, 2) * S(jac, i1, j1, k1));
  const double kz = 0.5 * (M(i, j, k, dir + 1, 3) * S(jac, i, j, k) + M(i1, j1, k1, dir + 1, 3) * S(jac, i1, j1, k1));
  fv[0] = 0;
  fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  const double um = 0.5 * (F(bv, i, j, k, 1) + F(bv, i1, j1, k1, 1));
  const double vm = 0.5 * (F(bv, i, j, k, 2) + F(bv, i1, j1, k1, 2));
  const double wm = 0.5 * (F(bv, i, j, k, 3) + F(bv, i1, j1, k1, 3));
  const double t_x = t_xi * xi_x + t_eta * eta_x + t_zeta * zeta_x;
  const double t_y = t_xi * xi_y + t_eta * eta_y + t_zeta * zeta_y;
  const double t_z = t_xi * xi_z + t_eta * eta_z + t_zeta * zeta_z;
  double conductivity = 0.5 * (S(thermal_conductivity, i, j, k) + S(thermal_conductivity, i1, j1, k1));
  if (p->turb_method == ORC_RANS) conductivity += 0.5 * (S(turb_therm_cond, i, j, k) + S(turb_therm_cond, i1, j1, k1));
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + conductivity * (kx * t_x + ky * t_y + kz * t_z);

  if (p->mix_model != ORC_AIR) {
    double turb_diffusivity = 0;
    if (p->turb_method == ORC_RANS) turb_diffusivity = mut / p->Sct;
    double h[NSMAX], diffusivity[NSMAX], y_x[NSMAX], y_y[NSMAX], y_z[NSMAX];
    const double tm = 0.5 * (F(bv, i, j, k, 5) + F(bv, i1, j1, k1, 5));
    orc_enthalpy(tm, h, p);
    double rho_uc = 0, rho_vc = 0, rho_wc = 0;
    for (int l = 0; l < ns; ++l) {
      diffusivity[l] = 0.5 * (F(rho_D, i, j, k, l) + F(rho_D, i1, j1, k1, l)) + turb_diffusivity;
      double y_xi, y_eta, y_zeta;
      DER(sv, l, 0, y_xi);
      DER(sv, l, 1, y_eta);
      DER(sv, l, 2, y_zeta);
      y_x[l] = y_xi * xi_x + y_eta * eta_x + y_zeta * zeta_x;
      y_y[l] = y_xi * xi_y + y_eta * eta_y + y_zeta * zeta_y;
      y_z[l] = y_xi * xi_z + y_eta * eta_z + y_zeta * zeta_z;
      const double DMulH = diffusivity[l] * h[l];
      fv[4] += kx * DMulH * y_x[l];
      fv[4] += ky * DMulH * y_y[l];
      fv[4] += kz * DMulH * y_z[l];
      rho_uc += diffusivity[l] * y_x[l];
      rho_vc += diffusivity[l] * y_y[l];
      rho_wc += diffusivity[l] * y_z[l];
    }
    for (int l = 0; l < ns; ++l)
      fv[5 + l] = diffusivity[l] * (kx * y_x[l] + ky * y_y[l] + kz * y_z[l]) -
                  0.5 * (F(sv, i, j, k, l) + F(sv, i1, j1, k1, l)) * (kx * rho_uc + ky * rho_vc + kz * rho_wc);
  }

  if (p->turb_method == ORC_RANS && p->rans_model != 1) {
    const int it = ns;
    double k_xi, k_eta, k_zeta, o_xi, o_eta, o_zeta;
    DER(sv, it, 0, k_xi);
    DER(sv, it, 1, k_eta);
    DER(sv, it, 2, k_zeta);
    DER(sv, it + 1, 0, o_xi);
    DER(sv, it + 1, 1, o_eta);
    DER(sv, it + 1, 2, o_zeta);
    const double k_x = k_xi * xi_x + k_eta * eta_x + k_zeta * zeta_x;
    const double k_y = k_xi * xi_y + k_eta * eta_y + k_zeta * zeta_y;
    const double k_z = k_xi * xi_z + k_eta * eta_z + k_zeta * zeta_z;
    const double omega_x = o_xi * xi_x + o_eta * eta_x + o_zeta * zeta_x;
    const double omega_y = o_xi * xi_y + o_eta * eta_y + o_zeta * zeta_y;
    const double omega_z = o_xi * xi_z + o_eta * eta_z + o_zeta * zeta_z;
    const double wall_dist = 0.5 * (S(wall_distance, i, j, k) + S(wall_distance, i1, j1, k1));
    double f1 = 1;
    if (wall_dist > 1e-25) {
      const double km = 0.5 * (F(sv, i, j, k, it) + F(sv, i1, j1, k1, it));
      const double omegam = 0.5 * (F(sv, i, j, k, it + 1) + F(sv, i1, j1, k1, it + 1));
      const double param1 = sqrt(km) / (0.09 * omegam * wall_dist);
      const double rhom = 0.5 * (F(bv, i, j, k, 0) + F(bv, i1, j1, k1, 0));
      const double d2 = wall_dist * wall_dist;
      const double param2 = 500 * mul / (rhom * d2 * omegam);
      const double CDkomega =
          dmax(1e-20, 2 * rhom * SST_SIGMA_OMEGA2 / omegam * (k_x * omega_x + k_y * omega_y + k_z * omega_z));
      const double param3 = 4 * rhom * SST_SIGMA_OMEGA2 * km / (CDkomega * d2);
      const double arg1 = dmin(dmax(param1, param2), param3);
      f1 = tanh(arg1 * arg1 * arg1 * arg1);
    }
    const double sigma_k = SST_SIGMA_K2 + SST_DELTA_SIGMA_K * f1;
    const double sigma_omega = SST_SIGMA_OMEGA2 + SST_DELTA_SIGMA_OMEGA * f1;
    fv[5 + ns] = (mul + mut * sigma_k) * (kx * k_x + ky * k_y + kz * k_z);
    fv[6 + ns] = (mul + mut * sigma_omega) * (kx * omega_x + ky * omega_y + kz * omega_z);
  }
#undef DER
}

/* compute_viscous_flux: src/SchemeSelector.cuh:172-280.  With
 * viscous_order == 0 the reference adds uninitialised shared memory; here the
 * viscous flux is zero (quirk A-1). */
void orc_viscous_flux(orc_zone *z, const orc_param *p) {
  if (p->viscous_scheme == 0) return;
  const int ext[3] = {z->mx, z->my, z->mz};
  const int nv = p->n_var;
  for (int dir = 0; dir < 3; ++dir) {
    if (dir == 2 && z->mz == 1) break;
    const int m = ext[dir];
    const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
#pragma omp parallel
    {
      double *fv = (double *)calloc((size_t)(m + 1) * nv, sizeof(double));
#pragma omp for collapse(2) schedule(static)
      for (int b = 0; b < ext[o2]; ++b)
        for (int a = 0; a < ext[o1]; ++a) {
          int idx[3];
          idx[o1] = a;
          idx[o2] = b;
          for (int f = -1; f < m; ++f) {
            idx[dir] = f;
            viscous_face(z, p, dir, idx[0], idx[1], idx[2], &fv[(f + 1) * nv]);
          }
          for (int c = 0; c < m; ++c) {
            idx[dir] = c;
            for (int l = 0; l < nv; ++l) F(dq, idx[0], idx[1], idx[2], l) += fv[(c + 1) * nv + l] - fv[c * nv + l];
          }
        }
      free(fv);
    }
  }
}

/* ---------------------------------------------------------------- time step */

/* local_time_step: src/TimeAdvanceFunc.cuh:21-71 */
void orc_local_time_step(orc_zone *z, const orc_param *p) {
  const int dim = z->mz == 1 ? 2 : 3;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const double grad_xi = sqrt(M(i, j, k, 1, 1) * M(i, j, k, 1, 1) + M(i, j, k, 1, 2) * M(i, j, k, 1, 2) +
                                    M(i, j, k, 1, 3) * M(i, j, k, 1, 3));
        const double grad_eta = sqrt(M(i, j, k, 2, 1) * M(i, j, k, 2, 1) + M(i, j, k, 2, 2) * M(i, j, k, 2, 2) +
                                     M(i, j, k, 2, 3) * M(i, j, k, 2, 3));
        const double grad_zeta = sqrt(M(i, j, k, 3, 1) * M(i, j, k, 3, 1) + M(i, j, k, 3, 2) * M(i, j, k, 3, 2) +
                                      M(i, j, k, 3, 3) * M(i, j, k, 3, 3));
        const double u = F(bv, i, j, k, 1), v = F(bv, i, j, k, 2), w = F(bv, i, j, k, 3);
        const double U = u * M(i, j, k, 1, 1) + v * M(i, j, k, 1, 2) + w * M(i, j, k, 1, 3);
        const double V = u * M(i, j, k, 2, 1) + v * M(i, j, k, 2, 2) + w * M(i, j, k, 2, 3);
        const double W = u * M(i, j, k, 3, 1) + v * M(i, j, k, 3, 2) + w * M(i, j, k, 3, 3);
        const double c = S(acoustic_speed, i, j, k);
        const double r0 = fabs(U) + c * grad_xi, r1 = fabs(V) + c * grad_eta;
        const double r2 = dim == 3 ? fabs(W) + c * grad_zeta : 0;
        F(inv_spectr_rad, i, j, k, 0) = r0;
        F(inv_spectr_rad, i, j, k, 1) = r1;
        F(inv_spectr_rad, i, j, k, 2) = r2;
        const double sri = r0 + r1 + r2;
        double gamma = ORC_GAMMA_AIR;
        if (p->mix_model != ORC_AIR) gamma = S(gamma, i, j, k);
        const double coeff_1 = dmax(gamma, 4.0 / 3.0) / F(bv, i, j, k, 0);
        double coeff_2 = S(mul, i, j, k) / p->Pr;
        if (p->turb_method == ORC_RANS) coeff_2 += S(mut, i, j, k) / p->Prt;
        double srv = grad_xi * grad_xi + grad_eta * grad_eta;
        if (dim == 3) srv += grad_zeta * grad_zeta;
        srv *= coeff_1 * coeff_2;
        if (z->visc_spectr_rad) S(visc_spectr_rad, i, j, k) = srv;
        S(dt_local, i, j, k) = p->cfl / (sri + srv);
      }
}

/* ---------------------------------------------------------------- implicit */

/* compute_jacobian_times_dq: src/DPLURHPP.cuh:78-125 (reads dq at the
 * neighbour (i,j,k) with that neighbour's metric row dir+1) */
static void jac_times_dq(const orc_zone *z, const orc_param *p, int i, int j, int k, int dir, double pm_sr,
                         double *out) {
  const double xi_x = M(i, j, k, dir + 1, 1), xi_y = M(i, j, k, dir + 1, 2), xi_z = M(i, j, k, dir + 1, 3);
  const double u = F(bv, i, j, k, 1), v = F(bv, i, j, k, 2), w = F(bv, i, j, k, 3);
  const double U = xi_x * u + xi_y * v + xi_z * w;
  const double lmd1 = U + pm_sr;
  const double e = 0.5 * S(vel, i, j, k) * S(vel, i, j, k);
  double gamma = ORC_GAMMA_AIR, b3 = 0, b4 = 0, h = 0;
  if (p->mix_model != ORC_AIR) {
    double enthalpy[NSMAX];
    const double t = F(bv, i, j, k, 5);
    orc_enthalpy(t, enthalpy, p);
    gamma = S(gamma, i, j, k);
    for (int l = 0; l < p->n_spec; ++l) {
      b3 += ORC_RU * t / p->mw[l] * F(dq, i, j, k, 5 + l);
      b4 += enthalpy[l] * F(dq, i, j, k, 5 + l);
      h += F(sv, i, j, k, l) * enthalpy[l];
    }
    b3 *= gamma;
    b4 *= gamma - 1;
    h += e;
  } else {
    h = gamma / (gamma - 1) * F(bv, i, j, k, 4) / F(bv, i, j, k, 0) + e;
  }
  const double b1 = xi_x * F(dq, i, j, k, 1) + xi_y * F(dq, i, j, k, 2) + xi_z * F(dq, i, j, k, 3) - U * F(dq, i, j, k, 0);
  const double b2 = (gamma - 1) * (e * F(dq, i, j, k, 0) - u * F(dq, i, j, k, 1) - v * F(dq, i, j, k, 2) -
                                   w * F(dq, i, j, k, 3) + F(dq, i, j, k, 4));
  out[0] = b1 + lmd1 * F(dq, i, j, k, 0);
  out[1] = u * b1 + xi_x * b2 + lmd1 * F(dq, i, j, k, 1) + xi_x * (b3 - b4);
  out[2] = v * b1 + xi_y * b2 + lmd1 * F(dq, i, j, k, 2) + xi_y * (b3 - b4);
  out[3] = w * b1 + xi_z * b2 + lmd1 * F(dq, i, j, k, 3) + xi_z * (b3 - b4);
  out[4] = h * b1 + U * b2 + lmd1 * F(dq, i, j, k, 4) + U * (b3 - b4);
  for (int l = 0; l < p->n_scalar; ++l) out[5 + l] = lmd1 * F(dq, i, j, k, l + 5) + F(sv, i, j, k, l) * b1;
}

/* compute_DQ_0: src/DPLURHPP.cuh:13-76 */
static void dq0_point(orc_zone *z, const orc_param *p, int i, int j, int k) {
  const double dt = S(dt_local, i, j, k);
  const double diag =
      1 + dt * (F(inv_spectr_rad, i, j, k, 0) + F(inv_spectr_rad, i, j, k, 1) + F(inv_spectr_rad, i, j, k, 2));
  const int ns = p->n_spec;
  if (p->mix_model != ORC_FR) {
    for (int l = 0; l < 5 + ns; ++l) F(dq, i, j, k, l) /= diag;
  } else {
    for (int l = 0; l < 5; ++l) F(dq, i, j, k, l) /= diag;
    if (p->chem_src_method == 1) {
      orc_epi_solve_dq(z, i, j, k, diag, p);
    } else if (p->chem_src_method == 2) {
      for (int l = 0; l < ns; ++l) F(dq, i, j, k, 5 + l) /= diag - dt * F(chem_src_jac, i, j, k, l);
    } else {
      for (int l = 0; l < ns; ++l) F(dq, i, j, k, l + 5) /= diag;
    }
  }
  if (p->turb_method == ORC_RANS) {
    if (p->turb_implicit == 1) {
      F(dq, i, j, k, ns + 5) /= diag - dt * F(turb_src_jac, i, j, k, 0);
      F(dq, i, j, k, ns + 6) /= diag - dt * F(turb_src_jac, i, j, k, 1);
    } else {
      F(dq, i, j, k, ns + 5) /= diag;
      F(dq, i, j, k, ns + 6) /= diag;
    }
  }
}

/* DPLUR_inner_iteration: src/DPLURHPP.cuh:127-249 */
static void dplur_sweep_point(orc_zone *z, const orc_param *p, int i, int j, int k) {
  double cj[NVMAX], dq_total[NVMAX];
  memset(dq_total, 0, sizeof(dq_total));
  const int nv = p->n_var;
  const int ext[3] = {z->mx, z->my, z->mz};
  if (i > 0) {
    jac_times_dq(z, p, i - 1, j, k, 0, F(inv_spectr_rad, i - 1, j, k, 0), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] += 0.5 * cj[l];
  }
  if (j > 0) {
    jac_times_dq(z, p, i, j - 1, k, 1, F(inv_spectr_rad, i, j - 1, k, 1), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] += 0.5 * cj[l];
  }
  if (k > 0) {
    jac_times_dq(z, p, i, j, k - 1, 2, F(inv_spectr_rad, i, j, k - 1, 2), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] += 0.5 * cj[l];
  }
  if (i != ext[0] - 1) {
    jac_times_dq(z, p, i + 1, j, k, 0, -F(inv_spectr_rad, i + 1, j, k, 0), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] -= 0.5 * cj[l];
  }
  if (j != ext[1] - 1) {
    jac_times_dq(z, p, i, j + 1, k, 1, -F(inv_spectr_rad, i, j + 1, k, 1), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] -= 0.5 * cj[l];
  }
  if (k != ext[2] - 1) {
    jac_times_dq(z, p, i, j, k + 1, 2, -F(inv_spectr_rad, i, j, k + 1, 2), cj);
    for (int l = 0; l < nv; ++l) dq_total[l] -= 0.5 * cj[l];
  }
  const double dt = S(dt_local, i, j, k);
  const double diag =
      1 + dt * (F(inv_spectr_rad, i, j, k, 0) + F(inv_spectr_rad, i, j, k, 1) + F(inv_spectr_rad, i, j, k, 2));
  const int ns = p->n_spec;
  if (p->mix_model != ORC_FR) {
    for (int l = 0; l < 5 + ns; ++l) F(dqk, i, j, k, l) = F(dq0, i, j, k, l) + dt * dq_total[l] / diag;
  } else {
    for (int l = 0; l < 5; ++l) F(dqk, i, j, k, l) = F(dq0, i, j, k, l) + dt * dq_total[l] / diag;
    if (p->chem_src_method == 1) {
      orc_epi_dqk(z, i, j, k, diag, dq_total, p);
    } else if (p->chem_src_method == 2) {
      for (int l = 0; l < ns; ++l)
        F(dqk, i, j, k, 5 + l) =
            F(dq0, i, j, k, 5 + l) + dt * dq_total[5 + l] / (diag - dt * F(chem_src_jac, i, j, k, l));
    } else {
      for (int l = 0; l < ns; ++l) F(dqk, i, j, k, 5 + l) = F(dq0, i, j, k, 5 + l) + dt * dq_total[5 + l] / diag;
    }
  }
  if (p->turb_method == ORC_RANS) {
    if (p->turb_implicit == 1) {
      F(dqk, i, j, k, ns + 5) =
          F(dq0, i, j, k, ns + 5) + dt * dq_total[5 + ns] / (diag - dt * F(turb_src_jac, i, j, k, 0));
      F(dqk, i, j, k, ns + 6) =
          F(dq0, i, j, k, ns + 6) + dt * dq_total[6 + ns] / (diag - dt * F(turb_src_jac, i, j, k, 1));
    } else {
      F(dqk, i, j, k, ns + 5) = F(dq0, i, j, k, ns + 5) + dt * dq_total[5 + ns] / diag;
      F(dqk, i, j, k, ns + 6) = F(dq0, i, j, k, ns + 6) + dt * dq_total[6 + ns] / diag;
    }
  }
}

/* implicit_treatment: src/ImplicitTreatmentHPP.cuh:10-59, DPLUR
 * src/DPLURHPP.cuh:251-273 (dq0 = copy of dq after DQ0; dq = dqk after each
 * sweep; only interior values are ever read, so copying the interior is
 * equivalent to the reference's whole-array D2D copies). */
void orc_implicit_treatment(orc_zone *z, const orc_param *p) {
  const int ns = p->n_spec;
  if (p->implicit_method == 0) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < z->mz; ++k)
      for (int j = 0; j < z->my; ++j)
        for (int i = 0; i < z->mx; ++i) {
          const double dt = S(dt_local, i, j, k);
          if (p->mix_model == ORC_FR && p->chem_src_method == 1) {
            orc_epi_solve_dq(z, i, j, k, 1.0, p);
          } else if (p->mix_model == ORC_FR && p->chem_src_method == 2) {
            for (int l = 0; l < ns; ++l) F(dq, i, j, k, 5 + l) /= 1 - dt * F(chem_src_jac, i, j, k, l);
          }
          if (p->turb_method == ORC_RANS && p->turb_implicit == 1) {
            F(dq, i, j, k, ns + 5) /= 1 - dt * F(turb_src_jac, i, j, k, 0);
            F(dq, i, j, k, ns + 6) /= 1 - dt * F(turb_src_jac, i, j, k, 1);
          }
        }
    return;
  }
  const int nv = p->n_var;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        dq0_point(z, p, i, j, k);
        for (int l = 0; l < nv; ++l) F(dq0, i, j, k, l) = F(dq, i, j, k, l);
      }
  for (int it = 0; it < p->dplur_inner_step; ++it) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < z->mz; ++k)
      for (int j = 0; j < z->my; ++j)
        for (int i = 0; i < z->mx; ++i) dplur_sweep_point(z, p, i, j, k);
    for (int k = 0; k < z->mz; ++k)
      for (int j = 0; j < z->my; ++j)
        for (int i = 0; i < z->mx; ++i)
          for (int l = 0; l < nv; ++l) F(dq, i, j, k, l) = F(dqk, i, j, k, l);
  }
}

/* ---------------------------------------------------------------- update */

/* compute_temperature: src/FieldOperation.cu:3-40 (Newton, stop
 * |1 - t1/t| < 1e-3, at most 1000 iterations) */
void orc_compute_temperature(orc_zone *z, int i, int j, int k, const orc_param *p, int *iters) {
  const int ns = p->n_spec;
  double y[NSMAX];
  for (int l = 0; l < ns; ++l) y[l] = F(sv, i, j, k, l);
  double mw = 0;
  for (int l = 0; l < ns; ++l) mw += y[l] / p->mw[l];
  mw = 1 / mw;
  const double gas_const = ORC_RU / mw;
  const double e = F(cv, i, j, k, 4) / F(cv, i, j, k, 0) -
                   0.5 * (F(bv, i, j, k, 1) * F(bv, i, j, k, 1) + F(bv, i, j, k, 2) * F(bv, i, j, k, 2) +
                          F(bv, i, j, k, 3) * F(bv, i, j, k, 3));
  double t;
  const int n = orc_newton_temperature(p, y, e, F(bv, i, j, k, 5), &t);
  if (iters) *iters = n;
  F(bv, i, j, k, 5) = t;
  F(bv, i, j, k, 4) = F(bv, i, j, k, 0) * t * gas_const;
}

int orc_newton_temperature(const orc_param *p, const double *y, double e, double t0, double *t_out) {
  const int ns = p->n_spec;
  double mw = 0;
  for (int l = 0; l < ns; ++l) mw += y[l] / p->mw[l];
  mw = 1 / mw;
  const double gas_const = ORC_RU / mw;
  double err = 1, t = t0;
  int iter = 0;
  double h_i[NSMAX], cp_i[NSMAX];
  while (err > 1e-3 && iter++ < 1000) {
    orc_enthalpy_and_cp(t, h_i, cp_i, p);
    double cp_tot = 0, h = 0;
    for (int l = 0; l < ns; ++l) {
      cp_tot += cp_i[l] * y[l];
      h += h_i[l] * y[l];
    }
    const double e_t = h - gas_const * t;
    const double cvv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cvv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  *t_out = t;
  return iter;
}

/* update_cv_and_bv: src/FieldOperation.cuh:139-183.  Returns the number of
 * cells whose Newton iteration hit the 1000-iteration cap. */
int orc_update_cv_and_bv(orc_zone *z, const orc_param *p) {
  int fails = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : fails)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const double dt_div_jac = S(dt_local, i, j, k) / S(jac, i, j, k);
        for (int l = 0; l < p->n_var; ++l) F(cv, i, j, k, l) += F(dq, i, j, k, l) * dt_div_jac;
        if (z->mz == 1) F(cv, i, j, k, 3) = 0;
        F(bv, i, j, k, 0) = F(cv, i, j, k, 0);
        const double density_inv = 1.0 / F(cv, i, j, k, 0);
        F(bv, i, j, k, 1) = F(cv, i, j, k, 1) * density_inv;
        F(bv, i, j, k, 2) = F(cv, i, j, k, 2) * density_inv;
        F(bv, i, j, k, 3) = F(cv, i, j, k, 3) * density_inv;
        double velocity = F(bv, i, j, k, 1) * F(bv, i, j, k, 1) + F(bv, i, j, k, 2) * F(bv, i, j, k, 2) +
                          F(bv, i, j, k, 3) * F(bv, i, j, k, 3);
        if (p->mix_model != ORC_AIR || p->turb_method == ORC_RANS)
          for (int l = 0; l < p->n_scalar; ++l) F(sv, i, j, k, l) = F(cv, i, j, k, 5 + l) * density_inv;
        if (p->mix_model != ORC_AIR) {
          int it = 0;
          orc_compute_temperature(z, i, j, k, p, &it);
          if (it > 1000) ++fails;
        } else {
          F(bv, i, j, k, 4) = (ORC_GAMMA_AIR - 1) * (F(cv, i, j, k, 4) - 0.5 * F(bv, i, j, k, 0) * velocity);
          F(bv, i, j, k, 5) = F(bv, i, j, k, 4) * ORC_MW_AIR * density_inv / ORC_RU;
        }
        S(vel, i, j, k) = sqrt(velocity);
      }
  return fails;
}

/* ---------------------------------------------------------------- limit_flow */

/* limit_flow: src/TimeAdvanceFunc.cu:34-224.  Every repair reads the field
 * as it was before the pass (snapshot semantics; the reference's parallel
 * kernel is order-dependent only where bad cells are adjacent, and equals this
 * everywhere else).  A cell's flow repair precedes its turbulence repair, as in
 * the reference.  The i-neighbour bound uses mz, as the reference does.
 * Returns the number of cells with a flow repair. */
/* running count of flow repairs over every orc_limit_flow call (the
 * device counter's counterpart), read and reset by orc_limit_flow_total */
static long orc_lf_total = 0;

int orc_limit_flow(orc_zone *z, const orc_param *p) {
  const int mx = z->mx, my = z->my, mz = z->mz, ns = p->n_spec, nsc = p->n_scalar, nv = p->n_var;
  const double *ll = p->limit_ll, *ul = p->limit_ul;
  const int nval = 6 + nsc + nv;
  const long ncell = (long)mx * my * mz;
  long *lidx = (long *)malloc(sizeof(long) * (size_t)ncell);
  double *lval = (double *)malloc(sizeof(double) * (size_t)ncell * nval);
  long n = 0;
  int repaired = 0;
  /* pass 1: the final state of every bad cell from the un-repaired field */
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        double var[5];
        int unphysical = 0;
        for (int l = 0; l < 5; ++l) {
          var[l] = F(bv, i, j, k, l);
          if (isnan(var[l]) || var[l] < ll[l] || var[l] > ul[l]) unphysical = 1;
        }
        int tbad = 0;
        if (p->rans_model == 2) {
          const double t0 = F(sv, i, j, k, ns), t1 = F(sv, i, j, k, ns + 1);
          tbad = isnan(t0) || isnan(t1) || t0 < 0 || t1 < 0;
        }
        if (!unphysical && !tbad) continue;
        double bvv[6], svv[NSMAX + 2], cvv[NVMAX];
        for (int l = 0; l < 6; ++l) bvv[l] = F(bv, i, j, k, l);
        for (int l = 0; l < nsc; ++l) svv[l] = F(sv, i, j, k, l);
        for (int l = 0; l < nv; ++l) cvv[l] = F(cv, i, j, k, l);
        if (unphysical) {
          ++repaired;
          double uv[5 + NSMAX + 2];
          memset(uv, 0, sizeof(uv));
          int kn = 0;
          for (int ka = -1; ka < 2; ++ka) {
            const int k1 = k + ka;
            if (k1 < 0 || k1 >= mz) continue;
            for (int ja = -1; ja < 2; ++ja) {
              const int j1 = j + ja;
              if (j1 < 0 || j1 >= my) continue;
              for (int ia = -1; ia < 2; ++ia) {
                const int i1 = i + ia;
                if (i1 < 0 || i1 >= mz) continue;
                int bad = 0;
                for (int l = 0; l < 5; ++l) {
                  const double b = F(bv, i1, j1, k1, l);
                  if (isnan(b) || b < ll[l] || b > ul[l]) bad = 1;
                }
                if (bad) continue;
                for (int l = 0; l < 5; ++l) uv[l] += F(bv, i1, j1, k1, l);
                for (int l = 0; l < nsc; ++l) uv[l + 5] += F(sv, i1, j1, k1, l);
                ++kn;
              }
            }
          }
          if (kn > 0) {
            const double kn_inv = 1.0 / kn;
            for (int l = 0; l < 5 + nsc; ++l) uv[l] *= kn_inv;
          } else {
            for (int l = 0; l < 5; ++l) uv[l] = dmin(dmax(var[l], ll[l]), ul[l]);
            for (int l = 0; l < nsc; ++l) uv[l + 5] = p->sv_inf[l];
          }
          for (int l = 0; l < 5; ++l) bvv[l] = uv[l];
          cvv[0] = uv[0];
          cvv[1] = uv[0] * uv[1];
          cvv[2] = uv[0] * uv[2];
          cvv[3] = uv[0] * uv[3];
          cvv[4] = 0.5 * uv[0] * (uv[1] * uv[1] + uv[2] * uv[2] + uv[3] * uv[3]);
          for (int l = 0; l < nsc; ++l) {
            svv[l] = uv[5 + l];
            cvv[5 + l] = uv[0] * uv[5 + l];
          }
          if (ns > 0) {
            double mw = 0;
            for (int l = 0; l < ns; ++l) mw += svv[l] / p->mw[l];
            mw = 1 / mw;
            bvv[5] = uv[4] * mw / (uv[0] * ORC_RU);
            double h[NSMAX];
            orc_enthalpy(bvv[5], h, p);
            for (int l = 0; l < ns; ++l) cvv[4] += h[l] * cvv[5 + l];
            cvv[4] -= bvv[4];
          } else {
            bvv[5] = uv[4] * ORC_MW_AIR / (uv[0] * ORC_RU);
            cvv[4] += uv[4] / (ORC_GAMMA_AIR - 1);
          }
        }
        if (p->rans_model == 2) {
          const double t0 = svv[ns], t1 = svv[ns + 1];
          if (isnan(t0) || isnan(t1) || t0 < 0 || t1 < 0) {
            double a0 = 0, a1 = 0;
            int kn = 0;
            for (int ka = -1; ka < 2; ++ka) {
              const int k1 = k + ka;
              if (k1 < 0 || k1 >= mz) continue;
              for (int ja = -1; ja < 2; ++ja) {
                const int j1 = j + ja;
                if (j1 < 0 || j1 >= my) continue;
                for (int ia = -1; ia < 2; ++ia) {
                  const int i1 = i + ia;
                  if (i1 < 0 || i1 >= mz) continue;
                  const double a = F(sv, i1, j1, k1, ns), b = F(sv, i1, j1, k1, ns + 1);
                  if (isnan(a) || isnan(b) || a < 0 || b < 0) continue;
                  a0 += a;
                  a1 += b;
                  ++kn;
                }
              }
            }
            if (kn > 0) {
              const double kn_inv = 1.0 / kn;
              a0 *= kn_inv;
              a1 *= kn_inv;
            } else {
              a0 = t0 < 0 ? p->sv_inf[ns] : t0;
              a1 = t1 < 0 ? p->sv_inf[ns + 1] : t1;
            }
            svv[ns] = a0;
            svv[ns + 1] = a1;
            cvv[ns + 5] = cvv[0] * a0;
            cvv[ns + 6] = cvv[0] * a1;
          }
        }
        lidx[n] = IX(z, i, j, k);
        double *o = &lval[n * nval];
        for (int l = 0; l < 6; ++l) o[l] = bvv[l];
        for (int l = 0; l < nsc; ++l) o[6 + l] = svv[l];
        for (int l = 0; l < nv; ++l) o[6 + nsc + l] = cvv[l];
        ++n;
      }
  /* pass 2: write back */
  for (long s = 0; s < n; ++s) {
    const long id = lidx[s];
    const double *o = &lval[s * nval];
    for (int l = 0; l < 6; ++l) z->bv[id + (long)l * z->plane] = o[l];
    for (int l = 0; l < nsc; ++l) z->sv[id + (long)l * z->plane] = o[6 + l];
    for (int l = 0; l < nv; ++l) z->cv[id + (long)l * z->plane] = o[6 + nsc + l];
  }
  free(lidx);
  free(lval);
  orc_lf_total += repaired;
  return repaired;
}

long orc_limit_flow_total(int reset) {
  const long t = orc_lf_total;
  if (reset) orc_lf_total = 0;
  return t;
}

/* ---------------------------------------------------------------- BCs */

static void apply_outflow(orc_zone *z, const orc_param *p, const orc_face *b) {
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i)
        for (int g = 1; g <= z->ngg; ++g) {
          const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
          for (int l = 0; l < 6; ++l) F(bv, gi, gj, gk, l) = F(bv, i, j, k, l);
          for (int l = 0; l < p->n_scalar; ++l) F(sv, gi, gj, gk, l) = F(sv, i, j, k, l);
          if (p->turb_method == ORC_RANS) S(mut, gi, gj, gk) = S(mut, i, j, k);
        }
}

static void apply_inflow(orc_zone *z, const orc_param *p, const orc_face *b, const orc_inflow *in) {
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i)
        for (int g = 1; g <= z->ngg; ++g) {
          const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
          F(bv, gi, gj, gk, 0) = in->density;
          F(bv, gi, gj, gk, 1) = in->u;
          F(bv, gi, gj, gk, 2) = in->v;
          F(bv, gi, gj, gk, 3) = in->w;
          F(bv, gi, gj, gk, 4) = in->pressure;
          F(bv, gi, gj, gk, 5) = in->temperature;
          for (int l = 0; l < p->n_scalar; ++l) F(sv, gi, gj, gk, l) = in->sv[l];
          if (p->turb_method == ORC_RANS) S(mut, gi, gj, gk) = in->mut;
        }
}

/* apply_symmetry: src/BoundCond.cuh:45-125 */
static void apply_symmetry(orc_zone *z, const orc_param *p, const orc_face *b) {
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        const int ii = i - dir[0], ij = j - dir[1], ik = k - dir[2];
        double k_x = M(i, j, k, b->face + 1, 1), k_y = M(i, j, k, b->face + 1, 2), k_z = M(i, j, k, b->face + 1, 3);
        const double k_mag = sqrt(k_x * k_x + k_y * k_y + k_z * k_z);
        k_x /= k_mag;
        k_y /= k_mag;
        k_z /= k_mag;
        const double u1 = F(bv, ii, ij, ik, 1), v1 = F(bv, ii, ij, ik, 2), w1 = F(bv, ii, ij, ik, 3);
        double u_k = k_x * u1 + k_y * v1 + k_z * w1;
        const double u_t = u1 - k_x * u_k, v_t = v1 - k_y * u_k, w_t = w1 - k_z * u_k;
        F(bv, i, j, k, 1) = u_t;
        F(bv, i, j, k, 2) = v_t;
        F(bv, i, j, k, 3) = w_t;
        S(vel, i, j, k) = sqrt(u_t * u_t + v_t * v_t + w_t * w_t);
        F(bv, i, j, k, 0) = F(bv, ii, ij, ik, 0);
        F(bv, i, j, k, 4) = F(bv, ii, ij, ik, 4);
        F(bv, i, j, k, 5) = F(bv, ii, ij, ik, 5);
        for (int l = 0; l < p->n_scalar; ++l) F(sv, i, j, k, l) = F(sv, ii, ij, ik, l);
        F(cv, i, j, k, 0) = F(bv, i, j, k, 0);
        F(cv, i, j, k, 1) = F(bv, i, j, k, 0) * F(bv, i, j, k, 1);
        F(cv, i, j, k, 2) = F(bv, i, j, k, 0) * F(bv, i, j, k, 2);
        F(cv, i, j, k, 3) = F(bv, i, j, k, 0) * F(bv, i, j, k, 3);
        F(cv, i, j, k, 4) = F(cv, ii, ij, ik, 4) - 0.5 * F(bv, ii, ij, ik, 0) * (u1 * u1 + v1 * v1 + w1 * w1) +
                            0.5 * F(bv, i, j, k, 0) * (u_t * u_t + v_t * v_t + w_t * w_t);
        for (int l = 0; l < p->n_scalar; ++l) F(cv, i, j, k, 5 + l) = F(bv, i, j, k, 0) * F(sv, i, j, k, l);
        for (int g = 1; g <= z->ngg; ++g) {
          const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
          const int si = i - g * dir[0], sj = j - g * dir[1], sk = k - g * dir[2];
          F(bv, gi, gj, gk, 0) = F(bv, si, sj, sk, 0);
          const double u = F(bv, si, sj, sk, 1), v = F(bv, si, sj, sk, 2), w = F(bv, si, sj, sk, 3);
          u_k = k_x * u + k_y * v + k_z * w;
          F(bv, gi, gj, gk, 1) = u - 2 * u_k * k_x;
          F(bv, gi, gj, gk, 2) = v - 2 * u_k * k_y;
          F(bv, gi, gj, gk, 3) = w - 2 * u_k * k_z;
          S(vel, gi, gj, gk) = sqrt(F(bv, gi, gj, gk, 1) * F(bv, gi, gj, gk, 1) +
                                    F(bv, gi, gj, gk, 2) * F(bv, gi, gj, gk, 2) +
                                    F(bv, gi, gj, gk, 3) * F(bv, gi, gj, gk, 3));
          F(bv, gi, gj, gk, 4) = F(bv, si, sj, sk, 4);
          F(bv, gi, gj, gk, 5) = F(bv, si, sj, sk, 5);
          for (int l = 0; l < p->n_scalar; ++l) F(sv, gi, gj, gk, l) = F(sv, si, sj, sk, l);
          if (p->turb_method == ORC_RANS) S(mut, gi, gj, gk) = S(mut, si, sj, sk);
        }
      }
}

/* apply_wall: src/BoundCond.cuh:196-319 */
static void apply_wall(orc_zone *z, const orc_param *p, const orc_face *b, const orc_wall *wall) {
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  const int ns = p->n_spec;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        double t_wall = wall->temperature;
        const int ii = i - dir[0], ij = j - dir[1], ik = k - dir[2];
        if (!wall->isothermal) t_wall = F(bv, ii, ij, ik, 5);
        const double pw = F(bv, ii, ij, ik, 4);
        double mw = ORC_MW_AIR;
        if (p->mix_model != ORC_AIR) {
          mw = 0;
          for (int l = 0; l < ns; ++l) {
            F(sv, i, j, k, l) = F(sv, ii, ij, ik, l);
            mw += F(sv, i, j, k, l) / p->mw[l];
          }
          mw = 1 / mw;
        }
        const double rho_wall = pw * mw / (t_wall * ORC_RU);
        F(bv, i, j, k, 0) = rho_wall;
        F(bv, i, j, k, 1) = 0;
        F(bv, i, j, k, 2) = 0;
        F(bv, i, j, k, 3) = 0;
        F(bv, i, j, k, 4) = pw;
        F(bv, i, j, k, 5) = t_wall;
        F(cv, i, j, k, 0) = rho_wall;
        F(cv, i, j, k, 1) = 0;
        F(cv, i, j, k, 2) = 0;
        F(cv, i, j, k, 3) = 0;
        for (int l = 0; l < ns; ++l) F(cv, i, j, k, l + 5) = F(sv, i, j, k, l) * rho_wall;
        total_energy(z, p, i, j, k);
        if (p->turb_method == ORC_RANS && p->rans_model == 2) {
          double mu_wall;
          if (p->mix_model != ORC_AIR) {
            double y[NSMAX];
            for (int l = 0; l < ns; ++l) y[l] = F(sv, i, j, k, l);
            mu_wall = orc_mixture_viscosity(p, t_wall, mw, y);
          } else {
            mu_wall = orc_sutherland(t_wall);
          }
          const double dy = S(wall_distance, ii, ij, ik);
          F(sv, i, j, k, ns) = 0;
          F(cv, i, j, k, ns + 5) = 0;
          if (dy > 1e-25)
            F(sv, i, j, k, ns + 1) = 60 * mu_wall / (rho_wall * SST_BETA_1 * dy * dy);
          else
            F(sv, i, j, k, ns + 1) = F(sv, ii, ij, ik, ns + 1);
          F(cv, i, j, k, ns + 6) = rho_wall * F(sv, i, j, k, ns + 1);
        }
        for (int g = 1; g <= z->ngg; ++g) {
          const int si = i - g * dir[0], sj = j - g * dir[1], sk = k - g * dir[2];
          const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
          const double u_i = F(bv, si, sj, sk, 1), v_i = F(bv, si, sj, sk, 2), w_i = F(bv, si, sj, sk, 3);
          const double p_i = F(bv, si, sj, sk, 4), t_i = F(bv, si, sj, sk, 5);
          double t_g = t_i;
          if (wall->isothermal) {
            t_g = 2 * t_wall - t_i;
            if (t_g <= 0.1 * t_wall) t_g = t_wall;
          }
          if (p->mix_model != ORC_AIR) {
            mw = 0;
            for (int l = 0; l < ns; ++l) {
              F(sv, gi, gj, gk, l) = F(sv, si, sj, sk, l);
              mw += F(sv, gi, gj, gk, l) / p->mw[l];
            }
            mw = 1 / mw;
          }
          const double rho_g = p_i * mw / (t_g * ORC_RU);
          F(bv, gi, gj, gk, 0) = rho_g;
          F(bv, gi, gj, gk, 1) = -u_i;
          F(bv, gi, gj, gk, 2) = -v_i;
          F(bv, gi, gj, gk, 3) = -w_i;
          F(bv, gi, gj, gk, 4) = p_i;
          F(bv, gi, gj, gk, 5) = t_g;
          if (p->turb_method == ORC_RANS && p->rans_model == 2) {
            F(sv, gi, gj, gk, ns) = 0;
            F(sv, gi, gj, gk, ns + 1) = F(sv, i, j, k, ns + 1);
            S(mut, gi, gj, gk) = 0;
          }
        }
      }
}

/* DBoundCond::apply_boundary_conditions: src/BoundCond.cuh:321-412 */
void orc_apply_bc(orc_zone *z, const orc_param *p, const orc_bc *bc) {
  for (int f = 0; f < bc->n_face; ++f) {
    const orc_face *b = &bc->face[f];
    switch (b->kind) {
      case ORC_BC_OUTFLOW: apply_outflow(z, p, b); break;
      case ORC_BC_INFLOW: apply_inflow(z, p, b, &bc->inflow[b->slot]); break;
      case ORC_BC_SYMMETRY: apply_symmetry(z, p, b); break;
      case ORC_BC_WALL: apply_wall(z, p, b, &bc->wall[b->slot]); break;
      default: break;
    }
  }
}

/* update_bv_1_point: src/FieldOperation.cuh:185-215 */
static void update_bv_1_point(orc_zone *z, const orc_param *p, int i, int j, int k) {
  F(bv, i, j, k, 0) = F(cv, i, j, k, 0);
  const double density_inv = 1.0 / F(cv, i, j, k, 0);
  F(bv, i, j, k, 1) = F(cv, i, j, k, 1) * density_inv;
  F(bv, i, j, k, 2) = F(cv, i, j, k, 2) * density_inv;
  F(bv, i, j, k, 3) = F(cv, i, j, k, 3) * density_inv;
  const double velocity = F(bv, i, j, k, 1) * F(bv, i, j, k, 1) + F(bv, i, j, k, 2) * F(bv, i, j, k, 2) +
                          F(bv, i, j, k, 3) * F(bv, i, j, k, 3);
  if (p->mix_model != ORC_AIR || p->turb_method == ORC_RANS)
    for (int l = 0; l < p->n_scalar; ++l) F(sv, i, j, k, l) = F(cv, i, j, k, 5 + l) * density_inv;
  if (p->mix_model != ORC_AIR) {
    orc_compute_temperature(z, i, j, k, p, NULL);
  } else {
    F(bv, i, j, k, 4) = (ORC_GAMMA_AIR - 1) * (F(cv, i, j, k, 4) - 0.5 * F(bv, i, j, k, 0) * velocity);
    F(bv, i, j, k, 5) = F(bv, i, j, k, 4) * ORC_MW_AIR * density_inv / ORC_RU;
  }
  S(vel, i, j, k) = sqrt(velocity);
}

/* inner_communication: src/DataCommunication.cuh:58-98, one face of block blk */
void orc_inner_communication(orc_zone *zones, int blk, const orc_inner_face *f, const orc_param *p) {
  orc_zone *z = &zones[blk];
  orc_zone *t = &zones[f->target_block];
  const int face_dir = f->direction > 0 ? f->range_start[f->face] : f->range_end[f->face];
  for (int n2 = 0; n2 < f->n_point[2]; ++n2)
    for (int n1 = 0; n1 < f->n_point[1]; ++n1)
      for (int n0 = 0; n0 < f->n_point[0]; ++n0) {
        const int n[3] = {n0, n1, n2};
        int idx[3], it[3], d[3];
        for (int q = 0; q < 3; ++q) {
          d[q] = f->loop_dir[q] * n[q];
          idx[q] = f->range_start[q] + d[q];
        }
        for (int q = 0; q < 3; ++q) it[q] = f->target_start[q] + f->target_loop_dir[q] * d[f->src_tar[q]];
        if (idx[f->face] == face_dir) {
          for (int l = 0; l < p->n_var; ++l) {
            double *a = &z->cv[IX(z, idx[0], idx[1], idx[2]) + (long)l * z->plane];
            double *b = &t->cv[IX(t, it[0], it[1], it[2]) + (long)l * t->plane];
            const double ave = 0.5 * (*b + *a);
            *a = ave;
            *b = ave;
          }
          update_bv_1_point(z, p, idx[0], idx[1], idx[2]);
        } else {
          for (int l = 0; l < 6; ++l)
            z->bv[IX(z, idx[0], idx[1], idx[2]) + (long)l * z->plane] = t->bv[IX(t, it[0], it[1], it[2]) + (long)l * t->plane];
          for (int l = 0; l < p->n_scalar; ++l)
            z->sv[IX(z, idx[0], idx[1], idx[2]) + (long)l * z->plane] = t->sv[IX(t, it[0], it[1], it[2]) + (long)l * t->plane];
        }
      }
}

/* eliminate_k_gradient: src/FieldOperation.cu:42-62.  A 2-D block (mz = 1)
 * repeats its k = 0 plane into the k ghost layers, so every k derivative
 * vanishes; i and j cover the ghost layers too. */
void orc_eliminate_k_gradient(orc_zone *z, const orc_param *p) {
  const int g = z->ngg;
  for (int j = -g; j < z->my + g; ++j)
    for (int i = -g; i < z->mx + g; ++i)
      for (int k = 1; k <= g; ++k) {
        for (int l = 0; l < 6; ++l) {
          F(bv, i, j, k, l) = F(bv, i, j, 0, l);
          F(bv, i, j, -k, l) = F(bv, i, j, 0, l);
        }
        for (int l = 0; l < p->n_scalar; ++l) {
          F(sv, i, j, k, l) = F(sv, i, j, 0, l);
          F(sv, i, j, -k, l) = F(sv, i, j, 0, l);
        }
      }
}

/* ---------------------------------------------------------------- residual */

/* store_last_step: src/TimeAdvanceFunc.cu:5-16 (bv_last interior, 4 planes) */
void orc_store_last_step(orc_zone *z) {
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        F(bv_last, i, j, k, 0) = F(bv, i, j, k, 0);
        F(bv_last, i, j, k, 1) = S(vel, i, j, k);
        F(bv_last, i, j, k, 2) = F(bv, i, j, k, 4);
        F(bv_last, i, j, k, 3) = F(bv, i, j, k, 5);
      }
}

/* compute_square_of_dbv: src/TimeAdvanceFunc.cu:18-32 */
void orc_square_of_dbv(orc_zone *z) {
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        double d;
        d = F(bv, i, j, k, 0) - F(bv_last, i, j, k, 0);
        F(bv_last, i, j, k, 0) = d * d;
        d = S(vel, i, j, k) - F(bv_last, i, j, k, 1);
        F(bv_last, i, j, k, 1) = d * d;
        d = F(bv, i, j, k, 4) - F(bv_last, i, j, k, 2);
        F(bv_last, i, j, k, 2) = d * d;
        d = F(bv, i, j, k, 5) - F(bv_last, i, j, k, 3);
        F(bv_last, i, j, k, 3) = d * d;
      }
}

/* reduction_of_dv_squared: src/Driver.cu:417-459 (sum over the interior) */
void orc_sum_dbv(const orc_zone *z, double out[4]) {
  for (int l = 0; l < 4; ++l) {
    double s = 0;
    for (int k = 0; k < z->mz; ++k)
      for (int j = 0; j < z->my; ++j)
        for (int i = 0; i < z->mx; ++i) s += F(bv_last, i, j, k, l);
    out[l] = s;
  }
}

/* compute_wall_distance: src/Driver.cu:461-482 (all points incl. ghosts) */
void orc_wall_distance(orc_zone *z, const double *wp, int n_wall_point) {
  const int g = z->ngg;
#pragma omp parallel for collapse(2)
  for (int k = -g; k < z->mz + g; ++k)
    for (int j = -g; j < z->my + g; ++j)
      for (int i = -g; i < z->mx + g; ++i) {
        const double x = S(x, i, j, k), y = S(y, i, j, k), zz = S(z, i, j, k);
        double wd = 1e+6;
        for (int l = 0; l < n_wall_point; ++l) {
          const double d = (x - wp[3 * l]) * (x - wp[3 * l]) + (y - wp[3 * l + 1]) * (y - wp[3 * l + 1]) +
                           (zz - wp[3 * l + 2]) * (zz - wp[3 * l + 2]);
          if (wd > d) wd = d;
        }
        S(wall_distance, i, j, k) = sqrt(wd);
      }
}

/* ---------------------------------------------------------------- one step */

double orc_steady_step(orc_zone *zones, int n_blocks, const orc_param *p, const orc_bc *bcs,
                       const orc_inner_face *inner, const int *n_inner, int step, int output_screen,
                       double res_scale[4], long n_grid_total, int *newton_fail) {
  int nf = 0;
  const int do_res = (step % output_screen == 0) || step == 1;
  for (int b = 0; b < n_blocks; ++b) {
    orc_zone *z = &zones[b];
    if (step % output_screen == 0) orc_store_last_step(z);
    memset(z->dq, 0, sizeof(double) * (size_t)z->plane * p->n_var);
    orc_compute_source(z, p);
    orc_inviscid_flux(z, p);
    orc_viscous_flux(z, p);
    orc_local_time_step(z, p);
    orc_implicit_treatment(z, p);
    nf += orc_update_cv_and_bv(z, p);
    orc_limit_flow(z, p);
    orc_apply_bc(z, p, &bcs[b]);
  }
  int off = 0;
  for (int b = 0; b < n_blocks; ++b) {
    for (int f = 0; f < n_inner[b]; ++f) orc_inner_communication(zones, b, &inner[off + f], p);
    off += n_inner[b];
  }
  for (int b = 0; b < n_blocks; ++b)
    if (zones[b].mz == 1) orc_eliminate_k_gradient(&zones[b], p);
  for (int b = 0; b < n_blocks; ++b) orc_update_physical_properties(&zones[b], p);
  if (newton_fail) *newton_fail = nf;
  if (!do_res) return -1;
  double res[4] = {0, 0, 0, 0};
  for (int b = 0; b < n_blocks; ++b) {
    double r[4];
    orc_square_of_dbv(&zones[b]);
    orc_sum_dbv(&zones[b], r);
    for (int l = 0; l < 4; ++l) res[l] += r[l];
  }
  for (int l = 0; l < 4; ++l) res[l] = sqrt(res[l] / (double)n_grid_total);
  if (step == 1)
    for (int l = 0; l < 4; ++l) res_scale[l] = res[l] < 1e-20 ? 1e-20 : res[l];
  for (int l = 0; l < 4; ++l) res[l] /= res_scale[l];
  double err_max = res[0];
  for (int l = 1; l < 4; ++l)
    if (fabs(res[l]) > err_max) err_max = res[l];
  return err_max;
}

int orc_omp_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Test access to the DPLUR convective-Jacobian product of one cell (the
 * product A_dir(Q)*dq + pm_sr*dq with dq read from z->dq). */
void orc_jac_times_dq_point(const orc_zone *z, const orc_param *p, int i, int j, int k, int dir, double pm_sr,
                            double *out) {
  jac_times_dq(z, p, i, j, k, dir, pm_sr, out);
}
