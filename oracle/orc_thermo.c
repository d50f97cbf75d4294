/* Thermodynamic and transport restatement.  TEST INFRASTRUCTURE ONLY. */
#include <math.h>

#include "orc_internal.h"

/* compute_enthalpy: src/Thermo.cu:5-23 (below T_low: h(T_low)+cp(T_low)(T-T_low)) */
void orc_enthalpy(double t, double *h, const orc_param *p) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < p->n_spec; ++i) {
    if (t < p->t_low[i]) {
      const double tt = p->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cp = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp * (t - tt);
    } else {
      const double *c = t < p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= ORC_RU / p->mw[i];
  }
}

/* compute_enthalpy_and_cp: src/Thermo.cu:25-45 */
void orc_enthalpy_and_cp(double t, double *h, double *cp, const orc_param *p) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < p->n_spec; ++i) {
    if (t < p->t_low[i]) {
      const double tt = p->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp[i] * (t - tt);
    } else {
      const double *c = t < p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= ORC_RU / p->mw[i];
    cp[i] *= ORC_RU / p->mw[i];
  }
}

/* compute_cp: src/Thermo.cu:47-61 */
void orc_cp(double t, double *cp, const orc_param *p) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
  for (int i = 0; i < p->n_spec; ++i) {
    if (t < p->t_low[i]) {
      const double tt = p->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double *c = t < p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= ORC_RU / p->mw[i];
  }
}

/* compute_gibbs_div_rt: src/Thermo.cu:63-79 (below T_low: evaluated at T_low) */
void orc_gibbs_div_rt(double t, const orc_param *p, double *g) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
  for (int i = 0; i < p->n_spec; ++i) {
    if (t < p->t_low[i]) {
      const double tt = p->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      const double *c = &p->low_temp_coeff[i * 7];
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const double *c = t < p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

/* Sutherland: src/Transport.cu:6-9 */
double orc_sutherland(double t) { return 1.716e-5 * pow(t / 273, 1.5) * (273 + 111) / (t + 111); }

/* Species viscosities (collision integral) and mole fractions:
 * src/Transport.cu:48-53 */
static void species_visc(const orc_param *p, double t, double mw_total, const double *y, double *x, double *vis) {
  for (int l = 0; l < p->n_spec; ++l) {
    x[l] = y[l] * mw_total / p->mw[l];
    const double t_dl = t * p->LJ_potent_inv[l];
    const double ci = 1.147 * pow(t_dl, -0.145) + pow(t_dl + 0.5, -2);
    vis[l] = p->vis_coeff[l] * sqrt(t) / ci;
  }
}

/* compute_transport_property: src/Transport.cu:40-84 (Wilke mixing,
 * Eucken-like conductivity, constant-Sc rho*D) */
void orc_transport_point(const orc_param *p, double t, double mw_total, const double *y, const double *cp,
                         double *mu_out, double *lambda_out, double *rhoD) {
  double x[ORC_MAX_SPEC], vis[ORC_MAX_SPEC];
  species_visc(p, t, mw_total, y, x, vis);
  const int ns = p->n_spec;
  double viscosity = 0, conductivity = 0;
  for (int m = 0; m < ns; ++m) {
    double vis_temp = 0;
    for (int n = 0; n < ns; ++n) {
      double pf = 1.0;
      if (m != n) {
        const double num = 1 + sqrt(vis[m] / vis[n]) * p->WjDivWi_to_One4th[m * ORC_MAX_SPEC + n];
        pf = num * num * p->sqrt_WiDivWjPl1Mul8[m * ORC_MAX_SPEC + n];
      }
      vis_temp += pf * x[n];
    }
    const double cond_temp = 1.065 * vis_temp - 0.065 * x[m];
    viscosity += vis[m] * x[m] / vis_temp;
    const double lambda = vis[m] * (cp[m] + 1.25 * ORC_RU / p->mw[m]);
    conductivity += lambda * x[m] / cond_temp;
  }
  *mu_out = viscosity;
  *lambda_out = conductivity;
  for (int l = 0; l < ns; ++l) {
    if (fabs(x[l] - 1) < 1e-3)
      rhoD[l] = viscosity / p->Sc;
    else
      rhoD[l] = (1 - y[l]) * viscosity / ((1 - x[l]) * p->Sc);
  }
}

/* compute_viscosity (device form): src/Transport.cu:86-115 */
double orc_mixture_viscosity(const orc_param *p, double t, double mw_total, const double *y) {
  double x[ORC_MAX_SPEC], vis[ORC_MAX_SPEC];
  species_visc(p, t, mw_total, y, x, vis);
  const int ns = p->n_spec;
  double viscosity = 0;
  for (int m = 0; m < ns; ++m) {
    double vis_temp = 0;
    for (int n = 0; n < ns; ++n) {
      double pf = 1.0;
      if (m != n) {
        const double num = 1 + sqrt(vis[m] / vis[n]) * p->WjDivWi_to_One4th[m * ORC_MAX_SPEC + n];
        pf = num * num * p->sqrt_WiDivWjPl1Mul8[m * ORC_MAX_SPEC + n];
      }
      vis_temp += pf * x[n];
    }
    viscosity += vis[m] * x[m] / vis_temp;
  }
  return viscosity;
}
