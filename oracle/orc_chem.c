/* Finite-rate chemistry restatement.  TEST INFRASTRUCTURE ONLY. */
#include <math.h>
#include <string.h>

#include "orc_internal.h"

#define NS ORC_MAX_SPEC
#define NR ORC_MAX_REAC

/* arrhenius: src/FiniteRateChem.cu:110-112 */
static inline double arrhenius(double t, double A, double b, double Ea) { return A * pow(t, b) * exp(-Ea / t); }

/* forward_reaction_rate: src/FiniteRateChem.cu:65-108 */
static void forward_rate(double t, double *kf, const double *c, const orc_param *p) {
  for (int i = 0; i < p->n_reac; ++i) {
    kf[i] = arrhenius(t, p->A[i], p->b[i], p->Ea[i]);
    const int type = p->reac_type[i];
    if (type == 3) kf[i] += arrhenius(t, p->A2[i], p->b2[i], p->Ea2[i]);
    if (type > 3) {
      double cc = 0;
      for (int l = 0; l < p->n_spec; ++l) cc += c[l] * p->third_body_coeff[i * NS + l];
      if (type == 4) {
        kf[i] *= cc;
      } else {
        const double kf_low = arrhenius(t, p->A2[i], p->b2[i], p->Ea2[i]);
        const double kf_high = kf[i];
        const double reduced_pressure = kf_low * cc / kf_high;
        double Fv = 1.0;
        if (type > 5) {
          double f_cent = (1 - p->troe_alpha[i]) * exp(-t / p->troe_t3[i]) + p->troe_alpha[i] * exp(-t / p->troe_t1[i]);
          if (type == 7) f_cent += exp(-p->troe_t2[i] / t);
          const double logFc = log10(f_cent);
          const double cc1 = -0.4 - 0.67 * logFc;
          const double n = 0.75 - 1.27 * logFc;
          const double logPr = log10(reduced_pressure);
          const double tempo = (logPr + cc1) / (n - 0.14 * (logPr + cc1));
          const double pp = logFc / (1.0 + tempo * tempo);
          Fv = pow(10, pp);
        }
        kf[i] = kf_high * reduced_pressure / (1.0 + reduced_pressure) * Fv;
      }
    }
  }
}

/* backward_reaction_rate: src/FiniteRateChem.cu:114-136 */
static void backward_rate(double t, const double *kf, const orc_param *p, double *kb) {
  double g[NS];
  memset(g, 0, sizeof(g));
  orc_gibbs_div_rt(t, p, g);
  const double temp_p = ORC_P_ATM / ORC_RU * 1e-3;
  const double temp_t = temp_p / t;
  for (int i = 0; i < p->n_reac; ++i) {
    if (p->reac_type[i] != 2) {
      double d_gibbs = 0;
      for (int l = 0; l < p->n_spec; ++l) d_gibbs += g[l] * (p->stoi_b[i * NS + l] - p->stoi_f[i * NS + l]);
      const double kc = pow(temp_t, p->reac_order[i]) * exp(-d_gibbs);
      kb[i] = kf[i] / kc;
    } else {
      kb[i] = arrhenius(t, p->A2[i], p->b2[i], p->Ea2[i]);
    }
  }
}

/* rate_of_progress + chemical_source + Jacobians:
 * src/FiniteRateChem.cu:138-209.  rho*Y, mol/cm3 internally. */
void orc_chem_source_point(const orc_param *p, double rho, double t, const double *y, double *omega,
                           double *omega_d, double *jac_epi, double *jac_da, double *kf_out, double *kb_out) {
  const int ns = p->n_spec, nr = p->n_reac;
  double c[NS], kf[NR], kb[NR], q1[NR], q2[NR];
  memset(c, 0, sizeof(c));
  for (int l = 0; l < ns; ++l) c[l] = rho * y[l] / p->mw[l] * 1e-3;
  forward_rate(t, kf, c, p);
  backward_rate(t, kf, p, kb);
  for (int i = 0; i < nr; ++i) {
    q1[i] = 1.0;
    q2[i] = 1.0;
    for (int j = 0; j < ns; ++j) {
      q1[i] *= pow(c[j], p->stoi_f[i * NS + j]);
      q2[i] *= pow(c[j], p->stoi_b[i * NS + j]);
    }
    q1[i] *= kf[i];
    q2[i] *= kb[i];
  }
  for (int i = 0; i < ns; ++i) {
    double creation = 0, od = 0;
    for (int j = 0; j < nr; ++j) {
      creation += q2[j] * p->stoi_f[j * NS + i] + q1[j] * p->stoi_b[j * NS + i];
      od += q1[j] * p->stoi_f[j * NS + i] + q2[j] * p->stoi_b[j * NS + i];
    }
    creation *= p->mw[i] * 1e+3;
    od *= p->mw[i] * 1e+3;
    omega[i] = creation - od;
    if (omega_d) omega_d[i] = od;
  }
  if (jac_epi) {
    for (int m = 0; m < ns; ++m)
      for (int n = 0; n < ns; ++n) {
        double zz = 0;
        if (y[n] > 1e-25) {
          for (int r = 0; r < nr; ++r)
            zz += (p->stoi_b[r * NS + m] - p->stoi_f[r * NS + m]) *
                  (p->stoi_f[r * NS + n] * q1[r] - p->stoi_b[r * NS + n] * q2[r]);
          zz /= rho * y[n];
        }
        jac_epi[m * ns + n] = p->mw[m] * zz * 1e+3;
      }
  }
  if (jac_da) {
    for (int l = 0; l < ns; ++l) {
      double od = 0;
      for (int j = 0; j < nr; ++j) od += q1[j] * p->stoi_f[j * NS + l] + q2[j] * p->stoi_b[j * NS + l];
      od *= p->mw[l] * 1e+3;
      jac_da[l] = y[l] > 1e-25 ? -od / (y[l] * rho) : 0;
    }
  }
  if (kf_out)
    for (int i = 0; i < nr; ++i) kf_out[i] = kf[i];
  if (kb_out)
    for (int i = 0; i < nr; ++i) kb_out[i] = kb[i];
}

/* finite_rate_chemistry: src/FiniteRateChem.cu:7-63 */
void orc_finite_rate_chemistry(orc_zone *z, int i, int j, int k, const orc_param *p) {
  const int ns = p->n_spec;
  double y[NS], omega[NS], jac[NS * NS], jd[NS];
  for (int l = 0; l < ns; ++l) y[l] = F(sv, i, j, k, l);
  const double rho = F(bv, i, j, k, 0), t = F(bv, i, j, k, 5);
  orc_chem_source_point(p, rho, t, y, omega, NULL, p->chem_src_method == 1 ? jac : NULL,
                        p->chem_src_method == 2 ? jd : NULL, NULL, NULL);
  for (int l = 0; l < ns; ++l) F(dq, i, j, k, l + 5) += S(jac, i, j, k) * omega[l];
  if (p->chem_src_method == 1)
    for (int q = 0; q < ns * ns; ++q) F(chem_src_jac, i, j, k, q) = jac[q];
  else if (p->chem_src_method == 2)
    for (int l = 0; l < ns; ++l) F(chem_src_jac, i, j, k, l) = jd[l];
}

/* solve_chem_system(lhs, rhs, dim): src/FiniteRateChem.cu:344-398 */
void orc_solve_lu(double *lhs, double *rhs, int dim) {
  int ipiv[NS];
  memset(ipiv, 0, sizeof(ipiv));
  for (int n = 0; n < dim; ++n) {
    int ik = n;
    for (int m = n; m < dim; ++m) {
      for (int t = 0; t < n; ++t) lhs[m * dim + n] -= lhs[m * dim + t] * lhs[t * dim + n];
      if (fabs(lhs[m * dim + n]) > fabs(lhs[ik * dim + n])) ik = m;
    }
    ipiv[n] = ik;
    if (ik != n)
      for (int t = 0; t < dim; ++t) {
        const double mid = lhs[ik * dim + t];
        lhs[ik * dim + t] = lhs[n * dim + t];
        lhs[n * dim + t] = mid;
      }
    for (int q = n + 1; q < dim; ++q)
      for (int t = 0; t < n; ++t) lhs[n * dim + q] -= lhs[n * dim + t] * lhs[t * dim + q];
    for (int m = n + 1; m < dim; ++m) lhs[m * dim + n] /= lhs[n * dim + n];
  }
  for (int m = 0; m < dim; ++m) {
    const int t = ipiv[m];
    if (t != m) {
      const double mid = rhs[t];
      rhs[t] = rhs[m];
      rhs[m] = mid;
    }
  }
  for (int m = 1; m < dim; ++m)
    for (int t = 0; t < m; ++t) rhs[m] -= lhs[m * dim + t] * rhs[t];
  rhs[dim - 1] /= lhs[dim * dim - 1];
  for (int m = dim - 2; m >= 0; --m) {
    for (int t = m + 1; t < dim; ++t) rhs[m] -= lhs[m * dim + t] * rhs[t];
    rhs[m] /= lhs[m * dim + m];
  }
}

/* EPI / EPI_for_dq0: src/FiniteRateChem.cu:211-253 — solve
 * (d*I - dt*J) x = dq_species in place (d = 1 explicit, diag for DPLUR) */
void orc_epi_solve_dq(orc_zone *z, int i, int j, int k, double d, const orc_param *p) {
  const int ns = p->n_spec;
  double lhs[NS * NS], rhs[NS];
  const double dt = S(dt_local, i, j, k);
  for (int m = 0; m < ns; ++m)
    for (int n = 0; n < ns; ++n)
      lhs[m * ns + n] = (m == n ? d : 0.0) - dt * F(chem_src_jac, i, j, k, m * ns + n);
  for (int l = 0; l < ns; ++l) rhs[l] = F(dq, i, j, k, 5 + l);
  orc_solve_lu(lhs, rhs, ns);
  for (int l = 0; l < ns; ++l) F(dq, i, j, k, 5 + l) = rhs[l];
}

/* EPI_for_dqk: src/FiniteRateChem.cu:255-284 */
void orc_epi_dqk(orc_zone *z, int i, int j, int k, double diag, const double *dq_total, const orc_param *p) {
  const int ns = p->n_spec;
  double lhs[NS * NS], rhs[NS];
  const double dt = S(dt_local, i, j, k);
  for (int m = 0; m < ns; ++m)
    for (int n = 0; n < ns; ++n)
      lhs[m * ns + n] = m == n ? diag - dt * F(chem_src_jac, i, j, k, m * ns + n) : -dt * F(chem_src_jac, i, j, k, m * ns + n);
  for (int l = 0; l < ns; ++l) rhs[l] = dq_total[5 + l] * dt;
  orc_solve_lu(lhs, rhs, ns);
  for (int l = 0; l < ns; ++l) F(dqk, i, j, k, l + 5) = F(dq0, i, j, k, l + 5) + rhs[l];
}

/* The plain point-implicit chemistry step: the reference's explicit-mode EPI
 * (src/FiniteRateChem.cu:211-234) applied to the chemical source alone. */
void orc_plain_chem_step(const orc_param *p, double t, double pres, const double *y, double dt, double *dy) {
  const int ns = p->n_spec;
  double mw = 0;
  for (int l = 0; l < ns; ++l) mw += y[l] / p->mw[l];
  mw = 1 / mw;
  const double rho = pres * mw / (ORC_RU * t);
  double omega[NS], jac[NS * NS], lhs[NS * NS], rhs[NS];
  orc_chem_source_point(p, rho, t, y, omega, NULL, jac, NULL, NULL, NULL);
  for (int m = 0; m < ns; ++m)
    for (int n = 0; n < ns; ++n) lhs[m * ns + n] = (m == n ? 1.0 : 0.0) - dt * jac[m * ns + n];
  for (int l = 0; l < ns; ++l) rhs[l] = dt * omega[l];
  orc_solve_lu(lhs, rhs, ns);
  for (int l = 0; l < ns; ++l) dy[l] = rhs[l] / rho;
}

void orc_plain_chem_batch(const orc_param *p, long n, const double *t, const double *pres, const double *y,
                          const double *dt, double *dy) {
  const int ns = p->n_spec;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) orc_plain_chem_step(p, t[c], pres[c], &y[c * ns], dt[c], &dy[c * ns]);
}
