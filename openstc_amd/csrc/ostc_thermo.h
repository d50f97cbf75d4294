// Device thermodynamics, transport and finite-rate chemistry for one point.
// T is the arithmetic type (float for the FP32 bulk path where the stage runs
// in FP32, double for chemistry and the Newton solve).
#pragma once

#include "ostc_dev.h"

namespace ostc {

// NASA-7 enthalpy h_k(T) [J/kg]; below T_low linear extrapolation with
// cp(T_low) (src/Thermo.cu:5-23).
template <typename T>
__device__ __forceinline__ void enthalpy(T t, T *h, const ostc_param *__restrict__ p, int ns) {
  const T t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < ns; ++i) {
    const double tl = p->t_low[i];
    if (t < (T)tl) {
      const T tt = (T)tl;
      const T tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      T hv = (T)c[0] * tt + (T)0.5 * (T)c[1] * tt2 + (T)c[2] * tt3 / (T)3 + (T)0.25 * (T)c[3] * tt4 +
             (T)0.2 * (T)c[4] * tt5 + (T)c[5];
      const T cp = (T)c[0] + (T)c[1] * tt + (T)c[2] * tt2 + (T)c[3] * tt3 + (T)c[4] * tt4;
      hv += cp * (t - tt);
      h[i] = hv * (T)(kRu / p->mw[i]);
    } else {
      const double *c = t < (T)p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      const T hv = (T)c[0] * t + (T)0.5 * (T)c[1] * t2 + (T)c[2] * t3 / (T)3 + (T)0.25 * (T)c[3] * t4 +
                   (T)0.2 * (T)c[4] * t5 + (T)c[5];
      h[i] = hv * (T)(kRu / p->mw[i]);
    }
  }
}

// h_k and cp_k (src/Thermo.cu:25-45)
template <typename T>
__device__ __forceinline__ void enthalpy_cp(T t, T *h, T *cp, const ostc_param *__restrict__ p, int ns) {
  const T t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < ns; ++i) {
    const T scale = (T)(kRu / p->mw[i]);
    if (t < (T)p->t_low[i]) {
      const T tt = (T)p->t_low[i];
      const T tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      T hv = (T)c[0] * tt + (T)0.5 * (T)c[1] * tt2 + (T)c[2] * tt3 / (T)3 + (T)0.25 * (T)c[3] * tt4 +
             (T)0.2 * (T)c[4] * tt5 + (T)c[5];
      const T cpv = (T)c[0] + (T)c[1] * tt + (T)c[2] * tt2 + (T)c[3] * tt3 + (T)c[4] * tt4;
      hv += cpv * (t - tt);
      h[i] = hv * scale;
      cp[i] = cpv * scale;
    } else {
      const double *c = t < (T)p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      const T hv = (T)c[0] * t + (T)0.5 * (T)c[1] * t2 + (T)c[2] * t3 / (T)3 + (T)0.25 * (T)c[3] * t4 +
                   (T)0.2 * (T)c[4] * t5 + (T)c[5];
      const T cpv = (T)c[0] + (T)c[1] * t + (T)c[2] * t2 + (T)c[3] * t3 + (T)c[4] * t4;
      h[i] = hv * scale;
      cp[i] = cpv * scale;
    }
  }
}

// cp_k (src/Thermo.cu:47-61)
template <typename T>
__device__ __forceinline__ void heat_capacity(T t, T *cp, const ostc_param *__restrict__ p, int ns) {
  const T t2 = t * t, t3 = t2 * t, t4 = t3 * t;
  for (int i = 0; i < ns; ++i) {
    T cpv;
    if (t < (T)p->t_low[i]) {
      const T tt = (T)p->t_low[i];
      const T tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double *c = &p->low_temp_coeff[i * 7];
      cpv = (T)c[0] + (T)c[1] * tt + (T)c[2] * tt2 + (T)c[3] * tt3 + (T)c[4] * tt4;
    } else {
      const double *c = t < (T)p->t_mid[i] ? &p->low_temp_coeff[i * 7] : &p->high_temp_coeff[i * 7];
      cpv = (T)c[0] + (T)c[1] * t + (T)c[2] * t2 + (T)c[3] * t3 + (T)c[4] * t4;
    }
    cp[i] = cpv * (T)(kRu / p->mw[i]);
  }
}

// Sutherland viscosity (src/Transport.cu:6-9)
template <typename T>
__device__ __forceinline__ T sutherland(T t) {
  const T r = t / (T)273;
  return (T)1.716e-5 * r * sqrt(r) * (T)(273 + 111) / (t + (T)111);
}

// Species viscosities from the collision integral and Wilke mixing
// (src/Transport.cu:40-115).  Writes mu; if lambda/rhoD non-null also the
// Eucken-like conductivity and constant-Schmidt rho*D.
template <typename T>
__device__ __forceinline__ void mixture_transport(const ostc_param *__restrict__ p, int ns, T t, T mw_total,
                                                  const T *y, const T *cp, T *mu, T *lambda, T *rhoD) {
  T x[kMaxSpec], vis[kMaxSpec];
  const T sqrt_t = sqrt(t);
  for (int l = 0; l < ns; ++l) {
    x[l] = y[l] * mw_total / (T)p->mw[l];
    const T t_dl = t * (T)p->LJ_potent_inv[l];
    const T ci = (T)1.147 * pow(t_dl, (T)-0.145) + (T)1 / ((t_dl + (T)0.5) * (t_dl + (T)0.5));
    vis[l] = (T)p->vis_coeff[l] * sqrt_t / ci;
  }
  T viscosity = 0, conductivity = 0;
  for (int m = 0; m < ns; ++m) {
    T vis_temp = 0;
    for (int n = 0; n < ns; ++n) {
      T pf = 1;
      if (m != n) {
        const T num = (T)1 + sqrt(vis[m] / vis[n]) * (T)p->WjDivWi_to_One4th[m * kMaxSpec + n];
        pf = num * num * (T)p->sqrt_WiDivWjPl1Mul8[m * kMaxSpec + n];
      }
      vis_temp += pf * x[n];
    }
    viscosity += vis[m] * x[m] / vis_temp;
    if (lambda) {
      const T cond_temp = (T)1.065 * vis_temp - (T)0.065 * x[m];
      const T lam = vis[m] * (cp[m] + (T)(1.25 * kRu / p->mw[m]));
      conductivity += lam * x[m] / cond_temp;
    }
  }
  *mu = viscosity;
  if (lambda) *lambda = conductivity;
  if (rhoD) {
    for (int l = 0; l < ns; ++l) {
      if (fabs(x[l] - (T)1) < (T)1e-3)
        rhoD[l] = viscosity / (T)p->Sc;
      else
        rhoD[l] = ((T)1 - y[l]) * viscosity / (((T)1 - x[l]) * (T)p->Sc);
    }
  }
}

// ---------------------------------------------------------------- chemistry
// All FP64 (Arrhenius / equilibrium constants / point-implicit Jacobian).

__device__ __forceinline__ double arrhenius(double t, double log_t, double A, double b, double Ea) {
  // A*T^b*exp(-Ea/T) evaluated as one exp of the combined exponent (b==0
  // skips the log term); src/FiniteRateChem.cu:110-112
  return b == 0.0 ? A * exp(-Ea / t) : A * exp(b * log_t - Ea / t);
}

// g_k/(R T) (src/Thermo.cu:63-79, evaluated at T_low below T_low)
__device__ __forceinline__ void gibbs_div_rt(double t, const ostc_param *__restrict__ p, int ns, double *g) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
  for (int i = 0; i < ns; ++i) {
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

// integer power for stoichiometric exponents (0..3); pow(c, 0) == 1
__device__ __forceinline__ double ipow(double c, int n) {
  double r = 1.0;
  for (int q = 0; q < n; ++q) r *= c;
  return r;
}

// Rates of progress q1 (forward), q2 (backward) in mol/cm3/s at (rho, T, Y):
// src/FiniteRateChem.cu:7-153.
template <int NS, int NR>
__device__ __forceinline__ void rates_of_progress(const ostc_param *__restrict__ p, int ns, int nr, double rho,
                                                  double t, const double *y, double *q1, double *q2) {
  double c[NS];
  for (int l = 0; l < ns; ++l) c[l] = rho * y[l] / p->mw[l] * 1e-3;
  double g[NS];
  gibbs_div_rt(t, p, ns, g);
  const double log_t = log(t);
  const double temp_t = kPatm / kRu * 1e-3 / t;
  for (int i = 0; i < nr; ++i) {
    const int type = p->reac_type[i];
    double kf = arrhenius(t, log_t, p->A[i], p->b[i], p->Ea[i]);
    if (type == 3) kf += arrhenius(t, log_t, p->A2[i], p->b2[i], p->Ea2[i]);
    if (type > 3) {
      double cc = 0;
      for (int l = 0; l < ns; ++l) cc += c[l] * p->third_body_coeff[i * kMaxSpec + l];
      if (type == 4) {
        kf *= cc;
      } else {
        const double kf_low = arrhenius(t, log_t, p->A2[i], p->b2[i], p->Ea2[i]);
        const double kf_high = kf;
        const double pr = kf_low * cc / kf_high;
        double fv = 1.0;
        if (type > 5) {
          double f_cent = (1 - p->troe_alpha[i]) * exp(-t / p->troe_t3[i]) + p->troe_alpha[i] * exp(-t / p->troe_t1[i]);
          if (type == 7) f_cent += exp(-p->troe_t2[i] / t);
          const double logFc = log10(f_cent);
          const double cc1 = -0.4 - 0.67 * logFc;
          const double n = 0.75 - 1.27 * logFc;
          const double logPr = log10(pr);
          const double tempo = (logPr + cc1) / (n - 0.14 * (logPr + cc1));
          fv = exp10(logFc / (1.0 + tempo * tempo));
        }
        kf = kf_high * pr / (1.0 + pr) * fv;
      }
    }
    double kb;
    if (type != 2) {
      double d_gibbs = 0;
      for (int l = 0; l < ns; ++l) {
        const int dn = p->stoi_b[i * kMaxSpec + l] - p->stoi_f[i * kMaxSpec + l];
        if (dn != 0) d_gibbs += g[l] * dn;
      }
      const int order = p->reac_order[i];
      // kb = kf / Kc with Kc = temp_t^order * exp(-d_gibbs)
      kb = kf * exp(d_gibbs) / ipow(order >= 0 ? temp_t : 1.0 / temp_t, order >= 0 ? order : -order);
    } else {
      kb = arrhenius(t, log_t, p->A2[i], p->b2[i], p->Ea2[i]);
    }
    double a = kf, b = kb;
    for (int l = 0; l < ns; ++l) {
      const int nf = p->stoi_f[i * kMaxSpec + l], nb = p->stoi_b[i * kMaxSpec + l];
      if (nf) a *= ipow(c[l], nf);
      if (nb) b *= ipow(c[l], nb);
    }
    q1[i] = a;
    q2[i] = b;
  }
}

// omega_k [kg/m3/s] and optionally the EPI Jacobian (row-major m*ns+n) or the
// DA diagonal (src/FiniteRateChem.cu:155-209)
template <int NS, int NR>
__device__ __forceinline__ void chem_source(const ostc_param *__restrict__ p, int ns, int nr, double rho, double t,
                                            const double *y, double *omega, double *jac_epi, double *jac_da) {
  double q1[NR], q2[NR];
  rates_of_progress<NS, NR>(p, ns, nr, rho, t, y, q1, q2);
  for (int i = 0; i < ns; ++i) {
    double creation = 0, od = 0;
    for (int j = 0; j < nr; ++j) {
      const int nf = p->stoi_f[j * kMaxSpec + i], nb = p->stoi_b[j * kMaxSpec + i];
      creation += q2[j] * nf + q1[j] * nb;
      od += q1[j] * nf + q2[j] * nb;
    }
    creation *= p->mw[i] * 1e+3;
    od *= p->mw[i] * 1e+3;
    omega[i] = creation - od;
    if (jac_da) jac_da[i] = y[i] > 1e-25 ? -od / (y[i] * rho) : 0.0;
  }
  if (jac_epi) {
    for (int m = 0; m < ns; ++m)
      for (int n = 0; n < ns; ++n) {
        double zz = 0;
        if (y[n] > 1e-25) {
          for (int r = 0; r < nr; ++r) {
            const int dm = p->stoi_b[r * kMaxSpec + m] - p->stoi_f[r * kMaxSpec + m];
            if (dm != 0) zz += dm * (p->stoi_f[r * kMaxSpec + n] * q1[r] - p->stoi_b[r * kMaxSpec + n] * q2[r]);
          }
          zz /= rho * y[n];
        }
        jac_epi[m * ns + n] = p->mw[m] * zz * 1e+3;
      }
  }
}

// Pivoted LU factorisation in place, as solve_chem_system
// (src/FiniteRateChem.cu:286-318): Crout-style column pivoting.
template <int NS>
__device__ __forceinline__ void lu_factor(double *a, int *piv, int dim) {
  for (int n = 0; n < dim; ++n) {
    int ik = n;
    for (int m = n; m < dim; ++m) {
      double v = a[m * dim + n];
      for (int t = 0; t < n; ++t) v -= a[m * dim + t] * a[t * dim + n];
      a[m * dim + n] = v;
      if (fabs(v) > fabs(a[ik * dim + n])) ik = m;
    }
    piv[n] = ik;
    if (ik != n)
      for (int t = 0; t < dim; ++t) {
        const double mid = a[ik * dim + t];
        a[ik * dim + t] = a[n * dim + t];
        a[n * dim + t] = mid;
      }
    for (int q = n + 1; q < dim; ++q) {
      double v = a[n * dim + q];
      for (int t = 0; t < n; ++t) v -= a[n * dim + t] * a[t * dim + q];
      a[n * dim + q] = v;
    }
    for (int m = n + 1; m < dim; ++m) a[m * dim + n] /= a[n * dim + n];
  }
}

// Forward/back substitution with the factors above (src/FiniteRateChem.cu:
// 320-341 / 377-397)
template <int NS>
__device__ __forceinline__ void lu_solve(const double *a, const int *piv, double *b, int dim) {
  for (int m = 0; m < dim; ++m) {
    const int t = piv[m];
    if (t != m) {
      const double mid = b[t];
      b[t] = b[m];
      b[m] = mid;
    }
  }
  for (int m = 1; m < dim; ++m)
    for (int t = 0; t < m; ++t) b[m] -= a[m * dim + t] * b[t];
  b[dim - 1] /= a[dim * dim - 1];
  for (int m = dim - 2; m >= 0; --m) {
    for (int t = m + 1; t < dim; ++t) b[m] -= a[m * dim + t] * b[t];
    b[m] /= a[m * dim + m];
  }
}

// The plain point-implicit chemistry step at fixed (T, p):
// (I - dt*J) d(rhoY) = dt*omega, dY = d(rhoY)/rho -- the reference's explicit
// EPI (src/FiniteRateChem.cu:211-234) applied to the chemical source alone.
template <int NS>
__device__ __forceinline__ void plain_chem_step(const ostc_param *__restrict__ p, double tc, double pc,
                                                const double *yy, double dtc, double *dy) {
  double mw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) mw += yy[l] / p->mw[l];
  mw = 1 / mw;
  const double rho = pc * mw / (kRu * tc);
  double om[NS], a[NS * NS];
  int piv[NS];
  chem_source<NS, kMaxReac>(p, NS, p->n_reac, rho, tc, yy, om, a, nullptr);
#pragma unroll
  for (int q = 0; q < NS * NS; ++q) a[q] = -dtc * a[q];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    a[q * NS + q] += 1.0;
    om[q] *= dtc;
  }
  lu_factor<NS>(a, piv, NS);
  lu_solve<NS>(a, piv, om, NS);
#pragma unroll
  for (int l = 0; l < NS; ++l) dy[l] = om[l] / rho;
}

}  // namespace ostc
