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
template <int NS>
__device__ __forceinline__ void gibbs_div_rt(double t, const ostc_param *__restrict__ p, double *g) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
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

// Forward / backward rate of progress (mol/cm3/s) of reaction i at
// concentrations c (mol/cm3) and g_k/RT: one iteration of
// src/FiniteRateChem.cu:7-153.  nf/nb: the reaction's stoichiometric
// coefficients (uniform across the wavefront).
template <int NS>
__device__ __forceinline__ void reaction_rates(const ostc_param *__restrict__ p, int i, double t, double log_t,
                                               double temp_t, const double *c, const double *g, const int *nf,
                                               const int *nb, double &q1, double &q2) {
  const int type = p->reac_type[i];
  double kf = arrhenius(t, log_t, p->A[i], p->b[i], p->Ea[i]);
  if (type == 3) kf += arrhenius(t, log_t, p->A2[i], p->b2[i], p->Ea2[i]);
  if (type > 3) {
    double cc = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cc += c[l] * p->third_body_coeff[i * kMaxSpec + l];
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
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const int dn = nb[l] - nf[l];
      if (dn != 0) d_gibbs += g[l] * dn;
    }
    const int order = p->reac_order[i];
    // kb = kf / Kc with Kc = temp_t^order * exp(-d_gibbs)
    kb = kf * exp(d_gibbs) / ipow(order >= 0 ? temp_t : 1.0 / temp_t, order >= 0 ? order : -order);
  } else {
    kb = arrhenius(t, log_t, p->A2[i], p->b2[i], p->Ea2[i]);
  }
  double a = kf, b = kb;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (nf[l]) a *= ipow(c[l], nf[l]);
    if (nb[l]) b *= ipow(c[l], nb[l]);
  }
  q1 = a;
  q2 = b;
}

// omega_k [kg/m3/s] and optionally the EPI Jacobian (row-major m*NS+n) or the
// DA diagonal (src/FiniteRateChem.cu:155-209).  One pass over the reactions:
// each reaction's q1/q2 is formed and immediately added to the production
// sums and the Jacobian (same summation order over reactions as the
// reference), so no per-reaction array is kept and every per-species array
// has compile-time indices (registers, no scratch).  JAC selects the
// Jacobian at compile time: 0 none, 1 EPI (jac = NS*NS matrix), 2 DA
// diagonal (jac = NS values).  (A run-time null test on a pointer to a local
// array would keep that array in private memory: in the private address
// space null is a valid address, so the comparison cannot be folded.)
template <int NS, int NR, int JAC>
__device__ __forceinline__ void chem_source(const ostc_param *__restrict__ p, int ns, int nr, double rho, double t,
                                            const double *y, double *omega, double *jac) {
  double *jac_epi = JAC == 1 ? jac : nullptr;
  double *jac_da = JAC == 2 ? jac : nullptr;
  (void)ns;
  double c[NS], g[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = rho * y[l] / p->mw[l] * 1e-3;
  gibbs_div_rt<NS>(t, p, g);
  const double log_t = log(t);
  const double temp_t = kPatm / kRu * 1e-3 / t;
  // the Jacobian sums accumulate in the caller's array and are scaled in place
  double creation[NS], od[NS];
  double *zz = jac_epi;
#pragma unroll
  for (int l = 0; l < NS; ++l) creation[l] = od[l] = 0;
  if constexpr (JAC == 1) {
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) zz[q] = 0;
  }
  for (int r = 0; r < nr && r < NR; ++r) {
    int nf[NS], nb[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      nf[l] = p->stoi_f[r * kMaxSpec + l];
      nb[l] = p->stoi_b[r * kMaxSpec + l];
    }
    double q1, q2;
    reaction_rates<NS>(p, r, t, log_t, temp_t, c, g, nf, nb, q1, q2);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      creation[i] += q2 * nf[i] + q1 * nb[i];
      od[i] += q1 * nf[i] + q2 * nb[i];
    }
    if constexpr (JAC == 1) {
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const int dm = nb[m] - nf[m];
        if (dm != 0) {
#pragma unroll
          for (int n = 0; n < NS; ++n) zz[m * NS + n] += dm * (nf[n] * q1 - nb[n] * q2);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double cr = creation[i] * (p->mw[i] * 1e+3), d = od[i] * (p->mw[i] * 1e+3);
    omega[i] = cr - d;
    if constexpr (JAC == 2) jac_da[i] = y[i] > 1e-25 ? -d / (y[i] * rho) : 0.0;
  }
  if constexpr (JAC == 1) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const bool on = y[n] > 1e-25;
      const double ry = rho * y[n];
#pragma unroll
      for (int m = 0; m < NS; ++m) zz[m * NS + n] = on ? p->mw[m] * (zz[m * NS + n] / ry) * 1e+3 : 0.0;
    }
  }
}

// In-place LU factorisation with partial pivoting of the reference
// (src/FiniteRateChem.cu:286-318): Crout order, pivot = first row of maximal
// |column entry|.  Row exchanges are written as selects over compile-time rows
// so the matrix stays in registers.
template <int NS>
__device__ __forceinline__ void lu_factor(double *a, int *piv, int dim) {
  (void)dim;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    int ik = n;
    double best = 0;
#pragma unroll
    for (int m = n; m < NS; ++m) {
      double v = a[m * NS + n];
#pragma unroll
      for (int t = 0; t < n; ++t) v -= a[m * NS + t] * a[t * NS + n];
      a[m * NS + n] = v;
      if (m == n) {
        best = fabs(v);
      } else if (fabs(v) > best) {
        best = fabs(v);
        ik = m;
      }
    }
    piv[n] = ik;
#pragma unroll
    for (int m = n + 1; m < NS; ++m) {
      const bool sw = ik == m;
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        const double x = a[m * NS + t], y = a[n * NS + t];
        a[m * NS + t] = sw ? y : x;
        a[n * NS + t] = sw ? x : y;
      }
    }
#pragma unroll
    for (int q = n + 1; q < NS; ++q) {
      double v = a[n * NS + q];
#pragma unroll
      for (int t = 0; t < n; ++t) v -= a[n * NS + t] * a[t * NS + q];
      a[n * NS + q] = v;
    }
#pragma unroll
    for (int m = n + 1; m < NS; ++m) a[m * NS + n] /= a[n * NS + n];
  }
}

// Forward/back substitution with the factors above (src/FiniteRateChem.cu:
// 320-341 / 377-397); the row exchanges as selects.
template <int NS>
__device__ __forceinline__ void lu_solve(const double *a, const int *piv, double *b, int dim) {
  (void)dim;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    const int t = piv[m];
#pragma unroll
    for (int r = m + 1; r < NS; ++r) {
      const bool sw = t == r;
      const double x = b[r], y = b[m];
      b[r] = sw ? y : x;
      b[m] = sw ? x : y;
    }
  }
#pragma unroll
  for (int m = 1; m < NS; ++m)
#pragma unroll
    for (int t = 0; t < m; ++t) b[m] -= a[m * NS + t] * b[t];
  b[NS - 1] /= a[NS * NS - 1];
#pragma unroll
  for (int m = NS - 2; m >= 0; --m) {
#pragma unroll
    for (int t = m + 1; t < NS; ++t) b[m] -= a[m * NS + t] * b[t];
    b[m] /= a[m * NS + m];
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
  chem_source<NS, kMaxReac, 1>(p, NS, p->n_reac, rho, tc, yy, om, a);
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
