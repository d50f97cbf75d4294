// Device thermodynamics, transport and finite-rate chemistry for one point.
// T is the arithmetic type (float for the FP32 bulk path where the stage runs
// in FP32, double for chemistry and the Newton solve).
#pragma once

#include "ostc_dev.h"

namespace ostc {

// FP32 tables derived from ostc_param when a context is created, stored in
// device memory right after the parameter block (thermo32(p)): the FP32 bulk
// stages evaluate h_k, cp_k and the mixture transport from these instead of
// converting the FP64 coefficients (and dividing R by W) at every use.
// Polynomials are in Horner form with the 1/2 .. 1/5 factors and R/W_k folded
// into the coefficients (the FP64 path keeps the reference's form).
struct ThermoF32 {
  float t_low[kMaxSpec], t_mid[kMaxSpec];
  float h[2][kMaxSpec][6];    // [low/high] R/W (a5, a0, a1/2, a2/3, a3/4, a4/5): h = a5 + t(a0 + t(a1/2 + ...))
  float cp[2][kMaxSpec][5];   // [low/high] R/W (a0 .. a4)
  float h_tl[kMaxSpec], cp_tl[kMaxSpec];  // h and cp at T_low (linear extrapolation below it)
  float r_over_w[kMaxSpec], inv_w[kMaxSpec], cp_lam[kMaxSpec];  // R/W, 1/W, 1.25 R/W
  float lj_inv[kMaxSpec], vis_coeff[kMaxSpec];
  float wj_wi4[kMaxSpec * kMaxSpec], sqrt_wi_wj8[kMaxSpec * kMaxSpec];
  float t_mid_min, t_mid_max, t_low_max;  // over the species: the wave-uniform range test
};
constexpr size_t kThermoF32Offset = (sizeof(ostc_param) + 255) / 256 * 256;
__host__ __device__ __forceinline__ const ThermoF32 *thermo32(const ostc_param *p) {
  return reinterpret_cast<const ThermoF32 *>(reinterpret_cast<const char *>(p) + kThermoF32Offset);
}
// host: fill the tables of one parameter block
void make_thermo32(const ostc_param &p, ThermoF32 &out);

// FP64 chemistry tables derived from ostc_param (after ThermoF32 in the same
// device block): stoichiometric coefficients as doubles (no integer
// conversions in the rate loops), 1/W_k, and per reaction the bit masks of
// the species that take part (forward or backward coefficient nonzero) and of
// those whose net coefficient is nonzero -- wave-uniform tests that skip the
// zero terms of the production sums and of the EPI Jacobian.
// One reaction's rate constants in one record (the parameter block keeps
// them in per-quantity arrays): a wave-uniform reaction index then reads them
// from one base address with wide scalar loads.
struct ReacF64 {
  double A, b, Ea, A2, b2, Ea2;                 // Arrhenius: main, second (low-pressure / reverse)
  double troe_alpha, troe_t1, troe_t2, troe_t3;
  int type, order;                               // reaction type, reaction order (for Kc)
  unsigned part, net;                            // ChemF64::part / net of the reaction
};
struct ChemF64 {
  ReacF64 reac[kMaxReac];
  double inv_mw[kMaxSpec];
  double nu_f[kMaxReac * kMaxSpec], nu_b[kMaxReac * kMaxSpec];
  unsigned part[kMaxReac], net[kMaxReac];
  // g_k/(R T) of the NASA-7 fit per range [low/high]: (a0 - a6, -a0, -a1/2,
  // -a2/6, -a3/12, -a4/20, a5): g = c0 + c1 ln T + T(c2 + T(c3 + T(c4 + T c5)))
  // + c6 / T; below T_low the value at T_low
  double g[2][kMaxSpec][7];
  double g_low[kMaxSpec];
  // species that take part in at least one reaction, in order: act[r] is the
  // r-th of them, rank[k] species k's position among them (-1: none -- its
  // row and column of the EPI Jacobian are zero), n_act their count
  int n_act;
  int act[kMaxSpec], rank[kMaxSpec];
  unsigned act_mask;  // bit k: species k reacts (rank[k] = popcount of the bits below k)
  // per species m, the reactions whose net coefficient of m is nonzero (the
  // terms of row m of the EPI Jacobian), in order (ints, not bytes: a scalar
  // load reads whole dwords, a byte would take a vector load)
  int jn[kMaxSpec];
  int jr[kMaxSpec][kMaxReac];
};
constexpr size_t kChemF64Offset = (kThermoF32Offset + sizeof(ThermoF32) + 255) / 256 * 256;
constexpr size_t kParamBlockBytes = kChemF64Offset + sizeof(ChemF64);
__host__ __device__ __forceinline__ const ChemF64 *chem64(const ostc_param *p) {
  return reinterpret_cast<const ChemF64 *>(reinterpret_cast<const char *>(p) + kChemF64Offset);
}
void make_chem64(const ostc_param &p, ChemF64 &out);

// 1/W_k and R/W_k at the stage's precision: FP32 from the tables (the same
// rounded values the FP64 expressions give), FP64 as the reference writes them
template <typename R>
__device__ __forceinline__ R inv_mw(const ostc_param *__restrict__ p, int l) {
  if constexpr (sizeof(R) == 4)
    return thermo32(p)->inv_w[l];
  else
    return 1.0 / p->mw[l];
}
template <typename R>
__device__ __forceinline__ R r_over_mw(const ostc_param *__restrict__ p, int l) {
  if constexpr (sizeof(R) == 4)
    return thermo32(p)->r_over_w[l];
  else
    return kRu / p->mw[l];
}

// NASA-7 enthalpy h_k(T) [J/kg]; below T_low linear extrapolation with
// cp(T_low) (src/Thermo.cu:5-23).
template <typename T>
__device__ __forceinline__ void enthalpy(T t, T *h, const ostc_param *__restrict__ p, int ns);

// per-lane choice between the low- and high-range coefficients of species i
// (both rows are loaded with wave-uniform addresses, then selected)
template <int N>
__device__ __forceinline__ void pick_coeffs(const float *lo, const float *hi, bool use_hi, float *c) {
#pragma unroll
  for (int q = 0; q < N; ++q) c[q] = use_hi ? hi[q] : lo[q];
}

// 1 if every active lane of the wave is in the high range of every species,
// 0 if every one is in the low range of every species and above every T_low,
// -1 otherwise.  In the first two cases the coefficients are wave-uniform
// (scalar operands), and the per-lane selects and the T_low test drop out:
// the same Horner polynomial on the same coefficients, in another
// instruction sequence (FP32 results can differ in the last bit).
__device__ __forceinline__ int uniform_range(float t, const ThermoF32 *tf) {
  if (__all(t >= tf->t_mid_max && t >= tf->t_low_max)) return 1;
  if (__all(t < tf->t_mid_min && t >= tf->t_low_max)) return 0;
  return -1;
}

template <>
__device__ __forceinline__ void enthalpy<float>(float t, float *h, const ostc_param *__restrict__ p, int ns) {
  const ThermoF32 *tf = thermo32(p);
  const int r = uniform_range(t, tf);
  if (r >= 0) {
    for (int i = 0; i < ns; ++i) {
      const float *c = tf->h[r][i];
      h[i] = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    }
    return;
  }
  for (int i = 0; i < ns; ++i) {
    float c[6];
    pick_coeffs<6>(tf->h[0][i], tf->h[1][i], t >= tf->t_mid[i], c);
    const float v = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    h[i] = t < tf->t_low[i] ? tf->h_tl[i] + tf->cp_tl[i] * (t - tf->t_low[i]) : v;
  }
}

template <>
__device__ __forceinline__ void enthalpy<double>(double t, double *h, const ostc_param *__restrict__ p, int ns) {
  using T = double;
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
__device__ __forceinline__ void enthalpy_cp(T t, T *h, T *cp, const ostc_param *__restrict__ p, int ns);

template <>
__device__ __forceinline__ void enthalpy_cp<float>(float t, float *h, float *cp, const ostc_param *__restrict__ p,
                                                   int ns) {
  const ThermoF32 *tf = thermo32(p);
  const int r = uniform_range(t, tf);
  if (r >= 0) {
    for (int i = 0; i < ns; ++i) {
      const float *c = tf->h[r][i], *d = tf->cp[r][i];
      h[i] = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
      cp[i] = d[0] + t * (d[1] + t * (d[2] + t * (d[3] + t * d[4])));
    }
    return;
  }
  for (int i = 0; i < ns; ++i) {
    const bool hi = t >= tf->t_mid[i];
    float c[6], d[5];
    pick_coeffs<6>(tf->h[0][i], tf->h[1][i], hi, c);
    pick_coeffs<5>(tf->cp[0][i], tf->cp[1][i], hi, d);
    const float hv = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    const float cv = d[0] + t * (d[1] + t * (d[2] + t * (d[3] + t * d[4])));
    const bool lo = t < tf->t_low[i];
    h[i] = lo ? tf->h_tl[i] + tf->cp_tl[i] * (t - tf->t_low[i]) : hv;
    cp[i] = lo ? tf->cp_tl[i] : cv;
  }
}

template <>
__device__ __forceinline__ void enthalpy_cp<double>(double t, double *h, double *cp, const ostc_param *__restrict__ p,
                                                    int ns) {
  using T = double;
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
__device__ __forceinline__ void heat_capacity(T t, T *cp, const ostc_param *__restrict__ p, int ns);

template <>
__device__ __forceinline__ void heat_capacity<float>(float t, float *cp, const ostc_param *__restrict__ p, int ns) {
  const ThermoF32 *tf = thermo32(p);
  const int r = uniform_range(t, tf);
  if (r >= 0) {
    for (int i = 0; i < ns; ++i) {
      const float *d = tf->cp[r][i];
      cp[i] = d[0] + t * (d[1] + t * (d[2] + t * (d[3] + t * d[4])));
    }
    return;
  }
  for (int i = 0; i < ns; ++i) {
    float d[5];
    pick_coeffs<5>(tf->cp[0][i], tf->cp[1][i], t >= tf->t_mid[i], d);
    const float cv = d[0] + t * (d[1] + t * (d[2] + t * (d[3] + t * d[4])));
    cp[i] = t < tf->t_low[i] ? tf->cp_tl[i] : cv;
  }
}

template <>
__device__ __forceinline__ void heat_capacity<double>(double t, double *cp, const ostc_param *__restrict__ p, int ns) {
  using T = double;
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
                                                  const T *y, const T *cp, T *mu, T *lambda, T *rhoD);

template <>
__device__ __forceinline__ void mixture_transport<float>(const ostc_param *__restrict__ p, int ns, float t,
                                                         float mw_total, const float *y, const float *cp, float *mu,
                                                         float *lambda, float *rhoD) {
  const ThermoF32 *tf = thermo32(p);
  float x[kMaxSpec], vis[kMaxSpec];
  const float sqrt_t = sqrtf(t);
  for (int l = 0; l < ns; ++l) {
    x[l] = y[l] * mw_total * tf->inv_w[l];
    const float t_dl = t * tf->lj_inv[l];
    const float ci = 1.147f * powf(t_dl, -0.145f) + 1.0f / ((t_dl + 0.5f) * (t_dl + 0.5f));
    vis[l] = tf->vis_coeff[l] * sqrt_t / ci;
  }
  // sqrt(mu_m / mu_n) as sqrt(mu_m) * (1 / sqrt(mu_n)): one square root and
  // one reciprocal per species instead of one of each per pair
  float sq[kMaxSpec], isq[kMaxSpec];
  for (int l = 0; l < ns; ++l) {
    sq[l] = sqrtf(vis[l]);
    isq[l] = 1.0f / sq[l];
  }
  float viscosity = 0, conductivity = 0;
  for (int m = 0; m < ns; ++m) {
    float vis_temp = 0;
    for (int n = 0; n < ns; ++n) {
      float pf = 1;
      if (m != n) {
        const float num = 1.0f + sq[m] * isq[n] * tf->wj_wi4[m * kMaxSpec + n];
        pf = num * num * tf->sqrt_wi_wj8[m * kMaxSpec + n];
      }
      vis_temp += pf * x[n];
    }
    viscosity += vis[m] * x[m] / vis_temp;
    if (lambda) {
      const float cond_temp = 1.065f * vis_temp - 0.065f * x[m];
      const float lam = vis[m] * (cp[m] + tf->cp_lam[m]);
      conductivity += lam * x[m] / cond_temp;
    }
  }
  *mu = viscosity;
  if (lambda) *lambda = conductivity;
  if (rhoD) {
    const float sc = (float)p->Sc;
    for (int l = 0; l < ns; ++l) {
      if (fabsf(x[l] - 1.0f) < 1e-3f)
        rhoD[l] = viscosity / sc;
      else
        rhoD[l] = (1.0f - y[l]) * viscosity / ((1.0f - x[l]) * sc);
    }
  }
}

template <>
__device__ __forceinline__ void mixture_transport<double>(const ostc_param *__restrict__ p, int ns, double t,
                                                          double mw_total, const double *y, const double *cp,
                                                          double *mu, double *lambda, double *rhoD) {
  using T = double;
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

__device__ __forceinline__ double arrhenius(double t_inv, double log_t, double A, double b, double Ea) {
  // A*T^b*exp(-Ea/T) evaluated as one exp of the combined exponent, with
  // 1/T formed once per cell (b==0 skips the log term);
  // src/FiniteRateChem.cu:110-112
  return b == 0.0 ? A * exp(-Ea * t_inv) : A * exp(b * log_t - Ea * t_inv);
}

// g_k/(R T) (src/Thermo.cu:63-79, evaluated at T_low below T_low), from the
// derived coefficients (no divisions; ln T and 1/T from the caller)
template <int NS>
__device__ __forceinline__ void gibbs_div_rt(double t, double log_t, double t_inv, const ostc_param *__restrict__ p,
                                             double *g) {
  const ChemF64 *ct = chem64(p);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const bool hi = t >= p->t_mid[i];
    double c[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) c[q] = hi ? ct->g[1][i][q] : ct->g[0][i][q];
    const double v = c[0] + c[1] * log_t + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))) + c[6] * t_inv;
    g[i] = t < p->t_low[i] ? ct->g_low[i] : v;
  }
}

// (int)x for a non-negative integer x < 2^52 held in a double (the
// stoichiometric coefficients), from its bit pattern: integer operations
// only, so a wave-uniform x stays in scalar registers (a float-to-int
// conversion would run on the vector unit)
__device__ __forceinline__ int small_int_of(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  if (u == 0) return 0;
  const int e = (int)(u >> 52) - 1023;  // x = 1.f * 2^e, 0 <= e < 52
  return (int)(((u & 0xFFFFFFFFFFFFFull) | 0x10000000000000ull) >> (52 - e));
}

// Forward / backward rate of progress (mol/cm3/s) of reaction i at
// concentrations c (mol/cm3) and g_k/RT: one iteration of
// src/FiniteRateChem.cu:7-153.  nf/nb: the reaction's stoichiometric
// coefficients as doubles, part: the species that take part (all uniform
// across the wavefront).
template <int NS>
__device__ __forceinline__ void reaction_rates(const ostc_param *__restrict__ p, int i, double t, double t_inv,
                                               double log_t, double temp_t, double temp_t_inv, const double *c,
                                               const double *g, const double *nf, const double *nb, unsigned part,
                                               double &q1, double &q2) {
  const ReacF64 &rc = chem64(p)->reac[i];
  const int type = rc.type;
  double kf = arrhenius(t_inv, log_t, rc.A, rc.b, rc.Ea);
  if (type == 3) kf += arrhenius(t_inv, log_t, rc.A2, rc.b2, rc.Ea2);
  if (type > 3) {
    double cc = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cc += c[l] * p->third_body_coeff[i * kMaxSpec + l];
    if (type == 4) {
      kf *= cc;
    } else {
      const double kf_low = arrhenius(t_inv, log_t, rc.A2, rc.b2, rc.Ea2);
      const double kf_high = kf;
      const double pr = kf_low * cc / kf_high;
      double fv = 1.0;
      if (type > 5) {
        double f_cent = (1 - rc.troe_alpha) * exp(-t / rc.troe_t3) + rc.troe_alpha * exp(-t / rc.troe_t1);
        if (type == 7) f_cent += exp(-rc.troe_t2 * t_inv);
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
    for (int l = 0; l < NS; ++l)
      if (part & (1u << l)) d_gibbs += g[l] * (nb[l] - nf[l]);
    const int order = rc.order;
    // kb = kf / Kc with Kc = temp_t^order * exp(-d_gibbs)
    const double base = order >= 0 ? temp_t_inv : temp_t;
    double pw = 1.0;
    for (int q = 0; q < (order >= 0 ? order : -order); ++q) pw *= base;
    kb = kf * exp(d_gibbs) * pw;
  } else {
    kb = arrhenius(t_inv, log_t, rc.A2, rc.b2, rc.Ea2);
  }
  double a = kf, b = kb;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (part & (1u << l)) {
      const int ef = small_int_of(nf[l]), eb = small_int_of(nb[l]);
      for (int q = 0; q < ef; ++q) a *= c[l];
      for (int q = 0; q < eb; ++q) b *= c[l];
    }
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
  const ChemF64 *ct = chem64(p);
  double c[NS], g[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = rho * y[l] * ct->inv_mw[l] * 1e-3;
  const double log_t = log(t), t_inv = 1.0 / t;
  gibbs_div_rt<NS>(t, log_t, t_inv, p, g);
  const double temp_t = kPatm / kRu * 1e-3 * t_inv, temp_t_inv = 1.0 / temp_t;
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
    double nf[NS], nb[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      nf[l] = ct->nu_f[r * kMaxSpec + l];
      nb[l] = ct->nu_b[r * kMaxSpec + l];
    }
    const unsigned part = ct->reac[r].part, net = ct->reac[r].net;
    double q1, q2;
    reaction_rates<NS>(p, r, t, t_inv, log_t, temp_t, temp_t_inv, c, g, nf, nb, part, q1, q2);
    (void)part;
    // dense over the species (terms with zero coefficients add exact zeros;
    // wave-uniform branches per species measured slower here)
    double v[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      creation[i] += q2 * nf[i] + q1 * nb[i];
      od[i] += q1 * nf[i] + q2 * nb[i];
      v[i] = nf[i] * q1 - nb[i] * q2;
    }
    if constexpr (JAC == 1) {
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        if (net & (1u << m)) {
          const double dm = nb[m] - nf[m];
#pragma unroll
          for (int n = 0; n < NS; ++n) zz[m * NS + n] += dm * v[n];
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
      const double rry = on ? 1.0 / (rho * y[n]) : 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) zz[m * NS + n] = on ? p->mw[m] * (zz[m * NS + n] * rry) * 1e+3 : 0.0;
    }
  }
}

// The same production rates and Jacobian in two passes: first every
// reaction's rates of progress, kept in LDS (q1 of reaction r at
// qs[r * stride], q2 at qs[(nr_max + r) * stride]; stride = the workgroup's
// thread count, qs already offset to this thread), then omega and the
// Jacobian.  Every loop over the reactions is a real loop with one copy of
// the rate law (rates of progress in registers need the reaction loops
// unrolled: 20 copies of the rate law, ~0.5 MB of code through a 64 KB
// instruction cache).  Sums run over the reactions in the same order as the
// one-pass form and, as it does, over all species (one that takes no part in
// a reaction adds exact zeros).  Up to 9 species the Jacobian is summed
// reaction-major with the whole matrix in registers (the kernels that call
// this leave 256 VGPRs per lane); above 9, row by row, row m visiting only
// the reactions that change species m (ChemF64::jr).  JAC: 1 EPI rows, each
// finished row to row_out(m, row); 3 (up to 9 species) the whole EPI matrix,
// row-major, in one call row_out(0, jac) that may overwrite it; 2 DA
// diagonal (row_out(m, &d) with the diagonal entry); 0 none.
// omega_out(k, omega_k) receives the production rates.
// JAC == 4 of chem_source_lds (up to 9 species): the EPI matrix first, to
// row_out(0, jac) (which may factor it in place), then the production rates
// to omega_out, then row_out(1, jac) with the matrix row_out(0) left: the
// caller keeps the block's factors in registers across the production sums
// (the same sums in the same order as the other modes).
template <int NS, class RowOut, class OmegaOut>
__device__ __forceinline__ void chem_jac_block(const ostc_param *__restrict__ p, const ChemF64 *ct, int nr,
                                               int nr_max, double rho, const double *y, const double *qs, int stride,
                                               OmegaOut omega_out, RowOut row_out) {
  static_assert(NS <= 9, "the whole matrix in registers up to 9 species");
  double zz[NS * NS];
  {
    double rry[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) rry[n] = y[n] > 1e-25 ? 1.0 / (rho * y[n]) : 0.0;
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) zz[q] = 0;
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      const double q1 = qs[r * stride], q2 = qs[(nr_max + r) * stride];
      const unsigned net = ct->reac[r].net;
      double nf[NS], nb[NS], v[NS];
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        nf[n] = ct->nu_f[r * kMaxSpec + n];
        nb[n] = ct->nu_b[r * kMaxSpec + n];
        v[n] = nf[n] * q1 - nb[n] * q2;
      }
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        if (net & (1u << m)) {
          const double dm = nb[m] - nf[m];
#pragma unroll
          for (int n = 0; n < NS; ++n) zz[m * NS + n] += dm * v[n];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NS; ++m) {
#pragma unroll
      for (int n = 0; n < NS; ++n)
        zz[m * NS + n] = y[n] > 1e-25 ? p->mw[m] * (zz[m * NS + n] * rry[n]) * 1e+3 : 0.0;
    }
  }
  row_out(0, zz);
  {
    double cr[NS], d[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) cr[i] = d[i] = 0;
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      const double q1 = qs[r * stride], q2 = qs[(nr_max + r) * stride];
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const double nf = ct->nu_f[r * kMaxSpec + i], nb = ct->nu_b[r * kMaxSpec + i];
        cr[i] += q2 * nf + q1 * nb;
        d[i] += q1 * nf + q2 * nb;
      }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) omega_out(i, cr[i] * (p->mw[i] * 1e+3) - d[i] * (p->mw[i] * 1e+3));
  }
  row_out(1, zz);
}

template <int NS, int JAC, class RowOut, class OmegaOut>
__device__ __forceinline__ void chem_source_lds(const ostc_param *__restrict__ p, int nr, int nr_max, double rho,
                                                double t, const double *y, double *qs, int stride,
                                                OmegaOut omega_out, RowOut row_out) {
  const ChemF64 *ct = chem64(p);
  {
    double c[NS], g[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) c[l] = rho * y[l] * ct->inv_mw[l] * 1e-3;
    const double log_t = log(t), t_inv = 1.0 / t;
    gibbs_div_rt<NS>(t, log_t, t_inv, p, g);
    const double temp_t = kPatm / kRu * 1e-3 * t_inv, temp_t_inv = 1.0 / temp_t;
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      double nf[NS], nb[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        nf[l] = ct->nu_f[r * kMaxSpec + l];
        nb[l] = ct->nu_b[r * kMaxSpec + l];
      }
      double q1, q2;
      reaction_rates<NS>(p, r, t, t_inv, log_t, temp_t, temp_t_inv, c, g, nf, nb, ct->reac[r].part, q1, q2);
      qs[r * stride] = q1;
      qs[(nr_max + r) * stride] = q2;
    }
  }
  if constexpr (JAC == 4) {
    chem_jac_block<NS>(p, ct, nr, nr_max, rho, y, qs, stride, omega_out, row_out);
    return;
  }
  {
    // dense over the species, as the one-pass form: a species that takes no
    // part in reaction r adds exact zeros (no per-species branches)
    double cr[NS], d[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) cr[i] = d[i] = 0;
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      const double q1 = qs[r * stride], q2 = qs[(nr_max + r) * stride];
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const double nf = ct->nu_f[r * kMaxSpec + i], nb = ct->nu_b[r * kMaxSpec + i];
        cr[i] += q2 * nf + q1 * nb;
        d[i] += q1 * nf + q2 * nb;
      }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const double crv = cr[i] * (p->mw[i] * 1e+3), dv = d[i] * (p->mw[i] * 1e+3);
      omega_out(i, crv - dv);
      if constexpr (JAC == 2) {
        const double jd = y[i] > 1e-25 ? -dv / (y[i] * rho) : 0.0;
        row_out(i, &jd);
      }
    }
  }
  static_assert(JAC != 3 || NS <= 9, "the whole-matrix output is for up to 9 species");
  if constexpr ((JAC == 1 || JAC == 3) && NS <= 9) {
    // reaction-major: each reaction's rates and coefficient rows are read
    // once and added to every row m it changes, all NS*NS sums in
    // registers; per entry the same terms in the same (reaction) order as
    // the row-major form below
    double rry[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) rry[n] = y[n] > 1e-25 ? 1.0 / (rho * y[n]) : 0.0;
    double zz[NS * NS];
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) zz[q] = 0;
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      const double q1 = qs[r * stride], q2 = qs[(nr_max + r) * stride];
      const unsigned net = ct->reac[r].net;
      double nf[NS], nb[NS], v[NS];
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        nf[n] = ct->nu_f[r * kMaxSpec + n];
        nb[n] = ct->nu_b[r * kMaxSpec + n];
        v[n] = nf[n] * q1 - nb[n] * q2;
      }
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        if (net & (1u << m)) {
          const double dm = nb[m] - nf[m];
#pragma unroll
          for (int n = 0; n < NS; ++n) zz[m * NS + n] += dm * v[n];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NS; ++m) {
#pragma unroll
      for (int n = 0; n < NS; ++n)
        zz[m * NS + n] = y[n] > 1e-25 ? p->mw[m] * (zz[m * NS + n] * rry[n]) * 1e+3 : 0.0;
    }
    if constexpr (JAC == 3) {
      row_out(0, zz);
    } else {
#pragma unroll
      for (int m = 0; m < NS; ++m) row_out(m, zz + m * NS);
    }
  } else if constexpr (JAC == 1) {
    double rry[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) rry[n] = y[n] > 1e-25 ? 1.0 / (rho * y[n]) : 0.0;
#pragma unroll 1
    for (int m = 0; m < NS; ++m) {
      double row[NS];
#pragma unroll
      for (int n = 0; n < NS; ++n) row[n] = 0;
      // the reactions that change species m, dense over the columns
      const int jn = ct->jn[m];
#pragma unroll 1
      for (int q = 0; q < jn; ++q) {
        const int r = ct->jr[m][q];
        const double dm = ct->nu_b[r * kMaxSpec + m] - ct->nu_f[r * kMaxSpec + m];
        const double q1 = qs[r * stride], q2 = qs[(nr_max + r) * stride];
#pragma unroll
        for (int n = 0; n < NS; ++n) row[n] += dm * (ct->nu_f[r * kMaxSpec + n] * q1 - ct->nu_b[r * kMaxSpec + n] * q2);
      }
#pragma unroll
      for (int n = 0; n < NS; ++n) row[n] = y[n] > 1e-25 ? p->mw[m] * (row[n] * rry[n]) * 1e+3 : 0.0;
      row_out(m, row);
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

// The same factorisation and solve with run-time loops over an array of
// stride n in private memory, for blocks too large to unroll into registers
// (the 20-species bucket): the same operations in the same order.
__device__ inline void lu_factor_rt(double *a, int *piv, int n) {
  for (int k = 0; k < n; ++k) {
    int ik = k;
    double best = 0;
    for (int m = k; m < n; ++m) {
      double v = a[m * n + k];
      for (int t = 0; t < k; ++t) v -= a[m * n + t] * a[t * n + k];
      a[m * n + k] = v;
      if (m == k) {
        best = fabs(v);
      } else if (fabs(v) > best) {
        best = fabs(v);
        ik = m;
      }
    }
    piv[k] = ik;
    if (ik != k)
      for (int t = 0; t < n; ++t) {
        const double x = a[ik * n + t];
        a[ik * n + t] = a[k * n + t];
        a[k * n + t] = x;
      }
    for (int q = k + 1; q < n; ++q) {
      double v = a[k * n + q];
      for (int t = 0; t < k; ++t) v -= a[k * n + t] * a[t * n + q];
      a[k * n + q] = v;
    }
    for (int m = k + 1; m < n; ++m) a[m * n + k] /= a[k * n + k];
  }
}
__device__ inline void lu_solve_rt(const double *a, const int *piv, double *b, int n) {
  for (int m = 0; m < n; ++m) {
    const int t = piv[m];
    if (t != m) {
      const double x = b[t];
      b[t] = b[m];
      b[m] = x;
    }
  }
  for (int m = 1; m < n; ++m)
    for (int t = 0; t < m; ++t) b[m] -= a[m * n + t] * b[t];
  b[n - 1] /= a[n * n - 1];
  for (int m = n - 2; m >= 0; --m) {
    for (int t = m + 1; t < n; ++t) b[m] -= a[m * n + t] * b[t];
    b[m] /= a[m * n + m];
  }
}

// LU of an NS x NS block held in a[NS * NS]: unrolled into registers up to 9
// species, run-time loops beyond
template <int NS>
__device__ __forceinline__ void lu_factor_any(double *a, int *piv) {
  if constexpr (NS > 9)
    lu_factor_rt(a, piv, NS);
  else
    lu_factor<NS>(a, piv, NS);
}
template <int NS>
__device__ __forceinline__ void lu_solve_any(const double *a, const int *piv, double *b) {
  if constexpr (NS > 9)
    lu_solve_rt(a, piv, b, NS);
  else
    lu_solve<NS>(a, piv, b, NS);
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
  // the one-pass rates: the two-pass form would hold the rates of every
  // reaction beside this whole matrix (2.1-3.2 ms against 1.2 ms measured
  // for 1.77 M cells)
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
