/*
 * FP64 oracle of the learned chemistry surrogate (TEST INFRASTRUCTURE ONLY;
 * see ostc_oracle.h).  The network, its features and its fallback checks have
 * no counterpart in the reference; the label it learns is the plain
 * point-implicit step of orc_chem.c.
 *
 * Parameter layout: per layer W[out][in] row-major, then b[out]; layers
 * 12->128, 128->128 (x3), 128->n_out.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ostc_oracle.h"

#define NH 128
#define NL 5
static const double GELU_C = 1.5957691216057308; /* 2*sqrt(2/pi) */
static const double GELU_A = 0.044715;

static int in_dim(int l) { return l == 0 ? 12 : NH; }
static int out_dim(int l, int n_out) { return l == 4 ? n_out : NH; }
static int layer_off(int l) { return l == 0 ? 0 : 1664 + (l - 1) * 16512; }

int orc_sur_n_params(int n_out) { return 1664 + 3 * 16512 + 129 * n_out; }

/* float(x) rounded to the nearest BF16, ties to even */
double orc_bf16(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (double)f; /* inf / nan */
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

static double gelu(double z) {
  const double u = GELU_C * (z + GELU_A * z * z * z);
  return z / (1.0 + exp(-u));
}
static double gelu_grad(double z) {
  const double u = GELU_C * (z + GELU_A * z * z * z);
  const double s = 1.0 / (1.0 + exp(-u));
  return s + z * s * (1.0 - s) * GELU_C * (1.0 + 3.0 * GELU_A * z * z);
}
/* hard-swish z*clamp(z/6 + 1/2, 0, 1); derivative (2z + 3)/6 inside (-3, 3) */
static double hswish(double z) {
  const double c = z / 6.0 + 0.5;
  return z * (c < 0.0 ? 0.0 : c > 1.0 ? 1.0 : c);
}
static double hswish_grad(double z) {
  if (z <= -3.0) return 0.0;
  if (z >= 3.0) return 1.0;
  return (2.0 * z + 3.0) / 6.0;
}
static double act(const orc_sur_desc *d, double z) { return d->activation == 1 ? hswish(z) : gelu(z); }
static double act_grad(const orc_sur_desc *d, double z) {
  return d->activation == 1 ? hswish_grad(z) : gelu_grad(z);
}

void orc_sur_features(const orc_sur_desc *d, double t, double p, const double *y, double dt, float *f) {
  const float inv_atm = (float)(1.0 / 101325.0);
  float v[12];
  v[0] = 1000.0f / (float)t;
  v[1] = logf((float)p * inv_atm);
  for (int k = 0; k < 9; ++k) {
    const float yk = fmaxf((float)y[k], 0.0f);
    v[2 + k] = (powf(yk, 0.1f) - 1.0f) * 10.0f;
  }
  v[11] = logf((float)dt);
  for (int q = 0; q < 12; ++q) {
    const float r = (float)(1.0 / (double)d->in_std[q]);
    f[q] = (v[q] - d->in_mean[q]) * r;
  }
}

/* forward of one sample; keeps pre-activations z[l][128] and layer inputs
 * h[l][128] (h[0] = features) for the backward */
static void forward_one(const orc_sur_desc *d, const float *w, int quant, const float *feat, double h[NL][NH],
                        double z[NL][NH], double *out) {
  for (int q = 0; q < 12; ++q) h[0][q] = quant ? orc_bf16(feat[q]) : (double)feat[q];
  for (int l = 0; l < NL; ++l) {
    const int ki = in_dim(l), mo = out_dim(l, d->n_out);
    const float *W = w + layer_off(l), *b = W + (size_t)mo * ki;
    for (int m = 0; m < mo; ++m) {
      double s = 0;
      for (int k = 0; k < ki; ++k) {
        const double wk = quant ? orc_bf16(W[m * ki + k]) : (double)W[m * ki + k];
        s += wk * h[l][k];
      }
      s += b[m];
      if (l < NL - 1) {
        z[l][m] = s;
        const double a = act(d, s);
        h[l + 1][m] = quant ? orc_bf16(a) : a;
      } else {
        out[m] = s;
      }
    }
  }
}

void orc_sur_forward(const orc_sur_desc *d, const float *w, int quant, long n, const double *t, const double *p,
                     const double *y, const double *dt, double *out) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    float feat[12];
    double h[NL][NH], z[NL][NH], o[16] = {0};
    orc_sur_features(d, t[c], p[c], &y[c * 9], dt[c], feat);
    forward_one(d, w, quant, feat, h, z, o);
    for (int q = 0; q < 16; ++q) out[c * 16 + q] = q < d->n_out ? o[q] : 0.0;
  }
}

double orc_sur_loss_grad(const orc_sur_desc *d, const float *w, int quant, long n, const double *t, const double *p,
                         const double *y, const double *dt, const float *target, double *grad) {
  const int np = orc_sur_n_params(d->n_out);
  memset(grad, 0, sizeof(double) * np);
  double loss = 0;
  double (*h)[NH] = malloc(sizeof(double) * NL * NH);
  double (*z)[NH] = malloc(sizeof(double) * NL * NH);
  double g[NH], gp[NH];
  for (long c = 0; c < n; ++c) {
    float feat[12];
    double o[16] = {0};
    orc_sur_features(d, t[c], p[c], &y[c * 9], dt[c], feat);
    forward_one(d, w, quant, feat, h, z, o);
    for (int q = 0; q < d->n_out; ++q) {
      const double e = o[q] - (double)target[c * 16 + q];
      loss += d->loss_w[q] * e * e / (double)n;
      g[q] = 2.0 * d->loss_w[q] * e / (double)n;
    }
    for (int l = NL - 1; l >= 0; --l) {
      const int ki = in_dim(l), mo = out_dim(l, d->n_out);
      const float *W = w + layer_off(l);
      double *GW = grad + layer_off(l), *Gb = GW + (size_t)mo * ki;
      for (int m = 0; m < mo; ++m) {
        for (int k = 0; k < ki; ++k) GW[m * ki + k] += g[m] * h[l][k];
        Gb[m] += g[m];
      }
      if (l == 0) break;
      for (int k = 0; k < ki; ++k) {
        double s = 0;
        for (int m = 0; m < mo; ++m) {
          const double wk = quant ? orc_bf16(W[m * ki + k]) : (double)W[m * ki + k];
          s += wk * g[m];
        }
        gp[k] = s * act_grad(d, z[l - 1][k]);
      }
      memcpy(g, gp, sizeof(double) * ki);
    }
  }
  free(h);
  free(z);
  return loss;
}

int orc_sur_decode_check(const orc_sur_desc *d, const float *o, double t, double p, const double *y, double dt,
                         double *dy) {
  if (d->kind == 0) {
    const double lg = (double)d->out_mean[0] + (double)d->out_std[0] * o[0];
    const double mag = pow(10.0, lg);
    for (int k = 0; k < 9; ++k) dy[k] = ((double)d->out_mean[k + 1] + (double)d->out_std[k + 1] * o[k + 1]) * mag;
  } else {
    for (int k = 0; k < 9; ++k) dy[k] = d->floor * sinh((double)d->out_mean[k] + (double)d->out_std[k] * o[k]);
  }
  int in = t >= d->t_min && t <= d->t_max && p >= d->p_min && p <= d->p_max && dt >= d->dt_min && dt <= d->dt_max;
  for (int k = 0; k < 9; ++k) in = in && y[k] >= -d->y_tol;
  if (!in) return 1;
  for (int k = 0; k < 9; ++k) {
    const double yn = y[k] + dy[k];
    if (!(yn >= -d->y_tol && yn <= 1.0 + d->y_tol && fabs(dy[k]) <= d->dy_max)) return 2;
  }
  double tot = 0, worst = 0;
  for (int j = 0; j < d->n_elem; ++j) {
    double r = 0;
    for (int k = 0; k < 9; ++k) {
      const double q = d->elem_w[j * 9 + k] * dy[k];
      r += q;
      tot += fabs(q);
    }
    if (fabs(r) > worst) worst = fabs(r);
  }
  return worst <= d->elem_tol * tot + d->elem_abs ? 0 : 3;
}
