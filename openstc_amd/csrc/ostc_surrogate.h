// Shared definitions of the learned chemistry surrogate: network geometry,
// parameter layout, device state, decoding of the network outputs and the
// per-cell fallback checks.  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/ostc.h"
#include "ostc_dev.h"

namespace ostc {

constexpr int kSurIn = OSTC_SUR_IN;   // 12 features
constexpr int kSurHidden = 128;
constexpr int kSurTiles = 108;        // BF16 A tiles of the LDS image (1 KB each)
constexpr int kSurBias = 528;         // 4 x 128 hidden biases + 16 output biases
constexpr int kSurSplits = 32;        // batch splits of the weight-gradient GEMMs
constexpr float kGeluC = 1.5957691216057308f;  // 2*sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLog2_10 = 3.3219280948873623f;
constexpr float kInvAtm = (float)(1.0 / 101325.0);
constexpr float kBoxCox = 0.1f;
constexpr float kInvBoxCox = 10.0f;
constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;
// The fallback list is split into kFbLists sub-lists, each with its own
// counter (and row of statistics) 128 bytes apart: wave w of the decode kernel
// appends to sub-list w % kFbLists, so the list atomics spread over many
// memory channels instead of serialising on one address.
constexpr int kFbLists = 64;
constexpr int kFbStride = 32;    // ints between two sub-list counters
constexpr int kStatStride = 16;  // u64 between two statistics rows
// capacity of one sub-list for n cells (its waves times 64 cells)
__host__ __device__ inline long fb_list_cap(long n) {
  const long waves = (n + 63) / 64;
  return (waves + kFbLists - 1) / kFbLists * 64;
}

// parameter layout: per layer W[out][in] row-major, then b[out]
__host__ __device__ inline int sur_in_dim(int l) { return l == 0 ? kSurIn : kSurHidden; }
__host__ __device__ inline int sur_out_dim(int l, int n_out) { return l == 4 ? n_out : kSurHidden; }
__host__ __device__ inline int sur_layer_offset(int l) {
  // 12*128+128 = 1664; 128*128+128 = 16512
  return l == 0 ? 0 : 1664 + (l - 1) * 16512;
}
__host__ __device__ inline int sur_n_params(int n_out) { return 1664 + 3 * 16512 + 129 * n_out; }

// device copy of ostc_surrogate_desc with reciprocal standard deviations
struct SurParams {
  int n_out, kind, n_elem, act;
  float in_mean[kSurIn], in_rstd[kSurIn];
  float out_mean[16], out_std[16];
  float elem_w[OSTC_SUR_MAX_ELEM * 9];
  float floor;
  float t_min, t_max, p_min, p_max, dt_min, dt_max;
  float y_tol, elem_tol, elem_abs;
  float dy_max;
};

struct SurState {
  SurParams params{};
  int n_params = 0;
  int cus = 256;
  long max_batch = 0;
  long adam_step = 0;
  float *master = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *wq = nullptr;
  void *packed = nullptr;  // kSurTiles * 64 lanes * 8 BF16
  float *bias = nullptr;   // kSurBias
  float *loss_w = nullptr; // 16
  int *fb_count = nullptr;                 // [kFbLists * kFbStride]
  long *fb_list = nullptr;                 // [kFbLists][fb_list_cap(fb_cap)]
  float *out12 = nullptr;  // [fb_cap][12] standardised outputs of the forward kernel
  long fb_cap = 0;
  unsigned long long *stats = nullptr;  // [kFbLists * kStatStride], 4 used per row
  // training buffers (max_batch rows)
  float *a0 = nullptr, *z = nullptr, *a = nullptr, *yout = nullptr, *g5 = nullptr;
  float *gbuf[2] = {nullptr, nullptr};
  float *part = nullptr;
  double *loss_part = nullptr, *loss_dev = nullptr;
};

// standardised network outputs -> dY (FP32)
__device__ __forceinline__ void sur_decode(const SurParams &s, const float *o, float *dy) {
  if (s.kind == 0) {
    const float lg = s.out_mean[0] + s.out_std[0] * o[0];
    const float mag = __builtin_amdgcn_exp2f(lg * kLog2_10);
#pragma unroll
    for (int k = 0; k < 9; ++k) dy[k] = (s.out_mean[k + 1] + s.out_std[k + 1] * o[k + 1]) * mag;
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) dy[k] = s.floor * sinhf(s.out_mean[k] + s.out_std[k] * o[k]);
  }
}

// 0 = accepted; 1 = input outside the trained box; 2 = Y+dY outside
// [-y_tol, 1+y_tol] or |dY_k| > dy_max; 3 = element balance
// sum_k a_jk dY_k/W_k violated.
__device__ __forceinline__ int sur_check(const SurParams &s, double T, double P, double DT, const double *Y,
                                         const float *dy) {
  bool in = (T >= s.t_min && T <= s.t_max && P >= s.p_min && P <= s.p_max && DT >= s.dt_min && DT <= s.dt_max);
#pragma unroll
  for (int k = 0; k < 9; ++k) in = in && (Y[k] >= -s.y_tol);
  if (!in) return 1;
  bool out = true;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const float yn = (float)Y[k] + dy[k];
    out = out && (yn >= -s.y_tol && yn <= 1.0f + s.y_tol) && fabsf(dy[k]) <= s.dy_max;
  }
  if (!out) return 2;
  float tot = 0.0f, worst = 0.0f;
  for (int j = 0; j < s.n_elem; ++j) {
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float q = s.elem_w[j * 9 + k] * dy[k];
      r += q;
      tot += fabsf(q);
    }
    worst = fmaxf(worst, fabsf(r));
  }
  return (worst <= s.elem_tol * tot + s.elem_abs) ? 0 : 3;
}

int sur_pack(const SurState &st, hipStream_t s);
int sur_infer_batch(const SurState &st, const ostc_param *dp, long n, const double *t, const double *p,
                    const double *y, const double *dt, double *dy, float *raw, hipStream_t s);
template <typename R>
int sur_infer_zone(const SurState &st, const ostc_param *dp, const ZoneView<R> &z, double *dy, hipStream_t s);
template <typename R>
int zone_plain_chem(const ostc_param *dp, const ZoneView<R> &z, double *dy, hipStream_t s);
int sur_gradients(const SurState &st, long nb, const double *t, const double *p, const double *y, const double *dt,
                  const int *idx, const float *target, float *grad, double *loss_dev, hipStream_t s);
int sur_adam(SurState &st, float lr, hipStream_t s);

}  // namespace ostc
