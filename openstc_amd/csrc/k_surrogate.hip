// Learned chemistry surrogate (no counterpart in the reference): a
// 12-128-128-128-128-n_out MLP that maps a chemistry state (T, p, Y_k, dt) to
// the plain point-implicit step's dY_k.  Hidden activation: GELU (tanh form)
// or hard-swish z*clamp(z/6 + 1/2, 0, 1), a compile-time parameter of the
// kernels chosen by the model's descriptor; hard-swish is two VALU operations
// per unit against GELU's four plus two transcendentals, which dominate the
// non-MFMA work of the forward pass.
//
// Inference: BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
// The network is evaluated in the transposed form Y^T = W X^T: a weight tile
// is the A operand (read from LDS, where the whole BF16 network lives, 108 KB)
// and a 16-cell slab of activations is the B operand.  The C layout of one
// layer (lane holds units 4g..4g+3 of a 16-unit tile for cell lane%16) is
// re-used directly as the B operand of the next layer: the contraction order
// inside each 32-wide k chunk is permuted consistently in the packed weights,
// so activations never leave registers between layers.
//
// The forward kernel writes the standardised outputs straight from the MFMA
// accumulators; a second, high-occupancy kernel (one thread per cell) decodes
// them, runs the checks and stores dY.  In the fused form the decode's global
// loads, stores and list atomics stalled the low-occupancy MLP waves for as
// long as the MLP itself.  Cells that fail the input-range, output-range or
// element-balance checks are appended to a list and re-done by the FP64 plain
// integrator in a third kernel (no serialisation inside a wavefront).
//
// Training: the same forward kernel (saving pre-activations and BF16-rounded
// activations), a weighted-MSE gradient kernel, FP32-MFMA
// (v_mfma_f32_16x16x4f32) GEMMs for the input and weight gradients
// (straight-through for the BF16 rounding), deterministic split reductions
// and Adam on FP32 master weights.
#include <algorithm>
#include <cmath>

#include "ostc_kernels.h"
#include "ostc_surrogate.h"
#include "ostc_thermo.h"

namespace ostc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNs = 9;
constexpr int kWaves = 12;             // waves per workgroup of the fused kernel (3 per SIMD)
constexpr int kTiles = 4;              // 16-cell slabs per wave pass (64 cells)
constexpr int kStageLd = 9;            // padded LDS staging row: 8 values per cell

// LDS hand-off between the lanes of one wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// GELU, tanh form written as z*sigmoid(u), u = 2*sqrt(2/pi)*(z + 0.044715 z^3):
// exp(-u) = 2^(z*(C + C*0.044715*z^2)) with C = -2*sqrt(2/pi)*log2(e), so one
// activation is 5 VALU operations and two transcendentals.
constexpr float kGeluE = -2.3022081983490023f;             // -2*sqrt(2/pi)*log2(e)
constexpr float kGeluE3 = -2.3022081983490023f * 0.044715f;
__device__ __forceinline__ float gelu_sig(float z) {
  const float q = fmaf(z * z, kGeluE3, kGeluE);
  return fast_rcp(1.0f + fast_exp2(z * q));
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float s = gelu_sig(z);
  return s + z * s * (1.0f - s) * kGeluC * (1.0f + 3.0f * kGeluA * z * z);
}

// hard-swish: the gate clamp(z/6 + 1/2, 0, 1) is one FMA with the clamp
// output modifier
constexpr float kSixth = 1.0f / 6.0f;
__device__ __forceinline__ float hswish_f(float z) {
  return z * __builtin_amdgcn_fmed3f(fmaf(z, kSixth, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ float hswish_grad_f(float z) {
  return z <= -3.0f ? 0.0f : (z >= 3.0f ? 1.0f : fmaf(z, 2.0f * kSixth, 0.5f));
}

template <int ACT>
__device__ __forceinline__ float act_f(float z) {
  if constexpr (ACT == OSTC_SUR_HSWISH)
    return hswish_f(z);
  else
    return gelu_f(z);
}
template <int ACT>
__device__ __forceinline__ float act_grad_f(float z) {
  if constexpr (ACT == OSTC_SUR_HSWISH)
    return hswish_grad_f(z);
  else
    return gelu_grad_f(z);
}

// two FP32 values -> packed BF16 pair (round to nearest even)
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  const bf16x2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(unsigned, v);
}

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

// ---- sources of chemistry states ----------------------------------------
// Each source reads one cell's state as stored (Raw: FP32 or FP64) and
// widens it to FP64 for the features and the checks.
template <typename Raw>
__device__ __forceinline__ void unpack_raw(const Raw &r, double &T, double &P, double *Y, double &DT) {
  T = (double)r.t;
  P = (double)r.p;
  DT = (double)r.dt;
#pragma unroll
  for (int k = 0; k < kNs; ++k) Y[k] = (double)r.y[k];
}

struct BatchSrc {
  const double *t, *p, *y, *dt;
  const int *idx;  // optional gather (training)
  long n;
  struct Raw {
    double t, p, dt, y[kNs];
  };
  __device__ void load_raw(long c, Raw &r) const {
    const long q = row(c);
    r.t = t[q];
    r.p = p[q];
    r.dt = dt[q];
#pragma unroll
    for (int k = 0; k < kNs; ++k) r.y[k] = y[q * kNs + k];
  }
  __device__ long row(long c) const { return idx ? (long)idx[c] : c; }
  __device__ void load(long c, double &T, double &P, double *Y, double &DT) const {
    Raw r;
    load_raw(c, r);
    unpack_raw(r, T, P, Y, DT);
  }
};

template <typename R>
struct ZoneSrc {
  const R *bv, *sv, *dt;
  int mx, my, g, nx, ny;
  long plane, n;
  struct Raw {
    R t, p, dt, y[kNs];
  };
  __device__ void load_raw(long c, Raw &r) const {
    const long q = at(c);
    r.t = bv[5 * plane + q];
    r.p = bv[4 * plane + q];
    r.dt = dt[q];
#pragma unroll
    for (int k = 0; k < kNs; ++k) r.y[k] = sv[k * plane + q];
  }
  __device__ long at(long c) const {
    const long i = c % mx, jk = c / mx;
    const long j = jk % my, k = jk / my;
    return ((k + g) * ny + (j + g)) * (long)nx + (i + g);
  }
  __device__ void load(long c, double &T, double &P, double *Y, double &DT) const {
    Raw r;
    load_raw(c, r);
    unpack_raw(r, T, P, Y, DT);
  }
};

// The 12 raw features in FP32, standardised: 1000/T, ln(p/1 atm),
// BoxCox_0.1(Y_k) = (Y^0.1 - 1)/0.1, ln dt (logarithms and powers through the
// hardware log2/exp2).
constexpr float kLn2 = 0.6931471805599453f;
__device__ __forceinline__ void features(const SurParams &s, double T, double P, const double *Y, double DT,
                                         float *f) {
  f[0] = 1000.0f / (float)T;
  f[1] = __builtin_amdgcn_logf((float)P * kInvAtm) * kLn2;
#pragma unroll
  for (int k = 0; k < kNs; ++k) {
    const float y = fmaxf((float)Y[k], 0.0f);
    f[2 + k] = (fast_exp2(kBoxCox * __builtin_amdgcn_logf(y)) - 1.0f) * kInvBoxCox;
  }
  f[11] = __builtin_amdgcn_logf((float)DT) * kLn2;
#pragma unroll
  for (int q = 0; q < kSurIn; ++q) f[q] = (f[q] - s.in_mean[q]) * s.in_rstd[q];
}

// ---- fused forward --------------------------------------------------------
struct TrainOut {
  float *a0;    // [n][16]  features (BF16-rounded), 1, 0...
  float *z;     // [4][n][128] pre-activations of the hidden layers
  float *a;     // [4][n][144] BF16-rounded activations, 1, 0...
  float *yout;  // [n][16] standardised outputs
};

struct InferOut {
  float *out12;                  // [n][12] standardised outputs (forward -> decode)
  double *dy;                    // [n][9]
  int *fb_count;                 // sub-list lengths (stride kFbStride)
  long *fb_list;                 // sub-lists of fallback cells, fb_cap apart
  long fb_cap;                   // capacity of one sub-list
  unsigned long long *stats;     // statistics rows (stride kStatStride)
  float *raw;                    // optional [n][16] standardised outputs
};

// activation of two accumulator quads (units 4g..4g+3 of two 16-unit tiles)
// packed into the next layer's B operand
template <int ACT>
__device__ __forceinline__ bf16x8 act_pack(const f32x4 &a0, const f32x4 &a1) {
  uint4 u;
  u.x = pack_bf16(act_f<ACT>(a0[0]), act_f<ACT>(a0[1]));
  u.y = pack_bf16(act_f<ACT>(a0[2]), act_f<ACT>(a0[3]));
  u.z = pack_bf16(act_f<ACT>(a1[0]), act_f<ACT>(a1[1]));
  u.w = pack_bf16(act_f<ACT>(a1[2]), act_f<ACT>(a1[3]));
  return __builtin_bit_cast(bf16x8, u);
}

// training: pre-activations z and BF16-rounded activations of hidden layer l
// for cell c, units u0..u0+3 (first tile) and u0+16..u0+19 (second tile)
__device__ __forceinline__ void save_layer(const TrainOut &tr, int l, long n, long c, int u0, const f32x4 &z0,
                                           const f32x4 &z1, const bf16x8 &h) {
  if (c >= n) return;
  float *zr = tr.z + ((long)l * n + c) * 128, *ar = tr.a + ((long)l * n + c) * 144;
  *reinterpret_cast<float4 *>(zr + u0) = make_float4(z0[0], z0[1], z0[2], z0[3]);
  *reinterpret_cast<float4 *>(zr + u0 + 16) = make_float4(z1[0], z1[1], z1[2], z1[3]);
  *reinterpret_cast<float4 *>(ar + u0) = make_float4((float)h[0], (float)h[1], (float)h[2], (float)h[3]);
  *reinterpret_cast<float4 *>(ar + u0 + 16) = make_float4((float)h[4], (float)h[5], (float)h[6], (float)h[7]);
}

// LDS image offsets (in 1 KB A tiles of 64 lanes x 8 BF16)
__device__ __forceinline__ int tile_l0(int mt) { return mt; }
__device__ __forceinline__ int tile_h(int l, int mt, int kc) { return 8 + (l - 1) * 32 + mt * 4 + kc; }
__device__ __forceinline__ int tile_out(int kc) { return 104 + kc; }

template <typename Src, bool TRAIN, int ACT>
__global__ __launch_bounds__(kWaves * 64) void k_sur_forward(SurParams s, Src src, const bf16x8 *__restrict__ wpk,
                                                            const float *__restrict__ bias, TrainOut tr,
                                                            InferOut io) {
  __shared__ bf16x8 sW[kSurTiles * 64];
  __shared__ __attribute__((aligned(16))) float sB[kSurBias];
  __shared__ float sStage[kWaves][64 * kStageLd];
  static_assert(kWaves * 64 >= kFbLists, "workgroup 0 clears every sub-list counter");
  if constexpr (!TRAIN) {
    // clear the fallback sub-list counters and statistics rows that the
    // decode kernel (next on the stream) appends to
    if (blockIdx.x == 0 && threadIdx.x < kFbLists && io.fb_count) {
      io.fb_count[threadIdx.x * kFbStride] = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) io.stats[threadIdx.x * kStatStride + r] = 0;
    }
  }
  for (int q = threadIdx.x; q < kSurTiles * 64; q += blockDim.x) sW[q] = wpk[q];
  for (int q = threadIdx.x; q < kSurBias; q += blockDim.x) sB[q] = bias[q];
  __syncthreads();

  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63, g = ln >> 4, cn = ln & 15;
  const long n = src.n;
  const long stride = (long)gridDim.x * kWaves * 64;
  for (long base = ((long)blockIdx.x * kWaves + wave) * 64; base < n; base += stride) {
    float *stg = sStage[wave];
    // ---- features, one lane per cell, staged through LDS in two halves
    // (rows hold 8 values: features 0..7, then 8..11)
    float f[kSurIn];
    {
      const long c = base + ln;
#pragma unroll
      for (int q = 0; q < kSurIn; ++q) f[q] = 0.0f;
      if (c < n) {
        double T, P, DT, Y[kNs];
        src.load(c, T, P, Y, DT);
        features(s, T, P, Y, DT, f);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) stg[ln * kStageLd + q] = f[q];
    }
    wave_sync();
    // ---- layer 0 operand: slots 8g..8g+7 of the 32-slot chunk = features
    // (g = 0: features 0..7, g = 1: features 8..11 then zeros, g > 1: zeros)
    bf16x8 b0[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
      const float *row = stg + (16 * t + cn) * kStageLd;
      uint4 u;
      u.x = g == 0 ? pack_bf16(row[0], row[1]) : 0u;
      u.y = g == 0 ? pack_bf16(row[2], row[3]) : 0u;
      u.z = g == 0 ? pack_bf16(row[4], row[5]) : 0u;
      u.w = g == 0 ? pack_bf16(row[6], row[7]) : 0u;
      b0[t] = __builtin_bit_cast(bf16x8, u);
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < kSurIn - 8; ++q) stg[ln * kStageLd + q] = f[8 + q];
    wave_sync();
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
      const float *row = stg + (16 * t + cn) * kStageLd;
      if (g == 1) {
        uint4 u;
        u.x = pack_bf16(row[0], row[1]);
        u.y = pack_bf16(row[2], row[3]);
        u.z = 0u;
        u.w = 0u;
        b0[t] = __builtin_bit_cast(bf16x8, u);
      }

      const long c = base + 16 * t + cn;
      if (TRAIN && c < n && g < 2) {
        float4 *dst = reinterpret_cast<float4 *>(tr.a0 + c * 16 + 8 * g);
        dst[0] = make_float4((float)b0[t][0], (float)b0[t][1], (float)b0[t][2], (float)b0[t][3]);
        if (g == 0)
          dst[1] = make_float4((float)b0[t][4], (float)b0[t][5], (float)b0[t][6], (float)b0[t][7]);
        else
          dst[1] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);  // bias column for the weight gradient
      }
    }
    // the staging rows are rewritten by the next pass (their reads above are
    // consumed before the first MFMA)
    __builtin_amdgcn_wave_barrier();

    bf16x8 bc[4][kTiles];  // current layer input: [k chunk][slab]
    // ---- layer 0: 12 (padded 32) -> 128; the bias enters as the MFMA's C
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
      f32x4 acc[2][kTiles];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const f32x4 bz = *reinterpret_cast<const f32x4 *>(&sB[32 * kc + 16 * h + 4 * g]);
        const bf16x8 a = sW[tile_l0(2 * kc + h) * 64 + ln];
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
          acc[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0[t], bz, 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < kTiles; ++t) {
        bc[kc][t] = act_pack<ACT>(acc[0][t], acc[1][t]);
        if (TRAIN) save_layer(tr, 0, n, base + 16 * t + cn, 32 * kc + 4 * g, acc[0][t], acc[1][t], bc[kc][t]);
      }
    }
    // ---- hidden layers 1..3: 128 -> 128
#pragma unroll
    for (int l = 1; l <= 3; ++l) {
      bf16x8 bn[4][kTiles];
      // the 32 A tiles of the layer in issue order q = (kc*4 + ki)*2 + h,
      // each read from LDS one step ahead of its MFMAs
      bf16x8 abuf[2];
      abuf[0] = sW[tile_h(l, 0, 0) * 64 + ln];
#pragma unroll
      for (int kc = 0; kc < 4; ++kc) {
        f32x4 acc[2][kTiles];
        f32x4 bz[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) bz[h] = *reinterpret_cast<const f32x4 *>(&sB[128 * l + 32 * kc + 16 * h + 4 * g]);
#pragma unroll
        for (int ki = 0; ki < 4; ++ki)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int q = (kc * 4 + ki) * 2 + h;
            if (q + 1 < 32) {
              const int qn = q + 1;
              abuf[qn & 1] = sW[tile_h(l, 2 * (qn >> 3) + (qn & 1), (qn >> 1) & 3) * 64 + ln];
            }
            const bf16x8 a = abuf[q & 1];
#pragma unroll
            for (int t = 0; t < kTiles; ++t)
              acc[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bc[ki][t], ki == 0 ? bz[h] : acc[h][t], 0, 0, 0);
          }
#pragma unroll
        for (int t = 0; t < kTiles; ++t) {
          bn[kc][t] = act_pack<ACT>(acc[0][t], acc[1][t]);
          if (TRAIN) save_layer(tr, l, n, base + 16 * t + cn, 32 * kc + 4 * g, acc[0][t], acc[1][t], bn[kc][t]);
        }
      }
#pragma unroll
      for (int kc = 0; kc < 4; ++kc)
#pragma unroll
        for (int t = 0; t < kTiles; ++t) bc[kc][t] = bn[kc][t];
    }
    if (TRAIN && g == 0) {
      // bias column (1, then zero padding) of the activations feeding layers 1..4
#pragma unroll
      for (int t = 0; t < kTiles; ++t) {
        const long c = base + 16 * t + cn;
        if (c < n)
          for (int l = 0; l < 4; ++l) {
            float4 *dst = reinterpret_cast<float4 *>(tr.a + ((long)l * n + c) * 144 + 128);
            dst[0] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
            dst[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            dst[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            dst[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          }
      }
    }
    // ---- output layer: 128 -> n_out (padded 16)
    f32x4 acc[kTiles];
    const f32x4 bo = *reinterpret_cast<const f32x4 *>(&sB[512 + 4 * g]);
#pragma unroll
    for (int ki = 0; ki < 4; ++ki) {
      const bf16x8 a = sW[tile_out(ki) * 64 + ln];
#pragma unroll
      for (int t = 0; t < kTiles; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bc[ki][t], ki == 0 ? bo : acc[t], 0, 0, 0);
    }
    if (TRAIN) {
#pragma unroll
      for (int t = 0; t < kTiles; ++t) {
        const long c = base + 16 * t + cn;
        if (c < n)
          *reinterpret_cast<float4 *>(tr.yout + c * 16 + 4 * g) =
              make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
      }
      continue;
    }
    // standardised outputs in the accumulator layout: lane (g, cn) holds
    // outputs 4g..4g+3 of cell 16t+cn (outputs 12..15 are padding)
    if (g < 3) {
#pragma unroll
      for (int t = 0; t < kTiles; ++t) {
        const long c = base + 16 * t + cn;
        if (c < n)
          *reinterpret_cast<float4 *>(io.out12 + c * 12 + 4 * g) =
              make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
      }
    }
  }
}

// Decode + checks, one thread per cell: accepted cells store dY, the others
// join the wave's fallback sub-list (one atomic per wave).
template <typename Src>
__global__ __launch_bounds__(256) void k_sur_decode(SurParams s, Src src, InferOut io) {
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  const int ln = threadIdx.x & 63;
  int why = 0;  // 0 accepted, 1 input range, 2 output range, 3 element balance
  if (c < src.n) {
    float o[16];
    const float4 *r4 = reinterpret_cast<const float4 *>(io.out12 + c * 12);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float4 v = r4[q];
      o[4 * q] = v.x;
      o[4 * q + 1] = v.y;
      o[4 * q + 2] = v.z;
      o[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int q = 12; q < 16; ++q) o[q] = 0.0f;
    double T, P, DT, Y[kNs];
    src.load(c, T, P, Y, DT);
    if (io.raw) {
#pragma unroll
      for (int q = 0; q < 16; q += 4)
        *reinterpret_cast<float4 *>(io.raw + c * 16 + q) = make_float4(o[q], o[q + 1], o[q + 2], o[q + 3]);
    }
    float dy[kNs];
    sur_decode(s, o, dy);
    why = sur_check(s, T, P, DT, Y, dy);
    if (why == 0) {
#pragma unroll
      for (int k = 0; k < kNs; ++k) io.dy[c * kNs + k] = (double)dy[k];
    }
  }
  const int list = (int)((c >> 6) % kFbLists);
  const unsigned long long fail = __ballot(why != 0);
  if (fail) {
    const int first = __ffsll((long long)fail) - 1;
    int at = 0;
    if (ln == first) at = atomicAdd(io.fb_count + list * kFbStride, __popcll(fail));
    at = __shfl(at, first);
    if (why != 0) io.fb_list[list * io.fb_cap + at + __popcll(fail & ((1ull << ln) - 1ull))] = c;
  }
  if (io.stats) {
    unsigned long long *row = io.stats + list * kStatStride;
    const unsigned long long valid = __ballot(c < src.n);
    if (ln == 0) atomicAdd(&row[0], (unsigned long long)__popcll(valid));
#pragma unroll
    for (int r = 1; r <= 3; ++r) {
      const unsigned long long m = __ballot(why == r);
      if (m && ln == 0) atomicAdd(&row[r], (unsigned long long)__popcll(m));
    }
  }
}

// Plain point-implicit step on the fallback sub-lists (grid-stride over their
// total, which lives on the device; entry q is found by a binary search over
// the sub-list offsets).
template <typename Src>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_sur_fallback(
    const ostc_param *__restrict__ p, Src src,
                                                      const int *__restrict__ count, const long *__restrict__ list,
                                                      long cap, double *__restrict__ dyo) {
  __shared__ int pre[kFbLists + 1];
  for (int l = threadIdx.x; l < kFbLists; l += blockDim.x) pre[l + 1] = count[l * kFbStride];
  __syncthreads();
  if (threadIdx.x == 0) {
    pre[0] = 0;
    for (int l = 1; l <= kFbLists; ++l) pre[l] += pre[l - 1];
  }
  __syncthreads();
  const int m = pre[kFbLists];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
    int lo = 0, hi = kFbLists;  // pre[lo] <= q < pre[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid] <= q)
        lo = mid;
      else
        hi = mid;
    }
    const long c = list[lo * cap + (q - pre[lo])];
    double T, P, DT, Y[kNs], dy[kNs];
    src.load(c, T, P, Y, DT);
    plain_chem_step<kNs>(p, T, P, Y, DT, dy);
#pragma unroll
    for (int k = 0; k < kNs; ++k) dyo[c * kNs + k] = dy[k];
  }
}

template <typename Src>
__global__ __launch_bounds__(256) void k_zone_plain(const ostc_param *__restrict__ p, Src src,
                                                    double *__restrict__ dyo) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= src.n) return;
  double T, P, DT, Y[kNs], dy[kNs];
  src.load(c, T, P, Y, DT);
  plain_chem_step<kNs>(p, T, P, Y, DT, dy);
#pragma unroll
  for (int k = 0; k < kNs; ++k) dyo[c * kNs + k] = dy[k];
}

// ---- training kernels -----------------------------------------------------
// dL/dY of L = sum_b sum_o w_o (y - t)^2 / B; per-block partial losses.
__global__ __launch_bounds__(256) void k_sur_loss(const float *__restrict__ yout, const float *__restrict__ target,
                                                  const int *__restrict__ idx, const float *__restrict__ w, int n_out,
                                                  long nb, float *__restrict__ g5, double *__restrict__ part) {
  __shared__ double red[256];
  const long q = (long)blockIdx.x * 256 + threadIdx.x;  // (sample, output) pair
  double l = 0;
  if (q < nb * 16) {
    const long b = q >> 4;
    const int o = (int)(q & 15);
    float gq = 0.0f;
    if (o < n_out) {
      const long r = idx ? (long)idx[b] : b;
      const float d = yout[b * 16 + o] - target[r * 16 + o];
      gq = 2.0f * w[o] * d / (float)nb;
      l = (double)w[o] * (double)d * (double)d / (double)nb;
    }
    g5[b * 16 + o] = gq;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int m,
                                                      double *__restrict__ out) {
  __shared__ double red[256];
  double v = 0;
  for (int q = threadIdx.x; q < m; q += 256) v += part[q];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// C[M x N] = sum_k A(m,k) B(k,n) on FP32 MFMA 16x16x4, one 16x16 tile per
// wave.  In each group of 4 MFMAs lane group g supplies k = kb + 4g + s
// (s = 0..3) to both operands, so a lane reads 4 consecutive k.
enum { GEMM_DGRAD = 0, GEMM_WGRAD = 1 };

struct GemmArgs {
  int M, N;        // multiples of 16
  long K;          // multiple of 16
  int splits;      // split of K (WGRAD)
  // DGRAD: A(m,k) = wq[k*lda + m] (k < kmax), B(k,n) = g[n*ldb + k];
  //        out[n*128 + m] = C * act'(z[n*128 + m])
  // WGRAD: A(m,k) = g[k*lda + m], B(k,n) = act[k*ldb + n]; out[s][m*N + n]
  const float *a, *b, *z;
  int lda, ldb, kmax;
  float *out;
};

template <int MODE, int ACT>
__global__ __launch_bounds__(256) void k_sur_gemm(GemmArgs ga) {
  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63, g = ln >> 4, cn = ln & 15;
  const long tiles_m = ga.M / 16, tiles_n = ga.N / 16;
  const long w = (long)blockIdx.x * 4 + wave;
  const long total = tiles_m * tiles_n * ga.splits;
  if (w >= total) return;
  const long mt = w % tiles_m, rest = w / tiles_m;
  const long nt = rest % tiles_n, sp = rest / tiles_n;
  const long cps = (ga.K / 16 + ga.splits - 1) / ga.splits;  // 16-wide k steps per split
  const long k0 = sp * cps * 16, k1 = k0 + cps * 16 < ga.K ? k0 + cps * 16 : ga.K;
  const long m = mt * 16 + cn, nn = nt * 16 + cn;
  f32x4 acc = {0, 0, 0, 0};
  for (long kb = k0; kb < k1; kb += 16) {
    float av[4], bv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const long k = kb + 4 * g + s;
      if (MODE == GEMM_DGRAD) {
        av[s] = (k < ga.kmax) ? ga.a[k * ga.lda + m] : 0.0f;
        bv[s] = ga.b[nn * ga.ldb + k];
      } else {
        av[s] = ga.a[k * ga.lda + m];
        bv[s] = ga.b[k * ga.ldb + nn];
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s], acc, 0, 0, 0);
  }
  const long mr = mt * 16 + 4 * g;  // rows of this lane's 4 results; column nn
  if (MODE == GEMM_DGRAD) {
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = acc[i] * act_grad_f<ACT>(ga.z[nn * 128 + mr + i]);
    *reinterpret_cast<float4 *>(ga.out + nn * 128 + mr) = make_float4(r[0], r[1], r[2], r[3]);
  } else {
    float *o = ga.out + sp * (long)ga.M * ga.N;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[(mr + i) * ga.N + nn] = acc[i];
  }
}

// grad[off + o*K + i] = sum_s part[s][o*N + i]; bias from column K.
__global__ __launch_bounds__(256) void k_sur_reduce(const float *__restrict__ part, int splits, int M, int N,
                                                    int m_used, int k_used, float *__restrict__ grad) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m_used * (k_used + 1)) return;
  const int o = q / (k_used + 1), i = q % (k_used + 1);
  float v = 0.0f;
  for (int s = 0; s < splits; ++s) v += part[(long)s * M * N + o * N + i];
  if (i < k_used)
    grad[o * k_used + i] = v;
  else
    grad[m_used * k_used + o] = v;
}

__global__ __launch_bounds__(256) void k_sur_adam(float *__restrict__ w, const float *__restrict__ gr,
                                                  float *__restrict__ m1, float *__restrict__ m2, int n, float lr,
                                                  float c1, float c2) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gq = gr[q];
  const float a = kAdamB1 * m1[q] + (1.0f - kAdamB1) * gq;
  const float b = kAdamB2 * m2[q] + (1.0f - kAdamB2) * gq * gq;
  m1[q] = a;
  m2[q] = b;
  w[q] -= lr * (a * c1) / (sqrtf(b * c2) + kAdamEps);
}

// master (FP32, [out][in] then bias per layer) -> LDS image of BF16 A tiles,
// FP32 bias image, and the BF16-rounded weights as FP32 for the backward.
__global__ __launch_bounds__(64) void k_sur_pack(const float *__restrict__ w, int n_out, bf16x8 *__restrict__ wpk,
                                                 float *__restrict__ bias, float *__restrict__ wq) {
  const int tile = blockIdx.x, ln = threadIdx.x, g = ln >> 4;
  int l, mt, kc;
  if (tile < 8) {
    l = 0, mt = tile, kc = 0;
  } else if (tile < 104) {
    l = 1 + (tile - 8) / 32;
    mt = ((tile - 8) % 32) / 4;
    kc = (tile - 8) % 4;
  } else {
    l = 4, mt = 0, kc = tile - 104;
  }
  const int kin = sur_in_dim(l), mout = sur_out_dim(l, n_out);
  const float *wl = w + sur_layer_offset(l);
  const int m = 16 * mt + (ln & 15);
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = (l == 0) ? 8 * g + j : 32 * kc + ((j < 4) ? 4 * g + j : 16 + 4 * g + (j - 4));
    v[j] = (__bf16)((m < mout && f < kin) ? wl[m * kin + f] : 0.0f);
  }
  wpk[tile * 64 + ln] = v;
  if (tile == 0) {
    for (int q = ln; q < kSurBias; q += 64) {
      const int ll = q < 512 ? q / 128 : 4, o = q < 512 ? q % 128 : q - 512;
      bias[q] = (o < sur_out_dim(ll, n_out)) ? w[sur_layer_offset(ll) + sur_out_dim(ll, n_out) * sur_in_dim(ll) + o]
                                              : 0.0f;
    }
  }
  // weights rounded to BF16 (as the forward sees them), one layer slice per tile
  if (tile < 5) {
    const int kk = sur_in_dim(tile), mm = sur_out_dim(tile, n_out);
    const float *src = w + sur_layer_offset(tile);
    float *dst = wq + sur_layer_offset(tile);
    for (int q = ln; q < kk * mm; q += 64) dst[q] = bf16_round(src[q]);
  }
}

unsigned grid_for(long waves) { return (unsigned)((waves + 3) / 4); }

}  // namespace

// ---------------------------------------------------------------------------
int sur_pack(const SurState &st, hipStream_t s) {
  hipLaunchKernelGGL(k_sur_pack, dim3(kSurTiles), dim3(64), 0, s, st.master, st.params.n_out,
                     reinterpret_cast<bf16x8 *>(st.packed), st.bias, st.wq);
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

static unsigned forward_grid(long n, int cus) {
  const long wgs = (n + kWaves * 64 - 1) / (kWaves * 64);
  return (unsigned)std::max<long>(1, std::min<long>(wgs, cus));
}

template <typename Src, bool TRAIN>
static void launch_forward(const SurState &st, const Src &src, const TrainOut &tr, const InferOut &io,
                           hipStream_t s) {
  const dim3 grid(forward_grid(src.n, st.cus)), block(kWaves * 64);
  const bf16x8 *wpk = reinterpret_cast<const bf16x8 *>(st.packed);
  if (st.params.act == OSTC_SUR_HSWISH)
    hipLaunchKernelGGL((k_sur_forward<Src, TRAIN, OSTC_SUR_HSWISH>), grid, block, 0, s, st.params, src, wpk,
                       st.bias, tr, io);
  else
    hipLaunchKernelGGL((k_sur_forward<Src, TRAIN, OSTC_SUR_GELU>), grid, block, 0, s, st.params, src, wpk, st.bias,
                       tr, io);
}

template <typename Src>
static int infer_impl(const SurState &st, const ostc_param *dp, const Src &src, double *dy, float *raw,
                      hipStream_t s) {
  TrainOut tr{};
  InferOut io{st.out12, dy, st.fb_count, st.fb_list, fb_list_cap(st.fb_cap), st.stats, raw};
  launch_forward<Src, false>(st, src, tr, io, s);
  hipLaunchKernelGGL(k_sur_decode<Src>, dim3((unsigned)((src.n + 255) / 256)), dim3(256), 0, s, st.params, src,
                     io);
  // one wave per workgroup (the kernel's launch bound), enough of them to
  // spread a typical fallback list over every CU; the kernel strides over
  // the device-side count
  const long fb_blocks = std::min<long>((src.n + 63) / 64, 16L * st.cus);
  hipLaunchKernelGGL(k_sur_fallback<Src>, dim3((unsigned)std::max<long>(1, fb_blocks)), dim3(64), 0, s, dp, src,
                     st.fb_count, st.fb_list, fb_list_cap(st.fb_cap), dy);
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

int sur_infer_batch(const SurState &st, const ostc_param *dp, long n, const double *t, const double *p,
                    const double *y, const double *dt, double *dy, float *raw, hipStream_t s) {
  if (n <= 0) return OSTC_OK;
  BatchSrc src{t, p, y, dt, nullptr, n};
  return infer_impl(st, dp, src, dy, raw, s);
}

template <typename R>
int sur_infer_zone(const SurState &st, const ostc_param *dp, const ZoneView<R> &z, double *dy, hipStream_t s) {
  ZoneSrc<R> src{z.bv, z.sv, z.dt, z.mx, z.my, z.g, z.nx, z.ny, z.plane, (long)z.mx * z.my * z.mz};
  return infer_impl(st, dp, src, dy, nullptr, s);
}

template <typename R>
int zone_plain_chem(const ostc_param *dp, const ZoneView<R> &z, double *dy, hipStream_t s) {
  ZoneSrc<R> src{z.bv, z.sv, z.dt, z.mx, z.my, z.g, z.nx, z.ny, z.plane, (long)z.mx * z.my * z.mz};
  hipLaunchKernelGGL(k_zone_plain<ZoneSrc<R>>, dim3((unsigned)((src.n + 255) / 256)), dim3(256), 0, s, dp, src, dy);
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

template int sur_infer_zone<float>(const SurState &, const ostc_param *, const ZoneView<float> &, double *,
                                   hipStream_t);
template int sur_infer_zone<double>(const SurState &, const ostc_param *, const ZoneView<double> &, double *,
                                    hipStream_t);
template int zone_plain_chem<float>(const ostc_param *, const ZoneView<float> &, double *, hipStream_t);
template int zone_plain_chem<double>(const ostc_param *, const ZoneView<double> &, double *, hipStream_t);

int sur_gradients(const SurState &st, long nb, const double *t, const double *p, const double *y, const double *dt,
                  const int *idx, const float *target, float *grad, double *loss_dev, hipStream_t s) {
  if (nb <= 0 || nb % 16 || nb > st.max_batch) return OSTC_ERR_ARG;
  BatchSrc src{t, p, y, dt, idx, nb};
  TrainOut tr{st.a0, st.z, st.a, st.yout};
  InferOut io{};
  launch_forward<BatchSrc, true>(st, src, tr, io, s);
  const unsigned lblocks = (unsigned)((nb * 16 + 255) / 256);
  hipLaunchKernelGGL(k_sur_loss, dim3(lblocks), dim3(256), 0, s, st.yout, target, idx, st.loss_w,
                     st.params.n_out, nb, st.g5, st.loss_part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, st.loss_part, (int)lblocks, loss_dev);
  const int n_out = st.params.n_out;
  const float *gcur = st.g5;
  int ldg = 16;
  for (int l = 4; l >= 0; --l) {
    const int kin = sur_in_dim(l), mout = sur_out_dim(l, n_out);
    const int M = (mout + 15) / 16 * 16;        // 128 or 16
    const int N = (kin + 1 + 15) / 16 * 16;     // 144 or 16
    const float *act = (l == 0) ? st.a0 : st.a + (long)(l - 1) * nb * 144;
    const int ldh = (l == 0) ? 16 : 144;
    // weight gradient (split over the batch)
    GemmArgs wg{};
    wg.M = M;
    wg.N = N;
    wg.K = nb;
    wg.splits = (int)std::max<long>(1, std::min<long>(nb / 256, kSurSplits));
    wg.a = gcur;
    wg.lda = ldg;
    wg.b = act;
    wg.ldb = ldh;
    wg.out = st.part;
    // (the weight gradient does not use the activation parameter)
    hipLaunchKernelGGL((k_sur_gemm<GEMM_WGRAD, OSTC_SUR_GELU>), dim3(grid_for((long)(M / 16) * (N / 16) * wg.splits)), dim3(256), 0,
                       s, wg);
    hipLaunchKernelGGL(k_sur_reduce, dim3((unsigned)((mout * (kin + 1) + 255) / 256)), dim3(256), 0, s, st.part,
                       wg.splits, M, N, mout, kin, grad + sur_layer_offset(l));
    if (l == 0) break;
    // input gradient through W (BF16-rounded) and the activation's derivative
    GemmArgs dg{};
    dg.M = 128;
    dg.N = (int)nb;
    dg.K = M;
    dg.splits = 1;
    dg.a = st.wq + sur_layer_offset(l);
    dg.lda = kin;
    dg.kmax = mout;
    dg.b = gcur;
    dg.ldb = ldg;
    dg.z = st.z + (long)(l - 1) * nb * 128;
    float *gnext = st.gbuf[l & 1];
    dg.out = gnext;
    const dim3 dgrid(grid_for((long)(128 / 16) * (nb / 16)));
    if (st.params.act == OSTC_SUR_HSWISH)
      hipLaunchKernelGGL((k_sur_gemm<GEMM_DGRAD, OSTC_SUR_HSWISH>), dgrid, dim3(256), 0, s, dg);
    else
      hipLaunchKernelGGL((k_sur_gemm<GEMM_DGRAD, OSTC_SUR_GELU>), dgrid, dim3(256), 0, s, dg);
    gcur = gnext;
    ldg = 128;
  }
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

int sur_adam(SurState &st, float lr, hipStream_t s) {
  st.adam_step += 1;
  const double c1 = 1.0 / (1.0 - std::pow((double)kAdamB1, (double)st.adam_step));
  const double c2 = 1.0 / (1.0 - std::pow((double)kAdamB2, (double)st.adam_step));
  const int n = st.n_params;
  hipLaunchKernelGGL(k_sur_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, st.master, st.grad, st.m1,
                     st.m2, n, lr, (float)c1, (float)c2);
  if (hipGetLastError() != hipSuccess) return OSTC_ERR_HIP;
  return sur_pack(st, s);
}

}  // namespace ostc
