// Device-side common definitions for the MI355X residual-and-update path.
// Zone views, index arithmetic, model constants and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/ostc.h"

namespace ostc {

// Physical constants: src/Constants.h:9-27
constexpr double kRu = 8314.462618;
constexpr double kRc = 1.98720425864083;
constexpr double kGammaAir = 1.4;
constexpr double kPatm = 101325.0;
constexpr double kMwAir = 28.9647;

// k-omega SST constants: src/SST.cuh:10-31
namespace sst {
constexpr double beta_star = 0.09;
constexpr double sqrt_beta_star = 0.3;
constexpr double kappa = 0.41;
constexpr double sigma_k1 = 0.85;
constexpr double sigma_omega1 = 0.5;
constexpr double beta_1 = 0.0750;
constexpr double a_1 = 0.31;
constexpr double gamma1 = beta_1 / beta_star - sigma_omega1 * kappa * kappa / sqrt_beta_star;
constexpr double sigma_k2 = 1;
constexpr double sigma_omega2 = 0.856;
constexpr double beta_2 = 0.0828;
constexpr double gamma2 = beta_2 / beta_star - sigma_omega2 * kappa * kappa / sqrt_beta_star;
constexpr double delta_sigma_k = sigma_k1 - sigma_k2;
constexpr double delta_sigma_omega = sigma_omega1 - sigma_omega2;
constexpr double delta_beta = beta_1 - beta_2;
constexpr double delta_gamma = gamma1 - gamma2;
}  // namespace sst

constexpr int kMaxSpec = OSTC_MAX_SPEC;
constexpr int kMaxReac = OSTC_MAX_REAC;
constexpr int kMaxVar = 5 + kMaxSpec + 2;

// Device view of one block.  All planes are padded by g on every side and
// laid out as the reference's VectorField3D (i fastest).  R is the storage
// type of the bulk fields; the chemistry Jacobian / LU stay double.
template <typename R>
struct ZoneView {
  int mx, my, mz, g;
  int nx, ny, nz;
  int64_t plane;
  // geometry
  const double *x, *y, *z;
  R *jac, *metric, *wall_dist;
  // state
  R *bv, *sv, *cv, *bv_last;
  R *vel, *sound, *mach, *mul, *lambda;
  R *rho_d, *gamma, *cp;
  R *mut, *lambda_t, *turb_jac;
  R *dq, *dq0, *dqk;
  R *spectr, *dt;
  double *chem_jac;   // ns*ns (EPI) or ns (DA) planes
  double *chem_lu;    // ns*ns planes: LU factors of the DPLUR chemistry block
  int *chem_piv;      // ns planes
  unsigned long long *counters;  // [0] newton cap hits, [1] limit_flow repairs, [2] surrogate fallbacks
  // limit_flow repair list: count, cell index and final (bv, sv, cv) per
  // entry; lf_cap entries (one per interior cell); an index of -1 marks an
  // entry whose cell turned out to need no repair
  int *lf_count;
  int64_t *lf_idx;
  R *lf_vals;
  int lf_cap;

  __host__ __device__ __forceinline__ int64_t idx(int i, int j, int k) const {
    return ((int64_t)(k + g) * ny + (j + g)) * nx + (i + g);
  }
  // metric(r,c), 1-based like gxl::Matrix<real,3,3,1>
  __device__ __forceinline__ R m(int64_t id, int r, int c) const {
    return metric[id + (int64_t)((r - 1) * 3 + (c - 1)) * plane];
  }
};

// Tiled storage of the DPLUR chemistry block's LU factors (the zones whose
// source kernel stores them, where they do not share the Jacobian's planes):
// entry q (of nq) of the cell at padded index id, in tiles of 64 consecutive
// cells -- a wave row -- each tile holding its cells' entry q as one run of 64
// doubles.  A wave's loads of the whole block read one contiguous stretch of
// nq * 512 B instead of nq runs a plane apart.  Indices reach
// ceil(P / 64) * 64 * nq, beyond P * nq by up to 63 * nq.
__host__ __device__ __forceinline__ int64_t lu_tiled(int64_t id, int q, int nq) {
  return (id & ~(int64_t)63) * nq + (int64_t)q * 64 + (id & 63);
}

template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }

// Block-id remapping so that each XCD (blocks are dealt round-robin to the 8
// XCDs) works on one contiguous slab of the domain: stencil halos between
// neighbouring tiles then hit the same L2.
// a/b: IEEE division in FP64; in the FP32 bulk path a times the hardware
// reciprocal (v_rcp_f32, 1 ulp), which replaces the ~10-instruction
// correctly rounded division sequence in the flux kernels.
template <typename R>
__device__ __forceinline__ R fdiv(R a, R b) {
  return a / b;
}
template <>
__device__ __forceinline__ float fdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
  constexpr unsigned kXcd = 8;
  const unsigned per = nblocks / kXcd;
  if (per == 0 || b >= per * kXcd) return b;
  return (b % kXcd) * per + b / kXcd;
}

// Workgroup coordinates after the remap above: workgroups are dispatched
// round-robin over the XCDs, so spreading contiguous ranges of the linear
// workgroup id over them keeps spatial neighbours (the j and k stencil rows)
// in one XCD's L2.
__device__ __forceinline__ uint3 xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned r = xcd_remap(b, n);
  return make_uint3(r % gx, (r / gx) % gy, r / (gx * gy));
}

}  // namespace ostc
