// Point chemistry kernels on batches of states (FP64): the plain
// point-implicit step that labels the surrogate's training data and serves as
// its GPU baseline, and the raw source/Jacobian evaluation used by tests.
#include "ostc_kernels.h"
#include "ostc_thermo.h"

namespace ostc {

namespace {

// Batch of plain point-implicit steps (plain_chem_step), AoS [cell][species].
template <int NS>
__global__ __launch_bounds__(256) void k_plain_step(const ostc_param *__restrict__ p, long n,
                                                    const double *__restrict__ t, const double *__restrict__ pr,
                                                    const double *__restrict__ y, const double *__restrict__ dt,
                                                    double *__restrict__ dy) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double yy[NS], out[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) yy[l] = y[c * NS + l];
  plain_chem_step<NS>(p, t[c], pr[c], yy, dt[c], out);
#pragma unroll
  for (int l = 0; l < NS; ++l) dy[c * NS + l] = out[l];
}

template <int NS>
__global__ __launch_bounds__(256) void k_source_batch(const ostc_param *__restrict__ p, long n,
                                                      const double *__restrict__ rho, const double *__restrict__ t,
                                                      const double *__restrict__ y, double *__restrict__ omega,
                                                      double *__restrict__ jac) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double yy[NS], om[NS], a[NS * NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) yy[l] = y[c * NS + l];
  if (jac)
    chem_source<NS, kMaxReac, 1>(p, NS, p->n_reac, rho[c], t[c], yy, om, a);
  else
    chem_source<NS, kMaxReac, 0>(p, NS, p->n_reac, rho[c], t[c], yy, om, nullptr);
#pragma unroll
  for (int l = 0; l < NS; ++l) omega[c * NS + l] = om[l];
  if (jac)
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) jac[c * NS * NS + q] = a[q];
}

}  // namespace

int chem_plain_step(const ostc_param *dp, int ns, int nr, long n, const double *t, const double *p, const double *y,
                    const double *dt, double *dy, hipStream_t st) {
  if (nr > kMaxReac) return OSTC_ERR_ARG;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (n <= 0) return OSTC_OK;
  if (ns == 9)
    hipLaunchKernelGGL(k_plain_step<9>, dim3(blocks), dim3(256), 0, st, dp, n, t, p, y, dt, dy);
  else
    return OSTC_ERR_ARG;
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

int chem_source_batch(const ostc_param *dp, int ns, int nr, long n, const double *rho, const double *t,
                      const double *y, double *omega, double *jac, hipStream_t st) {
  if (nr > kMaxReac) return OSTC_ERR_ARG;
  if (n <= 0) return OSTC_OK;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (ns == 9)
    hipLaunchKernelGGL(k_source_batch<9>, dim3(blocks), dim3(256), 0, st, dp, n, rho, t, y, omega, jac);
  else
    return OSTC_ERR_ARG;
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

}  // namespace ostc
