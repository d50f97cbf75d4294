/* Internal helpers of the FP64 oracle.  TEST INFRASTRUCTURE ONLY. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include "ostc_oracle.h"

/* Physical constants: src/Constants.h:9-27 */
#define ORC_RU 8314.462618
#define ORC_RC 1.98720425864083
#define ORC_GAMMA_AIR 1.4
#define ORC_P_ATM 101325.0
#define ORC_MW_AIR 28.9647

/* SST constants: src/SST.cuh:10-31 */
#define SST_BETA_STAR 0.09
#define SST_SQRT_BETA_STAR 0.3
#define SST_KAPPA 0.41
#define SST_SIGMA_K1 0.85
#define SST_SIGMA_OMEGA1 0.5
#define SST_BETA_1 0.0750
#define SST_A_1 0.31
#define SST_SIGMA_K2 1.0
#define SST_SIGMA_OMEGA2 0.856
#define SST_BETA_2 0.0828
#define SST_GAMMA1 (SST_BETA_1 / SST_BETA_STAR - SST_SIGMA_OMEGA1 * SST_KAPPA * SST_KAPPA / SST_SQRT_BETA_STAR)
#define SST_GAMMA2 (SST_BETA_2 / SST_BETA_STAR - SST_SIGMA_OMEGA2 * SST_KAPPA * SST_KAPPA / SST_SQRT_BETA_STAR)
#define SST_DELTA_SIGMA_K (SST_SIGMA_K1 - SST_SIGMA_K2)
#define SST_DELTA_SIGMA_OMEGA (SST_SIGMA_OMEGA1 - SST_SIGMA_OMEGA2)
#define SST_DELTA_BETA (SST_BETA_1 - SST_BETA_2)
#define SST_DELTA_GAMMA (SST_GAMMA1 - SST_GAMMA2)

static inline long IX(const orc_zone *z, int i, int j, int k) {
  return ((long)(k + z->ngg) * z->ny + (j + z->ngg)) * z->nx + (i + z->ngg);
}
/* arr(i,j,k,l) */
#define F(arr, i, j, k, l) (z->arr[IX(z, (i), (j), (k)) + (long)(l) * z->plane])
#define S(arr, i, j, k) (z->arr[IX(z, (i), (j), (k))])
/* metric(r,c), 1-based as the reference's gxl::Matrix<real,3,3,1> */
#define M(i, j, k, r, c) (z->metric[IX(z, (i), (j), (k)) + (long)(((r)-1) * 3 + ((c)-1)) * z->plane])

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

void orc_finite_rate_chemistry(orc_zone *z, int i, int j, int k, const orc_param *p);
void orc_epi_solve_dq(orc_zone *z, int i, int j, int k, double diag_or_one, const orc_param *p);
void orc_epi_dqk(orc_zone *z, int i, int j, int k, double diag, const double *dq_total, const orc_param *p);
void orc_compute_temperature(orc_zone *z, int i, int j, int k, const orc_param *p, int *iters);

#endif
