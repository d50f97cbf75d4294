/*
 * FP64 CPU oracle of OpenSTC's per-iteration residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * openstc_amd/csrc; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Every routine restates, in plain C, the arithmetic of the reference source
 * cited next to it (paths relative to the reference repository root).  Array
 * layout is the reference's VectorField3D/Array3D layout (src/gxl_lib/Array.hpp:
 * 83-89,110-130): SoA planes, i fastest, negative ghost indices down to -ngg:
 *   idx(i,j,k,l) = ((k+g)*ny + (j+g))*nx + (i+g) + l*plane,
 *   nx = mx+2g, ny = my+2g, nz = mz+2g, plane = nx*ny*nz.
 * Unlike the reference, every per-cell array here uses the same g (=ngg), the
 * metric is 9 SoA planes (plane (r-1)*3+(c-1) holds metric(r,c)) and the
 * inviscid spectral radius is 3 planes.
 *
 * Parity status per function is recorded in the table of pins in the
 * repository's design notes; functions without a closed-form / invariant /
 * brute-force check are marked "parity unpinned" next to their definition.
 */
#ifndef OSTC_ORACLE_H
#define OSTC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SPEC 20
#define ORC_MAX_REAC 50

/* Model enums: src/Define.h:14-27 */
enum { ORC_AIR = 0, ORC_MIXTURE = 1, ORC_FR = 2 };
enum { ORC_LAMINAR = 0, ORC_RANS = 1 };

/* Mirror of DParameter (src/DParameter.h:11-58) plus the model switches the
 * reference takes from Parameter.  Matrices are row-major with fixed strides
 * ORC_MAX_SPEC (species) and 7 (NASA-7 coefficients). */
typedef struct orc_param {
  int mix_model, turb_method;
  int inviscid_scheme, reconstruction, limiter, viscous_scheme;
  int rans_model, turb_implicit, chem_src_method;
  int implicit_method, dplur_inner_step;
  int n_spec, n_scalar, n_reac, n_var;
  double Pr, cfl, Sc, Prt, Sct;
  double mw[ORC_MAX_SPEC], t_low[ORC_MAX_SPEC], t_mid[ORC_MAX_SPEC], t_high[ORC_MAX_SPEC];
  double low_temp_coeff[ORC_MAX_SPEC * 7], high_temp_coeff[ORC_MAX_SPEC * 7];
  double LJ_potent_inv[ORC_MAX_SPEC], vis_coeff[ORC_MAX_SPEC];
  double WjDivWi_to_One4th[ORC_MAX_SPEC * ORC_MAX_SPEC];
  double sqrt_WiDivWjPl1Mul8[ORC_MAX_SPEC * ORC_MAX_SPEC];
  int reac_type[ORC_MAX_REAC], reac_order[ORC_MAX_REAC];
  int stoi_f[ORC_MAX_REAC * ORC_MAX_SPEC], stoi_b[ORC_MAX_REAC * ORC_MAX_SPEC];
  double A[ORC_MAX_REAC], b[ORC_MAX_REAC], Ea[ORC_MAX_REAC];
  double A2[ORC_MAX_REAC], b2[ORC_MAX_REAC], Ea2[ORC_MAX_REAC];
  double third_body_coeff[ORC_MAX_REAC * ORC_MAX_SPEC];
  double troe_alpha[ORC_MAX_REAC], troe_t3[ORC_MAX_REAC], troe_t1[ORC_MAX_REAC], troe_t2[ORC_MAX_REAC];
  double limit_ll[7], limit_ul[7], sv_inf[ORC_MAX_SPEC + 2];
} orc_param;

/* One block's state (src/Field.h:13-57).  All planes share the padding g. */
typedef struct orc_zone {
  int mx, my, mz, ngg;
  int nx, ny, nz;
  long plane;
  double *x, *y, *z, *jac, *metric, *wall_distance;
  double *bv, *sv, *cv, *bv_last;
  double *vel, *acoustic_speed, *mach, *mul, *thermal_conductivity;
  double *rho_D, *gamma, *cp, *chem_src_jac;
  double *mut, *turb_therm_cond, *turb_src_jac;
  double *dq, *dq0, *dqk;
  double *inv_spectr_rad, *visc_spectr_rad, *dt_local;
} orc_zone;

/* Boundary program: faces in the order the reference applies them
 * (outflow -> inflow -> symmetry -> wall, src/BoundCond.cuh:321-412), with the
 * ranges already expanded by Boundary::register_boundary (src/Mesh.cpp:14-32)
 * and, for walls, trimmed back (src/BoundCond.cu:215-230). */
enum { ORC_BC_WALL = 2, ORC_BC_SYMMETRY = 3, ORC_BC_INFLOW = 5, ORC_BC_OUTFLOW = 6 };
#define ORC_MAX_FACE 32
#define ORC_MAX_BCDATA 8
typedef struct orc_face {
  int range_start[3], range_end[3];
  int face, direction, kind, slot;
} orc_face;
typedef struct orc_inflow {
  double density, u, v, w, pressure, temperature, mut;
  double sv[ORC_MAX_SPEC + 2];
} orc_inflow;
typedef struct orc_wall {
  int isothermal;
  double temperature;
} orc_wall;
typedef struct orc_bc {
  int n_face;
  orc_face face[ORC_MAX_FACE];
  orc_inflow inflow[ORC_MAX_BCDATA];
  orc_wall wall[ORC_MAX_BCDATA];
} orc_bc;

/* Inner (block-to-block) face after InnerFace::register_boundary
 * (src/Mesh.cpp:44-178). */
typedef struct orc_inner_face {
  int range_start[3], range_end[3], face, direction;
  int target_start[3], target_end[3], target_block, target_face, target_direction;
  int src_tar[3], loop_dir[3], target_loop_dir[3], n_point[3];
} orc_inner_face;

void orc_zone_dims(orc_zone *z);

/* ---- geometry: src/Mesh.cpp ---- */
/* Interior coordinates in (ngg+1)-padded arrays -> ghost coordinates. */
void orc_ghost_grid(int mx, int my, int mz, int ngg, double *x, double *y, double *z);
/* the two halves of orc_ghost_grid: face layers, then edges and corners (a
 * mesh copies the inner-face ghost layers in between) */
void orc_ghost_faces(int mx, int my, int mz, int ngg, double *x, double *y, double *z);
void orc_ghost_edges(int mx, int my, int mz, int ngg, double *x, double *y, double *z);
/* Returns 0, or -1 on a non-positive interior Jacobian; *n_neg counts ghost
 * points repaired from their inner neighbour. */
int orc_jac_metric(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                   double *jac, double *metric, int *n_neg);
/* Ghost coordinates across one inner face copied from the target block
 * (src/Mesh.cpp:940-985); both arrays (ngg+1)-padded. */
void orc_inner_ghost_grid(const orc_inner_face *fc, int mx, int my, int mz, int ngg, double *x, double *y, double *z,
                          int tmx, int tmy, int tmz, const double *tx, const double *ty, const double *tz);
/* Copy the (ngg+1)-padded coordinates into ngg-padded arrays. */
void orc_trim_ghost(int mx, int my, int mz, int ngg, const double *src, double *dst);

/* ---- thermo / transport: src/Thermo.cu, src/Transport.cu ---- */
void orc_enthalpy(double t, double *h, const orc_param *p);
void orc_enthalpy_and_cp(double t, double *h, double *cp, const orc_param *p);
void orc_cp(double t, double *cp, const orc_param *p);
void orc_gibbs_div_rt(double t, const orc_param *p, double *g);
double orc_sutherland(double t);
/* Mixture viscosity, conductivity and rho*D per species at one point (cp per
 * species in J/(kg K)); viscosity alone (device form of the same rule). */
void orc_transport_point(const orc_param *p, double t, double mw_total, const double *y, const double *cp,
                         double *mu_out, double *lambda_out, double *rhoD);
double orc_mixture_viscosity(const orc_param *p, double t, double mw_total, const double *y);

/* ---- per-stage kernels (one call = the whole block) ---- */
void orc_compute_cv_from_bv(orc_zone *z, const orc_param *p);
void orc_update_physical_properties(orc_zone *z, const orc_param *p);
void orc_initialize_mut(orc_zone *z, const orc_param *p);
void orc_compute_source(orc_zone *z, const orc_param *p);
void orc_inviscid_flux(orc_zone *z, const orc_param *p);
void orc_inviscid_flux_dir(orc_zone *z, const orc_param *p, int dir);
/* One AUSM+ face flux (reconstruction from 4 points pv[4*n_var]) */
void orc_ausm_face(const orc_param *p, const double *pv, const double *ml, const double *mr, double jl, double jr,
                   double *fc);
void orc_viscous_flux(orc_zone *z, const orc_param *p);
void orc_local_time_step(orc_zone *z, const orc_param *p);
void orc_implicit_treatment(orc_zone *z, const orc_param *p);
/* DPLUR convective-Jacobian product of cell (i,j,k) in direction dir
 * (src/DPLURHPP.cuh:78-125) with dq read from z->dq (tests) */
void orc_jac_times_dq_point(const orc_zone *z, const orc_param *p, int i, int j, int k, int dir, double pm_sr,
                            double *out);
int orc_update_cv_and_bv(orc_zone *z, const orc_param *p);
int orc_limit_flow(orc_zone *z, const orc_param *p);
/* flow repairs made by every orc_limit_flow call so far; reset if asked */
long orc_limit_flow_total(int reset);
void orc_apply_bc(orc_zone *z, const orc_param *p, const orc_bc *bc);
void orc_inner_communication(orc_zone *zones, int blk, const orc_inner_face *f, const orc_param *p);
/* eliminate_k_gradient (src/FieldOperation.cu:42-62): 2-D blocks only */
void orc_eliminate_k_gradient(orc_zone *z, const orc_param *p);
void orc_store_last_step(orc_zone *z);
void orc_square_of_dbv(orc_zone *z);
void orc_sum_dbv(const orc_zone *z, double out[4]);
void orc_wall_distance(orc_zone *z, const double *wall_xyz, int n_wall_point);

/* One steady iteration of the reference Driver (src/Driver.cu:218-281) over
 * n_blocks blocks; returns the maximum normalised residual if this step
 * computes one (step%output_screen==0 or step==1), else -1.  res_scale[4] is
 * set on step 1 (src/Driver.cu:342-348).  inner faces per block are passed in
 * a flat array with per-block counts. */
double orc_steady_step(orc_zone *zones, int n_blocks, const orc_param *p, const orc_bc *bcs,
                       const orc_inner_face *inner, const int *n_inner, int step, int output_screen,
                       double res_scale[4], long n_grid_total, int *newton_fail);

/* ---- chemistry points: src/FiniteRateChem.cu ---- */
/* omega[ns] (kg/m3/s), omega_d[ns], jac[ns*ns] (EPI, row-major m*ns+n) */
void orc_chem_source_point(const orc_param *p, double rho, double t, const double *y, double *omega,
                           double *omega_d, double *jac_epi, double *jac_da, double *kf_out, double *kb_out);
/* In-place pivoted LU solve as solve_chem_system(lhs,rhs,dim) */
void orc_solve_lu(double *lhs, double *rhs, int dim);
/* The plain point-implicit chemistry step (label generator):
 * rho from EOS, (I - dt*J) d(rhoY) = dt*omega, dY = d(rhoY)/rho. */
void orc_plain_chem_step(const orc_param *p, double t, double pres, const double *y, double dt, double *dy);
void orc_plain_chem_batch(const orc_param *p, long n, const double *t, const double *pres, const double *y,
                          const double *dt, double *dy);
/* T from internal energy by the reference Newton iteration; returns iterations */
int orc_newton_temperature(const orc_param *p, const double *y, double e, double t0, double *t_out);

int orc_omp_threads(void);

/* ---- learned chemistry surrogate (no reference counterpart) ----
 * FP64 restatement of the product's MLP 12-128-128-128-128-n_out (hidden
 * activation by the descriptor: 0 = GELU tanh form, 1 = hard-swish
 * z*clamp(z/6 + 1/2, 0, 1)): features, forward, weighted-MSE gradient,
 * output decoding and the
 * fallback checks.  With quant = 1 the BF16 rounding points of the product's
 * inference (features, weights, hidden activations; round to nearest even)
 * are emulated and the backward treats them as identity (straight-through);
 * with quant = 0 everything is plain FP64 (for finite-difference checks).
 * The descriptor has the layout of ostc_surrogate_desc. */
typedef struct orc_sur_desc {
  int n_out, kind, n_elem, activation;
  float in_mean[12], in_std[12];
  float out_mean[16], out_std[16];
  float loss_w[16];
  float elem_w[4 * 9];
  float floor;
  float t_min, t_max, p_min, p_max, dt_min, dt_max;
  float y_tol, elem_tol, elem_abs;
  float dy_max;
} orc_sur_desc;

int orc_sur_n_params(int n_out);
double orc_bf16(double x);
/* standardised FP32 features of one state */
void orc_sur_features(const orc_sur_desc *d, double t, double p, const double *y, double dt, float *f);
/* standardised outputs out[n*16] (columns >= n_out are 0) */
void orc_sur_forward(const orc_sur_desc *d, const float *w, int quant, long n, const double *t, const double *p,
                     const double *y, const double *dt, double *out);
/* L = sum_b sum_o loss_w[o] (out_bo - target[b*16+o])^2 / n; grad[n_params] */
double orc_sur_loss_grad(const orc_sur_desc *d, const float *w, int quant, long n, const double *t, const double *p,
                         const double *y, const double *dt, const float *target, double *grad);
/* decode standardised outputs o[16] into dy[9] and return the fallback
 * reason (0 accepted, 1 input range, 2 output range, 3 element balance) */
int orc_sur_decode_check(const orc_sur_desc *d, const float *o, double t, double p, const double *y, double dt,
                         double *dy);

#ifdef __cplusplus
}
#endif
#endif
