/*
 * ostc.h -- C-ABI of the MI355X-native OpenSTC residual-and-update path.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * an int status (OSTC_OK == 0).  Streams are passed as void* (a hipStream_t,
 * or NULL for the null stream).  The library owns all device memory behind
 * ostc_ctx / ostc_zone handles; host data goes in and out through
 * ostc_zone_upload / ostc_zone_download in FP64 regardless of the zone's
 * storage precision.
 *
 * Each per-stage entry point replaces one stage of the reference's
 * Driver<mix,turb>::steady_simulation loop (src/Driver.cu:218-289); the
 * reference interface it stands in for is cited next to it.
 */
#ifndef OSTC_H
#define OSTC_H

#ifdef __cplusplus
extern "C" {
#endif

#define OSTC_MAX_SPEC 20
#define OSTC_MAX_REAC 50

enum ostc_status {
  OSTC_OK = 0,
  OSTC_ERR_ARG = 1,       /* invalid argument / unsupported configuration */
  OSTC_ERR_HIP = 2,       /* a HIP runtime call failed */
  OSTC_ERR_ALLOC = 3,     /* device allocation failed */
  OSTC_ERR_NAN = 4,       /* residual is NaN (reference aborts: src/Driver.cu:372-377) */
  OSTC_ERR_STATE = 5,     /* call out of order (e.g. zone not initialised) */
  OSTC_ERR_IO = 6         /* a file could not be read or written */
};

/* Mixture / turbulence models: src/Define.h:14-27 */
enum { OSTC_AIR = 0, OSTC_MIXTURE = 1, OSTC_FR = 2 };
enum { OSTC_LAMINAR = 0, OSTC_RANS = 1 };

/* Model constants: the device copy of the reference's DParameter
 * (src/DParameter.h:11-58, filled by src/DParameter.cu:4-101) plus the model
 * switches the reference reads from Parameter (implicit_method,
 * DPLUR_inner_step).  Species matrices use the fixed stride OSTC_MAX_SPEC;
 * NASA-7 coefficient rows use stride 7.  Mixtures take 1 to 20 species: the
 * kernels are instantiated for 9 and for 20, and a mechanism runs on the
 * smaller of the two that holds it, with inert species appended (zero mass
 * fraction, no reactions), which leaves every result of its own species
 * unchanged.  Zone fields are exchanged in the caller's layout (its species
 * count); ostc_zone_field_ptr exposes the device layout, with 9 or 20
 * species planes. */
typedef struct ostc_param {
  int mix_model, turb_method;
  int inviscid_scheme, reconstruction, limiter, viscous_scheme;
  int rans_model, turb_implicit, chem_src_method;
  int implicit_method, dplur_inner_step;
  int n_spec, n_scalar, n_reac, n_var;
  double Pr, cfl, Sc, Prt, Sct;
  double mw[OSTC_MAX_SPEC], t_low[OSTC_MAX_SPEC], t_mid[OSTC_MAX_SPEC], t_high[OSTC_MAX_SPEC];
  double low_temp_coeff[OSTC_MAX_SPEC * 7], high_temp_coeff[OSTC_MAX_SPEC * 7];
  double LJ_potent_inv[OSTC_MAX_SPEC], vis_coeff[OSTC_MAX_SPEC];
  double WjDivWi_to_One4th[OSTC_MAX_SPEC * OSTC_MAX_SPEC];
  double sqrt_WiDivWjPl1Mul8[OSTC_MAX_SPEC * OSTC_MAX_SPEC];
  int reac_type[OSTC_MAX_REAC], reac_order[OSTC_MAX_REAC];
  int stoi_f[OSTC_MAX_REAC * OSTC_MAX_SPEC], stoi_b[OSTC_MAX_REAC * OSTC_MAX_SPEC];
  double A[OSTC_MAX_REAC], b[OSTC_MAX_REAC], Ea[OSTC_MAX_REAC];
  double A2[OSTC_MAX_REAC], b2[OSTC_MAX_REAC], Ea2[OSTC_MAX_REAC];
  double third_body_coeff[OSTC_MAX_REAC * OSTC_MAX_SPEC];
  double troe_alpha[OSTC_MAX_REAC], troe_t3[OSTC_MAX_REAC], troe_t1[OSTC_MAX_REAC], troe_t2[OSTC_MAX_REAC];
  double limit_ll[7], limit_ul[7], sv_inf[OSTC_MAX_SPEC + 2];
} ostc_param;

/* Boundary program of one block: faces in application order (outflow ->
 * inflow -> symmetry -> wall, src/BoundCond.cuh:321-412) with ranges as
 * produced by Boundary::register_boundary (src/Mesh.cpp:14-32) and the wall
 * trimming of src/BoundCond.cu:215-230.  kind is the BC label 2/3/5/6; slot
 * indexes inflow[] (kind 5) or wall[] (kind 2). */
#define OSTC_MAX_FACE 32
#define OSTC_MAX_BCDATA 8
typedef struct ostc_face {
  int range_start[3], range_end[3];
  int face, direction, kind, slot;
} ostc_face;
typedef struct ostc_inflow { /* src/BoundCond.h:8-26 */
  double density, u, v, w, pressure, temperature, mut;
  double sv[OSTC_MAX_SPEC + 2];
} ostc_inflow;
typedef struct ostc_wall { /* src/BoundCond.h:28-38 */
  int isothermal;
  double temperature;
} ostc_wall;
typedef struct ostc_bc {
  int n_face;
  ostc_face face[OSTC_MAX_FACE];
  ostc_inflow inflow[OSTC_MAX_BCDATA];
  ostc_wall wall[OSTC_MAX_BCDATA];
} ostc_bc;

/* Inner (block-to-block) face after InnerFace::register_boundary
 * (src/Mesh.h:343-370, src/Mesh.cpp:44-178). */
typedef struct ostc_inner_face {
  int range_start[3], range_end[3], face, direction;
  int target_start[3], target_end[3], target_block, target_face, target_direction;
  int src_tar[3], loop_dir[3], target_loop_dir[3], n_point[3];
} ostc_inner_face;

/* Storage precision of the bulk fields of a zone.  OSTC_FP64 keeps every
 * field in double (the reference's real=double).  OSTC_FP32 stores the flow
 * fields, metrics and residuals in float; the chemistry Jacobian, its LU
 * factors, Arrhenius/equilibrium evaluation and the T-from-e Newton solve
 * stay in double. */
enum { OSTC_FP32 = 4, OSTC_FP64 = 8 };

typedef struct ostc_zone_desc {
  int mx, my, mz, ngg; /* interior extents and ghost layers (ngg = 2) */
  int precision;       /* OSTC_FP32 or OSTC_FP64 */
  int block_id;
} ostc_zone_desc;

/* Field identifiers for upload/download (layout: the reference's
 * VectorField3D, src/gxl_lib/Array.hpp:83-130, every field padded by ngg,
 * idx = ((k+g)*(my+2g) + (j+g))*(mx+2g) + (i+g) + l*plane).  METRIC has 9
 * planes, plane (r-1)*3+(c-1) = metric(r,c) of src/gxl_lib/Matrix.hpp:47-62. */
enum ostc_field {
  OSTC_F_X = 0, OSTC_F_Y, OSTC_F_Z, OSTC_F_JAC, OSTC_F_METRIC, OSTC_F_WALL_DIST,
  OSTC_F_BV, OSTC_F_SV, OSTC_F_CV, OSTC_F_BV_LAST,
  OSTC_F_VEL, OSTC_F_SOUND, OSTC_F_MACH, OSTC_F_MUL, OSTC_F_LAMBDA,
  OSTC_F_RHO_D, OSTC_F_GAMMA, OSTC_F_CP, OSTC_F_CHEM_JAC,
  OSTC_F_MUT, OSTC_F_LAMBDA_T, OSTC_F_TURB_JAC,
  OSTC_F_DQ, OSTC_F_DQ0, OSTC_F_DQK, OSTC_F_SPECTR, OSTC_F_DT,
  OSTC_F_COUNT
};

typedef struct ostc_ctx ostc_ctx;
typedef struct ostc_zone ostc_zone;
typedef struct ostc_surrogate ostc_surrogate;

/* ---- context and zones: stand in for Driver's ctor (src/Driver.cu:14-41),
 * Field::setup_device_memory (src/Field.cu:69-171) and the DParameter upload
 * (src/Driver.cu:37-39). */
int ostc_ctx_create(const ostc_param *param, ostc_ctx **out);
int ostc_ctx_destroy(ostc_ctx *ctx);
int ostc_zone_create(ostc_ctx *ctx, const ostc_zone_desc *desc, ostc_zone **out);
int ostc_zone_destroy(ostc_zone *zone);
/* n_planes of the field as allocated for this ctx/zone (0 if unused) */
int ostc_zone_field_planes(const ostc_zone *zone, int field, int *n_planes);
/* Points per plane (padded) */
long ostc_zone_plane(const ostc_zone *zone);
int ostc_zone_upload(ostc_zone *zone, int field, const double *host, int n_planes);
int ostc_zone_download(const ostc_zone *zone, int field, double *host, int n_planes);
/* Device pointer and element size of a field (for zero-copy callers). */
int ostc_zone_field_ptr(const ostc_zone *zone, int field, void **dev_ptr, int *elem_size);
int ostc_zone_set_bc(ostc_zone *zone, const ostc_bc *bc);
int ostc_sizeof_param(void);
int ostc_sizeof_bc(void);

/* ---- host-side setup (CPU, FP64) ---- */
/* Ghost coordinates (Mesh::init_ghost_grid, src/Mesh.cpp:789-943) and the
 * 2nd-order Jacobian/metrics (Block::compute_jac_metric, src/Mesh.cpp:
 * 259-521) of one block with no inner faces.  x/y/z_in: interior points,
 * i fastest (mx*my*mz).  Outputs are ngg-padded in the zone layout:
 * x/y/z_out (optional, 1 plane each), jac (1 plane), metric (9 planes).
 * Bit-exact with the reference's operation order. */
int ostc_block_geometry(int mx, int my, int mz, int ngg, const double *x_in, const double *y_in, const double *z_in,
                        double *x_out, double *y_out, double *z_out, double *jac, double *metric, int *n_neg_jac);

/* InnerFace ctor + register_boundary(ngg, dim) (src/Mesh.cpp:34-178): raw[13]
 * is one record of the reference's inner{rank}.txt read by
 * Mesh::read_inner_interface (src/Mesh.cpp:709-735): the 1-based ranges
 * i1 i2 j1 j2 k1 k2 of this block, those of the target block and the 1-based
 * target block id.  dim = 3, or 2 for a mesh of 2-D blocks (mz = 1). */
int ostc_register_inner_face(const int raw[13], int ngg, int dim, ostc_inner_face *out);

/* Geometry of a multi-block mesh (host, FP64): per block ghost grids
 * (Mesh::init_ghost_grid, src/Mesh.cpp:789-943), then the ghost layers across
 * inner faces copied from the neighbour block (Mesh::init_inner_ghost_grid,
 * src/Mesh.cpp:940-985), then Jacobian and metrics (src/Mesh.cpp:259-521).
 * Inner faces of block b are inner[off_b .. off_b+n_inner[b]). */
typedef struct ostc_block_coords {
  int mx, my, mz;
  const double *x, *y, *z;      /* interior coordinates, i fastest */
  double *x_out, *y_out, *z_out; /* optional: with ngg ghost layers */
  double *jac, *metric;          /* ngg-padded; metric 9 planes */
  int n_neg_jac;                 /* out: ghost points repaired */
} ostc_block_coords;
int ostc_mesh_geometry(int n_blocks, int ngg, ostc_block_coords *blocks, const ostc_inner_face *inner,
                       const int *n_inner);

/* CHEMKIN inputs (Species/Reaction constructors, src/ChemData.cpp:11-614):
 * fills mw, NASA-7 tables, T ranges, Lennard-Jones/viscosity constants, the
 * Wilke tables and (with_reactions != 0) the reaction arrays of *param, and
 * n_spec / n_reac.  The caller sets the remaining model switches. */
#define OSTC_MAX_ELEM 8
typedef struct ostc_mechanism_info {
  int n_spec, n_elem;
  char spec_names[OSTC_MAX_SPEC][24];
  char elem_names[OSTC_MAX_ELEM][4];
  int elem_comp[OSTC_MAX_SPEC * OSTC_MAX_ELEM]; /* [species][element], elements in file order */
} ostc_mechanism_info;
int ostc_read_chemkin(const char *mechanism, const char *thermo, const char *transport, int with_reactions,
                      ostc_param *param, ostc_mechanism_info *info);
const char *ostc_chemkin_error(void);

/* ---- case directory (host, CPU) ----
 * A directory laid out as the reference's working directory: input_files/
 * setup/{0_global_control,...,9_transport_property}.txt, input_files/grid/
 * "grid   0.grd", input_files/boundary_condition/"boundary   0.txt" and
 * "inner   0.txt", and the CHEMKIN files the setup names.  ostc_case_read
 * stands in for the reference's Parameter (src/Parameter.cpp:8-168),
 * Species/Reaction (src/ChemData.cpp:11-614), Mesh (src/Mesh.cpp:546-735),
 * DParameter's limits (src/DParameter.cu:78-100), the boundary programs of
 * DBoundCond::initialize_bc_on_GPU (src/BoundCond.cu:73-232) with the
 * inflow states of Inflow::Inflow (src/BoundCond.cpp:6-106), and
 * initialize_from_start (src/Initialize.cu:4-43, src/Field.cu:21-67).
 * Unsupported settings (2-D blocks, parallel faces, restart, unsteady, WENO,
 * SA) are rejected with OSTC_ERR_ARG and a message (ostc_case_error). */
#define OSTC_MAX_BLOCK 16
typedef struct ostc_case ostc_case;
typedef struct ostc_case_info {
  int n_block, ngg;
  int total_step, output_screen, output_file, n_wall_point;
  double convergence_criteria, grid_scale;
  int mx[OSTC_MAX_BLOCK], my[OSTC_MAX_BLOCK], mz[OSTC_MAX_BLOCK];
  int n_inner[OSTC_MAX_BLOCK], n_neg_jac[OSTC_MAX_BLOCK];
  /* restart (initial = 1): the step the flow-field file was written at
   * (output/message/step.txt, src/Initialize.cuh:206-210) and the residual
   * scales of that run (output/message/residual_scale.txt,
   * src/Driver.cu:25-31); 0 and 1 otherwise */
  int start_step, reserved;
  double res_scale[4];
  /* density and speed of the setup struct named freestream (rho_inf, v_inf
   * of src/BoundCond.cpp:96-99); the wall post-process divides by
   * 0.5 rho_inf v_inf^2 */
  double rho_inf, v_inf;
} ostc_case_info;
int ostc_case_read(const char *case_dir, ostc_case **out);
int ostc_case_destroy(ostc_case *c);
int ostc_case_get_info(const ostc_case *c, ostc_case_info *info);
int ostc_case_get_param(const ostc_case *c, ostc_param *out);
int ostc_case_get_mechanism(const ostc_case *c, ostc_mechanism_info *out);
int ostc_case_get_bc(const ostc_case *c, int block, ostc_bc *out);
/* registered inner faces of one block; *n_faces = their number */
int ostc_case_get_inner(const ostc_case *c, int block, ostc_inner_face *out, int max_faces, int *n_faces);
/* ngg-padded fields of a block in the zone layout: X, Y, Z, JAC, METRIC
 * (read) and the initial BV, SV (read, or overwrite when write != 0 -- the
 * way to start from fields the setup files cannot express). */
int ostc_case_field(ostc_case *c, int block, int field, double *host, int n_planes, int write);
/* wall points (x,y,z per point) used for the wall distance (src/Driver.cu:
 * 128-149); the count is ostc_case_info.n_wall_point */
int ostc_case_wall_points(const ostc_case *c, double *xyz, int max_points);
const char *ostc_case_error(void);

/* ---- time-loop driver: Driver<mix,turb> (src/Driver.cu:14-289) ----
 * create = the constructor (one zone per block, geometry, initial fields and
 * boundary programs uploaded; precision OSTC_FP32 or OSTC_FP64); initialize =
 * initialize_computation (src/Driver.cu:43-104) incl. the wall distance;
 * step = one iteration of steady_simulation with compute_residual on output
 * steps (res[4] normalised by the step-1 values, err_max their max,
 * src/Driver.cu:291-380); simulate = steady_simulation up to total_step or
 * convergence, with io != 0 printing the reference's screen lines and
 * writing history.dat, output/message/residual_scale.txt and
 * output/field/flowfield.plt under the case directory.  The driver keeps no
 * pointer to the case. */
typedef struct ostc_driver ostc_driver;
int ostc_driver_create(const ostc_case *c, int precision, ostc_driver **out);
int ostc_driver_initialize(ostc_driver *d);
int ostc_driver_step(ostc_driver *d, double res[4], int *res_computed, double *err_max);
int ostc_driver_simulate(ostc_driver *d, int io, int *steps_done, double *err_max);
/* the flow-field file of MPIIO::print_field (src/MPIIO.hpp:60-548): Tecplot
 * binary V112, one ORDERED zone per block without ghost layers, doubles;
 * variables x y z density u v w pressure temperature mach, species, tke
 * omega mut (RANS).  path NULL = <case dir>/output/field/flowfield.plt, the
 * file simulate(io != 0) writes every output_file steps and at convergence
 * and a setup with initial = 1 restarts from. */
int ostc_driver_write_flowfield(ostc_driver *d, const char *path);
/* Driver::post_process (src/Driver.cu:404-415): the procedures listed in the
 * setup's post_process array; 0 = wall_friction_heatflux_2d
 * (src/PostProcess.cu:6-47), which writes <case>/output/wall/
 * friction_heatflux.dat ("x cf qw" per i of the j = 0, k = 0 wall line, one
 * file rewritten per block so the last block's line remains, as the
 * reference does).  ostc_driver_simulate(io = 1) calls it after every
 * flow-field file. */
int ostc_driver_post_process(ostc_driver *d);
int ostc_driver_get_zone(ostc_driver *d, int block, ostc_zone **out);
int ostc_driver_get_ctx(ostc_driver *d, ostc_ctx **out);
int ostc_driver_get_step(const ostc_driver *d, int *step);
int ostc_driver_destroy(ostc_driver *d);
const char *ostc_driver_error(void);

/* ---- setup stages ---- */
/* compute_cv_from_bv<M,T> (src/FieldOperation.cuh:37-73) */
int ostc_compute_cv_from_bv(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* initialize_mut<M> (src/FieldOperation.cuh:111-137) */
int ostc_initialize_mut(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* compute_wall_distance (src/Driver.cu:461-482); wall_xyz holds n points */
int ostc_wall_distance(ostc_ctx *ctx, ostc_zone *zone, const double *wall_xyz, int n_wall_point, void *stream);

/* ---- post-processing ---- */
/* wall_friction_heatFlux_2d (src/PostProcess.cu:49-70): for i < mx on the
 * j = 0 wall of the k = 0 line, cf[i] = mu_l u_par |grad eta| / dyn_pressure
 * and qw[i] = lambda (T(i,1,0) - T(i,0,0)) |grad eta| (u_par: the velocity at
 * j = 1 along xi).  cf and qw are DEVICE arrays of mx doubles; dyn_pressure
 * is 0.5 rho_inf v_inf^2 of the freestream.  Stream-ordered. */
int ostc_wall_friction_heatflux(ostc_zone *zone, double dyn_pressure, double *cf, double *qw, void *stream);

/* ---- per-iteration stages (src/Driver.cu:232-272) ---- */
/* compute_source<M,T> (src/SourceTerm.cuh:11-40); writes dq (no memset needed) */
int ostc_compute_source(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* compute_inviscid_flux<M,T> (src/SchemeSelector.cuh:36-68) */
int ostc_inviscid_flux(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* compute_viscous_flux<M,T> (src/SchemeSelector.cuh:172-192) */
int ostc_viscous_flux(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* local_time_step<M,T> (src/TimeAdvanceFunc.cuh:21-71) */
int ostc_local_time_step(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* implicit_treatment<M,T> (src/ImplicitTreatmentHPP.cuh:10-59), DPLUR
 * (src/DPLURHPP.cuh:251-273) with param.dplur_inner_step sweeps */
int ostc_implicit_treatment(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* update_cv_and_bv<M,T> (src/FieldOperation.cuh:139-183) incl. the
 * T-from-e Newton solve (src/FieldOperation.cu:3-40) */
int ostc_update_cv_bv(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* limit_flow (src/TimeAdvanceFunc.cu:34-224) */
int ostc_limit_flow(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* DBoundCond::apply_boundary_conditions<M,T> (src/BoundCond.cuh:321-412) */
int ostc_apply_bc(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* inner_communication<M,T> for one face of zones[blk]
 * (src/DataCommunication.cuh:27-98) */
int ostc_inner_halo(ostc_ctx *ctx, ostc_zone **zones, int blk, const ostc_inner_face *face, void *stream);
/* update_physical_properties<M,T> (src/FieldOperation.cuh:75-109) */
int ostc_update_physical_properties(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* eliminate_k_gradient (src/FieldOperation.cu:42-62): a 2-D block (mz = 1)
 * copies its k = 0 plane of bv and sv into the k ghost layers; no-op for a
 * 3-D block.  ostc_steady_step calls it after the inner-face halos, as the
 * reference's loop does (src/Driver.cu:259-265); a driver calls it once
 * before the first boundary conditions (src/Driver.cu:59-65). */
int ostc_eliminate_k_gradient(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* store_last_step (src/TimeAdvanceFunc.cu:5-16) */
int ostc_store_last_step(ostc_ctx *ctx, ostc_zone *zone, void *stream);
/* compute_square_of_dbv + reduction_of_dv_squared (src/TimeAdvanceFunc.cu:
 * 18-32, src/Driver.cu:309-328): sums[4] of the squared changes of rho, |V|,
 * p, T over the interior of the given zones (synchronises the stream). */
int ostc_residual_sums(ostc_ctx *ctx, ostc_zone **zones, int n, double sums[4], void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the sums of
 * ostc_residual_sums (same kernels, same summation order, so the same bits),
 * written to the device array d_sums[4] on `stream` without synchronising and
 * without any allocation.  The partial sums go to a library-owned device
 * array of 2^20 doubles (8 MiB, one per GPU, held until the process ends):
 * 4 doubles per 64x4 cell tile of every zone, i.e. the sum over zones of
 * ceil(mx/64)*ceil(my/4)*mz must not exceed 262144, else OSTC_ERR_ARG and
 * nothing is launched.  Calls that overlap on one GPU (other streams, other
 * threads) share that array and race. */
int ostc_residual_sums_noalloc(ostc_ctx *ctx, ostc_zone **zones, int n, double *d_sums, void *stream);
/* Device-side counters since the last reset: [0] Newton iterations that hit
 * the 1000 cap, [1] cells repaired by limit_flow, [2] surrogate fallbacks
 * (synchronises the stream). */
int ostc_zone_counters(ostc_zone *zone, long long out[4], int reset, void *stream);

/* One steady iteration of Driver::steady_simulation over n zones
 * (src/Driver.cu:218-281) in the reference's stage order.  Stages are fused
 * with the same per-cell arithmetic: in SST zones without chemistry the
 * source kernel also forms the local time step (nothing the flux stages
 * write enters it), and in non-reacting zones the first DPLUR sweep forms
 * dq0 from dq as it loads it.  In EPI zones under DPLUR on the 9-species
 * kernels (at most 20 reactions) the finite-rate source follows the fluxes
 * in one kernel with dq0 and the chemistry block's factors: the species
 * source is added to the flux residual (the reference adds the fluxes to
 * it; the sum can differ in its last bit), and the EPI Jacobian is not
 * stored (OSTC_F_CHEM_JAC then holds the block's factors).  inner faces of block b are
 * inner[off_b .. off_b+n_inner[b]).  Returns the residual sums in
 * res_sums[4] when step%output_screen==0 or step==1 (else leaves them).
 * With n > 1 (and the stage profiler off) each zone's chain runs on a
 * library-owned stream forked from `stream` and joined back before the
 * inner-face exchange; zone 2k+1 starts once zone 2k's fluxes are done.
 * The result is the same as running the zones in order on `stream`. */
int ostc_steady_step(ostc_ctx *ctx, ostc_zone **zones, int n, const ostc_inner_face *inner, const int *n_inner,
                     int step, int output_screen, double res_sums[4], int *res_computed, void *stream);

/* ---- stage timing (no reference counterpart; the reference prints only the
 * wall time every output_screen steps, src/Driver.cu:382-401) ----
 * When enabled, every stage launched by the library is bracketed by HIP
 * events on its stream; ostc_profile_read synchronises and returns, per
 * stage, the summed milliseconds and the number of bracketed launches. */
int ostc_profile(ostc_ctx *ctx, int enable);
int ostc_profile_read(ostc_ctx *ctx, double *ms, long long *count, int n, int reset);
const char *ostc_profile_stage_name(int stage);
int ostc_profile_stage_count(void);

/* ---- chemistry points and the learned surrogate ---- */
/* Plain point-implicit chemistry step on a batch of states (FP64): rho from
 * EOS, (I - dt*J_chem) d(rhoY) = dt*omega (the reference's explicit EPI,
 * src/FiniteRateChem.cu:7-63,173-194,211-234, applied to the source alone).
 * All arrays are device pointers: t[n], p[n], y[n*ns], dt[n] -> dy[n*ns]. */
int ostc_chem_plain_step(ostc_ctx *ctx, long n, const double *t, const double *p, const double *y,
                         const double *dt, double *dy, void *stream);
/* omega[n*ns] (kg/m3/s) and the EPI Jacobian jac[n*ns*ns] at (rho,T,Y);
 * device pointers (src/FiniteRateChem.cu:7-63,173-194). */
int ostc_chem_source_batch(ostc_ctx *ctx, long n, const double *rho, const double *t, const double *y,
                           double *omega, double *jac, void *stream);

/* Chemistry step of every interior cell of a zone from its T, p, Y and local
 * dt (the zone's current state): mode 0 = the plain point-implicit step
 * (FP64), mode 1 = the learned surrogate with per-cell fallback to that step.
 * dy[n_interior*9] (device), cells in i-fastest order.  stats[4] (host, may be
 * NULL; mode 1 only): cells, fallbacks for input range, output range, element
 * balance.  No counterpart in the reference: it is the speed/accuracy harness
 * of the surrogate on the solver's own states. */
int ostc_chem_zone_step(ostc_ctx *ctx, ostc_zone *zone, ostc_surrogate *sur, int mode, double *dy,
                        long long stats[4], void *stream);

/* ---- learned chemistry surrogate (no counterpart in the reference) ----
 * MLP 12-128-128-128-128-n_out on the features
 * (1000/T, ln(p/1 atm), BoxCox_0.1(Y_k) for the 9 species, ln dt),
 * standardised with in_mean/in_std; its standardised outputs (out_mean,
 * out_std) encode the plain step's dY: kind 0 (n_out 10): log10 of
 * max(max_k|dY_k|, floor), then dY_k divided by that magnitude; kind 1
 * (n_out 9): asinh(dY_k/floor).  Inference: BF16 weights and activations,
 * FP32 accumulation (MFMA).  A cell falls back to the FP64 plain step when
 * its input lies outside [t_min,t_max] x [p_min,p_max] x [dt_min,dt_max]
 * (or Y_k < -y_tol), when Y_k+dY_k leaves [-y_tol, 1+y_tol] or a |dY_k|
 * exceeds dy_max (the output range the network is trusted with), or when an
 * element balance |sum_k elem_w[j][k] dY_k| exceeds
 * elem_tol * sum_jk |elem_w[j][k] dY_k| + elem_abs (elem_w = atoms/W_k).
 * Hidden activation: activation 0 = GELU (tanh form), 1 = hard-swish
 * z*clamp(z/6 + 1/2, 0, 1) (no transcendental: the cheaper inference).
 * Master weights are FP32, per layer W[out][in] row-major then b[out]. */
#define OSTC_SUR_GELU 0
#define OSTC_SUR_HSWISH 1
#define OSTC_SUR_IN 12
#define OSTC_SUR_MAX_OUT 16
#define OSTC_SUR_MAX_ELEM 4
typedef struct ostc_surrogate_desc {
  int n_out, kind, n_elem, activation;
  float in_mean[OSTC_SUR_IN], in_std[OSTC_SUR_IN];
  float out_mean[OSTC_SUR_MAX_OUT], out_std[OSTC_SUR_MAX_OUT];
  float loss_w[OSTC_SUR_MAX_OUT];             /* weights of the training loss */
  float elem_w[OSTC_SUR_MAX_ELEM * 9];        /* [element][species] */
  float floor;
  float t_min, t_max, p_min, p_max, dt_min, dt_max;
  float y_tol, elem_tol, elem_abs;
  float dy_max;                               /* trusted output range: |dY_k| <= dy_max */
} ostc_surrogate_desc;

int ostc_surrogate_n_params(int n_out);
/* weights: host FP32 [ostc_surrogate_n_params(n_out)]; max_batch: largest
 * training batch (multiple of 16; 0 = inference only). */
int ostc_surrogate_create(ostc_ctx *ctx, const ostc_surrogate_desc *desc, const float *weights, long max_batch,
                          ostc_surrogate **out);
int ostc_surrogate_destroy(ostc_surrogate *sur);
int ostc_surrogate_get_weights(ostc_surrogate *sur, float *weights);
int ostc_surrogate_set_weights(ostc_surrogate *sur, const float *weights);
/* dY of a batch (device arrays as ostc_chem_plain_step) with fallback;
 * stats as ostc_chem_zone_step. */
int ostc_surrogate_step(ostc_surrogate *sur, long n, const double *t, const double *p, const double *y,
                        const double *dt, double *dy, long long stats[4], void *stream);
/* Standardised network outputs raw[n*16] (device) of a batch, beside the
 * decoded dy (tests). */
int ostc_surrogate_forward(ostc_surrogate *sur, long n, const double *t, const double *p, const double *y,
                           const double *dt, float *raw, double *dy, void *stream);
/* Gradient of L = sum_b sum_o loss_w[o] (net_o(x_b) - target[r_b][o])^2 / nb
 * over the batch r_b = idx[b] (idx may be NULL: r_b = b) of the data set
 * (t, p, y, dt, target[*][16] standardised; device arrays) into grad
 * (device, n_params); *loss (host, may be NULL) = L.  nb: multiple of 16,
 * <= max_batch. */
int ostc_surrogate_gradients(ostc_surrogate *sur, long nb, const double *t, const double *p, const double *y,
                             const double *dt, const int *idx, const float *target, float *grad, double *loss,
                             void *stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8, rate lr) on that gradient. */
int ostc_surrogate_train_step(ostc_surrogate *sur, long nb, const double *t, const double *p, const double *y,
                              const double *dt, const int *idx, const float *target, float lr, double *loss,
                              void *stream);

/* ---- device memory utilities (no reference counterpart; used by callers
 * that hand batches of chemistry states to the library) ---- */
int ostc_device_alloc(void **ptr, long bytes);
int ostc_device_free(void *ptr);
int ostc_memcpy_h2d(void *dst, const void *src, long bytes);
int ostc_memcpy_d2h(void *dst, const void *src, long bytes);
/* Streaming copy of `bytes` (multiple of 16): the calibration kernel for HBM
 * traffic counters and the achievable-bandwidth reference. */
int ostc_stream_copy(void *dst, const void *src, long bytes, void *stream);

int ostc_device_sync(void);
const char *ostc_last_error(void);
const char *ostc_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
