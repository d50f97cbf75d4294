// Host-side launchers of the device kernels (one translation unit each).
#pragma once

#include "ostc_dev.h"

namespace ostc {

// Stage timing with HIP events on the launch stream (enabled per context by
// ostc_profile); prof_mark is a no-op when profiling is off.
enum ProfStage {
  P_SOURCE = 0, P_FLUX_I, P_FLUX_J, P_FLUX_K, P_DT, P_DQ0, P_DPLUR, P_EXPL_IMPL, P_UPDATE, P_LIMIT, P_BC,
  P_HALO, P_PROPS, P_STORE, P_RESID, P_FLUX_IJ, P_DPLUR_FIRST, P_DPLUR_CHEM, P_COUNT
};
void prof_mark(int stage, hipStream_t st, bool begin);

// Cell-parallel launch shape: 256 threads = 64 (i) x 4 (j) per workgroup.
struct CellGrid {
  dim3 grid, block;
};
inline CellGrid cell_grid(int ni, int nj, int nk) {
  CellGrid g;
  g.block = dim3(64, 4, 1);
  g.grid = dim3((ni + 63) / 64, (nj + 3) / 4, nk);
  return g;
}

// k_flux.hip
template <typename R>
// first_writer: the i direction writes dq instead of accumulating into it;
// keep_turb (with first_writer): it accumulates into the turbulence rows,
// which the SST source wrote, and writes the others
int flux_sweeps(const ZoneView<R> &z, const ostc_param *dp, int ns, int nt, bool inv, bool visc, bool first_writer,
                hipStream_t st, bool keep_turb = false);

// k_cell.hip
struct CellOpts {
  int ns, nt;      // species count, turbulence equations (0 or 2)
  int n_reac;      // reactions of the mechanism
  int mix_model;   // OSTC_AIR / OSTC_MIXTURE / OSTC_FR
  int chem;        // chem_src_method
  int rans_model, turb_implicit, implicit_method, dplur_steps;
  int dplur_wg;            // DPLUR sweep workgroup target (launch tuning; 0 = from the resident capacity)
  int n_act;               // species that take part in a reaction (the EPI block's size)
};
template <typename R>
int cv_from_bv(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st);
template <typename R>
int physical_properties(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st);
template <typename R>
int initialize_mut(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st);
// writes dq (first writer; zeros when the model has no source and force is set);
// with_dt: also the local time step (spectr, dt) and only the turbulence
// rows of dq (flux_sweeps then runs with first_writer and keep_turb), for
// the zones where
// source_has_dt or chem_after_fluxes holds -- the time step then skips
// local_time_step; in the second case the source kernel leaves out the
// chemistry and writes only the turbulence rows of dq (the fluxes then write
// the others: flux_sweeps with first_writer and keep_turb), and
// implicit_treatment, called with chem_late, adds the chemistry after the
// fluxes together with dq0 and the DPLUR chemistry block's factors
template <typename R>
int source_terms(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, bool force,
                 bool *wrote, bool with_dt = false);
inline bool source_has_dt(const CellOpts &o) { return o.nt > 0 && o.mix_model != OSTC_FR; }
// EPI chemistry under DPLUR on the 9-species kernels with the two-pass
// chemistry (at most 20 reactions)
inline bool chem_after_fluxes(const CellOpts &o) {
  return o.mix_model == OSTC_FR && o.chem == 1 && o.implicit_method == 1 && o.ns == 9 && o.n_reac <= 20 &&
         o.n_act > 0;
}
template <typename R>
int local_time_step(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st);
// leaves the final increment in *result (one of z.dq/z.dq0/z.dqk); chem_late
// (chem_after_fluxes zones whose source_terms ran with_dt): the chemistry
// source, its block's factors and dq0 are formed together after the fluxes
template <typename R>
int implicit_treatment(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, R **result,
                       bool chem_late = false);
// check_slot >= 0: also classify every cell for limit_flow into repair-list
// slot (check_slot & 1); then call limit_flow with the same slot.  Time steps
// alternate the slot.
template <typename R>
int update_cv_bv(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, const R *dq_final, hipStream_t st,
                 int check_slot = -1);
template <typename R>
int limit_flow(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, int slot = -1);
template <typename R>
int store_last_step(const ZoneView<R> &z, hipStream_t st);
// 2-D blocks: the k = 0 plane of bv/sv copied into the k ghost layers (no-op for mz > 1)
template <typename R>
int eliminate_k_gradient(const ZoneView<R> &z, const CellOpts &o, hipStream_t st);
// per-zone partial sums -> partial[nblocks*4]; returns nblocks
template <typename R>
int residual_partials(const ZoneView<R> &z, double *partial, hipStream_t st, int *nblocks);
int residual_finalize(const double *partial, int nparts, double *out4, hipStream_t st);
template <typename R>
int wall_distance(const ZoneView<R> &z, const double *wall_xyz, int n_wall, hipStream_t st);

// k_bc.hip
template <typename R>
int apply_bc(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, const ostc_bc &bc,
             const ostc_inflow *d_inflow, hipStream_t st);
template <typename R>
int inner_halo(const ZoneView<R> &z, const ZoneView<R> &t, const ostc_param *dp, const CellOpts &o,
               const ostc_inner_face &f, hipStream_t st);
// cf[i], qw[i] (i < mx, device FP64) along the j = 0, k = 0 wall line
template <typename R>
int wall_friction_heatflux(const ZoneView<R> &z, double dyn_pressure, double *cf, double *qw, hipStream_t st);

// k_chem.hip
int chem_plain_step(const ostc_param *dp, int ns, int nr, long n, const double *t, const double *p, const double *y,
                    const double *dt, double *dy, hipStream_t st);
int chem_source_batch(const ostc_param *dp, int ns, int nr, long n, const double *rho, const double *t,
                      const double *y, double *omega, double *jac, hipStream_t st);

}  // namespace ostc
