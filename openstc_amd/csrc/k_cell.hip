// Cell-parallel stages of the per-iteration path: properties, sources,
// local time step, point-implicit / DPLUR treatment, update with the
// T-from-e Newton solve, limit_flow, residual reduction, wall distance.
#include <algorithm>

#include "ostc_kernels.h"
#include "ostc_thermo.h"

namespace ostc {

namespace {

template <typename R>
struct CellIdx {
  int i, j, k;
  bool ok;
};

// interior + `halo` layers, workgroups of 64 x 4 over (i, j)
template <typename R>
__device__ __forceinline__ CellIdx<R> cell_index(const ZoneView<R> &z, int halo) {
  CellIdx<R> c;
  const uint3 b = xcd_block();
  c.i = (int)(b.x * blockDim.x + threadIdx.x) - halo;
  c.j = (int)(b.y * blockDim.y + threadIdx.y) - halo;
  c.k = (int)b.z - halo;
  c.ok = c.i < z.mx + halo && c.j < z.my + halo && c.k < z.mz + halo;
  return c;
}

// ------------------------------------------------------------------ setup

// compute_cv_from_bv (src/FieldOperation.cuh:37-73) incl. compute_total_energy
// (src/FieldOperation.cuh:15-35); all points incl. ghosts
template <typename R, int NS, int NT>
__global__ __launch_bounds__(256) void k_cv_from_bv(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const auto c = cell_index(z, z.g);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const R rho = z.bv[id], u = z.bv[P + id], v = z.bv[2 * P + id], w = z.bv[3 * P + id];
  z.cv[id] = rho;
  z.cv[P + id] = rho * u;
  z.cv[2 * P + id] = rho * v;
  z.cv[3 * P + id] = rho * w;
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) z.cv[(5 + l) * P + id] = rho * z.sv[l * P + id];
  const R vel2 = u * u + v * v + w * w;
  R e = (R)0.5 * rho * vel2;
  if constexpr (NS > 0) {
    R h[NS];
    enthalpy<R>(z.bv[5 * P + id], h, p, NS);
#pragma unroll
    for (int l = 0; l < NS; ++l) e += h[l] * z.cv[(5 + l) * P + id];
    e -= z.bv[4 * P + id];
  } else {
    e += z.bv[4 * P + id] / (R)(kGammaAir - 1);
  }
  z.cv[4 * P + id] = e;
  z.vel[id] = sqrt(vel2);
}

// update_physical_properties (src/FieldOperation.cuh:75-109); interior + 1.
// One thread per point of the (mx+2) x (my+2) x (mz+2) box, numbered i
// fastest, so that every wave but the last is full (rows of mx+2 points in
// 64-lane workgroup rows would leave most of a fifth wave idle at mx = 256).
template <typename R, int NS>
__global__ __launch_bounds__(256) void k_props(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const unsigned ni = (unsigned)z.mx + 2, nj = (unsigned)z.my + 2, nk = (unsigned)z.mz + 2;
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ni * nj * nk) return;
  const unsigned row = q / ni;
  const int ci = (int)(q - row * ni) - 1, cj = (int)(row % nj) - 1, ck = (int)(row / nj) - 1;
  const int64_t id = z.idx(ci, cj, ck), P = z.plane;
  const R t = z.bv[5 * P + id];
  if constexpr (NS > 0) {
    R cp[NS], y[NS];
    heat_capacity<R>(t, cp, p, NS);
    R mw = 0, cp_tot = 0, cvv = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = z.sv[l * P + id];
      mw += y[l] * inv_mw<R>(p, l);
      cp_tot += y[l] * cp[l];
      cvv += y[l] * (cp[l] - r_over_mw<R>(p, l));
    }
    mw = (R)1 / mw;
    const R gam = cp_tot / cvv;
    z.cp[id] = cp_tot;
    z.gamma[id] = gam;
    const R snd = sqrt(gam * (R)kRu * t / mw);
    z.sound[id] = snd;
    R mu, lam, rhod[NS];
    mixture_transport<R>(p, NS, t, mw, y, cp, &mu, &lam, rhod);
    z.mul[id] = mu;
    z.lambda[id] = lam;
#pragma unroll
    for (int l = 0; l < NS; ++l) z.rho_d[l * P + id] = rhod[l];
    z.mach[id] = z.vel[id] / snd;
  } else {
    const R c_temp = (R)(kGammaAir * kRu / kMwAir);
    const R cp = (R)(kGammaAir * kRu / kMwAir / (kGammaAir - 1));
    const R snd = sqrt(c_temp * t);
    z.sound[id] = snd;
    const R mu = sutherland<R>(t);
    z.mul[id] = mu;
    z.lambda[id] = mu * cp / (R)p->Pr;
    z.mach[id] = z.vel[id] / snd;
  }
}

// velocity-gradient helper: 0.5*(m_1c*d_i + m_2c*d_j + m_3c*d_k) of plane f
template <typename R>
__device__ __forceinline__ R cgrad(const R *__restrict__ f, int64_t id, const int64_t s[3], R m1, R m2, R m3) {
  return (R)0.5 * (m1 * (f[id + s[0]] - f[id - s[0]]) + m2 * (f[id + s[1]] - f[id - s[1]]) +
                   m3 * (f[id + s[2]] - f[id - s[2]]));
}

// SST::compute_mut (src/SST.cu:6-58): rho*k from cv; interior + 1
template <typename R, int NS>
__global__ __launch_bounds__(256) void k_init_mut(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const auto c = cell_index(z, 1);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const int64_t s[3] = {1, (int64_t)z.nx, (int64_t)z.nx * z.ny};
  const R t = z.bv[5 * P + id];
  R mul = sutherland<R>(t);
  if constexpr (NS > 0) {
    R y[NS], mw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = z.sv[l * P + id];
      mw += y[l] * inv_mw<R>(p, l);
    }
    mw = (R)1 / mw;
    mixture_transport<R>(p, NS, t, mw, y, (const R *)nullptr, &mul, (R *)nullptr, (R *)nullptr);
  }
  R m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = z.metric[q * P + id];
  const R *u = z.bv + P, *v = z.bv + 2 * P, *w = z.bv + 3 * P;
  const R u_y = cgrad(u, id, s, m[1], m[4], m[7]), u_z = cgrad(u, id, s, m[2], m[5], m[8]);
  const R v_x = cgrad(v, id, s, m[0], m[3], m[6]), v_z = cgrad(v, id, s, m[2], m[5], m[8]);
  const R w_x = cgrad(w, id, s, m[0], m[3], m[6]), w_y = cgrad(w, id, s, m[1], m[4], m[7]);
  const R rhoK = z.cv[(NS + 5) * P + id];
  const R tke = z.sv[NS * P + id], omega = z.sv[(NS + 1) * P + id];
  const R vort = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
  const R rho = z.bv[id];
  R f2 = 1;
  const R dy = z.wall_dist[id];
  if (dy > (R)1e-25) {
    const R param1 = (R)2 * sqrt(tke) / ((R)0.09 * omega * dy);
    const R param2 = (R)500 * mul / (rho * dy * dy * omega);
    const R arg2 = dmax(param1, param2);
    f2 = tanh(arg2 * arg2);
  }
  R mut = 0;
  const R den = dmax((R)sst::a_1 * omega, vort * f2);
  if (den > (R)1e-25) mut = (R)sst::a_1 * rhoK / den;
  z.mut[id] = mut;
}

// ------------------------------------------------------------------ dt

// local_time_step (src/TimeAdvanceFunc.cuh:21-71) of one cell from its metric
// m, velocity, density, laminar and eddy viscosity; the viscous radius is not
// stored (the reference never reads visc_spectr_rad).  Returns the time step
// and the sum of the three inviscid spectral radii it stored.
template <typename R>
struct CellDt {
  R dt, sri;
};
template <typename R, int NS, int NT>
__device__ __forceinline__ CellDt<R> local_dt_cell(const ZoneView<R> &z, const ostc_param *__restrict__ p,
                                                   int64_t id, const R *m, R u, R v, R w, R rho, R mul, R mut) {
  const int64_t P = z.plane;
  const R gx = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const R ge = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
  const R gz = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
  const R U = u * m[0] + v * m[1] + w * m[2];
  const R V = u * m[3] + v * m[4] + w * m[5];
  const R W = u * m[6] + v * m[7] + w * m[8];
  const R a = z.sound[id];
  const bool three = z.mz > 1;
  const R r0 = fabs(U) + a * gx, r1 = fabs(V) + a * ge, r2 = three ? fabs(W) + a * gz : (R)0;
  z.spectr[id] = r0;
  z.spectr[P + id] = r1;
  z.spectr[2 * P + id] = r2;
  const R sri = r0 + r1 + r2;
  R gam = (R)kGammaAir;
  if constexpr (NS > 0) gam = z.gamma[id];
  const R c1 = dmax(gam, (R)(4.0 / 3.0)) / rho;
  R c2 = mul / (R)p->Pr;
  if constexpr (NT > 0) c2 += mut / (R)p->Prt;
  R srv = gx * gx + ge * ge;
  if (three) srv += gz * gz;
  srv *= c1 * c2;
  const R dt = (R)p->cfl / (sri + srv);
  z.dt[id] = dt;
  return {dt, sri};
}

template <typename R, int NS, int NT>
__global__ __launch_bounds__(256) void k_local_dt(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  R m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = z.metric[q * P + id];
  local_dt_cell<R, NS, NT>(z, p, id, m, z.bv[P + id], z.bv[2 * P + id], z.bv[3 * P + id], z.bv[id], z.mul[id],
                           NT > 0 ? z.mut[id] : (R)0);
}

// The chemistry block of DPLUR, (diag I - dt J), in the space of the species
// that react.  A species that takes part in no reaction (N2 of H2-air, the
// inert species appended to smaller mechanisms) has a zero row and column in
// J: a factorisation of the whole block never picks its row as a pivot (its
// column entries are zero), its L and U entries stay zero and its U diagonal
// stays diag, every update it enters subtracts an exact zero, and the solve
// gives it b / diag.  So k_dq0 factors only the NA x NA block of the NA
// species that react (ChemF64::act, in their order) and the sweeps solve that
// block: the same values as the whole 9 x 9 factors (the sign of a zero
// aside), with NA^2 doubles per cell instead of 81.  NA == NS takes the whole
// block in the species' own order (the 20-species bucket, whose appended
// inert species stay in it as zero rows and columns).  Pivots are kept as
// positions in the block, four bits each (eight bits above 16 species),
// packed into 32-bit planes of chem_piv.
template <int NA>
struct PivPack {
  static constexpr int bits = NA > 16 ? 8 : 4, per = 32 / bits, words = (NA + per - 1) / per;
};
template <int NA>
__device__ __forceinline__ void piv_pack(const int *piv, unsigned *w) {
  using PP = PivPack<NA>;
#pragma unroll
  for (int q = 0; q < PP::words; ++q) w[q] = 0;
#pragma unroll
  for (int r = 0; r < NA; ++r) w[r / PP::per] |= (unsigned)piv[r] << (PP::bits * (r % PP::per));
}
template <int NA>
__device__ __forceinline__ void piv_unpack(const unsigned *w, int *piv) {
  using PP = PivPack<NA>;
#pragma unroll
  for (int r = 0; r < NA; ++r) piv[r] = (int)((w[r / PP::per] >> (PP::bits * (r % PP::per))) & ((1u << PP::bits) - 1));
}
// the r-th species of the block
template <int NS, int NA>
__device__ __forceinline__ int block_species(const ChemF64 *ct, int r) {
  if constexpr (NA == NS)
    return r;
  else
    return ct->act[r];
}

// The DPLUR chemistry block of one cell, (diag I - dt J) with diag = 1 + dt *
// (sum of the spectral radii), formed from its whole EPI Jacobian a (NS x NS,
// row-major) in place and factored as a whole (pivots as row indices).
template <typename R, int NS>
__device__ __forceinline__ void epi_factor_whole(R dt, R sri, double *a, int *piv) {
  static_assert(NS <= 9, "whole-block factors in registers up to 9 species");
  const R diag = (R)1 + dt * sri;
  const double dtd = (double)dt, dg = (double)diag;
  double d[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) d[q] = dg - dtd * a[q * NS + q];
#pragma unroll
  for (int q = 0; q < NS * NS; ++q) a[q] = -dtd * a[q];
#pragma unroll
  for (int q = 0; q < NS; ++q) a[q * NS + q] = d[q];
  lu_factor_any<NS>(a, piv);
}

// The whole block's factors a and pivots stored as the factors of the block
// of the species that react (tiled, lu_tiled) and their pivots, what the
// sweeps' block solve reads.  Every row and column of a species that reacts
// in no reaction is zero in J, so the whole factors restricted to the
// species that react are the block's own (the sign of a zero aside), and a
// pivot is always a reacting species' row; its position in the block is the
// number of reacting species before it (ChemF64::act_mask).  Up to 9
// species: the pivots take four bits each, PivPack's layout for blocks of up
// to 16.
template <typename R, int NS>
__device__ __forceinline__ void epi_store_factors(const ZoneView<R> &z, const ChemF64 *__restrict__ ct, int64_t id,
                                                  const double *a, const int *piv) {
  static_assert(NS <= 9, "whole-block factors in registers up to 9 species");
  const int64_t P = z.plane;
  const int na = ct->n_act;
  const unsigned amask = ct->act_mask;
  unsigned w0 = 0, w1 = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    const int rm = ct->rank[m];
    if (rm < 0) continue;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const int rn = ct->rank[n];
      if (rn >= 0) z.chem_lu[lu_tiled(id, rm * na + rn, na * na)] = a[m * NS + n];
    }
    const unsigned v = (unsigned)__popc(amask & ((1u << piv[m]) - 1u));
    if (rm < 8)
      w0 |= v << (4 * rm);
    else
      w1 |= v << (4 * (rm - 8));
  }
  z.chem_piv[id] = (int)w0;
  if (na > 8) z.chem_piv[P + id] = (int)w1;
}

// ------------------------------------------------------------------ source

// compute_source (src/SourceTerm.cuh:11-40): SST source and mut
// (src/SST.cu:60-178), turbulent conductivity, finite-rate chemistry
// (src/FiniteRateChem.cu:7-63, FP64).  First writer of dq.  DT (SST zones in
// the time step): also the local time step of the cell from the metric,
// density, viscosity and mut this kernel holds -- local_time_step reads
// nothing the flux stages between them write.
// CR: 0 no chemistry; > 0 the two-pass chemistry for mechanisms of at most
// CR reactions (rates of progress in 2 CR doubles of LDS per thread: 80 KB
// per workgroup at CR = 20, two workgroups per CU); < 0 the one-pass form (any
// reaction count).  The LDS variants ask for two waves per SIMD, the most
// their LDS allows, which leaves 256 VGPRs per lane (the EPI Jacobian's sums
// stay in registers); the others ask for three.  TURB_ROWS: only the
// turbulence rows of dq are written (zones whose fluxes write the others
// first: no chemistry here, its source follows the fluxes).
template <typename R, int NS, int NT, int CR, bool DT, bool TURB_ROWS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CR > 0 && NS > 0 ? 2 : 3))) void k_source(ZoneView<R> z,
                                                                                        const ostc_param *__restrict__ p) {
  constexpr bool CHEM = CR != 0;

  // CR > 0: the rates of progress of the 256 cells of the workgroup
  __shared__ double qs_mem[CR > 0 && NS > 0 ? 2 * CR * 256 : 1];
  constexpr int NV = 5 + NS + NT;
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  R dq[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) dq[l] = 0;
  const R J = z.jac[id];
  if constexpr (NT > 0) {
    const int64_t s[3] = {1, (int64_t)z.nx, (int64_t)z.nx * z.ny};
    R m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = z.metric[q * P + id];
    const R *u = z.bv + P, *v = z.bv + 2 * P, *w = z.bv + 3 * P;
    const R u_x = cgrad(u, id, s, m[0], m[3], m[6]), u_y = cgrad(u, id, s, m[1], m[4], m[7]),
            u_z = cgrad(u, id, s, m[2], m[5], m[8]);
    const R v_x = cgrad(v, id, s, m[0], m[3], m[6]), v_y = cgrad(v, id, s, m[1], m[4], m[7]),
            v_z = cgrad(v, id, s, m[2], m[5], m[8]);
    const R w_x = cgrad(w, id, s, m[0], m[3], m[6]), w_y = cgrad(w, id, s, m[1], m[4], m[7]),
            w_z = cgrad(w, id, s, m[2], m[5], m[8]);
    const R rho = z.bv[id];
    const R *kf = z.sv + NS * P, *of = z.sv + (NS + 1) * P;
    const R omega = of[id], tke = kf[id];
    const R k_x = cgrad(kf, id, s, m[0], m[3], m[6]), k_y = cgrad(kf, id, s, m[1], m[4], m[7]),
            k_z = cgrad(kf, id, s, m[2], m[5], m[8]);
    const R o_x = cgrad(of, id, s, m[0], m[3], m[6]), o_y = cgrad(of, id, s, m[1], m[4], m[7]),
            o_z = cgrad(of, id, s, m[2], m[5], m[8]);
    const R inter = (R)2 * rho * (R)sst::sigma_omega2 / omega * (k_x * o_x + k_y * o_y + k_z * o_z);
    const R rhoK = rho * tke;
    const R vort = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
    R f1 = 1, f2 = 1;
    const R dy = z.wall_dist[id];
    if (dy > (R)1e-25) {
      const R param1 = sqrt(tke) / ((R)0.09 * omega * dy);
      const R d2 = dy * dy;
      const R param2 = (R)500 * z.mul[id] / (rho * d2 * omega);
      const R arg2 = dmax((R)2 * param1, param2);
      f2 = tanh(arg2 * arg2);
      const R cd = dmax((R)1e-20, inter);
      const R param3 = (R)4 * rho * (R)sst::sigma_omega2 * tke / (cd * d2);
      const R arg1 = dmin(dmax(param1, param2), param3);
      f1 = tanh(arg1 * arg1 * arg1 * arg1);
    }
    R mut = 0;
    const R den = dmax((R)sst::a_1 * omega, vort * f2);
    if (den > (R)1e-25) mut = (R)sst::a_1 * rhoK / den;
    z.mut[id] = mut;
    const R beta = (R)sst::beta_2 + (R)sst::delta_beta * f1;
    if (mut > (R)1e-25) {
      const R divU = u_x + v_y + w_z;
      const R prod_k = mut * ((R)2 * (u_x * u_x + v_y * v_y + w_z * w_z) - (R)(2.0 / 3) * divU * divU +
                              (u_y + v_x) * (u_y + v_x) + (u_z + w_x) * (u_z + w_x) + (v_z + w_y) * (v_z + w_y)) -
                       (R)(2.0 / 3) * rhoK * divU;
      const R diss_k = (R)sst::beta_star * rhoK * omega;
      dq[NS + 5] = J * (prod_k - diss_k);
      const R gam = (R)sst::gamma2 + (R)sst::delta_gamma * f1;
      const R prod_o = gam * rho / mut * prod_k + ((R)1 - f1) * inter;
      const R diss_o = beta * rho * omega * omega;
      dq[NS + 6] = J * (prod_o - diss_o);
    }
    if (p->turb_implicit == 1) {
      z.turb_jac[id] = (R)(-2 * sst::beta_star) * omega;
      z.turb_jac[P + id] = (R)-2 * beta * omega;
    }
    if constexpr (NS > 0)
      z.lambda_t[id] = mut * z.cp[id] / (R)p->Prt;
    else
      z.lambda_t[id] = mut * (R)(kGammaAir * kRu / kMwAir / (kGammaAir - 1)) / (R)p->Prt;
    if constexpr (DT) local_dt_cell<R, NS, NT>(z, p, id, m, u[id], v[id], w[id], rho, z.mul[id], mut);
  } else if constexpr (DT) {
    R m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = z.metric[q * P + id];
    local_dt_cell<R, NS, NT>(z, p, id, m, z.bv[P + id], z.bv[2 * P + id], z.bv[3 * P + id], z.bv[id], z.mul[id],
                             (R)0);
  }
  if constexpr (CHEM && NS > 0) {
    double y[NS], omega[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) y[l] = (double)z.sv[l * P + id];
    const double rho = (double)z.bv[id], t = (double)z.bv[5 * P + id];
    const int meth = p->chem_src_method;
    if constexpr (CR > 0) {
      // the flow and turbulence rows first: nothing of the above stays live
      // through the chemistry
#pragma unroll
      for (int l = 0; l < 5; ++l) z.dq[l * P + id] = dq[l];
#pragma unroll
      for (int l = 5 + NS; l < NV; ++l) z.dq[l * P + id] = dq[l];
      double *jac = z.chem_jac;
      // rates of progress in LDS, this thread's column of qs_mem
      double *qs = qs_mem + threadIdx.x + 64 * threadIdx.y;
      const int nr = min(p->n_reac, CR);
      const auto om = [&](int l, double v) { z.dq[(5 + l) * P + id] = (R)((double)J * v); };
      if (meth == 1) {
        chem_source_lds<NS, 1>(p, nr, CR, rho, t, y, qs, 256, om, [&](int m, const double *row) {
#pragma unroll
          for (int n = 0; n < NS; ++n) jac[(m * NS + n) * P + id] = row[n];
        });
      } else if (meth == 2) {
        chem_source_lds<NS, 2>(p, nr, CR, rho, t, y, qs, 256, om,
                               [&](int m, const double *d) { jac[m * P + id] = *d; });
      } else {
        chem_source_lds<NS, 0>(p, nr, CR, rho, t, y, qs, 256, om, [&](int, const double *) {});
      }
      (void)omega;
      return;
    } else if (meth == 1) {
      double jac[NS * NS];
      chem_source<NS, kMaxReac, 1>(p, NS, p->n_reac, rho, t, y, omega, jac);
#pragma unroll
      for (int q = 0; q < NS * NS; ++q) z.chem_jac[q * P + id] = jac[q];
    } else if (meth == 2) {
      double jd[NS];
      chem_source<NS, kMaxReac, 2>(p, NS, p->n_reac, rho, t, y, omega, jd);
#pragma unroll
      for (int q = 0; q < NS; ++q) z.chem_jac[q * P + id] = jd[q];
    } else {
      chem_source<NS, kMaxReac, 0>(p, NS, p->n_reac, rho, t, y, omega, nullptr);
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) dq[5 + l] = (R)((double)J * omega[l]);
  }
  static_assert(!TURB_ROWS || !CHEM, "the species rows carry the chemistry source");
#pragma unroll
  for (int l = TURB_ROWS ? 5 + NS : 0; l < NV; ++l) z.dq[l * P + id] = dq[l];
}

// ------------------------------------------------------------------ implicit

// explicit-mode point-implicit treatment (src/ImplicitTreatmentHPP.cuh:10-59,
// EPI src/FiniteRateChem.cu:211-234, DA :400-412, SST src/SST.cu:180-194)
template <typename R, int NS, int NT, bool CHEM>
__global__ __launch_bounds__(256) void k_explicit_implicit(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const R dt = z.dt[id];
  if constexpr (CHEM && NS > 0) {
    const double dtd = (double)dt;
    if (p->chem_src_method == 1) {
      double a[NS * NS], b[NS];
      int piv[NS];
#pragma unroll
      for (int q = 0; q < NS * NS; ++q) a[q] = -dtd * z.chem_jac[q * P + id];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        a[q * NS + q] = 1.0 - dtd * z.chem_jac[(q * NS + q) * P + id];
        b[q] = (double)z.dq[(5 + q) * P + id];
      }
      lu_factor_any<NS>(a, piv);
      lu_solve_any<NS>(a, piv, b);
#pragma unroll
      for (int q = 0; q < NS; ++q) z.dq[(5 + q) * P + id] = (R)b[q];
    } else if (p->chem_src_method == 2) {
#pragma unroll
      for (int q = 0; q < NS; ++q)
        z.dq[(5 + q) * P + id] = (R)((double)z.dq[(5 + q) * P + id] / (1.0 - dtd * z.chem_jac[q * P + id]));
    }
  }
  if constexpr (NT > 0) {
    if (p->turb_implicit == 1) {
      z.dq[(NS + 5) * P + id] /= (R)1 - dt * z.turb_jac[id];
      z.dq[(NS + 6) * P + id] /= (R)1 - dt * z.turb_jac[P + id];
    }
  }
}

// compute_DQ_0 (src/DPLURHPP.cuh:13-76): dq0 = dq / diag with the chemistry
// (EPI_for_dq0, src/FiniteRateChem.cu:236-253, LU factors kept for the sweeps)
// and SST (src/SST.cu:196-204) blocks treated point-implicitly.
template <typename R, int NS, int NT, bool CHEM, int NA = NS>
__global__ __launch_bounds__(256) void k_dq0(ZoneView<R> z, const ostc_param *__restrict__ p) {
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const R dt = z.dt[id];
  const R diag = (R)1 + dt * (z.spectr[id] + z.spectr[P + id] + z.spectr[2 * P + id]);
  const int nflow = (CHEM && NS > 0) ? 5 : 5 + NS;
  for (int l = 0; l < nflow; ++l) z.dq0[l * P + id] = z.dq[l * P + id] / diag;
  if constexpr (CHEM && NS > 0) {
    const double dtd = (double)dt, dg = (double)diag;
    const int meth = p->chem_src_method;
    if (meth == 1) {
      // the block of the species that react (NA of them, n_act at run time)
      const ChemF64 *ct = chem64(p);
      double a[NA * NA], b[NA];
      int piv[NA];
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int m = block_species<NS, NA>(ct, r);
#pragma unroll
        for (int c2 = 0; c2 < NA; ++c2)
          a[r * NA + c2] = -dtd * z.chem_jac[(m * NS + block_species<NS, NA>(ct, c2)) * P + id];
        a[r * NA + r] = dg - dtd * z.chem_jac[(m * NS + m) * P + id];
        b[r] = (double)z.dq[(5 + m) * P + id];
      }
      lu_factor_any<NA>(a, piv);
#pragma unroll
      for (int q = 0; q < NA * NA; ++q) z.chem_lu[q * P + id] = a[q];
      unsigned w[PivPack<NA>::words];
      piv_pack<NA>(piv, w);
#pragma unroll
      for (int q = 0; q < PivPack<NA>::words; ++q) z.chem_piv[q * P + id] = (int)w[q];
      lu_solve_any<NA>(a, piv, b);
#pragma unroll
      for (int r = 0; r < NA; ++r) z.dq0[(5 + block_species<NS, NA>(ct, r)) * P + id] = (R)b[r];
      if constexpr (NA < NS) {
        // species that react in no reaction: b / diag
        for (int m = 0; m < NS; ++m)
          if (ct->rank[m] < 0) z.dq0[(5 + m) * P + id] = (R)((double)z.dq[(5 + m) * P + id] / dg);
      }
    } else if (meth == 2) {
#pragma unroll
      for (int q = 0; q < NS; ++q)
        z.dq0[(5 + q) * P + id] = (R)((double)z.dq[(5 + q) * P + id] / (dg - dtd * z.chem_jac[q * P + id]));
    } else {
#pragma unroll
      for (int q = 0; q < NS; ++q) z.dq0[(5 + q) * P + id] = z.dq[(5 + q) * P + id] / diag;
    }
  }
  if constexpr (NT > 0) {
    if (p->turb_implicit == 1) {
      z.dq0[(NS + 5) * P + id] = z.dq[(NS + 5) * P + id] / (diag - dt * z.turb_jac[id]);
      z.dq0[(NS + 6) * P + id] = z.dq[(NS + 6) * P + id] / (diag - dt * z.turb_jac[P + id]);
    } else {
      z.dq0[(NS + 5) * P + id] = z.dq[(NS + 5) * P + id] / diag;
      z.dq0[(NS + 6) * P + id] = z.dq[(NS + 6) * P + id] / diag;
    }
  }
}

// The finite-rate chemistry of an EPI zone under DPLUR placed after the flux
// stages (the time step of ostc_steady_step on the 9-species kernels, at most
// 20 reactions; the SST source and the local time step ran before the fluxes
// in k_source, which wrote zeros to the species rows of dq).  Per cell: the
// source J * omega (src/FiniteRateChem.cu:7-63) added to the species rows of
// dq, which hold the flux residual -- the reference adds the two the other
// way round, so the sum can differ in its last bit --; the EPI Jacobian
// (reaction-major, in registers); the block (diag I - dt J) factored whole
// and its reacting species' factors and pivots stored for the sweeps; dq0 of
// every row (compute_DQ_0 and EPI_for_dq0, src/DPLURHPP.cuh:13-76,
// src/FiniteRateChem.cu:236-253) from the same factors in registers.  The
// Jacobian is never stored, and k_dq0's pass over it goes.
template <typename R, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_chem_dq0(
    ZoneView<R> z, const ostc_param *__restrict__ p) {
  constexpr int NS = 9, CR = 20;
  __shared__ double qs_mem[2 * CR * 256];
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const R dt = z.dt[id];
  const R sri = z.spectr[id] + z.spectr[P + id] + z.spectr[2 * P + id];
  const R diag = (R)1 + dt * sri;
  // flow and turbulence rows as k_dq0 forms them
#pragma unroll
  for (int l = 0; l < 5; ++l) z.dq0[l * P + id] = z.dq[l * P + id] / diag;
  if constexpr (NT > 0) {
    if (p->turb_implicit == 1) {
      z.dq0[(NS + 5) * P + id] = z.dq[(NS + 5) * P + id] / (diag - dt * z.turb_jac[id]);
      z.dq0[(NS + 6) * P + id] = z.dq[(NS + 6) * P + id] / (diag - dt * z.turb_jac[P + id]);
    } else {
      z.dq0[(NS + 5) * P + id] = z.dq[(NS + 5) * P + id] / diag;
      z.dq0[(NS + 6) * P + id] = z.dq[(NS + 6) * P + id] / diag;
    }
  }
  const R J = z.jac[id];
  double y[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = (double)z.sv[l * P + id];
  const double rho = (double)z.bv[id], t = (double)z.bv[5 * P + id];
  double *qs = qs_mem + threadIdx.x + 64 * threadIdx.y;
  const int nr = min(p->n_reac, CR);
  // the block first (factored and stored, its factors kept in registers),
  // then the production rates: the species rows of dq with the source added,
  // at the bulk precision, then the solve
  int piv[NS];
  double b[NS];
  const auto om = [&](int l, double v) { b[l] = (double)(z.dq[(5 + l) * P + id] + (R)((double)J * v)); };
  chem_source_lds<NS, 4>(p, nr, CR, rho, t, y, qs, 256, om, [&](int phase, double *a) {
    if (phase == 0) {
      epi_factor_whole<R, NS>(dt, sri, a, piv);
      epi_store_factors<R, NS>(z, chem64(p), id, a, piv);
    } else {
      lu_solve_any<NS>(a, piv, b);
#pragma unroll
      for (int m = 0; m < NS; ++m) z.dq0[(5 + m) * P + id] = (R)b[m];
    }
  });
}

// Diagonal solve and write of one cell after a DPLUR sweep has gathered its
// off-diagonal sum acc (src/DPLURHPP.cuh:193-249); EPI_for_dqk
// (src/FiniteRateChem.cu:255-284) with the LU factors from k_dq0.  FIRST
// (non-reacting zones): the cell's dq0 comes in dq0v[0..NV) and is stored for
// the later sweeps; dq0v[NV], dq0v[NV+1] hold its SST Jacobian diagonal.
template <typename R, int NS, int NT, bool CHEM, bool FIRST>
__device__ __forceinline__ void dplur_finish(const ZoneView<R> &z, const ostc_param *__restrict__ p, int64_t id,
                                             const R *acc, R dt, R diag, R *__restrict__ dout, const R *dq0v) {
  static_assert(!(FIRST && CHEM && NS > 0), "the chemistry block needs k_dq0's LU factors");
  const int64_t P = z.plane;
  const int nflow = (CHEM && NS > 0) ? 5 : 5 + NS;
  auto q0 = [&](int l) -> R {
    if constexpr (FIRST) {
      z.dq0[l * P + id] = dq0v[l];
      return dq0v[l];
    } else {
      return z.dq0[l * P + id];
    }
  };
  for (int l = 0; l < nflow; ++l) dout[l * P + id] = q0(l) + dt * acc[l] / diag;
  if constexpr (CHEM && NS > 0) {
    const int meth = p->chem_src_method;
    const double dtd = (double)dt, dg = (double)diag;
    if (meth == 1) {
      // the species rows' off-diagonal sums; k_dplur_chem solves the block
#pragma unroll
      for (int q = 0; q < NS; ++q) dout[(5 + q) * P + id] = acc[5 + q];
    } else if (meth == 2) {
#pragma unroll
      for (int q = 0; q < NS; ++q)
        dout[(5 + q) * P + id] =
            (R)((double)z.dq0[(5 + q) * P + id] + dtd * (double)acc[5 + q] / (dg - dtd * z.chem_jac[q * P + id]));
    } else {
#pragma unroll
      for (int q = 0; q < NS; ++q) dout[(5 + q) * P + id] = z.dq0[(5 + q) * P + id] + dt * acc[5 + q] / diag;
    }
  }
  if constexpr (NT > 0) {
    if (p->turb_implicit == 1) {
      const R tj0 = FIRST ? dq0v[5 + NS + NT] : z.turb_jac[id];
      const R tj1 = FIRST ? dq0v[6 + NS + NT] : z.turb_jac[P + id];
      dout[(NS + 5) * P + id] = q0(NS + 5) + dt * acc[5 + NS] / (diag - dt * tj0);
      dout[(NS + 6) * P + id] = q0(NS + 6) + dt * acc[6 + NS] / (diag - dt * tj1);
    } else {
      dout[(NS + 5) * P + id] = q0(NS + 5) + dt * acc[5 + NS] / diag;
      dout[(NS + 6) * P + id] = q0(NS + 6) + dt * acc[6 + NS] / diag;
    }
  }
}

// Chemistry block of one DPLUR sweep of a reacting zone with EPI
// (EPI_for_dqk, src/FiniteRateChem.cu:255-284): the sweep kernel leaves the
// species rows' off-diagonal sums in dout; each cell solves its block with
// the LU factors k_dq0 kept, dout_s = dq0_s + (diag I - dt J)^-1 (dt acc_s),
// in place.  A kernel of its own: one streaming pass over the 9x9 FP64
// factors with every load issued at once, instead of loads that the sweep's
// neighbour exchange would leave exposed at the end of each plane.
template <typename R, int NS, int NA, bool TILED = false>
__global__ __launch_bounds__(256) void k_dplur_chem(ZoneView<R> z, const ostc_param *__restrict__ p,
                                                    R *__restrict__ dout) {
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const ChemF64 *ct = chem64(p);
  const R dt = z.dt[id];
  const double dtd = (double)dt;
  double a[NA * NA], b[NA];
  int piv[NA];
  unsigned w[PivPack<NA>::words];
#pragma unroll
  for (int q = 0; q < NA * NA; ++q) a[q] = z.chem_lu[TILED ? lu_tiled(id, q, NA * NA) : q * P + id];
#pragma unroll
  for (int q = 0; q < PivPack<NA>::words; ++q) w[q] = (unsigned)z.chem_piv[q * P + id];
  piv_unpack<NA>(w, piv);
#pragma unroll
  for (int r = 0; r < NA; ++r) b[r] = (double)dout[(5 + block_species<NS, NA>(ct, r)) * P + id] * dtd;
  lu_solve_any<NA>(a, piv, b);
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    const int m = block_species<NS, NA>(ct, r);
    dout[(5 + m) * P + id] = (R)((double)z.dq0[(5 + m) * P + id] + b[r]);
  }
  if constexpr (NA < NS) {
    // species that react in no reaction: their block row is diag, as k_dq0
    // formed it
    const double dg = (double)((R)1 + dt * (z.spectr[id] + z.spectr[P + id] + z.spectr[2 * P + id]));
    for (int m = 0; m < NS; ++m)
      if (ct->rank[m] < 0)
        dout[(5 + m) * P + id] = (R)((double)z.dq0[(5 + m) * P + id] + (double)dout[(5 + m) * P + id] * dtd / dg);
  }
}

// Per-cell inputs of the convective-Jacobian products of one DPLUR sweep
// (the neighbour part of src/DPLURHPP.cuh:127-191), loaded once per cell.
template <typename R, int NS, int NT>
struct JdCell {
  static constexpr int NV = 5 + NS + NT;
  R d[NV];
  R sv[NS + NT];
  R u, v, w, e, h, b2, bb;
  R dt, diag, tj0, tj1;  // first sweep only: the cell's dq0 scaling
};

// compute_DQ_0's diagonal scaling without a chemistry block
// (src/DPLURHPP.cuh:13-76, SST rows src/SST.cu:196-204): d = dq / diag, the
// same operations k_dq0 performs.  The first sweep of a non-reacting zone
// applies it to every dq it loads instead of reading a stored dq0, and keeps
// dt, diag and the SST Jacobian diagonal for the cell's own diagonal solve.
template <typename R, int NS, int NT>
__device__ __forceinline__ void dq0_scale(const ZoneView<R> &z, const ostc_param *__restrict__ p, int64_t q,
                                          JdCell<R, NS, NT> &c) {
  const int64_t P = z.plane;
  c.dt = z.dt[q];
  c.diag = (R)1 + c.dt * (z.spectr[q] + z.spectr[P + q] + z.spectr[2 * P + q]);
  c.tj0 = c.tj1 = 0;
#pragma unroll
  for (int l = 0; l < 5 + NS; ++l) c.d[l] = c.d[l] / c.diag;
  if constexpr (NT > 0) {
    if (p->turb_implicit == 1) {
      c.tj0 = z.turb_jac[q];
      c.tj1 = z.turb_jac[P + q];
      c.d[NS + 5] = c.d[NS + 5] / (c.diag - c.dt * c.tj0);
      c.d[NS + 6] = c.d[NS + 6] / (c.diag - c.dt * c.tj1);
    } else {
      c.d[NS + 5] = c.d[NS + 5] / c.diag;
      c.d[NS + 6] = c.d[NS + 6] / c.diag;
    }
  }
}

// FIRST: din is the right-hand side dq, scaled to dq0 on load
template <typename R, int NS, int NT, bool FIRST>
__device__ __forceinline__ void jd_load(const ZoneView<R> &z, const ostc_param *__restrict__ p,
                                        const R *__restrict__ din, int64_t q, JdCell<R, NS, NT> &c) {
  constexpr int NV = 5 + NS + NT;
  const int64_t P = z.plane;
#pragma unroll
  for (int l = 0; l < NV; ++l) c.d[l] = din[l * P + q];
  if constexpr (FIRST) dq0_scale<R, NS, NT>(z, p, q, c);
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) c.sv[l] = z.sv[l * P + q];
  c.u = z.bv[P + q];
  c.v = z.bv[2 * P + q];
  c.w = z.bv[3 * P + q];
  const R vel = z.vel[q];
  c.e = (R)0.5 * vel * vel;
  R gamma = (R)kGammaAir, b3 = 0, b4 = 0;
  if constexpr (NS > 0) {
    R hk[NS];
    const R t = z.bv[5 * P + q];
    enthalpy<R>(t, hk, p, NS);
    gamma = z.gamma[q];
    R h = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if constexpr (sizeof(R) == 4)
        b3 += t * r_over_mw<R>(p, l) * c.d[5 + l];
      else
        b3 += (R)kRu * t / (R)p->mw[l] * c.d[5 + l];
      b4 += hk[l] * c.d[5 + l];
      h += c.sv[l] * hk[l];
    }
    b3 *= gamma;
    b4 *= gamma - (R)1;
    c.h = h + c.e;
  } else {
    c.h = gamma / (gamma - (R)1) * z.bv[4 * P + q] / z.bv[q] + c.e;
  }
  c.bb = b3 - b4;
  c.b2 = (gamma - (R)1) * (c.e * c.d[0] - c.u * c.d[1] - c.v * c.d[2] - c.w * c.d[3] + c.d[4]);
}

// Contribution of cell q, seen as a neighbour across direction `dir`:
// w * (A_dir(q) dq(q) + pm_sr dq(q)); out = 0.5 * (JD + sr*d) (sign +1, the
// minus-side neighbour) or 0.5 * (JD - sr*d) (sign -1, the plus-side one).
template <typename R, int NS, int NT>
__device__ __forceinline__ void jd_dir(const ZoneView<R> &z, int64_t q, int dir, const JdCell<R, NS, NT> &c,
                                       R sr, R *plus, R *minus) {
  constexpr int NV = 5 + NS + NT;
  const int64_t P = z.plane;
  const R xi_x = z.metric[(dir * 3) * P + q], xi_y = z.metric[(dir * 3 + 1) * P + q],
          xi_z = z.metric[(dir * 3 + 2) * P + q];
  const R U = xi_x * c.u + xi_y * c.v + xi_z * c.w;
  const R b1 = xi_x * c.d[1] + xi_y * c.d[2] + xi_z * c.d[3] - U * c.d[0];
  R jd[NV];
  jd[0] = b1 + U * c.d[0];
  jd[1] = c.u * b1 + xi_x * c.b2 + U * c.d[1] + xi_x * c.bb;
  jd[2] = c.v * b1 + xi_y * c.b2 + U * c.d[2] + xi_y * c.bb;
  jd[3] = c.w * b1 + xi_z * c.b2 + U * c.d[3] + xi_z * c.bb;
  jd[4] = c.h * b1 + U * c.b2 + U * c.d[4] + U * c.bb;
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) jd[5 + l] = U * c.d[5 + l] + c.sv[l] * b1;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    if (plus) plus[l] = (R)0.5 * (jd[l] + sr * c.d[l]);
    if (minus) minus[l] = (R)0.5 * (jd[l] - sr * c.d[l]);
  }
}

// DPLUR_inner_iteration (src/DPLURHPP.cuh:127-249): one Jacobi sweep din ->
// dout.  A workgroup owns a 64 x 4 (i, j) tile and marches through a chunk of
// k planes: each cell's Jacobian products are formed once per direction and
// handed to its neighbours -- along i by lane shuffles inside the wave (one
// wave = one i row), along j through LDS, along k in registers from the
// previous plane.  Tile-edge neighbours outside the tile: the rows above and
// below by the edge rows themselves (whole-wave loads); the cells left and
// right of every row by 2*TJ lanes of an inner row's wave in one pass, handed
// over through LDS (single-lane halo loads per row would double the tile's vector
// memory instructions, which bound this kernel).  Off-diagonal sums are
// accumulated in the reference's order (i-, j-, k-, i+, j+, k+).  FIRST
// (first sweep of a non-reacting zone): din is the right-hand side dq, scaled
// to dq0 as it is loaded, and each cell's dq0 is stored for the later sweeps
// (no separate dq0 pass).
template <typename R, int NS, int NT, bool CHEM, int TJ, bool FIRST>
__global__ __launch_bounds__(64 * TJ) void k_dplur(ZoneView<R> z, const ostc_param *__restrict__ p,
                                               const R *__restrict__ din, R *__restrict__ dout, int kc) {
  constexpr int NV = 5 + NS + NT;
  __shared__ R sj[2][TJ][NV][64];  // [plus/minus][row][var][lane]
  __shared__ R si[2][TJ][NV];      // [left P_i / right M_i][row][var]
  const int tx = threadIdx.x, ty = threadIdx.y;
  const uint3 bk = xcd_block();
  const int i = (int)bk.x * 64 + tx, j = (int)bk.y * TJ + ty;
  const int k0 = (int)bk.z * kc, k1 = min(k0 + kc, z.mz);
  const bool ok = i < z.mx && j < z.my;
  const int ii = min(i, z.mx - 1), jj = min(j, z.my - 1);  // clamped (loads of idle lanes stay in bounds)
  const int64_t P = z.plane;
  const int64_t sk = (int64_t)z.nx * z.ny;
  const R *sr = z.spectr;
  JdCell<R, NS, NT> c;
  R pk[NV], acc[NV], accp[NV];
  R dq0p[FIRST ? NV + 2 : 1];  // FIRST: dq0 (+ SST diagonal) of the cell below, finished with the next plane
  R dtp = 0, diagp = 1;
  // P_k of the plane below the chunk
  if (k0 > 0) {
    const int64_t q = z.idx(ii, jj, k0 - 1);
    jd_load<R, NS, NT, FIRST>(z, p, din, q, c);
    jd_dir<R, NS, NT>(z, q, 2, c, sr[2 * P + q], pk, nullptr);
  }
  for (int k = k0; k < k1; ++k) {
    const int64_t id = z.idx(ii, jj, k);
    jd_load<R, NS, NT, FIRST>(z, p, din, id, c);
    R mk[NV], pnext[NV];
    jd_dir<R, NS, NT>(z, id, 2, c, sr[2 * P + id], pnext, mk);
    // finish the cell below: its plus-side k neighbour is this cell
    if (k > k0) {
#pragma unroll
      for (int l = 0; l < NV; ++l) accp[l] -= mk[l];
      if (ok) dplur_finish<R, NS, NT, CHEM, FIRST>(z, p, id - sk, accp, dtp, diagp, dout, dq0p);
    }
    // minus-side terms in the reference order: i-, j-, k-
    R pi[NV], mi[NV];
    jd_dir<R, NS, NT>(z, id, 0, c, sr[id], pi, mi);
    R pj[NV], mj[NV];
    jd_dir<R, NS, NT>(z, id, 1, c, sr[P + id], pj, mj);
    __syncthreads();  // previous plane's j exchange is consumed
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      sj[0][ty][l][tx] = pj[l];
      sj[1][ty][l][tx] = mj[l];
    }
    // i: neighbours' P_i / M_i from the adjacent lanes
    R ip[NV], im[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      ip[l] = __shfl_up(pi[l], 1);
      im[l] = __shfl_down(mi[l], 1);
    }
    if (ty == (TJ > 2 ? 1 : 0) && tx < 2 * TJ) {
      // lane 2r: the cell left of row r's segment (its P_i); lane 2r+1: the
      // cell right of it (its M_i).  An inner row does this (rows 0 and TJ-1
      // also evaluate the rows outside the tile), so that no wave carries
      // more than two cells' loads per plane.
      const int r = tx >> 1, side = tx & 1;
      const int ih = side ? (int)bk.x * 64 + 64 : (int)bk.x * 64 - 1;
      const int jr = min((int)bk.y * TJ + r, z.my - 1);
      if (ih >= 0 && ih < z.mx) {
        const int64_t q = z.idx(ih, jr, k);
        JdCell<R, NS, NT> e;
        R hp[NV], hm[NV];
        jd_load<R, NS, NT, FIRST>(z, p, din, q, e);
        jd_dir<R, NS, NT>(z, q, 0, e, sr[q], hp, hm);
#pragma unroll
        for (int l = 0; l < NV; ++l) si[side][r][l] = side ? hm[l] : hp[l];
      }
    }
    // j: rows at the tile edge evaluate the neighbour row themselves
    R jp[NV], jm[NV];
    if (ty == 0 && j > 0) {
      const int64_t q = z.idx(ii, j - 1, k);
      JdCell<R, NS, NT> e;
      jd_load<R, NS, NT, FIRST>(z, p, din, q, e);
      jd_dir<R, NS, NT>(z, q, 1, e, sr[P + q], jp, nullptr);
    }
    if (ty == TJ - 1 && j + 1 < z.my) {
      const int64_t q = z.idx(ii, j + 1, k);
      JdCell<R, NS, NT> e;
      jd_load<R, NS, NT, FIRST>(z, p, din, q, e);
      jd_dir<R, NS, NT>(z, q, 1, e, sr[P + q], nullptr, jm);
    }
    __syncthreads();
    if (tx == 0 && i > 0 && i - 1 < z.mx)
#pragma unroll
      for (int l = 0; l < NV; ++l) ip[l] = si[0][ty][l];
    if (tx == 63 && i + 1 < z.mx)
#pragma unroll
      for (int l = 0; l < NV; ++l) im[l] = si[1][ty][l];
    if (ty > 0)
#pragma unroll
      for (int l = 0; l < NV; ++l) jp[l] = sj[0][ty - 1][l][tx];
    if (ty < TJ - 1)
#pragma unroll
      for (int l = 0; l < NV; ++l) jm[l] = sj[1][ty + 1][l][tx];
    const bool hi = i > 0, hj = j > 0, hk = k > 0;
    const bool lo_i = i + 1 < z.mx, lo_j = j + 1 < z.my;
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      R a = 0;
      if (hi) a += ip[l];
      if (hj) a += jp[l];
      if (hk) a += pk[l];
      if (lo_i) a -= im[l];
      if (lo_j) a -= jm[l];
      acc[l] = a;
    }
    // this plane's plus-side k term arrives with the next plane
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      accp[l] = acc[l];
      pk[l] = pnext[l];
    }
    if constexpr (FIRST) {
      dtp = c.dt;
      diagp = c.diag;
#pragma unroll
      for (int l = 0; l < NV; ++l) dq0p[l] = c.d[l];
      dq0p[NV] = c.tj0;
      dq0p[NV + 1] = c.tj1;
    } else {
      dtp = z.dt[id];
      diagp = (R)1 + dtp * (sr[id] + sr[P + id] + sr[2 * P + id]);
    }
  }
  // last plane of the chunk: plus-side k neighbour from the plane above
  if (k1 > k0) {
    const int kl = k1 - 1;
    if (k1 < z.mz) {
      const int64_t q = z.idx(ii, jj, k1);
      jd_load<R, NS, NT, FIRST>(z, p, din, q, c);
      R mk[NV];
      jd_dir<R, NS, NT>(z, q, 2, c, sr[2 * P + q], nullptr, mk);
#pragma unroll
      for (int l = 0; l < NV; ++l) accp[l] -= mk[l];
    }
    if (ok) dplur_finish<R, NS, NT, CHEM, FIRST>(z, p, z.idx(ii, jj, kl), accp, dtp, diagp, dout, dq0p);
  }
}

// ------------------------------------------------------------------ update

// T from internal energy by the reference's Newton iteration
// (src/FieldOperation.cu:3-40), FP64; returns iterations used
template <int NS>
__device__ __forceinline__ int newton_t(const ostc_param *__restrict__ p, const double *y, double gas_const,
                                        double e, double &t) {
  double err = 1;
  int iter = 0;
  while (err > 1e-3 && iter++ < 1000) {
    double h[NS], cp[NS];
    enthalpy_cp<double>(t, h, cp, p, NS);
    double cp_tot = 0, hm = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp_tot += cp[l] * y[l];
      hm += h[l] * y[l];
    }
    const double e_t = hm - gas_const * t;
    const double cvv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cvv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return iter;
}

template <typename R, int NS, int NT>
__device__ __forceinline__ bool limit_bad_flow(const ostc_param *__restrict__ p, const R *var) {
  bool bad = false;
#pragma unroll
  for (int l = 0; l < 5; ++l)
    if (isnan(var[l]) || var[l] < (R)p->limit_ll[l] || var[l] > (R)p->limit_ul[l]) bad = true;
  return bad;
}

template <typename R>
__device__ __forceinline__ bool limit_bad_turb(const ostc_param *__restrict__ p, R t0, R t1) {
  return p->rans_model == 2 && (isnan(t0) || isnan(t1) || t0 < (R)0 || t1 < (R)0);
}

// update_cv_and_bv (src/FieldOperation.cuh:139-183).  With CHECK the cell is
// also classified as limit_flow's first step does (src/TimeAdvanceFunc.cu:
// 34-224) on the values just written, and a cell to repair is appended to
// the zone's repair list for k_limit_repair.
template <typename R, int NS, int NT, bool CHECK>
__global__ __launch_bounds__(256) void k_update(ZoneView<R> z, const ostc_param *__restrict__ p,
                                                const R *__restrict__ dqf, int slot) {
  constexpr int NV = 5 + NS + NT;
  const auto c = cell_index(z, 0);
  // the other list slot belongs to the previous time step, whose repairs are done
  if (CHECK && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0 &&
      threadIdx.z == 0)
    z.lf_count[slot ^ 1] = 0;
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const R dtj = z.dt[id] / z.jac[id];
  R cv[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) cv[l] = z.cv[l * P + id] + dqf[l * P + id] * dtj;
  if (z.mz == 1) cv[3] = 0;  // a 2-D block carries no z momentum
#pragma unroll
  for (int l = 0; l < NV; ++l) z.cv[l * P + id] = cv[l];
  const R rho = cv[0];
  const R rinv = (R)1.0 / rho;
  const R u = cv[1] * rinv, v = cv[2] * rinv, w = cv[3] * rinv;
  z.bv[id] = rho;
  z.bv[P + id] = u;
  z.bv[2 * P + id] = v;
  z.bv[3 * P + id] = w;
  const R vel2 = u * u + v * v + w * w;
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) z.sv[l * P + id] = cv[5 + l] * rinv;
  if constexpr (NS > 0) {
    double y[NS], mw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = (double)(cv[5 + l] * rinv);
      mw += y[l] / p->mw[l];
    }
    mw = 1 / mw;
    const double gas_const = kRu / mw;
    const double ud = (double)u, vd = (double)v, wd = (double)w;
    const double e = (double)cv[4] / (double)cv[0] - 0.5 * (ud * ud + vd * vd + wd * wd);
    double t = (double)z.bv[5 * P + id];
    const int it = newton_t<NS>(p, y, gas_const, e, t);
    if (it > 1000) atomicAdd(&z.counters[0], 1ull);
    const R tr = (R)t, pr = (R)((double)rho * t * gas_const);
    z.bv[5 * P + id] = tr;
    z.bv[4 * P + id] = pr;
    if constexpr (CHECK) {
      const R var[5] = {rho, u, v, w, pr};
      bool bad = limit_bad_flow<R, NS, NT>(p, var);
      if constexpr (NT > 0) bad = bad || limit_bad_turb<R>(p, cv[5 + NS] * rinv, cv[6 + NS] * rinv);
      if (bad) {
        const int e = atomicAdd(&z.lf_count[slot], 1);
        if (e < z.lf_cap) z.lf_idx[e] = id;
      }
    }
  } else {
    const R pr = (R)(kGammaAir - 1) * (cv[4] - (R)0.5 * rho * vel2);
    z.bv[4 * P + id] = pr;
    z.bv[5 * P + id] = pr * (R)kMwAir * rinv / (R)kRu;
    if constexpr (CHECK) {
      const R var[5] = {rho, u, v, w, pr};
      bool bad = limit_bad_flow<R, NS, NT>(p, var);
      if constexpr (NT > 0) bad = bad || limit_bad_turb<R>(p, cv[5 + NS] * rinv, cv[6 + NS] * rinv);
      if (bad) {
        const int e = atomicAdd(&z.lf_count[slot], 1);
        if (e < z.lf_cap) z.lf_idx[e] = id;
      }
    }
  }
  z.vel[id] = sqrt(vel2);
}

// ------------------------------------------------------------------ limit_flow

// limit_flow (src/TimeAdvanceFunc.cu:34-224), in two passes so that every
// repair reads the state as it was before the pass (the reference's single
// kernel lets a repaired cell be read by a neighbour in the same launch, an
// order-dependent race; here the result is deterministic and equal to the
// reference wherever bad cells are not adjacent).  The i-neighbour bound by
// mz is the reference's and is kept.
//   pass 1: classify; for bad cells compute the final bv/sv/cv into a list
//   pass 2: write the list back
// Repaired state of cell (i, j, k) into out[6 + NSC + NV] (bv, sv, cv);
// false if the cell needs no repair.  Reads the field only.
template <typename R, int NS, int NT>
__device__ bool limit_fix_cell(const ZoneView<R> &z, const ostc_param *__restrict__ p, int ci, int cj, int ck,
                               R *__restrict__ out) {
  constexpr int NSC = NS + NT, NV = 5 + NSC;
  struct {
    int i, j, k;
  } c{ci, cj, ck};
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  const double *ll = p->limit_ll, *ul = p->limit_ul;
  R var[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) var[l] = z.bv[l * P + id];
  const bool bad = limit_bad_flow<R, NS, NT>(p, var);
  bool tbad = false;
  if constexpr (NT > 0) tbad = limit_bad_turb<R>(p, z.sv[NS * P + id], z.sv[(NS + 1) * P + id]);
  if (!bad && !tbad) return false;
  R bv[6], sv[NSC > 0 ? NSC : 1], cv[NV];
#pragma unroll
  for (int l = 0; l < 6; ++l) bv[l] = z.bv[l * P + id];
#pragma unroll
  for (int l = 0; l < NSC; ++l) sv[l] = z.sv[l * P + id];
#pragma unroll
  for (int l = 0; l < NV; ++l) cv[l] = z.cv[l * P + id];
  if (bad) {
    atomicAdd(&z.counters[1], 1ull);
    R uv[5 + NSC];
#pragma unroll
    for (int l = 0; l < 5 + NSC; ++l) uv[l] = 0;
    int kn = 0;
    for (int ka = -1; ka < 2; ++ka) {
      const int k1 = c.k + ka;
      if (k1 < 0 || k1 >= z.mz) continue;
      for (int ja = -1; ja < 2; ++ja) {
        const int j1 = c.j + ja;
        if (j1 < 0 || j1 >= z.my) continue;
        for (int ia = -1; ia < 2; ++ia) {
          const int i1 = c.i + ia;
          if (i1 < 0 || i1 >= z.mz) continue;
          const int64_t n1 = z.idx(i1, j1, k1);
          bool nbad = false;
          R nb[5];
#pragma unroll
          for (int l = 0; l < 5; ++l) {
            nb[l] = z.bv[l * P + n1];
            if (isnan(nb[l]) || nb[l] < (R)ll[l] || nb[l] > (R)ul[l]) nbad = true;
          }
          if (nbad) continue;
#pragma unroll
          for (int l = 0; l < 5; ++l) uv[l] += nb[l];
#pragma unroll
          for (int l = 0; l < NSC; ++l) uv[5 + l] += z.sv[l * P + n1];
          ++kn;
        }
      }
    }
    if (kn > 0) {
      const R kinv = (R)1.0 / (R)kn;
#pragma unroll
      for (int l = 0; l < 5 + NSC; ++l) uv[l] *= kinv;
    } else {
#pragma unroll
      for (int l = 0; l < 5; ++l) uv[l] = dmin(dmax(var[l], (R)ll[l]), (R)ul[l]);
#pragma unroll
      for (int l = 0; l < NSC; ++l) uv[5 + l] = (R)p->sv_inf[l];
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) bv[l] = uv[l];
    cv[0] = uv[0];
    cv[1] = uv[0] * uv[1];
    cv[2] = uv[0] * uv[2];
    cv[3] = uv[0] * uv[3];
    R e = (R)0.5 * uv[0] * (uv[1] * uv[1] + uv[2] * uv[2] + uv[3] * uv[3]);
#pragma unroll
    for (int l = 0; l < NSC; ++l) {
      sv[l] = uv[5 + l];
      cv[5 + l] = uv[0] * uv[5 + l];
    }
    if constexpr (NS > 0) {
      R mw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) mw += uv[5 + l] * inv_mw<R>(p, l);
      mw = (R)1 / mw;
      const R t = uv[4] * mw / (uv[0] * (R)kRu);
      bv[5] = t;
      R h[NS];
      enthalpy<R>(t, h, p, NS);
#pragma unroll
      for (int l = 0; l < NS; ++l) e += h[l] * cv[5 + l];
      e -= uv[4];
    } else {
      bv[5] = uv[4] * (R)kMwAir / (uv[0] * (R)kRu);
      e += uv[4] / (R)(kGammaAir - 1);
    }
    cv[4] = e;
  }
  if constexpr (NT > 0) {
    const R t0 = sv[NS], t1 = sv[NS + 1];
    if (p->rans_model == 2 && (isnan(t0) || isnan(t1) || t0 < (R)0 || t1 < (R)0)) {
      R a0 = 0, a1 = 0;
      int kn = 0;
      for (int ka = -1; ka < 2; ++ka) {
        const int k1 = c.k + ka;
        if (k1 < 0 || k1 >= z.mz) continue;
        for (int ja = -1; ja < 2; ++ja) {
          const int j1 = c.j + ja;
          if (j1 < 0 || j1 >= z.my) continue;
          for (int ia = -1; ia < 2; ++ia) {
            const int i1 = c.i + ia;
            if (i1 < 0 || i1 >= z.mz) continue;
            const int64_t n1 = z.idx(i1, j1, k1);
            const R x0 = z.sv[NS * P + n1], x1 = z.sv[(NS + 1) * P + n1];
            if (isnan(x0) || isnan(x1) || x0 < (R)0 || x1 < (R)0) continue;
            a0 += x0;
            a1 += x1;
            ++kn;
          }
        }
      }
      if (kn > 0) {
        const R kinv = (R)1.0 / (R)kn;
        a0 *= kinv;
        a1 *= kinv;
      } else {
        a0 = t0 < (R)0 ? (R)p->sv_inf[NS] : t0;
        a1 = t1 < (R)0 ? (R)p->sv_inf[NS + 1] : t1;
      }
      sv[NS] = a0;
      sv[NS + 1] = a1;
      cv[NS + 5] = cv[0] * a0;
      cv[NS + 6] = cv[0] * a1;
    }
  }
#pragma unroll
  for (int l = 0; l < 6; ++l) out[l] = bv[l];
#pragma unroll
  for (int l = 0; l < NSC; ++l) out[6 + l] = sv[l];
#pragma unroll
  for (int l = 0; l < NV; ++l) out[6 + NSC + l] = cv[l];
  return true;
}

// pass 1 of the stand-alone limit_flow: classify every cell, repair into the list
template <typename R, int NS, int NT>
__global__ __launch_bounds__(256) void k_limit_check(ZoneView<R> z, const ostc_param *__restrict__ p) {
  constexpr int NSC = NS + NT, NV = 5 + NSC, NVAL = 6 + NSC + NV;
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  R vals[NVAL];
  if (!limit_fix_cell<R, NS, NT>(z, p, c.i, c.j, c.k, vals)) return;
  const int slot = atomicAdd(&z.lf_count[2], 1);
  if (slot >= z.lf_cap) return;
  z.lf_idx[slot] = z.idx(c.i, c.j, c.k);
  R *out = z.lf_vals + (int64_t)slot * NVAL;
#pragma unroll
  for (int l = 0; l < NVAL; ++l) out[l] = vals[l];
}

// pass 1 after a checking update: repair the listed cells (the update has
// already classified every cell).  An entry whose cell needs no repair after
// all (never, while the update and this kernel classify the same stored
// values) is marked -1 and skipped by the write-back.
template <typename R, int NS, int NT>
__global__ __launch_bounds__(256) void k_limit_repair(ZoneView<R> z, const ostc_param *__restrict__ p, int slot) {
  constexpr int NSC = NS + NT, NV = 5 + NSC, NVAL = 6 + NSC + NV;
  const int n = min(z.lf_count[slot], z.lf_cap);
  const int64_t nx = z.nx, ny = z.ny;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int64_t id = z.lf_idx[s];
    const int i = (int)(id % nx) - z.g, j = (int)((id / nx) % ny) - z.g, k = (int)(id / (nx * ny)) - z.g;
    R vals[NVAL];
    if (!limit_fix_cell<R, NS, NT>(z, p, i, j, k, vals)) {
      z.lf_idx[s] = -1;
      continue;
    }
    R *out = z.lf_vals + (int64_t)s * NVAL;
#pragma unroll
    for (int l = 0; l < NVAL; ++l) out[l] = vals[l];
  }
}

// pass 2: write the list of one slot back
template <typename R, int NS, int NT>
__global__ __launch_bounds__(256) void k_limit_apply(ZoneView<R> z, int slot) {
  constexpr int NSC = NS + NT, NV = 5 + NSC, NVAL = 6 + NSC + NV;
  const int n = min(z.lf_count[slot], z.lf_cap);
  const int64_t P = z.plane;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int64_t id = z.lf_idx[s];
    if (id < 0) continue;
    const R *v = z.lf_vals + (int64_t)s * NVAL;
#pragma unroll
    for (int l = 0; l < 6; ++l) z.bv[l * P + id] = v[l];
#pragma unroll
    for (int l = 0; l < NSC; ++l) z.sv[l * P + id] = v[6 + l];
#pragma unroll
    for (int l = 0; l < NV; ++l) z.cv[l * P + id] = v[6 + NSC + l];
  }
}

__global__ void k_zero_int(int *p) { *p = 0; }

// eliminate_k_gradient (src/FieldOperation.cu:42-62): a 2-D block (mz = 1)
// repeats its k = 0 plane of bv and sv into the k ghost layers, over the
// padded (i, j) range
template <typename R>
__global__ __launch_bounds__(256) void k_eliminate_k_gradient(ZoneView<R> z, int nsc) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - z.g;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - z.g;
  if (i >= z.mx + z.g || j >= z.my + z.g) return;
  const int64_t P = z.plane, src = z.idx(i, j, 0);
  for (int k = 1; k <= z.g; ++k) {
    const int64_t up = z.idx(i, j, k), dn = z.idx(i, j, -k);
    for (int l = 0; l < 6; ++l) {
      const R v = z.bv[l * P + src];
      z.bv[l * P + up] = v;
      z.bv[l * P + dn] = v;
    }
    for (int l = 0; l < nsc; ++l) {
      const R v = z.sv[l * P + src];
      z.sv[l * P + up] = v;
      z.sv[l * P + dn] = v;
    }
  }
}

// ------------------------------------------------------------------ residual

// store_last_step (src/TimeAdvanceFunc.cu:5-16)
template <typename R>
__global__ __launch_bounds__(256) void k_store_last(ZoneView<R> z) {
  const auto c = cell_index(z, 0);
  if (!c.ok) return;
  const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
  z.bv_last[id] = z.bv[id];
  z.bv_last[P + id] = z.vel[id];
  z.bv_last[2 * P + id] = z.bv[4 * P + id];
  z.bv_last[3 * P + id] = z.bv[5 * P + id];
}

// compute_square_of_dbv + first reduction pass (src/TimeAdvanceFunc.cu:18-32,
// src/Driver.cu:417-459): per-workgroup FP64 sums, fixed order.
template <typename R>
__global__ __launch_bounds__(256) void k_residual_partial(ZoneView<R> z, double *__restrict__ partial) {
  const auto c = cell_index(z, 0);
  double s[4] = {0, 0, 0, 0};
  if (c.ok) {
    const int64_t id = z.idx(c.i, c.j, c.k), P = z.plane;
    double d;
    d = (double)z.bv[id] - (double)z.bv_last[id];
    s[0] = d * d;
    d = (double)z.vel[id] - (double)z.bv_last[P + id];
    s[1] = d * d;
    d = (double)z.bv[4 * P + id] - (double)z.bv_last[2 * P + id];
    s[2] = d * d;
    d = (double)z.bv[5 * P + id] - (double)z.bv_last[3 * P + id];
    s[3] = d * d;
  }
  __shared__ double red[4][256];
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
#pragma unroll
  for (int l = 0; l < 4; ++l) red[l][t] = s[l];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int l = 0; l < 4; ++l) red[l][t] += red[l][t + w];
    __syncthreads();
  }
  if (t == 0) {
    const int64_t b = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
#pragma unroll
    for (int l = 0; l < 4; ++l) partial[b * 4 + l] = red[l][0];
  }
}

__global__ __launch_bounds__(256) void k_residual_final(const double *__restrict__ partial, int n,
                                                        double *__restrict__ out) {
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  double s[4] = {0, 0, 0, 0};
  for (int b = t; b < n; b += 256)
#pragma unroll
    for (int l = 0; l < 4; ++l) s[l] += partial[b * 4 + l];
#pragma unroll
  for (int l = 0; l < 4; ++l) red[l][t] = s[l];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int l = 0; l < 4; ++l) red[l][t] += red[l][t + w];
    __syncthreads();
  }
  if (t == 0)
#pragma unroll
    for (int l = 0; l < 4; ++l) out[l] = red[l][0];
}

// compute_wall_distance (src/Driver.cu:461-482): all points incl. ghosts.
// Wall points are staged through LDS in chunks.
template <typename R>
__global__ __launch_bounds__(256) void k_wall_dist(ZoneView<R> z, const double *__restrict__ wp, int n) {
  const auto c = cell_index(z, z.g);
  __shared__ double sw[3 * 256];
  double x = 0, y = 0, zz = 0, best = 1e+6;
  int64_t id = 0;
  if (c.ok) {
    id = z.idx(c.i, c.j, c.k);
    x = z.x[id];
    y = z.y[id];
    zz = z.z[id];
  }
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  for (int base = 0; base < n; base += 256) {
    const int cnt = n - base < 256 ? n - base : 256;
    __syncthreads();
    if (t < cnt) {
      sw[3 * t] = wp[3 * (base + t)];
      sw[3 * t + 1] = wp[3 * (base + t) + 1];
      sw[3 * t + 2] = wp[3 * (base + t) + 2];
    }
    __syncthreads();
    for (int l = 0; l < cnt; ++l) {
      const double dx = x - sw[3 * l], dy = y - sw[3 * l + 1], dz = zz - sw[3 * l + 2];
      const double d = dx * dx + dy * dy + dz * dz;
      if (best > d) best = d;
    }
  }
  if (c.ok) z.wall_dist[id] = (R)sqrt(best);
}

}  // namespace

// ------------------------------------------------------------------ launchers

// Workgroups of a DPLUR sweep kernel resident on the whole current device at
// once (registers and LDS decide it), cached per device.
template <typename R, int NS, int NT, bool CHEM, int TJ, bool FIRST>
int dplur_resident() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cached[dev] > 0) return cached[dev];
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dplur<R, NS, NT, CHEM, TJ, FIRST>, 64 * TJ, 0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  cached[dev] = per_cu * cus;
  return cached[dev];
}

// Planes per k chunk of a DPLUR sweep over `tiles` (i, j) tiles.  The chunks
// run in rounds of `resident` workgroups and a round lasts one chunk's march,
// so the sweep takes about rounds x planes-per-chunk plane steps: the count
// minimising that is taken (ties: fewer chunks, fewer boundary planes to
// reload), with at least 4 planes per chunk.  resident <= 0 (unknown) falls
// back to about 1024 workgroups.
inline int dplur_chunk_planes(int tiles, int mz, int resident) {
  const int nmax = std::max(1, (mz + 3) / 4);
  if (resident <= 0) {
    const int n = std::max(1, std::min((1024 + tiles - 1) / tiles, nmax));
    return (mz + n - 1) / n;
  }
  int best = mz;
  long long best_cost = -1;
  for (int n = 1; n <= nmax; ++n) {
    const int kc = (mz + n - 1) / n;
    const long long chunks = (mz + kc - 1) / kc;
    const long long rounds = ((long long)tiles * chunks + resident - 1) / resident;
    const long long cost = rounds * kc;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = kc;
    }
  }
  return best;
}

// species of the mechanism that react (the EPI block of DPLUR), 1..9
#define OSTC_DISPATCH_NA(o, CALL)                                  \
  do {                                                             \
    switch ((o).n_act) {                                           \
      case 1: { constexpr int NA_ = 1; CALL; } break;              \
      case 2: { constexpr int NA_ = 2; CALL; } break;              \
      case 3: { constexpr int NA_ = 3; CALL; } break;              \
      case 4: { constexpr int NA_ = 4; CALL; } break;              \
      case 5: { constexpr int NA_ = 5; CALL; } break;              \
      case 6: { constexpr int NA_ = 6; CALL; } break;              \
      case 7: { constexpr int NA_ = 7; CALL; } break;              \
      case 8: { constexpr int NA_ = 8; CALL; } break;              \
      case 9: { constexpr int NA_ = 9; CALL; } break;              \
      default: return OSTC_ERR_ARG;                                \
    }                                                              \
  } while (0)

#define OSTC_DISPATCH_NSNT(o, CALL)              \
  do {                                           \
    if ((o).ns == 0 && (o).nt == 0) {            \
      constexpr int NS_ = 0, NT_ = 0;            \
      CALL;                                      \
    } else if ((o).ns == 0 && (o).nt == 2) {     \
      constexpr int NS_ = 0, NT_ = 2;            \
      CALL;                                      \
    } else if ((o).ns == 9 && (o).nt == 0) {     \
      constexpr int NS_ = 9, NT_ = 0;            \
      CALL;                                      \
    } else if ((o).ns == 9 && (o).nt == 2) {     \
      constexpr int NS_ = 9, NT_ = 2;            \
      CALL;                                      \
    } else if ((o).ns == 20 && (o).nt == 0) {    \
      constexpr int NS_ = 20, NT_ = 0;           \
      CALL;                                      \
    } else if ((o).ns == 20 && (o).nt == 2) {    \
      constexpr int NS_ = 20, NT_ = 2;           \
      CALL;                                      \
    } else {                                     \
      return OSTC_ERR_ARG;                       \
    }                                            \
  } while (0)

static inline int last_err() { return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP; }

template <typename R>
int cv_from_bv(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st) {
  const auto g = cell_grid(z.mx + 2 * z.g, z.my + 2 * z.g, z.mz + 2 * z.g);
  OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_cv_from_bv<R, NS_, NT_>), g.grid, g.block, 0, st, z, dp));
  return last_err();
}

template <typename R>
int physical_properties(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st) {
  const long long n = (long long)(z.mx + 2) * (z.my + 2) * (z.mz + 2);
  if (n >= (1ll << 31)) return OSTC_ERR_ARG;  // the kernel numbers the points in 32 bits
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (o.ns == 0)
    hipLaunchKernelGGL((k_props<R, 0>), grid, block, 0, st, z, dp);
  else if (o.ns == 9)
    hipLaunchKernelGGL((k_props<R, 9>), grid, block, 0, st, z, dp);
  else if (o.ns == 20)
    hipLaunchKernelGGL((k_props<R, 20>), grid, block, 0, st, z, dp);
  else
    return OSTC_ERR_ARG;
  return last_err();
}

template <typename R>
int initialize_mut(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st) {
  if (o.nt == 0) return OSTC_OK;
  const auto g = cell_grid(z.mx + 2, z.my + 2, z.mz + 2);
  if (o.ns == 0)
    hipLaunchKernelGGL((k_init_mut<R, 0>), g.grid, g.block, 0, st, z, dp);
  else if (o.ns == 9)
    hipLaunchKernelGGL((k_init_mut<R, 9>), g.grid, g.block, 0, st, z, dp);
  else if (o.ns == 20)
    hipLaunchKernelGGL((k_init_mut<R, 20>), g.grid, g.block, 0, st, z, dp);
  else
    return OSTC_ERR_ARG;
  return last_err();
}

template <typename R>
int source_terms(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, bool force,
                 bool *wrote, bool with_dt) {
  const bool chem = o.mix_model == OSTC_FR;
  *wrote = false;
  if (with_dt && !source_has_dt(o) && !chem_after_fluxes(o)) return OSTC_ERR_ARG;
  if (o.nt == 0 && !chem && !force) return OSTC_OK;
  const auto g = cell_grid(z.mx, z.my, z.mz);
  if (chem) {
    // mechanisms of up to 20 reactions take the two-pass chemistry
    if (with_dt) {
      // chem_after_fluxes: the SST source and the time step only; the
      // chemistry follows the fluxes (k_chem_dq0)
      if (o.nt == 2)
        hipLaunchKernelGGL((k_source<R, 9, 2, 0, true, true>), g.grid, g.block, 0, st, z, dp);
      else
        hipLaunchKernelGGL((k_source<R, 9, 0, 0, true, true>), g.grid, g.block, 0, st, z, dp);
    } else if (o.n_reac <= 20)
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_source<R, NS_, NT_, 20, false>), g.grid, g.block, 0, st, z, dp));
    else
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_source<R, NS_, NT_, -1, false>), g.grid, g.block, 0, st, z, dp));
  } else if (with_dt) {
    // the SST rows of dq only: the fluxes write the others (keep_turb)
    if (o.ns == 0)
      hipLaunchKernelGGL((k_source<R, 0, 2, 0, true, true>), g.grid, g.block, 0, st, z, dp);
    else if (o.ns == 9)
      hipLaunchKernelGGL((k_source<R, 9, 2, 0, true, true>), g.grid, g.block, 0, st, z, dp);
    else
      hipLaunchKernelGGL((k_source<R, 20, 2, 0, true, true>), g.grid, g.block, 0, st, z, dp);
  } else {
    OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_source<R, NS_, NT_, 0, false>), g.grid, g.block, 0, st, z, dp));
  }
  *wrote = true;
  return last_err();
}

template <typename R>
int local_time_step(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st) {
  const auto g = cell_grid(z.mx, z.my, z.mz);
  OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_local_dt<R, NS_, NT_>), g.grid, g.block, 0, st, z, dp));
  return last_err();
}

template <typename R>
int implicit_treatment(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, R **result,
                       bool chem_late) {
  const auto g = cell_grid(z.mx, z.my, z.mz);
  const bool chem = o.mix_model == OSTC_FR;
  *result = z.dq;
  if (o.implicit_method == 0) {
    const bool need = (chem && o.chem != 0) || (o.nt > 0 && o.turb_implicit == 1);
    if (!need) return OSTC_OK;
    if (chem) {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_explicit_implicit<R, NS_, NT_, true>), g.grid, g.block, 0, st, z,
                                               dp));
    } else {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_explicit_implicit<R, NS_, NT_, false>), g.grid, g.block, 0, st,
                                               z, dp));
    }
    return last_err();
  }
  // DPLUR: dq -> dq0; sweeps ping-pong dq0 -> dqk -> dq -> dqk ...  Without
  // a chemistry block dq0 is a diagonal scaling, which the first sweep applies
  // to the dq it loads (and stores dq0 for the later sweeps): no dq0 pass.
  const bool fuse_dq0 = !chem && o.dplur_steps > 0;
  if (!fuse_dq0) {
    prof_mark(P_DQ0, st, true);
    if (chem) {
      if (o.ns == 9 && o.chem == 1 && chem_late) {
        // the chemistry after the fluxes, with dq0 (chem_after_fluxes)
        if (o.nt == 2)
          hipLaunchKernelGGL((k_chem_dq0<R, 2>), g.grid, g.block, 0, st, z, dp);
        else
          hipLaunchKernelGGL((k_chem_dq0<R, 0>), g.grid, g.block, 0, st, z, dp);
      } else if (o.ns == 9 && o.chem == 1) {
        // EPI block: factored in the space of the species that react
        if (o.nt == 2)
          OSTC_DISPATCH_NA(o, hipLaunchKernelGGL((k_dq0<R, 9, 2, true, NA_>), g.grid, g.block, 0, st, z, dp));
        else
          OSTC_DISPATCH_NA(o, hipLaunchKernelGGL((k_dq0<R, 9, 0, true, NA_>), g.grid, g.block, 0, st, z, dp));
      } else if (o.ns == 20 && o.chem == 1) {
        // the whole 20 x 20 block (non-reacting species as zero rows and
        // columns), factored with run-time loops
        if (o.nt == 2)
          hipLaunchKernelGGL((k_dq0<R, 20, 2, true, 20>), g.grid, g.block, 0, st, z, dp);
        else
          hipLaunchKernelGGL((k_dq0<R, 20, 0, true, 20>), g.grid, g.block, 0, st, z, dp);
      } else {
        OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_dq0<R, NS_, NT_, true>), g.grid, g.block, 0, st, z, dp));
      }
    } else {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_dq0<R, NS_, NT_, false>), g.grid, g.block, 0, st, z, dp));
    }
    prof_mark(P_DQ0, st, false);
    if (last_err()) return OSTC_ERR_HIP;
  }
  // k-marching sweeps: (64 x 4) tiles, k split into chunks sized from the
  // resident capacity of the kernel that runs most sweeps (or so that about
  // o.dplur_wg workgroups are in flight, when set); at least 4 planes per chunk
  constexpr int tj = 4;
  const int tiles = ((z.mx + 63) / 64) * ((z.my + tj - 1) / tj);
  int kc;
  if (o.dplur_wg > 0) {
    const int nkc = std::max(1, std::min((o.dplur_wg + tiles - 1) / tiles, (z.mz + 3) / 4));
    kc = (z.mz + nkc - 1) / nkc;
  } else {
    int resident = 0;
    if (chem)
      OSTC_DISPATCH_NSNT(o, resident = (dplur_resident<R, NS_, NT_, true, tj, false>()));
    else
      OSTC_DISPATCH_NSNT(o, resident = (dplur_resident<R, NS_, NT_, false, tj, false>()));
    kc = dplur_chunk_planes(tiles, z.mz, resident);
  }
  const dim3 sgrid((unsigned)((z.mx + 63) / 64), (unsigned)((z.my + tj - 1) / tj), (unsigned)((z.mz + kc - 1) / kc));
  const dim3 sblock(64, tj, 1);
  R *in = fuse_dq0 ? z.dq : z.dq0, *out = z.dqk;
  for (int s = 0; s < o.dplur_steps; ++s) {
    const int stage = fuse_dq0 && s == 0 ? P_DPLUR_FIRST : P_DPLUR;
    prof_mark(stage, st, true);
    if (chem) {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_dplur<R, NS_, NT_, true, tj, false>), sgrid, sblock, 0, st, z, dp,
                                               (const R *)in, out, kc));
    } else if (s == 0) {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_dplur<R, NS_, NT_, false, tj, true>), sgrid, sblock, 0, st, z, dp,
                                               (const R *)in, out, kc));
    } else {
      OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_dplur<R, NS_, NT_, false, tj, false>), sgrid, sblock, 0, st, z,
                                               dp, (const R *)in, out, kc));
    }
    prof_mark(stage, st, false);
    if (last_err()) return OSTC_ERR_HIP;
    if (chem && o.chem == 1 && o.ns > 0) {
      prof_mark(P_DPLUR_CHEM, st, true);
      if (o.ns == 9) {
        // factors stored by the source kernel are tiled (lu_tiled)
        if (chem_late)
          OSTC_DISPATCH_NA(o, hipLaunchKernelGGL((k_dplur_chem<R, 9, NA_, true>), g.grid, g.block, 0, st, z, dp, out));
        else
          OSTC_DISPATCH_NA(o, hipLaunchKernelGGL((k_dplur_chem<R, 9, NA_>), g.grid, g.block, 0, st, z, dp, out));
      } else if (o.ns == 20)
        hipLaunchKernelGGL((k_dplur_chem<R, 20, 20>), g.grid, g.block, 0, st, z, dp, out);
      else
        return OSTC_ERR_ARG;
      prof_mark(P_DPLUR_CHEM, st, false);
      if (last_err()) return OSTC_ERR_HIP;
    }
    in = out;
    out = (out == z.dqk) ? z.dq : z.dqk;
  }
  *result = in;
  return OSTC_OK;
}

// check_slot >= 0: also classify every cell for limit_flow into repair-list
// slot check_slot (0 or 1, alternating between time steps; the update clears
// the other slot), then call limit_flow with the same slot
template <typename R>
int update_cv_bv(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, const R *dq_final, hipStream_t st,
                 int check_slot) {
  const auto g = cell_grid(z.mx, z.my, z.mz);
  if (check_slot >= 0)
    OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_update<R, NS_, NT_, true>), g.grid, g.block, 0, st, z, dp, dq_final,
                                             check_slot & 1));
  else
    OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_update<R, NS_, NT_, false>), g.grid, g.block, 0, st, z, dp, dq_final,
                                             0));
  return last_err();
}

// limit_flow.  After a checking update (slot 0 or 1) only the listed cells
// are repaired; otherwise (slot < 0) every cell is classified first, into
// slot 2.
template <typename R>
int limit_flow(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, hipStream_t st, int slot) {
  if (slot >= 0) {
    slot &= 1;
    OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_limit_repair<R, NS_, NT_>), dim3(256), dim3(256), 0, st, z, dp,
                                             slot));
  } else {
    slot = 2;
    const auto g = cell_grid(z.mx, z.my, z.mz);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, z.lf_count + 2);
    OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_limit_check<R, NS_, NT_>), g.grid, g.block, 0, st, z, dp));
  }
  OSTC_DISPATCH_NSNT(o, hipLaunchKernelGGL((k_limit_apply<R, NS_, NT_>), dim3(256), dim3(256), 0, st, z, slot));
  return last_err();
}

template <typename R>
int eliminate_k_gradient(const ZoneView<R> &z, const CellOpts &o, hipStream_t st) {
  if (z.mz != 1) return OSTC_OK;
  const dim3 block(64, 4, 1);
  const dim3 grid((unsigned)((z.mx + 2 * z.g + 63) / 64), (unsigned)((z.my + 2 * z.g + 3) / 4), 1);
  hipLaunchKernelGGL((k_eliminate_k_gradient<R>), grid, block, 0, st, z, o.ns + o.nt);
  return last_err();
}

template <typename R>
int store_last_step(const ZoneView<R> &z, hipStream_t st) {
  const auto g = cell_grid(z.mx, z.my, z.mz);
  hipLaunchKernelGGL((k_store_last<R>), g.grid, g.block, 0, st, z);
  return last_err();
}

template <typename R>
int residual_partials(const ZoneView<R> &z, double *partial, hipStream_t st, int *nblocks) {
  const auto g = cell_grid(z.mx, z.my, z.mz);
  hipLaunchKernelGGL((k_residual_partial<R>), g.grid, g.block, 0, st, z, partial);
  *nblocks = (int)(g.grid.x * g.grid.y * g.grid.z);
  return last_err();
}

int residual_finalize(const double *partial, int nparts, double *out4, hipStream_t st) {
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, st, partial, nparts, out4);
  return last_err();
}

template <typename R>
int wall_distance(const ZoneView<R> &z, const double *wall_xyz, int n_wall, hipStream_t st) {
  const auto g = cell_grid(z.mx + 2 * z.g, z.my + 2 * z.g, z.mz + 2 * z.g);
  hipLaunchKernelGGL((k_wall_dist<R>), g.grid, g.block, 0, st, z, wall_xyz, n_wall);
  return last_err();
}

#define OSTC_INST(R)                                                                                            \
  template int cv_from_bv<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t);           \
  template int physical_properties<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t);  \
  template int initialize_mut<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t);       \
  template int source_terms<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t, bool,   \
                                bool *, bool);                                                                  \
  template int local_time_step<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t);      \
  template int implicit_treatment<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t,    \
                                     R **, bool);                                                               \
  template int update_cv_bv<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, const R *,            \
                               hipStream_t, int);                                                               \
  template int limit_flow<R>(const ZoneView<R> &, const ostc_param *, const CellOpts &, hipStream_t, int);      \
  template int store_last_step<R>(const ZoneView<R> &, hipStream_t);                                            \
  template int eliminate_k_gradient<R>(const ZoneView<R> &, const CellOpts &, hipStream_t);                      \
  template int residual_partials<R>(const ZoneView<R> &, double *, hipStream_t, int *);                         \
  template int wall_distance<R>(const ZoneView<R> &, const double *, int, hipStream_t);
OSTC_INST(float)
OSTC_INST(double)
#undef OSTC_INST

}  // namespace ostc
