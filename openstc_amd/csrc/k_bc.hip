// Boundary conditions (src/BoundCond.cuh:45-412) and the on-device inner-face
// halo exchange (src/DataCommunication.cuh:27-98).  Each face is one launch
// over its face points and layers, one thread per (point, layer).
#include <algorithm>

#include "ostc_kernels.h"
#include "ostc_thermo.h"

namespace ostc {

namespace {

struct FaceLaunch {
  int rs[3], re[3];
  int face, direction;
  int g0;  // first layer a launch handles: 0 = the face point, 1 = the first ghost layer
};

// Thread -> (face point, layer g): g = 0 is the face point itself, 1..ngg the
// ghost layers outside it.  Every (point, layer) pair is independent, so the
// layers run in parallel and each thread issues all its loads before its
// stores.  On an i-face four adjacent lanes take the layers of one row (the
// layers are adjacent in memory there); on j/k faces i varies fastest.
__device__ __forceinline__ bool face_layer(const FaceLaunch &f, int ngg, int p[3], int &g) {
  const int a = f.face == 0 ? 1 : 0, b = f.face == 2 ? 1 : 2;
  int ia, il;
  if (f.face == 0) {
    il = (int)(threadIdx.x & 3);
    ia = (int)(blockIdx.x * 16 + (threadIdx.x >> 2));
  } else {
    il = (int)blockIdx.z;
    ia = (int)(blockIdx.x * 64 + threadIdx.x);
  }
  const int ib = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  g = f.g0 + il;
  p[f.face] = f.rs[f.face];
  p[a] = f.rs[a] + ia;
  p[b] = f.rs[b] + ib;
  return g <= ngg && p[a] <= f.re[a] && p[b] <= f.re[b];
}

// compute_total_energy (src/FieldOperation.cuh:15-35) of a point whose bv and
// species cv are in registers
template <typename R, int NS>
__device__ __forceinline__ R total_energy_reg(const ostc_param *__restrict__ p, R rho, R u, R v, R w, R pr, R t,
                                              const R *cvs, R &vel) {
  const R vel2 = u * u + v * v + w * w;
  R e = (R)0.5 * rho * vel2;
  if constexpr (NS > 0) {
    R h[NS];
    enthalpy<R>(t, h, p, NS);
#pragma unroll
    for (int l = 0; l < NS; ++l) e += h[l] * cvs[l];
    e -= pr;
  } else {
    e += pr / (R)(kGammaAir - 1);
  }
  vel = sqrt(vel2);
  return e;
}

// apply_outflow (src/BoundCond.cuh:127-154): ghost layer g copies the face point
template <typename R, int NSC, bool RANS>
__global__ __launch_bounds__(256) void k_outflow(ZoneView<R> z, FaceLaunch f) {
  int pt[3], g;
  if (!face_layer(f, z.g, pt, g)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const int64_t P = z.plane, id = z.idx(pt[0], pt[1], pt[2]);
  const int64_t gi = z.idx(pt[0] + g * dir[0], pt[1] + g * dir[1], pt[2] + g * dir[2]);
  R bv[6], sv[NSC > 0 ? NSC : 1], mut = 0;
#pragma unroll
  for (int l = 0; l < 6; ++l) bv[l] = z.bv[l * P + id];
#pragma unroll
  for (int l = 0; l < NSC; ++l) sv[l] = z.sv[l * P + id];
  if (RANS) mut = z.mut[id];
#pragma unroll
  for (int l = 0; l < 6; ++l) z.bv[l * P + gi] = bv[l];
#pragma unroll
  for (int l = 0; l < NSC; ++l) z.sv[l * P + gi] = sv[l];
  if (RANS) z.mut[gi] = mut;
}

// apply_inflow (src/BoundCond.cuh:156-194): ghost layers take the inflow state
template <typename R, int NSC, bool RANS>
__global__ __launch_bounds__(256) void k_inflow(ZoneView<R> z, FaceLaunch f, const ostc_inflow *__restrict__ in) {
  int pt[3], g;
  if (!face_layer(f, z.g, pt, g)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const int64_t P = z.plane;
  const int64_t gi = z.idx(pt[0] + g * dir[0], pt[1] + g * dir[1], pt[2] + g * dir[2]);
  z.bv[gi] = (R)in->density;
  z.bv[P + gi] = (R)in->u;
  z.bv[2 * P + gi] = (R)in->v;
  z.bv[3 * P + gi] = (R)in->w;
  z.bv[4 * P + gi] = (R)in->pressure;
  z.bv[5 * P + gi] = (R)in->temperature;
#pragma unroll
  for (int l = 0; l < NSC; ++l) z.sv[l * P + gi] = (R)in->sv[l];
  if (RANS) z.mut[gi] = (R)in->mut;
}

// apply_symmetry (src/BoundCond.cuh:45-125): layer 0 = the face point (the
// interior neighbour's state with its normal velocity removed), layer g = the
// mirror of interior point -g with the normal velocity reversed
template <typename R, int NSC, bool RANS>
__global__ __launch_bounds__(256) void k_symmetry(ZoneView<R> z, FaceLaunch f) {
  int pt[3], g;
  if (!face_layer(f, z.g, pt, g)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const int64_t P = z.plane, id = z.idx(pt[0], pt[1], pt[2]);
  R kx = z.metric[(f.face * 3) * P + id], ky = z.metric[(f.face * 3 + 1) * P + id],
    kz = z.metric[(f.face * 3 + 2) * P + id];
  const R km = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= km;
  ky /= km;
  kz /= km;
  if (g == 0) {
    const int64_t in = z.idx(pt[0] - dir[0], pt[1] - dir[1], pt[2] - dir[2]);
    const R u1 = z.bv[P + in], v1 = z.bv[2 * P + in], w1 = z.bv[3 * P + in];
    const R rho = z.bv[in], pr = z.bv[4 * P + in], tt = z.bv[5 * P + in], e1 = z.cv[4 * P + in];
    R sv[NSC > 0 ? NSC : 1];
#pragma unroll
    for (int l = 0; l < NSC; ++l) sv[l] = z.sv[l * P + in];
    const R uk = kx * u1 + ky * v1 + kz * w1;
    const R ut = u1 - kx * uk, vt = v1 - ky * uk, wt = w1 - kz * uk;
    z.bv[P + id] = ut;
    z.bv[2 * P + id] = vt;
    z.bv[3 * P + id] = wt;
    z.vel[id] = sqrt(ut * ut + vt * vt + wt * wt);
    z.bv[id] = rho;
    z.bv[4 * P + id] = pr;
    z.bv[5 * P + id] = tt;
#pragma unroll
    for (int l = 0; l < NSC; ++l) z.sv[l * P + id] = sv[l];
    z.cv[id] = rho;
    z.cv[P + id] = rho * ut;
    z.cv[2 * P + id] = rho * vt;
    z.cv[3 * P + id] = rho * wt;
    z.cv[4 * P + id] = e1 - (R)0.5 * rho * (u1 * u1 + v1 * v1 + w1 * w1) + (R)0.5 * rho * (ut * ut + vt * vt + wt * wt);
#pragma unroll
    for (int l = 0; l < NSC; ++l) z.cv[(5 + l) * P + id] = rho * sv[l];
    return;
  }
  const int64_t gi = z.idx(pt[0] + g * dir[0], pt[1] + g * dir[1], pt[2] + g * dir[2]);
  const int64_t si = z.idx(pt[0] - g * dir[0], pt[1] - g * dir[1], pt[2] - g * dir[2]);
  const R rho = z.bv[si], u = z.bv[P + si], v = z.bv[2 * P + si], w = z.bv[3 * P + si];
  const R pr = z.bv[4 * P + si], tt = z.bv[5 * P + si];
  R sv[NSC > 0 ? NSC : 1], mut = 0;
#pragma unroll
  for (int l = 0; l < NSC; ++l) sv[l] = z.sv[l * P + si];
  if (RANS) mut = z.mut[si];
  const R uk = kx * u + ky * v + kz * w;
  const R ug = u - (R)2 * uk * kx, vg = v - (R)2 * uk * ky, wg = w - (R)2 * uk * kz;
  z.bv[gi] = rho;
  z.bv[P + gi] = ug;
  z.bv[2 * P + gi] = vg;
  z.bv[3 * P + gi] = wg;
  z.vel[gi] = sqrt(ug * ug + vg * vg + wg * wg);
  z.bv[4 * P + gi] = pr;
  z.bv[5 * P + gi] = tt;
#pragma unroll
  for (int l = 0; l < NSC; ++l) z.sv[l * P + gi] = sv[l];
  if (RANS) z.mut[gi] = mut;
}

// apply_wall (src/BoundCond.cuh:196-319): layer 0 = the wall point (no slip,
// pressure of the interior neighbour, wall or interior temperature, SST omega
// at the wall), layer g = the reflection of interior point -g; a ghost thread
// recomputes the wall omega it copies
template <typename R, int NS, bool RANS>
__global__ __launch_bounds__(256) void k_wall(ZoneView<R> z, FaceLaunch f, const ostc_param *__restrict__ p,
                                              int isothermal, double tw_in) {
  int pt[3], g;
  if (!face_layer(f, z.g, pt, g)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const int64_t P = z.plane, id = z.idx(pt[0], pt[1], pt[2]);
  const int64_t in = z.idx(pt[0] - dir[0], pt[1] - dir[1], pt[2] - dir[2]);
  R t_wall = (R)tw_in;
  if (!isothermal) t_wall = z.bv[5 * P + in];
  const R pw = z.bv[4 * P + in];
  R mw = (R)kMwAir;
  R yw[NS > 0 ? NS : 1];
  if constexpr (NS > 0) {
    mw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      yw[l] = z.sv[l * P + in];
      mw += yw[l] * inv_mw<R>(p, l);
    }
    mw = (R)1 / mw;
  }
  const R rho_w = pw * mw / (t_wall * (R)kRu);
  R om = 0;
  if constexpr (RANS) {
    R mu_w;
    if constexpr (NS > 0) {
      mixture_transport<R>(p, NS, t_wall, mw, yw, (const R *)nullptr, &mu_w, (R *)nullptr, (R *)nullptr);
    } else {
      mu_w = sutherland<R>(t_wall);
    }
    const R dy = z.wall_dist[in];
    if (dy > (R)1e-25)
      om = (R)60 * mu_w / (rho_w * (R)sst::beta_1 * dy * dy);
    else
      om = z.sv[(NS + 1) * P + in];
  }
  if (g == 0) {
    R cvs[NS > 0 ? NS : 1];
#pragma unroll
    for (int l = 0; l < NS; ++l) cvs[l] = yw[l] * rho_w;
    R vel;
    const R e = total_energy_reg<R, NS>(p, rho_w, (R)0, (R)0, (R)0, pw, t_wall, cvs, vel);
#pragma unroll
    for (int l = 0; l < NS; ++l) z.sv[l * P + id] = yw[l];
    z.bv[id] = rho_w;
    z.bv[P + id] = 0;
    z.bv[2 * P + id] = 0;
    z.bv[3 * P + id] = 0;
    z.bv[4 * P + id] = pw;
    z.bv[5 * P + id] = t_wall;
    z.cv[id] = rho_w;
    z.cv[P + id] = 0;
    z.cv[2 * P + id] = 0;
    z.cv[3 * P + id] = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) z.cv[(5 + l) * P + id] = cvs[l];
    z.cv[4 * P + id] = e;
    z.vel[id] = vel;
    if constexpr (RANS) {
      z.sv[NS * P + id] = 0;
      z.cv[(NS + 5) * P + id] = 0;
      z.sv[(NS + 1) * P + id] = om;
      z.cv[(NS + 6) * P + id] = rho_w * om;
    }
    return;
  }
  const int64_t si = z.idx(pt[0] - g * dir[0], pt[1] - g * dir[1], pt[2] - g * dir[2]);
  const int64_t gi = z.idx(pt[0] + g * dir[0], pt[1] + g * dir[1], pt[2] + g * dir[2]);
  const R ui = z.bv[P + si], vi = z.bv[2 * P + si], wi = z.bv[3 * P + si];
  const R pi = z.bv[4 * P + si], ti = z.bv[5 * P + si];
  R ys[NS > 0 ? NS : 1];
  R mg = (R)kMwAir;
  if constexpr (NS > 0) {
    mg = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      ys[l] = z.sv[l * P + si];
      mg += ys[l] * inv_mw<R>(p, l);
    }
    mg = (R)1 / mg;
  }
  R tg = ti;
  if (isothermal) {
    tg = (R)2 * t_wall - ti;
    if (tg <= (R)0.1 * t_wall) tg = t_wall;
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) z.sv[l * P + gi] = ys[l];
  z.bv[gi] = pi * mg / (tg * (R)kRu);
  z.bv[P + gi] = -ui;
  z.bv[2 * P + gi] = -vi;
  z.bv[3 * P + gi] = -wi;
  z.bv[4 * P + gi] = pi;
  z.bv[5 * P + gi] = tg;
  if constexpr (RANS) {
    z.sv[NS * P + gi] = 0;
    z.sv[(NS + 1) * P + gi] = om;
    z.mut[gi] = 0;
  }
}

// update_bv_1_point (src/FieldOperation.cuh:185-215) incl. the Newton solve
template <typename R, int NS, int NT>
__device__ __forceinline__ void update_bv_point(const ZoneView<R> &z, const ostc_param *__restrict__ p, int64_t id) {
  const int64_t P = z.plane;
  const R rho = z.cv[id];
  const R rinv = (R)1.0 / rho;
  z.bv[id] = rho;
  const R u = z.cv[P + id] * rinv, v = z.cv[2 * P + id] * rinv, w = z.cv[3 * P + id] * rinv;
  z.bv[P + id] = u;
  z.bv[2 * P + id] = v;
  z.bv[3 * P + id] = w;
  const R vel2 = u * u + v * v + w * w;
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) z.sv[l * P + id] = z.cv[(5 + l) * P + id] * rinv;
  if constexpr (NS > 0) {
    double y[NS], mw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = (double)z.sv[l * P + id];
      mw += y[l] / p->mw[l];
    }
    mw = 1 / mw;
    const double gc = kRu / mw;
    const double e = (double)z.cv[4 * P + id] / (double)rho -
                     0.5 * ((double)u * u + (double)v * v + (double)w * w);
    double t = (double)z.bv[5 * P + id];
    double err = 1;
    int iter = 0;
    while (err > 1e-3 && iter++ < 1000) {
      double h[NS], cp[NS];
      enthalpy_cp<double>(t, h, cp, p, NS);
      double cpt = 0, hm = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        cpt += cp[l] * y[l];
        hm += h[l] * y[l];
      }
      const double t1 = t - (hm - gc * t - e) / (cpt - gc);
      err = fabs(1 - t1 / t);
      t = t1;
    }
    z.bv[5 * P + id] = (R)t;
    z.bv[4 * P + id] = (R)((double)rho * t * gc);
  } else {
    const R pr = (R)(kGammaAir - 1) * (z.cv[4 * P + id] - (R)0.5 * rho * vel2);
    z.bv[4 * P + id] = pr;
    z.bv[5 * P + id] = pr * (R)kMwAir * rinv / (R)kRu;
  }
  z.vel[id] = sqrt(vel2);
}

// inner_communication (src/DataCommunication.cuh:58-98)
template <typename R, int NS, int NT>
__global__ void k_inner(ZoneView<R> z, ZoneView<R> t, const ostc_param *__restrict__ p, ostc_inner_face f) {
  constexpr int NV = 5 + NS + NT;
  const int n[3] = {(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(blockIdx.y * blockDim.y + threadIdx.y),
                    (int)(blockIdx.z * blockDim.z + threadIdx.z)};
  if (n[0] >= f.n_point[0] || n[1] >= f.n_point[1] || n[2] >= f.n_point[2]) return;
  int idx[3], it[3], d[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    d[q] = f.loop_dir[q] * n[q];
    idx[q] = f.range_start[q] + d[q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) it[q] = f.target_start[q] + f.target_loop_dir[q] * d[f.src_tar[q]];
  const int face_dir = f.direction > 0 ? f.range_start[f.face] : f.range_end[f.face];
  const int64_t a = z.idx(idx[0], idx[1], idx[2]), b = t.idx(it[0], it[1], it[2]);
  if (idx[f.face] == face_dir) {
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const R ave = (R)0.5 * (t.cv[l * t.plane + b] + z.cv[l * z.plane + a]);
      z.cv[l * z.plane + a] = ave;
      t.cv[l * t.plane + b] = ave;
    }
    update_bv_point<R, NS, NT>(z, p, a);
  } else {
#pragma unroll
    for (int l = 0; l < 6; ++l) z.bv[l * z.plane + a] = t.bv[l * t.plane + b];
#pragma unroll
    for (int l = 0; l < NS + NT; ++l) z.sv[l * z.plane + a] = t.sv[l * t.plane + b];
  }
}

// launch shape of face_layer: layers g0..ngg of every face point
dim3 face_grid(const FaceLaunch &f, int ngg, dim3 &block) {
  const int a = f.face == 0 ? 1 : 0, b = f.face == 2 ? 1 : 2;
  const int na = f.re[a] - f.rs[a] + 1, nb = f.re[b] - f.rs[b] + 1;
  const int nl = ngg + 1 - f.g0;
  block = dim3(64, 4, 1);
  if (f.face == 0) return dim3((na + 15) / 16, (nb + 3) / 4, 1);  // 4 lanes (layers) per point
  return dim3((na + 63) / 64, (nb + 3) / 4, nl);
}

// Wall friction coefficient and heat flux along the j = 0 wall of the k = 0
// line (wall_friction_heatFlux_2d, src/PostProcess.cu:49-70): per i,
//   u_par = (xi_x u + xi_y v) / |(xi_x, xi_y)|  at j = 1,
//   cf    = mu_l(i,0,0) u_par |(eta_x, eta_y)| / q_dyn,
//   qw    = lambda(i,0,0) (T(i,1,0) - T(i,0,0)) |(eta_x, eta_y)|,
// evaluated in FP64 from the stored fields (metrics at the wall point).
template <typename R>
__global__ __launch_bounds__(128) void k_wall_cf_qw(ZoneView<R> z, double dyn_pressure, double *__restrict__ cf,
                                                    double *__restrict__ qw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= z.mx) return;
  const int64_t w = z.idx(i, 0, 0), a = z.idx(i, 1, 0);
  const double xi_x = z.m(w, 1, 1), xi_y = z.m(w, 1, 2);
  const double eta_x = z.m(w, 2, 1), eta_y = z.m(w, 2, 2);
  const double u = z.bv[a + z.plane], v = z.bv[a + 2 * z.plane];
  const double u_par = (xi_x * u + xi_y * v) / sqrt(xi_x * xi_x + xi_y * xi_y);
  const double grad_eta = sqrt(eta_x * eta_x + eta_y * eta_y);
  cf[i] = (double)z.mul[w] * (u_par * grad_eta) / dyn_pressure;
  const double t1 = z.bv[a + 5 * z.plane], t0 = z.bv[w + 5 * z.plane];
  qw[i] = (double)z.lambda[w] * (t1 - t0) * grad_eta;
}

}  // namespace

template <typename R>
int apply_bc(const ZoneView<R> &z, const ostc_param *dp, const CellOpts &o, const ostc_bc &bc,
             const ostc_inflow *d_inflow, hipStream_t st) {
  for (int fi = 0; fi < bc.n_face; ++fi) {
    const ostc_face &bf = bc.face[fi];
    FaceLaunch f;
    for (int q = 0; q < 3; ++q) {
      f.rs[q] = bf.range_start[q];
      f.re[q] = bf.range_end[q];
    }
    f.face = bf.face;
    f.direction = bf.direction;
    f.g0 = (bf.kind == 5 || bf.kind == 6) ? 1 : 0;  // inflow/outflow write only the ghost layers
    if (z.g > 3) return OSTC_ERR_ARG;               // four lanes per point on i-faces
    dim3 block;
    const dim3 grid = face_grid(f, z.g, block);
    const int nsc = o.ns + o.nt;
    const bool rans = o.nt > 0;
#define OSTC_BC_SC(KERNEL, ...)                                                              \
  do {                                                                                       \
    if (nsc == 0 && !rans) hipLaunchKernelGGL((KERNEL<R, 0, false>), grid, block, 0, st, __VA_ARGS__); \
    else if (nsc == 2 && rans) hipLaunchKernelGGL((KERNEL<R, 2, true>), grid, block, 0, st, __VA_ARGS__); \
    else if (nsc == 9 && !rans) hipLaunchKernelGGL((KERNEL<R, 9, false>), grid, block, 0, st, __VA_ARGS__); \
    else if (nsc == 11 && rans) hipLaunchKernelGGL((KERNEL<R, 11, true>), grid, block, 0, st, __VA_ARGS__); \
    else if (nsc == 20 && !rans) hipLaunchKernelGGL((KERNEL<R, 20, false>), grid, block, 0, st, __VA_ARGS__); \
    else if (nsc == 22 && rans) hipLaunchKernelGGL((KERNEL<R, 22, true>), grid, block, 0, st, __VA_ARGS__); \
    else return OSTC_ERR_ARG;                                                                \
  } while (0)
    switch (bf.kind) {
      case 6: OSTC_BC_SC(k_outflow, z, f); break;
      case 5: OSTC_BC_SC(k_inflow, z, f, d_inflow + bf.slot); break;
      case 3: OSTC_BC_SC(k_symmetry, z, f); break;
      case 2: {
        const ostc_wall &w = bc.wall[bf.slot];
        if (o.ns == 0 && !rans)
          hipLaunchKernelGGL((k_wall<R, 0, false>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else if (o.ns == 0 && rans)
          hipLaunchKernelGGL((k_wall<R, 0, true>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else if (o.ns == 9 && !rans)
          hipLaunchKernelGGL((k_wall<R, 9, false>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else if (o.ns == 9 && rans)
          hipLaunchKernelGGL((k_wall<R, 9, true>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else if (o.ns == 20 && !rans)
          hipLaunchKernelGGL((k_wall<R, 20, false>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else if (o.ns == 20 && rans)
          hipLaunchKernelGGL((k_wall<R, 20, true>), grid, block, 0, st, z, f, dp, w.isothermal, w.temperature);
        else
          return OSTC_ERR_ARG;
        break;
      }
      default: return OSTC_ERR_ARG;
    }
#undef OSTC_BC_SC
    if (hipGetLastError() != hipSuccess) return OSTC_ERR_HIP;
  }
  return OSTC_OK;
}

template <typename R>
int inner_halo(const ZoneView<R> &z, const ZoneView<R> &t, const ostc_param *dp, const CellOpts &o,
               const ostc_inner_face &f, hipStream_t st) {
  unsigned tpb[3], bpg[3];
  const int fd = f.face;
  if (fd >= 0 && fd < 3 && f.n_point[fd] <= 2 * z.g + 1) {
    // all layers of a face point in one workgroup, on adjacent lanes: along
    // i (an i face) a row's layers then share their cache lines; i first
    // among the other two directions, where it is one of them
    tpb[fd] = f.n_point[fd];
    int budget = 256 / (int)tpb[fd];
    for (int q = 0; q < 3; ++q) {
      if (q == fd) continue;
      const int want = q == 0 ? 64 : 16;
      tpb[q] = (unsigned)std::max(1, std::min(std::min(want, budget), f.n_point[q]));
      budget = std::max(1, budget / (int)tpb[q]);
    }
  } else {
    for (int q = 0; q < 3; ++q) tpb[q] = f.n_point[q] <= 2 * z.g + 1 ? 1 : 16;
  }
  for (int q = 0; q < 3; ++q) bpg[q] = (f.n_point[q] - 1) / tpb[q] + 1;
  const dim3 grid(bpg[0], bpg[1], bpg[2]), block(tpb[0], tpb[1], tpb[2]);
  if (o.ns == 0 && o.nt == 0)
    hipLaunchKernelGGL((k_inner<R, 0, 0>), grid, block, 0, st, z, t, dp, f);
  else if (o.ns == 0 && o.nt == 2)
    hipLaunchKernelGGL((k_inner<R, 0, 2>), grid, block, 0, st, z, t, dp, f);
  else if (o.ns == 9 && o.nt == 0)
    hipLaunchKernelGGL((k_inner<R, 9, 0>), grid, block, 0, st, z, t, dp, f);
  else if (o.ns == 9 && o.nt == 2)
    hipLaunchKernelGGL((k_inner<R, 9, 2>), grid, block, 0, st, z, t, dp, f);
  else if (o.ns == 20 && o.nt == 0)
    hipLaunchKernelGGL((k_inner<R, 20, 0>), grid, block, 0, st, z, t, dp, f);
  else if (o.ns == 20 && o.nt == 2)
    hipLaunchKernelGGL((k_inner<R, 20, 2>), grid, block, 0, st, z, t, dp, f);
  else
    return OSTC_ERR_ARG;
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

template <typename R>
int wall_friction_heatflux(const ZoneView<R> &z, double dyn_pressure, double *cf, double *qw, hipStream_t st) {
  hipLaunchKernelGGL((k_wall_cf_qw<R>), dim3((z.mx + 127) / 128), dim3(128), 0, st, z, dyn_pressure, cf, qw);
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

template int apply_bc<float>(const ZoneView<float> &, const ostc_param *, const CellOpts &, const ostc_bc &,
                             const ostc_inflow *, hipStream_t);
template int apply_bc<double>(const ZoneView<double> &, const ostc_param *, const CellOpts &, const ostc_bc &,
                              const ostc_inflow *, hipStream_t);
template int inner_halo<float>(const ZoneView<float> &, const ZoneView<float> &, const ostc_param *,
                               const CellOpts &, const ostc_inner_face &, hipStream_t);
template int inner_halo<double>(const ZoneView<double> &, const ZoneView<double> &, const ostc_param *,
                                const CellOpts &, const ostc_inner_face &, hipStream_t);
template int wall_friction_heatflux<float>(const ZoneView<float> &, double, double *, double *, hipStream_t);
template int wall_friction_heatflux<double>(const ZoneView<double> &, double, double *, double *, hipStream_t);

}  // namespace ostc
