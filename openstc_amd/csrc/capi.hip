// C-ABI of the residual-and-update path: contexts, zones, field I/O and the
// per-stage entry points declared in include/ostc.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ostc_kernels.h"
#include "ostc_surrogate.h"
#include "ostc_thermo.h"

using namespace ostc;

namespace {

thread_local std::string g_err;

int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char *where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return OSTC_ERR_HIP;
}
#define HIP_TRY(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #call);   \
  } while (0)

// Event-based stage profiler (one active context at a time).
struct Profiler {
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> recs;
  hipEvent_t open[ostc::P_COUNT] = {};
  size_t next = 0;
  hipEvent_t get() {
    if (next == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[next++];
  }
  ~Profiler() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
Profiler *g_prof = nullptr;

// Partial sums of ostc_residual_sums_noalloc: a library-owned device array,
// one copy per GPU, held until the process ends, so that entry point
// allocates nothing.  In host code `scratch` is only a stand-in for the
// symbol: kernels and copies get the per-device address that
// hipGetSymbolAddress returns (g_scratch_ptr), never the name itself.
constexpr long kScratchDoubles = 1L << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Look the scratch array's address up on the current device (once per
// device; ostc_ctx_create calls this with the context's device current).
int scratch_setup(double **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(OSTC_ERR_ARG, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(OSTC_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  if (out) *out = g_scratch_ptr[dev];
  return OSTC_OK;
}

}  // namespace

namespace ostc {
void prof_mark(int stage, hipStream_t st, bool begin) {
  if (!g_prof || stage < 0 || stage >= P_COUNT) return;
  hipEvent_t e = g_prof->get();
  if (!e) return;
  (void)hipEventRecord(e, st);
  if (begin) {
    g_prof->open[stage] = e;
  } else if (g_prof->open[stage]) {
    g_prof->recs.push_back({stage, g_prof->open[stage], e});
    g_prof->open[stage] = nullptr;
  }
}
}  // namespace ostc

struct ostc_ctx {
  // host copy of the parameter block the kernels use: a mechanism of fewer
  // than its species bucket (9 or 20) is padded to that many (see pad_species)
  ostc_param host{};
  int user_ns = 0;  // species of the caller's mechanism
  int pad = 0;      // inert species added after them
  ostc_param *dev = nullptr;
  CellOpts opts{};
  double *partial = nullptr;  // residual partial sums
  size_t partial_cap = 0;      // in doubles
  double *out4 = nullptr;      // device [4]
  double *host4 = nullptr;     // pinned host [4]
  // ostc_steady_step on several zones: one stream per zone (default
  // priority) forked from and joined back to the caller's stream, so that
  // the kernels of independent zones can share the GPU
  std::vector<hipStream_t> zstream;
  std::vector<hipEvent_t> zdone;
  std::vector<hipEvent_t> zmid;  // a zone's fluxes are done
  hipEvent_t fork = nullptr;
};

struct ostc_zone {
  ostc_ctx *ctx = nullptr;
  ostc_zone_desc desc{};
  int64_t plane = 0;
  int esize = 8;
  void *ptr[OSTC_F_COUNT] = {};
  int planes[OSTC_F_COUNT] = {};
  int fsize[OSTC_F_COUNT] = {};  // element size per field
  int *piv = nullptr;
  unsigned long long *counters = nullptr;
  int *lf_count = nullptr;
  int64_t *lf_idx = nullptr;
  void *lf_vals = nullptr;
  // repair-list slot (0/1) of the next checking update in ostc_steady_step:
  // the zone alternates it itself, so callers may repeat or skip step numbers
  int lf_slot = 0;
  ostc_bc bc{};
  bool has_bc = false;
  ostc_inflow *d_inflow = nullptr;
  ZoneView<float> vf{};
  ZoneView<double> vd{};
};

namespace {

template <typename R>
ZoneView<R> make_view(ostc_zone *z) {
  ZoneView<R> v{};
  v.mx = z->desc.mx;
  v.my = z->desc.my;
  v.mz = z->desc.mz;
  v.g = z->desc.ngg;
  v.nx = v.mx + 2 * v.g;
  v.ny = v.my + 2 * v.g;
  v.nz = v.mz + 2 * v.g;
  v.plane = z->plane;
  auto P = [&](int f) { return static_cast<R *>(z->ptr[f]); };
  v.x = static_cast<const double *>(z->ptr[OSTC_F_X]);
  v.y = static_cast<const double *>(z->ptr[OSTC_F_Y]);
  v.z = static_cast<const double *>(z->ptr[OSTC_F_Z]);
  v.jac = P(OSTC_F_JAC);
  v.metric = P(OSTC_F_METRIC);
  v.wall_dist = P(OSTC_F_WALL_DIST);
  v.bv = P(OSTC_F_BV);
  v.sv = P(OSTC_F_SV);
  v.cv = P(OSTC_F_CV);
  v.bv_last = P(OSTC_F_BV_LAST);
  v.vel = P(OSTC_F_VEL);
  v.sound = P(OSTC_F_SOUND);
  v.mach = P(OSTC_F_MACH);
  v.mul = P(OSTC_F_MUL);
  v.lambda = P(OSTC_F_LAMBDA);
  v.rho_d = P(OSTC_F_RHO_D);
  v.gamma = P(OSTC_F_GAMMA);
  v.cp = P(OSTC_F_CP);
  v.mut = P(OSTC_F_MUT);
  v.lambda_t = P(OSTC_F_LAMBDA_T);
  v.turb_jac = P(OSTC_F_TURB_JAC);
  v.dq = P(OSTC_F_DQ);
  v.dq0 = P(OSTC_F_DQ0);
  v.dqk = P(OSTC_F_DQK);
  v.spectr = P(OSTC_F_SPECTR);
  v.dt = P(OSTC_F_DT);
  v.chem_jac = static_cast<double *>(z->ptr[OSTC_F_CHEM_JAC]);
  v.chem_lu = v.chem_jac;  // factors overwrite the Jacobian (it is not needed afterwards)
  v.chem_piv = z->piv;
  v.counters = z->counters;
  v.lf_count = z->lf_count;
  v.lf_idx = z->lf_idx;
  v.lf_vals = static_cast<R *>(z->lf_vals);
  v.lf_cap = (int)((int64_t)z->desc.mx * z->desc.my * z->desc.mz);
  return v;
}

bool fp64(const ostc_zone *z) { return z->desc.precision == OSTC_FP64; }
hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

// species counts the mixture kernels are instantiated for: a mechanism of
// n species runs on the smallest bucket that holds it, padded with inert
// species (pad_species)
constexpr int kSpeciesBuckets[] = {9, 20};
int species_bucket(int ns) {
  for (int b : kSpeciesBuckets)
    if (ns <= b) return b;
  return 0;
}

int check_param(const ostc_param &p) {
  if (p.mix_model < 0 || p.mix_model > 2) return fail(OSTC_ERR_ARG, "mix_model must be 0..2");
  if (p.turb_method == OSTC_RANS && p.rans_model != 2) return fail(OSTC_ERR_ARG, "only the k-omega SST RANS model");
  if (p.mix_model != OSTC_AIR && (p.n_spec < 1 || p.n_spec > OSTC_MAX_SPEC))
    return fail(OSTC_ERR_ARG, "mixtures of 1 to 20 species (the kernels are instantiated for 9 and 20; "
                              "other mechanisms are padded with inert species)");
  if (p.mix_model == OSTC_AIR && p.n_spec != 0) return fail(OSTC_ERR_ARG, "air has no species");
  const int nt = p.turb_method == OSTC_RANS ? 2 : 0;
  if (p.n_scalar != p.n_spec + nt || p.n_var != 5 + p.n_spec + nt)
    return fail(OSTC_ERR_ARG, "n_scalar/n_var inconsistent with n_spec and the turbulence model");
  if (p.limiter != 0 && p.limiter != 1) return fail(OSTC_ERR_ARG, "only minmod limiters (0, 1)");
  if (p.n_reac > OSTC_MAX_REAC || p.n_spec > OSTC_MAX_SPEC) return fail(OSTC_ERR_ARG, "mechanism too large");
  if (p.implicit_method != 0 && p.implicit_method != 1) return fail(OSTC_ERR_ARG, "implicit_method 0 or 1");
  return OSTC_OK;
}

}  // namespace

namespace {
// the increment left by the last implicit treatment, per zone
template <typename R>
int implicit_and_update(ostc_ctx *c, const ZoneView<R> &v, hipStream_t st, bool do_update) {
  R *res = nullptr;
  if (int r = implicit_treatment(v, c->dev, c->opts, st, &res)) return r;
  if (!do_update) {
    if (res != v.dq) {
      const size_t bytes = (size_t)v.plane * c->host.n_var * sizeof(R);
      if (hipMemcpyAsync(v.dq, res, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return OSTC_ERR_HIP;
    }
    return OSTC_OK;
  }
  return update_cv_bv(v, c->dev, c->opts, res, st);
}
}  // namespace

namespace ostc {
void make_thermo32(const ostc_param &p, ThermoF32 &out) {
  std::memset(&out, 0, sizeof(out));
  for (int i = 0; i < p.n_spec && i < kMaxSpec; ++i) {
    const double rw = p.mw[i] > 0 ? kRu / p.mw[i] : 0.0;
    out.t_low[i] = (float)p.t_low[i];
    out.t_mid[i] = (float)p.t_mid[i];
    for (int r = 0; r < 2; ++r) {
      const double *a = r == 0 ? &p.low_temp_coeff[i * 7] : &p.high_temp_coeff[i * 7];
      const double hc[6] = {a[5], a[0], a[1] / 2, a[2] / 3, a[3] / 4, a[4] / 5};
      for (int q = 0; q < 6; ++q) out.h[r][i][q] = (float)(rw * hc[q]);
      for (int q = 0; q < 5; ++q) out.cp[r][i][q] = (float)(rw * a[q]);
    }
    const double tl = p.t_low[i], *a = &p.low_temp_coeff[i * 7];
    const double h_tl = a[0] * tl + a[1] / 2 * tl * tl + a[2] / 3 * tl * tl * tl + a[3] / 4 * tl * tl * tl * tl +
                        a[4] / 5 * tl * tl * tl * tl * tl + a[5];
    const double cp_tl = a[0] + a[1] * tl + a[2] * tl * tl + a[3] * tl * tl * tl + a[4] * tl * tl * tl * tl;
    out.h_tl[i] = (float)(rw * h_tl);
    out.cp_tl[i] = (float)(rw * cp_tl);
    out.r_over_w[i] = (float)rw;
    out.inv_w[i] = p.mw[i] > 0 ? (float)(1.0 / p.mw[i]) : 0.0f;
    out.cp_lam[i] = (float)(1.25 * rw);
    out.lj_inv[i] = (float)p.LJ_potent_inv[i];
    out.vis_coeff[i] = (float)p.vis_coeff[i];
  }
  for (int q = 0; q < kMaxSpec * kMaxSpec; ++q) {
    out.wj_wi4[q] = (float)p.WjDivWi_to_One4th[q];
    out.sqrt_wi_wj8[q] = (float)p.sqrt_WiDivWjPl1Mul8[q];
  }
  // range bounds over the species for the wave-uniform range test (with no
  // species the test never passes)
  out.t_mid_min = 3.0e38f;
  out.t_mid_max = out.t_low_max = -3.0e38f;
  for (int i = 0; i < p.n_spec && i < kMaxSpec; ++i) {
    out.t_mid_min = std::min(out.t_mid_min, out.t_mid[i]);
    out.t_mid_max = std::max(out.t_mid_max, out.t_mid[i]);
    out.t_low_max = std::max(out.t_low_max, out.t_low[i]);
  }
  if (p.n_spec <= 0) out.t_mid_max = out.t_low_max = 3.0e38f;
}
// A mechanism of ns < 9 species runs on the 9-species kernels with inert
// species ns..8 appended: zero mass fraction everywhere, no reactions, zero
// NASA-7 coefficients, W = 1 and a positive viscosity coefficient.  Every
// term they enter is an exact zero (a product with their mass fraction,
// concentration or gradient, or a zero stoichiometric coefficient), so the
// real species see exactly the mechanism they were given.  The caller's
// freestream scalars [Y_0..Y_ns-1, k, omega] move to [Y..., 0.., k, omega].
ostc_param pad_species(const ostc_param &u, int bucket) {
  ostc_param p = u;
  const int ns = u.n_spec, nt = u.n_scalar - u.n_spec;
  for (int l = ns; l < bucket; ++l) {
    p.mw[l] = 1.0;
    p.t_low[l] = 200.0;
    p.t_mid[l] = 1000.0;
    p.t_high[l] = 6000.0;
    for (int q = 0; q < 7; ++q) p.low_temp_coeff[l * 7 + q] = p.high_temp_coeff[l * 7 + q] = 0.0;
    p.LJ_potent_inv[l] = 0.01;
    p.vis_coeff[l] = 1e-6;
  }
  for (int m = 0; m < bucket; ++m)
    for (int n = 0; n < bucket; ++n) {
      if (m < ns && n < ns) continue;
      // the Wilke factors of pairs with an inert species (finite; they only
      // ever multiply its zero mole fraction)
      p.WjDivWi_to_One4th[m * kMaxSpec + n] = std::pow(p.mw[n] / p.mw[m], 0.25);
      p.sqrt_WiDivWjPl1Mul8[m * kMaxSpec + n] = 1.0 / std::sqrt(8.0 * (1.0 + p.mw[m] / p.mw[n]));
    }
  for (int r = 0; r < u.n_reac && r < kMaxReac; ++r)
    for (int l = ns; l < bucket; ++l) {
      p.stoi_f[r * kMaxSpec + l] = p.stoi_b[r * kMaxSpec + l] = 0;
      p.third_body_coeff[r * kMaxSpec + l] = 0.0;
    }
  for (int l = 0; l < kMaxSpec + 2; ++l) p.sv_inf[l] = 0.0;
  for (int l = 0; l < ns; ++l) p.sv_inf[l] = u.sv_inf[l];
  for (int l = 0; l < nt; ++l) p.sv_inf[bucket + l] = u.sv_inf[ns + l];
  p.n_spec = bucket;
  p.n_scalar = bucket + nt;
  p.n_var = 5 + bucket + nt;
  return p;
}

void make_chem64(const ostc_param &p, ChemF64 &out) {
  std::memset(&out, 0, sizeof(out));
  for (int r = 0; r < p.n_reac && r < kMaxReac; ++r) {
    ReacF64 &q = out.reac[r];
    q.A = p.A[r];
    q.b = p.b[r];
    q.Ea = p.Ea[r];
    q.A2 = p.A2[r];
    q.b2 = p.b2[r];
    q.Ea2 = p.Ea2[r];
    q.troe_alpha = p.troe_alpha[r];
    q.troe_t1 = p.troe_t1[r];
    q.troe_t2 = p.troe_t2[r];
    q.troe_t3 = p.troe_t3[r];
    q.type = p.reac_type[r];
    q.order = p.reac_order[r];
  }
  for (int i = 0; i < p.n_spec && i < kMaxSpec; ++i) out.inv_mw[i] = p.mw[i] > 0 ? 1.0 / p.mw[i] : 0.0;
  for (int r = 0; r < p.n_reac && r < kMaxReac; ++r)
    for (int l = 0; l < p.n_spec && l < kMaxSpec; ++l) {
      const int f = p.stoi_f[r * kMaxSpec + l], b = p.stoi_b[r * kMaxSpec + l];
      out.nu_f[r * kMaxSpec + l] = f;
      out.nu_b[r * kMaxSpec + l] = b;
      if (f != 0 || b != 0) out.part[r] |= 1u << l;
      if (f != b) out.net[r] |= 1u << l;
    }
  for (int r = 0; r < p.n_reac && r < kMaxReac; ++r) {
    out.reac[r].part = out.part[r];
    out.reac[r].net = out.net[r];
  }
  for (int l = 0; l < kMaxSpec; ++l)
    for (int r = 0; r < p.n_reac && r < kMaxReac; ++r)
      if (out.net[r] & (1u << l)) out.jr[l][out.jn[l]++] = r;
  unsigned any = 0;
  for (int r = 0; r < p.n_reac && r < kMaxReac; ++r) any |= out.part[r];
  for (int l = 0; l < kMaxSpec; ++l) out.rank[l] = -1;
  for (int l = 0; l < p.n_spec && l < kMaxSpec; ++l)
    if (any & (1u << l)) {
      out.act[out.n_act] = l;
      out.act_mask |= 1u << l;
      out.rank[l] = out.n_act++;
    }
  for (int i = 0; i < p.n_spec && i < kMaxSpec; ++i) {
    for (int r = 0; r < 2; ++r) {
      const double *c = r == 0 ? &p.low_temp_coeff[i * 7] : &p.high_temp_coeff[i * 7];
      const double g[7] = {c[0] - c[6], -c[0], -0.5 * c[1], -c[2] / 6.0, -c[3] / 12.0, -c[4] * 0.05, c[5]};
      for (int q = 0; q < 7; ++q) out.g[r][i][q] = g[q];
    }
    const double tt = p.t_low[i], *c = &p.low_temp_coeff[i * 7];
    out.g_low[i] = c[0] * (1.0 - std::log(tt)) - 0.5 * c[1] * tt - c[2] * tt * tt / 6.0 -
                   c[3] * tt * tt * tt / 12.0 - c[4] * tt * tt * tt * tt * 0.05 + c[5] / tt - c[6];
  }
}
}  // namespace ostc

extern "C" {

const char *ostc_last_error(void) { return g_err.c_str(); }

const char *ostc_build_info(void) {
  return "openstc_amd libostc: HIP gfx950, FP32/FP64 zones, air and mixtures of 1-20 species (9- and 20-species kernels)";
}

int ostc_sizeof_param(void) { return (int)sizeof(ostc_param); }
int ostc_sizeof_bc(void) { return (int)sizeof(ostc_bc); }

int ostc_device_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return OSTC_OK;
}


int ostc_ctx_create(const ostc_param *param, ostc_ctx **out) {
  if (!param || !out) return fail(OSTC_ERR_ARG, "null argument");
  if (int r = check_param(*param)) return r;
  auto *c = new (std::nothrow) ostc_ctx;
  if (!c) return fail(OSTC_ERR_ALLOC, "host allocation");
  c->user_ns = param->n_spec;
  const int bucket = param->mix_model != OSTC_AIR ? species_bucket(param->n_spec) : 0;
  c->pad = (param->mix_model != OSTC_AIR && param->n_spec < bucket) ? bucket - param->n_spec : 0;
  c->host = c->pad ? pad_species(*param, bucket) : *param;
  param = &c->host;
  // the parameter block, followed by its FP32 thermodynamic and FP64
  // chemistry tables
  hipError_t e = hipMalloc(&c->dev, kParamBlockBytes);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipMalloc(param)");
  }
  e = hipMemcpy(c->dev, param, sizeof(ostc_param), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    ThermoF32 tf;
    make_thermo32(*param, tf);
    e = hipMemcpy(reinterpret_cast<char *>(c->dev) + kThermoF32Offset, &tf, sizeof(tf), hipMemcpyHostToDevice);
  }
  int n_act = 0;
  if (e == hipSuccess) {
    ChemF64 cf;
    make_chem64(*param, cf);
    n_act = cf.n_act;
    e = hipMemcpy(reinterpret_cast<char *>(c->dev) + kChemF64Offset, &cf, sizeof(cf), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && n_act == 0 && param->mix_model == OSTC_FR && param->chem_src_method == 1) {
    ostc_ctx_destroy(c);
    return fail(OSTC_ERR_ARG, "finite-rate chemistry with EPI: no species takes part in a reaction");
  }
  if (e == hipSuccess) e = hipMalloc(&c->out4, 4 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&c->host4, 4 * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    ostc_ctx_destroy(c);
    return hip_fail(e, "ctx setup");
  }
  if (int r = scratch_setup(nullptr)) {
    ostc_ctx_destroy(c);
    return r;
  }
  CellOpts &o = c->opts;
  o.ns = param->n_spec;
  o.nt = param->turb_method == OSTC_RANS ? 2 : 0;
  o.n_reac = param->n_reac;
  o.mix_model = param->mix_model;
  o.chem = param->chem_src_method;
  o.rans_model = param->rans_model;
  o.turb_implicit = param->turb_implicit;
  o.implicit_method = param->implicit_method;
  o.dplur_steps = param->dplur_inner_step;
  o.n_act = n_act;
  // launch tuning of the DPLUR sweep: 0 sizes its k chunks from the device's
  // resident workgroup capacity; the environment can fix a workgroup target
  o.dplur_wg = 0;
  if (const char *e = std::getenv("OSTC_DPLUR_WG")) o.dplur_wg = std::max(64, std::atoi(e));
  *out = c;
  return OSTC_OK;
}

int ostc_ctx_destroy(ostc_ctx *ctx) {
  if (!ctx) return OSTC_OK;
  for (auto e : ctx->zdone) (void)hipEventDestroy(e);
  for (auto e : ctx->zmid) (void)hipEventDestroy(e);
  for (auto q : ctx->zstream) (void)hipStreamDestroy(q);
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->dev) (void)hipFree(ctx->dev);
  if (ctx->partial) (void)hipFree(ctx->partial);
  if (ctx->out4) (void)hipFree(ctx->out4);
  if (ctx->host4) (void)hipHostFree(ctx->host4);
  delete ctx;
  return OSTC_OK;
}

int ostc_zone_create(ostc_ctx *ctx, const ostc_zone_desc *d, ostc_zone **out) {
  if (!ctx || !d || !out) return fail(OSTC_ERR_ARG, "null argument");
  if (d->mx < 2 || d->my < 2 || d->mz < 1)
    return fail(OSTC_ERR_ARG, "blocks need mx, my >= 2 and mz >= 1 (mz = 1: a 2-D block)");
  if (d->ngg != 2) return fail(OSTC_ERR_ARG, "ngg must be 2 (MUSCL/NND-2 stencils)");
  if (d->precision != OSTC_FP32 && d->precision != OSTC_FP64) return fail(OSTC_ERR_ARG, "precision 4 or 8");
  auto *z = new (std::nothrow) ostc_zone;
  if (!z) return fail(OSTC_ERR_ALLOC, "host allocation");
  z->ctx = ctx;
  z->desc = *d;
  const int64_t nx = d->mx + 2 * d->ngg, ny = d->my + 2 * d->ngg, nz = d->mz + 2 * d->ngg;
  z->plane = nx * ny * nz;
  z->esize = d->precision;
  const ostc_param &p = ctx->host;
  const int ns = p.n_spec, nv = p.n_var, nsc = p.n_scalar;
  const bool rans = p.turb_method == OSTC_RANS;
  const bool fr = p.mix_model == OSTC_FR;
  const bool dplur = p.implicit_method == 1;
  int pl[OSTC_F_COUNT] = {};
  pl[OSTC_F_X] = pl[OSTC_F_Y] = pl[OSTC_F_Z] = 1;
  pl[OSTC_F_JAC] = 1;
  pl[OSTC_F_METRIC] = 9;
  pl[OSTC_F_WALL_DIST] = rans ? 1 : 0;
  pl[OSTC_F_BV] = 6;
  pl[OSTC_F_SV] = nsc;
  pl[OSTC_F_CV] = nv;
  pl[OSTC_F_BV_LAST] = 4;
  pl[OSTC_F_VEL] = pl[OSTC_F_SOUND] = pl[OSTC_F_MACH] = pl[OSTC_F_MUL] = pl[OSTC_F_LAMBDA] = 1;
  pl[OSTC_F_RHO_D] = ns;
  pl[OSTC_F_GAMMA] = pl[OSTC_F_CP] = ns > 0 ? 1 : 0;
  pl[OSTC_F_CHEM_JAC] = fr ? (p.chem_src_method == 1 ? ns * ns : (p.chem_src_method == 2 ? ns : 0)) : 0;
  pl[OSTC_F_MUT] = pl[OSTC_F_LAMBDA_T] = rans ? 1 : 0;
  pl[OSTC_F_TURB_JAC] = rans ? 2 : 0;
  pl[OSTC_F_DQ] = nv;
  pl[OSTC_F_DQ0] = pl[OSTC_F_DQK] = dplur ? nv : 0;
  pl[OSTC_F_SPECTR] = 3;
  pl[OSTC_F_DT] = 1;
  for (int f = 0; f < OSTC_F_COUNT; ++f) {
    z->planes[f] = pl[f];
    z->fsize[f] = (f == OSTC_F_X || f == OSTC_F_Y || f == OSTC_F_Z || f == OSTC_F_CHEM_JAC) ? 8 : z->esize;
    if (pl[f] == 0) continue;
    size_t bytes = (size_t)z->plane * pl[f] * z->fsize[f];
    // the LU factors share these planes and may be stored in tiles of 64
    // cells, which reach up to 63 cells' entries beyond them
    if (f == OSTC_F_CHEM_JAC) bytes += (size_t)64 * pl[f] * z->fsize[f];
    hipError_t e = hipMalloc(&z->ptr[f], bytes);
    if (e == hipSuccess) e = hipMemset(z->ptr[f], 0, bytes);
    if (e != hipSuccess) {
      ostc_zone_destroy(z);
      return hip_fail(e, "hipMalloc(zone field)");
    }
  }
  hipError_t e = hipSuccess;
  if (fr && p.chem_src_method == 1 && dplur) {
    e = hipMalloc(&z->piv, (size_t)z->plane * ns * sizeof(int));
    if (e == hipSuccess) e = hipMemset(z->piv, 0, (size_t)z->plane * ns * sizeof(int));
  }
  if (e == hipSuccess) e = hipMalloc(&z->counters, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(z->counters, 0, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&z->d_inflow, OSTC_MAX_BCDATA * sizeof(ostc_inflow));
  {
    // limit_flow repair list, sized for every interior cell
    const size_t ncell = (size_t)d->mx * d->my * d->mz;
    const size_t nval = 6 + nsc + nv;
    // three list counters: slots 0/1 alternate between time steps, slot 2 is
    // the stand-alone limit_flow's
    if (e == hipSuccess) e = hipMalloc(&z->lf_count, 3 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(z->lf_count, 0, 3 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&z->lf_idx, ncell * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&z->lf_vals, ncell * nval * z->esize);
  }
  if (e != hipSuccess) {
    ostc_zone_destroy(z);
    return hip_fail(e, "zone setup");
  }
  z->vf = make_view<float>(z);
  z->vd = make_view<double>(z);
  *out = z;
  return OSTC_OK;
}

int ostc_zone_destroy(ostc_zone *z) {
  if (!z) return OSTC_OK;
  for (int f = 0; f < OSTC_F_COUNT; ++f)
    if (z->ptr[f]) (void)hipFree(z->ptr[f]);
  if (z->piv) (void)hipFree(z->piv);
  if (z->counters) (void)hipFree(z->counters);
  if (z->d_inflow) (void)hipFree(z->d_inflow);
  if (z->lf_count) (void)hipFree(z->lf_count);
  if (z->lf_idx) (void)hipFree(z->lf_idx);
  if (z->lf_vals) (void)hipFree(z->lf_vals);
  delete z;
  return OSTC_OK;
}

long ostc_zone_plane(const ostc_zone *z) { return z ? (long)z->plane : 0; }

// Planes of field f as the caller sees them: with a padded mechanism the
// species-indexed fields hold the caller's species count.
static int user_planes(const ostc_zone *z, int f) {
  const int pad = z->ctx->pad, n = z->planes[f];
  if (!pad || n == 0) return n;
  const int b = z->ctx->host.n_spec, u = z->ctx->user_ns;
  switch (f) {
    case OSTC_F_SV:
    case OSTC_F_CV:
    case OSTC_F_DQ:
    case OSTC_F_DQ0:
    case OSTC_F_DQK:
    case OSTC_F_RHO_D:
      return n - pad;
    case OSTC_F_CHEM_JAC:
      return n == b * b ? u * u : u;
    default:
      return n;
  }
}

// The device plane of the caller's plane l of field f: species planes keep
// their index, the planes after them (turbulence) move past the inert ones,
// and the EPI Jacobian's (m, n) entry moves from m*ns + n to m*9 + n.
static int device_plane(const ostc_zone *z, int f, int l) {
  const int pad = z->ctx->pad;
  if (!pad) return l;
  const int b = z->ctx->host.n_spec, u = z->ctx->user_ns;
  switch (f) {
    case OSTC_F_SV:
      return l < u ? l : l + pad;
    case OSTC_F_CV:
    case OSTC_F_DQ:
    case OSTC_F_DQ0:
    case OSTC_F_DQK:
      return l < 5 + u ? l : l + pad;
    case OSTC_F_CHEM_JAC:
      return z->planes[f] == b * b ? (l / u) * b + (l % u) : l;
    default:
      return l;
  }
}

int ostc_zone_field_planes(const ostc_zone *z, int f, int *n) {
  if (!z || f < 0 || f >= OSTC_F_COUNT || !n) return fail(OSTC_ERR_ARG, "bad field");
  *n = user_planes(z, f);
  return OSTC_OK;
}

int ostc_zone_field_ptr(const ostc_zone *z, int f, void **dev, int *es) {
  if (!z || f < 0 || f >= OSTC_F_COUNT) return fail(OSTC_ERR_ARG, "bad field");
  if (dev) *dev = z->ptr[f];
  if (es) *es = z->fsize[f];
  return OSTC_OK;
}

int ostc_zone_upload(ostc_zone *z, int f, const double *host, int n_planes) {
  if (!z || f < 0 || f >= OSTC_F_COUNT || !host) return fail(OSTC_ERR_ARG, "bad field");
  if (n_planes > user_planes(z, f) || n_planes < 0) return fail(OSTC_ERR_ARG, "too many planes for this field");
  const size_t pn = (size_t)z->plane, n = pn * n_planes;
  if (n == 0) return OSTC_OK;
  std::vector<float> tmp(z->fsize[f] == 8 ? 0 : pn);
  for (int l = 0; l < n_planes; ++l) {
    char *dst = static_cast<char *>(z->ptr[f]) + (size_t)device_plane(z, f, l) * pn * z->fsize[f];
    const double *src = host + (size_t)l * pn;
    if (z->fsize[f] == 8) {
      HIP_TRY(hipMemcpy(dst, src, pn * 8, hipMemcpyHostToDevice));
    } else {
      for (size_t q = 0; q < pn; ++q) tmp[q] = (float)src[q];
      HIP_TRY(hipMemcpy(dst, tmp.data(), pn * 4, hipMemcpyHostToDevice));
    }
  }
  return OSTC_OK;
}

int ostc_zone_download(const ostc_zone *z, int f, double *host, int n_planes) {
  if (!z || f < 0 || f >= OSTC_F_COUNT || !host) return fail(OSTC_ERR_ARG, "bad field");
  if (n_planes > user_planes(z, f) || n_planes < 0) return fail(OSTC_ERR_ARG, "too many planes for this field");
  const size_t pn = (size_t)z->plane, n = pn * n_planes;
  if (n == 0) return OSTC_OK;
  HIP_TRY(hipDeviceSynchronize());
  std::vector<float> tmp(z->fsize[f] == 8 ? 0 : pn);
  for (int l = 0; l < n_planes; ++l) {
    const char *src = static_cast<const char *>(z->ptr[f]) + (size_t)device_plane(z, f, l) * pn * z->fsize[f];
    double *dst = host + (size_t)l * pn;
    if (z->fsize[f] == 8) {
      HIP_TRY(hipMemcpy(dst, src, pn * 8, hipMemcpyDeviceToHost));
    } else {
      HIP_TRY(hipMemcpy(tmp.data(), src, pn * 4, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < pn; ++q) dst[q] = tmp[q];
    }
  }
  return OSTC_OK;
}

int ostc_zone_set_bc(ostc_zone *z, const ostc_bc *bc) {
  if (!z || !bc) return fail(OSTC_ERR_ARG, "null argument");
  if (bc->n_face < 0 || bc->n_face > OSTC_MAX_FACE) return fail(OSTC_ERR_ARG, "n_face out of range");
  const int g = z->desc.ngg;
  const int lo[3] = {-g, -g, -g}, hi[3] = {z->desc.mx - 1 + g, z->desc.my - 1 + g, z->desc.mz - 1 + g};
  for (int f = 0; f < bc->n_face; ++f) {
    const ostc_face &b = bc->face[f];
    if (b.face < 0 || b.face > 2 || (b.direction != 1 && b.direction != -1))
      return fail(OSTC_ERR_ARG, "bad face orientation");
    for (int q = 0; q < 3; ++q)
      if (b.range_start[q] < lo[q] || b.range_end[q] > hi[q] || b.range_start[q] > b.range_end[q])
        return fail(OSTC_ERR_ARG, "face range outside the padded block");
    const int ext[3] = {z->desc.mx, z->desc.my, z->desc.mz};
    const int at = b.range_start[b.face];
    if (b.range_end[b.face] != at || (b.direction < 0 ? at != 0 : at != ext[b.face] - 1))
      return fail(OSTC_ERR_ARG, "face must lie on the block boundary");
    if ((b.kind == 5 || b.kind == 2) && (b.slot < 0 || b.slot >= OSTC_MAX_BCDATA))
      return fail(OSTC_ERR_ARG, "bad BC slot");
    if (b.kind != 2 && b.kind != 3 && b.kind != 5 && b.kind != 6) return fail(OSTC_ERR_ARG, "unknown BC kind");
  }
  z->bc = *bc;
  if (z->ctx->pad) {
    // inflow scalars [Y_0..Y_ns-1, k, omega] -> the padded layout
    const int u = z->ctx->user_ns, b = z->ctx->host.n_spec, nt = z->ctx->host.n_scalar - b;
    for (auto &in : z->bc.inflow) {
      double sv[OSTC_MAX_SPEC + 2] = {};
      for (int l = 0; l < u; ++l) sv[l] = in.sv[l];
      for (int l = 0; l < nt; ++l) sv[b + l] = in.sv[u + l];
      std::memcpy(in.sv, sv, sizeof(sv));
    }
  }
  z->has_bc = true;
  HIP_TRY(hipMemcpy(z->d_inflow, z->bc.inflow, sizeof(z->bc.inflow), hipMemcpyHostToDevice));
  return OSTC_OK;
}

#define ZONE_CALL(z, EXPR_F, EXPR_D) (fp64(z) ? (EXPR_D) : (EXPR_F))

int ostc_compute_cv_from_bv(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, cv_from_bv(z->vf, c->dev, c->opts, S(s)), cv_from_bv(z->vd, c->dev, c->opts, S(s)));
}

int ostc_initialize_mut(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, initialize_mut(z->vf, c->dev, c->opts, S(s)), initialize_mut(z->vd, c->dev, c->opts, S(s)));
}

int ostc_wall_distance(ostc_ctx *c, ostc_zone *z, const double *wall_xyz, int n, void *s) {
  if (!c || !z || (n > 0 && !wall_xyz) || n < 0) return fail(OSTC_ERR_ARG, "bad wall points");
  if (!z->ptr[OSTC_F_WALL_DIST]) return fail(OSTC_ERR_STATE, "zone has no wall distance (laminar)");
  double *d = nullptr;
  if (n > 0) {
    HIP_TRY(hipMalloc(&d, (size_t)n * 3 * sizeof(double)));
    HIP_TRY(hipMemcpy(d, wall_xyz, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
  }
  int r = ZONE_CALL(z, wall_distance(z->vf, d, n, S(s)), wall_distance(z->vd, d, n, S(s)));
  HIP_TRY(hipStreamSynchronize(S(s)));
  if (d) (void)hipFree(d);
  return r;
}

int ostc_wall_friction_heatflux(ostc_zone *z, double dyn_pressure, double *cf, double *qw, void *s) {
  if (!z || !cf || !qw) return fail(OSTC_ERR_ARG, "null argument");
  if (!(dyn_pressure > 0)) return fail(OSTC_ERR_ARG, "dyn_pressure must be positive");
  return ZONE_CALL(z, wall_friction_heatflux(z->vf, dyn_pressure, cf, qw, S(s)),
                   wall_friction_heatflux(z->vd, dyn_pressure, cf, qw, S(s)));
}

int ostc_compute_source(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  bool wrote = false;
  return ZONE_CALL(z, source_terms(z->vf, c->dev, c->opts, S(s), true, &wrote),
                   source_terms(z->vd, c->dev, c->opts, S(s), true, &wrote));
}

int ostc_inviscid_flux(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  const int ns = c->opts.ns, nt = c->opts.nt;
  return ZONE_CALL(z, flux_sweeps(z->vf, c->dev, ns, nt, true, false, false, S(s)),
                   flux_sweeps(z->vd, c->dev, ns, nt, true, false, false, S(s)));
}

int ostc_viscous_flux(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  if (c->host.viscous_scheme == 0) return OSTC_OK;
  const int ns = c->opts.ns, nt = c->opts.nt;
  return ZONE_CALL(z, flux_sweeps(z->vf, c->dev, ns, nt, false, true, false, S(s)),
                   flux_sweeps(z->vd, c->dev, ns, nt, false, true, false, S(s)));
}

int ostc_local_time_step(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, local_time_step(z->vf, c->dev, c->opts, S(s)), local_time_step(z->vd, c->dev, c->opts, S(s)));
}


int ostc_implicit_treatment(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  // leaves the final increment in dq, as the reference does after its copies
  return ZONE_CALL(z, implicit_and_update(c, z->vf, S(s), false), implicit_and_update(c, z->vd, S(s), false));
}

int ostc_update_cv_bv(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, update_cv_bv(z->vf, c->dev, c->opts, (const float *)z->vf.dq, S(s)),
                   update_cv_bv(z->vd, c->dev, c->opts, (const double *)z->vd.dq, S(s)));
}

int ostc_limit_flow(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, limit_flow(z->vf, c->dev, c->opts, S(s)), limit_flow(z->vd, c->dev, c->opts, S(s)));
}

int ostc_apply_bc(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  if (!z->has_bc) return OSTC_OK;
  return ZONE_CALL(z, apply_bc(z->vf, c->dev, c->opts, z->bc, z->d_inflow, S(s)),
                   apply_bc(z->vd, c->dev, c->opts, z->bc, z->d_inflow, S(s)));
}

int ostc_inner_halo(ostc_ctx *c, ostc_zone **zones, int blk, const ostc_inner_face *f, void *s) {
  if (!c || !zones || !f || blk < 0) return fail(OSTC_ERR_ARG, "null argument");
  ostc_zone *z = zones[blk], *t = zones[f->target_block];
  if (!z || !t || z->desc.precision != t->desc.precision) return fail(OSTC_ERR_ARG, "bad inner face blocks");
  return ZONE_CALL(z, inner_halo(z->vf, t->vf, c->dev, c->opts, *f, S(s)),
                   inner_halo(z->vd, t->vd, c->dev, c->opts, *f, S(s)));
}

int ostc_update_physical_properties(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, physical_properties(z->vf, c->dev, c->opts, S(s)),
                   physical_properties(z->vd, c->dev, c->opts, S(s)));
}

int ostc_eliminate_k_gradient(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, eliminate_k_gradient(z->vf, c->opts, S(s)), eliminate_k_gradient(z->vd, c->opts, S(s)));
}

int ostc_store_last_step(ostc_ctx *c, ostc_zone *z, void *s) {
  if (!c || !z) return fail(OSTC_ERR_ARG, "null argument");
  return ZONE_CALL(z, store_last_step(z->vf, S(s)), store_last_step(z->vd, S(s)));
}

int ostc_residual_sums(ostc_ctx *c, ostc_zone **zones, int n, double sums[4], void *s) {
  if (!c || !zones || n <= 0 || !sums) return fail(OSTC_ERR_ARG, "null argument");
  size_t need = 0;
  for (int b = 0; b < n; ++b) {
    const auto g = cell_grid(zones[b]->desc.mx, zones[b]->desc.my, zones[b]->desc.mz);
    need += (size_t)g.grid.x * g.grid.y * g.grid.z * 4;
  }
  if (need > c->partial_cap) {
    HIP_TRY(hipStreamSynchronize(S(s)));
    if (c->partial) (void)hipFree(c->partial);
    c->partial = nullptr;
    HIP_TRY(hipMalloc(&c->partial, need * sizeof(double)));
    c->partial_cap = need;
  }
  int off = 0;
  for (int b = 0; b < n; ++b) {
    int nb = 0;
    ostc_zone *z = zones[b];
    int r = ZONE_CALL(z, residual_partials(z->vf, c->partial + (size_t)off * 4, S(s), &nb),
                      residual_partials(z->vd, c->partial + (size_t)off * 4, S(s), &nb));
    if (r) return r;
    off += nb;
  }
  if (int r = residual_finalize(c->partial, off, c->out4, S(s))) return r;
  HIP_TRY(hipMemcpyAsync(c->host4, c->out4, 4 * sizeof(double), hipMemcpyDeviceToHost, S(s)));
  HIP_TRY(hipStreamSynchronize(S(s)));
  for (int l = 0; l < 4; ++l) sums[l] = c->host4[l];
  return OSTC_OK;
}

int ostc_residual_sums_noalloc(ostc_ctx *c, ostc_zone **zones, int n, double *d_sums, void *s) {
  if (!c || !zones || n <= 0 || !d_sums) return fail(OSTC_ERR_ARG, "null argument");
  long need = 0;
  for (int b = 0; b < n; ++b) {
    if (!zones[b]) return fail(OSTC_ERR_ARG, "null zone");
    const auto g = cell_grid(zones[b]->desc.mx, zones[b]->desc.my, zones[b]->desc.mz);
    need += (long)g.grid.x * g.grid.y * g.grid.z * 4;
    if (need > kScratchDoubles)
      return fail(OSTC_ERR_ARG, "residual partial sums exceed the 8 MiB scratch array (2^20 doubles: "
                                "4 per 64x4 cell tile of every zone)");
  }
  double *ptr = nullptr;
  if (int r = scratch_setup(&ptr)) return r;
  int off = 0;
  for (int b = 0; b < n; ++b) {
    int nb = 0;
    ostc_zone *z = zones[b];
    int r = ZONE_CALL(z, residual_partials(z->vf, ptr + (size_t)off * 4, S(s), &nb),
                      residual_partials(z->vd, ptr + (size_t)off * 4, S(s), &nb));
    if (r) return r;
    off += nb;
  }
  return residual_finalize(ptr, off, d_sums, S(s));
}

int ostc_zone_counters(ostc_zone *z, long long out[4], int reset, void *s) {
  if (!z || !out) return fail(OSTC_ERR_ARG, "null argument");
  unsigned long long h[4];
  HIP_TRY(hipStreamSynchronize(S(s)));
  HIP_TRY(hipMemcpy(h, z->counters, sizeof(h), hipMemcpyDeviceToHost));
  for (int l = 0; l < 4; ++l) out[l] = (long long)h[l];
  if (reset) HIP_TRY(hipMemset(z->counters, 0, sizeof(h)));
  return OSTC_OK;
}

// Zone streams for ostc_steady_step (created on first use, n of them)
static int zone_streams(ostc_ctx *c, int n) {
  if (!c->fork) HIP_TRY(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  while ((int)c->zstream.size() < n) {
    hipStream_t q = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    c->zstream.push_back(q);
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->zdone.push_back(e);
    hipEvent_t m = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&m, hipEventDisableTiming));
    c->zmid.push_back(m);
  }
  return OSTC_OK;
}

// Everything of one iteration that touches only zone z (src/Driver.cu:
// 224-257): source, fluxes, local dt, implicit treatment, update,
// limit_flow, boundary conditions.
static int zone_iteration(ostc_ctx *c, ostc_zone *z, int step, int output_screen, hipStream_t st,
                          hipEvent_t fluxes_done = nullptr) {
  const CellOpts &o = c->opts;
  const bool visc = c->host.viscous_scheme != 0;
  void *s = (void *)st;
  int r = OSTC_OK;
  if (step % output_screen == 0) {
    prof_mark(P_STORE, st, true);
    r = ostc_store_last_step(c, z, s);
    prof_mark(P_STORE, st, false);
  }
  if (r) return r;
  bool wrote = false;
  // SST zones without chemistry: the source kernel also forms the local
  // time step (it holds the metric, density, viscosity and mut); EPI zones
  // under DPLUR (9-species kernels): so does the SST-only source kernel, and
  // the chemistry source follows the fluxes, fused with dq0 and the
  // factorisation of the chemistry block (no Jacobian stored)
  const bool chem_late = chem_after_fluxes(o);
  const bool dt_in_source = source_has_dt(o) || chem_late;
  prof_mark(P_SOURCE, st, true);
  r = ZONE_CALL(z, source_terms(z->vf, c->dev, o, st, false, &wrote, dt_in_source),
                source_terms(z->vd, c->dev, o, st, false, &wrote, dt_in_source));
  prof_mark(P_SOURCE, st, false);
  if (r) return r;
  // a source kernel that forms the time step writes only the turbulence
  // rows of dq (none without SST; the chemistry, if any, follows the
  // fluxes): the first flux direction writes the others
  const bool flux_first = !wrote || dt_in_source, keep_turb = dt_in_source && o.nt > 0;
  r = ZONE_CALL(z, flux_sweeps(z->vf, c->dev, o.ns, o.nt, true, visc, flux_first, st, keep_turb),
                flux_sweeps(z->vd, c->dev, o.ns, o.nt, true, visc, flux_first, st, keep_turb));
  if (r) return r;
  if (!dt_in_source) {
    prof_mark(P_DT, st, true);
    r = ZONE_CALL(z, local_time_step(z->vf, c->dev, o, st), local_time_step(z->vd, c->dev, o, st));
    prof_mark(P_DT, st, false);
    if (r) return r;
  }
  if (fluxes_done) HIP_TRY(hipEventRecord(fluxes_done, st));
  if (o.implicit_method == 0) prof_mark(P_EXPL_IMPL, st, true);
  {
    // implicit treatment, then the update with the T-from-e solve
    const bool dplur = o.implicit_method == 1;
    if (fp64(z)) {
      double *res = nullptr;
      r = implicit_treatment(z->vd, c->dev, o, st, &res, chem_late);
      if (!dplur) prof_mark(P_EXPL_IMPL, st, false);
      if (r) return r;
      prof_mark(P_UPDATE, st, true);
      r = update_cv_bv(z->vd, c->dev, o, res, st, z->lf_slot);
    } else {
      float *res = nullptr;
      r = implicit_treatment(z->vf, c->dev, o, st, &res, chem_late);
      if (!dplur) prof_mark(P_EXPL_IMPL, st, false);
      if (r) return r;
      prof_mark(P_UPDATE, st, true);
      r = update_cv_bv(z->vf, c->dev, o, res, st, z->lf_slot);
    }
    prof_mark(P_UPDATE, st, false);
  }
  if (r) return r;
  prof_mark(P_LIMIT, st, true);
  r = ZONE_CALL(z, limit_flow(z->vf, c->dev, o, st, z->lf_slot), limit_flow(z->vd, c->dev, o, st, z->lf_slot));
  prof_mark(P_LIMIT, st, false);
  if (r) return r;
  // the update cleared the other slot for the next step's list
  z->lf_slot ^= 1;
  prof_mark(P_BC, st, true);
  r = ostc_apply_bc(c, z, s);
  prof_mark(P_BC, st, false);
  return r;
}

int ostc_steady_step(ostc_ctx *c, ostc_zone **zones, int n, const ostc_inner_face *inner, const int *n_inner,
                     int step, int output_screen, double res_sums[4], int *res_computed, void *s) {
  if (!c || !zones || n <= 0 || output_screen <= 0) return fail(OSTC_ERR_ARG, "bad arguments");
  const hipStream_t st = S(s);
  // Several zones: each zone's chain on its own stream, forked from the
  // caller's stream and joined back before the inner-face exchange (the
  // zones share nothing until then).  With the stage profiler on, the zones
  // run one after another on the caller's stream, so that the stage events
  // bracket one kernel at a time.
  if (n > 1 && !g_prof) {
    if (int r = zone_streams(c, n)) return r;
    HIP_TRY(hipEventRecord(c->fork, st));
    for (int b = 0; b < n; ++b) {
      HIP_TRY(hipStreamWaitEvent(c->zstream[b], c->fork, 0));
      // zones in pairs, the second starting when the first's fluxes are
      // done: its source kernel (LDS-bound, vector registers to spare) then
      // meets the first zone's HBM-bound block solves rather than another
      // source kernel
      if (b & 1) HIP_TRY(hipStreamWaitEvent(c->zstream[b], c->zmid[b - 1], 0));
      if (int r = zone_iteration(c, zones[b], step, output_screen, c->zstream[b], c->zmid[b])) return r;
      HIP_TRY(hipEventRecord(c->zdone[b], c->zstream[b]));
    }
    for (int b = 0; b < n; ++b) HIP_TRY(hipStreamWaitEvent(st, c->zdone[b], 0));
  } else {
    for (int b = 0; b < n; ++b)
      if (int r = zone_iteration(c, zones[b], step, output_screen, st)) return r;
  }
  if (inner && n_inner) {
    prof_mark(P_HALO, st, true);
    int off = 0;
    for (int b = 0; b < n; ++b) {
      for (int f = 0; f < n_inner[b]; ++f)
        if (int r = ostc_inner_halo(c, zones, b, &inner[off + f], s)) return r;
      off += n_inner[b];
    }
    prof_mark(P_HALO, st, false);
  }
  // 2-D blocks: k ghost layers from the k = 0 plane (src/Driver.cu:259-265)
  for (int b = 0; b < n; ++b)
    if (zones[b]->desc.mz == 1)
      if (int r = ostc_eliminate_k_gradient(c, zones[b], s)) return r;
  prof_mark(P_PROPS, st, true);
  for (int b = 0; b < n; ++b)
    if (int r = ostc_update_physical_properties(c, zones[b], s)) return r;
  prof_mark(P_PROPS, st, false);
  const bool do_res = (step % output_screen == 0) || step == 1;
  if (res_computed) *res_computed = do_res ? 1 : 0;
  if (do_res && res_sums) {
    prof_mark(P_RESID, st, true);
    int r = ostc_residual_sums(c, zones, n, res_sums, s);
    prof_mark(P_RESID, st, false);
    return r;
  }
  return OSTC_OK;
}

static const char *kStageNames[P_COUNT] = {"source", "flux_i", "flux_j", "flux_k", "local_dt", "dplur_dq0",
                                           "dplur_sweep", "explicit_implicit", "update", "limit_flow", "bc",
                                           "inner_halo", "properties", "store_last", "residual", "flux_ij",
                                           "dplur_sweep_first", "dplur_chem"};

int ostc_profile(ostc_ctx *c, int enable) {
  (void)c;
  if (enable) {
    if (!g_prof) g_prof = new Profiler;
    g_prof->recs.clear();
    g_prof->next = 0;
  } else if (g_prof) {
    delete g_prof;
    g_prof = nullptr;
  }
  return OSTC_OK;
}

int ostc_profile_read(ostc_ctx *c, double *ms, long long *count, int n, int reset) {
  (void)c;
  if (!ms || !count || n <= 0) return fail(OSTC_ERR_ARG, "bad arguments");
  for (int q = 0; q < n; ++q) {
    ms[q] = 0;
    count[q] = 0;
  }
  if (!g_prof) return OSTC_OK;
  for (auto &r : g_prof->recs) {
    HIP_TRY(hipEventSynchronize(r.b));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
    if (r.stage < n) {
      ms[r.stage] += t;
      count[r.stage] += 1;
    }
  }
  if (reset) {
    g_prof->recs.clear();
    g_prof->next = 0;
  }
  return OSTC_OK;
}

const char *ostc_profile_stage_name(int stage) {
  return (stage >= 0 && stage < P_COUNT) ? kStageNames[stage] : "";
}

int ostc_profile_stage_count(void) { return P_COUNT; }

int ostc_chem_plain_step(ostc_ctx *c, long n, const double *t, const double *p, const double *y, const double *dt,
                         double *dy, void *s) {
  if (!c || n < 0 || (n > 0 && (!t || !p || !y || !dt || !dy))) return fail(OSTC_ERR_ARG, "bad arguments");
  if (c->user_ns != 9 || c->host.n_reac <= 0) return fail(OSTC_ERR_ARG, "needs the 9-species mechanism");
  return chem_plain_step(c->dev, c->host.n_spec, c->host.n_reac, n, t, p, y, dt, dy, S(s));
}

int ostc_chem_source_batch(ostc_ctx *c, long n, const double *rho, const double *t, const double *y, double *omega,
                           double *jac, void *s) {
  if (!c || n < 0 || (n > 0 && (!rho || !t || !y || !omega))) return fail(OSTC_ERR_ARG, "bad arguments");
  if (c->user_ns != 9) return fail(OSTC_ERR_ARG, "needs the 9-species mechanism");
  return chem_source_batch(c->dev, c->host.n_spec, c->host.n_reac, n, rho, t, y, omega, jac, S(s));
}

// ---- learned chemistry surrogate ------------------------------------------
}  // extern "C"

struct ostc_surrogate {
  ostc_ctx *ctx = nullptr;
  ostc::SurState st;
  std::vector<void *> owned;
};

namespace {

template <typename T>
hipError_t sur_alloc(ostc_surrogate *s, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    s->owned.push_back(q);
    *p = static_cast<T *>(q);
  }
  return e;
}

int sur_check_desc(const ostc_surrogate_desc &d) {
  if (d.kind == 0 && d.n_out != 10) return fail(OSTC_ERR_ARG, "surrogate kind 0 has 10 outputs");
  if (d.kind == 1 && d.n_out != 9) return fail(OSTC_ERR_ARG, "surrogate kind 1 has 9 outputs");
  if (d.kind < 0 || d.kind > 1) return fail(OSTC_ERR_ARG, "surrogate kind must be 0 or 1");
  if (d.n_elem < 0 || d.n_elem > OSTC_SUR_MAX_ELEM) return fail(OSTC_ERR_ARG, "n_elem out of range");
  if (d.activation != OSTC_SUR_GELU && d.activation != OSTC_SUR_HSWISH)
    return fail(OSTC_ERR_ARG, "surrogate activation must be 0 (GELU) or 1 (hard-swish)");
  if (!(d.dy_max > 0)) return fail(OSTC_ERR_ARG, "dy_max must be positive");
  for (int f = 0; f < OSTC_SUR_IN; ++f)
    if (!(d.in_std[f] > 0)) return fail(OSTC_ERR_ARG, "in_std must be positive");
  return OSTC_OK;
}

ostc::SurParams sur_params(const ostc_surrogate_desc &d) {
  ostc::SurParams p{};
  p.n_out = d.n_out;
  p.kind = d.kind;
  p.n_elem = d.n_elem;
  p.act = d.activation;
  for (int f = 0; f < OSTC_SUR_IN; ++f) {
    p.in_mean[f] = d.in_mean[f];
    p.in_rstd[f] = (float)(1.0 / (double)d.in_std[f]);
  }
  for (int o = 0; o < 16; ++o) {
    p.out_mean[o] = d.out_mean[o];
    p.out_std[o] = d.out_std[o];
  }
  for (int q = 0; q < OSTC_SUR_MAX_ELEM * 9; ++q) p.elem_w[q] = d.elem_w[q];
  p.floor = d.floor;
  p.t_min = d.t_min, p.t_max = d.t_max, p.p_min = d.p_min, p.p_max = d.p_max;
  p.dt_min = d.dt_min, p.dt_max = d.dt_max;
  p.y_tol = d.y_tol, p.elem_tol = d.elem_tol, p.elem_abs = d.elem_abs;
  p.dy_max = d.dy_max;
  return p;
}

int sur_ensure_fallback(ostc_surrogate *s, long n) {
  if (n <= s->st.fb_cap) return OSTC_OK;
  HIP_TRY(hipDeviceSynchronize());
  for (void *buf : {(void *)s->st.fb_list, (void *)s->st.out12}) {
    if (!buf) continue;
    for (auto &q : s->owned)
      if (q == buf) q = nullptr;
    HIP_TRY(hipFree(buf));
  }
  s->st.fb_list = nullptr;
  s->st.out12 = nullptr;
  s->st.fb_cap = 0;
  HIP_TRY(sur_alloc(s, &s->st.fb_list, (size_t)ostc::kFbLists * ostc::fb_list_cap(n)));
  HIP_TRY(sur_alloc(s, &s->st.out12, (size_t)n * 12));
  s->st.fb_cap = n;
  return OSTC_OK;
}

int sur_read_stats(ostc_surrogate *s, long long stats[4], hipStream_t st) {
  if (!stats) return OSTC_OK;
  unsigned long long h[ostc::kFbLists * ostc::kStatStride];
  HIP_TRY(hipMemcpyAsync(h, s->st.stats, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int q = 0; q < 4; ++q) {
    unsigned long long v = 0;
    for (int l = 0; l < ostc::kFbLists; ++l) v += h[l * ostc::kStatStride + q];
    stats[q] = (long long)v;
  }
  return OSTC_OK;
}

}  // namespace

extern "C" {

int ostc_surrogate_n_params(int n_out) { return (n_out > 0 && n_out <= 16) ? ostc::sur_n_params(n_out) : 0; }

int ostc_surrogate_create(ostc_ctx *ctx, const ostc_surrogate_desc *desc, const float *weights, long max_batch,
                          ostc_surrogate **out) {
  if (!ctx || !desc || !weights || !out || max_batch < 0 || max_batch % 16)
    return fail(OSTC_ERR_ARG, "bad arguments");
  if (ctx->user_ns != 9 || ctx->host.n_reac <= 0) return fail(OSTC_ERR_ARG, "needs the 9-species mechanism");
  int st = sur_check_desc(*desc);
  if (st) return st;
  auto *s = new (std::nothrow) ostc_surrogate;
  if (!s) return fail(OSTC_ERR_ALLOC, "host allocation");
  s->ctx = ctx;
  ostc::SurState &S = s->st;
  S.params = sur_params(*desc);
  S.n_params = ostc::sur_n_params(desc->n_out);
  S.max_batch = max_batch;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&S.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || S.cus <= 0)
    S.cus = 256;
  const size_t np = (size_t)S.n_params;
  hipError_t e = hipSuccess;
  auto A = [&](auto **p, size_t n) {
    if (e == hipSuccess) e = sur_alloc(s, p, n);
  };
  A(&S.master, np);
  A(&S.grad, np);
  A(&S.m1, np);
  A(&S.m2, np);
  A(&S.wq, np);
  char *pk = nullptr;
  A(&pk, (size_t)ostc::kSurTiles * 64 * 16);
  S.packed = pk;
  A(&S.bias, (size_t)ostc::kSurBias);
  A(&S.loss_w, 16);
  A(&S.fb_count, (size_t)ostc::kFbLists * ostc::kFbStride);
  A(&S.stats, (size_t)ostc::kFbLists * ostc::kStatStride);
  if (max_batch > 0) {
    const size_t B = (size_t)max_batch;
    A(&S.a0, B * 16);
    A(&S.z, 4 * B * 128);
    A(&S.a, 4 * B * 144);
    A(&S.yout, B * 16);
    A(&S.g5, B * 16);
    A(&S.gbuf[0], B * 128);
    A(&S.gbuf[1], B * 128);
    A(&S.part, (size_t)ostc::kSurSplits * 128 * 144);
    A(&S.loss_part, (B * 16 + 255) / 256);
    A(&S.loss_dev, 1);
  }
  if (e == hipSuccess) e = hipMemcpy(S.master, weights, np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(S.m1, 0, np * sizeof(float));
  if (e == hipSuccess) e = hipMemset(S.m2, 0, np * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(S.loss_w, desc->loss_w, 16 * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    ostc_surrogate_destroy(s);
    return hip_fail(e, "surrogate setup");
  }
  st = ostc::sur_pack(S, nullptr);
  if (st == OSTC_OK && hipDeviceSynchronize() != hipSuccess) st = OSTC_ERR_HIP;
  if (st) {
    ostc_surrogate_destroy(s);
    return fail(st, "surrogate weight packing failed");
  }
  *out = s;
  return OSTC_OK;
}

int ostc_surrogate_destroy(ostc_surrogate *s) {
  if (!s) return OSTC_OK;
  (void)hipDeviceSynchronize();
  for (void *q : s->owned)
    if (q) (void)hipFree(q);
  delete s;
  return OSTC_OK;
}

int ostc_surrogate_get_weights(ostc_surrogate *s, float *w) {
  if (!s || !w) return fail(OSTC_ERR_ARG, "bad arguments");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(w, s->st.master, (size_t)s->st.n_params * sizeof(float), hipMemcpyDeviceToHost));
  return OSTC_OK;
}

int ostc_surrogate_set_weights(ostc_surrogate *s, const float *w) {
  if (!s || !w) return fail(OSTC_ERR_ARG, "bad arguments");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(s->st.master, w, (size_t)s->st.n_params * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(s->st.m1, 0, (size_t)s->st.n_params * sizeof(float)));
  HIP_TRY(hipMemset(s->st.m2, 0, (size_t)s->st.n_params * sizeof(float)));
  s->st.adam_step = 0;
  int st = ostc::sur_pack(s->st, nullptr);
  if (st) return fail(st, "surrogate weight packing failed");
  HIP_TRY(hipDeviceSynchronize());
  return OSTC_OK;
}

int ostc_surrogate_step(ostc_surrogate *s, long n, const double *t, const double *p, const double *y,
                        const double *dt, double *dy, long long stats[4], void *strm) {
  if (!s || n < 0 || (n > 0 && (!t || !p || !y || !dt || !dy))) return fail(OSTC_ERR_ARG, "bad arguments");
  int st = sur_ensure_fallback(s, n);
  if (st) return st;
  st = ostc::sur_infer_batch(s->st, s->ctx->dev, n, t, p, y, dt, dy, nullptr, S(strm));
  if (st) return fail(st, "surrogate inference launch failed");
  return sur_read_stats(s, stats, S(strm));
}

int ostc_surrogate_forward(ostc_surrogate *s, long n, const double *t, const double *p, const double *y,
                           const double *dt, float *raw, double *dy, void *strm) {
  if (!s || n < 0 || (n > 0 && (!t || !p || !y || !dt || !raw || !dy))) return fail(OSTC_ERR_ARG, "bad arguments");
  int st = sur_ensure_fallback(s, n);
  if (st) return st;
  st = ostc::sur_infer_batch(s->st, s->ctx->dev, n, t, p, y, dt, dy, raw, S(strm));
  return st ? fail(st, "surrogate inference launch failed") : OSTC_OK;
}

int ostc_surrogate_gradients(ostc_surrogate *s, long nb, const double *t, const double *p, const double *y,
                             const double *dt, const int *idx, const float *target, float *grad, double *loss,
                             void *strm) {
  if (!s || !t || !p || !y || !dt || !target || !grad) return fail(OSTC_ERR_ARG, "bad arguments");
  if (nb <= 0 || nb % 16 || nb > s->st.max_batch)
    return fail(OSTC_ERR_ARG, "batch must be a positive multiple of 16 within max_batch");
  int st = ostc::sur_gradients(s->st, nb, t, p, y, dt, idx, target, grad, s->st.loss_dev, S(strm));
  if (st) return fail(st, "surrogate gradient launch failed");
  if (loss) {
    HIP_TRY(hipMemcpyAsync(loss, s->st.loss_dev, sizeof(double), hipMemcpyDeviceToHost, S(strm)));
    HIP_TRY(hipStreamSynchronize(S(strm)));
  }
  return OSTC_OK;
}

int ostc_surrogate_train_step(ostc_surrogate *s, long nb, const double *t, const double *p, const double *y,
                              const double *dt, const int *idx, const float *target, float lr, double *loss,
                              void *strm) {
  if (!s) return fail(OSTC_ERR_ARG, "bad arguments");
  int st = ostc_surrogate_gradients(s, nb, t, p, y, dt, idx, target, s->st.grad, loss, strm);
  if (st) return st;
  st = ostc::sur_adam(s->st, lr, S(strm));
  return st ? fail(st, "surrogate update launch failed") : OSTC_OK;
}

int ostc_chem_zone_step(ostc_ctx *c, ostc_zone *z, ostc_surrogate *s, int mode, double *dy, long long stats[4],
                        void *strm) {
  if (!c || !z || !dy || (mode != 0 && mode != 1) || (mode == 1 && !s)) return fail(OSTC_ERR_ARG, "bad arguments");
  if (c->user_ns != 9 || c->host.n_reac <= 0) return fail(OSTC_ERR_ARG, "needs the 9-species mechanism");
  const long n = (long)z->desc.mx * z->desc.my * z->desc.mz;
  int st;
  if (mode == 0) {
    st = fp64(z) ? ostc::zone_plain_chem(c->dev, z->vd, dy, S(strm)) : ostc::zone_plain_chem(c->dev, z->vf, dy, S(strm));
    return st ? fail(st, "plain chemistry launch failed") : OSTC_OK;
  }
  st = sur_ensure_fallback(s, n);
  if (st) return st;
  st = fp64(z) ? ostc::sur_infer_zone(s->st, c->dev, z->vd, dy, S(strm))
               : ostc::sur_infer_zone(s->st, c->dev, z->vf, dy, S(strm));
  if (st) return fail(st, "surrogate inference launch failed");
  return sur_read_stats(s, stats, S(strm));
}

}  // extern "C"
