// Time-loop driver over the C-ABI: the reference's Driver<mix,turb>
// (src/Driver.cu:14-289) for a case read by ostc_case_read.
//  * ostc_driver_create: the constructor (src/Driver.cu:14-41) -- context
//    with the model constants, one zone per block with its geometry, initial
//    fields and boundary program;
//  * ostc_driver_initialize: initialize_computation (:43-104) with the wall
//    distance of acquire_wall_distance (:122-189);
//  * ostc_driver_step / ostc_driver_simulate: steady_simulation (:191-289),
//    compute_residual (:291-380) and steady_screen_output (:382-401);
//  * ostc_driver_write_flowfield: MPIIO::print_field (src/MPIIO.hpp), the
//    flow-field file simulate writes every output_file steps and at
//    convergence (src/Driver.cu:283-286).
// Model dispatch (src/main.cu:22-64) is the runtime model of ostc_param.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "ostc_case.h"

namespace {

thread_local std::string g_drv_err;

int drv_fail(int st, const std::string &msg) {
  g_drv_err = msg + (st == OSTC_ERR_ARG || st == OSTC_ERR_STATE || st == OSTC_ERR_IO ? std::string()
                                                                                      : std::string(": ") + ostc_last_error());
  return st;
}

#define DRV_TRY(call, what)                      \
  do {                                           \
    const int st_ = (call);                      \
    if (st_ != OSTC_OK) {                        \
      g_drv_err = std::string(what) + ": " + ostc_last_error(); \
      return st_;                                \
    }                                            \
  } while (0)

}  // namespace

struct ostc_driver {
  std::string dir;
  ostc_param param{};
  int total_step = 0, output_screen = 1, step = 0;
  double convergence = 0;
  long n_grid = 0;
  ostc_ctx *ctx = nullptr;
  std::vector<ostc_zone *> zones;
  std::vector<ostc_inner_face> inner;
  std::vector<int> n_inner;
  std::vector<double> wall_xyz;
  std::vector<std::string> spec_names;  // mechanism order
  std::vector<int> dims;                // mx, my, mz per block
  int ngg = 2, output_file = 1;
  double dyn_pressure = 0;               // 0.5 rho_inf v_inf^2
  std::vector<int> post_process;          // setup array post_process
  double res[4] = {0, 0, 0, 0}, res_scale[4] = {1, 1, 1, 1};
  bool initialized = false;
  std::chrono::steady_clock::time_point t0, t_last;

  ~ostc_driver() {
    for (auto *z : zones) ostc_zone_destroy(z);
    if (ctx) ostc_ctx_destroy(ctx);
  }
};

extern "C" {

int ostc_driver_create(const ostc_case *c, int precision, ostc_driver **out) {
  if (!c || !out) return drv_fail(OSTC_ERR_ARG, "null argument");
  if (precision != OSTC_FP32 && precision != OSTC_FP64) return drv_fail(OSTC_ERR_ARG, "precision must be 4 or 8");
  *out = nullptr;
  auto *d = new ostc_driver;
  d->dir = c->dir;
  d->param = c->param;
  d->total_step = c->total_step;
  d->output_screen = c->output_screen;
  d->convergence = c->convergence;
  d->wall_xyz = c->wall_xyz;
  d->output_file = c->output_file;
  d->dyn_pressure = 0.5 * c->rho_inf * c->v_inf * c->v_inf;
  d->post_process = c->post_process;
  // restart: continue the step count and keep the earlier run's residual
  // scales (src/Driver.cu:25-31,197-198)
  d->step = c->start_step;
  d->total_step = c->total_step + c->start_step;
  for (int l = 0; l < 4; ++l) d->res_scale[l] = c->res_scale[l];
  for (int k = 0; k < c->param.n_spec; ++k) d->spec_names.emplace_back(c->mech.spec_names[k]);
  auto cleanup = [&](int st) {
    delete d;
    return st;
  };
  if (int st = ostc_ctx_create(&d->param, &d->ctx)) {
    drv_fail(st, "ostc_ctx_create");
    return cleanup(st);
  }
  const int g = c->ngg;
  d->ngg = g;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const ostc_host::CaseBlock &blk = c->blocks[b];
    ostc_zone_desc desc{blk.mx, blk.my, blk.mz, g, precision, (int)b};
    ostc_zone *z = nullptr;
    if (int st = ostc_zone_create(d->ctx, &desc, &z)) {
      drv_fail(st, "ostc_zone_create");
      return cleanup(st);
    }
    d->zones.push_back(z);
    struct Up {
      int field;
      const std::vector<double> *arr;
      int planes;
    } ups[] = {{OSTC_F_X, &blk.xg, 1},     {OSTC_F_Y, &blk.yg, 1},  {OSTC_F_Z, &blk.zg, 1},
               {OSTC_F_JAC, &blk.jac, 1},  {OSTC_F_METRIC, &blk.metric, 9}, {OSTC_F_BV, &blk.bv, 6},
               {OSTC_F_SV, &blk.sv, c->param.n_scalar}};
    for (auto &u : ups) {
      if (u.planes == 0) continue;
      if (int st = ostc_zone_upload(z, u.field, u.arr->data(), u.planes)) {
        drv_fail(st, "ostc_zone_upload");
        return cleanup(st);
      }
    }
    if (int st = ostc_zone_set_bc(z, &blk.bc)) {
      drv_fail(st, "ostc_zone_set_bc");
      return cleanup(st);
    }
    d->dims.insert(d->dims.end(), {blk.mx, blk.my, blk.mz});
    d->n_inner.push_back((int)blk.inner.size());
    d->inner.insert(d->inner.end(), blk.inner.begin(), blk.inner.end());
    d->n_grid += (long)blk.mx * blk.my * blk.mz;
  }
  *out = d;
  return OSTC_OK;
}

int ostc_driver_initialize(ostc_driver *d) {
  if (!d) return drv_fail(OSTC_ERR_ARG, "null argument");
  const bool rans = d->param.turb_method == OSTC_RANS;
  if (rans) {
    const int n = (int)(d->wall_xyz.size() / 3);
    for (auto *z : d->zones)
      DRV_TRY(ostc_wall_distance(d->ctx, z, n ? d->wall_xyz.data() : nullptr, n, nullptr), "wall distance");
  }
  // 2-D blocks: k ghost layers from the k = 0 plane first (src/Driver.cu:59-65)
  for (auto *z : d->zones) DRV_TRY(ostc_eliminate_k_gradient(d->ctx, z, nullptr), "eliminate_k_gradient");
  for (auto *z : d->zones) DRV_TRY(ostc_apply_bc(d->ctx, z, nullptr), "apply_bc");
  for (auto *z : d->zones) {
    DRV_TRY(ostc_compute_cv_from_bv(d->ctx, z, nullptr), "compute_cv_from_bv");
    if (rans) DRV_TRY(ostc_initialize_mut(d->ctx, z, nullptr), "initialize_mut");
  }
  size_t off = 0;
  for (size_t b = 0; b < d->zones.size(); ++b) {
    for (int f = 0; f < d->n_inner[b]; ++f)
      DRV_TRY(ostc_inner_halo(d->ctx, d->zones.data(), (int)b, &d->inner[off + f], nullptr), "inner_halo");
    off += d->n_inner[b];
  }
  for (auto *z : d->zones) DRV_TRY(ostc_update_physical_properties(d->ctx, z, nullptr), "properties");
  DRV_TRY(ostc_device_sync(), "synchronise");
  d->initialized = true;
  d->t0 = d->t_last = std::chrono::steady_clock::now();
  return OSTC_OK;
}

int ostc_driver_step(ostc_driver *d, double res[4], int *res_computed, double *err_max) {
  if (!d) return drv_fail(OSTC_ERR_ARG, "null argument");
  if (!d->initialized) return drv_fail(OSTC_ERR_STATE, "ostc_driver_initialize has not run");
  ++d->step;
  double sums[4] = {0, 0, 0, 0};
  int computed = 0;
  const int *n_inner = d->n_inner.data();
  const ostc_inner_face *inner = d->inner.empty() ? nullptr : d->inner.data();
  DRV_TRY(ostc_steady_step(d->ctx, d->zones.data(), (int)d->zones.size(), inner, n_inner, d->step, d->output_screen,
                           sums, &computed, nullptr),
          "ostc_steady_step");
  if (res_computed) *res_computed = computed;
  if (!computed) return OSTC_OK;
  for (int l = 0; l < 4; ++l) d->res[l] = std::sqrt(sums[l] / (double)d->n_grid);
  if (d->step == 1)
    for (int l = 0; l < 4; ++l) d->res_scale[l] = d->res[l] < 1e-20 ? 1e-20 : d->res[l];
  for (int l = 0; l < 4; ++l) d->res[l] /= d->res_scale[l];
  double err = d->res[0];
  for (int l = 1; l < 4; ++l)
    if (std::abs(d->res[l]) > err) err = d->res[l];
  if (res)
    for (int l = 0; l < 4; ++l) res[l] = d->res[l];
  if (err_max) *err_max = err;
  if (std::isnan(err)) return drv_fail(OSTC_ERR_NAN, "NaN residual at step " + std::to_string(d->step));
  return OSTC_OK;
}

int ostc_driver_simulate(ostc_driver *d, int io, int *steps_done, double *err_max) {
  if (!d) return drv_fail(OSTC_ERR_ARG, "null argument");
  namespace fs = std::filesystem;
  const std::string history = d->dir + "/history.dat";
  if (io && d->step == 0) {
    if (FILE *h = std::fopen(history.c_str(), "w")) {
      std::fprintf(h, "step\terror_max\n");
      std::fclose(h);
    }
  }
  double err = 0;
  int done = 0;
  while (d->step < d->total_step) {
    int computed = 0;
    double e = 0;
    const int st = ostc_driver_step(d, nullptr, &computed, &e);
    ++done;
    if (st != OSTC_OK) {
      if (io && st == OSTC_ERR_NAN) std::printf("Nan occurred in step %d. Stop simulation.\n", d->step);
      if (steps_done) *steps_done = done;
      return st;
    }
    if (computed) err = e;
    const bool converged = computed && err < d->convergence;
    if (io && (d->step % d->output_file == 0 || converged)) {
      int wst = ostc_driver_write_flowfield(d, nullptr);
      if (wst == OSTC_OK) wst = ostc_driver_post_process(d);
      if (wst != OSTC_OK) {
        if (steps_done) *steps_done = done;
        return wst;
      }
    }
    if (!computed) continue;
    if (io) {
      if (d->step == 1) {
        std::error_code ec;
        fs::create_directories(d->dir + "/output/message", ec);
        if (FILE *f = std::fopen((d->dir + "/output/message/residual_scale.txt").c_str(), "w")) {
          for (double v : d->res_scale) std::fprintf(f, "%g\n", v);
          std::fclose(f);
        }
      }
      const auto now = std::chrono::steady_clock::now();
      const double step_time = std::chrono::duration<double>(now - d->t_last).count();
      const double elapsed = std::chrono::duration<double>(now - d->t0).count();
      d->t_last = now;
      if (FILE *h = std::fopen(history.c_str(), "a")) {
        std::fprintf(h, "%d\t%11.4e\n", d->step, err);
        std::fclose(h);
      }
      std::printf("\n%38s    converged to: %11.4e\n", "rho", d->res[0]);
      std::printf("  n=%8d,                       V     converged to: %11.4e   \n", d->step, d->res[1]);
      std::printf("  n=%8d,                       p     converged to: %11.4e   \n", d->step, d->res[2]);
      std::printf("%38s    converged to: %11.4e\n", "T ", d->res[3]);
      std::printf("Wall time since the last output is %16.8fs\n", step_time);
      std::printf("Total elapsed wall time is %16.8fs\n", elapsed);
      std::fflush(stdout);
    }
    if (err < d->convergence) break;
  }
  DRV_TRY(ostc_device_sync(), "synchronise");
  if (steps_done) *steps_done = done;
  if (err_max) *err_max = err;
  return OSTC_OK;
}

int ostc_driver_write_flowfield(ostc_driver *d, const char *path) {
  if (!d) return drv_fail(OSTC_ERR_ARG, "null argument");
  const std::string file = path ? std::string(path) : d->dir + "/output/field/flowfield.plt";
  const bool rans = d->param.turb_method == OSTC_RANS;
  const int ns = d->param.n_spec, nsc = d->param.n_scalar;
  std::vector<std::string> names{"x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"};
  names.insert(names.end(), d->spec_names.begin(), d->spec_names.end());
  if (rans) {
    names.emplace_back("tke");
    names.emplace_back("omega");
    names.emplace_back("mut");
  }
  std::vector<ostc_host::PltZone> zones;
  for (size_t b = 0; b < d->zones.size(); ++b) {
    ostc_zone *z = d->zones[b];
    const int g = d->ngg, mx = d->dims[3 * b], my = d->dims[3 * b + 1], mz = d->dims[3 * b + 2];
    const long nx = mx + 2 * g, ny = my + 2 * g, plane = nx * ny * (mz + 2 * g);
    std::vector<double> buf((size_t)plane * std::max(6, nsc));
    ostc_host::PltZone pz;
    pz.mx = mx, pz.my = my, pz.mz = mz;
    auto interior = [&](const double *src) {
      std::vector<double> out((size_t)mx * my * mz);
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < mx; ++i)
            out[((size_t)k * my + j) * mx + i] = src[((long)(k + g) * ny + (j + g)) * nx + (i + g)];
      return out;
    };
    for (int f : {OSTC_F_X, OSTC_F_Y, OSTC_F_Z}) {
      DRV_TRY(ostc_zone_download(z, f, buf.data(), 1), "download grid");
      pz.vars.push_back(interior(buf.data()));
    }
    DRV_TRY(ostc_zone_download(z, OSTC_F_BV, buf.data(), 6), "download bv");
    for (int l = 0; l < 6; ++l) pz.vars.push_back(interior(buf.data() + (size_t)l * plane));
    DRV_TRY(ostc_zone_download(z, OSTC_F_MACH, buf.data(), 1), "download mach");
    pz.vars.push_back(interior(buf.data()));
    if (nsc > 0) {
      DRV_TRY(ostc_zone_download(z, OSTC_F_SV, buf.data(), nsc), "download sv");
      for (int l = 0; l < ns + (rans ? 2 : 0); ++l) pz.vars.push_back(interior(buf.data() + (size_t)l * plane));
    }
    if (rans) {
      DRV_TRY(ostc_zone_download(z, OSTC_F_MUT, buf.data(), 1), "download mut");
      pz.vars.push_back(interior(buf.data()));
    }
    zones.push_back(std::move(pz));
  }
  std::string err;
  if (ostc_host::write_plt(file, names, zones, &err)) return drv_fail(OSTC_ERR_IO, err);
  if (!path) {
    // print_field also records the step (src/MPIIO.hpp:421-426)
    std::error_code ec;
    std::filesystem::create_directories(d->dir + "/output/message", ec);
    FILE *f = std::fopen((d->dir + "/output/message/step.txt").c_str(), "w");
    if (!f) return drv_fail(OSTC_ERR_IO, "cannot write output/message/step.txt");
    std::fprintf(f, "%d", d->step);
    std::fclose(f);
  }
  return OSTC_OK;
}

int ostc_driver_post_process(ostc_driver *d) {
  if (!d) return drv_fail(OSTC_ERR_ARG, "null argument");
  for (int proc : d->post_process) {
    if (proc != 0) continue;  // the reference knows procedure 0 only
    // wall_friction_heatflux_2d (src/PostProcess.cu:6-47)
    const std::string dir = d->dir + "/output/wall";
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    size_t mx_max = 0;
    for (size_t b = 0; b < d->zones.size(); ++b) mx_max = std::max(mx_max, (size_t)d->dims[3 * b]);
    void *dev = nullptr;
    DRV_TRY(ostc_device_alloc(&dev, (long)(2 * mx_max * sizeof(double))), "allocate cf/qw");
    double *cf = static_cast<double *>(dev), *qw = cf + mx_max;
    std::vector<double> h_cf(mx_max), h_qw(mx_max), xg;
    int st = OSTC_OK;
    for (size_t b = 0; b < d->zones.size() && st == OSTC_OK; ++b) {
      ostc_zone *z = d->zones[b];
      const int g = d->ngg, mx = d->dims[3 * b], my = d->dims[3 * b + 1], mz = d->dims[3 * b + 2];
      const long nx = mx + 2 * g, ny = my + 2 * g;
      xg.resize((size_t)nx * ny * (mz + 2 * g));
      if ((st = ostc_wall_friction_heatflux(z, d->dyn_pressure, cf, qw, nullptr)) != OSTC_OK ||
          (st = ostc_memcpy_d2h(h_cf.data(), cf, (long)(mx * sizeof(double)))) != OSTC_OK ||
          (st = ostc_memcpy_d2h(h_qw.data(), qw, (long)(mx * sizeof(double)))) != OSTC_OK ||
          (st = ostc_zone_download(z, OSTC_F_X, xg.data(), 1)) != OSTC_OK) {
        drv_fail(st, "wall friction/heat flux");
        break;
      }
      // one file, rewritten for every block (the reference's behaviour)
      FILE *f = std::fopen((dir + "/friction_heatflux.dat").c_str(), "w");
      if (!f) {
        st = drv_fail(OSTC_ERR_IO, "cannot write " + dir + "/friction_heatflux.dat");
        break;
      }
      std::fprintf(f, "variables = \"x\", \"cf\", \"qw\"\n");
      // std::ostream's default floating-point output: %g, 6 significant digits
      for (int i = 0; i < mx; ++i)
        std::fprintf(f, "%g\t%g\t%g\n", xg[((long)g * ny + g) * nx + i + g], h_cf[i], h_qw[i]);
      std::fclose(f);
    }
    ostc_device_free(dev);
    if (st != OSTC_OK) return st;
  }
  return OSTC_OK;
}

int ostc_driver_get_zone(ostc_driver *d, int block, ostc_zone **out) {
  if (!d || !out || block < 0 || block >= (int)d->zones.size()) return drv_fail(OSTC_ERR_ARG, "bad block");
  *out = d->zones[block];
  return OSTC_OK;
}

int ostc_driver_get_ctx(ostc_driver *d, ostc_ctx **out) {
  if (!d || !out) return drv_fail(OSTC_ERR_ARG, "null argument");
  *out = d->ctx;
  return OSTC_OK;
}

int ostc_driver_get_step(const ostc_driver *d, int *step) {
  if (!d || !step) return drv_fail(OSTC_ERR_ARG, "null argument");
  *step = d->step;
  return OSTC_OK;
}

int ostc_driver_destroy(ostc_driver *d) {
  delete d;
  return OSTC_OK;
}

const char *ostc_driver_error(void) { return g_drv_err.c_str(); }

}  // extern "C"
