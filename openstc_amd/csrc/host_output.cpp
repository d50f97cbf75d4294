// Flow-field file of the time loop and the restart that reads it back.
//
// Format: the Tecplot binary (V112) file the reference's parallel writer
// produces (src/MPIIO.hpp:60-176 header, :196-548 data), one ORDERED zone
// per block, point-located data in double precision, no ghost layers:
//   "#!TDV112", 1, file type 0, title, n_var, variable names,
//   per zone: 299.0f, "zone <b>", -1, -2, 0.0 (solution time), -1, 0, 0, 0, 0,
//             IMax, JMax, KMax, 0 (no auxiliary data);
//   357.0f;
//   per zone: 299.0f, n_var x 2 (double), 0, 0, -1, n_var (min, max) pairs,
//             then every variable as IMax*JMax*KMax doubles, i fastest.
// Strings are one int32 per character and a terminating 0 (gxl::write_str,
// src/gxl_lib/MyString.cpp:106-116).  Variables: x y z density u v w
// pressure temperature mach, the species in mechanism order, tke omega
// (SST), mut (RANS) -- src/MPIIO.hpp acquire_variable_names.
//
// The restart (setup `initial = 1`, src/Initialize.cuh:58-200) reads
// <case>/output/field/flowfield.plt and maps variables by their upper-case
// label (src/Initialize.cuh:218-272): DENSITY/ROE/RHO, U, V, W, P/PRESSURE,
// T/TEMPERATURE, species names, K/TKE and OMEGA; other labels are skipped.
// Variables the file lacks keep the setup's inflow initialisation, as the
// reference does for species and turbulence.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "ostc_case.h"

namespace ostc_host {

namespace {

struct Writer {
  FILE *f;
  bool ok = true;
  void raw(const void *p, size_t n) { ok = ok && std::fwrite(p, 1, n, f) == n; }
  void i32(int32_t v) { raw(&v, 4); }
  void f32(float v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  void str(const std::string &s) {
    for (unsigned char ch : s) i32((int32_t)ch);
    i32(0);
  }
};

struct Reader {
  const std::vector<unsigned char> &b;
  size_t at = 0;
  bool ok = true;
  bool take(void *p, size_t n) {
    if (!ok || at + n > b.size()) return ok = false;
    std::memcpy(p, b.data() + at, n);
    at += n;
    return true;
  }
  int32_t i32() {
    int32_t v = 0;
    take(&v, 4);
    return v;
  }
  float f32() {
    float v = 0;
    take(&v, 4);
    return v;
  }
  double f64() {
    double v = 0;
    take(&v, 8);
    return v;
  }
  std::string str() {
    std::string s;
    while (ok) {
      const int32_t ch = i32();
      if (ch == 0) break;
      s += (char)ch;
      if (s.size() > 4096) ok = false;
    }
    return s;
  }
};

}  // namespace

int write_plt(const std::string &path, const std::vector<std::string> &names, const std::vector<PltZone> &zones,
              std::string *err) {
  std::error_code ec;
  const std::filesystem::path dir = std::filesystem::path(path).parent_path();
  if (!dir.empty()) std::filesystem::create_directories(dir, ec);
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    if (err) *err = "cannot open " + path + " for writing";
    return 1;
  }
  Writer w{f};
  const int32_t nv = (int32_t)names.size();
  w.raw("#!TDV112", 8);
  w.i32(1);
  w.i32(0);
  w.str("Solution file");
  w.i32(nv);
  for (const auto &n : names) w.str(n);
  for (size_t b = 0; b < zones.size(); ++b) {
    w.f32(299.0f);
    w.str("zone " + std::to_string(b));
    w.i32(-1);
    w.i32(-2);
    w.f64(0.0);
    w.i32(-1);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(zones[b].mx);
    w.i32(zones[b].my);
    w.i32(zones[b].mz);
    w.i32(0);
  }
  w.f32(357.0f);
  for (const PltZone &z : zones) {
    w.f32(299.0f);
    for (int v = 0; v < nv; ++v) w.i32(2);
    w.i32(0);
    w.i32(0);
    w.i32(-1);
    for (int v = 0; v < nv; ++v) {
      const auto &a = z.vars[v];
      double lo = a.empty() ? 0.0 : a[0], hi = lo;
      for (double x : a) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
      }
      w.f64(lo);
      w.f64(hi);
    }
    for (int v = 0; v < nv; ++v) w.raw(z.vars[v].data(), z.vars[v].size() * sizeof(double));
  }
  const bool ok = w.ok;
  if (std::fclose(f) != 0 || !ok) {
    if (err) *err = "write error on " + path;
    return 1;
  }
  return 0;
}

int read_plt(const std::string &path, std::vector<std::string> *names, std::vector<PltZone> *zones,
             std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "cannot open " + path;
    return 1;
  }
  std::vector<unsigned char> buf;
  unsigned char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
  std::fclose(f);
  Reader r{buf};
  auto fail = [&](const std::string &m) {
    if (err) *err = path + ": " + m;
    return 1;
  };
  char magic[9] = {0};
  if (!r.take(magic, 8) || std::string(magic, 5) != "#!TDV") return fail("not a Tecplot binary file");
  r.i32();  // byte order
  r.i32();  // file type
  r.str();  // title
  const int32_t nv = r.i32();
  if (!r.ok || nv <= 0 || nv > 4096) return fail("bad variable count");
  names->clear();
  for (int v = 0; v < nv; ++v) names->push_back(r.str());
  zones->clear();
  while (r.ok) {
    const float marker = r.f32();
    if (marker == 357.0f) break;
    if (marker != 299.0f) return fail("bad zone marker in the header");
    PltZone z;
    r.str();       // zone name
    r.i32();       // parent zone
    r.i32();       // strand id
    r.f64();       // solution time
    r.i32();       // zone color
    const int32_t type = r.i32();
    const int32_t loc = r.i32();
    r.i32();       // raw face neighbours
    r.i32();       // miscellaneous face connections
    if (type != 0 || loc != 0) return fail("only ORDERED zones with node data are read");
    z.mx = r.i32();
    z.my = r.i32();
    z.mz = r.i32();
    while (r.ok && r.i32() == 1) {  // auxiliary name/value pairs
      r.str();
      r.i32();
      r.str();
    }
    if (z.mx <= 0 || z.my <= 0 || z.mz <= 0) return fail("bad zone size");
    zones->push_back(std::move(z));
  }
  if (!r.ok) return fail("truncated header");
  for (PltZone &z : *zones) {
    if (r.f32() != 299.0f) return fail("bad zone marker in the data section");
    std::vector<int32_t> fmt(nv);
    for (int v = 0; v < nv; ++v) fmt[v] = r.i32();
    if (r.i32() != 0) return fail("passive variables are not read");
    if (r.i32() != 0) return fail("shared variables are not read");
    r.i32();  // connectivity sharing
    for (int v = 0; v < nv; ++v) {
      r.f64();
      r.f64();
    }
    const size_t np = (size_t)z.mx * z.my * z.mz;
    z.vars.assign(nv, std::vector<double>(np));
    for (int v = 0; v < nv; ++v) {
      if (fmt[v] == 2) {
        if (!r.take(z.vars[v].data(), np * 8)) break;
      } else if (fmt[v] == 1) {
        for (size_t q = 0; q < np && r.ok; ++q) z.vars[v][q] = (double)r.f32();
      } else {
        return fail("variable data must be float or double");
      }
    }
    if (!r.ok) return fail("truncated data section");
  }
  return 0;
}

namespace {
std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}
}  // namespace

int restart_fields(ostc_case &c, std::string *err) {
  // the step the file was written at and that run's residual scales (the
  // reference reads both without checking; missing files leave 0 and 1)
  if (FILE *f = std::fopen((c.dir + "/output/message/step.txt").c_str(), "r")) {
    if (std::fscanf(f, "%d", &c.start_step) != 1) c.start_step = 0;
    std::fclose(f);
  }
  if (FILE *f = std::fopen((c.dir + "/output/message/residual_scale.txt").c_str(), "r")) {
    double v[4];
    if (std::fscanf(f, "%lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3]) == 4)
      for (int l = 0; l < 4; ++l) c.res_scale[l] = v[l];
    std::fclose(f);
  }
  const std::string path = c.dir + "/output/field/flowfield.plt";
  std::vector<std::string> names;
  std::vector<PltZone> zones;
  if (read_plt(path, &names, &zones, err)) return 1;
  if (zones.size() != c.blocks.size()) {
    if (err) *err = path + ": " + std::to_string(zones.size()) + " zones for " + std::to_string(c.blocks.size()) +
                    " blocks";
    return 1;
  }
  const int ns = c.param.n_spec, nt = c.param.turb_method == OSTC_RANS ? 2 : 0;
  // label of every file variable: 0..5 basic variables, 6.. scalars, -1 skipped
  std::vector<int> label(names.size(), -1);
  for (size_t v = 0; v < names.size(); ++v) {
    const std::string n = upper(names[v]);
    if (n == "DENSITY" || n == "ROE" || n == "RHO") label[v] = 0;
    else if (n == "U") label[v] = 1;
    else if (n == "V") label[v] = 2;
    else if (n == "W") label[v] = 3;
    else if (n == "P" || n == "PRESSURE") label[v] = 4;
    else if (n == "T" || n == "TEMPERATURE") label[v] = 5;
    else if (nt == 2 && (n == "K" || n == "TKE")) label[v] = 6 + ns;
    else if (nt == 2 && n == "OMEGA") label[v] = 7 + ns;
    else
      for (int k = 0; k < ns; ++k)
        if (n == upper(c.mech.spec_names[k])) label[v] = 6 + k;
  }
  const int g = c.ngg;
  for (size_t b = 0; b < zones.size(); ++b) {
    CaseBlock &blk = c.blocks[b];
    const PltZone &z = zones[b];
    const int gf = (z.mx - blk.mx) / 2;
    if (gf < 0 || gf > g || z.mx != blk.mx + 2 * gf || z.my != blk.my + 2 * gf || z.mz != blk.mz + 2 * gf) {
      if (err) *err = path + ": zone " + std::to_string(b) + " does not match block " + std::to_string(b);
      return 1;
    }
    const long nx = blk.mx + 2 * g, ny = blk.my + 2 * g, plane = blk.plane(g);
    for (size_t v = 0; v < names.size(); ++v) {
      const int l = label[v];
      if (l < 0) continue;
      double *dst = l < 6 ? blk.bv.data() + (size_t)l * plane : blk.sv.data() + (size_t)(l - 6) * plane;
      const auto &src = z.vars[v];
      for (int k = 0; k < z.mz; ++k)
        for (int j = 0; j < z.my; ++j)
          for (int i = 0; i < z.mx; ++i) {
            const long q = ((long)(k + g - gf) * ny + (j + g - gf)) * nx + (i + g - gf);
            dst[q] = src[((size_t)k * z.my + j) * z.mx + i];
          }
    }
  }
  return 0;
}

}  // namespace ostc_host
