// Case reader: a directory laid out as the reference's working directory
// (input_files/setup, input_files/grid, input_files/boundary_condition and
// the CHEMKIN files named in the setup) becomes an ostc_case: model
// constants, block geometry, boundary programs, inflow states, initial fields
// and wall points.  Host C++, FP64, no GPU.
//
// Behaviour follows the reference's own readers:
//  * setup files: Parameter::read_one_file (src/Parameter.cpp:59-168) and the
//    derived counts of Parameter::read_param_from_file (:21-57);
//  * mesh: Mesh::read_grid / read_boundary / read_inner_interface
//    (src/Mesh.cpp:597-735), gridScale (:567-573), ghost grids and metrics
//    through ostc_mesh_geometry;
//  * boundary programs: DBoundCond::initialize_bc_on_GPU and
//    link_bc_to_boundaries (src/BoundCond.cu:73-232), application order of
//    apply_boundary_conditions (src/BoundCond.cuh:321-412);
//  * inflow states: Inflow::Inflow (src/BoundCond.cpp:6-106), walls
//    Wall::Wall (:131-145);
//  * DParameter limits (src/DParameter.cu:78-100);
//  * initial field: initialize_from_start (src/Initialize.cu:4-43) and
//    Field::initialize_basic_variables (src/Field.cu:21-67);
//  * wall points: Driver::acquire_wall_distance (src/Driver.cu:122-149).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "ostc_case.h"

namespace {

thread_local std::string g_case_err;

constexpr double kRu = 8314.462618;   // src/Constants.h:16
constexpr double kMwAir = 28.9647;    // src/Constants.h:26
constexpr double kGammaAir = 1.4;     // src/Constants.h:20

int case_fail(const std::string &msg) {
  g_case_err = msg;
  return OSTC_ERR_ARG;
}

bool starts_with(const std::string &s, const char *p) { return s.rfind(p, 0) == 0; }

std::string ltrim(const std::string &s) {
  const auto p = s.find_first_not_of(" \t\r\n\v\f");
  return p == std::string::npos ? std::string() : s.substr(p);
}

}  // namespace

namespace ostc_host {

// Parameter::read_one_file: "type name = value" lines; "array <type> name {
// v v ... }" possibly over several lines ("//" ends a line's values); "struct
// name { ... }" closed by a line starting with "}".  Lines starting with "//"
// or "!" and empty lines are skipped; other unknown first words are ignored.
bool Setup::read_file(const std::string &path, std::string *err) {
  std::ifstream f(path);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  std::string input;
  auto bad = [&](const std::string &what) {
    if (err) *err = path + ": " + what;
    return false;
  };
  while (std::getline(f, input)) {
    if (input.empty() || starts_with(input, "//") || starts_with(input, "!")) continue;
    std::istringstream line(input);
    std::string type, key, temp;
    line >> type >> key >> temp;
    if (type == "int") {
      int v = 0;
      line >> v;
      ints.emplace(key, v);
    } else if (type == "real") {
      double v = 0;
      line >> v;
      reals.emplace(key, v);
    } else if (type == "bool") {
      bool v = false;
      line >> v;
      bools.emplace(key, v);
    } else if (type == "string") {
      std::string v;
      line >> v;
      strings.emplace(key, v);
    } else if (type == "array") {
      // key = element type, temp = name.  The reference consumes one token
      // after the name for int/real arrays ("{") and two for string arrays,
      // so a string value written on the opening line is dropped.
      const std::string name = temp;
      std::string tok;
      line >> tok;
      if (key == "string") line >> tok;
      std::vector<std::string> vals;
      bool done = false;
      std::istringstream *ls = &line;
      std::istringstream next;
      for (;;) {
        while (*ls >> tok) {
          if (tok == "}") {
            done = true;
            break;
          }
          if (tok == "//") break;
          vals.push_back(tok);
        }
        if (done) break;
        std::string more;
        if (!std::getline(f, more)) break;
        next.clear();
        next.str(ltrim(more));
        ls = &next;
      }
      if (key == "int") {
        std::vector<int> a;
        for (auto &v : vals) {
          char *end = nullptr;
          const long x = std::strtol(v.c_str(), &end, 10);
          if (end == v.c_str()) return bad("array " + name + ": not an integer: " + v);
          a.push_back((int)x);
        }
        int_arrays.emplace(name, a);
      } else if (key == "real") {
        std::vector<double> a;
        for (auto &v : vals) {
          char *end = nullptr;
          const double x = std::strtod(v.c_str(), &end);
          if (end == v.c_str()) return bad("array " + name + ": not a number: " + v);
          a.push_back(x);
        }
        real_arrays.emplace(name, a);
      } else if (key == "string") {
        string_arrays.emplace(name, vals);
      }
    } else if (type == "struct") {
      SetupStruct st;
      std::string more;
      while (std::getline(f, more)) {
        std::istringstream ls(ltrim(more));
        std::string k;
        if (!(ls >> k)) continue;
        if (k == "}") break;
        std::string name, eq;
        if (k == "string") {
          SetupValue v;
          v.kind = SetupValue::STRING;
          ls >> name >> eq >> v.s;
          st.emplace(name, v);
        } else if (k == "int") {
          SetupValue v;
          v.kind = SetupValue::INT;
          ls >> name >> eq >> v.i;
          st.emplace(name, v);
        } else if (k == "real") {
          SetupValue v;
          v.kind = SetupValue::REAL;
          ls >> name >> eq >> v.r;
          st.emplace(name, v);
        }
      }
      structs.emplace(key, st);
    }
  }
  return true;
}

}  // namespace ostc_host

namespace {

using ostc_host::CaseBlock;
using ostc_host::InflowState;
using ostc_host::Setup;
using ostc_host::SetupStruct;

struct Need {
  const Setup &s;
  std::string missing;
  int i(const char *k) {
    auto it = s.ints.find(k);
    if (it == s.ints.end()) {
      if (missing.empty()) missing = std::string("int ") + k;
      return 0;
    }
    return it->second;
  }
  double r(const char *k) {
    auto it = s.reals.find(k);
    if (it == s.reals.end()) {
      if (missing.empty()) missing = std::string("real ") + k;
      return 0;
    }
    return it->second;
  }
  bool b(const char *k) {
    auto it = s.bools.find(k);
    if (it == s.bools.end()) {
      if (missing.empty()) missing = std::string("bool ") + k;
      return false;
    }
    return it->second;
  }
  std::string str(const char *k) {
    auto it = s.strings.find(k);
    if (it == s.strings.end()) {
      if (missing.empty()) missing = std::string("string ") + k;
      return std::string();
    }
    return it->second;
  }
};

// cp_k(T) in J/(kg K): Species::compute_cp (src/ChemData.cpp:100-117)
void species_cp(const ostc_param &p, double temp, double *cp) {
  const double t2 = temp * temp, t3 = t2 * temp, t4 = t3 * temp;
  for (int i = 0; i < p.n_spec; ++i) {
    const double *c;
    if (temp < p.t_low[i]) {
      const double tt = p.t_low[i], tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      c = p.low_temp_coeff + 7 * i;
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      c = (temp < p.t_mid[i] ? p.low_temp_coeff : p.high_temp_coeff) + 7 * i;
      cp[i] = c[0] + c[1] * temp + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= kRu / p.mw[i];
  }
}

// compute_viscosity (host form, src/Transport.cu:11-38): Wilke's rule
double mixture_viscosity(const ostc_param &p, double t, double mw_total, const double *y) {
  const int ns = p.n_spec;
  double x[OSTC_MAX_SPEC], vis[OSTC_MAX_SPEC];
  for (int i = 0; i < ns; ++i) {
    x[i] = y[i] * mw_total / p.mw[i];
    const double t_dl = t * p.LJ_potent_inv[i];
    const double ci = 1.147 * std::pow(t_dl, -0.145) + std::pow(t_dl + 0.5, -2);
    vis[i] = p.vis_coeff[i] * std::sqrt(t) / ci;
  }
  double pf[OSTC_MAX_SPEC][OSTC_MAX_SPEC];
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      if (i == j) {
        pf[i][j] = 1.0;
      } else {
        const double num = 1 + std::sqrt(vis[i] / vis[j]) * p.WjDivWi_to_One4th[i * OSTC_MAX_SPEC + j];
        pf[i][j] = num * num * p.sqrt_WiDivWjPl1Mul8[i * OSTC_MAX_SPEC + j];
      }
    }
  double mu = 0;
  for (int i = 0; i < ns; ++i) {
    double s = 0;
    for (int j = 0; j < ns; ++j) s += pf[i][j] * x[j];
    mu += vis[i] * x[i] / s;
  }
  return mu;
}

double sutherland(double t) { return 1.716e-5 * std::pow(t / 273, 1.5) * (273 + 111) / (t + 111); }

struct CaseCtx {
  ostc_case *c;
  bool sst = false;
  bool have_inf = false;
  double rho_inf = 0, v_inf = 0, p_inf = 0;
  std::vector<double> sv_inf;
};

// Inflow::Inflow (src/BoundCond.cpp:6-106) for the struct named `name`
int make_inflow(CaseCtx &cx, const std::string &name, InflowState *out) {
  const ostc_case &c = *cx.c;
  auto it = c.setup.structs.find(name);
  if (it == c.setup.structs.end()) return case_fail("inflow struct '" + name + "' is not defined in the setup files");
  const SetupStruct &info = it->second;
  InflowState st;
  st.name = name;
  auto num = [&](const char *k, double *dst) {
    auto f = info.find(k);
    if (f != info.end()) *dst = f->second.number();
  };
  auto lab = info.find("label");
  if (lab != info.end()) st.label = (int)lab->second.number();
  num("mach", &st.mach);
  num("pressure", &st.pressure);
  num("temperature", &st.temperature);
  num("velocity", &st.velocity);
  num("density", &st.density);
  num("u", &st.u);
  num("v", &st.v);
  num("w", &st.w);
  const ostc_param &p = c.param;
  const int n_scalar = p.n_scalar, ns = p.n_spec;
  st.sv.assign(n_scalar, 0.0);
  st.mw = kMwAir;
  if (ns > 0) {
    for (int i = 0; i < ns; ++i) {
      auto f = info.find(c.mech.spec_names[i]);
      if (f != info.end()) st.sv[i] = f->second.number();
    }
    double mw = 0;
    for (int i = 0; i < ns; ++i) mw += st.sv[i] / p.mw[i];
    st.mw = 1 / mw;
  }
  if (st.temperature < 0) st.temperature = st.pressure * st.mw / (st.density * kRu);
  st.viscosity = ns > 0 ? mixture_viscosity(p, st.temperature, st.mw, st.sv.data()) : sutherland(st.temperature);
  double gamma = kGammaAir;
  if (ns > 0) {
    double cpi[OSTC_MAX_SPEC];
    species_cp(p, st.temperature, cpi);
    double cp = 0, cv = 0;
    for (int i = 0; i < ns; ++i) {
      cp += st.sv[i] * cpi[i];
      cv += st.sv[i] * (cpi[i] - kRu / p.mw[i]);
    }
    gamma = cp / cv;
  }
  const double sound = std::sqrt(gamma * kRu / st.mw * st.temperature);
  if (st.mach < 0)
    st.mach = st.velocity / sound;
  else
    st.velocity = st.mach * sound;
  st.u *= st.velocity;
  st.v *= st.velocity;
  st.w *= st.velocity;
  if (st.density < 0) st.density = st.pressure * st.mw / (kRu * st.temperature);
  if (cx.sst) {
    auto r = info.find("turb_viscosity_ratio");
    if (r == info.end()) return case_fail("inflow struct '" + name + "' needs turb_viscosity_ratio for SST");
    st.mut = r->second.number() * st.viscosity;
    auto ti = info.find("turbulence_intensity");
    if (ti != info.end()) {
      const double inten = ti->second.number();
      st.sv[ns] = 1.5 * st.velocity * st.velocity * inten * inten;
      st.sv[ns + 1] = st.density * st.sv[ns] / st.mut;
    }
  }
  if (name == "freestream") {
    cx.have_inf = true;
    cx.rho_inf = st.density;
    cx.v_inf = st.velocity;
    cx.p_inf = st.pressure;
    cx.sv_inf.assign(st.sv.begin(), st.sv.begin() + n_scalar);
  }
  *out = st;
  return OSTC_OK;
}

// whitespace-separated numbers of a whole file
struct NumReader {
  std::string buf;
  size_t pos = 0;
  bool open(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    buf = ss.str();
    return true;
  }
  bool line(std::string *out) {
    if (pos >= buf.size()) return false;
    size_t e = buf.find('\n', pos);
    if (e == std::string::npos) e = buf.size();
    *out = buf.substr(pos, e - pos);
    pos = e + 1;
    return true;
  }
  bool number(double *v) {
    while (pos < buf.size() && std::isspace((unsigned char)buf[pos])) ++pos;
    if (pos >= buf.size()) return false;
    const char *s = buf.c_str() + pos;
    char *end = nullptr;
    *v = std::strtod(s, &end);
    if (end == s) return false;
    pos += (size_t)(end - s);
    return true;
  }
};

std::string rank_file(const std::string &dir, const char *stem, const char *ext) {
  char name[64];
  std::snprintf(name, sizeof(name), "%s%4d.%s", stem, 0, ext);  // fmt "{:>4}" of rank 0
  return dir + "/" + name;
}

// Mesh::read_grid (src/Mesh.cpp:597-640)
int read_grid(ostc_case &c, const std::string &path) {
  NumReader r;
  if (!r.open(path)) return case_fail("cannot open grid file " + path);
  std::string ln;
  if (!r.line(&ln)) return case_fail(path + ": empty");
  const int nb = std::atoi(ln.c_str());
  if (nb <= 0 || nb > OSTC_MAX_BLOCK) return case_fail(path + ": bad block count " + ln);
  c.blocks.resize(nb);
  for (int b = 0; b < nb; ++b) {
    if (!r.line(&ln)) return case_fail(path + ": missing block dimensions");
    std::istringstream ls(ln);
    CaseBlock &blk = c.blocks[b];
    ls >> blk.mx >> blk.my >> blk.mz;
    if (blk.mx < 2 || blk.my < 2 || blk.mz < 1) return case_fail(path + ": bad block dimensions " + ln);
  }
  for (int b = 0; b < nb; ++b) {
    CaseBlock &blk = c.blocks[b];
    const size_t n = (size_t)blk.mx * blk.my * blk.mz;
    for (auto *arr : {&blk.x, &blk.y, &blk.z}) {
      arr->resize(n);
      for (size_t q = 0; q < n; ++q)
        if (!r.number(&(*arr)[q])) return case_fail(path + ": too few coordinates");
    }
  }
  return OSTC_OK;
}

// Mesh::read_boundary (src/Mesh.cpp:676-696)
int read_boundary(ostc_case &c, const std::string &path) {
  std::ifstream f(path);
  if (!f) return case_fail("cannot open boundary file " + path);
  std::string ln;
  for (auto &blk : c.blocks) {
    if (!std::getline(f, ln)) return case_fail(path + ": missing face count");
    const int n = std::atoi(ln.c_str());
    for (int q = 0; q < n; ++q) {
      if (!std::getline(f, ln)) return case_fail(path + ": missing face record");
      std::istringstream ls(ln);
      std::array<int, 7> rec{};
      for (auto &v : rec)
        if (!(ls >> v)) return case_fail(path + ": bad face record: " + ln);
      blk.boundary.push_back(rec);
    }
  }
  return OSTC_OK;
}

// Mesh::read_inner_interface (src/Mesh.cpp:698-735); a missing file means
// no inner faces
int read_inner(ostc_case &c, const std::string &path) {
  std::ifstream f(path);
  if (!f) return OSTC_OK;
  std::string ln;
  std::getline(f, ln);
  std::vector<int> n_face(c.blocks.size(), 0);
  for (auto &n : n_face)
    if (!(f >> n)) return case_fail(path + ": missing inner-face counts");
  std::getline(f, ln);
  std::getline(f, ln);
  for (size_t b = 0; b < c.blocks.size(); ++b) {
    for (int q = 0; q < n_face[b]; ++q) {
      int a[7], t[7];
      std::string l1, l2;
      if (!std::getline(f, l1) || !std::getline(f, l2)) return case_fail(path + ": missing inner-face record");
      std::istringstream s1(l1), s2(l2);
      for (int &v : a)
        if (!(s1 >> v)) return case_fail(path + ": bad inner-face record: " + l1);
      for (int &v : t)
        if (!(s2 >> v)) return case_fail(path + ": bad inner-face record: " + l2);
      std::array<int, 13> raw{a[0], a[1], a[2], a[3], a[4], a[5], t[0], t[1], t[2], t[3], t[4], t[5], t[6]};
      c.blocks[b].inner_raw.push_back(raw);
      ostc_inner_face face{};
      const int st = ostc_register_inner_face(raw.data(), c.ngg, c.dimension, &face);
      if (st != OSTC_OK) return case_fail(path + ": inner face cannot be registered: " + l1);
      if (face.target_block < 0 || face.target_block >= (int)c.blocks.size())
        return case_fail(path + ": inner face targets a missing block");
      c.blocks[b].inner.push_back(face);
    }
  }
  return OSTC_OK;
}

// Boundary::register_boundary (src/Mesh.cpp:14-32); a 2-D face spans the
// k ghost layers [-ngg, ngg]
void register_face(const std::array<int, 7> &rec, int ngg, int dim, int rs[3], int re[3], int *face, int *dir) {
  rs[0] = rec[0], re[0] = rec[1], rs[1] = rec[2], re[1] = rec[3], rs[2] = rec[4], re[2] = rec[5];
  *face = 0;
  *dir = 1;
  for (int q = 0; q < dim; ++q)
    if (rs[q] == re[q]) {
      *face = q;
      if (rs[q] == 0) *dir = -1;
      break;
    }
  for (int q = 0; q < dim; ++q) {
    if (q == *face) continue;
    if (re[q] < rs[q]) std::swap(rs[q], re[q]);
    rs[q] -= ngg;
    re[q] += ngg;
  }
  if (dim == 2) {
    rs[2] = -ngg;
    re[2] = ngg;
  }
}

enum { T_WALL = 0, T_INFLOW, T_OUTFLOW, T_SYMMETRY };

// DBoundCond::initialize_bc_on_GPU + link_bc_to_boundaries (src/BoundCond.cu:
// 73-232) and Driver::acquire_wall_distance's wall points (src/Driver.cu:
// 128-149)
int boundary_programs(CaseCtx &cx) {
  ostc_case &c = *cx.c;
  std::vector<int> labels;
  for (auto &blk : c.blocks)
    for (auto &rec : blk.boundary)
      if (std::find(labels.begin(), labels.end(), rec[6]) == labels.end()) labels.push_back(rec[6]);
  std::vector<int> idx[4];
  auto names_it = c.setup.string_arrays.find("boundary_conditions");
  static const std::vector<std::string> none;
  const std::vector<std::string> &names = names_it == c.setup.string_arrays.end() ? none : names_it->second;
  for (auto &nm : names) {
    auto s = c.setup.structs.find(nm);
    if (s == c.setup.structs.end()) return case_fail("boundary condition '" + nm + "' has no struct");
    auto lab = s->second.find("label");
    auto typ = s->second.find("type");
    if (lab == s->second.end() || typ == s->second.end())
      return case_fail("boundary condition '" + nm + "' needs label and type");
    const int label = (int)lab->second.number();
    auto it = std::find(labels.begin(), labels.end(), label);
    if (it == labels.end()) continue;
    labels.erase(it);
    const std::string &t = typ->second.s;
    if (t == "wall") idx[T_WALL].push_back(label);
    else if (t == "inflow") idx[T_INFLOW].push_back(label);
    else if (t == "outflow") idx[T_OUTFLOW].push_back(label);
    else if (t == "symmetry") idx[T_SYMMETRY].push_back(label);
  }
  for (int lab : labels) {
    if (lab == 2) idx[T_WALL].push_back(lab);
    else if (lab == 3) idx[T_SYMMETRY].push_back(lab);
    else if (lab == 5) idx[T_INFLOW].push_back(lab);
    else if (lab == 6) idx[T_OUTFLOW].push_back(lab);
  }
  if (idx[T_INFLOW].size() > OSTC_MAX_BCDATA || idx[T_WALL].size() > OSTC_MAX_BCDATA)
    return case_fail("too many inflow or wall conditions");
  // inflow states: the first struct with the label (register_bc<Inflow>)
  std::vector<InflowState> inflows;
  for (int lab : idx[T_INFLOW]) {
    std::string found;
    for (auto &nm : names) {
      auto &st = c.setup.structs.at(nm);
      if ((int)st.at("label").number() == lab) {
        found = nm;
        break;
      }
    }
    if (found.empty()) return case_fail("inflow label " + std::to_string(lab) + " has no struct in boundary_conditions");
    InflowState s;
    if (int e = make_inflow(cx, found, &s)) return e;
    inflows.push_back(s);
  }
  // walls: the last struct with the label (register_bc<Wall> does not stop
  // at the first match); defaults isothermal 300 K (src/BoundCond.h:28-38)
  std::vector<ostc_wall> walls;
  for (int lab : idx[T_WALL]) {
    ostc_wall w{1, 300.0};
    for (auto &nm : names) {
      auto &st = c.setup.structs.at(nm);
      if ((int)st.at("label").number() != lab) continue;
      w = ostc_wall{1, 300.0};
      auto th = st.find("thermal_type");
      if (th != st.end()) w.isothermal = th->second.s == "isothermal" ? 1 : 0;
      if (w.isothermal) {
        auto tw = st.find("temperature");
        if (tw != st.end()) w.temperature = tw->second.number();
      }
    }
    walls.push_back(w);
  }
  c.inflows = inflows;
  static const int code[4] = {2, 5, 6, 3};
  static const int order[4] = {T_OUTFLOW, T_INFLOW, T_SYMMETRY, T_WALL};
  for (auto &blk : c.blocks) {
    ostc_bc &bc = blk.bc;
    std::memset(&bc, 0, sizeof(bc));
    int n = 0;
    for (int t : order)
      for (size_t slot = 0; slot < idx[t].size(); ++slot)
        for (auto &rec : blk.boundary) {
          if (rec[6] != idx[t][slot]) continue;
          if (n >= OSTC_MAX_FACE) return case_fail("too many boundary faces in one block");
          ostc_face &f = bc.face[n++];
          register_face(rec, c.ngg, c.dimension, f.range_start, f.range_end, &f.face, &f.direction);
          if (t == T_WALL)
            for (int q = 0; q < 3; ++q) {
              if (q == f.face) continue;
              f.range_start[q] += c.ngg;
              f.range_end[q] -= c.ngg;
            }
          f.kind = code[t];
          f.slot = (int)slot;
        }
    bc.n_face = n;
    for (size_t s = 0; s < inflows.size(); ++s) {
      ostc_inflow &iv = bc.inflow[s];
      const InflowState &st = inflows[s];
      iv.density = st.density, iv.u = st.u, iv.v = st.v, iv.w = st.w;
      iv.pressure = st.pressure, iv.temperature = st.temperature, iv.mut = st.mut;
      for (size_t l = 0; l < st.sv.size() && l < OSTC_MAX_SPEC + 2; ++l) iv.sv[l] = st.sv[l];
    }
    for (size_t s = 0; s < walls.size(); ++s) bc.wall[s] = walls[s];
  }
  // wall points: wall conditions in order, blocks in order, faces in file
  // order, over the face's own (sorted) index range
  c.wall_xyz.clear();
  for (int lab : idx[T_WALL])
    for (auto &blk : c.blocks)
      for (auto &rec : blk.boundary) {
        if (rec[6] != lab) continue;
        int rs[3], re[3], face, dir;
        register_face(rec, c.ngg, c.dimension, rs, re, &face, &dir);
        for (int q = 0; q < 3; ++q)
          if (q != face) rs[q] += c.ngg, re[q] -= c.ngg;
        for (int k = rs[2]; k <= re[2]; ++k)
          for (int j = rs[1]; j <= re[1]; ++j)
            for (int i = rs[0]; i <= re[0]; ++i) {
              if (i < 0 || j < 0 || k < 0 || i >= blk.mx || j >= blk.my || k >= blk.mz)
                return case_fail("wall face outside its block");
              const size_t q = ((size_t)k * blk.my + j) * blk.mx + i;
              c.wall_xyz.push_back(blk.x[q]);
              c.wall_xyz.push_back(blk.y[q]);
              c.wall_xyz.push_back(blk.z[q]);
            }
      }
  return OSTC_OK;
}

// initialize_from_start (src/Initialize.cu:4-43) + Field::
// initialize_basic_variables (src/Field.cu:21-67).  Patch l (init_cond_{l-1})
// is tested against its own box; the reference indexes the boxes one ahead
// of the states (reading past its box vectors for the last patch).
int initial_fields(CaseCtx &cx) {
  ostc_case &c = *cx.c;
  Need need{c.setup, {}};
  const int groups = need.i("groups_init");
  const std::string def = need.str("default_init");
  if (!need.missing.empty()) return case_fail("setup is missing " + need.missing);
  std::vector<InflowState> st(1);
  if (int e = make_inflow(cx, def, &st[0])) return e;
  struct Box {
    double x0, x1, y0, y1, z0, z1;
  };
  std::vector<Box> boxes(1);
  for (int l = 0; l + 1 < groups; ++l) {
    const std::string pname = "init_cond_" + std::to_string(l);
    auto it = c.setup.structs.find(pname);
    if (it == c.setup.structs.end()) return case_fail("setup is missing struct " + pname);
    const SetupStruct &pc = it->second;
    Box bx{};
    const char *keys[6] = {"x0", "x1", "y0", "y1", "z0", "z1"};
    double *dst[6] = {&bx.x0, &bx.x1, &bx.y0, &bx.y1, &bx.z0, &bx.z1};
    for (int q = 0; q < 6; ++q) {
      auto f = pc.find(keys[q]);
      if (f == pc.end()) return case_fail(pname + " is missing " + keys[q]);
      *dst[q] = f->second.number();
    }
    InflowState s;
    auto lab = pc.find("label");
    if (lab != pc.end() && lab->second.kind == ostc_host::SetupValue::STRING) {
      if (int e = make_inflow(cx, lab->second.s, &s)) return e;
    } else {
      if (int e = make_inflow(cx, pname, &s)) return e;
    }
    st.push_back(s);
    boxes.push_back(bx);
  }
  const int g = c.ngg, nsc = c.param.n_scalar;
  for (auto &blk : c.blocks) {
    const long plane = blk.plane(g);
    blk.bv.assign(6 * plane, 0.0);
    blk.sv.assign((size_t)std::max(nsc, 0) * plane, 0.0);
    const int nx = blk.mx + 2 * g, ny = blk.my + 2 * g, nz = blk.mz + 2 * g;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long q = ((long)k * ny + j) * nx + i;
          const double x = blk.xg[q], y = blk.yg[q], z = blk.zg[q];
          size_t pick = 0;
          for (size_t l = 1; l < st.size(); ++l) {
            const Box &b = boxes[l];
            if (x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1 && z >= b.z0 && z <= b.z1) {
              pick = l;
              break;
            }
          }
          const InflowState &s = st[pick];
          const double v[6] = {s.density, s.u, s.v, s.w, s.pressure, s.temperature};
          for (int l = 0; l < 6; ++l) blk.bv[l * plane + q] = v[l];
          for (int l = 0; l < nsc; ++l) blk.sv[l * plane + q] = s.sv[l];
        }
  }
  return OSTC_OK;
}

int read_case(const char *dir_c, ostc_case *c) {
  c->dir = dir_c;
  const std::string in = c->dir + "/input_files";
  static const char *files[9] = {"0_global_control.txt", "1_grid_information.txt", "2_scheme.txt",
                                 "3_species_reactions.txt", "4_laminar_turbulent.txt", "5_boundary_condition.txt",
                                 "6_post_process.txt", "8_initialization.txt", "9_transport_property.txt"};
  int n_read = 0;
  for (const char *f : files) {
    std::string err;
    if (c->setup.read_file(in + "/setup/" + f, &err))
      ++n_read;
    else if (err.rfind("cannot open", 0) != 0)
      return case_fail(err);
  }
  if (n_read == 0) return case_fail("no setup files under " + in + "/setup");
  Setup &s = c->setup;
  Need need{s, {}};
  // global control and grid information
  const bool parallel = need.b("parallel");
  c->initial = need.i("initial");
  c->total_step = need.i("total_step");
  c->output_file = need.i("output_file");
  c->output_screen = need.i("output_screen");
  c->dimension = need.i("dimension");
  c->grid_scale = need.r("gridScale");
  const int wall_method = need.i("wall_distance");
  // scheme
  c->steady = need.b("steady") ? 1 : 0;
  ostc_param &p = c->param;
  std::memset(&p, 0, sizeof(p));
  p.cfl = need.r("cfl");
  p.implicit_method = need.i("implicit_method");
  p.dplur_inner_step = p.implicit_method == 1 ? need.i("DPLUR_inner_step") : 0;
  c->convergence = need.r("convergence_criteria");
  p.inviscid_scheme = need.i("inviscid_scheme");
  p.reconstruction = need.i("reconstruction");
  p.limiter = need.i("limiter");
  p.viscous_scheme = need.i("viscous_order");
  // species / turbulence
  const bool species = need.b("species");
  const bool reaction = species ? need.b("reaction") : false;
  const bool turbulence = need.b("turbulence");
  int turb_method = need.i("turbulence_method");
  int rans_model = turbulence ? need.i("RANS_model") : 0;
  if (!turbulence) turb_method = 0;
  p.turb_implicit = turbulence ? need.i("turb_implicit") : 1;
  p.chem_src_method = species && reaction ? need.i("chemSrcMethod") : 0;
  // transport
  p.Pr = need.r("prandtl_number");
  p.Sc = need.r("schmidt_number");
  p.Prt = need.r("turbulent_prandtl_number");
  p.Sct = need.r("turbulent_schmidt_number");
  if (!need.missing.empty()) return case_fail("setup is missing " + need.missing);
  if (parallel) return case_fail("parallel = 1: one process drives every block here (no parallel faces)");
  if (!c->steady) return case_fail("steady = 0: only the steady time loop is on this path");
  if (c->initial != 0 && c->initial != 1)
    return case_fail("initial must be 0 (from the setup's inflows) or 1 (restart from output/field/flowfield.plt)");
  if (c->dimension != 2 && c->dimension != 3) return case_fail("dimension must be 2 or 3");
  if (p.inviscid_scheme / 10 == 5) return case_fail("WENO schemes are not on this path");
  if (p.inviscid_scheme != 3) return case_fail("inviscid_scheme must be 3 (AUSM+)");
  if (p.reconstruction < 1 || p.reconstruction > 3) return case_fail("reconstruction must be 1, 2 or 3");
  if (p.limiter != 0 && p.limiter != 1) return case_fail("limiter must be 0 or 1 (minmod)");
  if (p.viscous_scheme != 0 && p.viscous_scheme != 2) return case_fail("viscous_order must be 0 or 2");
  if (p.implicit_method != 0 && p.implicit_method != 1) return case_fail("implicit_method must be 0 or 1 (DPLUR)");
  if (turb_method == 1 && rans_model != 2) return case_fail("RANS_model must be 2 (k-omega SST)");
  if (wall_method != 0 && turb_method == 1) return case_fail("wall_distance = 1 (read) is not supported");
  if (c->output_screen <= 0 || c->output_file <= 0) return case_fail("output_screen and output_file must be positive");
  p.turb_method = turb_method == 1 ? OSTC_RANS : OSTC_LAMINAR;
  p.rans_model = turb_method == 1 ? 2 : 0;
  const int n_turb = turb_method == 1 ? 2 : 0;
  p.mix_model = species ? (reaction ? OSTC_FR : OSTC_MIXTURE) : OSTC_AIR;
  if (species) {
    auto str = [&](const char *k) {
      auto it = s.strings.find(k);
      return it == s.strings.end() ? std::string() : it->second;
    };
    const std::string mech = str("mechanism_file"), therm = str("therm_file"), tran = str("transport_file");
    if (mech.empty() || therm.empty() || tran.empty())
      return case_fail("species = 1 needs mechanism_file, therm_file and transport_file");
    const int st = ostc_read_chemkin((in + "/" + mech).c_str(), (in + "/" + therm).c_str(), (in + "/" + tran).c_str(),
                                     reaction ? 1 : 0, &p, &c->mech);
    if (st != OSTC_OK) return case_fail(std::string("CHEMKIN: ") + ostc_chemkin_error());
  }
  p.n_scalar = p.n_spec + n_turb;
  p.n_var = 5 + p.n_scalar;
  c->ngg = 2;
  // mesh
  const std::string bcdir = in + "/boundary_condition";
  if (int e = read_grid(*c, rank_file(in + "/grid", "grid", "grd"))) return e;
  // the grid decides: a first block with mz = 1 makes the mesh 2-D
  // (src/Mesh.cpp:611-613); every block must then be one k plane
  if (c->blocks[0].mz == 1) c->dimension = 2;
  for (auto &blk : c->blocks) {
    if (c->dimension == 2 && blk.mz != 1) return case_fail("dimension = 2 (a 2-D mesh) needs mz = 1 in every block");
    if (c->dimension == 3 && blk.mz < 3) return case_fail("3-D blocks need mz >= 3");
  }
  if (std::abs(c->grid_scale - 1.0) > 1e-15)
    for (auto &blk : c->blocks)
      for (auto *arr : {&blk.x, &blk.y, &blk.z})
        for (auto &v : *arr) v *= c->grid_scale;
  if (int e = read_boundary(*c, rank_file(bcdir, "boundary", "txt"))) return e;
  if (int e = read_inner(*c, rank_file(bcdir, "inner", "txt"))) return e;
  // geometry (ghost grids incl. inner faces, Jacobian, metrics)
  const int nb = (int)c->blocks.size(), g = c->ngg;
  std::vector<ostc_block_coords> bc(nb);
  std::vector<ostc_inner_face> inner;
  std::vector<int> n_inner(nb);
  for (int b = 0; b < nb; ++b) {
    CaseBlock &blk = c->blocks[b];
    const long plane = blk.plane(g);
    blk.xg.assign(plane, 0.0), blk.yg.assign(plane, 0.0), blk.zg.assign(plane, 0.0);
    blk.jac.assign(plane, 0.0), blk.metric.assign(9 * plane, 0.0);
    ostc_block_coords &q = bc[b];
    q.mx = blk.mx, q.my = blk.my, q.mz = blk.mz;
    q.x = blk.x.data(), q.y = blk.y.data(), q.z = blk.z.data();
    q.x_out = blk.xg.data(), q.y_out = blk.yg.data(), q.z_out = blk.zg.data();
    q.jac = blk.jac.data(), q.metric = blk.metric.data();
    n_inner[b] = (int)blk.inner.size();
    inner.insert(inner.end(), blk.inner.begin(), blk.inner.end());
  }
  if (inner.empty()) inner.resize(1);
  if (int st = ostc_mesh_geometry(nb, g, bc.data(), inner.data(), n_inner.data())) {
    case_fail(std::string("mesh geometry: ") + ostc_last_error());
    return st;
  }
  for (int b = 0; b < nb; ++b) c->blocks[b].n_neg_jac = bc[b].n_neg_jac;
  // inflows, boundary programs, wall points, DParameter limits, initial field
  CaseCtx cx{c};
  cx.sst = p.rans_model == 2;
  if (int e = boundary_programs(cx)) return e;
  if (int e = initial_fields(cx)) return e;
  if (c->initial == 1) {
    std::string err;
    if (ostc_host::restart_fields(*c, &err)) return case_fail("restart: " + err);
  }
  if (!cx.have_inf) {
    InflowState fs;
    if (int e = make_inflow(cx, "freestream", &fs)) {
      case_fail("a struct named 'freestream' is required: it defines rho_inf, v_inf, p_inf and sv_inf");
      return e;
    }
  }
  c->rho_inf = cx.rho_inf;
  c->v_inf = cx.v_inf;
  if (auto pp = s.int_arrays.find("post_process"); pp != s.int_arrays.end()) c->post_process = pp->second;
  p.limit_ll[0] = 1e-6 * cx.rho_inf;
  p.limit_ul[0] = 1e+3 * cx.rho_inf;
  for (int l = 1; l < 4; ++l) {
    p.limit_ll[l] = -1e+3 * cx.v_inf;
    p.limit_ul[l] = 1e+3 * cx.v_inf;
  }
  p.limit_ll[4] = 1e-6 * cx.p_inf;
  p.limit_ul[4] = 1e+3 * cx.p_inf;
  if (p.rans_model == 2) {
    p.limit_ul[5] = std::numeric_limits<double>::max();
    p.limit_ul[6] = std::numeric_limits<double>::max();
  }
  for (int l = 0; l < p.n_scalar && l < (int)cx.sv_inf.size(); ++l) p.sv_inf[l] = cx.sv_inf[l];
  return OSTC_OK;
}

}  // namespace

extern "C" {

int ostc_case_read(const char *dir, ostc_case **out) {
  if (!dir || !out) return case_fail("null argument");
  *out = nullptr;
  auto *c = new ostc_case;
  const int st = read_case(dir, c);
  if (st != OSTC_OK) {
    delete c;
    return st;
  }
  *out = c;
  return OSTC_OK;
}

int ostc_case_destroy(ostc_case *c) {
  delete c;
  return OSTC_OK;
}

int ostc_case_get_info(const ostc_case *c, ostc_case_info *info) {
  if (!c || !info) return case_fail("null argument");
  std::memset(info, 0, sizeof(*info));
  info->n_block = (int)c->blocks.size();
  info->ngg = c->ngg;
  info->total_step = c->total_step;
  info->output_screen = c->output_screen;
  info->output_file = c->output_file;
  info->n_wall_point = (int)(c->wall_xyz.size() / 3);
  info->convergence_criteria = c->convergence;
  info->grid_scale = c->grid_scale;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    info->mx[b] = c->blocks[b].mx;
    info->my[b] = c->blocks[b].my;
    info->mz[b] = c->blocks[b].mz;
    info->n_inner[b] = (int)c->blocks[b].inner.size();
    info->n_neg_jac[b] = c->blocks[b].n_neg_jac;
  }
  info->start_step = c->start_step;
  for (int l = 0; l < 4; ++l) info->res_scale[l] = c->res_scale[l];
  info->rho_inf = c->rho_inf;
  info->v_inf = c->v_inf;
  return OSTC_OK;
}

int ostc_case_get_param(const ostc_case *c, ostc_param *out) {
  if (!c || !out) return case_fail("null argument");
  *out = c->param;
  return OSTC_OK;
}

int ostc_case_get_mechanism(const ostc_case *c, ostc_mechanism_info *out) {
  if (!c || !out) return case_fail("null argument");
  *out = c->mech;
  return OSTC_OK;
}

int ostc_case_get_bc(const ostc_case *c, int block, ostc_bc *out) {
  if (!c || !out || block < 0 || block >= (int)c->blocks.size()) return case_fail("bad block");
  *out = c->blocks[block].bc;
  return OSTC_OK;
}

int ostc_case_get_inner(const ostc_case *c, int block, ostc_inner_face *out, int max_faces, int *n_faces) {
  if (!c || block < 0 || block >= (int)c->blocks.size() || !n_faces) return case_fail("bad block");
  const auto &f = c->blocks[block].inner;
  *n_faces = (int)f.size();
  if (out)
    for (int q = 0; q < (int)f.size() && q < max_faces; ++q) out[q] = f[q];
  return OSTC_OK;
}

int ostc_case_field(ostc_case *c, int block, int field, double *host, int n_planes, int write) {
  if (!c || !host || block < 0 || block >= (int)c->blocks.size()) return case_fail("bad block");
  CaseBlock &blk = c->blocks[block];
  std::vector<double> *arr = nullptr;
  switch (field) {
    case OSTC_F_X: arr = &blk.xg; break;
    case OSTC_F_Y: arr = &blk.yg; break;
    case OSTC_F_Z: arr = &blk.zg; break;
    case OSTC_F_JAC: arr = &blk.jac; break;
    case OSTC_F_METRIC: arr = &blk.metric; break;
    case OSTC_F_BV: arr = &blk.bv; break;
    case OSTC_F_SV: arr = &blk.sv; break;
    default: return case_fail("field not held by a case (X, Y, Z, JAC, METRIC, BV, SV)");
  }
  const long plane = blk.plane(c->ngg);
  if (n_planes <= 0 || (size_t)n_planes * plane != arr->size()) return case_fail("plane count does not match");
  if (write) {
    if (field != OSTC_F_BV && field != OSTC_F_SV) return case_fail("only BV and SV can be overwritten");
    std::memcpy(arr->data(), host, arr->size() * sizeof(double));
  } else {
    std::memcpy(host, arr->data(), arr->size() * sizeof(double));
  }
  return OSTC_OK;
}

int ostc_case_wall_points(const ostc_case *c, double *xyz, int max_points) {
  if (!c || (!xyz && max_points > 0)) return case_fail("null argument");
  const int n = (int)(c->wall_xyz.size() / 3);
  for (int q = 0; q < 3 * std::min(n, max_points); ++q) xyz[q] = c->wall_xyz[q];
  return OSTC_OK;
}

const char *ostc_case_error(void) { return g_case_err.c_str(); }

}  // extern "C"
