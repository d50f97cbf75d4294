// Host-side case description shared by the case reader (host_case.cpp) and
// the time-loop driver (host_driver.cpp): the setup-file dictionary, the
// blocks of the mesh with their geometry, boundary programs and initial
// fields, and the model constants.
#pragma once

#include <array>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ostc.h"

namespace ostc_host {

// One value inside a setup-file struct: "int", "real" or "string".
struct SetupValue {
  enum Kind { INT, REAL, STRING } kind = INT;
  long long i = 0;
  double r = 0;
  std::string s;
  double number() const { return kind == REAL ? r : (double)i; }
};
using SetupStruct = std::map<std::string, SetupValue>;

// The key/value dictionary of the reference's setup files
// (Parameter::read_one_file, src/Parameter.cpp:59-168).  The first
// definition of a name wins, as with the reference's unordered_map::emplace.
struct Setup {
  std::unordered_map<std::string, int> ints;
  std::unordered_map<std::string, double> reals;
  std::unordered_map<std::string, bool> bools;
  std::unordered_map<std::string, std::string> strings;
  std::unordered_map<std::string, std::vector<int>> int_arrays;
  std::unordered_map<std::string, std::vector<double>> real_arrays;
  std::unordered_map<std::string, std::vector<std::string>> string_arrays;
  std::unordered_map<std::string, SetupStruct> structs;

  // false (and *err set) if the file cannot be opened or a value is malformed
  bool read_file(const std::string &path, std::string *err);
};

// Inflow state of one setup struct (Inflow::Inflow, src/BoundCond.cpp:6-106).
struct InflowState {
  std::string name;
  int label = 5;
  double mach = -1, pressure = 101325, temperature = -1, velocity = 0, density = -1;
  double u = 1, v = 0, w = 0, mw = 0, viscosity = 0, mut = 0;
  std::vector<double> sv;
};

struct CaseBlock {
  int mx = 0, my = 0, mz = 0;
  std::vector<double> x, y, z;                  // interior points, i fastest
  std::vector<std::array<int, 7>> boundary;     // boundary-file records
  std::vector<std::array<int, 13>> inner_raw;   // inner-file records
  std::vector<ostc_inner_face> inner;           // registered inner faces
  std::vector<double> xg, yg, zg, jac, metric;  // ngg-padded geometry
  std::vector<double> bv, sv;                   // ngg-padded initial fields
  ostc_bc bc{};
  int n_neg_jac = 0;
  long plane(int ngg) const { return (long)(mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg); }
};

// one ORDERED zone of a Tecplot binary flow-field file: vars[v][i + mx*(j + my*k)]
struct PltZone {
  int mx = 0, my = 0, mz = 0;
  std::vector<std::vector<double>> vars;
};

}  // namespace ostc_host

struct ostc_case;

namespace ostc_host {
// the reference's flow-field file (host_output.cpp); 0 = ok, else *err says why
int write_plt(const std::string &path, const std::vector<std::string> &names, const std::vector<PltZone> &zones,
              std::string *err);
int read_plt(const std::string &path, std::vector<std::string> *names, std::vector<PltZone> *zones,
             std::string *err);
// setup `initial = 1`: overwrite the initial bv/sv from <dir>/output/field/flowfield.plt
int restart_fields(ostc_case &c, std::string *err);
}  // namespace ostc_host

struct ostc_case {
  std::string dir;  // the directory that holds input_files/
  ostc_host::Setup setup;
  ostc_param param{};
  ostc_mechanism_info mech{};
  int ngg = 2, dimension = 3;
  int total_step = 0, output_screen = 1, output_file = 1, initial = 0, steady = 1;
  int start_step = 0;                            // restart: step of the flow-field file
  double res_scale[4] = {1, 1, 1, 1};            // restart: residual scales of that run
  double convergence = 0, grid_scale = 1;
  double rho_inf = 0, v_inf = 0;                 // the freestream struct's density and speed
  std::vector<int> post_process;                // setup array post_process (procedure labels)
  std::vector<ostc_host::CaseBlock> blocks;
  std::vector<ostc_host::InflowState> inflows;  // BC slots, in slot order
  std::vector<double> wall_xyz;                 // wall points for the wall distance
};
