// Command-line runner: the reference's main (src/main.cu:7-67) over the
// C-ABI.  Reads the case directory (default: the working directory, which
// holds input_files/ as for the reference), builds the driver, initialises
// the flow field and runs the steady time loop with the reference's screen
// and history output.
//   ostc_run [case_dir] [--fp32] [--steps N] [--quiet]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ostc.h"

int main(int argc, char **argv) {
  std::string dir = ".";
  int precision = OSTC_FP64, io = 1, steps = -1;
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--fp32")) {
      precision = OSTC_FP32;
    } else if (!std::strcmp(argv[a], "--quiet")) {
      io = 0;
    } else if (!std::strcmp(argv[a], "--steps") && a + 1 < argc) {
      steps = std::atoi(argv[++a]);
    } else if (argv[a][0] == '-') {
      std::fprintf(stderr, "usage: %s [case_dir] [--fp32] [--steps N] [--quiet]\n", argv[0]);
      return 2;
    } else {
      dir = argv[a];
    }
  }
  ostc_case *c = nullptr;
  if (ostc_case_read(dir.c_str(), &c) != OSTC_OK) {
    std::fprintf(stderr, "case %s: %s\n", dir.c_str(), ostc_case_error());
    return 1;
  }
  ostc_case_info info;
  ostc_case_get_info(c, &info);
  long cells = 0;
  for (int b = 0; b < info.n_block; ++b) cells += (long)info.mx[b] * info.my[b] * info.mz[b];
  std::printf("%s\nBlocks: %d, total grid number: %ld, precision: %s\n", ostc_build_info(), info.n_block, cells,
              precision == OSTC_FP32 ? "FP32 bulk + FP64 chemistry/Newton" : "FP64");
  ostc_driver *d = nullptr;
  if (ostc_driver_create(c, precision, &d) != OSTC_OK) {
    std::fprintf(stderr, "driver: %s\n", ostc_driver_error());
    ostc_case_destroy(c);
    return 1;
  }
  ostc_case_destroy(c);
  if (ostc_driver_initialize(d) != OSTC_OK) {
    std::fprintf(stderr, "initialize: %s\n", ostc_driver_error());
    ostc_driver_destroy(d);
    return 1;
  }
  std::printf("The flowfield is completely initialized on the GPU.\n");
  int done = 0, st = OSTC_OK;
  double err = 0;
  const auto t0 = std::chrono::steady_clock::now();
  if (steps > 0) {
    for (int s = 0; s < steps && st == OSTC_OK; ++s, ++done) st = ostc_driver_step(d, nullptr, nullptr, &err);
    if (st == OSTC_OK) st = ostc_device_sync();
  } else {
    st = ostc_driver_simulate(d, io, &done, &err);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st != OSTC_OK) {
    std::fprintf(stderr, "time loop: %s\n", ostc_driver_error());
    ostc_driver_destroy(d);
    return 1;
  }
  std::printf("%d iterations in %.3f s: %.2f Mcell-iter/s, last residual %.4e\n", done, secs,
              secs > 0 ? (double)cells * done / secs * 1e-6 : 0.0, err);
  ostc_driver_destroy(d);
  return 0;
}
