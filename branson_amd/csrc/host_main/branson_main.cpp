// branson_main.cpp -- command-line driver: runs a Branson XML deck through
// the MI355X engine and prints the reference's per-step diagnostics
// (src/main.cc:39-154, src/imc_state.h:264-284).
//
//   branson_hip <input.xml> [--fp32] [--history] [--groups N]
//               [--photons N] [--device D] [--device-physics]
//
// Photons are tracked by persistent waves unless --history asks for one
// work-item per history (the same events, about half the throughput);
// --persistent is accepted and changes nothing.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "branson_host.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("Usage: branson_hip <path_to_input_file> [--fp32] [--history] "
                "[--groups N] [--photons N] [--device D] [--device-physics]\n");
    return EXIT_FAILURE;
  }
  bh_driver_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.engine.algorithm = BH_ALGO_PERSISTENT;
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--fp32")) cfg.engine.precision = BH_PRECISION_FP32_TRACK;
    else if (!std::strcmp(argv[i], "--persistent")) cfg.engine.algorithm = BH_ALGO_PERSISTENT;
    else if (!std::strcmp(argv[i], "--history")) cfg.engine.algorithm = BH_ALGO_HISTORY;
    else if (!std::strcmp(argv[i], "--groups") && i + 1 < argc) cfg.n_groups = (uint32_t)std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--photons") && i + 1 < argc) cfg.n_photons = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) cfg.device = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--device-physics")) cfg.device_physics = 1;
    else {
      std::fprintf(stderr, "unknown argument %s\n", argv[i]);
      return EXIT_FAILURE;
    }
  }
  bh_driver* d = nullptr;
  if (bh_driver_create(argv[1], &cfg, &d) != BH_OK) {
    std::fprintf(stderr, "%s\n", bh_driver_last_error(d));
    bh_driver_destroy(d);
    return EXIT_FAILURE;
  }
  std::printf("Branson replicated IMC on MI355X -- cells: %u, photons/step: %llu, seed: %u\n",
              bh_driver_n_cells(d), (unsigned long long)bh_driver_n_photons(d), bh_driver_seed(d));
  double total_transport = 0.0;
  unsigned long long total_hist = 0, total_segs = 0;
  while (!bh_driver_finished(d)) {
    bh_driver_report r;
    if (bh_driver_step(d, &r) != BH_OK) {
      std::fprintf(stderr, "%s\n", bh_driver_last_error(d));
      bh_driver_destroy(d);
      return EXIT_FAILURE;
    }
    std::printf("****************************************"
                "****************************************\n");
    std::printf("Step: %u  Start Time: %g  End Time: %g  dt: %g\n", r.step, r.time, r.time + r.dt,
                r.dt);
    std::printf("Total Photons transported: %llu\n", (unsigned long long)r.transported);
    std::printf("Emission E: %g, Source E: %g, Absorption E: %g, Exit E: %g\n", r.emission_E,
                r.source_E, r.absorbed_E, r.exit_E);
    std::printf("Pre census E: %g Post census E: %g Post census Size: %llu\n", r.pre_census_E,
                r.post_census_E, (unsigned long long)r.census_size);
    std::printf("Pre mat E: %g Post mat E: %g\n", r.pre_mat_E, r.post_mat_E);
    std::printf("Radiation conservation: %g\n", r.rad_conservation);
    std::printf("Material conservation: %g\n", r.mat_conservation);
    const double dev_s = 1e-3 * (r.source_ms + r.transport_ms + r.census_ms);
    std::printf("Transport time (device source+transport+census): %g s, segments: %llu\n", dev_s,
                (unsigned long long)r.n_segments);
    total_transport += dev_s;
    total_hist += r.transported;
    total_segs += r.n_segments;
  }
  std::printf("****************************************"
              "****************************************\n");
  std::printf("Total transport: %g\n", total_transport);
  std::printf("Histories per second: %g, segments per second: %g\n",
              total_transport > 0 ? total_hist / total_transport : 0.0,
              total_transport > 0 ? total_segs / total_transport : 0.0);
  bh_driver_destroy(d);
  return EXIT_SUCCESS;
}
