// fastscape_cli.cpp -- drop-in command line for the reference executables
// (main of fastscape_RB.cpp:414-449):
//
//   fastscape_mi355x <Dimension> <Steps> <Output Name> <Seed>
//
// Same arguments, same RichDEM-prefixed stdout (A/C/h/m/p/t lines,
// RB:431-444, parsed by tests/d.make_tabular), same ArcGIS ASCII output.
// Options may precede the positional arguments:
//   --preset=reference|benchmark   model constants (default: reference)
//   --device=N                     HIP device ordinal (default: 0)
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>
#include <chrono>

#include "fastscape.hpp"

#ifndef GIT_HASH
#define GIT_HASH "unknown"
#endif

// Algorithmic HBM bytes of one cell-step: receivers 9 (h 8 + direction 1),
// donors 2, order 10, accumulation 21, uplift + erosion 37; and MI355X's
// HBM3E peak.
constexpr double kBytesPerCellStep = 79.0;
constexpr double kHbmPeakBytesPerS = 8.0e12;

int main(int argc, char **argv) {
  int preset = FS_PRESET_REFERENCE, device = 0;
  std::vector<char *> args{argv[0]};
  for (int i = 1; i < argc; i++) {
    if (std::strncmp(argv[i], "--preset=", 9) == 0) {
      const std::string v = argv[i] + 9;
      if (v == "reference") preset = FS_PRESET_REFERENCE;
      else if (v == "benchmark") preset = FS_PRESET_BENCHMARK;
      else { std::cerr << "unknown preset " << v << std::endl; return -1; }
    } else if (std::strncmp(argv[i], "--device=", 9) == 0) {
      device = std::stoi(argv[i] + 9);
    } else {
      args.push_back(argv[i]);
    }
  }
  if (args.size() != 5) {
    std::cerr << "Syntax: " << argv[0] << " <Dimension> <Steps> <Output Name> <Seed>" << std::endl;
    return -1;
  }
  try {
    const int width = std::stoi(args[1]);
    const int height = std::stoi(args[1]);
    const int nstep = std::stoi(args[2]);
    const std::string output_name = args[3];
    const auto rand_seed = std::stoul(args[4]);

    fastscape::seed_rand(rand_seed);
    std::cout << "A FastScape RB MI355X" << std::endl;
    std::cout << "C Richard Barnes TODO" << std::endl;
    std::cout << "h git_hash    = " << GIT_HASH << std::endl;
    std::cout << "m Random seed = " << rand_seed << std::endl;
    std::cout << "m Preset = " << (preset == FS_PRESET_REFERENCE ? "reference" : "benchmark") << std::endl;

    const auto t0 = std::chrono::steady_clock::now();
    fastscape::FastScape tm(width, height, fastscape::preset(preset), device);
    tm.run(nstep);
    const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                        std::chrono::steady_clock::now() - t0).count();
    std::cout << "t Total calculation time    = " << std::setw(15) << us << " microseconds" << std::endl;
    // metadata lines (tag m: a column for d.make_tabular) -- no reference
    // counterpart: throughput, drainage depth and which step path ran
    uint64_t tile_steps = 0, level_steps = 0, halo16_steps = 0, halo32_steps = 0;
    int32_t nlevel = 0;
    fastscape::check(fs_path_stats(tm.handle(), &tile_steps, &level_steps), "fs_path_stats");
    fastscape::check(fs_tile_stats(tm.handle(), &halo16_steps, &halo32_steps), "fs_tile_stats");
    fastscape::check(fs_get_levels(tm.handle(), nullptr, 0, &nlevel), "fs_get_levels");
    uint64_t t[FS_NUM_TIMERS];
    fastscape::check(fs_get_timers(tm.handle(), t), "fs_get_timers");
    const double cell_steps = static_cast<double>(width) * height * (nstep + 1);
    const double run_us = static_cast<double>(t[FS_T_OVERALL] - t[FS_T_INITIALIZE]);
    const double step_rate = run_us > 0 ? cell_steps / (run_us * 1e-6) : 0.0;
    std::cout << "m cell_steps_per_s = " << (us > 0 ? cell_steps / (us * 1e-6) : 0.0) << std::endl;
    // the time-steps alone (Overall minus Initialize), against the HBM
    // roofline of the time-step's algorithmic bytes per cell-step
    std::cout << "m step_cell_steps_per_s = " << step_rate << std::endl;
    std::cout << "m roofline_bytes_per_cell_step = " << kBytesPerCellStep << std::endl;
    std::cout << "m hbm_roofline_frac = " << step_rate * kBytesPerCellStep / kHbmPeakBytesPerS
              << std::endl;
    std::cout << "m nlevel_last_step = " << nlevel << std::endl;
    std::cout << "m deepest_level_last_step = " << (nlevel >= 3 ? nlevel - 3 : 0) << std::endl;
    std::cout << "m tile_path_steps = " << tile_steps << std::endl;
    std::cout << "m level_path_steps = " << level_steps << std::endl;
    std::cout << "m tile_halo32_steps = " << halo32_steps << std::endl;
    fastscape::PrintDEM(output_name, tm.getH(), width, height);
  } catch (const fastscape::Error &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return e.status();
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return -1;
  }
  return 0;
}
