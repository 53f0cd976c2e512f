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
    const double cell_steps = static_cast<double>(width) * height * (nstep + 1);
    std::cout << "m cell_steps_per_s = " << (us > 0 ? cell_steps / (us * 1e-6) : 0.0) << std::endl;
    std::cout << "m nlevel_last_step = " << nlevel << std::endl;
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
