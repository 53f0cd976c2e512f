// ref_harness.cpp -- drives the reference FastScape model, compiled from its
// own unmodified sources where they lie (see oracle/Makefile), to produce the
// golden fixtures in tests/golden/ and the timed CPU baseline.
//
// TEST INFRASTRUCTURE ONLY.  The reference translation unit named by
// REF_SOURCE is included after every header it uses, with three
// preprocessor switches that apply to that translation unit only:
//   private -> public   so the harness can call the stage methods one at a
//                       time and read the work arrays;
//   const   -> (empty)  so the model constants (RB:58-65), which the
//                       reference hard-codes, can be set to the benchmark
//                       preset after construction;
//   main    -> reference_main   so the harness supplies its own entry point.
// No line of the reference is copied.
//
// Usage:
//   ref_harness dump <W> <H> <nsteps> <seed> <preset> <outdir>
//       per-stage arrays of the first <nsteps> time-steps as raw
//       little-endian binaries (int32 / float64) under <outdir>
//   ref_harness run  <W> <H> <nstep> <seed> <preset> <dem_out>
//       construct + run(nstep) (nstep+1 time-steps, RB:366) + PrintDEM,
//       printing the reference's own stdout and "m run_seconds = <s>"
// preset 0 = the reference's constants, 1 = the benchmark constants
// (K=1e-5, n=1, m=0.4, U=2e-3, dt=1000, dr={100, 100*SQRT2}, area=1e4).

#include <algorithm>
#include <array>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fenv.h>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "random.hpp"
#include "CumulativeTimer.hpp"

#define private public
#define const
#define main reference_main
#include REF_SOURCE
#undef main
#undef const
#undef private

namespace {

template <class M>
void set_preset(M &tm, int preset) {
  if (preset == 0) return;
  tm.keq = 1e-5;
  tm.neq = 1;
  tm.meq = 0.4;
  tm.ueq = 2e-3;
  tm.dt = 1000.;
  tm.tol = 1e-3;
  tm.cell_area = 1e4;
  for (int i = 0; i < 8; i++) tm.dr[i] = (i & 1) ? 100.0 * tm.SQRT2 : 100.0;
}

template <class T>
void dump(const std::string &dir, const std::string &name, const T *p, size_t n) {
  const std::string path = dir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  if (n && std::fwrite(p, sizeof(T), n, f) != n)
    throw std::runtime_error("short write " + path);
  std::fclose(f);
}

int do_dump(int W, int H, int nsteps, unsigned long seed, int preset,
            const std::string &outdir) {
  seed_rand(seed);
  REF_CLASS tm(W, H);
  set_preset(tm, preset);
  const int size = W * H;
  dump(outdir, "h_init", tm.h.data(), (size_t)size);
  // run() prologue (RB:346-362); the levels buffer is made large enough for
  // any grid (the reference sizes it 2W+2H), which changes no content.
  tm.accum.resize(size);
  tm.rec.resize(size);
  tm.ndon.resize(size);
  tm.donor.resize(8 * (size_t)size);
  tm.stack.resize(size);
  tm.levels.resize(std::max(2 * W + 2 * H, size + 3));
#ifdef REF_HAS_WIDTHS
  tm.stack_width = size;
  tm.level_width = size + 3;
#endif
  for (int i = 0; i < size; i++) tm.rec[i] = tm.NO_FLOW;
  for (int s = 0; s < nsteps; s++) {
    const std::string sfx = "_" + std::to_string(s);
    tm.ComputeReceivers();
    dump(outdir, "rec" + sfx, tm.rec.data(), (size_t)size);
    tm.ComputeDonors();
    dump(outdir, "ndon" + sfx, tm.ndon.data(), (size_t)size);
    dump(outdir, "donor" + sfx, tm.donor.data(), 8 * (size_t)size);
    tm.GenerateOrder();
    dump(outdir, "stack" + sfx, tm.stack.data(), (size_t)size);
    dump(outdir, "levels" + sfx, tm.levels.data(), (size_t)tm.nlevel);
    tm.ComputeFlowAcc();
    dump(outdir, "accum" + sfx, tm.accum.data(), (size_t)size);
    tm.AddUplift();
    dump(outdir, "hup" + sfx, tm.h.data(), (size_t)size);
    tm.Erode();
    dump(outdir, "h" + sfx, tm.h.data(), (size_t)size);
  }
  return 0;
}

int do_run(int W, int H, int nstep, unsigned long seed, int preset,
           const std::string &dem) {
  seed_rand(seed);
  REF_CLASS tm(W, H);
  set_preset(tm, preset);
  const auto t0 = std::chrono::steady_clock::now();
  tm.run(nstep);
  const auto t1 = std::chrono::steady_clock::now();
  std::cout << "m run_seconds = " << std::setprecision(9)
            << std::chrono::duration<double>(t1 - t0).count() << std::endl;
#ifdef _OPENMP
  std::cout << "m omp_threads = " << omp_get_max_threads() << std::endl;
#endif
  if (dem != "-") PrintDEM(dem, tm.getH(), W, H);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 8) {
    std::cerr << "usage: " << argv[0]
              << " dump|run <W> <H> <nsteps> <seed> <preset> <out>" << std::endl;
    return 2;
  }
  const std::string mode = argv[1];
  const int W = std::stoi(argv[2]), H = std::stoi(argv[3]), n = std::stoi(argv[4]);
  const unsigned long seed = std::stoul(argv[5]);
  const int preset = std::stoi(argv[6]);
  if (W < 5 || H < 5 || n < 0 || (preset != 0 && preset != 1)) {
    std::cerr << "bad arguments" << std::endl;
    return 2;
  }
  if (mode == "dump") return do_dump(W, H, n, seed, preset, argv[7]);
  if (mode == "run") return do_run(W, H, n, seed, preset, argv[7]);
  std::cerr << "unknown mode " << mode << std::endl;
  return 2;
}
