// fastscape.hpp -- C++ drop-in for the reference's FastScape class surface
// (class FastScape_RB, fastscape_RB.cpp:46-406), backed by the MI355X
// library through the C ABI in fastscape.h.
//
//   fastscape::seed_rand(seed);                 // random.cpp:19-35
//   fastscape::FastScape tm(width, height);     // RB:135-151
//   tm.run(nstep);                              // RB:341-398
//   fastscape::PrintDEM(name, tm.getH(), width, height);   // RB:20-40
//
// Differences from the reference, all deliberate:
//   * constants are a parameter block (fs_params) instead of const members;
//     the default is the reference's own preset (RB:58-65);
//   * errors throw fastscape::Error carrying an FS_ERR_* status instead of
//     failing an assert (RB:253);
//   * getH() refreshes a host copy from the device and returns it.
#ifndef FASTSCAPE_HPP
#define FASTSCAPE_HPP

#include <cstdint>
#include <iomanip>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "fastscape.h"

namespace fastscape {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &where)
      : std::runtime_error(where + ": " + fs_strerror(status) +
                           (fs_last_error()[0] ? std::string(" (") + fs_last_error() + ")" : "")),
        status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char *where) {
  if (status != FS_OK) throw Error(status, where);
}

inline fs_params preset(int which) {
  fs_params p;
  check(fs_params_preset(which, &p), "fs_params_preset");
  return p;
}

// seed_rand (random.cpp:19-35) seeds thread t's engine with seed*t and the
// terrain is drawn on thread 0, so every nonzero seed yields the terrain of
// mt19937 seeded with 0; seed 0 draws the engine seed from random_device.
inline uint32_t &engine_seed_slot() {
  static uint32_t s = 0;
  return s;
}
inline void seed_rand(unsigned long seed) {
  engine_seed_slot() = seed == 0 ? static_cast<uint32_t>(std::random_device{}()) : 0u;
}

class FastScape {
 public:
  FastScape(int width, int height, const fs_params &params = preset(FS_PRESET_REFERENCE),
            int device = 0)
      : width_(width), height_(height) {
    check(fs_create(width, height, &params, engine_seed_slot(), device, &m_), "fs_create");
  }
  FastScape(const FastScape &) = delete;
  FastScape &operator=(const FastScape &) = delete;
  ~FastScape() {
    if (m_) fs_destroy(m_);
  }

  // run(nstep) (RB:341-398): nstep+1 time-steps, "p Step" progress every 20
  // steps (RB:374-375) and the timer report (RB:380-387) on stdout.
  void run(const int nstep, std::ostream &out = std::cout) {
    if (nstep < 0) check(FS_ERR_INVALID_ARGUMENT, "run");
    int s = 0;   // index of the next step to run
    while (s <= nstep) {
      const int next = ((s + 19) / 20) * 20;         // next step that prints
      const int end = next < nstep ? next : nstep;   // run steps s..end
      check(fs_step(m_, end - s + 1), "fs_step");
      if (end == next) out << "p Step = " << end << std::endl;
      s = end + 1;
    }
    uint64_t t[FS_NUM_TIMERS];
    check(fs_get_timers(m_, t), "fs_get_timers");
    out << "t Step1: Initialize         = " << std::setw(15) << t[FS_T_INITIALIZE] << " microseconds" << std::endl;
    out << "t Step2: DetermineReceivers = " << std::setw(15) << t[FS_T_RECEIVERS] << " microseconds" << std::endl;
    out << "t Step3: DetermineDonors    = " << std::setw(15) << t[FS_T_DONORS] << " microseconds" << std::endl;
    out << "t Step4: GenerateOrder      = " << std::setw(15) << t[FS_T_ORDER] << " microseconds" << std::endl;
    out << "t Step5: FlowAcc            = " << std::setw(15) << t[FS_T_FLOWACC] << " microseconds" << std::endl;
    out << "t Step6: Uplift             = " << std::setw(15) << t[FS_T_UPLIFT] << " microseconds" << std::endl;
    out << "t Step7: Erosion            = " << std::setw(15) << t[FS_T_EROSION] << " microseconds" << std::endl;
    out << "t Overall                   = " << std::setw(15) << t[FS_T_OVERALL] << " microseconds" << std::endl;
  }

  // getH() (RB:403-405): the elevation field, copied from the device.
  std::vector<double> &getH() {
    h_.resize(static_cast<size_t>(width_) * height_);
    check(fs_get_h(m_, h_.data(), h_.size()), "fs_get_h");
    return h_;
  }
  void setH(const std::vector<double> &h) { check(fs_set_h(m_, h.data(), h.size()), "fs_set_h"); }

  int width() const { return width_; }
  int height() const { return height_; }
  fs_model *handle() { return m_; }

 private:
  int width_, height_;
  fs_model *m_ = nullptr;
  std::vector<double> h_;
};

// PrintDEM (RB:20-40).
inline void PrintDEM(const std::string &filename, const std::vector<double> &h, int width,
                     int height) {
  if (h.size() != static_cast<size_t>(width) * height) check(FS_ERR_BUFFER_SIZE, "PrintDEM");
  check(fs_write_dem(filename.c_str(), h.data(), width, height), "PrintDEM");
}

}  // namespace fastscape

#endif
