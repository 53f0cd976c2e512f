// fs_host.cpp -- host-side pieces of the FastScape library: the seeded
// terrain generator, the exponent table for the erosion factor, the DEM
// writer and the status strings.
#include "fs_internal.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace fs {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

int validate_params(int W, int H, const fs_params *p) {
  if (!p) return FS_ERR_INVALID_ARGUMENT;
  if (W < 5 || H < 5 || static_cast<long long>(W) * H >= (1ll << 31)) return FS_ERR_DIMENSIONS;
  if (!(p->dx > 0) || !(p->cell_area > 0) || !(p->tol > 0) || !std::isfinite(p->keq) ||
      !std::isfinite(p->neq) || !std::isfinite(p->meq) || !std::isfinite(p->ueq) ||
      !std::isfinite(p->dt))
    return FS_ERR_INVALID_ARGUMENT;
  // accum = count*cell_area must stay an exact integer-valued double
  if (static_cast<double>(W) * H * p->cell_area > 9007199254740992.0) return FS_ERR_INVALID_ARGUMENT;
  return FS_OK;
}

// mt19937 (Matsumoto & Nishimura 1998), the engine behind random.hpp:22.
struct Mt19937 {
  uint32_t mt[624];
  int mti;
  explicit Mt19937(uint32_t seed) {
    mt[0] = seed;
    for (int i = 1; i < 624; i++)
      mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
    mti = 624;
  }
  void twist() {
    for (int k = 0; k < 624 - 397; k++) {
      const uint32_t y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
      mt[k] = mt[k + 397] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    for (int k = 624 - 397; k < 623; k++) {
      const uint32_t y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
      mt[k] = mt[k + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    const uint32_t y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
    mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    mti = 0;
  }
  uint32_t next() {
    if (mti >= 624) twist();
    uint32_t y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
};

// GenerateRandomTerrain (RB:111-130) with uniform_rand_real(0,1)
// (random.cpp:45-49): libstdc++'s generate_canonical<double,53> over a 32-bit
// engine is (x0 + x1*2^32, rounded to double) / 2^64, with 1.0 mapped to the
// largest double below 1; the distribution's *(1-0)+0 is exact.
void generate_terrain(double *h, int W, int H, uint32_t engine_seed) {
  Mt19937 e(engine_seed);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double sum = static_cast<double>(e.next());
      sum += static_cast<double>(e.next()) * 4294967296.0;
      double r = sum / 18446744073709551616.0;
      if (r >= 1.0) r = std::nextafter(1.0, 0.0);
      const bool ring = x <= 1 || y <= 1 || x >= W - 2 || y >= H - 2;
      h[static_cast<size_t>(y) * W + x] = ring ? 0.0 : r;
    }
}

// Table of the erosion factor fact = keq*dt*pow(accum,meq)/pow(dr,neq)
// (RB:315) for every value accum can take: accum is always k*cell_area for an
// upstream cell count k, so tab[2k] (cardinal receiver, dr = dx) and
// tab[2k+1] (diagonal, dr = dx*SQRT2) hold the factor evaluated with the host
// libm and the reference's order of operations, ((keq*dt)*A^m)/dr^n, for
// k = 0..kmax.  Uses at most max_threads host threads; with one, the
// calling thread fills the table and no thread is started.
void fill_fact_table(double *tab, size_t kmax, double cell_area, double meq, double kdt,
                     double lenpow_card, double lenpow_diag, unsigned max_threads) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min({nt, 16u, std::max(1u, max_threads)}));
  const size_t n = kmax + 1;
  if (n < (1u << 16)) nt = 1;
  auto fill = [=](size_t a, size_t b) {
    for (size_t k = a; k < b; k++) {
      const double am = std::pow(static_cast<double>(k) * cell_area, meq);
      tab[2 * k] = (kdt * am) / lenpow_card;
      tab[2 * k + 1] = (kdt * am) / lenpow_diag;
    }
  };
  if (nt == 1) {
    fill(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    pool.emplace_back(fill, a, b);
  }
  for (auto &th : pool) th.join();
}

// PrintDEM (RB:20-40): header of fixed values, then rows y = 1..H-2 of the
// cells x = 1..W-2, each value in the default ostream format of a double
// (printf "%.6g") followed by a space, each row followed by a newline.
int write_dem(const char *filename, const double *h, int W, int H) {
  FILE *f = std::fopen(filename, "wb");
  if (!f) {
    set_last_error(std::string("cannot open ") + filename + ": " + std::strerror(errno));
    return FS_ERR_IO;
  }
  std::vector<char> buf;
  buf.reserve(static_cast<size_t>(W) * 16 + 64);
  char tmp[64];
  std::fprintf(f, "ncols %d\nnrows %d\n", W - 2, H - 2);
  std::fputs("xllcorner 637500.000\nyllcorner 206000.000\ncellsize 500.000\n"
             "NODATA_value -9999\n", f);
  bool ok = true;
  for (int y = 1; y < H - 1 && ok; y++) {
    buf.clear();
    const double *row = h + static_cast<size_t>(y) * W;
    for (int x = 1; x < W - 1; x++) {
      const int len = std::snprintf(tmp, sizeof(tmp), "%.6g ", row[x]);
      buf.insert(buf.end(), tmp, tmp + len);
    }
    buf.push_back('\n');
    ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) {
    set_last_error(std::string("write failed: ") + filename);
    return FS_ERR_IO;
  }
  return FS_OK;
}

const char *status_string(int s) {
  switch (s) {
    case FS_OK: return "success";
    case FS_ERR_INVALID_ARGUMENT: return "invalid argument";
    case FS_ERR_DIMENSIONS: return "grid dimensions out of range (need 5 <= W,H and W*H < 2^31)";
    case FS_ERR_OUT_OF_MEMORY: return "out of memory";
    case FS_ERR_HIP: return "HIP runtime error";
    case FS_ERR_LEVEL_OVERFLOW: return "drainage deeper than FS_MAX_LEVELS";
    case FS_ERR_NEWTON: return "Newton iteration did not converge";
    case FS_ERR_NO_DEVICE: return "no usable HIP device";
    case FS_ERR_BUFFER_SIZE: return "buffer size does not match the grid";
    case FS_ERR_INTERNAL: return "internal consistency check failed";
    case FS_ERR_IO: return "I/O error";
    default: return "unknown status";
  }
}

}  // namespace fs

extern "C" int fs_write_dem(const char *filename, const double *h, int width, int height) {
  if (!filename || !h) return FS_ERR_INVALID_ARGUMENT;
  if (width < 5 || height < 5) return FS_ERR_DIMENSIONS;
  return fs::write_dem(filename, h, width, height);
}

extern "C" int fs_generate_terrain(double *h, int width, int height, uint32_t engine_seed) {
  if (!h) return FS_ERR_INVALID_ARGUMENT;
  if (width < 5 || height < 5 ||
      static_cast<long long>(width) * height >= (1ll << 31))
    return FS_ERR_DIMENSIONS;
  fs::generate_terrain(h, width, height, engine_seed);
  return FS_OK;
}

extern "C" const char *fs_strerror(int status) { return fs::status_string(status); }
extern "C" const char *fs_last_error(void) { return fs::last_error(); }
