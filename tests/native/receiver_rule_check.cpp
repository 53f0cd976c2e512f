// CPU check of fs::receiver_fast (division-free decision) and
// fs::receiver_two_div (its fallback) against fs::receiver_reference (the
// reference's eight-division loop, fastscape_RB.cpp:160-182) on random,
// tied, near-tied and special inputs.  Built by tests/test_receiver_rule.py
// with hipcc's host pass.  Prints "cases mismatches".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "fs_receiver.h"

static uint64_t s = 0x9E3779B97F4A7C15ull;
static inline uint64_t rnd() {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return s;
}
static inline double u01() { return static_cast<double>(rnd() >> 11) * 0x1p-53; }
static inline double step_ulps(double x, int k) {
  for (; k > 0; k--) x = std::nextafter(x, std::numeric_limits<double>::infinity());
  for (; k < 0; k++) x = std::nextafter(x, -std::numeric_limits<double>::infinity());
  return x;
}

int main(int argc, char **argv) {
  const long n_cases = argc > 1 ? std::atol(argv[1]) : 20000000;
  const double SQRT2 = 1.414213562373095048801688724209698078569671875376948;
  // the last four lie outside the division-free decision's divisor range;
  // 2^123 and 2^200 make quotients of drops >= 2^-900 subnormal
  const double dcs[8] = {1.0, 100.0, 30.0, 0.001, 1e-30, 1e30, 0x1.3p123, 0x1p200};
  const double specials[10] = {0.0, -0.0, std::numeric_limits<double>::infinity(),
                               -std::numeric_limits<double>::infinity(),
                               std::numeric_limits<double>::quiet_NaN(), 4.9e-324, 1e-310,
                               1e308, -1e308, 1.0};
  long bad = 0;
  for (long t = 0; t < n_cases; t++) {
    const double dc = dcs[t % 8], dd = dc * SQRT2;
    double hc, nb[8];
    const int kind = static_cast<int>(rnd() % 7);
    if (kind == 0) {                         // uniform
      hc = u01();
      for (double &v : nb) v = u01();
    } else if (kind == 1) {                  // coarse grid: exact ties everywhere
      hc = static_cast<double>(rnd() % 9) / 8;
      for (double &v : nb) v = static_cast<double>(rnd() % 9) / 8;
    } else if (kind == 2) {                  // a few ulps apart
      const double base = u01() * 1000.0;
      hc = base + 1.0 + u01();
      for (double &v : nb) v = step_ulps(base, static_cast<int>(rnd() % 9) - 4);
    } else if (kind == 3) {                  // cardinal vs diagonal near-ties
      hc = u01() * 10.0 + 5.0;
      const double drop = u01() * 3.0 + 0.1;
      for (int n = 0; n < 8; n++) {
        const double target = (n & 1) ? drop * SQRT2 : drop;
        nb[n] = step_ulps(hc - target, static_cast<int>(rnd() % 5) - 2);
      }
    } else if (kind == 4) {                  // special values mixed in
      hc = (rnd() % 3 == 0) ? specials[rnd() % 10] : u01();
      for (double &v : nb) v = (rnd() % 3 == 0) ? specials[rnd() % 10] : u01();
    } else if (kind == 5) {                  // wide dynamic range
      const int e1 = static_cast<int>(rnd() % 2000) - 1000;
      hc = std::ldexp(u01() + 0.5, e1);
      for (double &v : nb) v = std::ldexp(u01() + 0.5, e1 - static_cast<int>(rnd() % 60));
    } else {                                 // tiny near-tied drops (subnormal quotients)
      const double drop = std::ldexp(u01() + 0.5, -860 - static_cast<int>(rnd() % 120));
      hc = 0.0;
      for (int n = 0; n < 8; n++) {
        const double target = (n & 1) ? drop * SQRT2 : drop;
        nb[n] = -step_ulps(target, static_cast<int>(rnd() % 9) - 4);
      }
    }
    const int a = fs::receiver_reference(hc, nb, dc, dd);
    const int b = fs::receiver_fast(hc, nb, dc, dd);
    const int c = fs::receiver_two_div(hc, nb, dc, dd);
    if (a != b || a != c) {
      if (bad < 10) {
        std::printf("mismatch kind %d dc %g: ref %d fast %d two-div %d hc %a nb", kind, dc, a, b, c,
                    hc);
        for (double v : nb) std::printf(" %a", v);
        std::printf("\n");
      }
      bad++;
    }
  }
  std::printf("%ld %ld\n", n_cases, bad);
  return bad ? 1 : 0;
}
