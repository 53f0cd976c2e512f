// fs_receiver.h -- the D8 receiver choice of ComputeReceivers
// (fastscape_RB.cpp:160-182) with two IEEE divisions per cell instead of
// eight.  Compiled into the device kernels and, through hipcc's host pass,
// into the CPU equivalence test (tests/test_receiver_rule.py).
#ifndef FS_RECEIVER_H
#define FS_RECEIVER_H

#include <hip/hip_runtime.h>

namespace fs {

constexpr int kNoReceiver = 0xFF;

// The reference rule, operation for operation: directions n = 0..7 (W, NW,
// N, NE, E, SE, S, SW) in order, slope = (hc - h[n]) / dr[n], strict '>'
// from 0, so ties keep the lowest direction and no positive slope means no
// receiver.  dr is dc for even n and dd for odd n.
__host__ __device__ inline int receiver_reference(double hc, const double nb[8], double dc,
                                                  double dd) {
  double ms = 0.0;
  int mn = kNoReceiver;
  for (int n = 0; n < 8; n++) {
    const double s = (hc - nb[n]) / ((n & 1) ? dd : dc);
    if (s > ms) {
      ms = s;
      mn = n;
    }
  }
  return mn;
}

// The best direction of one distance group (even n, divisor dc; or odd n,
// divisor dd): the lowest n whose rounded quotient equals the group's
// largest one.  Division by a fixed positive dr is monotone under correct
// rounding, so that largest quotient is q = RN(A/dr) for the group's
// largest drop A, and a smaller drop x can round to the same quotient only
// when A - x <= dr*ulp(q) <= 2^-52*A; drops within 2^-50*A of A are divided
// exactly, everything else is ruled out without a division.  Tiny A (where
// q could be subnormal) and an overflowing q divide every drop.  *q_out = 0 when the group has
// no positive slope.  NaN drops never compare greater, as in the reference.
__host__ __device__ inline int group_best(const double d[8], int first, double dr, double *q_out) {
  double A = -__builtin_huge_val();
  for (int n = first; n < 8; n += 2)
    if (d[n] > A) A = d[n];
  *q_out = 0.0;
  if (!(A > 0.0)) return kNoReceiver;
  const double q = A / dr;
  if (!(q > 0.0)) return kNoReceiver;
  *q_out = q;
  // subnormal or overflowing quotients: the relative window does not hold
  const bool exact_all = A < 0x1p-900 || !(q < __builtin_huge_val());
  const double window = A - A * 0x1p-50;
  for (int n = first; n < 8; n += 2) {
    if (d[n] == A) return n;
    if ((exact_all || d[n] >= window) && d[n] / dr == q) return n;
  }
  return kNoReceiver;   // unreachable: the maximum itself matches
}

__host__ __device__ inline int receiver_fast(double hc, const double nb[8], double dc, double dd) {
  double d[8];
  for (int n = 0; n < 8; n++) d[n] = hc - nb[n];
  double qc, qd;
  const int nc = group_best(d, 0, dc, &qc);
  const int nd = group_best(d, 1, dd, &qd);
  if (nc == kNoReceiver && nd == kNoReceiver) return kNoReceiver;
  if (nd == kNoReceiver || qc > qd) return nc;
  if (nc == kNoReceiver || qd > qc) return nd;
  return nc < nd ? nc : nd;
}

}  // namespace fs

#endif
