// fs_receiver.h -- the D8 receiver choice of ComputeReceivers
// (fastscape_RB.cpp:160-182), usually without any division and never with
// more than the two-division fallback plus near-tie checks.  Compiled into
// the device kernels and, through hipcc's host pass, into the CPU
// equivalence test (tests/test_receiver_rule.py).
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
// exactly, everything else is ruled out without a division.  The window
// needs q to be a normal number and A*2^-50 not to underflow: a subnormal or
// overflowing q, or a tiny A, divides every drop.  *q_out = 0 when the group
// has no positive slope.  NaN drops never compare greater, as in the
// reference.
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
  const bool exact_all =
      A < 0x1p-900 || !(q >= 0x1p-1022) || !(q < __builtin_huge_val());
  const double window = A - A * 0x1p-50;
  for (int n = first; n < 8; n += 2) {
    if (d[n] == A) return n;
    if ((exact_all || d[n] >= window) && d[n] / dr == q) return n;
  }
  return kNoReceiver;   // unreachable: the maximum itself matches
}

// Two IEEE divisions per cell (one per distance group) plus the near-tie
// checks of group_best; correct for every input, including NaN, infinities,
// subnormal and overflowing quotients.
__host__ __device__ inline int receiver_two_div(double hc, const double nb[8], double dc,
                                                double dd) {
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

// The group's best direction when it can be told without dividing: the
// first drop within 2^-50*A of the group's largest drop A is the best
// direction if it is A itself (earlier drops are too far below A to round to
// the same quotient).  -1 when that first drop is a smaller near-tie, which
// only a division can settle.
__host__ __device__ inline int group_first(const double d[8], int first, double A) {
  const double window = A - A * 0x1p-50;
  for (int n = first; n < 8; n += 2)
    if (d[n] >= window) return d[n] == A ? n : -1;
  return -1;   // unreachable: A itself is in the window
}

// The reference's direction, decided without division whenever every
// quotient involved is a normal number and the two groups' slopes are not
// near-tied.  For drops Ac, Ad and divisors dc, dd in range, P = RN(Ac*dd)
// and Q = RN(Ad*dc) carry relative errors of at most 2^-53 each, so
// P > RN(Q*(1 + 2^-48)) gives Ac/dc > (Ad/dd)*(1 + 2^-49), hence
// RN(Ac/dc) > RN(Ad/dd) strictly: the cardinal group's quotient wins without
// a tie.  Anything closer (or out of range) goes to receiver_two_div.
__host__ __device__ inline int receiver_fast(double hc, const double nb[8], double dc, double dd) {
  double d[8];
  for (int n = 0; n < 8; n++) d[n] = hc - nb[n];
  // Group maxima.  fmax ignores NaN drops, as the reference's strict '>'
  // does; it may return +0 where '>' keeps -0, but a group only counts when
  // its maximum is positive.
  const double Ac = fmax(fmax(d[0], d[2]), fmax(d[4], d[6]));
  const double Ad = fmax(fmax(d[1], d[3]), fmax(d[5], d[7]));
  const bool pc = Ac > 0.0, pd = Ad > 0.0;
  if (!pc && !pd) return kNoReceiver;
  // quotients of drops in [2^-900, 2^900] by divisors in [2^-60, 2^60] are
  // normal and nonzero: a positive drop is a positive slope
  const bool okc = !pc || (Ac >= 0x1p-900 && Ac <= 0x1p900);
  const bool okd = !pd || (Ad >= 0x1p-900 && Ad <= 0x1p900);
  const bool okr = dc >= 0x1p-60 && dc <= 0x1p60 && dd >= 0x1p-60 && dd <= 0x1p60;
  if (okc && okd && okr) {
    const int bc = pc ? group_first(d, 0, Ac) : kNoReceiver;
    const int bd = pd ? group_first(d, 1, Ad) : kNoReceiver;
    if (bc >= 0 && bd >= 0) {
      if (!pd) return bc;
      if (!pc) return bd;
      const double P = Ac * dd, Q = Ad * dc;
      if (P > Q * (1.0 + 0x1p-48)) return bc;
      if (Q > P * (1.0 + 0x1p-48)) return bd;
    }
  }
  return receiver_two_div(hc, nb, dc, dd);
}

}  // namespace fs

#endif
