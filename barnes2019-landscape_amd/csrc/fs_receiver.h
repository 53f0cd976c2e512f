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

// True when the divisors admit the division-free decision of receiver_fast
// (a property of the model's constants: callers may evaluate it once).
__host__ __device__ inline bool divisors_in_range(double dc, double dd) {
  return dc >= 0x1p-60 && dc <= 0x1p60 && dd >= 0x1p-60 && dd <= 0x1p60;
}

// The reference's direction, decided without division whenever every
// quotient involved is a normal number and no two slopes are near-tied;
// *ok is false when that decision does not hold and the result is to be
// replaced by receiver_two_div's.
//
// Within a group, RN(d/dr) is monotone in d, so the group's steepest slope
// is RN(A/dr) for its largest drop A, and a smaller drop can round to the
// same quotient only when it lies within dr*ulp(q) <= 2^-52*A*(1 + 2^-53) of
// A.  The window bound w = RN(A*(1 - 2^-50)) lies below A - 2^-51*A (the
// product is exact up to half an ulp of A), so drops below w never tie with
// A.  If the lowest-index drop in [w, A] is A itself, that index is the
// group's answer; otherwise a near-tie needs the divisions.  Between the
// groups, for drops Ac, Ad and divisors dc, dd in range, P = RN(Ac*dd) and
// Q = RN(Ad*dc) carry relative errors of at most 2^-53 each, so
// P > RN(Q*(1 + 2^-48)) gives Ac/dc > (Ad/dd)*(1 + 2^-49), hence
// RN(Ac/dc) > RN(Ad/dd) strictly: the cardinal group wins without a tie (and
// symmetrically).  Anything closer, or out of range, is left to
// receiver_two_div.
//
// Written as straight-line selects with no branch, so callers can decide a
// run of cells and fix up the rare undecided ones afterwards.
__host__ __device__ inline int receiver_try(double hc, const double nb[8], double dc, double dd,
                                            bool div_ok, bool *ok) {
  double d[8];
  for (int n = 0; n < 8; n++) d[n] = hc - nb[n];
  // Group maxima.  fmax ignores NaN drops, as the reference's strict '>'
  // does; it may return +0 where '>' keeps -0, but a group only counts when
  // its maximum is positive.
  const double Ac = fmax(fmax(d[0], d[2]), fmax(d[4], d[6]));
  const double Ad = fmax(fmax(d[1], d[3]), fmax(d[5], d[7]));
  const bool pc = Ac > 0.0, pd = Ad > 0.0;
  const double wc = Ac * (1.0 - 0x1p-50), wd = Ad * (1.0 - 0x1p-50);
  // lowest in-window index of each group, and whether that drop is A
  const bool c0 = d[0] >= wc, c2 = d[2] >= wc, c4 = d[4] >= wc;
  const bool d1 = d[1] >= wd, d3 = d[3] >= wd, d5 = d[5] >= wd;
  const int ic = c0 ? 0 : c2 ? 2 : c4 ? 4 : 6;
  const int id = d1 ? 1 : d3 ? 3 : d5 ? 5 : 7;
  const bool fc = (d[0] == Ac) | (!c0 & ((d[2] == Ac) | (!c2 & ((d[4] == Ac) | (!c4 &
                  (d[6] == Ac))))));
  const bool fd = (d[1] == Ad) | (!d1 & ((d[3] == Ad) | (!d3 & ((d[5] == Ad) | (!d5 &
                  (d[7] == Ad))))));
  // quotients of drops in [2^-900, 2^900] by divisors in [2^-60, 2^60] are
  // normal and nonzero: a positive drop is a positive slope
  const bool rc = (Ac >= 0x1p-900) & (Ac <= 0x1p900);
  const bool rd = (Ad >= 0x1p-900) & (Ad <= 0x1p900);
  const double P = Ac * dd, Q = Ad * dc;
  const bool cw = P > Q * (1.0 + 0x1p-48), dw = Q > P * (1.0 + 0x1p-48);
  *ok = div_ok & (!pc | (rc & fc)) & (!pd | (rd & fd)) & (!(pc & pd) | cw | dw);
  const int best = (pc & (!pd | cw)) ? ic : id;
  return (!pc & !pd) ? kNoReceiver : best;
}

__host__ __device__ inline int receiver_fast(double hc, const double nb[8], double dc, double dd,
                                             bool div_ok) {
  bool ok;
  const int r = receiver_try(hc, nb, dc, dd, div_ok, &ok);
  return ok ? r : receiver_two_div(hc, nb, dc, dd);
}

__host__ __device__ inline int receiver_fast(double hc, const double nb[8], double dc, double dd) {
  return receiver_fast(hc, nb, dc, dd, divisors_in_range(dc, dd));
}

}  // namespace fs

#endif
