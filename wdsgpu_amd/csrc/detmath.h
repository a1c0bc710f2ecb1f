// Deterministic pow and log of the power laws (pipe, pump, emitter,
// Darcy-Weisbach friction).
//
// The reference evaluates these with the platform libm (dgga/solver.py:
// 58-69), whose last bit differs between CPU and GPU libraries and is then
// amplified by the Newton iteration.  This build fixes one algorithm, a
// fixed sequence of IEEE-754 double operations that the CPU checker repeats
// operation for operation, so both compute the same link coefficients bit
// for bit:
//
//   log x  x = m 2^k, m in [sqrt(1/2), sqrt(2)) (exact); s = (m - 1) / (m + 1)
//          in double-double; log m = 2 atanh s = 2 s + 2 s^3 P(s^2), P the
//          atanh series to s^22; k ln 2 with ln 2 split in a 32-bit head and
//          a tail (Cody & Waite, Software Manual for the Elementary
//          Functions, 1980); sums in double-double (Dekker 1971).
//   exp z  n = rint(z / ln 2), r = z - n ln 2 (exact head), e^r = 1 + r +
//          r^2/2 + r^3/6 + r^4 R(r), the first four terms in double-double,
//          the Taylor tail R to r^14 in double, times 2^n.
//   pow    x^y = exp(y log x) with y log x in double-double; y == 1 -> x,
//          y == 0 -> 1, x == 0 -> 0 (y > 0) or inf (y < 0).
//
// Products are rounded separately from sums (the library is compiled with
// -ffp-contract=off); the only fused multiply-add is TwoProduct's exact
// error term.  Within one ulp of the correctly rounded result.
#pragma once
#include <hip/hip_runtime.h>

namespace wds {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dm_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ dd dm_two_prod(double a, double b) {
  const double p = a * b;
  return dd{p, __builtin_fma(a, b, -p)};
}

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
constexpr double kInvLn2 = 1.44269504088896338700e+00;

__device__ __noinline__ dd det_log_dd(double x) {
  int ke;
  double m = frexp(x, &ke);  // x = m 2^k, m in [0.5, 1)
  m = m * 2.0;
  ke -= 1;
  if (!(m < 1.4142135623730951)) {
    m = m * 0.5;
    ke += 1;
  }
  const double k = (double)ke;
  const double f = m - 1.0;
  const dd u = dm_two_sum(m, 1.0);
  const double s_hi = f / u.hi;
  const dd pq = dm_two_prod(s_hi, u.hi);
  const double s_lo = (((f - pq.hi) - pq.lo) - s_hi * u.lo) / u.hi;
  const double z = s_hi * s_hi;
  double P = 1.0 / 23.0;
  P = P * z + 1.0 / 21.0;
  P = P * z + 1.0 / 19.0;
  P = P * z + 1.0 / 17.0;
  P = P * z + 1.0 / 15.0;
  P = P * z + 1.0 / 13.0;
  P = P * z + 1.0 / 11.0;
  P = P * z + 1.0 / 9.0;
  P = P * z + 1.0 / 7.0;
  P = P * z + 1.0 / 5.0;
  P = P * z + 1.0 / 3.0;
  const double tail = ((2.0 * s_hi) * z) * P;
  const dd a = dm_two_sum(k * kLn2Hi, 2.0 * s_hi);
  const double lo = ((k * kLn2Lo + 2.0 * s_lo) + tail) + a.lo;
  return dm_two_sum(a.hi, lo);
}

__device__ __forceinline__ double det_log(double x) {
  const dd l = det_log_dd(x);
  return l.hi + l.lo;
}

__device__ __noinline__ double det_exp_dd(double z_hi, double z_lo) {
  const double n = rint(z_hi * kInvLn2);
  const double a = z_hi - n * kLn2Hi;
  const dd r = dm_two_sum(a, z_lo - n * kLn2Lo);
  dd h = dm_two_prod(r.hi, r.hi);
  h.lo = h.lo + (2.0 * r.hi) * r.lo;
  h.hi = h.hi * 0.5;
  h.lo = h.lo * 0.5;
  dd c = dm_two_prod(r.hi, h.hi);
  c.lo = c.lo + r.hi * h.lo;
  const double t_hi = c.hi / 3.0;
  const dd p3 = dm_two_prod(t_hi, 3.0);
  const double t_lo = (((c.hi - p3.hi) - p3.lo) + c.lo) / 3.0;
  // 1/4! .. 1/14! (double roundings of the exact reciprocals)
  double R = 1.0 / 87178291200.0;
  R = R * r.hi + 1.0 / 6227020800.0;
  R = R * r.hi + 1.0 / 479001600.0;
  R = R * r.hi + 1.0 / 39916800.0;
  R = R * r.hi + 1.0 / 3628800.0;
  R = R * r.hi + 1.0 / 362880.0;
  R = R * r.hi + 1.0 / 40320.0;
  R = R * r.hi + 1.0 / 5040.0;
  R = R * r.hi + 1.0 / 720.0;
  R = R * r.hi + 1.0 / 120.0;
  R = R * r.hi + 1.0 / 24.0;
  const double t4 = (h.hi * h.hi) * (4.0 * R);
  const dd s = dm_two_sum(r.hi, h.hi);
  const dd u = dm_two_sum(s.hi, t_hi);
  const double s_lo = ((((s.lo + u.lo) + r.lo) + h.lo) + t_lo) + t4;
  const dd s2 = dm_two_sum(u.hi, s_lo);
  const dd o = dm_two_sum(1.0, s2.hi);
  return ldexp(o.hi + (o.lo + s2.lo), (int)n);
}

__device__ __forceinline__ double det_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_huge_val();
  const dd l = det_log_dd(x);
  const dd p = dm_two_prod(y, l.hi);
  const dd z = dm_two_sum(p.hi, p.lo + y * l.lo);
  return det_exp_dd(z.hi, z.lo);
}

}  // namespace wds
