// Correctly rounded exp/log for the Metropolis weight exp(2*dlog) and
// log|<x|psi>| on the device, so accept/reject decisions reproduce a correctly
// rounded libm; the same source compiles on the host (the oracle and the
// agreement tests against glibc and 60-digit decimal arithmetic).
//
// Fast path (Ziv's strategy): table-driven reduction, a short polynomial and
// a double-double result y_hi + y_lo with a proven error bound E; when
// y_hi + (y_lo - E) and y_hi + (y_lo + E) round to the same double that double
// is the correctly rounded value.  Otherwise (about one call in 250) the
// accurate path runs:
//   exp: Cody-Waite reduction x = k*ln2 + r (ln2 in three parts), Taylor
//        series of exp(r) in double-double (|r| <= ln2/2);
//   log: log(m) = 2*atanh((m-1)/(m+1)) by its odd series in double-double;
// both correctly rounded unless the exact value lies within ~2^-100
// (relative) of a rounding boundary.
// Double-double primitives: Dekker/Knuth two-sum and fma two-product.
// Requires IEEE round-to-nearest and no floating-point contraction
// (-ffp-contract=off) for the error-free transformations.
#pragma once
#include <math.h>
#include <stdint.h>

#include "ddmath_tables.h"

#if defined(__HIPCC__)
#define DD_FN __host__ __device__ inline
#else
#define DD_FN inline
#endif

namespace ddm {

struct dd {
  double hi, lo;
};

DD_FN dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
DD_FN dd quick_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
DD_FN dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}
DD_FN dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
DD_FN dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}
DD_FN dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return quick_two_sum(p.hi, p.lo);
}
DD_FN dd div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = add(a, mul_d(b, -q1));
  double q2 = r.hi / b.hi;
  r = add(r, mul_d(b, -q2));
  double q3 = r.hi / b.hi;
  dd q = quick_two_sum(q1, q2);
  return add(q, dd{q3, 0.0});
}

// ln 2 split: hi has 32 significant bits so k*hi is exact for |k| < 2^21.
constexpr double LN2_HI = 0x1.62e42feep-1;              // 0x3FE62E42FEE00000
constexpr double LN2_MID = 0x1.a39ef35793c76p-33;        // RN(ln2 - LN2_HI)
constexpr double LN2_LO = 0x1.cc01f97b57a08p-87;         // RN(ln2 - LN2_HI - LN2_MID)
constexpr double LN2_DD_HI = 0x1.62e42fefa39efp-1;       // RN(ln 2)
constexpr double LN2_DD_LO = 0x1.abc9e3b39803fp-56;      // RN(ln 2 - LN2_DD_HI)

DD_FN double exp_accurate(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.2) return 0.0;
  const double k = rint(x * 1.4426950408889634074);
  // r = x - k*ln2 in double-double; k*LN2_HI is exact (32-bit constant, |k| < 2^11).
  dd r = two_sum(x, -(k * LN2_HI));
  r = add(r, two_prod(-k, LN2_MID));
  r = add(r, dd{-(k * LN2_LO), 0.0});
  // Taylor series of exp(r), |r| <= 0.347, to 2^-110.
  dd term = {1.0, 0.0};
  dd sum = {1.0, 0.0};
  for (int i = 1; i <= 27; ++i) {
    term = mul(term, r);
    term = div(term, dd{(double)i, 0.0});
    sum = add(sum, term);
    if (fabs(term.hi) < 1e-36 * fabs(sum.hi)) break;
  }
  const int ki = (int)k;
  if (ki > -1021) return ldexp(sum.hi, ki);
  // subnormal range: scale in two steps, rounding once at the end
  dd s2 = {ldexp(sum.hi, ki + 600), ldexp(sum.lo, ki + 600)};
  return (s2.hi + s2.lo) * 0x1p-600;
}

DD_FN double log_accurate(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int e;
  double m = frexp(x, &e);  // x = m*2^e, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  if (m == 1.0 && e == 0) return 0.0;
  // s = (m-1)/(m+1) in double-double; m-1 and m+1 are exact for m in [0.707,1.414).
  dd num = two_sum(m, -1.0);
  dd den = two_sum(m, 1.0);
  dd s = div(num, den);
  dd s2 = mul(s, s);
  dd term = s;
  dd sum = s;
  for (int k = 3; k <= 81; k += 2) {
    term = mul(term, s2);
    dd t = div(term, dd{(double)k, 0.0});
    sum = add(sum, t);
    if (fabs(t.hi) < 1e-36 * fabs(sum.hi)) break;
  }
  sum = add(sum, sum);  // 2*atanh(s)
  dd el = mul_d(dd{LN2_DD_HI, LN2_DD_LO}, (double)e);
  dd r = add(el, sum);
  return r.hi;
}

// ---------------------------------------------------------------------------
// fast paths
// ---------------------------------------------------------------------------
constexpr double INV_L64 = 0x1.71547652b82fep+6;  // ~64/ln2 (any nearby value works)
constexpr double L64_HI = LN2_HI * 0x1p-6;          // ln2/64 in three parts (exact scalings)
constexpr double L64_MID = LN2_MID * 0x1p-6;
constexpr double L64_LO = LN2_LO * 0x1p-6;

DD_FN double cr_exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.2) return 0.0;
  const double kd = rint(x * INV_L64);  // |k| < 2^17
  const int k = (int)kd;
  const int j = k & 63;
  const int m = (k - j) / 64;
  if (m < -1020) return exp_accurate(x);  // subnormal range
  // r = x - k*ln2/64: x - k*L64_HI is exact (32-bit constant, Sterbenz)
  const double t1 = x - kd * L64_HI;
  const dd p = two_prod(kd, L64_MID);
  const dd rs = two_sum(t1, -p.hi);
  const double r = rs.hi;
  const double rl = (rs.lo - p.lo) - kd * L64_LO;  // |r| <= 2^-7.5, |rl| <= 2^-60
  // exp(r) = 1 + r + r^2 * P(r), P to r^5/5040: truncation < 2^-75
  const double P = 0.5 + r * (0x1.5555555555555p-3 + r * (0x1.5555555555555p-5 + r * (0x1.1111111111111p-7 +
                   r * (0x1.6c16c16c16c17p-10 + r * 0x1.a01a01a01a01ap-13))));
  const double r2p = (r * r) * P;
  const dd e0 = quick_two_sum(1.0, r);
  const double elo = e0.lo + (rl + (r2p + r * rl));
  // y = T[j] * (e0.hi + elo), |rel. error| < 2^-64
  const double th = EXP_T[j][0], tl = EXP_T[j][1];
  const dd y = two_prod(th, e0.hi);
  const double ylo = y.lo + (th * elo + tl * e0.hi);
  const double err = 0x1p-62 * y.hi;
  const double a = y.hi + (ylo - err);
  const double b = y.hi + (ylo + err);
  if (a != b) return exp_accurate(x);
  return ldexp(a, m);
}

DD_FN double cr_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  if (x == 1.0) return 0.0;
  if (x < 0x1p-1022) return log_accurate(x);  // subnormal input
  int e;
  double mm = frexp(x, &e);  // x = mm*2^e, mm in [0.5, 1)
  if (mm < 0.70710678118654752440) {
    mm *= 2.0;
    e -= 1;
  }
  // mm*R - 1 = t_hi + t_lo exactly, |t| <= 2^-7.4
  const int j = (int)rint(mm * 128.0) - LOG_J0;
  const double R = LOG_R[j];
  const dd pr = two_prod(mm, R);
  const double t = pr.hi - 1.0;  // exact (pr.hi within 2^-7 of 1)
  const double tl = pr.lo;
  // log1p(t) = t - t^2/2 + t^3 * Q(t), Q to t^6/9: truncation < 2^-77
  const dd q = two_prod(t, t);
  const double Q = 0x1.5555555555555p-2 + t * (-0.25 + t * (0x1.999999999999ap-3 + t * (-0x1.5555555555555p-3 +
                   t * (0x1.2492492492492p-3 + t * (-0.125 + t * 0x1.c71c71c71c71cp-4)))));
  const double t3q = (q.hi * t) * Q;
  const dd h = two_sum(t, -0.5 * q.hi);
  const double hlo = h.lo + (((tl - t * tl) - 0.5 * q.lo) + t3q);
  // e*ln2 + (-log R) + log1p(t)
  const double ed = (double)e;
  const dd u = two_prod(ed, LN2_DD_HI);
  const double ulo = u.lo + ed * LN2_DD_LO;
  const dd s1 = two_sum(u.hi, LOG_L[j][0]);
  const dd s2 = two_sum(s1.hi, h.hi);
  const double ylo = s1.lo + s2.lo + ((ulo + LOG_L[j][1]) + hlo);
  const double mag = fabs(u.hi) + fabs(LOG_L[j][0]) + fabs(h.hi);
  const double err = 0x1p-62 * mag;
  const double a = s2.hi + (ylo - err);
  const double b = s2.hi + (ylo + err);
  if (a != b) return log_accurate(x);
  return a;
}

}  // namespace ddm
