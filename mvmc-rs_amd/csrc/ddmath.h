// Double-double exp/log, correctly rounded except when the exact value lies
// within ~2^-100 (relative) of a rounding boundary.  Used for the Metropolis
// weight exp(2*dlog) and log|<x|psi>| on the device so that accept/reject
// decisions reproduce a correctly rounded libm; the same source compiles on
// the host for the agreement test against glibc.
//
// Method: Cody-Waite reduction x = k*ln2 + r with ln2 in three parts, then
// exp(r) by Taylor series in double-double arithmetic (|r| <= ln2/2);
// log(m) = 2*atanh((m-1)/(m+1)) by its odd series in double-double.
// Double-double primitives: Dekker/Knuth two-sum and fma two-product.
// Requires IEEE round-to-nearest and no floating-point contraction
// (-ffp-contract=off) for the error-free transformations.
#pragma once
#include <math.h>
#include <stdint.h>

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

DD_FN double cr_exp(double x) {
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

DD_FN double cr_log(double x) {
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

}  // namespace ddm
