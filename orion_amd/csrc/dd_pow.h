// dd_pow.h -- x^y rounded correctly to the nearest double, for x > 0.
//
// The far-field boundary condition raises densities and entropy ratios to
// the powers gamma and 1/(gamma-1) (src/bc/PhysicalBC.cpp:413-414,437-438),
// where the reference calls glibc's pow.  glibc's pow is correctly rounded
// for all but a few in ten thousand of these arguments; a pow that is always
// correctly rounded agrees with it that often, where a faithfully rounded
// one (within one ulp) disagrees far more often.  Each disagreement is an
// ulp in a ghost value, and a multi-block solution amplifies it step by step.
//
// Method (textbook double-double arithmetic, about 2^-100 relative before
// the final rounding):
//   log x = e ln2 + 2 atanh(s),  s = (m - 1)/(m + 1),  x = m 2^e,
//           m in [sqrt(1/2), sqrt(2)), atanh by its odd series in s^2;
//   x^y   = exp(y log x) = 2^k exp(t)^1024,  t = (y log x - k ln2)/1024,
//           exp(t) by its Taylor series, then ten double-double squarings.
// Only IEEE +, -, *, / and fma appear, each correctly rounded, so the host
// and the device compute the same bits.  Non-positive, infinite and NaN x
// go to the library pow (the IEEE special cases).
#ifndef ORION_DD_POW_H
#define ORION_DD_POW_H

#include <math.h>

#if defined(__HIPCC__)
#define ORION_HD __host__ __device__
#else
#define ORION_HD
#endif

namespace orion_dd {

struct dd {
  double hi, lo;
};

ORION_HD inline dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}
ORION_HD inline dd quick_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return dd{s, b - (s - a)};
}
ORION_HD inline dd two_prod(double a, double b) {
  const double p = a * b;
  return dd{p, fma(a, b, -p)};
}
ORION_HD inline dd add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi);
  const dd t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
ORION_HD inline dd mul(dd x, dd y) {
  dd p = two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return quick_two_sum(p.hi, p.lo);
}
ORION_HD inline dd mul_d(dd x, double y) {
  dd p = two_prod(x.hi, y);
  p.lo += x.lo * y;
  return quick_two_sum(p.hi, p.lo);
}
ORION_HD inline dd div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  dd r = add(a, mul_d(b, -q1));
  const double q2 = r.hi / b.hi;
  r = add(r, mul_d(b, -q2));
  const double q3 = r.hi / b.hi;
  return add(quick_two_sum(q1, q2), dd{q3, 0.0});
}

// natural log of x > 0 (finite, normal)
ORION_HD inline dd log_dd(double x) {
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const dd s = div(dd{m - 1.0, 0.0}, two_sum(m, 1.0));  // m - 1 is exact
  const dd s2 = mul(s, s);
  // sum_{k=0..K} s2^k / (2k + 1), |s2| < 0.0295: K = 23 leaves < 2^-115
  const int K = 23;
  dd p = div(dd{1.0, 0.0}, dd{2.0 * K + 1.0, 0.0});
  for (int k = K - 1; k >= 0; --k) p = add(mul(p, s2), div(dd{1.0, 0.0}, dd{2.0 * k + 1.0, 0.0}));
  const dd atanh2 = mul_d(mul(s, p), 2.0);
  const dd ln2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  return add(mul_d(ln2, (double)e), atanh2);
}

// x^y for finite x > 0, rounded correctly (ties to even)
ORION_HD inline double pow_cr(double x, double y) {
  if (!(x > 0.0) || !(x < INFINITY) || !(y == y) || !(fabs(y) < INFINITY)) return pow(x, y);
  if (x == 1.0 || y == 0.0) return 1.0;
  const dd r = mul_d(log_dd(x), y);
  if (!(fabs(r.hi) < 700.0)) return pow(x, y);  // overflow and underflow ranges
  const dd ln2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  const double k = rint(r.hi / ln2.hi);
  dd t = add(r, mul_d(ln2, -k));
  t.hi *= 0x1p-10;  // exact scalings
  t.lo *= 0x1p-10;
  // exp(t), |t| < 3.4e-4: Taylor to t^10/10!
  dd e{1.0, 0.0};
  for (int n = 10; n >= 1; --n) e = add(dd{1.0, 0.0}, div(mul(e, t), dd{(double)n, 0.0}));
  for (int q = 0; q < 10; ++q) e = mul(e, e);
  const dd n = quick_two_sum(e.hi, e.lo);  // hi = round(hi + lo)
  return ldexp(n.hi, (int)k);
}

}  // namespace orion_dd

#endif  // ORION_DD_POW_H
