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
// The series' coefficients 1/(2k+1) and 1/n are double-double constants.
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

// 1/(2k+1), k = 0..23, and 1/n, n = 1..10, as double-doubles (hi the
// nearest double, lo the nearest double to the remainder: exact to about
// 2^-106), so that the series multiply where they would divide.
#if defined(__HIPCC__)
#define ORION_DD_CONST __device__ __constant__
#else
#define ORION_DD_CONST
#endif
static ORION_DD_CONST const double RCP_ODD[24][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.5555555555555p-2, 0x1.5555555555555p-56},
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},
    {0x1.2492492492492p-3, 0x1.2492492492492p-57},
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
    {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
    {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
    {0x1.1111111111111p-4, 0x1.1111111111111p-60},
    {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},
    {0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59},
    {0x1.8618618618618p-5, 0x1.8618618618618p-59},
    {0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60},
    {0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61},
    {0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59},
    {0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61},
    {0x1.0842108421084p-5, 0x1.0842108421084p-60},
    {0x1.f07c1f07c1f08p-6, -0x1.f07c1f07c1f08p-61},
    {0x1.d41d41d41d41dp-6, 0x1.0750750750750p-60},
    {0x1.bacf914c1bad0p-6, -0x1.bacf914c1bad0p-60},
    {0x1.a41a41a41a41ap-6, 0x1.0690690690690p-60},
    {0x1.8f9c18f9c18fap-6, -0x1.f3831f3831f38p-61},
    {0x1.7d05f417d05f4p-6, 0x1.7d05f417d05f4p-62},
    {0x1.6c16c16c16c17p-6, -0x1.f49f49f49f49fp-61},
    {0x1.5c9882b931057p-6, 0x1.310572620ae4cp-61}};
static ORION_DD_CONST const double RCP_N[10][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p-1, 0x0.0p+0},
    {0x1.5555555555555p-2, 0x1.5555555555555p-56},
    {0x1.0000000000000p-2, 0x0.0p+0},
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.2492492492492p-3, 0x1.2492492492492p-57},
    {0x1.0000000000000p-3, 0x0.0p+0},
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
    {0x1.999999999999ap-4, -0x1.999999999999ap-58}};

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
  dd p = dd{RCP_ODD[K][0], RCP_ODD[K][1]};
  for (int k = K - 1; k >= 0; --k) p = add(mul(p, s2), dd{RCP_ODD[k][0], RCP_ODD[k][1]});
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
  for (int n = 10; n >= 1; --n) e = add(dd{1.0, 0.0}, mul(mul(e, t), dd{RCP_N[n - 1][0], RCP_N[n - 1][1]}));
  for (int q = 0; q < 10; ++q) e = mul(e, e);
  const dd n = quick_two_sum(e.hi, e.lo);  // hi = round(hi + lo)
  return ldexp(n.hi, (int)k);
}

}  // namespace orion_dd

#endif  // ORION_DD_POW_H
