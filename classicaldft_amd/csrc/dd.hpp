// Double-double arithmetic (value = hi + lo, |lo| <= ulp(hi)/2, about 106
// significant bits) for the one-time weight set-up, host and device.
//
// Error-free transformations: two_sum (Knuth), two_prod through fma (exact
// product on any IEEE fma unit).  Addition, multiplication and division are the
// accurate double-double algorithms of Shewchuk / Bailey's QD library
// (relative error a few 2^-106); sqrt is one Newton step from the double root;
// atan2 corrects the double angle t0 by atan of the residual rotation
// (y cos t0 - x sin t0) / (x cos t0 + y sin t0), with cos / sin of t0 from
// their Taylor series at t0 / 16 and four angle doublings.
#pragma once
#include <cmath>

#ifndef CDFT_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define CDFT_HD __host__ __device__
#else
#define CDFT_HD
#endif
#endif

// Error-free transformations need every product and sum rounded on its own: a
// multiply fused into a later add (the AMDGPU back end fuses across statements
// under the default -ffp-contract=fast) breaks quick_two_sum after two_prod.
// Every function of this file and of wcells.hpp starts with CDFT_EXACT_FP.
#ifndef CDFT_EXACT_FP
#if defined(__clang__)
#define CDFT_EXACT_FP _Pragma("clang fp contract(off)")
#else
#define CDFT_EXACT_FP
#endif
#endif

namespace cdft {

struct dd {
  double hi, lo;
  CDFT_HD dd() : hi(0.0), lo(0.0) {}
  CDFT_HD dd(double h) : hi(h), lo(0.0) {}
  CDFT_HD dd(double h, double l) : hi(h), lo(l) {}
};

CDFT_HD inline dd two_sum(double a, double b) {
  CDFT_EXACT_FP
  const double s = a + b;
  const double bb = s - a;
  return dd(s, (a - (s - bb)) + (b - bb));
}
CDFT_HD inline dd quick_two_sum(double a, double b) {  // |a| >= |b|
  CDFT_EXACT_FP
  const double s = a + b;
  return dd(s, b - (s - a));
}
CDFT_HD inline dd two_prod(double a, double b) {
  CDFT_EXACT_FP
  const double p = a * b;
  return dd(p, std::fma(a, b, -p));
}

CDFT_HD inline dd operator+(dd a, dd b) {
  CDFT_EXACT_FP
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
CDFT_HD inline dd operator-(dd a) { return dd(-a.hi, -a.lo); }
CDFT_HD inline dd operator-(dd a, dd b) { return a + (-b); }
CDFT_HD inline dd operator*(dd a, dd b) {
  CDFT_EXACT_FP
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}
CDFT_HD inline dd operator/(dd a, dd b) {
  CDFT_EXACT_FP
  const double q1 = a.hi / b.hi;
  dd r = a - b * dd(q1);
  const double q2 = r.hi / b.hi;
  r = r - b * dd(q2);
  const double q3 = r.hi / b.hi;
  dd q = quick_two_sum(q1, q2);
  return q + dd(q3);
}
CDFT_HD inline dd operator+(dd a, double b) { return a + dd(b); }
CDFT_HD inline dd operator+(double a, dd b) { return dd(a) + b; }
CDFT_HD inline dd operator-(dd a, double b) { return a - dd(b); }
CDFT_HD inline dd operator-(double a, dd b) { return dd(a) - b; }
CDFT_HD inline dd operator*(dd a, double b) { return a * dd(b); }
CDFT_HD inline dd operator*(double a, dd b) { return dd(a) * b; }
CDFT_HD inline dd operator/(dd a, double b) { return a / dd(b); }
CDFT_HD inline dd operator/(double a, dd b) { return dd(a) / b; }
CDFT_HD inline dd& operator+=(dd& a, dd b) { return a = a + b; }
CDFT_HD inline dd& operator-=(dd& a, dd b) { return a = a - b; }
CDFT_HD inline bool operator<(dd a, dd b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
CDFT_HD inline bool operator<(dd a, double b) { return a < dd(b); }
CDFT_HD inline bool operator>(dd a, dd b) { return b < a; }
CDFT_HD inline bool operator>(dd a, double b) { return dd(b) < a; }
CDFT_HD inline bool operator<=(dd a, double b) { return !(dd(b) < a); }

CDFT_HD inline double to_double(double a) { return a; }
CDFT_HD inline double to_double(dd a) {
  CDFT_EXACT_FP
  return a.hi + a.lo;
}

// pi as a double-double
CDFT_HD inline dd dd_pi() { return dd(3.141592653589793116, 1.2246467991473532e-16); }

CDFT_HD inline dd dd_sqrt(dd a) {
  CDFT_EXACT_FP
  if (a.hi <= 0.0) return dd(0.0);
  const double x = std::sqrt(a.hi);
  const dd r = a - two_prod(x, x);  // a - x^2, exact part first
  return quick_two_sum(x, r.hi / (2.0 * x));
}

// cos and sin of a double-double angle |t| <= 4.
CDFT_HD inline void dd_sincos(dd t, dd& s, dd& c) {
  CDFT_EXACT_FP
  const dd h = t * 0.0625;  // exact scaling
  const dd h2 = h * h;
  // Horner from the term of order 23 (|h| <= 0.25: h^25/25! < 1e-40)
  dd ss(0.0), cc(0.0);
  for (int k = 23; k >= 1; k -= 2) ss = dd(1.0) - h2 * ss / dd(double((k + 1) * (k + 2)));
  for (int k = 22; k >= 0; k -= 2) cc = dd(1.0) - h2 * cc / dd(double((k + 1) * (k + 2)));
  // ss = 1 - h^2/(2*3)(1 - ...) : the loops above build sin(h)/h and cos(h)
  s = h * ss;
  c = cc;
  for (int i = 0; i < 4; ++i) {  // double the angle four times
    const dd s2 = 2.0 * s * c;
    const dd c2 = c * c - s * s;
    s = s2;
    c = c2;
  }
}

CDFT_HD inline dd dd_atan2(dd y, dd x) {
  CDFT_EXACT_FP
  const double t0 = std::atan2(y.hi, x.hi);
  if (y.hi == 0.0 && x.hi >= 0.0) return dd(0.0);
  dd s, c;
  dd_sincos(dd(t0), s, c);
  const dd num = y * c - x * s;  // r sin(t - t0)
  const dd den = x * c + y * s;  // r cos(t - t0)
  const double d = to_double(num / den);
  // atan(d) = d - d^3/3 to far below 2^-106 of t for |d| < 1e-15
  return dd(t0) + dd(d - d * d * d / 3.0);
}

}  // namespace cdft
