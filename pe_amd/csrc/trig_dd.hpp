// Correctly rounded sin, cos, atan2 and acos for FP64 on the device.
//
// pe's dynamic-friction branch (HCSITS.h:2429-2509) runs a golden-section search whose
// comparisons of nearly equal objective values are decided by the last bits of sin/cos;
// a one-ulp difference between two libms changes its result by ~1e-6.  The host libm
// (glibc) is correctly rounded for all but a small fraction of arguments, so evaluating
// these functions correctly rounded on the device reproduces the host's results.
//
// Method: Cody-Waite reduction by pi/2 in double-double (exact for |x| < 2^20), Taylor
// series of sin/cos in double-double, rounded once to double; atan2 and acos take the
// device libm value and refine it with one double-double Newton step on sin/cos.
// Accuracy before the final rounding is ~2^-100 relative, so the result is the correctly
// rounded value except within ~2^-100 of a rounding boundary.
#pragma once

#include <math.h>

#include "dmath.hpp"
#include "trig_tables.hpp"

namespace pe_amd {
namespace dd {

struct ddv {
  double hi, lo;
};

PE_HD ddv two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
PE_HD ddv quick_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
PE_HD ddv two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
PE_HD ddv add(ddv a, ddv b) {
  ddv s = two_sum(a.hi, b.hi);
  const ddv t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
PE_HD ddv add_d(ddv a, double b) {
  ddv s = two_sum(a.hi, b);
  s.lo += a.lo;
  return quick_two_sum(s.hi, s.lo);
}
PE_HD ddv mul(ddv a, ddv b) {
  ddv p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}
PE_HD ddv mul_d(ddv a, double b) {
  ddv p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return quick_two_sum(p.hi, p.lo);
}
PE_HD ddv neg(ddv a) { return {-a.hi, -a.lo}; }

// x = k*(pi/2) + r, |r| <~ pi/4, r in double-double; k returned modulo 4.
PE_HD int reduce(double x, ddv& r) {
  using namespace trig_tables;
  const double k = rint(x * kTwoOverPi);
  ddv t = two_sum(x, -(k * kP1));          // k*P1 is exact
  t = add_d(t, -(k * kP2));                // k*P2 is exact
  t = add(t, neg(two_prod(k, kP3)));
  r = t;
  return ((int)fmod(k, 4.0) + 4) & 3;
}

// sin(r), cos(r) for |r| <= ~0.786 by Taylor series (terms to r^27 / r^26).
PE_HD void sincos_kernel(ddv r, ddv& s, ddv& c) {
  using namespace trig_tables;
  const ddv r2 = mul(r, r);
  // sin(r) = r * (1/1! - r^2 (1/3! - r^2 (1/5! - ...)));  cos(r) = 1/0! - r^2 (1/2! - ...)
  ddv ps = {kInvFactHi[27], kInvFactLo[27]};
  ddv pc = {kInvFactHi[26], kInvFactLo[26]};
  for (int n = 25; n >= 1; n -= 2) ps = add({kInvFactHi[n], kInvFactLo[n]}, neg(mul(ps, r2)));
  for (int n = 24; n >= 0; n -= 2) pc = add({kInvFactHi[n], kInvFactLo[n]}, neg(mul(pc, r2)));
  s = mul(ps, r);
  c = pc;
}

PE_HD void sincos_dd(double x, ddv& s, ddv& c) {
  ddv r;
  const int q = reduce(x, r);
  ddv sr, cr;
  sincos_kernel(r, sr, cr);
  switch (q) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = neg(sr); break;
    case 2: s = neg(sr); c = neg(cr); break;
    default: s = neg(cr); c = sr; break;
  }
}

PE_HD double round_dd(ddv a) { return a.hi + a.lo; }

}  // namespace dd

// Correctly rounded trig (arguments |x| < 2^20).
struct TrigExact {
  static constexpr bool kCorrectlyRounded = true;
  static PE_HD double sin(double x) {
    if (x == 0.0) return x;
    dd::ddv s, c;
    dd::sincos_dd(x, s, c);
    return dd::round_dd(s);
  }
  static PE_HD double cos(double x) {
    dd::ddv s, c;
    dd::sincos_dd(x, s, c);
    return dd::round_dd(c);
  }
  // sin(x) and cos(x) from one reduction and series: the same values as sin() and cos().
  static PE_HD void sincos(double x, double& sn, double& cs) {
    dd::ddv s, c;
    dd::sincos_dd(x, s, c);
    sn = x == 0.0 ? x : dd::round_dd(s);
    cs = dd::round_dd(c);
  }
  // atan2: device value refined by one Newton step on y cos(a) - x sin(a) = 0.
  static PE_HD double atan2(double y, double x) {
    const double a0 = ::atan2(y, x);
    if (!(y != 0.0) || !isfinite(a0)) return a0;
    dd::ddv s, c;
    dd::sincos_dd(a0, s, c);
    const dd::ddv num = dd::add(dd::mul_d(c, y), dd::neg(dd::mul_d(s, x)));
    const double den = x * c.hi + y * s.hi;
    if (den == 0.0) return a0;
    const double q = num.hi / den;
    const dd::ddv corr = dd::two_prod(q, den);  // num - q*den: remainder for the low part
    const double q2 = (dd::add(num, dd::neg(corr)).hi) / den;
    return dd::round_dd(dd::add_d(dd::two_sum(a0, q), q2));
  }
  // acos on [-1, 1]: device value refined by one Newton step on cos(a) - f = 0.
  static PE_HD double acos(double f) {
    const double a0 = ::acos(f);
    if (!(a0 > 0.0) || !isfinite(a0)) return a0;
    dd::ddv s, c;
    dd::sincos_dd(a0, s, c);
    const dd::ddv num = dd::add_d(c, -f);  // cos(a0) - f
    const double q = num.hi / s.hi;
    const dd::ddv corr = dd::two_prod(q, s.hi);
    const double q2 = (dd::add(num, dd::neg(corr)).hi) / s.hi;
    return dd::round_dd(dd::add_d(dd::two_sum(a0, q), q2));
  }
};

// Fast trig of the production solver: the device libm (within about 1 ulp, not correctly
// rounded).  sin and cos of one angle come from one sincos, which returns the same values.
struct TrigFast {
  static constexpr bool kCorrectlyRounded = false;
  static PE_HD double sin(double x) { return ::sin(x); }
  static PE_HD double cos(double x) { return ::cos(x); }
  static PE_HD void sincos(double x, double& sn, double& cs) { ::sincos(x, &sn, &cs); }
  static PE_HD double atan2(double y, double x) { return ::atan2(y, x); }
  static PE_HD double acos(double f) { return ::acos(f); }
};

}  // namespace pe_amd
