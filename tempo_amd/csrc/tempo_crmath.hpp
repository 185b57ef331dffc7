// Nearly correctly rounded exp, log, log10 and pow in FP64 for the FP64 twin.
//
// The reference's FP64 build calls the host libm (glibc), whose exp, log and
// pow return the correctly rounded result in all but a tiny fraction of
// arguments.  The GPU's own libm is accurate to about one ulp, which differs
// from glibc in roughly one pow call in four and, through threshold
// branches, occasionally moves a column's result far more than an ulp.  These
// routines evaluate the function in double-double arithmetic with a relative
// error near 2^-77 and round once, so they agree with a correctly rounded
// libm except when the exact value lies within ~2^-77 relative of a rounding
// boundary.
//
// Methods (standard, from the literature):
//   error-free transformations two_sum / two_prod (Knuth; Dekker 1971, with
//   fused multiply-add for the product);
//   exp: Cody-Waite reduction by ln2/64 with a three-part constant, table of
//   2^(j/64) in double-double (Tang 1989), degree-8 Taylor polynomial whose
//   cubic and higher terms are summed in double;
//   log: frexp reduction to [1/sqrt2, sqrt2), a one-ulp first approximation
//   refined by one Newton step y + (m e^-y - 1) with e^-y in double-double;
//   pow: exp of y*log(x) with the product carried in double-double.
// Arguments outside the ranges below (non-positive, non-finite, results near
// overflow or in the subnormal range) fall back to the device libm, whose
// special-value behaviour is the IEEE one the reference also follows.
#pragma once

#include <hip/hip_runtime.h>

namespace tempo {
namespace crm {

#include "tempo_crmath_tables.inc"

struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ dd fast_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return {s, b - (s - a)};
}

__device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}

// e^(zh + zl) = 2^k * (hi + lo) for |zh| <= 708, |zl| <= ulp(zh)
__device__ __forceinline__ dd exp_dd(double zh, double zl, int& k) {
  const double n = rint(zh * kInvLn2x64);
  const int ni = int(n);
  const double r1 = fma(-n, kLn2o64C1, zh);  // exact: n*c1 exact, Sterbenz
  const dd p2 = two_prod(n, kLn2o64C2);
  dd r = two_sum(r1, -p2.hi);
  r.lo = r.lo - p2.lo - n * kLn2o64C3 + zl;
  r = fast_two_sum(r.hi, r.lo);
  const int j = ni & 63;
  k = (ni - j) / 64;
  const double x = r.hi;
  // e^x - 1 = x + x^2/2 + x^3 P(x), |x| <= ln2/128
  dd s = two_prod(x, x);
  s.hi *= 0.5;
  s.lo *= 0.5;
  const double P = 1.0 / 6.0 + x * (1.0 / 24.0 + x * (1.0 / 120.0 + x * (1.0 / 720.0 + x * (1.0 / 5040.0 +
                                                                                           x * (1.0 / 40320.0)))));
  const double q = x * x * x * P;
  dd a = fast_two_sum(x, s.hi);
  const double lo = a.lo + s.lo + q + r.lo + r.lo * x;
  const dd pm1 = fast_two_sum(a.hi, lo);
  // 2^(j/64) * (1 + pm1)
  const double th = kExp2TabHi[j], tl = kExp2TabLo[j];
  const dd m1 = two_prod(th, pm1.hi);
  const double lo2 = tl + m1.lo + th * pm1.lo + tl * pm1.hi;
  dd h = two_sum(th, m1.hi);
  h.lo += lo2;
  return fast_two_sum(h.hi, h.lo);
}

// log(x) = hi + lo for positive normal finite x
__device__ __forceinline__ dd log_dd(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double y0 = log(m);
  int k;
  dd E = exp_dd(-y0, 0.0, k);
  E.hi = ldexp(E.hi, k);
  E.lo = ldexp(E.lo, k);
  const dd P = two_prod(m, E.hi);
  const double u = (P.hi - 1.0) + (P.lo + m * E.lo);  // m e^-y0 - 1, |u| ~ 1e-16
  const double ed = double(e);
  const dd h = two_sum(ed * kLn2Hi, y0);
  const double lo = h.lo + u + ed * kLn2Lo;
  return fast_two_sum(h.hi, lo);
}

__device__ __forceinline__ double scale_result(const dd& r, int k) {
  // r.hi is already the double nearest r.hi + r.lo; the scaling is exact in
  // the normal range the callers guarantee
  return ldexp(r.hi, k);
}

__device__ __forceinline__ double exp_cr(double x) {
  if (!(x > -708.0 && x < 709.0)) return exp(x);
  int k;
  const dd r = exp_dd(x, 0.0, k);
  return scale_result(r, k);
}

__device__ __forceinline__ bool normal_positive(double x) {
  return x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308;
}

__device__ __forceinline__ double log_cr(double x) {
  if (!normal_positive(x)) return log(x);
  if (x == 1.0) return 0.0;
  const dd r = log_dd(x);
  return r.hi;
}

__device__ __forceinline__ double log10_cr(double x) {
  if (!normal_positive(x)) return log10(x);
  if (x == 1.0) return 0.0;
  const dd l = log_dd(x);
  const dd p = two_prod(l.hi, kInvLn10Hi);
  const double lo = p.lo + l.hi * kInvLn10Lo + l.lo * kInvLn10Hi;
  return fast_two_sum(p.hi, lo).hi;
}

__device__ __forceinline__ double pow_cr(double x, double y) {
  if (!normal_positive(x) || !(y == y) || y == 0.0 || x == 1.0 || !(fabs(y) < 1.0e300)) return pow(x, y);
  const dd l = log_dd(x);
  const dd p = two_prod(y, l.hi);
  const dd z = fast_two_sum(p.hi, p.lo + y * l.lo);
  if (!(z.hi > -708.0 && z.hi < 709.0)) return pow(x, y);
  int k;
  const dd r = exp_dd(z.hi, z.lo, k);
  return scale_result(r, k);
}

}  // namespace crm
}  // namespace tempo
