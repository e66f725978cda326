// Double-double arithmetic and correctly rounded elementary functions.
//
// The f and J accumulators use double-double in place of CAMP's 80-bit
// long double production/loss sums (time_derivative.h:21-28,
// Jacobian.h:31-40): error-free transforms (Knuth TwoSum, FMA TwoProd) carry
// ~106 bits, more than x87 extended precision.
//
// The reference computes its rate constants, SIMPOL and aqueous-equilibrium
// constants and ZSR molalities with the C library's exp/log/pow/log10, which
// on the host are correctly rounded for all but a small fraction of
// arguments.  The device library's versions differ from them by an ulp far
// more often, and where the reference subtracts nearly equal terms (gas and
// particle fluxes near equilibrium) one ulp in a constant becomes a large
// relative change of f.  The cr_* functions below evaluate in double-double
// (argument reduction by ln 2, Taylor series, repeated squaring; log by one
// Newton step on exp) and round once, so they return the correctly rounded
// double except for arguments within ~2^-90 of a rounding boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace camp {

struct dd {
  double hi, lo;
};

__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__host__ __device__ __forceinline__ dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
__host__ __device__ __forceinline__ dd two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}
__host__ __device__ __forceinline__ dd dd_from(double a) { return {a, 0.0}; }
__host__ __device__ __forceinline__ double dd_round(dd a) { return a.hi + a.lo; }
__host__ __device__ __forceinline__ dd dd_mul_d(dd a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = fma(a.lo, b, e);
  return fast_two_sum(p, e);
}
__host__ __device__ __forceinline__ dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  double lo = s.lo + (a.lo + b.lo);
  return fast_two_sum(s.hi, lo);
}
__host__ __device__ __forceinline__ dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo = fma(a.hi, b.lo, p.lo);
  p.lo = fma(a.lo, b.hi, p.lo);
  return fast_two_sum(p.hi, p.lo);
}
__host__ __device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
__host__ __device__ __forceinline__ dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
// a / b (two Newton corrections of the double quotient)
__host__ __device__ __forceinline__ dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul(b, dd_from(q1)));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul(b, dd_from(q2)));
  double q3 = r.hi / b.hi;
  dd q = fast_two_sum(q1, q2);
  return dd_add(q, dd_from(q3));
}
__host__ __device__ __forceinline__ dd dd_div_d(dd a, double b) { return dd_div(a, dd_from(b)); }
__host__ __device__ __forceinline__ bool dd_gt(dd a, double b) {
  return a.hi > b || (a.hi == b && a.lo > 0.0);
}
__host__ __device__ __forceinline__ dd dd_ldexp(dd a, int e) { return {ldexp(a.hi, e), ldexp(a.lo, e)}; }

// ln 2 and ln 10 to 106 bits, and 1/n! (n = 3..10), each as hi + lo
constexpr double kLn2Hi = 0.6931471805599453, kLn2Lo = 2.3190468138462996e-17;
constexpr double kLn10Hi = 2.302585092994046, kLn10Lo = -2.1707562233822494e-16;

__host__ __device__ __forceinline__ dd inv_factorial(int n) {
  switch (n) {
    case 1: return {1.0, 0.0};
    case 2: return {0.5, 0.0};
    case 3: return {0.16666666666666666, 9.25185853854297e-18};
    case 4: return {0.041666666666666664, 2.3129646346357427e-18};
    case 5: return {0.008333333333333333, 1.1564823173178714e-19};
    case 6: return {0.001388888888888889, -5.300543954373577e-20};
    case 7: return {0.0001984126984126984, 1.7209558293420705e-22};
    case 8: return {2.48015873015873e-05, 2.1511947866775882e-23};
    case 9: return {2.7557319223985893e-06, -1.858393274046472e-22};
    default: return {2.755731922398589e-07, 2.3767714622250297e-23};
  }
}

// e^x for a double-double x with |x.hi| <= 709: returns (e^x) / 2^k as
// 1 + u (u in double-double) and k, so that callers round once.
// x = k ln2 + r, |r| <= ln2/2; e^(r/512) - 1 by its Taylor series to
// degree 10 (truncation < 2^-110), then nine squarings of (1 + u)
// carried as u <- 2u + u^2.
__host__ __device__ inline __noinline__ dd dd_expm1_reduced(dd x, int* k_out) {
  double k = rint(x.hi / kLn2Hi);
  dd r = dd_sub(x, dd_mul_d(dd{kLn2Hi, kLn2Lo}, k));
  r = dd_ldexp(r, -9);
  dd p = inv_factorial(10);
  for (int j = 9; j >= 1; --j) p = dd_add(dd_mul(p, r), inv_factorial(j));
  dd u = dd_mul(p, r);
  for (int s = 0; s < 9; ++s) u = dd_add(dd_ldexp(u, 1), dd_mul(u, u));
  *k_out = (int)k;
  return u;
}

__host__ __device__ inline __noinline__ dd dd_exp(dd x) {
  int k;
  dd u = dd_expm1_reduced(x, &k);
  return dd_ldexp(dd_add(dd_from(1.0), u), k);
}

// correctly rounded e^x
__host__ __device__ inline __noinline__ double cr_exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.8) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  if (x == 0.0) return 1.0;
  int k;
  dd u = dd_expm1_reduced(dd_from(x), &k);
  return ldexp(dd_round(dd_add(dd_from(1.0), u)), k);
}

// natural log of a positive finite double, to ~2^-100: x = m 2^e, then
// one Newton step y = y0 + m e^(-y0) - 1 (minus half its square) on the
// library's log(m), whose error is squared away
__host__ __device__ inline __noinline__ dd dd_log(double x) {
  int e = 0;
  double m = x;
  if (!(x > 0.5 && x < 2.0)) m = frexp(x, &e);
  double y0 = log(m);
  dd t = dd_mul_d(dd_exp(dd_from(-y0)), m);
  dd d = dd_sub(t, dd_from(1.0));
  d = dd_sub(d, dd_mul_d(dd_mul(d, d), 0.5));
  dd y = dd_add(dd_from(y0), d);
  if (e != 0) y = dd_add(y, dd_mul_d(dd{kLn2Hi, kLn2Lo}, (double)e));
  return y;
}

__host__ __device__ inline __noinline__ double cr_log(double x) {
  if (!(x > 0.0) || x == HUGE_VAL) return log(x);
  if (x == 1.0) return 0.0;
  return dd_round(dd_log(x));
}

__host__ __device__ inline __noinline__ double cr_log10(double x) {
  if (!(x > 0.0) || x == HUGE_VAL) return log10(x);
  if (x == 1.0) return 0.0;
  return dd_round(dd_div(dd_log(x), dd{kLn10Hi, kLn10Lo}));
}

// correctly rounded x^n for a small non-negative integer n (the integer
// branch of cr_pow, inline for the aerosol code's per-evaluation powers)
__host__ __device__ inline double cr_pow(double x, double y);
__host__ __device__ __forceinline__ double cr_powi(double x, int n) {
  if (n < 0) return cr_pow(x, (double)n);
  if (n == 0) return 1.0;
  if (n == 1) return x;
  if (n == 2) return x * x;
  double ax = fabs(x);
  dd p = dd_from(ax);
  for (int i = 1; i < n; ++i) p = dd_mul_d(p, ax);
  double r = dd_round(p);
  return (x < 0.0 && (n & 1)) ? -r : r;
}

// correctly rounded x^y for the cases the restated formulas meet: any y
// for x > 0, integer y for x < 0, and the exact special cases
__host__ __device__ inline __noinline__ double cr_pow(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  bool yint = y == rint(y);
  if (x == 0.0 || !(x == x) || !(y == y) || fabs(x) == HUGE_VAL) return pow(x, y);
  if (x < 0.0 && !yint) return pow(x, y);
  double ax = fabs(x);
  bool neg = x < 0.0 && yint && fmod(fabs(y), 2.0) == 1.0;
  double r;
  if (yint && fabs(y) <= 16.0) {
    int n = (int)fabs(y);
    dd p = dd_from(ax);
    for (int i = 1; i < n; ++i) p = dd_mul_d(p, ax);
    r = y > 0.0 ? dd_round(p) : dd_round(dd_div(dd_from(1.0), p));
  } else {
    dd z = dd_mul_d(dd_log(ax), y);
    if (z.hi > 709.8) return neg ? -HUGE_VAL : HUGE_VAL;
    if (z.hi < -745.2) return neg ? -0.0 : 0.0;
    int k;
    dd u = dd_expm1_reduced(z, &k);
    r = ldexp(dd_round(dd_add(dd_from(1.0), u)), k);
  }
  return neg ? -r : r;
}

}  // namespace camp
