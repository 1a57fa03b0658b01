// Double-double arithmetic: a value is the unevaluated sum hi + lo of two
// doubles with |lo| <= ulp(hi) / 2, about 106 bits of significand.
//
// Built from the error-free transformations: Knuth's two-sum (any a, b),
// Dekker's fast two-sum (|a| >= |b|) and the exact product a * b = p + e with
// e = fma(a, b, -p).  The addition is the accurate one (both halves summed
// with error-free transformations), so cancellation between the operands
// keeps full relative precision; multiplication, division and square root
// are the usual one-correction forms (Dekker 1971; Hida, Li and Bailey,
// "Library for double-double and quad-double arithmetic", 2007).
//
// The library compiles with -ffp-contract=off, so every operation below is
// evaluated as written; fma() is an explicit fused multiply-add.
#pragma once
#include <hip/hip_runtime.h>

namespace hg {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dd_of(double a) { return dd{a, 0.0}; }
__device__ __forceinline__ double dd_val(dd a) { return a.hi + a.lo; }

__device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ dd fast_two_sum(double a, double b) {
  double s = a + b;
  return dd{s, b - (s - a)};
}

__device__ __forceinline__ dd two_prod(double a, double b) {
  double p = a * b;
  return dd{p, fma(a, b, -p)};
}

__device__ __forceinline__ dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }

__device__ __forceinline__ dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}

// reciprocal of b to about one ulp: the hardware estimate and one Newton step
__device__ __forceinline__ double rcp_1ulp(double b) {
  double r = __builtin_amdgcn_rcp(b);
  return fma(r, fma(-b, r, 1.0), r);
}

// a / b: three quotient terms, each from the remainder of the previous ones
__device__ __forceinline__ dd dd_div(dd a, dd b) {
  const double r = rcp_1ulp(b.hi);
  double q1 = a.hi * r;
  dd x = dd_sub(a, dd_mul_d(b, q1));
  double q2 = x.hi * r;
  x = dd_sub(x, dd_mul_d(b, q2));
  double q3 = x.hi * r;
  dd q = fast_two_sum(q1, q2);
  return dd_add(q, dd_of(q3));
}

// sqrt(a) for a > 0: one correction of the double square root
__device__ __forceinline__ dd dd_sqrt(dd a) {
  double s = sqrt(a.hi);
  dd e = dd_sub(a, two_prod(s, s));
  return fast_two_sum(s, e.hi * (0.5 * rcp_1ulp(s)));
}

}  // namespace hg
