// Correctly rounded FP64 division with the reciprocal part of the
// hardware's sequence shared by several numerators of one divisor.
#pragma once
#include <hip/hip_runtime.h>

namespace hg {

// Division through the steps of the hardware's correctly rounded FP64
// division with the reciprocal part done once per divisor.  hipcc lowers
// x / d to v_div_scale of d and of x, v_rcp_f64, two Newton steps on the
// reciprocal, the quotient, its residual, v_div_fmas and v_div_fixup.  For
// a divisor of magnitude in [2^-300, 2^300] (div_den_ok) and a numerator
// that is zero or of magnitude in [2^-660, 2^300] (div_num_ok), v_div_scale
// changes neither operand nor sets the flag that makes v_div_fmas rescale,
// and v_div_fixup returns the corrected quotient: rcp_refined and div_core
// are then the whole sequence and give x / d bit for bit, except that a zero
// numerator gives +0 for both signs.  Callers check the ranges and divide
// in full otherwise; hg_check_division compares both ways on the device.
__device__ __forceinline__ double rcp_refined(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double div_core(double x, double d, double y) {
  const double q = x * y;
  const double r = fma(-d, q, x);
  return fma(r, y, q);
}
__device__ __forceinline__ bool div_num_ok(double x) {
  const double a = fabs(x);
  return a == 0.0 || (a >= 0x1p-660 && a <= 0x1p300);
}

__device__ __forceinline__ bool div_den_ok(double d) { return d >= 0x1p-300 && d <= 0x1p300; }

// x[k] / d for K numerators of one positive divisor, none of them -0.  SD
// false: K full divisions.  The shuffle kernel along i runs faster with the
// shared reciprocal, the marching kernels (whose loop the range checks and
// their branches lengthen) without.
template <int K, bool SD>
__device__ __forceinline__ void div_shared(double *x, double d) {
  if (!SD) {
#pragma unroll
    for (int k = 0; k < K; k++) x[k] /= d;
    return;
  }
  const double y = rcp_refined(d);
  double q[K];
  bool ok = div_den_ok(d);
#pragma unroll
  for (int k = 0; k < K; k++) {
    q[k] = div_core(x[k], d, y);
    ok = ok && div_num_ok(x[k]);
  }
  if (ok) {
#pragma unroll
    for (int k = 0; k < K; k++) x[k] = q[k];
  } else {
#pragma unroll
    for (int k = 0; k < K; k++) x[k] /= d;
  }
}

// the mixed-sign numerators n[0..3] of one constant divisor d, whose
// reciprocal yd = rcp_refined(d) the caller forms once: q[k] = n[k] / d
// bit for bit, except that a zero quotient is +0 for both signs of n[k]
__device__ __forceinline__ void div_const4(const double *n, double d, double yd, double *q) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    q[k] = div_core(n[k], d, yd);
    ok = ok && div_num_ok(n[k]);
  }
  if (!ok) {
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = n[k] / d;
  }
}

}  // namespace hg
