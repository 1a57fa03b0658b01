// Device self-check of the shared-reciprocal division (hg_div.h) against the
// hardware's own x / d on generated operands: exponents over the whole range
// (zeros, denormals and the values outside the checked ranges included),
// random and structured mantissas, both signs, and the ranges' edges.
#include <hip/hip_runtime.h>

#include "harmgpu.h"
#include "hg_div.h"
#include "hg_host.h"

using namespace hg;

namespace {

__device__ __forceinline__ uint64_t mixh(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// an operand from hash h: the biased exponent anywhere, inside the divisor's
// range, around the numerator's range or at the edges of both; a random
// mantissa, one just below or just above a power of two, or a pattern
__device__ __forceinline__ double operand(uint64_t h, bool positive) {
  const uint64_t h2 = mixh(h);
  int e;
  switch ((h >> 60) & 3) {
    case 0: e = (int)((h >> 40) % 2047); break;
    case 1: e = 1023 - 300 + (int)((h >> 40) % 601); break;
    case 2: e = 1023 - 680 + (int)((h >> 40) % 1000); break;
    default: {
      const int edges[8] = {1023 - 300, 1023 + 300, 1023 - 660, 1023 + 299,
                            1023 - 301, 1023 - 661, 1023 + 301, 1};
      e = edges[(h >> 40) % 8] + (int)((h >> 36) & 1);
    }
  }
  uint64_t m;
  switch ((h >> 56) & 3) {
    case 0: m = h2 & 0xFFFFFFFFFFFFFull; break;
    case 1: m = 0xFFFFFFFFFFFFFull - (h2 & 0xFF); break;
    case 2: m = h2 & 0xFF; break;
    default: m = h2 & 0x5555555555555ull; break;
  }
  uint64_t bits = ((uint64_t)e << 52) | m;
  if ((h & 0x3FF) == 7) bits = 0;
  if (!positive && ((h >> 12) & 1)) bits |= 0x8000000000000000ull;
  return __longlong_as_double((long long)bits);
}

__device__ __forceinline__ bool same(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b) || (a != a && b != b);
}

// out: [0] mismatches of div_shared<3, true> (non-negative numerators, one
// positive divisor), [1] of 1 / d through div_shared<1, true>, [2] of
// div_const4 by 12 (mixed signs; a zero quotient may differ in sign only),
// [3] operand sets div_shared took its shortcut for
__global__ void k_check_div(int64_t n, uint64_t seed, unsigned long long *out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  unsigned long long bad = 0, badr = 0, badc = 0, fast = 0;
  if (t < n) {
    const uint64_t h = mixh(seed ^ mixh((uint64_t)t));
    const double d = fabs(operand(h, true));
    double x[3], ref[3];
    bool ok = div_den_ok(d);
    for (int k = 0; k < 3; k++) {
      x[k] = fabs(operand(mixh(h + 1 + k), true));
      ref[k] = x[k] / d;
      ok = ok && div_num_ok(x[k]);
    }
    fast = ok ? 1ull : 0ull;
    div_shared<3, true>(x, d);
    for (int k = 0; k < 3; k++) bad += !same(x[k], ref[k]);
    double r[1] = {1.0};
    div_shared<1, true>(r, d);
    badr += !same(r[0], 1.0 / d);
    double nn[4], q[4];
    for (int k = 0; k < 4; k++) nn[k] = operand(mixh(h + 11 + k), false);
    div_const4(nn, 12., rcp_refined(12.), q);
    for (int k = 0; k < 4; k++) {
      const double rq = nn[k] / 12.;
      badc += !(same(q[k], rq) || (q[k] == 0.0 && rq == 0.0));
    }
  }
  if (bad) atomicAdd(&out[0], bad);
  if (badr) atomicAdd(&out[1], badr);
  if (badc) atomicAdd(&out[2], badc);
  if (fast) atomicAdd(&out[3], fast);
}

}  // namespace

extern "C" int hg_check_division(int64_t n, uint64_t seed, int64_t result[4]) {
  if (n < 0 || n > ((int64_t)1 << 32) || !result) {
    set_error("hg_check_division: bad arguments");
    return 1;
  }
  unsigned long long *out = nullptr;
  hipError_t e = hipMalloc((void **)&out, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(out, 0, 4 * sizeof(unsigned long long));
  if (e == hipSuccess && n > 0)
    hipLaunchKernelGGL(k_check_div, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, seed, out);
  if (e == hipSuccess) e = hipGetLastError();
  unsigned long long h[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
  if (out) (void)hipFree(out);
  if (e != hipSuccess) {
    set_error("hg_check_division: %s", hipGetErrorString(e));
    return 1;
  }
  for (int k = 0; k < 4; k++) result[k] = (int64_t)h[k];
  return 0;
}
