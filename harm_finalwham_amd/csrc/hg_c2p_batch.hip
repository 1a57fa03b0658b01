// Batched con2prim on device memory: the generator of the benchmark
// distribution and the mixed-precision persistent inversion kernel.
//
//   k_make_c2p_batch  one thread per inversion: Kerr-Schild metric at a random
//                     (r, theta), random primitives, their conserved variables
//                     (primtoU, FP64), a perturbed guess
//   k_c2p_mixed       persistent grid; each wavefront takes 64 inversions at
//                     a time from a device counter until the batch is done,
//                     so wavefronts whose zones need more Newton steps do not
//                     hold up the rest (utoprimgen_mixed per lane)
#include <hip/hip_runtime.h>

#include <cmath>

#include "harmgpu.h"
#include "hg_c2p.h"
#include "hg_host.h"

using namespace hg;

namespace {

// counter-based random numbers: item idx, stream s -> uniform in [0, 1)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double urand(uint64_t seed, uint64_t idx, int s) {
  uint64_t x = mix64(seed ^ mix64(idx * 64ull + (uint64_t)s));
  return (double)(x >> 11) * 0x1.0p-53;
}
// Box-Muller from streams s, s+1
__device__ __forceinline__ double nrand(uint64_t seed, uint64_t idx, int s) {
  double u1 = 1.0 - urand(seed, idx, s);  // (0, 1]
  double u2 = urand(seed, idx, s + 1);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

// spatial norm with the covariant metric
__device__ __forceinline__ double snorm2(const double g[4][4], const double *v) {
  double s = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) s += g[i][j] * v[i - 1] * v[j - 1];
  return s;
}

__global__ void __launch_bounds__(256) k_make_c2p_batch(int64_t n, double gam, double a, uint64_t seed,
                                                        double *__restrict__ U, double *__restrict__ gcov,
                                                        double *__restrict__ gcon, double *__restrict__ gdet,
                                                        double *__restrict__ gcovpert, double *__restrict__ p_true,
                                                        double *__restrict__ p_guess) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  // Kerr-Schild metric (metric.c:697, :863) at r U[1.3,40], theta U(0.01, pi-0.01)
  double r = 1.3 + (40.0 - 1.3) * urand(seed, t, 0);
  double th = 0.01 + (M_PI - 0.02) * urand(seed, t, 1);
  double s = sin(th), c = cos(th), s2 = s * s;
  double rho2 = r * r + a * a * c * c;
  Geom G;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) G.gcov[i][j] = G.gcon[i][j] = 0.;
  G.gcov[0][0] = -1.0 + 2.0 * r / rho2;
  G.gcov[0][1] = G.gcov[1][0] = 2.0 * r / rho2;
  G.gcov[0][3] = G.gcov[3][0] = -2.0 * a * r * s2 / rho2;
  G.gcov[1][1] = 1.0 + 2.0 * r / rho2;
  G.gcov[1][3] = G.gcov[3][1] = -a * s2 * (1.0 + 2.0 * r / rho2);
  G.gcov[2][2] = rho2;
  G.gcov[3][3] = s2 * (rho2 + a * a * s2 * (1.0 + 2.0 * r / rho2));
  G.gcon[0][0] = -(1.0 + 2.0 * r / rho2);
  G.gcon[0][1] = G.gcon[1][0] = 2.0 * r / rho2;
  G.gcon[1][1] = (r * r - 2.0 * r + a * a) / rho2;
  G.gcon[1][3] = G.gcon[3][1] = a / rho2;
  G.gcon[2][2] = 1.0 / rho2;
  G.gcon[3][3] = 1.0 / (rho2 * s2);
  G.g = rho2 * s;
  G.gpert0 = 2.0 * r / rho2;
  // primitives: rho log-U[1e-6,1], u log-U[1e-8,1e2], Lorentz factor
  // log-U[1,50] with an isotropic direction, b^2/rho log-U[1e-4,1e3]
  double rho = pow(10.0, -6.0 + 6.0 * urand(seed, t, 2));
  double u = pow(10.0, -8.0 + 10.0 * urand(seed, t, 3));
  double lor = pow(10.0, log10(50.0) * urand(seed, t, 4));
  double sig = pow(10.0, -4.0 + 7.0 * urand(seed, t, 5));
  double ev[3] = {nrand(seed, t, 6), nrand(seed, t, 8), nrand(seed, t, 10)};
  double eb[3] = {nrand(seed, t, 12), nrand(seed, t, 14), nrand(seed, t, 16)};
  double nv = sqrt(snorm2(G.gcov, ev)), nb = sqrt(snorm2(G.gcov, eb));
  double ut[3];
  for (int i = 0; i < 3; i++) {
    ev[i] /= nv;
    eb[i] /= nb;
    ut[i] = sqrt(lor * lor - 1.0) * ev[i];
  }
  // b^2 of the unit-amplitude field (phys.c:849 get_state), rescaled to b^2/rho = sig
  double pr[NPR] = {rho, u, ut[0], ut[1], ut[2], eb[0], eb[1], eb[2]};
  State q;
  get_state(pr, G, q);
  double bsq1 = q.bcon[0] * q.bcov[0] + q.bcon[1] * q.bcov[1] + q.bcon[2] * q.bcov[2] + q.bcon[3] * q.bcov[3];
  double amp = sqrt(sig * rho / bsq1);
  for (int i = 0; i < 3; i++) pr[B1 + i] = amp * eb[i];
  double Uc[NPR];
  primtoU(pr, G, gam, Uc);
  for (int pl = 0; pl < NPR; pl++) {
    U[pl * n + t] = Uc[pl];
    p_true[pl * n + t] = pr[pl];
  }
  // guess: u and the velocities perturbed by 1e-2 N(0,1)
  double gu[NPR];
  for (int pl = 0; pl < NPR; pl++) gu[pl] = pr[pl];
  gu[UU] *= 1.0 + 1e-2 * nrand(seed, t, 18);
  gu[U1] *= 1.0 + 1e-2 * nrand(seed, t, 20);
  gu[U2] *= 1.0 + 1e-2 * nrand(seed, t, 22);
  gu[U3] *= 1.0 + 1e-2 * nrand(seed, t, 24);
  for (int pl = 0; pl < NPR; pl++) p_guess[pl * n + t] = gu[pl];
  for (int a4 = 0; a4 < 16; a4++) {
    gcov[a4 * n + t] = G.gcov[a4 / 4][a4 % 4];
    gcon[a4 * n + t] = G.gcon[a4 / 4][a4 % 4];
  }
  gdet[t] = G.g;
  gcovpert[t] = G.gpert0;
}

__global__ void k_reset2(unsigned long long *c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
}

__global__ void __launch_bounds__(256) k_c2p_mixed(int64_t n, double gam, const double *__restrict__ U,
                                                   const double *__restrict__ gcov,
                                                   const double *__restrict__ gcon,
                                                   const double *__restrict__ gdet,
                                                   const double *__restrict__ gcovpert, double *__restrict__ pr,
                                                   int *__restrict__ pflag, int *__restrict__ iters,
                                                   unsigned long long *__restrict__ queue) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&queue[0], 64ull);
    base = __shfl(base, 0, 64);
    if (base >= (unsigned long long)n) break;
    int64_t t = (int64_t)base + lane;
    bool fb = false;
    if (t < n) {
      Geom G;
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          G.gcov[a][c] = gcov[(a * 4 + c) * n + t];
          G.gcon[a][c] = gcon[(a * 4 + c) * n + t];
        }
      G.gpert0 = gcovpert[t];
      G.g = gdet[t];
      double u[NPR], p[NPR];
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) {
        u[pl] = U[pl * n + t];
        p[pl] = pr[pl * n + t];
      }
      int it = 0;
      int flag = utoprimgen_mixed(u, G, gam, p, it, fb);
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) pr[pl * n + t] = p[pl];
      pflag[t] = flag;
      if (iters) iters[t] = it;
    }
    unsigned long long nfb = __popcll(__ballot(fb));
    if (lane == 0 && nfb) atomicAdd(&queue[1], nfb);
  }
}

unsigned long long *g_queue[64];  // per device: {work counter, fallbacks of the last mixed batch}

}  // namespace

extern "C" int hg_make_c2p_batch(int64_t n, double gam, double a, uint64_t seed, double *U, double *gcov,
                                 double *gcon, double *gdet, double *gcovpert, double *p_true, double *p_guess,
                                 void *stream) {
  if (n < 0) { set_error("hg_make_c2p_batch: n < 0"); return 1; }
  if (n == 0) return 0;
  if (!U || !gcov || !gcon || !gdet || !gcovpert || !p_true || !p_guess) {
    set_error("hg_make_c2p_batch: null argument");
    return 1;
  }
  if (!(fabs(a) < 1.0)) { set_error("hg_make_c2p_batch: need |a| < 1"); return 1; }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_make_c2p_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, gam, a, seed, U,
                     gcov, gcon, gdet, gcovpert, p_true, p_guess);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { set_error("hg_make_c2p_batch: %s", hipGetErrorString(e)); return 1; }
  return 0;
}

namespace hg {
int launch_c2p_batch_mixed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                           const double *gdet, const double *gcovpert, double *pr, int *pflag, int *iters,
                           hipStream_t s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess && (dev < 0 || dev >= 64)) { set_error("hg_utoprim_batch: device index %d", dev); return 1; }
  if (e == hipSuccess && !g_queue[dev]) e = hipMalloc((void **)&g_queue[dev], 2 * sizeof(unsigned long long));
  int ncu = 0, per_cu = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_c2p_mixed, 256, 0);
  if (e != hipSuccess) { set_error("hg_utoprim_batch: %s", hipGetErrorString(e)); return 1; }
  if (per_cu < 1) per_cu = 1;
  int64_t want = (n + 255) / 256;
  int64_t blocks = (int64_t)ncu * per_cu;
  if (blocks > want) blocks = want;
  hipLaunchKernelGGL(k_reset2, dim3(1), dim3(64), 0, s, g_queue[dev]);
  hipLaunchKernelGGL(k_c2p_mixed, dim3((unsigned)blocks), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, gcovpert,
                     pr, pflag, iters, g_queue[dev]);
  return 0;
}
}  // namespace hg

extern "C" int hg_c2p_fallbacks(int64_t *nfallback) {
  if (!nfallback) { set_error("hg_c2p_fallbacks: null argument"); return 1; }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  *nfallback = 0;
  if (e != hipSuccess || dev < 0 || dev >= 64 || !g_queue[dev]) return 0;
  unsigned long long h[2];
  e = hipMemcpy(h, g_queue[dev], sizeof h, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { set_error("hg_c2p_fallbacks: %s", hipGetErrorString(e)); return 1; }
  *nfallback = (int64_t)h[1];
  return 0;
}
