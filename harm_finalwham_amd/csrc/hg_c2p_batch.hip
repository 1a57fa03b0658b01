// Batched con2prim on device memory (utoprimgen.c:23 over independent
// zones): the generator of the benchmark distribution and the inversion
// kernels of hg_utoprim_batch.
//
//   k_make_c2p_batch  one thread per inversion: Kerr-Schild metric at a random
//                     (r, theta), random primitives, their conserved variables
//                     (primtoU, FP64), a perturbed guess
//   k_c2p_batch       one thread per zone: the reference inversion
//                     (HG_C2P_FP64), or the reference inversion with its root
//                     converged and the primitives formed in double-double
//                     (HG_C2P_REFINED)
//   k_c2p_mixed       persistent work queue: each wavefront takes 4 x 64
//                     zones at a time from a device counter; per lane the
//                     FP64 set-up, the FP32 Newton iteration, the polish and
//                     the double-double primitives; a zone that does not
//                     converge joins a compacted list instead
//   k_c2p_listed      persistent over that list: the HG_C2P_REFINED
//                     inversion from the original guess
//
// Memory per zone: the inversion reads U (8), the upper triangles of the
// symmetric gcov and gcon (10 + 10), gdet and the guess's u and u~^i (4),
// and writes the field (3), the solved rho, u, u~^i (5, only where the
// reference assigns them), the failure code and the iteration count:
// 336 bytes.  gcovpert and the guess's rho and field are not read (the
// inversion does not use them).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <mutex>

#include "harmgpu.h"
#include "hg_c2p.h"
#include "hg_c2p_io.h"
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

template <bool REFINED>
__global__ void __launch_bounds__(256) k_c2p_batch(int64_t n, double gam, const double *__restrict__ U,
                                                   const double *__restrict__ gcov, const double *__restrict__ gcon,
                                                   const double *__restrict__ gdet, double *__restrict__ pr,
                                                   int *__restrict__ pflag, int *__restrict__ iters) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  Geom G;
  double u[NPR], p[NPR];
  load_batch_zone(n, t, U, gcov, gcon, gdet, pr, G, u, p);
  int it = 0, flag;
  if (REFINED) {
    bool pol;
    flag = utoprimgen_refined(u, G, gam, p, it, pol);
  } else {
    flag = utoprimgen_fp64(u, G, gam, p, it);
  }
  store_batch_zone(n, t, pr, p, code_assigns(flag));
  pflag[t] = flag;
  if (iters) iters[t] = it;
}

__global__ void k_c2p_reset(unsigned long long *ctr) {
  if (threadIdx.x < C2P_NCTR) ctr[threadIdx.x] = 0ull;
}

constexpr int MIXED_ROUNDS = 4;  // 64-zone rounds per work-queue grab

__global__ void __launch_bounds__(256) k_c2p_mixed(int64_t n, double gam, const double *__restrict__ U,
                                                   const double *__restrict__ gcov,
                                                   const double *__restrict__ gcon,
                                                   const double *__restrict__ gdet, double *__restrict__ pr,
                                                   int *__restrict__ pflag, int *__restrict__ iters,
                                                   unsigned long long *__restrict__ ctr,
                                                   uint32_t *__restrict__ list) {
  const int lane = threadIdx.x & 63;
  const double igr = (gam - 1.0) / gam;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&ctr[C2P_WORK], (unsigned long long)(64 * MIXED_ROUNDS));
    base = __shfl(base, 0, 64);
    if (base >= (unsigned long long)n) break;
    for (int r = 0; r < MIXED_ROUNDS; r++) {
      const int64_t t = (int64_t)base + 64 * r + lane;
      bool fb = false;
      if (t < n) {
        Geom G;
        double u[NPR], p[NPR], Ug[NPR], Ut[NPR];
        load_batch_zone(n, t, U, gcov, gcon, gdet, pr, G, u, p);
#pragma unroll
        for (int i = 0; i < NPR; i++) Ug[i] = u[i] / G.g;
        Ug[UU] -= Ug[RHO];
#pragma unroll
        for (int i = B1; i <= B3; i++) p[i] = Ug[i];
        const double alpha = 1.0 / sqrt(-G.gcon[0][0]);
#pragma unroll
        for (int i = 0; i < NPR; i++) Ut[i] = alpha * Ug[i];
        C2PInv c;
        double W_last;
        int it = 0, flag = c2p_setup(Ut, G, p, gam, c, W_last);
        if (flag == 0) {
          double x0, x1;
          newton_start(c, W_last, x0, x1);
          fb = !mixed_zone(c, igr, x0, x1, p, it, flag);
        }
        if (!fb) {
          store_batch_zone(n, t, pr, p, code_assigns(flag));
          pflag[t] = flag;
          if (iters) iters[t] = it;
        }
      }
      list_append(fb, t, ctr, list);
    }
  }
}

// the listed zones: HG_C2P_REFINED from the guess still in pr
__global__ void __launch_bounds__(256) k_c2p_listed(int64_t n, double gam, const double *__restrict__ U,
                                                    const double *__restrict__ gcov,
                                                    const double *__restrict__ gcon,
                                                    const double *__restrict__ gdet, double *__restrict__ pr,
                                                    int *__restrict__ pflag, int *__restrict__ iters,
                                                    unsigned long long *__restrict__ ctr,
                                                    const uint32_t *__restrict__ list) {
  const int lane = threadIdx.x & 63;
  const unsigned long long len = ctr[C2P_LISTLEN];
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&ctr[C2P_LISTWORK], 64ull);
    base = __shfl(base, 0, 64);
    if (base >= len) break;
    if (base + lane < len) {
      const int64_t t = list[base + lane];
      Geom G;
      double u[NPR], p[NPR];
      load_batch_zone(n, t, U, gcov, gcon, gdet, pr, G, u, p);
      int it = 0;
      bool pol;
      int flag = utoprimgen_refined(u, G, gam, p, it, pol);
      store_batch_zone(n, t, pr, p, code_assigns(flag));
      pflag[t] = flag;
      if (iters) iters[t] = it;
    }
  }
}

std::atomic<long long> g_last_fallbacks[64];  // per device: listed zones of the last mixed batch

// blocks of 256 threads that stay resident on the whole device
template <typename K>
static int persistent_blocks(K kernel, int64_t want, int64_t &blocks) {
  int dev = 0, ncu = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0);
  if (e != hipSuccess) {
    set_error("hg_utoprim_batch: %s", hipGetErrorString(e));
    return 1;
  }
  if (per_cu < 1) per_cu = 1;
  blocks = (int64_t)ncu * per_cu;
  if (blocks > want) blocks = want;
  if (blocks < 1) blocks = 1;
  return 0;
}

}  // namespace

namespace hg {

size_t c2p_scratch_bytes(int64_t n) { return C2P_NCTR * sizeof(unsigned long long) + (size_t)n * sizeof(uint32_t); }

int launch_c2p_batch(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                     const double *gdet, double *pr, int *pflag, int *iters, bool refined, hipStream_t s) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (refined)
    hipLaunchKernelGGL(k_c2p_batch<true>, dim3(nb), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, pr, pflag, iters);
  else
    hipLaunchKernelGGL(k_c2p_batch<false>, dim3(nb), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, pr, pflag, iters);
  return 0;
}

int launch_c2p_batch_mixed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                           const double *gdet, double *pr, int *pflag, int *iters, void *scratch, hipStream_t s) {
  unsigned long long *ctr = (unsigned long long *)scratch;
  uint32_t *list = (uint32_t *)(ctr + C2P_NCTR);
  int64_t b1 = 0, b2 = 0;
  if (persistent_blocks(k_c2p_mixed, (n + 255) / 256, b1)) return 1;
  if (persistent_blocks(k_c2p_listed, (n + 255) / 256, b2)) return 1;
  hipLaunchKernelGGL(k_c2p_reset, dim3(1), dim3(64), 0, s, ctr);
  hipLaunchKernelGGL(k_c2p_mixed, dim3((unsigned)b1), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, pr, pflag, iters,
                     ctr, list);
  hipLaunchKernelGGL(k_c2p_listed, dim3((unsigned)b2), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, pr, pflag,
                     iters, ctr, (const uint32_t *)list);
  return 0;
}

void launch_c2p_reset(void *scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_c2p_reset, dim3(1), dim3(64), 0, s, (unsigned long long *)scratch);
}

int launch_c2p_listed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                      const double *gdet, double *pr, int *pflag, int *iters, void *scratch, hipStream_t s) {
  unsigned long long *ctr = (unsigned long long *)scratch;
  int64_t b = 0;
  if (persistent_blocks(k_c2p_listed, (n + 255) / 256, b)) return 1;
  hipLaunchKernelGGL(k_c2p_listed, dim3((unsigned)b), dim3(256), 0, s, n, gam, U, gcov, gcon, gdet, pr, pflag, iters,
                     ctr, (const uint32_t *)(ctr + C2P_NCTR));
  return 0;
}

void c2p_note_fallbacks(long long nfb) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) g_last_fallbacks[dev] = nfb;
}

}  // namespace hg

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

// the list holds 32-bit zone indices
static int check_batch_args(const char *who, int64_t n, const double *U, const double *gcov, const double *gcon,
                            const double *gdet, const double *pr, const int32_t *pflag) {
  if (n < 0) { set_error("%s: n < 0", who); return 1; }
  if (n > (int64_t)0xFFFFFFFFll) { set_error("%s: n > 2^32 - 1", who); return 1; }
  if (n > 0 && (!U || !gcov || !gcon || !gdet || !pr || !pflag)) {
    set_error("%s: null argument", who);
    return 1;
  }
  return 0;
}

extern "C" int hg_utoprim_batch(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                                const double *gdet, const double *gcovpert, double *pr, int32_t *pflag,
                                int32_t *iters, int mode, void *stream) {
  (void)gcovpert;
  if (check_batch_args("hg_utoprim_batch", n, U, gcov, gcon, gdet, pr, pflag)) return 1;
  if (mode != HG_C2P_FP64 && mode != HG_C2P_MIXED && mode != HG_C2P_REFINED) {
    set_error("hg_utoprim_batch: unknown mode %d", mode);
    return 1;
  }
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (mode != HG_C2P_MIXED) {
    launch_c2p_batch(n, gam, U, gcov, gcon, gdet, pr, pflag, iters, mode == HG_C2P_REFINED, s);
  } else {
    // scratch of this call only, stream-ordered: concurrent calls on other
    // streams have their own counters and lists
    void *scr = nullptr;
    e = hipMallocAsync(&scr, c2p_scratch_bytes(n), s);
    if (e != hipSuccess) { set_error("hg_utoprim_batch: scratch: %s", hipGetErrorString(e)); return 1; }
    int rc = launch_c2p_batch_mixed(n, gam, U, gcov, gcon, gdet, pr, pflag, iters, scr, s);
    unsigned long long ctr[C2P_NCTR] = {0, 0, 0};
    if (!rc) e = hipMemcpyAsync(ctr, scr, sizeof ctr, hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipFreeAsync(scr, s);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (rc) return 1;
    if (e == hipSuccess) c2p_note_fallbacks((long long)ctr[C2P_LISTLEN]);
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { set_error("hg_utoprim_batch: %s", hipGetErrorString(e)); return 1; }
  return 0;
}

// hg_utoprim_batch_noalloc: the scratch of HG_C2P_MIXED (counters, then the
// 32-bit list) in a static device array; host code uses only the per-device
// address returned by hipGetSymbolAddress, never the array's name
static __device__ __align__(256) unsigned char scratch[C2P_NCTR * sizeof(unsigned long long) +
                                                       (size_t)HG_C2P_NOALLOC_MAX_N * sizeof(uint32_t)];
static void *g_scratch_ptr[64];
static std::mutex g_scratch_mu;

static int noalloc_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { set_error("hg_utoprim_batch_noalloc: %s", hipGetErrorString(e)); return 1; }
  if (dev < 0 || dev >= 64) { set_error("hg_utoprim_batch_noalloc: device %d out of range", dev); return 1; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("hg_utoprim_batch_noalloc: scratch address: %s", hipGetErrorString(e));
      return 1;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" int hg_utoprim_batch_noalloc(int64_t n, double gam, const double *U, const double *gcov,
                                        const double *gcon, const double *gdet, const double *gcovpert,
                                        double *pr, int32_t *pflag, int32_t *iters, int mode, void *stream) {
  (void)gcovpert;
  if (check_batch_args("hg_utoprim_batch_noalloc", n, U, gcov, gcon, gdet, pr, pflag)) return 1;
  if (mode != HG_C2P_FP64 && mode != HG_C2P_MIXED && mode != HG_C2P_REFINED) {
    set_error("hg_utoprim_batch_noalloc: unknown mode %d", mode);
    return 1;
  }
  if (n > HG_C2P_NOALLOC_MAX_N) {
    set_error("hg_utoprim_batch_noalloc: n = %lld exceeds the static scratch (%lld zones)", (long long)n,
              (long long)HG_C2P_NOALLOC_MAX_N);
    return 1;
  }
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (mode != HG_C2P_MIXED) {
    launch_c2p_batch(n, gam, U, gcov, gcon, gdet, pr, pflag, iters, mode == HG_C2P_REFINED, s);
  } else {
    void *scr = nullptr;
    if (noalloc_scratch(&scr)) return 1;
    // the counters are cleared by k_c2p_reset inside the launch sequence
    if (launch_c2p_batch_mixed(n, gam, U, gcov, gcon, gdet, pr, pflag, iters, scr, s)) return 1;
  }
  // under stream capture the launches are only recorded: no wait, and the
  // fallback count (hg_c2p_fallbacks) is not updated
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  e = hipStreamIsCapturing(s, &cs);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess && cs == hipStreamCaptureStatusNone) {
    if (mode == HG_C2P_MIXED) {
      unsigned long long ctr[C2P_NCTR] = {0, 0, 0};
      void *scr = nullptr;
      if (noalloc_scratch(&scr)) return 1;
      e = hipMemcpyAsync(ctr, scr, sizeof ctr, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e == hipSuccess) c2p_note_fallbacks((long long)ctr[C2P_LISTLEN]);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) { set_error("hg_utoprim_batch_noalloc: %s", hipGetErrorString(e)); return 1; }
  return 0;
}

extern "C" int hg_c2p_fallbacks(int64_t *nfallback) {
  if (!nfallback) { set_error("hg_c2p_fallbacks: null argument"); return 1; }
  int dev = 0;
  *nfallback = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  *nfallback = (int64_t)g_last_fallbacks[dev].load();
  return 0;
}
