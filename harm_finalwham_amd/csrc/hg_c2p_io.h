// Device helpers shared by the batched inversion kernels (hg_c2p_batch.hip,
// hg_mlp.hip): loading and storing one zone of a structure-of-arrays batch,
// the compacted list of zones left to the reference path, and the
// mixed-precision inversion of one zone.
#pragma once
#include <hip/hip_runtime.h>

#include "hg_c2p.h"
#include "hg_host.h"

namespace hg {

// one zone of a batch: the symmetric metric from its upper triangle (the
// reference's metric tables are symmetric), gdet, U, the guess's u and u~^i
__device__ __forceinline__ void load_batch_zone(int64_t n, int64_t t, const double *__restrict__ U,
                                                const double *__restrict__ gcov, const double *__restrict__ gcon,
                                                const double *__restrict__ gdet, const double *__restrict__ pr,
                                                Geom &G, double *u, double *p) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = a; c < 4; c++) {
      G.gcov[a][c] = G.gcov[c][a] = gcov[(a * 4 + c) * n + t];
      G.gcon[a][c] = G.gcon[c][a] = gcon[(a * 4 + c) * n + t];
    }
  G.gpert0 = 0.;
  G.g = gdet[t];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) u[pl] = U[pl * n + t];
  p[RHO] = 0.;
#pragma unroll
  for (int pl = UU; pl <= U3; pl++) p[pl] = pr[pl * n + t];
#pragma unroll
  for (int pl = B1; pl <= B3; pl++) p[pl] = 0.;
}

// the solution: the field always, rho .. u~^3 where the inversion assigned
// them (elsewhere the guess stays in place, as the reference leaves it)
__device__ __forceinline__ void store_batch_zone(int64_t n, int64_t t, double *__restrict__ pr, const double *p,
                                                 bool assigned) {
  if (assigned)
#pragma unroll
    for (int pl = RHO; pl <= U3; pl++) pr[pl * n + t] = p[pl];
#pragma unroll
  for (int pl = B1; pl <= B3; pl++) pr[pl * n + t] = p[pl];
}

// the reference assigns rho .. u~^3 for codes 0 and the sign failures
// 9, 10, 11 (utoprim_2d_final.c:418-431), not for the others
__device__ __forceinline__ bool code_assigns(int code) {
  return code == 0 || code == FAIL_RHONEG || code == FAIL_UNEG || code == FAIL_RHOUNEG;
}

// lane l of the wavefront appends zone t to the list when fb is set (one
// atomic per wavefront)
__device__ __forceinline__ void list_append(bool fb, int64_t t, unsigned long long *ctr, uint32_t *list) {
  const int lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(fb);
  if (!m) return;
  const int first = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == first) base = atomicAdd(&ctr[C2P_LISTLEN], (unsigned long long)__popcll(m));
  base = __shfl(base, first, 64);
  if (fb) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)t;
}

// the FP32 bulk, the polish and the primitives of one zone whose set-up
// succeeded; false: the zone goes to the list
__device__ __forceinline__ bool mixed_zone(const C2PInv &c, double igr, double x0, double x1, double *p, int &it,
                                           int &flag) {
  if (!newton_f32(c, igr, x0, x1, it)) return false;
  if (!c2p_polish(c, igr, x0, x1, POLISH_MAX_F64, it)) return false;
  // the primitives from the converged W with the reference's FP64 formulas
  // (utoprim_2d_final.c:335-441): rounding of 1 - vsq moves them by at most
  // ~1e-12 of the double-double values HG_C2P_REFINED forms
  double out[NPR];
  bool assign;
  flag = c2p_finish(c, x0, 0, igr, out, assign);
  if (assign)
#pragma unroll
    for (int i = 0; i < B1; i++) p[i] = out[i];
  return true;
}

// ---------------------------------------------------------------------------
// features of one inversion from its FP64 set-up (scale free: ratios to D)
// ---------------------------------------------------------------------------
// (ratios formed in FP64, logarithms in FP32: the network sees ~16 bits)
__device__ __forceinline__ void c2p_features(const C2PInv &c, double W_last, double alpha, double gam, float *x) {
  const double tiny = 1e-300;
  double D = c.D;
  double E = fmax(-c.Qdotn, tiny);
  double iD = 1.0 / D;
  x[0] = logf((float)(E * iD));                               // energy per rest mass
  x[1] = log1pf((float)(c.Qtsq / (E * E)));                   // momentum^2 / energy^2
  x[2] = log1pf((float)(c.Bsq * iD));                         // magnetisation
  x[3] = (float)(c.QdotBsq / (c.Bsq * c.Qtsq + tiny));        // cos^2 (S, B)
  x[4] = log10f((float)D);                                    // scale
  x[5] = (float)alpha;                                        // lapse
  x[6] = logf((float)(fmax(W_last, tiny) * iD));              // the plain guess
  x[7] = (float)gam;                                          // adiabatic index
}

// the FP64 set-up shared by features and the learned inversion
__device__ __forceinline__ int c2p_setup_zone(const double *U, const Geom &G, double gam, const double *guess,
                                          C2PInv &c, double &W_last, double &alpha, double *Ug) {
  double Ut[NPR];
#pragma unroll
  for (int i = 0; i < NPR; i++) Ug[i] = U[i] / G.g;
  Ug[UU] -= Ug[RHO];
  alpha = 1.0 / sqrt(-G.gcon[0][0]);
#pragma unroll
  for (int i = 0; i < NPR; i++) Ut[i] = alpha * Ug[i];
  return c2p_setup(Ut, G, guess, gam, c, W_last);
}

// The start of the mixed iteration from the network's prediction y =
// log(W / W_guess): the plain start from the previous primitives unless
// |y| > LEARNED_MIN_CORR (the network's error on zones the plain start suits
// is well below that); the learned start W_guess e^y is then checked by the
// FP64 Newton step its residual gives, and taken, with that step as the
// first Newton step (the reference loop's guards on W and vsq), when the
// step is finite and moves W by at most LEARNED_MAX_STEP of itself.
// Returns the evaluations spent (0 or 1); *used tells whether the learned
// start was taken.
constexpr float LEARNED_MIN_CORR = 0.25f;
constexpr double LEARNED_MAX_STEP = 0.5;
__device__ __forceinline__ int learned_start(const C2PInv &c, double igr, double W_last, float y, double &x0,
                                             double &x1, bool &used) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  const bool tried = fabsf(y) > LEARNED_MIN_CORR;
  const double Wm = W_last * (double)__expf(y);
  bool cand = tried && isfinite(Wm) && Wm > 0.;
  double d0 = 0., d1 = 0., fm = 0., dfm = 0., xm1 = 0.;
  if (cand) {
    const double vm = vsq_of_W(c, Wm);
    xm1 = (vm > 1.) ? vsq_big : vm;
    newton_dx(c, igr, Wm, xm1, d0, d1, fm, dfm);
    cand = isfinite(d0) && isfinite(d1) && fabs(d0) <= LEARNED_MAX_STEP * Wm;
  }
  if (cand) {
    x0 = fabs(Wm + d0);
    x0 = (x0 / c.w > gsq_big) ? Wm : x0;
    x1 = xm1 + d1;
    x1 = (x1 > 1.) ? vsq_big : x1;
  } else {
    newton_start(c, W_last, x0, x1);
  }
  used = cand;
  return tried ? 1 : 0;
}

}  // namespace hg
