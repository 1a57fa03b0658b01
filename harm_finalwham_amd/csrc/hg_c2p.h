// Device Noble et al. (2006) "2D" conserved-to-primitive inversion.
//
// utoprimgen.c:23 Utoprimgen -> utoprim_2d_final.c:34 Utoprim_2d_final ->
// utoprim_2d_final.c:152 Utoprim_new_body -> :573 find_root_2D_gen ->
// :736 general_newton_raphson (no line search) with :1189 func_vsq2 and the
// ideal-gas EOS of utoprim_jon_eos.c / eos.c.  FP64 throughout: same
// operations in the same order as the reference, hence the same iterates,
// iteration counts and failure codes.
#pragma once
#include "hg_phys.h"

namespace hg {

constexpr int MAX_NEWT_ITER = 200;                 // u2p_defs.h:38
constexpr double NEWT_TOL = 1E3 * DBL_EPS;         // u2p_defs.h:39
constexpr double MIN_NEWT_TOL = 1E-6;              // u2p_defs.h:47
constexpr double MAXNEGUTSQ = 1E-10;               // u2p_defs.h:123
constexpr double GAMMAFAIL = 100.0 * 25.0;         // initbase.c:471-476
constexpr double FAIL_VAL = 1.e30;

// Invariants of one inversion (the file statics of utoprim_2d_final.c:13-15)
struct C2PInv {
  double Bsq, QdotB, QdotBsq, Qtsq, Qdotn, D, w;
  double Qcon[4], ncon[4], Bcon[4];
  double Bprim[3];  // alpha * B^i
};

// utoprim_2d_final.c:457 vsq_calc
__device__ __forceinline__ double vsq_of_W(const C2PInv &c, double W) {
  double Wsq = W * W;
  double Xsq = (c.Bsq + W) * (c.Bsq + W);
  return ((Wsq * c.Qtsq + c.QdotBsq * (c.Bsq + 2. * W)) / (Wsq * Xsq));
}

// One Newton step of utoprim_2d_final.c:1189 func_vsq2 (with the EOS of
// utoprim_jon_eos.c:107,230,346); returns dx[0..1] and f, df.
__device__ __forceinline__ void newton_dx(const C2PInv &c, double igr, double W, double vsq,
                                          double &dx0, double &dx1, double &f, double &df) {
  double Wsq = W * W, D = c.D, Bsq = c.Bsq, QdotBsq = c.QdotBsq;
  double Wp = W - D;
  double gamma = 1.0 / sqrt(1.0 - vsq);
  double gammasq = gamma * gamma;
  double wmrho0 = Wp / gammasq - D * vsq / (1.0 + gamma);
  double p_tmp = igr * wmrho0;
  double dPdW = 0.0 * 0.0 + (1.0 - vsq) * igr;
  double Wp2 = W - D;
  double gamma2 = 1.0 / sqrt(1.0 - vsq);
  double dwmrho0dvsq = (D * (gamma2 * 0.5 - 1.0) - Wp2);
  double drho0dvsq = -D * gamma2 * 0.5;
  double dPdvsq = drho0dvsq * 0.0 + dwmrho0dvsq * igr;
  double t2 = -0.5 * Bsq + dPdvsq;
  double t3 = Bsq + W;
  double t4 = t3 * t3;
  double t9 = 1 / Wsq;
  double t11 = c.Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
  double t16 = QdotBsq * t9;
  double t18 = -c.Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  double t21 = 1 / t3;
  double t23 = 1 / W;
  double t24 = t16 * t23;
  double t25 = -1.0 + dPdW - t24;
  double t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  double t36 = 1 / t35;
  dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
  double t40 = (vsq + t24) * t3;
  dx1 = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
  df = -t11 * t11 - t18 * t18;
  f = -0.5 * df;
}

// Setup of Utoprim_new_body up to the Newton guess (utoprim_2d_final.c:152-300).
// U is alpha * U_geomfree; prim holds the guess.  Returns a failure code (4)
// when the guess is unusable, else 0, and sets W_last and the invariants.
__device__ __forceinline__ int c2p_setup(const double *U, const Geom &G, const double *prim,
                                         double gam, C2PInv &c, double &W_last) {
  double Bcov[4], Qcov[4], ncov[4];
  c.Bcon[0] = 0.;
  c.Bcon[1] = U[B1];
  c.Bcon[2] = U[B2];
  c.Bcon[3] = U[B3];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    Bcov[i] = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) Bcov[i] += G.gcov[i][j] * c.Bcon[j];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) Qcov[i] = U[UU + i];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    c.Qcon[i] = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) c.Qcon[i] += G.gcon[i][j] * Qcov[j];
  }
  c.Bsq = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++) c.Bsq += c.Bcon[i] * Bcov[i];
  c.QdotB = 0.;
#pragma unroll
  for (int i = 0; i < 4; i++) c.QdotB += Qcov[i] * c.Bcon[i];
  c.QdotBsq = c.QdotB * c.QdotB;
  double lapse = sqrt(-1. / G.gcon[0][0]);
  ncov[0] = -lapse;
  ncov[1] = 0.;
  ncov[2] = 0.;
  ncov[3] = 0.;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    c.ncon[i] = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) c.ncon[i] += G.gcon[i][j] * ncov[j];
  }
  c.Qdotn = c.Qcon[0] * ncov[0];
  double Qsq = 0.;
#pragma unroll
  for (int i = 0; i < 4; i++) Qsq += Qcov[i] * c.Qcon[i];
  c.Qtsq = Qsq + c.Qdotn * c.Qdotn;
  c.D = U[RHO];

  double utsq = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) utsq += G.gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
  if ((utsq < 0.) && (fabs(utsq) < MAXNEGUTSQ)) utsq = 0.0;
  const double utsq_too_big = (GAMMAFAIL - 1.0) * (GAMMAFAIL - 1.0);
  if (utsq < 0. || utsq > utsq_too_big) return FAIL_CONVGUESSUTSQ;
  double gammasq = 1. + utsq;
  double gamma = sqrt(gammasq);
  double rho0 = c.D / gamma;
  double u = prim[UU];
  double p = (gam - 1.) * u;
  double w = rho0 + u + p;
  W_last = w * gammasq;
  int i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * c.Bsq) - c.QdotBsq * (2. * W_last + c.Bsq)) <=
          W_last * W_last * (c.Qtsq - c.Bsq * c.Bsq)) &&
         (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  c.w = w;
  return 0;
}

// Post-solve of Utoprim_new_body (utoprim_2d_final.c:335-441): from the
// converged W to primitives.  Writes prim[0..4] when the inversion returns 0.
__device__ __forceinline__ int c2p_finish(const C2PInv &c, double W, int retval, double igr,
                                          double *prim_out, bool &assign) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  const double maxnegvsq = (1.0 - 1.0 / (MAXNEGUTSQ + 1.0));
  assign = false;
  if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
  double Wtest = W / c.w;
  assign = true;  // remaining failures return 0 from Utoprim_new_body: guess kept
  if (Wtest <= 0. || Wtest > gsq_big) { assign = false; return FAIL_CONVW; }
  double vsq = vsq_of_W(c, W);
  if ((vsq < 0.) && (fabs(vsq) < maxnegvsq)) vsq = 0.0;
  if ((vsq > vsq_big) || (vsq < 0.0)) { assign = false; return FAIL_CONVUTSQ; }
  double gtmp = sqrt(1. - vsq);
  double gamma = 1. / gtmp;
  double rho0 = c.D * gtmp;
  double w = W * (1. - vsq);
  double p = igr * (w - rho0);
  double u = w - (rho0 + p);
  int flag = 0;
  if ((rho0 <= 0.) || (u < 0.)) {
    if ((rho0 <= 0.) && (u >= 0.)) flag = FAIL_RHONEG;
    if ((rho0 > 0.) && (u < 0.)) flag = FAIL_UNEG;
    if ((rho0 <= 0.) && (u < 0.)) flag = FAIL_RHOUNEG;
  }
  prim_out[RHO] = rho0;
  prim_out[UU] = u;
#pragma unroll
  for (int i = 1; i < 4; i++) {
    double Qtcon = c.Qcon[i] + c.ncon[i] * c.Qdotn;
    prim_out[U1 + i - 1] = gamma / (W + c.Bsq) * (Qtcon + c.QdotB * c.Bcon[i] / W);
  }
  return flag;
}

// Full FP64 inversion of one zone (utoprimgen.c:23 with UTOPRIM2DFINAL).
// U: evolved conserved variables (with gdet); pr: guess in, solution out.
__device__ __forceinline__ int utoprimgen_fp64(const double *U, const Geom &G, double gam, double *pr,
                                               int &iters) {
  double Ug[NPR], Ut[NPR];
#pragma unroll
  for (int i = 0; i < NPR; i++) Ug[i] = U[i] / G.g;
  Ug[UU] -= Ug[RHO];
#pragma unroll
  for (int i = B1; i <= B3; i++) pr[i] = Ug[i];
  double alpha = 1.0 / sqrt(-G.gcon[0][0]);
#pragma unroll
  for (int i = 0; i < NPR; i++) Ut[i] = alpha * Ug[i];
  const double igr = (gam - 1.0) / gam;
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  C2PInv c;
  double W_last;
  iters = 0;
  int code = c2p_setup(Ut, G, pr, gam, c, W_last);
  if (code) return code;
  // find_root_2D_gen + general_newton_raphson without line search
  double x0 = fabs(W_last);
  double v0 = vsq_of_W(c, W_last);
  double x1 = (v0 > 1.) ? vsq_big : v0;
  double f = 1., df = 1., errx = 1.;
  int n_iter = 0;
  bool keep = true;
  while (keep) {
    iters++;
    double dx0, dx1;
    newton_dx(c, igr, x0, x1, dx0, dx1, f, df);
    double xo0 = x0;
    x0 += dx0;
    x1 += dx1;
    errx = (x0 == 0.) ? fabs(dx0) : fabs(dx0 / x0);
    x0 = fabs(x0);
    x0 = (x0 / c.w > gsq_big) ? xo0 : x0;
    x1 = (x1 > 1.) ? vsq_big : x1;
    if ((fabs(errx) <= NEWT_TOL) || (n_iter >= (MAX_NEWT_ITER - 1))) keep = false;
    n_iter++;
  }
  int retval = 0;
  if (!isfinite(f) || !isfinite(df) || !isfinite(x0) || !isfinite(x1)) retval = 1;
  else if (fabs(errx) > MIN_NEWT_TOL) retval = 1;
  double out[NPR];
  bool assign;
  code = c2p_finish(c, x0, retval, igr, out, assign);
  if (assign) {
#pragma unroll
    for (int i = 0; i < B1; i++) pr[i] = out[i];
  }
  return code;
}

}  // namespace hg

namespace hg {

// ---------------------------------------------------------------------------
// Mixed precision: the Newton iteration of utoprimgen_fp64 run in FP32 from
// the same FP64 set-up (metric contractions, invariants, guess), then FP64
// Newton polishes from the FP32 root until the reference's tolerance holds,
// then the FP64 post-solve.  Whenever the FP32 phase does not converge, meets
// a non-finite value, or the polish needs more than MIXED_MAX_POLISH steps,
// the zone is inverted again by utoprimgen_fp64 from the original guess, so
// every hard case keeps the reference's iterates and failure codes.
// ---------------------------------------------------------------------------
constexpr int MIXED_MAX_F32 = 40;       // FP32 Newton steps before falling back
constexpr float MIXED_TOL_F32 = 2e-6f;  // |dW/W| that ends the FP32 phase
constexpr int MIXED_MAX_POLISH = 4;     // FP64 steps allowed after it (2 expected)

struct C2PInvF {
  float Bsq, QdotBsq, Qtsq, Qdotn, D;
};

// newton_dx in FP32 (same algebra) with the hardware reciprocal and
// reciprocal square root (about 1 ulp): the FP32 phase only has to bring W
// within MIXED_TOL_F32 of the root, the FP64 polish does the rest
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ void newton_dx_f32(const C2PInvF &c, float igr, float W, float vsq, float &dx0,
                                              float &dx1) {
  float Wsq = W * W, D = c.D, Bsq = c.Bsq, QdotBsq = c.QdotBsq;
  float Wp = W - D;
  float gamma = __builtin_amdgcn_rsqf(1.0f - vsq);
  float gammasq = gamma * gamma;
  float wmrho0 = Wp * frcp(gammasq) - D * vsq * frcp(1.0f + gamma);
  float p_tmp = igr * wmrho0;
  float dPdW = (1.0f - vsq) * igr;
  float dwmrho0dvsq = (D * (gamma * 0.5f - 1.0f) - Wp);
  float dPdvsq = dwmrho0dvsq * igr;
  float t2 = -0.5f * Bsq + dPdvsq;
  float t3 = Bsq + W;
  float t4 = t3 * t3;
  float t9 = frcp(Wsq);
  float t11 = c.Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0f * W) * t9;
  float t16 = QdotBsq * t9;
  float t18 = -c.Qdotn - 0.5f * Bsq * (1.0f + vsq) + 0.5f * t16 - W + p_tmp;
  float t21 = frcp(t3);
  float t23 = frcp(W);
  float t24 = t16 * t23;
  float t25 = -1.0f + dPdW - t24;
  float t35 = t25 * t3 + (Bsq - 2.0f * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  float t36 = frcp(t35);
  dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
  float t40 = (vsq + t24) * t3;
  dx1 = -(-t25 * t11 - 2.0f * t40 * t18) * t21 * t36;
}

// Returns the failure code; iters counts FP32 + FP64 Newton steps (or the
// FP64 path's steps after a fallback); *fell_back tells which path answered.
__device__ __forceinline__ int utoprimgen_mixed(const double *U, const Geom &G, double gam, double *pr,
                                                int &iters, bool &fell_back) {
  double guess[NPR];
#pragma unroll
  for (int i = 0; i < NPR; i++) guess[i] = pr[i];
  fell_back = false;
  double Ug[NPR], Ut[NPR];
#pragma unroll
  for (int i = 0; i < NPR; i++) Ug[i] = U[i] / G.g;
  Ug[UU] -= Ug[RHO];
  double alpha = 1.0 / sqrt(-G.gcon[0][0]);
#pragma unroll
  for (int i = 0; i < NPR; i++) Ut[i] = alpha * Ug[i];
  const double igr = (gam - 1.0) / gam;
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  C2PInv c;
  double W_last;
  iters = 0;
  int code = c2p_setup(Ut, G, guess, gam, c, W_last);
  bool ok = (code == 0);
  double x0 = 0., x1 = 0.;
  if (ok) {
    // FP32 bulk
    C2PInvF cf = {(float)c.Bsq, (float)c.QdotBsq, (float)c.Qtsq, (float)c.Qdotn, (float)c.D};
    const float igrf = (float)igr, wf = (float)c.w;
    const float gsq_bigf = (float)gsq_big, vsq_bigf = (float)vsq_big;
    float w0 = (float)fabs(W_last);
    double v0 = vsq_of_W(c, W_last);
    float w1 = (float)((v0 > 1.) ? vsq_big : v0);
    int n32 = 0;
    bool conv = false;
    while (n32 < MIXED_MAX_F32) {
      float d0, d1;
      newton_dx_f32(cf, igrf, w0, w1, d0, d1);
      float wo = w0;
      w0 += d0;
      w1 += d1;
      float err = (w0 == 0.f) ? fabsf(d0) : fabsf(d0 * frcp(w0));
      w0 = fabsf(w0);
      w0 = (w0 > gsq_bigf * wf) ? wo : w0;
      w1 = (w1 > 1.f) ? vsq_bigf : w1;
      n32++;
      if (!isfinite(w0) || !isfinite(w1)) break;
      if (err <= MIXED_TOL_F32) { conv = true; break; }
    }
    ok = conv;
    iters = n32;
    x0 = (double)w0;
    x1 = (double)w1;
  }
  if (ok) {
    // FP64 polish with the reference's stopping rule
    double f = 1., df = 1., errx = 1.;
    int np = 0;
    bool done = false;
    while (np < MIXED_MAX_POLISH) {
      double dx0, dx1;
      newton_dx(c, igr, x0, x1, dx0, dx1, f, df);
      double xo0 = x0;
      x0 += dx0;
      x1 += dx1;
      errx = (x0 == 0.) ? fabs(dx0) : fabs(dx0 / x0);
      x0 = fabs(x0);
      x0 = (x0 / c.w > gsq_big) ? xo0 : x0;
      x1 = (x1 > 1.) ? vsq_big : x1;
      np++;
      if (fabs(errx) <= NEWT_TOL) { done = true; break; }
    }
    iters += np;
    ok = done && isfinite(f) && isfinite(df) && isfinite(x0) && isfinite(x1);
  }
  if (!ok) {
    // the reference path from the original guess
    fell_back = true;
#pragma unroll
    for (int i = 0; i < NPR; i++) pr[i] = guess[i];
    return utoprimgen_fp64(U, G, gam, pr, iters);
  }
#pragma unroll
  for (int i = B1; i <= B3; i++) pr[i] = Ug[i];
  double out[NPR];
  bool assign;
  code = c2p_finish(c, x0, 0, igr, out, assign);
  if (assign) {
#pragma unroll
    for (int i = 0; i < B1; i++) pr[i] = out[i];
  }
  return code;
}

}  // namespace hg
