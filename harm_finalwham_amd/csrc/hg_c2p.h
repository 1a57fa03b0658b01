// Device Noble et al. (2006) "2D" conserved-to-primitive inversion.
//
// utoprimgen.c:23 Utoprimgen -> utoprim_2d_final.c:34 Utoprim_2d_final ->
// utoprim_2d_final.c:152 Utoprim_new_body -> :573 find_root_2D_gen ->
// :736 general_newton_raphson (no line search) with :1189 func_vsq2 and the
// ideal-gas EOS of utoprim_jon_eos.c / eos.c.  FP64 throughout: same
// operations in the same order as the reference, hence the same iterates,
// iteration counts and failure codes.
#pragma once
#include "hg_dd.h"
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
// The core also returns the invariants and the final Newton iterate (x0 = W,
// x1 = vsq); newton is false when the set-up failed before the iteration.
__device__ __forceinline__ int utoprimgen_fp64_core(const double *U, const Geom &G, double gam, double *pr,
                                                    int &iters, C2PInv &c, double &x0, double &x1,
                                                    bool &newton) {
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
  double W_last;
  iters = 0;
  newton = false;
  int code = c2p_setup(Ut, G, pr, gam, c, W_last);
  if (code) return code;
  newton = true;
  // find_root_2D_gen + general_newton_raphson without line search
  x0 = fabs(W_last);
  double v0 = vsq_of_W(c, W_last);
  x1 = (v0 > 1.) ? vsq_big : v0;
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

__device__ __forceinline__ int utoprimgen_fp64(const double *U, const Geom &G, double gam, double *pr,
                                               int &iters) {
  C2PInv c;
  double x0 = 0., x1 = 0.;
  bool newton;
  return utoprimgen_fp64_core(U, G, gam, pr, iters, c, x0, x1, newton);
}

}  // namespace hg

namespace hg {

// ---------------------------------------------------------------------------
// The converged root and the mixed-precision path.
//
// The reference iteration stops at |dW/W| <= NEWT_TOL with W where the FP64
// residual is still within rounding noise of zero; at Lorentz factors near 50
// and magnetisations near 1e3 that noise moves the primitives by up to ~1e-9
// (vsq is recomputed from W and 1 / (1 - vsq) amplifies it), and two FP64
// iterations from different starts land at different points of the noise.
// The converged root of the same FP64 system (the invariants exactly as the
// reference sets them up) is found by Newton steps whose residuals are
// evaluated in double-double arithmetic (hg_dd.h) with the FP64 Jacobian of
// newton_dx, and the primitives are formed from it in double-double:
//   c2p_polish     FP64 steps to |dW/W| <= 1e-9, then double-double steps
//                  until one moves W by at most 1e-12 of itself (false when
//                  either is not reached).  Newton converges quadratically
//                  there: a step of 1e-12 W leaves an error near 1e-24 W, so
//                  that step's W is the root to the last bit and needs no
//                  further step to confirm it
//   c2p_finish_dd  c2p_finish with vsq, 1 - vsq, rho0, w, p, u and the
//                  velocities in double-double, rounded once
// The mixed-precision path runs the Newton iteration in FP32 from the
// reference's FP64 set-up and start, to |dW/W| <= MIXED_TOL_F32, then polishes
// and finishes as above; a zone whose FP32 phase or polish does not converge
// is inverted again by the reference path (and polished), by the caller.
// ---------------------------------------------------------------------------
constexpr int MIXED_MAX_F32 = 40;         // FP32 Newton steps before falling back
constexpr float MIXED_TOL_F32 = 2e-6f;    // |dW/W| that ends the FP32 phase
constexpr float MIXED_STALL_F32 = 1e-3f;  // below it, a step that shrank less than 4x ends it too
constexpr int POLISH_MAX_F64 = 8;         // FP64 steps of the polish (to |dW/W| <= 1e-9); from the
                                          // FP32 floor at Lorentz factors near 50 (|dW/W| ~ 1e-3) it
                                          // can take six
constexpr int POLISH_MAX_DD = 3;          // double-double steps (until one is <= 1e-12 W)

struct C2PInvF {
  float Bsq, QdotBsq, Qtsq, Qdotn, D;
};

// newton_dx in FP32 (same algebra) with the hardware reciprocal and
// reciprocal square root (about 1 ulp): the FP32 phase only has to bring W
// within MIXED_TOL_F32 of the root, the polish does the rest
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

// The start of utoprimgen_fp64's Newton loop from the set-up's W_last
__device__ __forceinline__ void newton_start(const C2PInv &c, double W_last, double &x0, double &x1) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  x0 = fabs(W_last);
  double v0 = vsq_of_W(c, W_last);
  x1 = (v0 > 1.) ? vsq_big : v0;
}

// FP32 Newton from (x0, x1) with the guards of the reference loop; true when
// |dW/W| <= MIXED_TOL_F32 was reached with finite iterates; iters counts steps
__device__ __forceinline__ bool newton_f32(const C2PInv &c, double igr, double &x0, double &x1, int &iters) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  C2PInvF cf = {(float)c.Bsq, (float)c.QdotBsq, (float)c.Qtsq, (float)c.Qdotn, (float)c.D};
  const float igrf = (float)igr, wf = (float)c.w;
  const float gsq_bigf = (float)gsq_big, vsq_bigf = (float)vsq_big;
  float w0 = (float)x0, w1 = (float)x1;
  bool conv = false;
  int n32 = 0;
  float errp = 1e30f;
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
    // converged; or near the root (|dW/W| <= MIXED_STALL_F32) with the
    // steps no longer shrinking quadratically: FP32 rounding of the
    // residual (amplified by 1 / (1 - vsq)) sets the floor, the FP64 polish
    // takes it from here
    if (err <= MIXED_TOL_F32 || (err <= MIXED_STALL_F32 && err > 0.25f * errp)) {
      conv = true;
      break;
    }
    errp = err;
  }
  iters += n32;
  x0 = (double)w0;
  x1 = (double)w1;
  return conv;
}

// One Newton step with the residuals t11, t18 of func_vsq2 in double-double
// and the FP64 Jacobian (newton_dx's terms)
__device__ __forceinline__ void newton_dx_dd(const C2PInv &c, double igr, double W, double vsq, double &dx0,
                                             double &dx1) {
  const double Wsq = W * W, D = c.D, Bsq = c.Bsq, QdotBsq = c.QdotBsq;
  const double gamma = 1.0 / sqrt(1.0 - vsq);
  const double dPdW = (1.0 - vsq) * igr;
  const double dPdvsq = (D * (gamma * 0.5 - 1.0) - (W - D)) * igr;
  const double t2 = -0.5 * Bsq + dPdvsq, t3 = Bsq + W, t4 = t3 * t3, t9 = 1 / Wsq;
  const double t16 = QdotBsq * t9, t21 = 1 / t3, t23 = 1 / W, t24 = t16 * t23, t25 = -1.0 + dPdW - t24;
  const double t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  const double t36 = 1 / t35;
  // residuals: t11 = Qtsq - vsq (Bsq + W)^2 + QdotBsq (Bsq + 2 W) / W^2,
  // t18 = -Qdotn - Bsq (1 + vsq) / 2 + QdotBsq / (2 W^2) - W + p(W, vsq)
  const dd Wd = dd_of(W), vd = dd_of(vsq);
  const dd T3 = two_sum(Bsq, W);
  const dd q = dd_div(dd_of(QdotBsq), two_prod(W, W));
  const dd r11 = dd_add(dd_sub(dd_of(c.Qtsq), dd_mul(vd, dd_mul(T3, T3))), dd_mul(q, two_sum(Bsq, 2.0 * W)));
  const dd omv = two_sum(1.0, -vsq);
  const dd gd = dd_div(dd_of(1.0), dd_sqrt(omv));
  const dd wm = dd_sub(dd_mul(two_sum(W, -D), omv), dd_div(two_prod(vsq, D), dd_add(dd_of(1.0), gd)));
  const dd r18 = dd_add(dd_add(dd_add(dd_of(-c.Qdotn), dd_mul_d(two_sum(1.0, vsq), -0.5 * Bsq)), dd_mul_d(q, 0.5)),
                        dd_add(dd_neg(Wd), dd_mul_d(wm, igr)));
  const double t11 = dd_val(r11), t18 = dd_val(r18);
  dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
  const double t40 = (vsq + t24) * t3;
  dx1 = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
}

// The polish: FP64 steps (at most n64; none when n64 = 0) until |dW/W| <=
// 1e-9, then double-double steps until one moves W by <= 1e-12 W.  True when both
// tolerances were reached with finite iterates.
__device__ __forceinline__ bool c2p_polish(const C2PInv &c, double igr, double &x0, double &x1, int n64,
                                           int &iters) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  bool ok = (n64 == 0);
  for (int k = 0; k < n64 && !ok; k++) {
    double dx0, dx1, f, df;
    newton_dx(c, igr, x0, x1, dx0, dx1, f, df);
    double xo0 = x0;
    x0 = fabs(x0 + dx0);
    x0 = (x0 / c.w > gsq_big) ? xo0 : x0;
    x1 += dx1;
    x1 = (x1 > 1.) ? vsq_big : x1;
    iters++;
    ok = fabs(x0 - xo0) <= 1e-9 * x0;
  }
  if (!ok) return false;
  ok = false;
  for (int k = 0; k < POLISH_MAX_DD && !ok; k++) {
    double dx0, dx1;
    newton_dx_dd(c, igr, x0, x1, dx0, dx1);
    double xo0 = x0;
    x0 = fabs(x0 + dx0);
    x0 = (x0 / c.w > gsq_big) ? xo0 : x0;
    x1 += dx1;
    x1 = (x1 > 1.) ? vsq_big : x1;
    iters++;
    ok = fabs(x0 - xo0) <= 1e-12 * x0;
  }
  return ok && isfinite(x0) && isfinite(x1);
}

// c2p_finish from a converged W with the primitives formed in double-double
// (the checks and failure codes of c2p_finish, on the rounded values)
__device__ __forceinline__ int c2p_finish_dd(const C2PInv &c, double W, double igr, double *prim_out,
                                             bool &assign) {
  const double gsq_big = GAMMAFAIL * GAMMAFAIL;
  const double vsq_big = (gsq_big - 1.0) / (gsq_big);
  const double maxnegvsq = (1.0 - 1.0 / (MAXNEGUTSQ + 1.0));
  assign = false;
  if (W == FAIL_VAL) return 101;
  double Wtest = W / c.w;
  assign = true;
  if (Wtest <= 0. || Wtest > gsq_big) { assign = false; return FAIL_CONVW; }
  const dd Wsq = two_prod(W, W), X = two_sum(c.Bsq, W);
  dd vd = dd_div(dd_add(dd_mul_d(Wsq, c.Qtsq), dd_mul_d(two_sum(c.Bsq, 2.0 * W), c.QdotBsq)),
                 dd_mul(Wsq, dd_mul(X, X)));
  double vsq = dd_val(vd);
  if ((vsq < 0.) && (fabs(vsq) < maxnegvsq)) {
    vsq = 0.0;
    vd = dd_of(0.0);
  }
  if ((vsq > vsq_big) || (vsq < 0.0)) { assign = false; return FAIL_CONVUTSQ; }
  const dd omv = dd_sub(dd_of(1.0), vd);
  const dd gt = dd_sqrt(omv);
  const dd rho0 = dd_mul_d(gt, c.D);
  const dd w = dd_mul_d(omv, W);
  const dd p = dd_mul_d(dd_sub(w, rho0), igr);
  const dd u = dd_sub(w, dd_add(rho0, p));
  const double r = dd_val(rho0), uu = dd_val(u);
  int flag = 0;
  if ((r <= 0.) || (uu < 0.)) {
    if ((r <= 0.) && (uu >= 0.)) flag = FAIL_RHONEG;
    if ((r > 0.) && (uu < 0.)) flag = FAIL_UNEG;
    if ((r <= 0.) && (uu < 0.)) flag = FAIL_RHOUNEG;
  }
  prim_out[RHO] = r;
  prim_out[UU] = uu;
  const dd f = dd_div(dd_div(dd_of(1.0), gt), X);  // gamma / (W + Bsq)
  const dd iW = dd_div(dd_of(1.0), dd_of(W));
#pragma unroll
  for (int i = 1; i < 4; i++) {
    dd Qt = dd_add(dd_of(c.Qcon[i]), two_prod(c.ncon[i], c.Qdotn));
    dd v = dd_add(Qt, dd_mul(two_prod(c.QdotB, c.Bcon[i]), iW));
    prim_out[U1 + i - 1] = dd_val(dd_mul(f, v));
  }
  return flag;
}

// The reference inversion of one zone with a converged root: utoprimgen_fp64
// (its failure code, its iteration count and, for a failed zone, its
// values), then, for a zone it solves, the polish from its final iterate and
// the primitives in double-double; these replace the reference's values
// when the polish converges.  *polished tells whether they did.
__device__ __forceinline__ int utoprimgen_refined(const double *U, const Geom &G, double gam, double *pr,
                                                  int &iters, bool &polished) {
  C2PInv c;
  double x0 = 0., x1 = 0.;
  bool newton;
  polished = false;
  int code = utoprimgen_fp64_core(U, G, gam, pr, iters, c, x0, x1, newton);
  if (code != 0 || !newton) return code;
  const double igr = (gam - 1.0) / gam;
  int it = 0;
  if (!c2p_polish(c, igr, x0, x1, POLISH_MAX_F64, it)) return code;
  double out[NPR];
  bool assign;
  if (c2p_finish_dd(c, x0, igr, out, assign) == 0 && assign) {
#pragma unroll
    for (int i = 0; i < B1; i++) pr[i] = out[i];
    iters += it;
    polished = true;
  }
  return code;
}

}  // namespace hg
