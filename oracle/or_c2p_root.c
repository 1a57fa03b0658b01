/*
 * or_c2p_root.c -- the converged root of the reference's Noble 2D system, in
 * x87 extended precision.  TEST INFRASTRUCTURE ONLY (the checker of the
 * library's HG_C2P_REFINED, HG_C2P_MIXED and learned-start inversions; the
 * product never links it).
 *
 * For a zone the reference inverts (code 0 from or_utoprimgen, which follows
 * utoprimgen.c:23 -> utoprim_2d_final.c:152 Utoprim_new_body), the invariants
 * of the Newton system are formed exactly as Utoprim_new_body forms them, in
 * FP64 (utoprim_2d_final.c:218-271: Bcov, Qcon, Bsq, QdotB, ncon, Qdotn,
 * Qtsq, D).  The root (W, vsq) of the residuals of func_vsq2
 * (utoprim_2d_final.c:1232-1254) for those FP64 invariants is then found by
 * Newton's method in long double (64-bit significand) from the reference's
 * solution, and the primitives are formed from W as Utoprim_new_body does
 * (utoprim_2d_final.c:335-431: vsq from W by vsq_calc, rho0, w, p, u and the
 * velocities), in long double, rounded once to double.
 *
 * The reference iteration stops where its FP64 residual is within rounding
 * noise (|dW/W| <= 1e3 eps), which at high Lorentz factor and magnetisation
 * leaves its primitives up to ~1e-9 away from this root.
 */
#include <math.h>
#include <stdint.h>

#include "harm_oracle.h"

typedef long double ld;

typedef struct {
  double Bsq, QdotB, QdotBsq, Qtsq, Qdotn, D;
  double Qcon[4], ncon[4], Bcon[4];
} inv64;

/* the invariants of Utoprim_new_body in its FP64 order, from U = alpha U_geomfree */
static void invariants(const double *U, const or_geom *g, inv64 *c)
{
  double Bcov[4], Qcov[4], ncov[4], Qsq;
  int i, j;
  c->Bcon[0] = 0.;
  for (i = 1; i < 4; i++) c->Bcon[i] = U[OR_B1 + i - 1];
  for (i = 0; i < 4; i++) {
    Bcov[i] = 0.;
    for (j = 0; j < 4; j++) Bcov[i] += g->gcov[i][j] * c->Bcon[j];
  }
  for (i = 0; i < 4; i++) Qcov[i] = U[OR_UU + i];
  for (i = 0; i < 4; i++) {
    c->Qcon[i] = 0.;
    for (j = 0; j < 4; j++) c->Qcon[i] += g->gcon[i][j] * Qcov[j];
  }
  c->Bsq = 0.;
  for (i = 1; i < 4; i++) c->Bsq += c->Bcon[i] * Bcov[i];
  c->QdotB = 0.;
  for (i = 0; i < 4; i++) c->QdotB += Qcov[i] * c->Bcon[i];
  c->QdotBsq = c->QdotB * c->QdotB;
  ncov[0] = -sqrt(-1. / g->gcon[0][0]);
  ncov[1] = ncov[2] = ncov[3] = 0.;
  for (i = 0; i < 4; i++) {
    c->ncon[i] = 0.;
    for (j = 0; j < 4; j++) c->ncon[i] += g->gcon[i][j] * ncov[j];
  }
  c->Qdotn = c->Qcon[0] * ncov[0];
  Qsq = 0.;
  for (i = 0; i < 4; i++) Qsq += Qcov[i] * c->Qcon[i];
  c->Qtsq = Qsq + c->Qdotn * c->Qdotn;
  c->D = U[OR_RHO];
}

/* residuals t11, t18 of func_vsq2 and the Newton step, all in long double */
static void step_ld(const inv64 *c, ld igr, ld W, ld vsq, ld *dx0, ld *dx1, ld *r0, ld *r1)
{
  ld D = c->D, Bsq = c->Bsq, QdotBsq = c->QdotBsq;
  ld Wsq = W * W, Wp = W - D, gamma = 1.0L / sqrtl(1.0L - vsq), gammasq = gamma * gamma;
  ld wmrho0 = Wp / gammasq - D * vsq / (1.0L + gamma);
  ld p_tmp = igr * wmrho0, dPdW = (1.0L - vsq) * igr;
  ld dPdvsq = (D * (gamma * 0.5L - 1.0L) - Wp) * igr;
  ld t2 = -0.5L * Bsq + dPdvsq, t3 = Bsq + W, t4 = t3 * t3, t9 = 1.0L / Wsq;
  ld t11 = (ld)c->Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0L * W) * t9;
  ld t16 = QdotBsq * t9;
  ld t18 = -(ld)c->Qdotn - 0.5L * Bsq * (1.0L + vsq) + 0.5L * t16 - W + p_tmp;
  ld t21 = 1.0L / t3, t23 = 1.0L / W, t24 = t16 * t23, t25 = -1.0L + dPdW - t24;
  ld t35 = t25 * t3 + (Bsq - 2.0L * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  ld t36 = 1.0L / t35, t40 = (vsq + t24) * t3;
  *dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
  *dx1 = -(-t25 * t11 - 2.0L * t40 * t18) * t21 * t36;
  *r0 = t11;
  *r1 = t18;
}

/*
 * n zones (records as or_c2p_batch: gcov[16], gcon[16], gdet, gcovpert[4]
 * per zone, U[8], pr[8]): pr_inout holds the guess and receives the
 * reference's solution where the reference fails, the converged root's
 * primitives where it succeeds; pflag the reference's code; Wv[2n] (may be
 * NULL) the root (W, vsq); resid[2n] (may be NULL) the residuals t11 / (Bsq
 * + W)^2 and t18 / (|Qdotn| + Bsq + W) at the root.  start: 0 = from the
 * reference's solution, 1 = from the primitives given in start_prims[8n].
 */
void or_c2p_converged_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                            const double *gcovpert, const double *U, double gam, double *pr_inout,
                            int32_t *pflag, int start, const double *start_prims, double *Wv, double *resid)
{
  int64_t z;
  for (z = 0; z < n; z++) {
    or_geom g;
    double Ug[OR_NPR], Ut[OR_NPR], alpha, utsq;
    const double *ps;
    inv64 c;
    int fl, it, a, b, k;
    ld igr = ((ld)gam - 1.0L) / (ld)gam, W, vsq, dx0, dx1, r0 = 0, r1 = 0;
    for (a = 0; a < 4; a++) {
      for (b = 0; b < 4; b++) {
        g.gcov[a][b] = gcov[z * 16 + a * 4 + b];
        g.gcon[a][b] = gcon[z * 16 + a * 4 + b];
      }
      g.gcovpert[a] = gcovpert[z * 4 + a];
    }
    g.g = gdet[z];
    or_utoprimgen(&U[z * 8], &g, gam, &pr_inout[z * 8], &fl, &it);
    pflag[z] = fl;
    if (Wv) Wv[2 * z] = Wv[2 * z + 1] = 0.;
    if (resid) resid[2 * z] = resid[2 * z + 1] = 0.;
    if (fl != 0) continue;
    /* utoprimgen.c:76-85 and utoprim_2d_final.c:60-70: U_geomfree, alpha U */
    for (k = 0; k < OR_NPR; k++) Ug[k] = U[z * 8 + k] / g.g;
    Ug[OR_UU] -= Ug[OR_RHO];
    alpha = 1.0 / sqrt(-g.gcon[0][0]);
    for (k = 0; k < OR_NPR; k++) Ut[k] = alpha * Ug[k];
    invariants(Ut, &g, &c);
    /* start: W and vsq of the given primitives */
    ps = start ? &start_prims[z * 8] : &pr_inout[z * 8];
    utsq = 0.;
    for (a = 1; a < 4; a++)
      for (b = 1; b < 4; b++) utsq += g.gcov[a][b] * ps[OR_U1 + a - 1] * ps[OR_U1 + b - 1];
    W = ((ld)ps[OR_RHO] + (ld)gam * (ld)ps[OR_UU]) * (1.0L + (ld)utsq);
    vsq = (ld)utsq / (1.0L + (ld)utsq);
    for (k = 0; k < 12; k++) {
      step_ld(&c, igr, W, vsq, &dx0, &dx1, &r0, &r1);
      W += dx0;
      vsq += dx1;
    }
    step_ld(&c, igr, W, vsq, &dx0, &dx1, &r0, &r1);
    {
      /* utoprim_2d_final.c:335-431 from W, in long double */
      ld Bsq = c.Bsq, Wsq = W * W, Xsq = (Bsq + W) * (Bsq + W);
      ld v = (Wsq * (ld)c.Qtsq + (ld)c.QdotBsq * (Bsq + 2.0L * W)) / (Wsq * Xsq);
      ld gtmp = sqrtl(1.0L - v), gamma = 1.0L / gtmp, rho0 = (ld)c.D * gtmp, w = W * (1.0L - v);
      ld p = igr * (w - rho0), u = w - (rho0 + p);
      double *o = &pr_inout[z * 8];
      o[OR_RHO] = (double)rho0;
      o[OR_UU] = (double)u;
      for (a = 1; a < 4; a++) {
        ld Qtcon = (ld)c.Qcon[a] + (ld)c.ncon[a] * (ld)c.Qdotn;
        o[OR_U1 + a - 1] = (double)(gamma / (W + Bsq) * (Qtcon + (ld)c.QdotB * (ld)c.Bcon[a] / W));
      }
    }
    if (Wv) {
      Wv[2 * z] = (double)W;
      Wv[2 * z + 1] = (double)vsq;
    }
    if (resid) {
      resid[2 * z] = (double)(fabsl(r0) / (((ld)c.Bsq + W) * ((ld)c.Bsq + W)));
      resid[2 * z + 1] = (double)(fabsl(r1) / (fabsl((ld)c.Qdotn) + (ld)c.Bsq + W));
    }
  }
}
