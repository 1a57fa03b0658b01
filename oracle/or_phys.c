/*
 * or_phys.c -- point physics and the Noble 2D inversion of
 * pseudotensor/harm_finalwham, restated in C.  TEST INFRASTRUCTURE ONLY
 * (the checker; the product never links it).
 *
 * Operation order follows the reference line by line so that results agree
 * bit for bit with the reference's gcc -O2 x86-64 build (no FMA).
 */
#include "harm_oracle.h"
#include <float.h>
#include <math.h>

#define TT 0

/* phys.c:1797-1809: max/min as the reference defines them */
static double rmax(double a, double b) { if (a > b) return a; else return b; }

/* phys.c:766 lower_vec (upper-triangle indexing of gcov) */
static void lower_vec(const double *ucon, const or_geom *g, double *ucov)
{
  const double (*gc)[4] = g->gcov;
  ucov[0] = gc[0][0] * ucon[0] + gc[0][1] * ucon[1] + gc[0][2] * ucon[2] + gc[0][3] * ucon[3];
  ucov[1] = gc[0][1] * ucon[0] + gc[1][1] * ucon[1] + gc[1][2] * ucon[2] + gc[1][3] * ucon[3];
  ucov[2] = gc[0][2] * ucon[0] + gc[1][2] * ucon[1] + gc[2][2] * ucon[2] + gc[2][3] * ucon[3];
  ucov[3] = gc[0][3] * ucon[0] + gc[1][3] * ucon[1] + gc[2][3] * ucon[2] + gc[3][3] * ucon[3];
}

/* phys.c:825 raise_vec (upper-triangle indexing of gcon) */
static void raise_vec(const double *ucov, const or_geom *g, double *ucon)
{
  const double (*gn)[4] = g->gcon;
  ucon[0] = gn[0][0] * ucov[0] + gn[0][1] * ucov[1] + gn[0][2] * ucov[2] + gn[0][3] * ucov[3];
  ucon[1] = gn[0][1] * ucov[0] + gn[1][1] * ucov[1] + gn[1][2] * ucov[2] + gn[1][3] * ucov[3];
  ucon[2] = gn[0][2] * ucov[0] + gn[1][2] * ucov[1] + gn[2][2] * ucov[2] + gn[2][3] * ucov[3];
  ucon[3] = gn[0][3] * ucov[0] + gn[1][3] * ucov[1] + gn[2][3] * ucov[2] + gn[3][3] * ucov[3];
}

/* phys.c:1069 qsq_calc + phys.c:1040 gamma_calc_fromuconrel (JONCHECKS) */
static int gamma_calc(const double *pr, const or_geom *g, double *gamma)
{
  const double (*gc)[4] = g->gcov;
  double u1 = pr[OR_U1], u2 = pr[OR_U2], u3 = pr[OR_U3];
  double qsq = +gc[1][1] * u1 * u1 + gc[2][2] * u2 * u2 + gc[3][3] * u3 * u3 +
               2. * (gc[1][2] * u1 * u2 + gc[1][3] * u1 * u3 + gc[2][3] * u2 * u3);
  if (qsq < 0.0) {
    if (qsq > -DBL_EPSILON * 100) qsq = 0.0;
    else return 1;
  }
  *gamma = sqrt(1. + qsq);
  return 0;
}

int or_gamma_calc(const double *pr, const or_geom *g, double *gamma) { return gamma_calc(pr, g, gamma); }

/* phys.c:1006 ucon_calc_rel4vel -> phys.c:987 ucon_calc_rel4vel_fromuconrel */
static int ucon_calc(const double *pr, const or_geom *g, double *ucon)
{
  double alpha, gamma, beta[4];
  int j;
  alpha = 1. / sqrt(-g->gcon[TT][TT]);
  for (j = 1; j < 4; j++) beta[j] = g->gcon[TT][j] * alpha * alpha;
  if (gamma_calc(pr, g, &gamma)) return 1;
  ucon[TT] = gamma / alpha;
  for (j = 1; j < 4; j++) ucon[j] = pr[OR_U1 + j - 1] - ucon[TT] * beta[j];
  return 0;
}

/* phys.c:309 bcon_calc */
static void bcon_calc(const double *pr, const double *ucon, const double *ucov, double *bcon)
{
  int j;
  bcon[TT] = pr[OR_B1] * ucov[1] + pr[OR_B2] * ucov[2] + pr[OR_B3] * ucov[3];
  for (j = 1; j <= 3; j++) bcon[j] = (pr[OR_B1 - 1 + j] + bcon[TT] * ucon[j]) / ucon[TT];
}

/* phys.c:849 get_state */
int or_get_state(const double *pr, const or_geom *g, or_state *q)
{
  if (ucon_calc(pr, g, q->ucon)) return 1;
  lower_vec(q->ucon, g, q->ucov);
  bcon_calc(pr, q->ucon, q->ucov, q->bcon);
  lower_vec(q->bcon, g, q->bcov);
  return 0;
}

/* phys.c:541 compute_1plusud0: 1+u_t without cancellation */
static double compute_1plusud0(const or_state *q, const or_geom *g)
{
  int j, k;
  double vcon[4], vsq, gvtt, alpha, uu0, plus1gv00;
  for (j = 1; j < 4; j++) vcon[j] = q->ucon[j] / q->ucon[TT];
  plus1gv00 = g->gcovpert[TT];
  vsq = g->gcovpert[TT];
  for (j = 1; j < 4; j++) vsq += 2.0 * g->gcov[TT][j] * vcon[j];
  for (j = 1; j < 4; j++)
    for (k = 1; k < 4; k++) vsq += g->gcov[j][k] * vcon[j] * vcon[k];
  gvtt = g->gcov[TT][TT];
  alpha = 0.0;
  for (j = 1; j < 4; j++) alpha += g->gcov[j][TT] * q->ucon[j];
  uu0 = q->ucon[TT];
  return alpha + ((1.0 - gvtt) * plus1gv00 - uu0 * uu0 * vsq * gvtt * gvtt) / (1.0 - gvtt * uu0);
}

/* phys.ffde.c:1707 lower_A */
static void lower_A(const double Acon[4][4], const or_geom *g, double Acov[4][4])
{
  int j, k;
  const double (*gc)[4] = g->gcov;
  for (j = 0; j < 4; j++)
    for (k = 0; k < 4; k++) Acov[j][k] = 0.;
  for (j = 0; j < 4; j++)
    for (k = 0; k < 4; k++) {
      Acov[0][1] += (gc[0][j]) * (gc[1][k]) * Acon[j][k];
      Acov[0][2] += (gc[0][j]) * (gc[2][k]) * Acon[j][k];
      Acov[0][3] += (gc[0][j]) * (gc[3][k]) * Acon[j][k];
      Acov[1][2] += (gc[1][j]) * (gc[2][k]) * Acon[j][k];
      Acov[2][3] += (gc[2][j]) * (gc[3][k]) * Acon[j][k];
      Acov[3][1] += (gc[3][j]) * (gc[1][k]) * Acon[j][k];
    }
  Acov[1][0] = -Acov[0][1];
  Acov[2][0] = -Acov[0][2];
  Acov[3][0] = -Acov[0][3];
  Acov[2][1] = -Acov[1][2];
  Acov[3][2] = -Acov[2][3];
  Acov[1][3] = -Acov[3][1];
}

/* phys.c:178 Mcon_calc (PRIMMAXWELL) */
static void Mcon_calc(const double *pr, const or_state *q, double Mcon[4][4])
{
  int j, k;
  for (j = 0; j < 4; j++) Mcon[j][j] = 0.0;
  for (k = 1; k < 4; k++) {
    Mcon[k][0] = pr[OR_B1 + k - 1];
    Mcon[0][k] = -Mcon[k][0];
  }
  for (j = 1; j < 4; j++)
    for (k = 1; k < 4; k++)
      Mcon[j][k] = (pr[OR_B1 + j - 1] * q->ucon[k] - pr[OR_B1 + k - 1] * q->ucon[j]) / q->ucon[0];
}

/* phys.c:442 mhd_calc_norestmass: T^dir_nu with rest mass removed from T^t_t,
   EM part from phys.ffde.c:1483 ffdestresstensor_dir */
static void mhd_calc(const double *pr, int dir, const or_geom *g, const or_state *q, double gam,
                     double *mhd)
{
  int j, k, i;
  double rho = pr[OR_RHO], u = pr[OR_UU];
  double P = (gam - 1.) * u; /* eos.c:58 pressure_rho0_u_grmhd */
  double w = P + rho + u;
  double eta = w, ptot = P;
  double plus1ud0 = compute_1plusud0(q, g);
  double Mcon[4][4], Mcov[4][4], TEM[4], Msq;

  j = 0;
  mhd[j] = (P + u) * q->ucon[dir] * q->ucov[j] + rho * q->ucon[dir] * plus1ud0 +
           ptot * ((dir == j) ? 1. : 0.);
  for (j = 1; j < 4; j++) mhd[j] = eta * q->ucon[dir] * q->ucov[j] + ptot * ((dir == j) ? 1. : 0.);

  Mcon_calc(pr, q, Mcon);
  lower_A(Mcon, g, Mcov);
  for (k = 0; k < 4; k++) TEM[k] = 0.;
  Msq = 0.;
  for (j = 0; j < 4; j++)
    for (k = 0; k < 4; k++) Msq += Mcov[j][k] * Mcon[j][k];
  for (k = 0; k < 4; k++)
    for (i = 0; i < 4; i++) TEM[k] += Mcon[dir][i] * Mcov[k][i];
  TEM[dir] -= 0.25 * Msq;
  for (j = 0; j < 4; j++) mhd[j] += TEM[j];
}

/* phys.c:16 primtoflux with returntype UEVOLVE (gdet prefactor, UtoU phys.c:243) */
void or_primtoflux(const double *pr, const or_state *q, int dir, const or_geom *g, double gam,
                   double *flux)
{
  double fl[OR_NPR];
  int pl;
  fl[OR_RHO] = pr[OR_RHO] * q->ucon[dir]; /* phys.c:58 massflux_calc */
  mhd_calc(pr, dir, g, q, gam, &fl[OR_UU]);
  /* phys.c:143 dualfullfaraday_calc (PRIMMAXWELL) */
  if (dir > 0) {
    fl[OR_B1] = (pr[OR_B1] * q->ucon[dir] - pr[OR_B1 + dir - 1] * q->ucon[1]) / q->ucon[0];
    fl[OR_B2] = (pr[OR_B2] * q->ucon[dir] - pr[OR_B1 + dir - 1] * q->ucon[2]) / q->ucon[0];
    fl[OR_B3] = (pr[OR_B3] * q->ucon[dir] - pr[OR_B1 + dir - 1] * q->ucon[3]) / q->ucon[0];
  } else {
    fl[OR_B1] = pr[OR_B1];
    fl[OR_B2] = pr[OR_B2];
    fl[OR_B3] = pr[OR_B3];
  }
  for (pl = 0; pl < OR_NPR; pl++) flux[pl] = fl[pl] * g->g;
}

void or_primtoU(const double *pr, const or_geom *g, double gam, double *U)
{
  or_state q;
  or_get_state(pr, g, &q);
  or_primtoflux(pr, &q, 0, g, gam, U);
}

/* vchar.c:213 simplefast */
static int simplefast(int dir, const or_geom *g, const or_state *q, double cms2, double *vmin,
                      double *vmax)
{
  double Acov[4], Bcov[4], Acon[4], Bcon[4];
  double Asq, Bsq, Au, Bu, AB, Au2, Bu2, AuBu, A, B, discr, vm, vp;
  int j;
  for (j = 0; j < 4; j++) Acov[j] = 0.;
  Acov[dir] = 1.;
  raise_vec(Acov, g, Acon);
  for (j = 0; j < 4; j++) Bcov[j] = 0.;
  Bcov[TT] = 1.;
  raise_vec(Bcov, g, Bcon);
  Asq = Acon[0] * Acov[0] + Acon[1] * Acov[1] + Acon[2] * Acov[2] + Acon[3] * Acov[3];
  Bsq = Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3];
  Au = Acov[0] * q->ucon[0] + Acov[1] * q->ucon[1] + Acov[2] * q->ucon[2] + Acov[3] * q->ucon[3];
  Bu = Bcov[0] * q->ucon[0] + Bcov[1] * q->ucon[1] + Bcov[2] * q->ucon[2] + Bcov[3] * q->ucon[3];
  AB = Acon[0] * Bcov[0] + Acon[1] * Bcov[1] + Acon[2] * Bcov[2] + Acon[3] * Bcov[3];
  Au2 = Au * Au;
  Bu2 = Bu * Bu;
  AuBu = Au * Bu;
  A = Bu2 * (1.0 - cms2) - Bsq * cms2;
  B = 2. * (AuBu * (1.0 - cms2) - AB * cms2);
  discr = 4.0 * cms2 *
          ((AB * AB - Asq * Bsq) * cms2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (cms2 - 1.0));
  if ((discr < 0.0) && (discr > -1E-10)) discr = 0.0;
  else if (discr < 0.0) return 1;
  discr = sqrt(discr);
  vm = -(-B + discr) / (2. * A);
  vp = -(-B - discr) / (2. * A);
  if (!isfinite(vm) || !isfinite(vp)) return 1;
  if (vp < vm) { *vmax = vm; *vmin = vp; }
  else { *vmax = vp; *vmin = vm; }
  return 0;
}

/* vchar.c:37 vchar (TRUEFAST 0, no boundary smoother, no group velocity) */
int or_vchar(const double *pr, const or_state *q, int dir, const or_geom *g, double gam,
             double *vmax, double *vmin)
{
  double bsq, WW, EF, va2, cs2, cms2, rho, u, P;
  bsq = q->bcon[0] * q->bcov[0] + q->bcon[1] * q->bcov[1] + q->bcon[2] * q->bcov[2] +
        q->bcon[3] * q->bcov[3];
  rho = pr[OR_RHO];
  u = pr[OR_UU];
  if (u < 0) u = 0;
  if (rho < 0) rho = 0;
  P = (gam - 1.) * u;
  WW = rho + u + P;
  if (WW > 0.0) {
    /* eos.c:87 cs2_compute_grmhd */
    double pressure = (gam - 1.) * u;
    double h = rho + u + pressure;
    cs2 = gam * pressure / h;
  } else WW = cs2 = 0.0;
  if (cs2 >= 1.0) cs2 = 1.0 - 1e-9;
  else if (cs2 <= 0.0) cs2 = 0.0;
  EF = bsq + WW;
  if (EF > 0.0) va2 = bsq / EF;
  else EF = va2 = 0.0;
  cms2 = cs2 + va2 - cs2 * va2;
  if (cms2 < 0.0) {
    if (cms2 / (cs2 + va2) > -DBL_EPSILON * 100.0) cms2 = 0.0;
    else return 1;
  }
  if (cms2 > 1.) {
    if (cms2 < 1.0 + DBL_EPSILON * 100.0) cms2 = 1.0;
    else if (rho < 0.0) cms2 = 1.0;
    else return 1;
  }
  (void)EF;
  return simplefast(dir, g, q, cms2, vmin, vmax);
}

/* phys.c:578 source -> phys.c:623 source_conn (REMOVERESTMASSFROMUU==2, WITHGDET) */
void or_source(const double *pr, const or_geom *g, const double conn[4][4][4], double gam,
               double *dU)
{
  or_state q;
  double mhd[4][4], dUconn[OR_NPR];
  int j, k, l, pl;
  or_get_state(pr, g, &q);
  for (j = 0; j < 4; j++) mhd_calc(pr, j, g, &q, gam, mhd[j]);
  for (pl = 0; pl < OR_NPR; pl++) dUconn[pl] = 0;
  for (l = 0; l < 4; l++)
    for (j = 0; j < 4; j++)
      for (k = 0; k < 4; k++) dUconn[OR_UU + l] += mhd[j][k] * conn[k][l][j];
  for (l = 0; l < 4; l++)
    for (j = 0; j < 4; j++) dUconn[OR_UU + l] += -pr[OR_RHO] * q.ucon[j] * conn[TT][l][j];
  for (pl = 0; pl < OR_NPR; pl++) dU[pl] = 0.0 + g->g * dUconn[pl];
}

/* ===================================================================== */
/* Noble et al. 2006 "2D" inversion: utoprim_2d_final.c                   */
/* ===================================================================== */

#define MAX_NEWT_ITER 200               /* u2p_defs.h:38 */
#define NEWT_TOL (1E3 * DBL_EPSILON)    /* u2p_defs.h:39 */
#define MIN_NEWT_TOL (1E-6)             /* u2p_defs.h:47 */
#define MAXNEGUTSQ (1E-10)              /* u2p_defs.h:123 */
#define MAXNEGVSQ (1.0 - 1.0 / (MAXNEGUTSQ + 1.0))
#define FAIL_VAL (1.e30)

typedef struct {
  double Bsq, QdotBsq, Qtsq, Qdotn, D, wglobal, gam;
  double gammasq_too_big, utsq_too_big, vsq_too_big;
  int iters;
} c2p_ctx;

/* utoprim_2d_final.c:457 vsq_calc */
static double vsq_calc(const c2p_ctx *c, double W)
{
  double Wsq = W * W;
  double Xsq = (c->Bsq + W) * (c->Bsq + W);
  return ((Wsq * c->Qtsq + c->QdotBsq * (c->Bsq + 2. * W)) / (Wsq * Xsq));
}

/* utoprim_2d_final.c:1189 func_vsq2 with the EOS of utoprim_jon_eos.c:107,230,346 */
static void func_vsq2(const c2p_ctx *c, const double x[2], double dx[2], double *f, double *df)
{
  double W = x[0], vsq = x[1], Wsq = W * W;
  double D = c->D, Bsq = c->Bsq, QdotBsq = c->QdotBsq, Qtsq = c->Qtsq, Qdotn = c->Qdotn;
  double igr = (c->gam - 1.0) / c->gam; /* eos.c IGAMMAR and compute_idwmrho0dp_grmhd */
  double p_tmp, dPdW, dPdvsq;
  double t2, t3, t4, t9, t11, t16, t18, t21, t23, t24, t25, t35, t36, t40;
  double resid0, resid1;
  {
    /* pressure_Wp_vsq (utoprim_jon_eos.c:107) */
    double Wp = W - D;
    double gamma = 1.0 / sqrt(1.0 - vsq);
    double gammasq = gamma * gamma;
    double wmrho0 = Wp / gammasq - D * vsq / (1.0 + gamma);
    p_tmp = igr * wmrho0;
  }
  {
    /* dpdWp_calc_vsq (utoprim_jon_eos.c:230) */
    double dwmrho0dW = 1.0 - vsq;
    double drho0dW = 0.0, idrho0dp = 0.0;
    dPdW = drho0dW * idrho0dp + dwmrho0dW * igr;
  }
  {
    /* dpdvsq_calc_Wp (utoprim_jon_eos.c:346) */
    double Wp = W - D;
    double gamma = 1.0 / sqrt(1.0 - vsq);
    double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
    double drho0dvsq = -D * gamma * 0.5;
    double idrho0dp = 0.0;
    dPdvsq = drho0dvsq * idrho0dp + dwmrho0dvsq * igr;
  }
  t2 = -0.5 * Bsq + dPdvsq;
  t3 = Bsq + W;
  t4 = t3 * t3;
  t9 = 1 / Wsq;
  t11 = Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
  t16 = QdotBsq * t9;
  t18 = -Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  t21 = 1 / t3;
  t23 = 1 / W;
  t24 = t16 * t23;
  t25 = -1.0 + dPdW - t24;
  t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  t36 = 1 / t35;
  dx[0] = -(t2 * t11 + t4 * t18) * t21 * t36;
  t40 = (vsq + t24) * t3;
  dx[1] = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
  resid0 = t11;
  resid1 = t18;
  *df = -resid0 * resid0 - resid1 * resid1;
  *f = -0.5 * (*df);
}

/* utoprim_2d_final.c:736 general_newton_raphson, no line search (USE_LINE_SEARCH -2) */
static int general_newton_raphson(c2p_ctx *c, double x[2])
{
  double f = 1., df = 1., dx[2], x_old[2], errx = 1.;
  int n_iter = 0, keep_iterating = 1, id;
  while (keep_iterating) {
    c->iters++;
    func_vsq2(c, x, dx, &f, &df);
    errx = 0.;
    for (id = 0; id < 2; id++) x_old[id] = x[id];
    for (id = 0; id < 2; id++) x[id] += dx[id];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    /* validate_x, utoprim_2d_final.c:708 */
    x[0] = fabs(x[0]);
    x[0] = (x[0] / c->wglobal > c->gammasq_too_big) ? x_old[0] : x[0];
    x[1] = (x[1] > 1.) ? c->vsq_too_big : x[1];
    if ((fabs(errx) <= NEWT_TOL) || (n_iter >= (MAX_NEWT_ITER - 1))) keep_iterating = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df) || !isfinite(x[0]) || !isfinite(x[1])) return 1;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

/* utoprim_2d_final.c:152 Utoprim_new_body; U is alpha*U_geomfree, prim's B are alpha*B */
static int new_body(c2p_ctx *c, const double *U, const or_geom *g, double *prim, int *pflag)
{
  double Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4], Qtcon[4];
  double QdotB, Qsq, utsq, gammasq, gamma, rho0, u, p, w, W_last, W, Wtest, vsq, gtmp, lapse;
  double x[2];
  double igr = (c->gam - 1.0) / c->gam;
  int i, j, i_increase, retval;

  *pflag = OR_FAIL_NONE;
  for (i = OR_B1; i <= OR_B3; i++) prim[i] = U[i];
  Bcon[0] = 0.;
  for (i = 1; i < 4; i++) Bcon[i] = U[OR_B1 + i - 1];
  /* u2p_util.c lower_g / raise_g: full-matrix sums from 0. */
  for (i = 0; i < 4; i++) {
    Bcov[i] = 0.;
    for (j = 0; j < 4; j++) Bcov[i] += g->gcov[i][j] * Bcon[j];
  }
  for (i = 0; i < 4; i++) Qcov[i] = U[OR_UU + i];
  for (i = 0; i < 4; i++) {
    Qcon[i] = 0.;
    for (j = 0; j < 4; j++) Qcon[i] += g->gcon[i][j] * Qcov[j];
  }
  c->Bsq = 0.;
  for (i = 1; i < 4; i++) c->Bsq += Bcon[i] * Bcov[i];
  QdotB = 0.;
  for (i = 0; i < 4; i++) QdotB += Qcov[i] * Bcon[i];
  c->QdotBsq = QdotB * QdotB;
  lapse = sqrt(-1. / g->gcon[0][0]); /* u2p_util.c ncov_calc */
  ncov[0] = -lapse;
  ncov[1] = 0.;
  ncov[2] = 0.;
  ncov[3] = 0.;
  for (i = 0; i < 4; i++) {
    ncon[i] = 0.;
    for (j = 0; j < 4; j++) ncon[i] += g->gcon[i][j] * ncov[j];
  }
  c->Qdotn = Qcon[0] * ncov[0];
  Qsq = 0.;
  for (i = 0; i < 4; i++) Qsq += Qcov[i] * Qcon[i];
  c->Qtsq = Qsq + c->Qdotn * c->Qdotn;
  c->D = U[OR_RHO];

  utsq = 0.;
  for (i = 1; i < 4; i++)
    for (j = 1; j < 4; j++) utsq += g->gcov[i][j] * prim[OR_U1 + i - 1] * prim[OR_U1 + j - 1];
  if ((utsq < 0.) && (fabs(utsq) < MAXNEGUTSQ)) utsq = 0.0;
  if (utsq < 0. || utsq > c->utsq_too_big) {
    *pflag = OR_FAIL_CONVGUESSUTSQ;
    return 1;
  }
  gammasq = 1. + utsq;
  gamma = sqrt(gammasq);
  rho0 = c->D / gamma;
  u = prim[OR_UU];
  p = (c->gam - 1.) * u;
  w = rho0 + u + p;
  W_last = w * gammasq;

  i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * c->Bsq) - c->QdotBsq * (2. * W_last + c->Bsq)) <=
          W_last * W_last * (c->Qtsq - c->Bsq * c->Bsq)) &&
         (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  c->wglobal = w;

  /* find_root_2D_gen, utoprim_2d_final.c:573 */
  x[0] = fabs(W_last);
  {
    double v = vsq_calc(c, W_last); /* x1_of_x0 */
    x[1] = (v > 1.) ? c->vsq_too_big : v;
  }
  retval = general_newton_raphson(c, x);
  W = x[0];

  if ((retval != 0) || (W == FAIL_VAL)) {
    *pflag = retval * 100 + 1;
    return retval;
  }
  Wtest = W / c->wglobal;
  if (Wtest <= 0. || Wtest > c->gammasq_too_big) {
    *pflag = OR_FAIL_CONVW;
    return retval;
  }
  vsq = vsq_calc(c, W);
  if ((vsq < 0.) && (fabs(vsq) < MAXNEGVSQ)) vsq = 0.0;
  if ((vsq > c->vsq_too_big) || (vsq < 0.0)) {
    *pflag = OR_FAIL_CONVUTSQ;
    return retval;
  }
  gtmp = sqrt(1. - vsq);
  gamma = 1. / gtmp;
  rho0 = c->D * gtmp;
  w = W * (1. - vsq);
  p = igr * (w - rho0); /* eos.c:550 pressure_rho0_w -> pressure_wmrho0_grmhd */
  u = w - (rho0 + p);
  if ((rho0 <= 0.) || (u < 0.)) {
    if ((rho0 <= 0.) && (u >= 0.)) *pflag = OR_FAIL_RHONEG;
    if ((rho0 > 0.) && (u < 0.)) *pflag = OR_FAIL_UNEG;
    if ((rho0 <= 0.) && (u < 0.)) *pflag = OR_FAIL_RHOUNEG;
    /* UTOPRIMRETURNADJUSTED: keep going and assign */
  }
  prim[OR_RHO] = rho0;
  prim[OR_UU] = u;
  for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * c->Qdotn;
  for (i = 1; i < 4; i++)
    prim[OR_U1 + i - 1] = gamma / (W + c->Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
  for (i = OR_B1; i <= OR_B3; i++) prim[i] = U[i];
  return retval;
}

/* utoprimgen.c:23 Utoprimgen (energy equation, UTOPRIM2DFINAL) ->
   utoprim_2d_final.c:34 Utoprim_2d_final */
int or_utoprimgen(const double *U, const or_geom *g, double gam, double *pr, int *pflag, int *iters)
{
  double Ug[OR_NPR], U_tmp[OR_NPR], prim_tmp[OR_NPR], alpha;
  double gammafail = 100.0 * 25.0; /* GAMMAFAIL = 100*GAMMAMAX, initbase.c:471-476 */
  c2p_ctx c;
  int i, ret;

  /* UtoU(UEVOLVE -> UNOTHING): divide by the EOM factor (phys.c:243) */
  for (i = 0; i < OR_NPR; i++) Ug[i] = U[i] / g->g;
  Ug[OR_UU] -= Ug[OR_RHO]; /* utoprimgen.c:76-85 */

  c.gam = gam;
  c.iters = 0;
  c.gammasq_too_big = gammafail * gammafail;                       /* u2p_defs.h:108 */
  c.utsq_too_big = (gammafail - 1.0) * (gammafail - 1.0);          /* u2p_defs.h:110 */
  c.vsq_too_big = (c.gammasq_too_big - 1.0) / (c.gammasq_too_big); /* u2p_defs.h:117 */

  for (i = OR_B1; i <= OR_B3; i++) pr[i] = Ug[i];
  alpha = 1.0 / sqrt(-g->gcon[0][0]);
  for (i = OR_RHO; i <= OR_B3; i++) U_tmp[i] = alpha * Ug[i];
  for (i = 0; i < OR_B1; i++) prim_tmp[i] = pr[i];
  for (i = OR_B1; i <= OR_B3; i++) prim_tmp[i] = alpha * pr[i];
  ret = new_body(&c, U_tmp, g, prim_tmp, pflag);
  if (ret == 0)
    for (i = 0; i < OR_B1; i++) pr[i] = prim_tmp[i];
  *iters = c.iters;
  return 0;
}

/* fixup.c:1859 limit_gamma (GAMMAERGOLIMIT 0) */
int or_limit_gamma(double gammamax, double *pr, const or_geom *g)
{
  double gamma, pref, f;
  if (gammamax <= 1.0) return 0;
  if (gamma_calc(pr, g, &gamma)) return 1;
  if ((gamma > gammamax && (gamma != 1.0))) {
    pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
    if (pref < 0.0) return 1;
    f = sqrt(pref);
    pr[OR_U1] *= f;
    pr[OR_U2] *= f;
    pr[OR_U3] *= f;
    return -1;
  }
  return 0;
}

/* fixup.c:341 fixup1zone with set_density_floors rescaletype 0 (fixup.c:1562);
   rhofloor/ufloor are RHOMIN*r^-1.5 and UUMIN*r^-2.5 at the zone centre */
int or_fixup1zone(double *pr, const or_geom *g, double rhofloor, double ufloor, double gammamax)
{
  double prf[2];
  int changed = 0;
  prf[0] = rhofloor;
  prf[1] = ufloor;
  if (prf[0] < 1.e-20) prf[0] = 1.e-20; /* RHOMINLIMIT, initbase.c:461 */
  if (prf[1] < 1.e-20) prf[1] = 1.e-20; /* UUMINLIMIT */
  if (prf[0] > pr[OR_RHO]) { pr[OR_RHO] = prf[0]; changed = 1; }
  if (prf[1] > pr[OR_UU]) { pr[OR_UU] = prf[1]; changed = 1; }
  if (or_limit_gamma(gammamax, pr, g) == -1) changed = 1;
  return changed;
}

/* fixup.c:181 diag_fixup, finite-volume branch (fixup.c:240-248): the
   carried conserved state u takes U(pr) - U(pr0), each U passed through
   UtoU from UDIAG to UEVOLVE (divided, then multiplied by gdet).  The
   update sits inside ENERREGIONLOOP (fixup.c:208) and an active zone is
   "correctable" in every region, so it is applied once per energy region:
   NUMENERREGIONS is 3 with the torus file's DOJETDIAG 1 (global.h:436) */
void or_diag_fixup_fv(const double *pr0, const double *pr, const or_geom *g, double gam, double *u)
{
  double Ui[OR_NPR], Uf[OR_NPR], Upre[OR_NPR], Upost[OR_NPR];
  int pl, region;
  or_primtoU(pr0, g, gam, Ui);
  or_primtoU(pr, g, gam, Uf);
  for (pl = 0; pl < OR_NPR; pl++) {
    Upre[pl] = (Ui[pl] / g->g) * g->g;
    Upost[pl] = (Uf[pl] / g->g) * g->g;
  }
  for (region = 0; region < 3; region++)
    for (pl = 0; pl < OR_NPR; pl++) u[pl] += Upost[pl] - Upre[pl];
}

/* fixup.c:2271 inflow_check_rel4vel for the radial direction at zone centres;
   side<0: inner boundary (fix u^r>0), side>0: outer boundary (fix u^r<0) */
int or_inflow_check_r(double *pr, const or_geom *g, int side, double gammamax)
{
  double ucon[4], gamma, alpha, betacon, vsq;
  int j, k;
  if (ucon_calc(pr, g, ucon)) return 1;
  if (!((side < 0 && ucon[1] > 0.) || (side > 0 && ucon[1] < 0.))) return 0;
  if (gamma_calc(pr, g, &gamma)) return 1;
  pr[OR_U1] /= gamma;
  pr[OR_U2] /= gamma;
  pr[OR_U3] /= gamma;
  alpha = 1. / sqrt(-g->gcon[0][0]);
  betacon = g->gcon[0][1] * alpha * alpha;
  pr[OR_U1] = betacon / alpha;
  vsq = 0.;
  for (j = 1; j < 4; j++)
    for (k = 1; k < 4; k++) vsq += g->gcov[j][k] * pr[OR_U1 + j - 1] * pr[OR_U1 + k - 1];
  if (vsq < 0.0) {
    if (vsq > -DBL_EPSILON * 100.0) vsq = 0.0;
    else return 1;
  }
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gammamax * gammamax);
  gamma = 1. / sqrt(1. - vsq);
  pr[OR_U1] *= gamma;
  pr[OR_U2] *= gamma;
  pr[OR_U3] *= gamma;
  return -1;
}

/* ---- batch wrappers ---- */
static void load_geom(or_geom *g, int64_t i, const double *gcov, const double *gcon,
                      const double *gdet, const double *gcovpert)
{
  int a, b;
  for (a = 0; a < 4; a++) {
    for (b = 0; b < 4; b++) {
      g->gcov[a][b] = gcov[i * 16 + a * 4 + b];
      g->gcon[a][b] = gcon[i * 16 + a * 4 + b];
    }
    g->gcovpert[a] = gcovpert[i * 4 + a];
  }
  g->g = gdet[i];
}

void or_c2p_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                  const double *gcovpert, const double *U, double gam, double *pr_inout,
                  int32_t *pflag, int32_t *iters)
{
  int64_t i;
  for (i = 0; i < n; i++) {
    or_geom g;
    int fl, it;
    load_geom(&g, i, gcov, gcon, gdet, gcovpert);
    or_utoprimgen(&U[i * 8], &g, gam, &pr_inout[i * 8], &fl, &it);
    pflag[i] = fl;
    iters[i] = it;
  }
}

void or_primtoU_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                      const double *gcovpert, const double *pr, double gam, double *U)
{
  int64_t i;
  for (i = 0; i < n; i++) {
    or_geom g;
    load_geom(&g, i, gcov, gcon, gdet, gcovpert);
    or_primtoU(&pr[i * 8], &g, gam, &U[i * 8]);
  }
}

/* fixup.c:1562-1568 set_density_floors, rescaletype 0, at zone centres of radius r */
void or_floors(int64_t n, const double *r, double rhomin, double uumin, double *out)
{
  int64_t i;
  for (i = 0; i < n; i++) {
    out[2 * i] = rhomin * pow(r[i], -1.5);
    out[2 * i + 1] = uumin * pow(r[i], -2.5);
  }
}
