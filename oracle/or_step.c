/*
 * or_step.c -- grid-level restatement of one RK2 step of
 * pseudotensor/harm_finalwham (TIMEORDER 2, NOENOFLUX, FLUXCTTOTH):
 * ghost fill, reconstruction, Riemann fluxes, flux-CT, conserved update,
 * Noble 2D inversion, fixup1zone and post_fixup.  TEST INFRASTRUCTURE ONLY.
 */
#include "or_internal.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BIG (1.e+50) /* global.h:1040 */

static double rmax(double a, double b) { if (a > b) return a; else return b; } /* phys.c:1797 */
static double rmin(double a, double b) { if (a > b) return b; else return a; } /* phys.c:1804 */
static double rsign(double a) { if (a > 0.0) return 1.0; else if (a < 0.0) return -1.0; else return 0.0; } /* phys.c:1778 */

/* global.h:361-362 */
#define MINMODB(a, b) ((fabs(a) < fabs(b)) ? (a) : (b))
#define MINMOD(a, b) (((a) * (b) <= 0) ? 0.0 : MINMODB(a, b))

static int dir_active(const or_dims *d, int dir)
{
  return (dir == 1) ? (d->n1 > 1) : (dir == 2) ? (d->n2 > 1) : (d->n3 > 1);
}

/* ------------------------------------------------------------------ */
/* bound_prim: bounds.fishmon.c:29 bound_prim_user (EXTRAP 0, POLEDEATH 1) */
/* ------------------------------------------------------------------ */
int or_bound_prim(const or_grid *G, double *p)
{
  or_dims d = or_get_dims(G);
  int i, j, k, pl;
  or_geom g;
  if (d.n1 > 1) {
    /* inner radial boundary: copy i=0, then inflow check and gamma limit */
    for (j = 0; j < d.n2; j++)
      for (k = 0; k < d.n3; k++) {
        for (i = -d.nb1; i <= -1; i++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, 0, j, k) * 8 + pl];
        for (i = -d.nb1; i <= -1; i++) {
          or_load_geom(G, &d, i, j, OL_CENT, &g);
          if (or_inflow_check_r(&p[or_zidx(&d, i, j, k) * 8], &g, -1, G->gammamax) > 0) return 1;
          if (or_limit_gamma(G->gammamax, &p[or_zidx(&d, i, j, k) * 8], &g) > 0) return 1;
        }
      }
    /* outer radial boundary */
    for (j = 0; j < d.n2; j++)
      for (k = 0; k < d.n3; k++) {
        for (i = d.n1; i < d.n1 + d.nb1; i++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, d.n1 - 1, j, k) * 8 + pl];
        for (i = d.n1; i < d.n1 + d.nb1; i++) {
          or_load_geom(G, &d, i, j, OL_CENT, &g);
          if (or_inflow_check_r(&p[or_zidx(&d, i, j, k) * 8], &g, +1, G->gammamax) > 0) return 1;
          if (or_limit_gamma(G->gammamax, &p[or_zidx(&d, i, j, k) * 8], &g) > 0) return 1;
        }
      }
  }
  if (d.n2 > 1) {
    /* polar axis: mirror copy (POSDEFMETRIC 0 leaves signs) then POLEDEATH */
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (k = 0; k < d.n3; k++)
        for (j = -d.nb2; j <= -1; j++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, i, -j - 1, k) * 8 + pl];
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (k = 0; k < d.n3; k++)
        for (j = -d.nb2; j < 0 + 1; j++) {
          p[or_zidx(&d, i, j, k) * 8 + OR_U2] *= 0;
          p[or_zidx(&d, i, j, k) * 8 + OR_B2] *= 0.;
        }
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (k = 0; k < d.n3; k++)
        for (j = d.n2; j < d.n2 + d.nb2; j++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, i, 2 * d.n2 - 1 - j, k) * 8 + pl];
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (k = 0; k < d.n3; k++)
        for (j = d.n2 - 1; j <= d.n2 - 1 + d.nb2; j++) {
          p[or_zidx(&d, i, j, k) * 8 + OR_U2] *= 0;
          p[or_zidx(&d, i, j, k) * 8 + OR_B2] *= 0.;
        }
  }
  if (d.n3 > 1) {
    /* periodic phi */
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (j = -d.nb2; j < d.n2 + d.nb2; j++) {
        for (k = -d.nb3; k <= -1; k++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, i, j, d.n3 + k) * 8 + pl];
        for (k = d.n3; k < d.n3 + d.nb3; k++)
          for (pl = 0; pl < OR_NPR; pl++)
            p[or_zidx(&d, i, j, k) * 8 + pl] = p[or_zidx(&d, i, j, k - d.n3) * 8 + pl];
      }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* reconstruction                                                      */
/* ------------------------------------------------------------------ */

/* interppoint.c:387 slope_lim_3points: MC (3), VANL (1), MINM (2), DONOR (0) */
static double slope3(int lim, double yl, double yc, double yr)
{
  double Dqm, Dqp, Dqc, s;
  if (lim == 3) {
    Dqm = 2.0 * (yc - yl);
    Dqp = 2.0 * (yr - yc);
    Dqc = 0.5 * (yr - yl);
    return MINMOD(Dqc, MINMOD(Dqm, Dqp));
  }
  if (lim == 1) {
    Dqm = (yc - yl);
    Dqp = (yr - yc);
    s = Dqm * Dqp;
    if (s <= 0.) return 0.;
    return (2.0 * s / (Dqm + Dqp));
  }
  if (lim == 2) {
    Dqm = 1.0 * (yc - yl);
    Dqp = 1.0 * (yr - yc);
    Dqc = 0.5 * (yr - yl);
    return MINMOD(MINMOD(Dqm, Dqc), Dqp);
  }
  return 0.0;
}

/* interppoint.c:24 slope_lim_pointtype + interppoint.c:387 slope_lim_3points;
   dq over the range of set_interppoint_loop (interppoint.c:159) */
static void mc_slopes(const or_dims *d, const double *p, int dir, int slim, double *dq)
{
  int i, j, k, pl;
  int is = -d->nb1, ie = d->n1 + d->nb1 - 1, js = -d->nb2, je = d->n2 + d->nb2 - 1;
  int ks = -d->nb3, ke = d->n3 + d->nb3 - 1;
  int di = (dir == 1), dj = (dir == 2), dk = (dir == 3);
  if (dir == 1) { is = -1; ie = d->n1; }
  if (dir == 2) { js = -1; je = d->n2; }
  if (dir == 3) { ks = -1; ke = d->n3; }
  for (i = is; i <= ie; i++)
    for (j = js; j <= je; j++)
      for (k = ks; k <= ke; k++)
        for (pl = 0; pl < OR_NPR; pl++) {
          double yl = p[or_zidx(d, i - di, j - dj, k - dk) * 8 + pl];
          double yc = p[or_zidx(d, i, j, k) * 8 + pl];
          double yr = p[or_zidx(d, i + di, j + dj, k + dk) * 8 + pl];
          dq[or_zidx(d, i, j, k) * 8 + pl] = slope3(slim, yl, yc, yr);
        }
}

/* interppoint.c:738 para4, PARA2LIM MC (global.h:426 for WHICHPARA PARA4);
   y points at the zone, y[-2..2] its stencil */
static void para4(const double *y, double *lout, double *rout)
{
  double dq0[3], *dq = dq0 + 1;
  double Dqm, Dqp, Dqc, s, l, r, qa, qd, qe;
  int mm;
  for (mm = -1; mm <= 1; mm++) {
    Dqm = 2.0 * (y[mm] - y[mm - 1]);
    Dqp = 2.0 * (y[mm + 1] - y[mm]);
    Dqc = 0.5 * (y[mm + 1] - y[mm - 1]);
    s = Dqm * Dqp;
    if (s <= 0.) dq[mm] = 0.;
    else dq[mm] = rmin(fabs(Dqc), rmin(fabs(Dqm), fabs(Dqp))) * rsign(Dqc);
  }
  l = 0.5 * (y[0] + y[-1]) - (dq[0] - dq[-1]) / 8.0;
  r = 0.5 * (y[1] + y[0]) - (dq[1] - dq[0]) / 8.0;
  l = rmax(rmin(y[0], y[-1]), l);
  l = rmin(rmax(y[0], y[-1]), l);
  r = rmax(rmin(y[0], y[1]), r);
  r = rmin(rmax(y[0], y[1]), r);
  qa = (r - y[0]) * (y[0] - l);
  qd = (r - l);
  qe = 6.0 * (y[0] - 0.5 * (l + r));
  if (qa <= 0.) {
    l = y[0];
    r = y[0];
  } else {
    if (qd * (qd - qe) < 0.0) l = 3.0 * y[0] - 2.0 * r;
    if (qd * (qd + qe) < 0.0) r = 3.0 * y[0] - 2.0 * l;
  }
  *lout = l;
  *rout = r;
}

/* interppoint.c:24 slope_lim_pointtype + interppoint.c:271 slope_lim_point
   (PARA): pleft/pright of every zone of set_interppoint_loop (interppoint.c:159) */
static void para_states(const or_dims *d, const double *p, int dir, double *pleft, double *pright)
{
  int i, j, k, pl, l;
  int is = -d->nb1, ie = d->n1 + d->nb1 - 1, js = -d->nb2, je = d->n2 + d->nb2 - 1;
  int ks = -d->nb3, ke = d->n3 + d->nb3 - 1;
  int di = (dir == 1), dj = (dir == 2), dk = (dir == 3);
  double y[5];
  if (dir == 1) { is = -1; ie = d->n1; }
  if (dir == 2) { js = -1; je = d->n2; }
  if (dir == 3) { ks = -1; ke = d->n3; }
  for (i = is; i <= ie; i++)
    for (j = js; j <= je; j++)
      for (k = ks; k <= ke; k++)
        for (pl = 0; pl < OR_NPR; pl++) {
          ptrdiff_t z = or_zidx(d, i, j, k);
          for (l = -2; l <= 2; l++) y[l + 2] = p[or_zidx(d, i + l * di, j + l * dj, k + l * dk) * 8 + pl];
          para4(y + 2, &pleft[z * 8 + pl], &pright[z * 8 + pl]);
        }
}

int or_reconstruct(const or_grid *G, const double *p, int dir, double *pleft, double *pright)
{
  or_dims d = or_get_dims(G);
  if (G->lim == 0) {
    mc_slopes(&d, p, dir, (G->slim >= 0 && G->slim <= 2) ? G->slim : 3, pleft);
    return 0;
  }
  if (G->lim == 2) {
    para_states(&d, p, dir, pleft, pright);
    return 0;
  }
  /* WENO5BND: one call per pencil, pencils over the full transverse extent
     (set_interp_loop, interpline.c:1205) */
  {
    int i, j, k;
    if (dir == 1) {
      for (j = -d.nb2; j < d.n2 + d.nb2; j++)
        for (k = -d.nb3; k < d.n3 + d.nb3; k++) or_weno5bnd_line(G, &d, p, 1, 0, j, k, pleft, pright);
    } else if (dir == 2) {
      for (i = -d.nb1; i < d.n1 + d.nb1; i++)
        for (k = -d.nb3; k < d.n3 + d.nb3; k++) or_weno5bnd_line(G, &d, p, 2, i, 0, k, pleft, pright);
    } else {
      for (i = -d.nb1; i < d.n1 + d.nb1; i++)
        for (j = -d.nb2; j < d.n2 + d.nb2; j++) or_weno5bnd_line(G, &d, p, 3, i, j, 0, pleft, pright);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* fluxes                                                              */
/* ------------------------------------------------------------------ */

/* fluxcompute.c:232 hllflux_compute, LaxF when cmax + cmin = 0 */
static void hllflux(double cmin, double cmax, double ctop, const double *U_l, const double *U_r, const double *F_l,
                    const double *F_r, double *F)
{
  int pl;
  if (cmax + cmin != 0.0) {
    for (pl = 0; pl < OR_NPR; pl++)
      F[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
  } else {
    for (pl = 0; pl < OR_NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
  }
}

/* fluxcompute.c:100 flux_compute for FORCEFLUX (2), MUSTAFLUX (3) and
   HLLLAXF1FLUX (4).  FORCE: forceflux_compute (fluxcompute.c:298) with
   cforce = ctop and GFORCE, MUSTACOEF 1.  MUSTA: mustaflux_compute
   (fluxcompute.c:398) -> musta1flux_compute (fluxcompute.c:538) with
   WHICHFLUX MUSTAHLL, one iteration, MUSTAVERSION 0; the energy-momentum
   fluxes are then the original HLL ones (fluxcompute.c:527-536). */
static int other_flux(const or_grid *G, int dir, const or_geom *g, const double *p_l, const double *p_r,
                      const double *U_l, const double *U_r, const double *F_l, const double *F_r, double cmin,
                      double cmax, double ctop, double *F)
{
  int pl, fl, it;
  if (G->flux == 4) {
    double dPoP = fabs((p_r[OR_UU] - p_l[OR_UU]) / (0.5 * (p_r[OR_UU] + p_l[OR_UU])));
    double f_s = (1.0 - 0.0) / (0.3 - 0.0) * (dPoP - 0.0) + 0.0;
    if (f_s > 1.0) f_s = 1.0;
    if (f_s < 0.0) f_s = 0.0;
    if (cmax + cmin != 0.0) {
      for (pl = 0; pl < OR_NPR; pl++)
        F[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin) * f_s +
                (0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]))) * (1.0 - f_s);
    } else {
      for (pl = 0; pl < OR_NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
    }
    return 0;
  }
  if (G->flux == 2) {
    double cforce = ctop, umid[OR_NPR], pmid[OR_NPR], fmid[OR_NPR];
    or_state q;
    for (pl = 0; pl < OR_NPR; pl++) umid[pl] = 0.5 * ((U_l[pl] + U_r[pl]) - (F_r[pl] - F_l[pl]) / cforce);
    for (pl = 0; pl < OR_NPR; pl++) pmid[pl] = 0.5 * (p_l[pl] + p_r[pl]);
    or_utoprimgen(umid, g, G->gam, pmid, &fl, &it);
    if (fl) {
      for (pl = 0; pl < OR_NPR; pl++) pmid[pl] = p_l[pl];
      or_utoprimgen(umid, g, G->gam, pmid, &fl, &it);
      if (fl) {
        for (pl = 0; pl < OR_NPR; pl++) pmid[pl] = p_r[pl];
        or_utoprimgen(umid, g, G->gam, pmid, &fl, &it);
      }
    }
    if (fl) {
      hllflux(cmin, cmax, ctop, U_l, U_r, F_l, F_r, F);
      return 0;
    }
    if (or_get_state(pmid, g, &q)) return 1;
    or_primtoflux(pmid, &q, dir, g, G->gam, fmid);
    for (pl = 0; pl < OR_NPR; pl++)
      F[pl] = ((fmid[pl]) + (1.0) * (0.5 * ((F_l[pl] + F_r[pl]) - cforce * (U_r[pl] - U_l[pl])))) / (1.0 + (1.0));
    return 0;
  }
  {
    double Fh[OR_NPR], Ul[OR_NPR], Ur[OR_NPR], pl_[OR_NPR], pr_[OR_NPR], Fl[OR_NPR], Fr[OR_NPR];
    double cmusta = ctop, cmax_l, cmin_l, cmax_r, cmin_r, lmin, lmax;
    or_state ql, qr;
    int domusta = 1;
    hllflux(cmin, cmax, ctop, U_l, U_r, F_l, F_r, Fh);
    for (pl = 0; pl < OR_NPR; pl++) {
      Ul[pl] = U_l[pl];
      Ur[pl] = U_r[pl];
      pl_[pl] = p_l[pl];
      pr_[pl] = p_r[pl];
    }
    for (pl = 0; pl < OR_NPR; pl++) {
      Ul[pl] -= (1.0) * (Fh[pl] - F_l[pl]) / cmusta;
      Ur[pl] -= (1.0) * (F_r[pl] - Fh[pl]) / cmusta;
    }
    or_utoprimgen(Ul, g, G->gam, pl_, &fl, &it);
    if (fl) domusta = 0;
    if (domusta) {
      or_utoprimgen(Ur, g, G->gam, pr_, &fl, &it);
      if (fl) domusta = 0;
    }
    if (domusta) {
      if (or_get_state(pl_, g, &ql)) return 1;
      or_primtoflux(pl_, &ql, dir, g, G->gam, Fl);
      if (or_get_state(pr_, g, &qr)) return 1;
      or_primtoflux(pr_, &qr, dir, g, G->gam, Fr);
      if (or_vchar(pl_, &ql, dir, g, G->gam, &cmax_l, &cmin_l)) return 1;
      if (or_vchar(pr_, &qr, dir, g, G->gam, &cmax_r, &cmin_r)) return 1;
      lmin = rmin(cmin_l, cmin_r);
      lmax = rmax(cmax_l, cmax_r);
      lmin = fabs(rmax(0., -lmin));
      lmax = fabs(rmax(0., lmax));
      hllflux(lmin, lmax, rmax(lmax, lmin), Ul, Ur, Fl, Fr, F);
    } else {
      for (pl = 0; pl < OR_NPR; pl++) F[pl] = Fh[pl];
    }
    for (pl = OR_UU; pl <= OR_U3; pl++) F[pl] = Fh[pl];
  }
  return 0;
}

/* fluxcompute.c:12 flux_compute_general (p2SFUevolve, get_wavespeeds, HLL/LaxF) */
static int flux_compute_general(const or_grid *G, int dir, const or_geom *g, const double *p_l,
                                const double *p_r, double *F, double *ctop)
{
  or_state ql, qr;
  double F_l[OR_NPR], F_r[OR_NPR], U_l[OR_NPR], U_r[OR_NPR];
  double cmax_l, cmin_l, cmax_r, cmin_r, lmin, lmax, cmin, cmax;
  int pl;
  if (or_get_state(p_l, g, &ql)) return 1;
  or_primtoflux(p_l, &ql, dir, g, G->gam, F_l);
  or_primtoflux(p_l, &ql, 0, g, G->gam, U_l);
  if (or_get_state(p_r, g, &qr)) return 1;
  or_primtoflux(p_r, &qr, dir, g, G->gam, F_r);
  or_primtoflux(p_r, &qr, 0, g, G->gam, U_r);
  /* wavespeeds.c:12 get_wavespeeds -> vchar.c:37, wavespeeds.c:325 cminmax_calc */
  if (or_vchar(p_l, &ql, dir, g, G->gam, &cmax_l, &cmin_l)) return 1;
  if (or_vchar(p_r, &qr, dir, g, G->gam, &cmax_r, &cmin_r)) return 1;
  lmin = rmin(cmin_l, cmin_r);
  lmax = rmax(cmax_l, cmax_r);
  lmin = fabs(rmax(0., -lmin));
  lmax = fabs(rmax(0., lmax));
  cmin = lmin;
  cmax = lmax;
  *ctop = rmax(lmax, lmin);
  if (G->flux == 0) {
    /* fluxcompute.c:78 LAXFCOMPUTE */
    for (pl = 0; pl < OR_NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - *ctop * (U_r[pl] - U_l[pl]));
  } else if (G->flux >= 2) {
    return other_flux(G, dir, g, p_l, p_r, U_l, U_r, F_l, F_r, cmin, cmax, *ctop, F);
  } else {
    /* fluxcompute.c:223 hllflux_compute */
    if (cmax + cmin != 0.0) {
      for (pl = 0; pl < OR_NPR; pl++)
        F[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
    } else {
      for (pl = 0; pl < OR_NPR; pl++)
        F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - *ctop * (U_r[pl] - U_l[pl]));
    }
  }
  return 0;
}

/* flux.c:469 fluxcalc_standard for one direction (faces of interp_loop_set,
   initbase.c:1027-1054) */
static int fluxcalc_dir(const or_grid *G, const or_dims *d, const double *p, int dir, double *F,
                        double *ndt, double *pl_buf, double *pr_buf)
{
  int i, j, k, pl;
  int di = (dir == 1), dj = (dir == 2), dk = (dir == 3);
  int s1 = (d->n1 > 1), s2 = (d->n2 > 1), s3 = (d->n3 > 1);
  int is = -s1, ie = d->n1 - 1 + s1, js = -s2, je = d->n2 - 1 + s2, ks = -s3, ke = d->n3 - 1 + s3;
  int loc = (dir == 1) ? OL_FACE1 : (dir == 2) ? OL_FACE2 : OL_FACE3;
  or_geom g;
  double p_l[OR_NPR], p_r[OR_NPR], ctop, dtij;

  if (G->fv) {
    /* initbase.c:949-993 with avgo = 0: every transverse ghost layer */
    is = -d->nb1; ie = d->n1 + d->nb1 - 1;
    js = -d->nb2; je = d->n2 + d->nb2 - 1;
    ks = -d->nb3; ke = d->n3 + d->nb3 - 1;
  }
  if (dir == 1) { is = 0; ie = d->n1 - 1 + s1; }
  if (dir == 2) { js = 0; je = d->n2 - 1 + s2; }
  if (dir == 3) { ks = 0; ke = d->n3 - 1 + s3; }
  *ndt = BIG;
  or_reconstruct(G, p, dir, pl_buf, pr_buf);
  for (i = is; i <= ie; i++)
    for (j = js; j <= je; j++)
      for (k = ks; k <= ke; k++) {
        ptrdiff_t zm = or_zidx(d, i - di, j - dj, k - dk), z = or_zidx(d, i, j, k);
        /* flux.c:753 getp2interplr */
        for (pl = 0; pl < OR_NPR; pl++) {
          if (G->lim == 0) {
            p_l[pl] = p[zm * 8 + pl] + 0.5 * pl_buf[zm * 8 + pl];
            p_r[pl] = p[z * 8 + pl] - 0.5 * pl_buf[z * 8 + pl];
          } else {
            p_l[pl] = pr_buf[zm * 8 + pl];
            p_r[pl] = pl_buf[z * 8 + pl];
          }
        }
        or_load_geom(G, d, i, j, loc, &g);
        /* flux.c:629 BDIRCONT */
        p_l[OR_B1 + dir - 1] = p_r[OR_B1 + dir - 1] = 0.5 * (p_l[OR_B1 + dir - 1] + p_r[OR_B1 + dir - 1]);
        if (flux_compute_general(G, dir, &g, p_l, p_r, &F[z * 8], &ctop)) return 1;
        dtij = G->cour * G->dx[dir] / ctop;
        if (dtij < *ndt) *ndt = dtij;
      }
  return 0;
}

/* fluxct.c:10 flux_ct, FLUXCTTOTH with CORNGDETVERSION 1 */
static void flux_ct(const or_grid *G, const or_dims *d, double *F1, double *F2, double *F3)
{
  int i, j, k, s1 = (d->n1 > 1), s2 = (d->n2 > 1), s3 = (d->n3 > 1);
  size_t nz = (size_t)d->m1 * d->m2 * d->m3;
  double *emf1 = calloc(nz, sizeof(double)), *emf2 = calloc(nz, sizeof(double)),
         *emf3 = calloc(nz, sizeof(double));
  double c1 = (s2 && s3) ? 0.25 : 0.5, c2 = (s1 && s3) ? 0.25 : 0.5, c3 = (s1 && s2) ? 0.25 : 0.5;
#define ZI(a, b, c) or_zidx(d, (a), (b), (c))
#define KM1(k) (s3 ? (k)-1 : (k))
#define JM1(j) (s2 ? (j)-1 : (j))
#define IM1(i) (s1 ? (i)-1 : (i))
#define KP1(k) (s3 ? (k) + 1 : (k))
#define JP1(j) (s2 ? (j) + 1 : (j))
#define IP1(i) (s1 ? (i) + 1 : (i))
  /* strip gdet from the field fluxes at the faces, fluxct.c:74-96 (only where
     the corner EMFs below read them) */
  for (i = -s1; i <= d->n1 - 1 + s1; i++)
    for (j = -s2; j <= d->n2 - 1 + s2; j++)
      for (k = -s3; k <= d->n3 - 1 + s3; k++) {
        ptrdiff_t z = ZI(i, j, k);
        int b;
        for (b = OR_B1; b <= OR_B3; b++) {
          if (s1) F1[z * 8 + b] /= or_gdet(G, d, i, j, OL_FACE1);
          if (s2) F2[z * 8 + b] /= or_gdet(G, d, i, j, OL_FACE2);
          if (s3) F3[z * 8 + b] /= or_gdet(G, d, i, j, OL_FACE3);
        }
      }
  /* corner EMFs (fluxct.c:148-198) times gdet at the corner (fluxct.c:202-218),
     over the corners the active-zone fluxes read */
  for (i = 0; i <= d->n1 - 1 + s1; i++)
    for (j = 0; j <= d->n2 - 1 + s2; j++)
      for (k = 0; k <= d->n3 - 1 + s3; k++) {
        ptrdiff_t z = ZI(i, j, k);
        double e1 = 0.0, e2 = 0.0, e3 = 0.0;
        if (s2 || s3) {
          if (s2 && s3)
            e1 = c1 * (+F2[z * 8 + OR_B3] + F2[ZI(i, j, KM1(k)) * 8 + OR_B3] - F3[z * 8 + OR_B2] -
                       F3[ZI(i, JM1(j), k) * 8 + OR_B2]);
          else if (s2)
            e1 = c1 * (+F2[z * 8 + OR_B3] + F2[ZI(i, j, KM1(k)) * 8 + OR_B3]);
          else
            e1 = c1 * (-F3[z * 8 + OR_B2] - F3[ZI(i, JM1(j), k) * 8 + OR_B2]);
        }
        if (s1 || s3) {
          if (s1 && s3)
            e2 = c2 * (+F3[z * 8 + OR_B1] + F3[ZI(IM1(i), j, k) * 8 + OR_B1] - F1[z * 8 + OR_B3] -
                       F1[ZI(i, j, KM1(k)) * 8 + OR_B3]);
          else if (s3)
            e2 = c2 * (+F3[z * 8 + OR_B1] + F3[ZI(IM1(i), j, k) * 8 + OR_B1]);
          else
            e2 = c2 * (-F1[z * 8 + OR_B3] - F1[ZI(i, j, KM1(k)) * 8 + OR_B3]);
        }
        if (s1 || s2) {
          if (s1 && s2)
            e3 = c3 * (+F1[z * 8 + OR_B2] + F1[ZI(i, JM1(j), k) * 8 + OR_B2] - F2[z * 8 + OR_B1] -
                       F2[ZI(IM1(i), j, k) * 8 + OR_B1]);
          else if (s1)
            e3 = c3 * (+F1[z * 8 + OR_B2] + F1[ZI(i, JM1(j), k) * 8 + OR_B2]);
          else
            e3 = c3 * (-F2[z * 8 + OR_B1] - F2[ZI(IM1(i), j, k) * 8 + OR_B1]);
        }
        emf1[z] = e1 * or_gdet(G, d, i, j, OL_CORN1);
        emf2[z] = e2 * or_gdet(G, d, i, j, OL_CORN2);
        emf3[z] = e3 * or_gdet(G, d, i, j, OL_CORN3);
      }
  /* rewrite the field fluxes from the EMFs, fluxct.c:720-781 */
  for (i = 0; i <= d->n1 - 1 + s1; i++)
    for (j = 0; j <= d->n2 - 1 + s2; j++)
      for (k = 0; k <= d->n3 - 1 + s3; k++) {
        ptrdiff_t z = ZI(i, j, k);
        if (s1 && j < d->n2 && k < d->n3) {
          F1[z * 8 + OR_B1] = 0.;
          F1[z * 8 + OR_B2] = 0.5 * (emf3[z] + emf3[ZI(i, JP1(j), k)]);
          F1[z * 8 + OR_B3] = -0.5 * (emf2[z] + emf2[ZI(i, j, KP1(k))]);
        }
        if (s2 && i < d->n1 && k < d->n3) {
          F2[z * 8 + OR_B1] = -0.5 * (emf3[z] + emf3[ZI(IP1(i), j, k)]);
          F2[z * 8 + OR_B2] = 0.;
          F2[z * 8 + OR_B3] = 0.5 * (emf1[z] + emf1[ZI(i, j, KP1(k))]);
        }
        if (s3 && i < d->n1 && j < d->n2) {
          F3[z * 8 + OR_B1] = 0.5 * (emf2[z] + emf2[ZI(IP1(i), j, k)]);
          F3[z * 8 + OR_B2] = -0.5 * (emf1[z] + emf1[ZI(i, JP1(j), k)]);
          F3[z * 8 + OR_B3] = 0.;
        }
      }
  free(emf1);
  free(emf2);
  free(emf3);
}

int or_fluxcalc(const or_grid *G, const double *p, double *F1, double *F2, double *F3, double *ndt)
{
  or_dims d = or_get_dims(G);
  size_t nz = (size_t)d.m1 * d.m2 * d.m3;
  double *bl = calloc(nz * 8, sizeof(double)), *br = calloc(nz * 8, sizeof(double));
  int rc = 0, dir;
  double *Fd[4] = {NULL, F1, F2, F3};
  for (dir = 1; dir <= 3; dir++) {
    ndt[dir - 1] = BIG;
    if (!dir_active(&d, dir)) continue;
    if (fluxcalc_dir(G, &d, p, dir, Fd[dir], &ndt[dir - 1], bl, br)) { rc = 1; break; }
  }
  if (!rc) flux_ct(G, &d, F1, F2, F3);
  free(bl);
  free(br);
  return rc;
}

/* ------------------------------------------------------------------ */
/* advance.c:58 advance_standard + step_ch.c:35 step_ch_simplempi       */
/* ------------------------------------------------------------------ */
/* advance.c:245 advance_finitevolume with avgscheme DONOR differs in
   taking Ui from the carried conserved state ui (ucons here) instead of
   U(pi), in its flux loops, and in handing ucum on to the next step */
static int advance(const or_grid *G, const double *pi, const double *pb, double *pf, double *uf,
                   double *ucum, double *ucons, const double CUf[4], const double Cunew[3], int timeorder,
                   double dt, double *ndt_out, int *pflag, int64_t *nfail, int64_t *iters)
{
  or_dims d = or_get_dims(G);
  size_t nz = (size_t)d.m1 * d.m2 * d.m3;
  double *F1 = calloc(nz * 8, sizeof(double)), *F2 = calloc(nz * 8, sizeof(double)),
         *F3 = calloc(nz * 8, sizeof(double));
  double ndt[3], idx1 = 1.0 / G->dx[1], idx2 = 1.0 / G->dx[2], idx3 = 1.0 / G->dx[3];
  int i, j, k, pl, rc = 0;
  int nto = (G->timeorder == 1 || G->timeorder == 3 || G->timeorder == 4) ? G->timeorder : 2;
  int finalstep = (timeorder == nto - 1);

  if (or_fluxcalc(G, pb, F1, F2, F3, ndt)) { rc = 1; goto done; }
  for (i = 0; i < d.n1; i++)
    for (j = 0; j < d.n2; j++)
      for (k = 0; k < d.n3; k++) {
        ptrdiff_t z = or_zidx(&d, i, j, k);
        or_geom g;
        double Ui[OR_NPR], dUgeom[OR_NPR], dU1[OR_NPR], dU2[OR_NPR], dU3[OR_NPR], dUr[OR_NPR];
        double Uf[OR_NPR], conn[4][4][4];
        int fl, it;
        if (timeorder == 0)
          for (pl = 0; pl < OR_NPR; pl++) uf[z * 8 + pl] = ucum[z * 8 + pl] = 0.0;
        or_load_geom(G, &d, i, j, OL_CENT, &g);
        if (ucons) memcpy(Ui, &ucons[z * 8], sizeof Ui);
        else or_primtoU(&pi[z * 8], &g, G->gam, Ui);
        memcpy(conn, &G->conn[((ptrdiff_t)(i + d.nb1) * d.m2 + (j + d.nb2)) * 64], sizeof(conn));
        or_source(&pb[z * 8], &g, (const double(*)[4][4])conn, G->gam, dUgeom);
        /* advance.c:741 flux2dUavg (FLUXB != FLUXCD) */
        for (pl = 0; pl < OR_NPR; pl++) {
          dU1[pl] = 0;
          dU2[pl] = 0;
          dU3[pl] = 0;
          if (d.n1 > 1) dU1[pl] = (-(F1[or_zidx(&d, i + 1, j, k) * 8 + pl] - F1[z * 8 + pl]) * idx1);
          if (d.n2 > 1) dU2[pl] = (-(F2[or_zidx(&d, i, j + 1, k) * 8 + pl] - F2[z * 8 + pl]) * idx2);
          if (d.n3 > 1) dU3[pl] = (-(F3[or_zidx(&d, i, j, k + 1) * 8 + pl] - F3[z * 8 + pl]) * idx3);
          dUr[pl] = dU1[pl] + dU2[pl] + dU3[pl];
        }
        /* advance.c:922 dUtoU */
        for (pl = 0; pl < OR_NPR; pl++)
          uf[z * 8 + pl] = Uf[pl] =
              CUf[0] * Ui[pl] + CUf[1] * uf[z * 8 + pl] + CUf[2] * dt * (dUr[pl] + dUgeom[pl]);
        for (pl = 0; pl < OR_NPR; pl++)
          ucum[z * 8 + pl] += Cunew[0] * Ui[pl] + Cunew[1] * dt * (dUr[pl] + dUgeom[pl]) + Cunew[2] * Uf[pl];
        for (pl = 0; pl < OR_NPR; pl++) pf[z * 8 + pl] = pb[z * 8 + pl];
        if (finalstep) or_utoprimgen(&ucum[z * 8], &g, G->gam, &pf[z * 8], &fl, &it);
        else or_utoprimgen(Uf, &g, G->gam, &pf[z * 8], &fl, &it);
        pflag[z] = fl;
        if (fl) (*nfail)++;
        *iters += it;
        if (finalstep) {
          const double *fz = &G->floors[((ptrdiff_t)(i + d.nb1) * d.m2 + (j + d.nb2)) * 2];
          double pr0[OR_NPR];
          memcpy(pr0, &pf[z * 8], sizeof pr0);
          /* advance.c:322-330: the next step's ui is this step's ucum,
             which fixup1zone's diag_fixup (fixup.c:443) moves with the floors */
          if (ucons) memcpy(&ucons[z * 8], &ucum[z * 8], 8 * sizeof(double));
          if (or_fixup1zone(&pf[z * 8], &g, fz[0], fz[1], G->gammamax) && ucons)
            or_diag_fixup_fv(pr0, &pf[z * 8], &g, G->gam, &ucons[z * 8]);
        }
      }
  *ndt_out = G->defcon * 1. / (1. / ndt[0] + 1. / ndt[1] + 1. / ndt[2]);
done:
  free(F1);
  free(F2);
  free(F3);
  return rc;
}

static int step_impl(const or_grid *G, double *p, double *ucons, double *unew, double *dt_io, int64_t *nfail,
                     int64_t *iters)
{
  or_dims d = or_get_dims(G);
  size_t nz = (size_t)d.m1 * d.m2 * d.m3;
  double *pk = malloc(nz * 8 * sizeof(double));
  double *uf = calloc(nz * 8, sizeof(double)), *ucum = calloc(nz * 8, sizeof(double));
  int *pflag = calloc(nz, sizeof(int));
  double *pk2 = NULL;
  /* step_ch.c:506-520, TIMEORDER 2 midpoint */
  const double CUf2[2][4] = {{1.0, 0.0, 0.5, 0.0}, {1.0, 0.0, 1.0, 0.0}};
  const double Cunew2[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
  /* step_ch.c:451-474, TIMEORDER 4 */
  const double CUf4[4][4] = {{1.0, 0.0, 0.5, 0.0}, {1.0, 0.0, 0.5, 0.0}, {1.0, 0.0, 1.0, 0.0}, {1.0, 0.0, 1.0, 0.0}};
  const double Cunew4[4][3] = {{1.0, 1.0 / 6.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 6.0, 0.0}};
  double dt = *dt_io, ndt = BIG;
  int rc = 0, t;
  memcpy(pk, p, nz * 8 * sizeof(double));
  if (G->timeorder == 1 || G->timeorder == 3) {
    /* step_ch.c:476-494 TVD RK3: pb = p, pk, pk2; pf = pk, pk2, p;
       step_ch.c:546-557 Euler: pb = pf = p */
    const double CUf3[3][4] = {{1.0, 0.0, 1.0, 0.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, 0.0},
                               {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.0}};
    const double Cunew3[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    double *pbb[3], *pff[3];
    int n = G->timeorder;
    pk2 = malloc(nz * 8 * sizeof(double));
    memcpy(pk2, p, nz * 8 * sizeof(double));
    pbb[0] = p; pbb[1] = pk; pbb[2] = pk2;
    pff[0] = pk; pff[1] = pk2; pff[2] = p;
    if (n == 1) pff[0] = p;
    for (t = 0; t < n; t++) {
      const double *cn = (n == 1) ? Cunew3[2] : Cunew3[t];
      if (advance(G, p, pbb[t], pff[t], uf, ucum, ucons, CUf3[t], cn, t, dt, &ndt, pflag, nfail, iters) ||
          or_bound_prim(G, pff[t]) || or_post_fixup_fv(G, pff[t], pflag, t == n - 1, ucons) ||
          or_bound_prim(G, pff[t])) {
        rc = 1;
        goto done;
      }
    }
  } else if (G->timeorder == 4) {
    /* pi = p; pb = p, pk, pk2, pk; pf = pk, pk2, pk, p */
    double *pbb[4], *pff[4];
    pk2 = malloc(nz * 8 * sizeof(double));
    memcpy(pk2, p, nz * 8 * sizeof(double));
    pbb[0] = p; pbb[1] = pk; pbb[2] = pk2; pbb[3] = pk;
    pff[0] = pk; pff[1] = pk2; pff[2] = pk; pff[3] = p;
    for (t = 0; t < 4; t++) {
      if (advance(G, p, pbb[t], pff[t], uf, ucum, ucons, CUf4[t], Cunew4[t], t, dt, &ndt, pflag, nfail, iters) ||
          or_bound_prim(G, pff[t]) || or_post_fixup_fv(G, pff[t], pflag, t == 3, ucons) || or_bound_prim(G, pff[t])) {
        rc = 1;
        goto done;
      }
    }
  } else {
    /* substep 0: pi=p pb=p pf=pk */
    /* each substep: advance, bound_prim, post_fixup, bound_prim (step_ch.c:123-157) */
    if (advance(G, p, p, pk, uf, ucum, ucons, CUf2[0], Cunew2[0], 0, dt, &ndt, pflag, nfail, iters)) { rc = 1; goto done; }
    if (or_bound_prim(G, pk) || or_post_fixup_fv(G, pk, pflag, 0, ucons) || or_bound_prim(G, pk)) { rc = 1; goto done; }
    /* substep 1: pi=p pb=pk pf=p */
    if (advance(G, p, pk, p, uf, ucum, ucons, CUf2[1], Cunew2[1], 1, dt, &ndt, pflag, nfail, iters)) { rc = 1; goto done; }
    if (or_bound_prim(G, p) || or_post_fixup_fv(G, p, pflag, 1, ucons) || or_bound_prim(G, p)) { rc = 1; goto done; }
  }
  /* step_ch.c:243-247 */
  if (ndt > G->safe * dt) ndt = G->safe * dt;
  *dt_io = ndt;
  /* the unew restart_write dumps (restart.c:240): ucum of the last substep */
  if (unew) memcpy(unew, ucum, nz * 8 * sizeof(double));
done:
  free(pk);
  free(pk2);
  free(uf);
  free(ucum);
  free(pflag);
  return rc;
}

int or_step(const or_grid *G, double *p, double *dt_io, int64_t *nfail, int64_t *iters)
{
  if (G->fv) return 1;
  return step_impl(G, p, NULL, NULL, dt_io, nfail, iters);
}

int or_step_unew(const or_grid *G, double *p, double *unew, double *dt_io, int64_t *nfail, int64_t *iters)
{
  if (G->fv || !unew) return 1;
  return step_impl(G, p, NULL, unew, dt_io, nfail, iters);
}

int or_step_fv(const or_grid *G, double *p, double *ucons, double *dt_io, int64_t *nfail, int64_t *iters)
{
  if (!G->fv || !ucons) return 1;
  return step_impl(G, p, ucons, NULL, dt_io, nfail, iters);
}

void or_cons_from_prims(const or_grid *G, const double *p, double *ucons)
{
  or_dims d = or_get_dims(G);
  int i, j, k;
  for (i = 0; i < d.n1; i++)
    for (j = 0; j < d.n2; j++)
      for (k = 0; k < d.n3; k++) {
        ptrdiff_t z = or_zidx(&d, i, j, k);
        or_geom g;
        or_load_geom(G, &d, i, j, OL_CENT, &g);
        or_primtoU(&p[z * 8], &g, G->gam, &ucons[z * 8]);
      }
}
