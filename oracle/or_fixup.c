/*
 * or_fixup.c -- restatement of the reference's post_fixup (fixup.c:37) for
 * the torus problem: bound_pflag (boundsint.c:6 bound_pflag_user),
 * fixup_checksolution (fixup.c:460, CHECKSOLUTION 1) and fixup_utoprim
 * (fixup.c:669, UTOPRIMADJUST UTOPRIMAVG, GENERALAVERAGE 1, CAUSALAVG 0,
 * ORDERINDEPENDENT 1, TODONOGOOD NOGOODAVERAGE, WHICHPTOUSEWHENNOGOOD 0,
 * HOWTOAVG_WHEN_U2AVG MAXUPOSAVG, STEPOVERNEG* -1, DOCOUNTNEG* 1,
 * ADJUSTCONSERVEDQUANTITY 0).  Loops run in the reference's ZLOOP order
 * (i outer, k inner) and every flag update is visible to later zones, as in
 * the reference.  TEST INFRASTRUCTURE ONLY.
 */
#include "or_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FAILFIXED (-1)          /* global.h:782 */
#define NOFAIL 0                /* global.h:783 */
#define FAIL_CONVRET 101
#define FAIL_CONV 1
#define FAIL_CONVW 2
#define FAIL_CONVUTSQ 3
#define FAIL_CONVGUESSUTSQ 4
#define FAIL_CONVUTSQVERYBAD 5
#define FAIL_BADINVERTCOMPARE 6
#define FAIL_NANGUESS 7
#define FAIL_NANRESULT 8
#define FAIL_RHONEG 9
#define FAIL_UNEG 10
#define FAIL_RHOUNEG 11
#define FAIL_GAMMAPERC 12
#define FAIL_UPERC 13
#define FAIL_U2AVG1 14
#define FAIL_U2AVG2 15
#define GAMMAPERCDIFFMAX 2.0    /* init.fishmon.h:178 */
#define UPERCDIFFMAX 10.0       /* init.fishmon.h:179 */
#define MAXVOTES 8              /* fixup.c:451 */

/* boundsint.c:6 bound_pflag_user for BCtype X1 OUTFLOW, X2 POLARAXIS,
   X3 PERIODIC (init.fishmon.c:140-145), in the reference's order:
   periodic x3, outflow x1, then the polar mirror in x2 */
void or_bound_pflag(const or_grid *G, int *pf)
{
  or_dims d = or_get_dims(G);
  int i, j, k;
  if (d.n3 > 1)
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (j = -d.nb2; j < d.n2 + d.nb2; j++) {
        for (k = -d.nb3; k <= -1; k++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, i, j, d.n3 + k)];
        for (k = d.n3; k < d.n3 + d.nb3; k++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, i, j, k - d.n3)];
      }
  if (d.n1 > 1)
    for (j = -d.nb2; j < d.n2 + d.nb2; j++)
      for (k = -d.nb3; k < d.n3 + d.nb3; k++) {
        for (i = -d.nb1; i <= -1; i++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, 0, j, k)];
        for (i = d.n1; i < d.n1 + d.nb1; i++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, d.n1 - 1, j, k)];
      }
  if (d.n2 > 1)
    for (i = -d.nb1; i < d.n1 + d.nb1; i++)
      for (k = -d.nb3; k < d.n3 + d.nb3; k++) {
        for (j = -d.nb2; j <= -1; j++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, i, -1 - j, k)];
        for (j = d.n2; j < d.n2 + d.nb2; j++) pf[or_zidx(&d, i, j, k)] = pf[or_zidx(&d, i, 2 * d.n2 - 1 - j, k)];
      }
}

/* fixup.c:460 fixup_checksolution */
static int checksolution(const or_grid *G, const or_dims *d, const double *p, int *pf)
{
  int s1 = d->n1 > 1, s2 = d->n2 > 1, s3 = d->n3 > 1;
  int i, j, k, l;
  /* gammacheck over DQZLOOP: one layer beyond the active zones */
  int e1 = d->n1 + 2 * s1, e2 = d->n2 + 2 * s2, e3 = d->n3 + 2 * s3;
  double *gc = malloc(sizeof(double) * e1 * e2 * e3);
#define GC(a, b, c) gc[(((a) + s1) * e2 + ((b) + s2)) * e3 + ((c) + s3)]
  for (i = -s1; i < d->n1 + s1; i++)
    for (j = -s2; j < d->n2 + s2; j++)
      for (k = -s3; k < d->n3 + s3; k++) {
        or_geom g;
        or_load_geom(G, d, i, j, OL_CENT, &g);
        if (or_gamma_calc(&p[or_zidx(d, i, j, k) * 8], &g, &GC(i, j, k))) { free(gc); return 1; }
      }
  for (i = 0; i < d->n1; i++)
    for (j = 0; j < d->n2; j++)
      for (k = 0; k < d->n3; k++) {
        if (GC(i, j, k) >= 2.0) {
          int ip = i + s1, im = i - s1, jp = j + s2, jm = j - s2;
          /* neighbours in the reference's order; entry 6 repeats entry 4 */
          int ni[MAXVOTES] = {i, i, ip, im, ip, ip, ip, im};
          int nj[MAXVOTES] = {jp, jm, j, j, jp, jm, jp, jm};
          double percdiff[MAXVOTES], percdiff2[MAXVOTES];
          double uu = p[or_zidx(d, i, j, k) * 8 + OR_UU], gg = GC(i, j, k);
          int vote0 = 0, vote1 = 0, numvotes = 0;
          for (l = 0; l < MAXVOTES; l++) {
            int good = pf[or_zidx(d, ni[l], nj[l], k)] <= NOFAIL;
            percdiff[l] = good ? (GC(ni[l], nj[l], k) / gg) : -1;
            percdiff2[l] = good ? (p[or_zidx(d, ni[l], nj[l], k) * 8 + OR_UU] / uu) : -1;
          }
          for (l = 0; l < MAXVOTES; l++) {
            if (percdiff[0] >= 0.0) { /* the reference tests entry 0 for every vote */
              if ((fabs(percdiff[l]) > GAMMAPERCDIFFMAX) || (fabs(percdiff[l]) < 1.0 / GAMMAPERCDIFFMAX)) vote0++;
              if ((fabs(percdiff2[l]) > UPERCDIFFMAX) || (fabs(percdiff2[l]) < 1.0 / UPERCDIFFMAX)) vote1++;
              numvotes++;
            }
          }
          if (vote0 > numvotes * 0.5) pf[or_zidx(d, i, j, k)] = FAIL_GAMMAPERC;
          if (vote1 > numvotes * 0.5) pf[or_zidx(d, i, j, k)] = FAIL_UPERC;
        }
      }
#undef GC
  free(gc);
  return 0;
}

/* the accounting branch of fixup_utoprim (fixup.c:1345-1420): whether a
   zone's flag is reset to FAILFIXED; -1 = no, 1 = yes, 2 = unknown code */
static int counted(int f, int finalstep)
{
  if (f == NOFAIL) return -1;
  if (f == FAIL_CONV || f == FAIL_CONVGUESSUTSQ || f >= FAIL_CONVRET || f == FAIL_CONVW ||
      f == FAIL_CONVUTSQVERYBAD || f == FAIL_NANGUESS || f == FAIL_NANRESULT || f == FAIL_BADINVERTCOMPARE ||
      f == FAIL_CONVUTSQ)
    return 1;
  if (f == FAIL_RHONEG || f == FAIL_UNEG || f == FAIL_U2AVG1 || f == FAIL_U2AVG2 || f == FAIL_RHOUNEG)
    return finalstep ? 1 : -1;
  if (f == FAIL_GAMMAPERC || f == FAIL_UPERC) return 1;
  return 2;
}

/* fixup.c:669 fixup_utoprim */
static int fixup_utoprim(const or_grid *G, const or_dims *d, double *pv, int *pf, int finalstep, double *ucons)
{
  int s1 = d->n1 > 1, s2 = d->n2 > 1, s3 = d->n3 > 1;
  size_t nz = (size_t)d->m1 * d->m2 * d->m3;
  double *pa = malloc(nz * 8 * sizeof(double)); /* ptoavg: the values before any fix */
  int i, j, k, pl, qq;
  memcpy(pa, pv, nz * 8 * sizeof(double));
  for (i = 0; i < d->n1; i++)
    for (j = 0; j < d->n2; j++)
      for (k = 0; k < d->n3; k++) {
        ptrdiff_t z = or_zidx(d, i, j, k);
        int f = pf[z];
        if (f != 0) {
          int fixed = 0;
          or_geom g;
          or_load_geom(G, d, i, j, OL_CENT, &g);
          /* STEPOVERNEG* == -1: negative densities are left as they are */
          if (f == FAIL_UNEG || f == FAIL_RHONEG || f == FAIL_RHOUNEG) fixed = 1;
          if (!fixed) {
            double gamma, mysum[2][OR_NPR], lastmin[OR_NPR], avg0[OR_NPR], avg1[OR_NPR];
            int startpl, endpl, looptype, numavg0 = 0, numavg1 = 0, numavg = 0;
            int rnx = s1 ? 3 : 1, rny = s2 ? 3 : 1, numupairs;
            if (or_gamma_calc(&pa[z * 8], &g, &gamma)) { free(pa); return 1; }
            if (gamma > G->gammamax) gamma = G->gammamax;
            if (f == FAIL_U2AVG1 || f == FAIL_U2AVG2) { startpl = OR_UU; endpl = OR_UU; looptype = 1; }
            else { startpl = OR_RHO; endpl = OR_U3; looptype = 0; }
            for (pl = startpl; pl <= endpl; pl++) { mysum[0][pl] = 0.0; mysum[1][pl] = 0.0; lastmin[pl] = 1E50; }
            numupairs = (int)(pow(3, s1 + s2 + s3) - 1) / 2;
            for (qq = 0; qq < numupairs; qq++) {
              int ii = (qq % rnx) - s1, jj = ((qq % (rnx * rny)) / rnx) - s2, kk = (qq / (rnx * rny)) - s3;
              ptrdiff_t za = or_zidx(d, i + ii, j + jj, k + kk), zb = or_zidx(d, i - ii, j - jj, k - kk);
              int thisnotfail = pf[za] <= NOFAIL, thatnotfail = pf[zb] <= NOFAIL;
              numavg1 += (pa[za * 8 + OR_UU] > 0.0);
              numavg1 += (pa[zb * 8 + OR_UU] > 0.0);
              numavg0 += thisnotfail + thatnotfail;
              if (looptype == 0)
                for (pl = startpl; pl <= endpl; pl++)
                  mysum[qq % 2][pl] += pa[za * 8 + pl] * thisnotfail + pa[zb * 8 + pl] * thatnotfail;
              if (looptype == 1)
                for (pl = startpl; pl <= endpl; pl++) {
                  double ft = pa[za * 8 + pl];
                  if (ft >= 0.0) lastmin[pl] = (lastmin[pl] > ft) ? ft : lastmin[pl];
                  ft = pa[zb * 8 + pl];
                  if (ft >= 0.0) lastmin[pl] = (lastmin[pl] > ft) ? ft : lastmin[pl];
                }
            }
            if (looptype == 0) {
              if (numavg0 != 0)
                for (pl = startpl; pl <= endpl; pl++) avg0[pl] = (mysum[0][pl] + mysum[1][pl]) / ((double)(numavg0));
              if (numavg0 != 0)
                for (pl = startpl; pl <= endpl; pl++) pv[z * 8 + pl] = avg0[pl];
              numavg = numavg0;
            } else {
              for (pl = startpl; pl <= endpl; pl++) avg1[pl] = lastmin[pl];
              if (numavg1 != 0)
                for (pl = startpl; pl <= endpl; pl++) pv[z * 8 + pl] = avg1[pl];
              numavg = numavg1;
            }
            if (f == FAIL_U2AVG2) numavg++;
            if (numavg == 0) {
              /* NOGOODAVERAGE: fixup.c:1288, AVG4_1 and AVG4_2 of fixup.c:579-580 */
              int ip = i + s1, im = i - s1, jp = j + s2, jm = j - s2;
              for (pl = startpl; pl <= endpl; pl++) {
                double a41 = 0.25 * ((pa[or_zidx(d, i, jp, k) * 8 + pl] + pa[or_zidx(d, i, jm, k) * 8 + pl]) +
                                     (pa[or_zidx(d, im, j, k) * 8 + pl] + pa[or_zidx(d, ip, j, k) * 8 + pl]));
                double a42 = 0.25 * ((pa[or_zidx(d, ip, jp, k) * 8 + pl] + pa[or_zidx(d, ip, jm, k) * 8 + pl]) +
                                     (pa[or_zidx(d, im, jp, k) * 8 + pl] + pa[or_zidx(d, im, jm, k) * 8 + pl]));
                pv[z * 8 + pl] = 0.5 * (a41 + a42);
              }
            }
            if (or_limit_gamma(gamma, &pv[z * 8], &g) >= 1) { free(pa); return 1; }
          }
          {
            int c = counted(f, finalstep);
            if (c == 2) { free(pa); return 1; }
            if (c == 1) pf[z] = FAILFIXED;
            /* fixup.c:1428 diag_fixup, which moves the finite-volume
               conserved state on the final substep */
            if (c == 1 && ucons && finalstep) or_diag_fixup_fv(&pa[z * 8], &pv[z * 8], &g, G->gam, &ucons[z * 8]);
          }
        }
      }
  free(pa);
  return 0;
}

/* fixup.c:37 post_fixup with SIMULBCCALC -1: one stage */
int or_post_fixup(const or_grid *G, double *p, int *pflag, int finalstep)
{
  or_dims d = or_get_dims(G);
  or_bound_pflag(G, pflag);
  if (checksolution(G, &d, p, pflag)) return 1;
  return fixup_utoprim(G, &d, p, pflag, finalstep, NULL);
}

int or_post_fixup_fv(const or_grid *G, double *p, int *pflag, int finalstep, double *ucons)
{
  or_dims d = or_get_dims(G);
  or_bound_pflag(G, pflag);
  if (checksolution(G, &d, p, pflag)) return 1;
  return fixup_utoprim(G, &d, p, pflag, finalstep, ucons);
}
