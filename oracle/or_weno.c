/*
 * or_weno.c -- WENO5BND interface reconstruction of pseudotensor/harm_finalwham
 * (lim = WENO5BND with NOENOFLUX, VARTOINTERP PRIMTOINTERP, DEFAULT reduction),
 * restated in C.  TEST INFRASTRUCTURE ONLY (the checker).
 *
 * Per 1-D pencil (interpline.c:517 slope_lim_linetype_c2e):
 *   df lines and the dP/P indicator      interpline.c:2112, :2148
 *   smooth-monotonicity indicator        smonointerp.c:15 (C2E branch)
 *   order-3 WENO weights                 reconstructeno.c:1406, :1687, :1878, :1948
 *   weight ratios -> lower-order frac.   reconstructeno.c:1962
 *   post/pre-shock dP reduction          reconstructeno.c:650
 *   optimised (point) weights            reconstructeno.c:1617
 *   order-3 / order-2 blend              reconstructeno.c:2087
 *   stencil evaluation                   reconstructeno.c:730 eno_cvt
 *   blend with the ENO5 value            reconstructeno.c:470-490
 * The shock indicator (interpline.c:2025) is computed by the reference but
 * only read by the PPM reduction type, so it is not restated.
 */
#include "or_internal.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BIG (1.e+50)
#define SMALL (1.e-50)
#define WENO_EPSILON (1.e-26)
#define SQRT_WENO_EPSILON (1.e-13)
#define MIN_TRANSITION_RATIO (1.3)
#define MAX_TRANSITION_RATIO (1.6)
#define REDUCE_DP_FRACTION (0.5)
#define MAXLINE 4096

#define MAX(a, b) (((a) > (b)) ? (a) : (b)) /* global.h:750 */
#define MIN(a, b) (((a) < (b)) ? (a) : (b)) /* global.h:751 */
#define MINMODB(a, b) ((fabs(a) < fabs(b)) ? (a) : (b))
#define MINMOD(a, b) (((a) * (b) <= 0) ? 0.0 : MINMODB(a, b))

static double fsign(double a) { if (a > 0.0) return 1.0; else if (a < 0.0) return -1.0; else return 0.0; } /* phys.c:1778 */

/* reconstructeno.c:1920 get_sum_of_elements (SYMMETRIZE_WENO 1) */
static double sum_sym(int n, const double *a)
{
  int i, half = n / 2, odd = n % 2;
  double sum = 0.0;
  for (i = 0; i < half; i++) sum += a[i] + a[n - 1 - i];
  sum += a[i] * odd;
  return sum;
}

/* reconstructeno.c:1948 normalize_array */
static double normalize_array(int n, double eps, double *a)
{
  double sum = sum_sym(n, a);
  int i;
  for (i = 0; i < n; i++) a[i] /= (sum + eps);
  return sum;
}

/* reconstructeno_static.h c2l/c2r tables (Lagrange point interpolation to the cell faces) */
static const double c2l3[3][3] = {{15. / 8., -5. / 4., 3. / 8.}, {3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}};
static const double c2r3[3][3] = {{3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}, {3. / 8., -5. / 4., 15. / 8.}};
static const double c2l2[2][2] = {{3. / 2., -1. / 2.}, {1. / 2., 1. / 2.}};
static const double c2r2[2][2] = {{1. / 2., 1. / 2.}, {-1. / 2., 3. / 2.}};
static const double c2l5_mid[5] = {-5. / 128., 15. / 32., 45. / 64., -5. / 32., 3. / 128.};
static const double c2r5_mid[5] = {3. / 128., -5. / 32., 45. / 64., 15. / 32., -5. / 128.};
/* optimal point weights, reconstructeno_static.h c2e_optimal_weights_left/rightface */
static const double optl[2][3] = {{1. / 4., 3. / 4, 0.0}, {1. / 16., 5. / 8., 5. / 16.}};
static const double optr[2][3] = {{3. / 4., 1. / 4., 0.0}, {5. / 16., 5. / 8., 1. / 16.}};

typedef struct {
  double w[9];  /* unoptimised [0..order), left-face [order..2order), right-face [2order..3order) */
  int order, len;
  double lof, lof_dpp;
} wts;

/* reconstructeno.c:1406 compute_cf_stencil_weights (orders 2 and 3); u points at the cell */
static void cf_weights(int order, int min_index, int max_index, double mono0, double mono1, const double *u, wts *W)
{
  double si[3], fd[3], sd[3];
  int shift, i;
  int min_shift = order - 1 - max_index, max_shift = -min_index;
  W->lof = 0.;
  W->order = order;
  W->len = 3 * order;
  if (mono0 == 1) {
    double simpleweight = 1.0 / order;
    for (i = 0; i < order; i++) W->w[i] = simpleweight;
    return;
  } else if (mono0 == 0 && mono1 != 0) {
    W->lof = 1.0;
  }
  if (order == 2) {
    fd[0] = u[1] - u[0];
    fd[1] = u[0] - u[-1];
    for (shift = 0; shift <= order - 1; shift++) si[shift] = fd[shift] * fd[shift];
  } else {
    fd[0] = -(3 * u[0] + u[2]) + 4 * u[1];
    fd[1] = u[1] - u[-1];
    fd[2] = (3 * u[0] + u[-2]) - 4 * u[-1];
    sd[0] = (u[0] + u[2]) - 2 * u[1];
    sd[1] = (u[-1] + u[1]) - 2 * u[0];
    sd[2] = (u[0] + u[-2]) - 2 * u[-1];
    for (shift = 0; shift <= order - 1; shift++)
      si[shift] = 13. / 12. * sd[shift] * sd[shift] + 1. / 4. * fd[shift] * fd[shift];
  }
  {
    /* reconstructeno.c:1687 desensitise_smoothness_indicators (USE_NORMU) */
    double u2[5], normu, e;
    for (i = -order + 1; i < order; i++) u2[i + order - 1] = u[i] * u[i];
    normu = sum_sym(2 * order - 1, u2);
    e = normu * WENO_EPSILON;
    for (shift = 0; shift < order; shift++) si[shift] += e;
  }
  for (shift = 0; shift < order; shift++)
    if (shift < min_shift || shift > max_shift) si[shift] = 0.0;
  normalize_array(order, DBL_MIN, si);
  /* reconstructeno.c:1878 compute_weights_from_smoothness_indicators */
  for (shift = 0; shift <= order - 1; shift++) {
    double t = si[shift] + WENO_EPSILON / 1.;
    t *= t;
    W->w[shift] = 1. / t;
  }
  for (shift = 0; shift <= order - 1; shift++)
    if (shift < min_shift || shift > max_shift) W->w[shift] = 0.0;
  normalize_array(order, 0, W->w);
}

/* reconstructeno.c:1617 compute_optimized_cf_stencil_weights */
static void cf_optimized(int order, double mono0, wts *W)
{
  int shift;
  W->order = order;
  W->len = 3 * order;
  if (mono0 == 1) return;
  if (mono0 == -1) return;
  for (shift = 0; shift <= order - 1; shift++) {
    W->w[order + shift] = W->w[shift] * optl[order - 2][shift];
    W->w[2 * order + shift] = W->w[shift] * optr[order - 2][shift];
  }
  normalize_array(order, DBL_MIN, &W->w[order]);
  normalize_array(order, DBL_MIN, &W->w[2 * order]);
}

/* reconstructeno.c:730 eno_cvt; right = 0 for the left face (CVT_C2L), 1 for CVT_C2R */
static double eno_cvt(int right, const wts *W, const double *u)
{
  int order = W->order, shift, j;
  const double *sw = W->w + order * (right ? 2 : 1);
  double pu[3], su[3];
  for (shift = 0; shift < order; shift++) {
    for (j = 0; j < order; j++) {
      double m = (order == 3) ? (right ? c2r3[shift][j] : c2l3[shift][j]) : (right ? c2r2[shift][j] : c2l2[shift][j]);
      pu[j] = m * u[j - shift];
    }
    su[shift] = 0.0;
    if (sw[shift] != 0.0) su[shift] = sw[shift] * sum_sym(order, pu);
  }
  return sum_sym(order, su);
}

/* smonointerp.c:439 check_for_cusp_new */
static int check_for_cusp(const double *y, const double *df, const double *ddf)
{
  double norm, f1a, f1, f2, f3, f3a, f4, f4a, f5a, f5, f6, f7, f7a, f8, f8a;
  norm = (fabs(y[-2]) + fabs(y[-1]) + fabs(y[0]) + fabs(y[1]) + fabs(y[2]) + SMALL);
  norm = norm * SQRT_WENO_EPSILON;
  f1a = fabs(df[0]) - norm;
  f1 = ddf[-1] * fsign(df[0]) - norm;
  f2 = -ddf[0] * fsign(df[0]) - norm;
  f3 = -df[2] * fsign(df[0]) - norm;
  f3a = fabs(-df[2] - df[0]) - 0.25 * (fabs(df[2]) + fabs(df[0]));
  f4 = -df[1] * fsign(df[0]) - norm;
  f4a = fabs(-df[1] - df[0]) - 0.25 * (fabs(df[1]) + fabs(df[0]));
  if (f1a > 0 && f1 > 0 && f2 > 0 && ((f3 > 0 && f3a > 0) || (f4 > 0 && f4a > 0))) return 1;
  f5a = fabs(df[1]) - norm;
  f5 = -ddf[1] * fsign(df[1]) - norm;
  f6 = ddf[0] * fsign(df[1]) - norm;
  f7 = -df[-1] * fsign(df[1]) - norm;
  f7a = fabs(-df[-1] - df[1]) - 0.25 * (fabs(df[-1]) + fabs(df[1]));
  f8 = -df[0] * fsign(df[1]) - norm;
  f8a = fabs(-df[0] - df[1]) - 0.25 * (fabs(df[0]) + fabs(df[1]));
  if (f5a > 0 && f5 > 0 && f6 > 0 && ((f7 > 0 && f7a > 0) || (f8 > 0 && f8a > 0))) return 1;
  return 0;
}

/* smonointerp.c:296 compute_mono_indicator_point_eno5 */
static double mono_point_eno5(const double *yin, double epsilon)
{
  double ld[4], rd[4], mm, minabs = BIG, norm, cutoff, x;
  int n;
  ld[1] = (((-3 * yin[-2] + yin[2]) + (-10 * yin[-1] - 6 * yin[1])) + 18 * yin[0]) / 12.;
  ld[2] = (((11 * yin[-2] - yin[2]) + (-20 * yin[-1] + 4 * yin[1])) + 6 * yin[0]) / 12.;
  ld[3] = (((-3 * yin[-2] - yin[2]) + (10 * yin[-1] + 6 * yin[1])) - 12 * yin[0]) / 2.;
  rd[1] = (((-yin[-2] + 3 * yin[2]) + (6 * yin[-1] + 10 * yin[1])) - 18 * yin[0]) / 12.;
  rd[2] = (((-yin[-2] + 11 * yin[2]) + (4 * yin[-1] - 20 * yin[1])) + 6 * yin[0]) / 12.;
  rd[3] = (((yin[-2] + 3 * yin[2]) - (6 * yin[-1] + 10 * yin[1])) + 12 * yin[0]) / 2.;
  ld[1] = MINMOD(ld[1], yin[-1] - yin[-2]);
  rd[1] = MINMOD(rd[1], yin[2] - yin[1]);
  for (n = 1; n <= 3; n++) {
    mm = MINMOD(ld[n], rd[n]);
    minabs = MIN(minabs, fabs(mm));
  }
  norm = fabs(yin[0]) + ((fabs(yin[-1]) + fabs(yin[1])) + (fabs(yin[-2]) + fabs(yin[2])));
  cutoff = norm * epsilon + DBL_MIN;
  /* smonointerp.c:429 transition_function(x, 0, cutoff) */
  x = (minabs - 0.0) / (cutoff - 0.0);
  return MAX(0., MIN(1., x));
}

/* reconstructeno.c:826 c2e_simple_eno(5, ...) : the ENO5 point value at a face */
static double eno5_face(int right, const double *y)
{
  double dres[5];
  int k;
  for (k = 0; k < 5; k++) dres[k] = (right ? c2r5_mid[k] : c2l5_mid[k]) * y[k - 2];
  return sum_sym(5, dres);
}

/* interpline.c:2148 compute_df_line_new: the dP/P-like jump indicator at point num */
static double dpp_indicator(const or_grid *G, const or_dims *d, const double *p, int dir, int i0, int j0, int k0,
                            int num)
{
  double par[3], upar[3], rho = 0.0;
  int s;
  for (s = -1; s <= 1; s++) {
    int ii = i0 + (dir == 1) * (num + s), jj = j0 + (dir == 2) * (num + s), kk = k0 + (dir == 3) * (num + s);
    double pr[8];
    or_geom g;
    or_state q;
    int pl;
    for (pl = 0; pl < 8; pl++) pr[pl] = p[or_zidx(d, ii, jj, kk) * 8 + pl];
    or_load_geom(G, d, ii, jj, OL_CENT, &g);
    or_get_state(pr, &g, &q);
    {
      /* phys.c:541 compute_1plusud0 */
      double vcon[4], vsq, gvtt, alpha, uu0, p1;
      int a, b;
      for (a = 1; a < 4; a++) vcon[a] = q.ucon[a] / q.ucon[0];
      vsq = g.gcovpert[0];
      for (a = 1; a < 4; a++) vsq += 2.0 * g.gcov[0][a] * vcon[a];
      for (a = 1; a < 4; a++)
        for (b = 1; b < 4; b++) vsq += g.gcov[a][b] * vcon[a] * vcon[b];
      gvtt = g.gcov[0][0];
      alpha = 0.0;
      for (a = 1; a < 4; a++) alpha += g.gcov[a][0] * q.ucon[a];
      uu0 = q.ucon[0];
      p1 = alpha + ((1.0 - gvtt) * g.gcovpert[0] - uu0 * uu0 * vsq * gvtt * gvtt) / (1.0 - gvtt * uu0);
      par[s + 1] = p1 * (q.ucon[0]);
    }
    upar[s + 1] = G->gam * pr[OR_UU] * (q.ucon[0] * q.ucov[0]) + (G->gam - 1.0) * pr[OR_UU];
    if (s == 0) rho = pr[OR_RHO];
  }
  return (fabs(rho * (par[2] - par[0])) + fabs(upar[2] - upar[0])) / (fabs(rho * par[1]) + fabs(upar[1]) + DBL_MIN);
}

/* One pencil along dir through (i,j,k) (the dir index is ignored).  Writes
   pleft/pright (layout of p) for the cells -1..n of the pencil. */
void or_weno5bnd_line(const or_grid *G, const or_dims *d, const double *p, int dir, int i, int j, int k,
                      double *pleft, double *pright)
{
  int n = (dir == 1) ? d->n1 : (dir == 2) ? d->n2 : d->n3;
  int nbd = (dir == 1) ? d->nb1 : (dir == 2) ? d->nb2 : d->nb3;
  int bs = -nbd, be = n + nbd - 1, ps = -1, pe = n;
  static double ybuf[MAXLINE], dfb[2][MAXLINE], dPb[MAXLINE], m0b[MAXLINE], m1b[MAXLINE], eL[MAXLINE], eR[MAXLINE];
  static wts Wb[MAXLINE];
  double *y = ybuf + nbd, *df0 = dfb[0] + nbd, *df1 = dfb[1] + nbd, *dP = dPb + nbd;
  double *mono0 = m0b + nbd, *mono1 = m1b + nbd, *yl = eL + nbd, *yr = eR + nbd;
  wts *W = Wb + nbd;
  int pl, s;
  const int order = 3;  /* (preforder+1)/2 with preforder 5, interpline.c:1063 */
  int ws = bs + order - 1, wf = be - (order - 1);

  /* dP line (interpline.c:1180 via compute_df_line_new), one per pencil */
  for (s = bs + 1; s <= be - 1; s++) dP[s] = dpp_indicator(G, d, p, dir, i, j, k, s);

  for (pl = 0; pl < OR_NPR; pl++) {
    for (s = bs; s <= be; s++) {
      int ii = (dir == 1) ? s : i, jj = (dir == 2) ? s : j, kk = (dir == 3) ? s : k;
      y[s] = p[or_zidx(d, ii, jj, kk) * 8 + pl];
    }
    /* interpline.c:2112 compute_df_line */
    for (s = bs + 1; s <= be; s++) df0[s] = y[s] - y[s - 1];
    for (s = bs + 1; s <= be - 1; s++) df1[s] = df0[s + 1] - df0[s];
    /* smonointerp.c:15, C2E branch */
    for (s = bs + order - 1; s <= be - order + 1; s++) {
      if (check_for_cusp(&y[s], &df0[s], &df1[s])) {
        mono0[s] = 0;
        mono1[s] = 1;
      } else {
        double ind = mono_point_eno5(&y[s], SQRT_WENO_EPSILON);
        mono0[s] = ind;
        mono1[s] = ind;
        if (ind != 0) {
          yl[s] = eno5_face(0, &y[s]);
          yr[s] = eno5_face(1, &y[s]);
        }
      }
    }
    /* reconstructeno.c:261 eno_line_reconstruct: unoptimised weights */
    for (s = ws; s <= wf; s++) {
      cf_weights(order, bs - s, be - s, mono0[s], mono1[s], &y[s], &W[s]);
      W[s].lof = (mono0[s] == 1) ? 0. : W[s].lof;
      /* reconstructeno.c:702 mono_adjust_unoptimized_weights */
      if (W[s].order == 3 && mono0[s] > 0.0 && mono0[s] < 1.0) {
        double simpleweight = 1.0 / 3.0;
        int c;
        for (c = 0; c < 3; c++) W[s].w[c] = (1.0 - mono0[s]) * W[s].w[c] + mono0[s] * simpleweight;
      }
    }
    /* reconstructeno.c:1962 compute_weights_ratios */
    for (s = ps - 1; s <= pe + 1; s++) {
      double wl = BIG, wr = BIG, lof, wmin;
      int il = -(order - 1), ir = order - 1, rbs = bs - s, rbf = be - s;
      if (il >= rbs + order - 1 && il <= rbf - (order - 1))
        wl = (W[s].w[order - 1] + WENO_EPSILON) / (W[s - 2].w[0] + WENO_EPSILON);
      if (ir >= rbs + order - 1 && ir <= rbf - (order - 1))
        wr = (W[s].w[0] + WENO_EPSILON) / (W[s + 2].w[order - 1] + WENO_EPSILON);
      wmin = MIN(wl, wr);
      lof = (wmin - MIN_TRANSITION_RATIO) / (MAX_TRANSITION_RATIO - MIN_TRANSITION_RATIO);
      lof = MAX(0., MIN(lof, 1.0));
      W[s].lof = MAX(lof, W[s].lof);
      W[s].lof_dpp = W[s].lof;
    }
    /* reconstructeno.c:650 apply_additional_reduction_to_weights */
    for (s = ps; s <= pe; s++) {
      double lof = W[s].lof, dPi = dP[s];
      double shock = MIN((fabs(dPi) - 0.5 * REDUCE_DP_FRACTION) / (0.5 * REDUCE_DP_FRACTION), 1.);
      lof = MAX(lof, shock * W[s + 1].lof_dpp);
      lof = MAX(lof, shock * W[s - 1].lof_dpp);
      W[s].lof = lof;
    }
    for (s = ws; s <= wf; s++) cf_optimized(order, mono0[s], &W[s]);
    /* outputs: choose_weno_order (reconstructeno.c:2087) and eno_cvt */
    for (s = ps; s <= pe; s++) {
      double lof = W[s].lof, out[2];
      int c, side, nw = 1;
      wts hi = W[s], lo;
      if (0.0 < lof) {
        for (c = order; c < hi.len; c++) hi.w[c] *= (1. - lof);
        hi.order = order;
        cf_weights(2, bs - s, be - s, mono0[s], mono1[s], &y[s], &lo);
        cf_optimized(2, mono0[s], &lo);
        for (c = 0; c < lo.len; c++) lo.w[c] *= lof;
        nw = 2;
      }
      for (side = 0; side < 2; side++) {
        double *yo = side ? &yr[s] : &yl[s];
        if (mono0[s] == 0.0) *yo = 0.0;
        if (mono0[s] >= 0.0 && mono0[s] < 1.0) {
          double ov = 0.0;
          ov += eno_cvt(side, &hi, &y[s]);
          if (nw == 2) ov += eno_cvt(side, &lo, &y[s]);
          *yo = (1.0 - mono0[s]) * ov + mono0[s] * *yo;
        }
        out[side] = *yo;
      }
      {
        int ii = (dir == 1) ? s : i, jj = (dir == 2) ? s : j, kk = (dir == 3) ? s : k;
        ptrdiff_t z = or_zidx(d, ii, jj, kk);
        pleft[z * 8 + pl] = out[0];
        pright[z * 8 + pl] = out[1];
      }
    }
  }
}
