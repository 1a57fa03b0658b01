// WENO5BND interface reconstruction on gfx950 (interpline.c:517
// slope_lim_linetype_c2e with lim = WENO5BND, PRIMTOINTERP, DEFAULT reduction).
//
// The pencil algorithm has data dependencies along the line, evaluated in
// stages per point and primitive:
//   A  dP/P jump indicator (interpline.c:2148) from the zone's p u^t (1+u_t)
//      and the enthalpy-like term at the neighbours; the two per-zone terms do
//      not depend on the direction and are computed once per substep
//      (k_weno_prep), the indicator itself once per zone and direction
//      (k_weno_dpi), not once per primitive
//   C  smooth-monotonicity indicator + order-3 weights (smonointerp.c:15,
//      reconstructeno.c:1406, :702), from y[s-2..s+2]
//   D  weight ratios -> lower-order fraction (reconstructeno.c:1962), from
//      the weights at s-2, s, s+2
//   E  post/pre-shock reduction, optimised weights, order-3/order-2 blend,
//      stencil evaluation and the ENO5 blend (reconstructeno.c:650, :1617,
//      :2087, :730, :470-490) -> pleft/pright, from y[s-2..s+2] and the
//      fractions at s-1, s, s+1
// Two kernel shapes, both with the radial index i on consecutive lanes so
// that every global load and store is coalesced:
//   along i (k_weno_x): one lane per point and primitive; the s+-1, s+-2
//     neighbours come from other lanes of the wavefront (cross-lane
//     permutes), a 64-lane wavefront yields 58 interface pairs
//   along j or k (k_weno_m): one lane per (i, transverse index, primitive)
//     marches along the line keeping the stage values of the last points
//     in registers
// Only the pencils and points the fluxes read are computed: transverse
// indices -1..n (every ghost row for the finite-volume flux loops) and
// points -1..n along the line (weights at -4..n+3).
// All arithmetic is FP64 in the reference's operation order.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "hg_div.h"
#include "hg_kernels.h"
#include "hg_phys.h"

namespace hg {

constexpr double W_BIG = 1.e+50;
constexpr double W_SMALL = 1.e-50;
constexpr double WENO_EPS = 1.e-26;
constexpr double SQRT_WENO_EPS = 1.e-13;
constexpr double MIN_TR = 1.3, MAX_TR = 1.6;
constexpr double REDUCE_DP = 0.5;

__device__ __forceinline__ double wmax(double a, double b) { return ((a) > (b)) ? (a) : (b); }  // global.h:750
__device__ __forceinline__ double wmin(double a, double b) { return ((a) < (b)) ? (a) : (b); }  // global.h:751
__device__ __forceinline__ double wminmod(double a, double b) {
  return ((a) * (b) <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? (a) : (b));
}
__device__ __forceinline__ double wsign(double a) { return (a > 0.0) ? 1.0 : ((a < 0.0) ? -1.0 : 0.0); }

// reconstructeno.c:1920 get_sum_of_elements, symmetric pairing
__device__ __forceinline__ double sum3(double a0, double a1, double a2) {
  double s = 0.0;
  s += a0 + a2;
  s += a1 * 1;
  return s;
}
__device__ __forceinline__ double sum2(double a0, double a1) {
  double s = 0.0;
  s += a0 + a1;
  s += a1 * 0;
  return s;
}
__device__ __forceinline__ double sum5(const double *a) {
  double s = 0.0;
  s += a[0] + a[4];
  s += a[1] + a[3];
  s += a[2] * 1;
  return s;
}

// order-3 unoptimised weights (reconstructeno.c:1406) for a point far from
// the line ends; u points at the cell.  Returns the cusp-driven lower-order
// fraction (1 for a cusp, else 0).
template <bool SD>
__device__ __forceinline__ double weights3(double mono0, double mono1, const double *u, double *w) {
  if (mono0 == 1) {
    double sw = 1.0 / 3;
    w[0] = sw; w[1] = sw; w[2] = sw;
    return 0.;
  }
  double lof = (mono0 == 0 && mono1 != 0) ? 1.0 : 0.;
  double fd0 = -(3 * u[0] + u[2]) + 4 * u[1];
  double fd1 = u[1] - u[-1];
  double fd2 = (3 * u[0] + u[-2]) - 4 * u[-1];
  double sd0 = (u[0] + u[2]) - 2 * u[1];
  double sd1 = (u[-1] + u[1]) - 2 * u[0];
  double sd2 = (u[0] + u[-2]) - 2 * u[-1];
  double si[3];
  si[0] = 13. / 12. * sd0 * sd0 + 1. / 4. * fd0 * fd0;
  si[1] = 13. / 12. * sd1 * sd1 + 1. / 4. * fd1 * fd1;
  si[2] = 13. / 12. * sd2 * sd2 + 1. / 4. * fd2 * fd2;
  double u2[5];
#pragma unroll
  for (int i = -2; i < 3; i++) u2[i + 2] = u[i] * u[i];
  double e = sum5(u2) * WENO_EPS;
#pragma unroll
  for (int s = 0; s < 3; s++) si[s] += e;
  double ss = sum3(si[0], si[1], si[2]);
  // sums of squares and positive terms: no numerator is -0
  div_shared<3, SD>(si, ss + DBL_MIN);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double t = si[s] + WENO_EPS / 1.;
    t *= t;
    w[s] = 1.;
    div_shared<1, SD>(&w[s], t);
  }
  double sw = sum3(w[0], w[1], w[2]);
  div_shared<3, SD>(w, sw + 0);
  return lof;
}

// order-2 unoptimised + optimised weights (reconstructeno.c:1406 and :1617 with order 2)
template <bool SD>
__device__ __forceinline__ void weights2(double mono0, const double *u, double *wl, double *wr) {
  double w[2];
  if (mono0 == 1) {
    // equal weights; compute_optimized_cf_stencil_weights returns without
    // setting the optimised ones, which are then never read (mono0 == 1)
    wl[0] = wl[1] = wr[0] = wr[1] = 0.0;
    return;
  }
  double fd0 = u[1] - u[0];
  double fd1 = u[0] - u[-1];
  double si[2] = {fd0 * fd0, fd1 * fd1};
  double u2[3] = {u[-1] * u[-1], u[0] * u[0], u[1] * u[1]};
  double e = sum3(u2[0], u2[1], u2[2]) * WENO_EPS;
  si[0] += e;
  si[1] += e;
  double ss = sum2(si[0], si[1]);
  // squares, positive sums and their products with positive constants: no
  // numerator is -0
  div_shared<2, SD>(si, ss + DBL_MIN);
#pragma unroll
  for (int s = 0; s < 2; s++) {
    double t = si[s] + WENO_EPS / 1.;
    t *= t;
    w[s] = 1.;
    div_shared<1, SD>(&w[s], t);
  }
  double sw = sum2(w[0], w[1]);
  div_shared<2, SD>(w, sw + 0);
  wl[0] = w[0] * (1. / 4.);
  wl[1] = w[1] * (3. / 4);
  wr[0] = w[0] * (3. / 4.);
  wr[1] = w[1] * (1. / 4.);
  double sl = sum2(wl[0], wl[1]);
  div_shared<2, SD>(wl, sl + DBL_MIN);
  double sr = sum2(wr[0], wr[1]);
  div_shared<2, SD>(wr, sr + DBL_MIN);
}

// reconstructeno.c:730 eno_cvt for order 3, left (C2L) or right (C2R) face
__device__ __forceinline__ double cvt3(bool right, const double *sw, const double *u) {
  double su[3];
  // rows of the center-to-face matrices (Lagrange interpolation on 3 points)
  const double L[3][3] = {{15. / 8., -5. / 4., 3. / 8.}, {3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}};
  const double R[3][3] = {{3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}, {3. / 8., -5. / 4., 15. / 8.}};
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double pu0 = (right ? R[s][0] : L[s][0]) * u[0 - s];
    double pu1 = (right ? R[s][1] : L[s][1]) * u[1 - s];
    double pu2 = (right ? R[s][2] : L[s][2]) * u[2 - s];
    su[s] = 0.0;
    if (sw[s] != 0.0) su[s] = sw[s] * sum3(pu0, pu1, pu2);
  }
  return sum3(su[0], su[1], su[2]);
}

__device__ __forceinline__ double cvt2(bool right, const double *sw, const double *u) {
  double su[2];
  const double L[2][2] = {{3. / 2., -1. / 2.}, {1. / 2., 1. / 2.}};
  const double R[2][2] = {{1. / 2., 1. / 2.}, {-1. / 2., 3. / 2.}};
#pragma unroll
  for (int s = 0; s < 2; s++) {
    double pu0 = (right ? R[s][0] : L[s][0]) * u[0 - s];
    double pu1 = (right ? R[s][1] : L[s][1]) * u[1 - s];
    su[s] = 0.0;
    if (sw[s] != 0.0) su[s] = sw[s] * sum2(pu0, pu1);
  }
  return sum2(su[0], su[1]);
}

// smonointerp.c:439 check_for_cusp_new; y points at the cell
__device__ __forceinline__ bool cusp(const double *y) {
  double df[4], ddf[3];  // df[-1..2] -> df[0..3], ddf[-1..1] -> ddf[0..2]
#pragma unroll
  for (int a = -1; a <= 2; a++) df[a + 1] = y[a] - y[a - 1];
#pragma unroll
  for (int a = -1; a <= 1; a++) ddf[a + 1] = df[a + 2] - df[a + 1];
  double norm = (fabs(y[-2]) + fabs(y[-1]) + fabs(y[0]) + fabs(y[1]) + fabs(y[2]) + W_SMALL);
  norm = norm * SQRT_WENO_EPS;
  double d_1 = df[0], d0 = df[1], d1 = df[2], d2 = df[3];
  double f1a = fabs(d0) - norm;
  double f1 = ddf[0] * wsign(d0) - norm;
  double f2 = -ddf[1] * wsign(d0) - norm;
  double f3 = -d2 * wsign(d0) - norm;
  double f3a = fabs(-d2 - d0) - 0.25 * (fabs(d2) + fabs(d0));
  double f4 = -d1 * wsign(d0) - norm;
  double f4a = fabs(-d1 - d0) - 0.25 * (fabs(d1) + fabs(d0));
  if (f1a > 0 && f1 > 0 && f2 > 0 && ((f3 > 0 && f3a > 0) || (f4 > 0 && f4a > 0))) return true;
  double f5a = fabs(d1) - norm;
  double f5 = -ddf[2] * wsign(d1) - norm;
  double f6 = ddf[1] * wsign(d1) - norm;
  double f7 = -d_1 * wsign(d1) - norm;
  double f7a = fabs(-d_1 - d1) - 0.25 * (fabs(d_1) + fabs(d1));
  double f8 = -d0 * wsign(d1) - norm;
  double f8a = fabs(-d0 - d1) - 0.25 * (fabs(d0) + fabs(d1));
  if (f5a > 0 && f5 > 0 && f6 > 0 && ((f7 > 0 && f7a > 0) || (f8 > 0 && f8a > 0))) return true;
  return false;
}

// smonointerp.c:296 compute_mono_indicator_point_eno5
__device__ __forceinline__ double mono_eno5(const double *yin) {
  // the four divisions by 12 (their zero quotients' signs cannot reach the
  // result: the minmods below return a literal 0 for a zero argument)
  const double n[4] = {(((-3 * yin[-2] + yin[2]) + (-10 * yin[-1] - 6 * yin[1])) + 18 * yin[0]),
                        (((11 * yin[-2] - yin[2]) + (-20 * yin[-1] + 4 * yin[1])) + 6 * yin[0]),
                        (((-yin[-2] + 3 * yin[2]) + (6 * yin[-1] + 10 * yin[1])) - 18 * yin[0]),
                        (((-yin[-2] + 11 * yin[2]) + (4 * yin[-1] - 20 * yin[1])) + 6 * yin[0])};
  double q[4];
  div_const4(n, 12., rcp_refined(12.), q);
  double ld1 = q[0], ld2 = q[1], rd1 = q[2], rd2 = q[3];
  double ld3 = (((-3 * yin[-2] - yin[2]) + (10 * yin[-1] + 6 * yin[1])) - 12 * yin[0]) / 2.;
  double rd3 = (((yin[-2] + 3 * yin[2]) - (6 * yin[-1] + 10 * yin[1])) + 12 * yin[0]) / 2.;
  ld1 = wminmod(ld1, yin[-1] - yin[-2]);
  rd1 = wminmod(rd1, yin[2] - yin[1]);
  double minabs = W_BIG;
  minabs = wmin(minabs, fabs(wminmod(ld1, rd1)));
  minabs = wmin(minabs, fabs(wminmod(ld2, rd2)));
  minabs = wmin(minabs, fabs(wminmod(ld3, rd3)));
  double norm = fabs(yin[0]) + ((fabs(yin[-1]) + fabs(yin[1])) + (fabs(yin[-2]) + fabs(yin[2])));
  double cutoff = norm * SQRT_WENO_EPS + DBL_MIN;
  double x = (minabs - 0.0) / (cutoff - 0.0);
  return wmax(0., wmin(1., x));
}

// reconstructeno.c:826 c2e_simple_eno(5): ENO5 face value
__device__ __forceinline__ double eno5(bool right, const double *y) {
  const double L[5] = {-5. / 128., 15. / 32., 45. / 64., -5. / 32., 3. / 128.};
  const double R[5] = {3. / 128., -5. / 32., 45. / 64., 15. / 32., -5. / 128.};
  double d[5];
#pragma unroll
  for (int k = 0; k < 5; k++) d[k] = (right ? R[k] : L[k]) * y[k - 2];
  return sum5(d);
}

// ---------------------------------------------------------------------------
// the stages at one point
// ---------------------------------------------------------------------------
// C: y points at the cell, y[-2..2] are read
template <bool SD>
__device__ __forceinline__ void stage_c(const double *y, double w[3], double &m0, double &lof0) {
  double m1;
  if (cusp(y)) { m0 = 0; m1 = 1; }
  else { m0 = mono_eno5(y); m1 = m0; }
  lof0 = weights3<SD>(m0, m1, y, w);
  if (m0 > 0.0 && m0 < 1.0) {  // reconstructeno.c:702 mono_adjust_unoptimized_weights
    double simpleweight = 1.0 / 3.0;
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = (1.0 - m0) * w[c] + m0 * simpleweight;
  }
}

// D: reconstructeno.c:1962 compute_weights_ratios and the cusp fraction
__device__ __forceinline__ double stage_d(double w2_s, double w0_sm2, double w0_s, double w2_sp2, double lof0) {
  double wl = (w2_s + WENO_EPS) / (w0_sm2 + WENO_EPS);
  double wr = (w0_s + WENO_EPS) / (w2_sp2 + WENO_EPS);
  double lof = (wmin(wl, wr) - MIN_TR) / (MAX_TR - MIN_TR);
  lof = wmax(0., wmin(lof, 1.0));
  return wmax(lof, lof0);
}

// A: interpline.c:2148 compute_df_line_new, dP/P at s from s-1, s, s+1
__device__ __forceinline__ double dp_jump(double rho, double parm, double par0, double parp, double uparm,
                                          double upar0, double uparp) {
  return (fabs(rho * (parp - parm)) + fabs(uparp - uparm)) / (fabs(rho * par0) + fabs(upar0) + DBL_MIN);
}

// E: y points at the cell; lofm/lof0/lofp are the stage-D fractions at s-1, s, s+1
template <bool SD>
__device__ __forceinline__ void stage_e(const double *y, double m0, double w0, double w1, double w2, double lofm,
                                        double lof0, double lofp, double dPi, double &outl, double &outr) {
  double shock = wmin((fabs(dPi) - 0.5 * REDUCE_DP) / (0.5 * REDUCE_DP), 1.);
  double lof = lof0;
  lof = wmax(lof, shock * lofp);
  lof = wmax(lof, shock * lofm);
  if (m0 == 1) {
    outl = eno5(false, y);
    outr = eno5(true, y);
    return;
  }
  // compute_optimized_cf_stencil_weights (order 3)
  double wl[3] = {w0 * (1. / 16.), w1 * (5. / 8.), w2 * (5. / 16.)};
  double wr[3] = {w0 * (5. / 16.), w1 * (5. / 8.), w2 * (1. / 16.)};
  double sl = sum3(wl[0], wl[1], wl[2]), sr = sum3(wr[0], wr[1], wr[2]);
  div_shared<3, SD>(wl, sl + DBL_MIN);
  div_shared<3, SD>(wr, sr + DBL_MIN);
  double l2[2] = {0., 0.}, r2[2] = {0., 0.};
  bool two = (0.0 < lof);
  if (two) {
#pragma unroll
    for (int c = 0; c < 3; c++) { wl[c] *= (1. - lof); wr[c] *= (1. - lof); }
    weights2<SD>(m0, y, l2, r2);
    l2[0] *= lof; l2[1] *= lof; r2[0] *= lof; r2[1] *= lof;
  }
  double outv[2];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double prev = 0.0;
    if (m0 != 0.0) prev = eno5(side == 1, y);
    double ov = 0.0;
    ov += cvt3(side == 1, side ? wr : wl, y);
    if (two) ov += cvt2(side == 1, side ? r2 : l2, y);
    outv[side] = (1.0 - m0) * ov + m0 * prev;
  }
  outl = outv[0];
  outr = outv[1];
}

// ---------------------------------------------------------------------------
// stage A inputs per zone: par = (1 + u_t) u^t, upar = gam u u^t u_t + (gam-1) u
// over the box [-2, n+1] of every direction with more than one zone
// ---------------------------------------------------------------------------
__device__ __forceinline__ void weno_prep_body(const double *__restrict__ p, const double *__restrict__ geo,
                                               double *__restrict__ aux, const Dims &d, double gam, int wide,
                                               unsigned long long *__restrict__ ndt_reset, unsigned bid) {
  // the substep's flux kernels take the minimum of their ndt into these
  if (ndt_reset && bid == 0 && threadIdx.x < 3)
    ndt_reset[threadIdx.x] = (unsigned long long)__double_as_longlong(1.e+50);
  // wide (the finite-volume flux loops): every ghost layer
  const int e1 = d.n1 > 1 ? (wide ? d.nb1 : 2) : 0, e2 = d.n2 > 1 ? (wide ? d.nb2 : 2) : 0;
  const int e3 = d.n3 > 1 ? (wide ? d.nb3 : 2) : 0;
  const int b1 = d.n1 + 2 * e1, b2 = d.n2 + 2 * e2, b3 = d.n3 + 2 * e3;
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)b1 * b2 * b3) return;
  int i = -e1 + (int)(t % b1);
  int64_t r = t / b1;
  int k = -e3 + (int)(r % b3), j = -e2 + (int)(r / b3);
  int64_t z = d.idx(i, j, k);
  double pr[NPR];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pr[pl] = p[z + pl * d.spl];
  Geom G;
  {
    const double *b = geo + (int64_t)GL_CENT * GCOMP * d.gplane + d.gidx(i, j);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        G.gcov[a][c] = b[(a * 4 + c) * d.gplane];
        G.gcon[a][c] = b[(16 + a * 4 + c) * d.gplane];
      }
    G.gpert0 = b[32 * d.gplane];
    G.g = b[33 * d.gplane];
  }
  State q;
  get_state(pr, G, q);
  double p1 = one_plus_ud0(q, G);
  aux[z] = p1 * (q.ucon[0]);
  aux[z + d.spl] = gam * pr[UU] * (q.ucon[0] * q.ucov[0]) + (gam - 1.0) * pr[UU];
}

__global__ void __launch_bounds__(256) k_weno_prep(const double *__restrict__ p, const double *__restrict__ geo,
                                                   double *__restrict__ aux, Dims d, double gam, int wide,
                                                   unsigned long long *__restrict__ ndt_reset) {
  weno_prep_body(p, geo, aux, d, gam, wide, ndt_reset, blockIdx.x);
}

// ---------------------------------------------------------------------------
// stage A along DIR (DIR 0: every direction of the grid): dP/P at every
// zone of the box above whose two neighbours along that direction are in the
// box too, into aux plane 1 + direction.  It does not depend on the
// primitive reconstructed, so the reconstruction kernels of all 8 primitives
// read it from here.  When the directions run concurrently each direction's
// runs on that direction's stream, ahead of its reconstruction; otherwise
// one launch serves all three.
// Work-groups go to the 8 XCDs round robin by linear index; each XCD takes
// a contiguous eighth of the box instead, so that the neighbours a zone
// reads were fetched into that XCD's L2 by the blocks just before it.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void weno_dpi_body(const double *__restrict__ p, double *__restrict__ aux, const Dims &d,
                                              int wide, unsigned bid, unsigned nblk) {
  const int e1 = d.n1 > 1 ? (wide ? d.nb1 : 2) : 0, e2 = d.n2 > 1 ? (wide ? d.nb2 : 2) : 0;
  const int e3 = d.n3 > 1 ? (wide ? d.nb3 : 2) : 0;
  const int b1 = d.n1 + 2 * e1, b2 = d.n2 + 2 * e2, b3 = d.n3 + 2 * e3;
  const unsigned xcd = bid & 7u, q = bid >> 3, base = nblk >> 3, rem = nblk & 7u;
  const unsigned lb = xcd * base + (xcd < rem ? xcd : rem) + q;  // a bijection of [0, nblk)
  int64_t t = lb * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)b1 * b2 * b3) return;
  int i = -e1 + (int)(t % b1);
  int64_t r = t / b1;
  int k = -e3 + (int)(r % b3), j = -e2 + (int)(r / b3);
  const int64_t z = d.idx(i, j, k), P = d.spl;
  const double *par = aux, *upar = aux + P;
#pragma unroll
  for (int dir = 1; dir <= 3; dir++) {
    if (DIR != 0 && dir != DIR) continue;
    const int s = (dir == 1) ? i : (dir == 2) ? j : k, e = (dir == 1) ? e1 : (dir == 2) ? e2 : e3;
    const int n = (dir == 1) ? d.n1 : (dir == 2) ? d.n2 : d.n3;
    const int64_t sd = (dir == 1) ? 1 : (dir == 2) ? d.sj : d.sk;
    if (n > 1 && s > -e && s < n + e - 1)
      aux[(1 + dir) * P + z] =
          dp_jump(p[RHO * P + z], par[z - sd], par[z], par[z + sd], upar[z - sd], upar[z], upar[z + sd]);
  }
}

template <int DIR>
__global__ void __launch_bounds__(256) k_weno_dpi(const double *__restrict__ p, double *__restrict__ aux, Dims d,
                                                  int wide) {
  weno_dpi_body<DIR>(p, aux, d, wide, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// along i: one lane per point.  The rows (primitive fastest, then j, then
// k) are laid end to end, each as the n1 + 8 lanes of its points -4 .. n1+3
// (the outputs -1 .. n1 and the three points either side that stage C and D
// need); wavefront w takes lanes 58 w .. 58 w + 63 of that sequence and
// writes the outputs on its lanes 3 .. 60, whose neighbours within three
// lanes then belong to the same row and wavefront.  A short row therefore
// shares its wavefront with the next one instead of leaving lanes idle.
// ---------------------------------------------------------------------------
constexpr int XOUT = 58;

// wavefronts of the i-direction reconstruction (transverse halo widths s2, s3)
__host__ __device__ inline int64_t weno_x_waves(const Dims &d, int s2, int s3) {
  const int64_t rows = (int64_t)NPR * (d.n2 + 2 * s2) * (d.n3 + 2 * s3);
  return (rows * (d.n1 + 8) + XOUT - 1) / XOUT;
}

__device__ __forceinline__ void weno_x_body(const double *__restrict__ p, const double *__restrict__ dpi,
                                            double *__restrict__ pleft, double *__restrict__ pright, const Dims &d,
                                            int wide, unsigned bid) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int n = d.n1;
  // transverse rows: -1..n, or every ghost row for the finite-volume loops
  const int s2 = d.n2 > 1 ? (wide ? d.nb2 : 1) : 0, s3 = d.n3 > 1 ? (wide ? d.nb3 : 1) : 0;
  const int c1 = d.n2 + 2 * s2, c2 = d.n3 + 2 * s3;
  const int lr = n + 8;
  const int64_t nrows = (int64_t)NPR * c1 * c2;
  if (wid * XOUT >= nrows * lr) return;  // uniform per wavefront
  const int64_t v = wid * XOUT + lane;
  int64_t r = v / lr;
  const int s = (int)(v - r * lr) - 4;
  const bool live = r < nrows;
  if (!live) r = nrows - 1;  // lanes past the last row: in-range indices, no work
  const int pl = (int)(r % NPR);
  r /= NPR;
  const int j = -s2 + (int)(r % c1), k = -s3 + (int)(r / c1);
  const int64_t zl = d.idx(0, j, k);
  const double *yl = p + (int64_t)pl * d.spl + zl;
  double y[5] = {0., 0., 0., 0., 0.}, w[3] = {1., 1., 1.}, m0 = 0., lof0 = 0.;
  if (live) {
#pragma unroll
    for (int a = 0; a < 5; a++) y[a] = yl[s - 2 + a];
    stage_c<true>(y + 2, w, m0, lof0);
  }
  // stage D needs the weights of s-2 and s+2 (lanes -2 and +2)
  double w0m2 = __shfl_up(w[0], 2, 64);
  double w2p2 = __shfl_down(w[2], 2, 64);
  double lofd = stage_d(w[2], w0m2, w[0], w2p2, lof0);
  double lofm = __shfl_up(lofd, 1, 64);
  double lofp = __shfl_down(lofd, 1, 64);
  if (live && lane >= 3 && lane < 3 + XOUT && s >= -1 && s <= n) {
    const double dPi = dpi[zl + s];
    double ol, orr;
    stage_e<true>(y + 2, m0, w[0], w[1], w[2], lofm, lofd, lofp, dPi, ol, orr);
    int64_t z = (int64_t)pl * d.spl + zl + s;
    pleft[z] = ol;
    pright[z] = orr;
  }
}

__global__ void __launch_bounds__(256) k_weno_x(const double *__restrict__ p, const double *__restrict__ dpi,
                                                double *__restrict__ pleft, double *__restrict__ pright, Dims d,
                                                int wide) {
  weno_x_body(p, dpi, pleft, pright, d, wide, blockIdx.x);
}

// ---------------------------------------------------------------------------
// along j (DIR 2) or k (DIR 3): lane = (i, transverse index) flattened with i
// fastest; each lane marches over its segment of the line: at step t it
// evaluates C(t), D(t-2) and E(t-3), keeping the last values in registers.
// DIR 1 marches along i with lane = (j, k): no halo lanes, but every load of
// a wavefront touches 64 rows (the rows stay in cache over the next steps).
// ---------------------------------------------------------------------------
// A row address that is the same in every lane, pinned to scalar registers
// as a global-memory pointer: loads and stores take it as their scalar base
// and add the lane's 32-bit offset, where the compiler would otherwise keep a
// 64-bit per-lane pointer in two vector registers for every array it walks.
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gchar *uniform_row(const void *base, int64_t off) {
  const uint64_t v = (uint64_t)base + (uint64_t)off;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (gchar *)(((uint64_t)hi << 32) | lo);
}

// Block order: the 8 primitive blocks of one (lane block, segment) unit read
// the same dP/P indicators, so they are dispatched back to back to one XCD
// (work-groups go to the 8 XCDs round robin by linear index): block b runs
// on XCD b % 8, takes primitive (b / 8) % 8 of unit (b / 64) * 8 + b % 8.
// Addressing: the primitive plane and the line position t are uniform over
// the block, so every load and store is a uniform row address plus one
// 32-bit per-lane byte offset (the context keeps a plane under 4 GiB).
// PACK (the grouped launches of small grids): the lanes of a block run over
// (primitive, a, o) with a fastest, nlb blocks of 256 lanes over all
// NPR * ca * co of them, so that the two halo lanes of a short row do not
// take a wavefront of their own per primitive; the primitive plane then
// joins the 32-bit per-lane offset (8 planes under 4 GiB).
template <int DIR, bool PACK = false>
__device__ __forceinline__ void weno_m_body(const double *__restrict__ p, const double *__restrict__ dpi,
                                            double *__restrict__ pleft, double *__restrict__ pright, const Dims &d,
                                            int nseg, int seglen, int wide, int nlb, unsigned bid) {
  const int n = (DIR == 1) ? d.n1 : (DIR == 2) ? d.n2 : d.n3;
  // the lane's two transverse indices: a (fastest) and o, over -1..n or,
  // for the finite-volume loops, every ghost row
  const int h1 = d.n1 > 1 ? (wide ? d.nb1 : 1) : 0, h2 = d.n2 > 1 ? (wide ? d.nb2 : 1) : 0;
  const int h3 = d.n3 > 1 ? (wide ? d.nb3 : 1) : 0;
  const int sa = (DIR == 1) ? h2 : h1;
  const int so = (DIR == 2) ? h3 : (DIR == 3) ? h2 : h3;
  const int ca = ((DIR == 1) ? d.n2 : d.n1) + 2 * sa;
  const int co = ((DIR == 2) ? d.n3 : (DIR == 3) ? d.n2 : d.n3) + 2 * so;
  const unsigned slot = bid >> 3;
  const unsigned unit = PACK ? bid : (slot >> 3) * 8u + (bid & 7u);
  if (unit >= (unsigned)(nlb * nseg)) return;
  const int seg = (int)(unit / (unsigned)nlb);
  int64_t g = (int64_t)(unit % (unsigned)nlb) * blockDim.x + threadIdx.x;
  int pl = (int)(slot & 7u);
  if (PACK) {
    if (g >= (int64_t)NPR * ca * co) return;
    pl = (int)(g / ((int64_t)ca * co));
    g -= (int64_t)pl * ca * co;
  } else if (g >= (int64_t)ca * co) {
    return;
  }
  const int a_ = -sa + (int)(g % ca);
  const int o = -so + (int)(g / ca);
  const int e0 = -1 + seg * seglen;
  const int e1 = (e0 + seglen - 1 < n) ? e0 + seglen - 1 : n;
  const int64_t zl = (DIR == 1) ? d.idx(0, a_, o) : (DIR == 2) ? d.idx(a_, 0, o) : d.idx(a_, o, 0);
  const int64_t sd = (DIR == 1) ? 1 : (DIR == 2) ? d.sj : d.sk;
  // lowest point read along the line (y at e0-5); row offsets count from it
  // so that the uniform part of every address is non-negative
  const int lo = e0 - 5;
  // per-lane byte offsets: zd into the indicator, zb into the primitive's
  // plane (PACK: from the first plane)
  const uint32_t zd = (uint32_t)((zl + (int64_t)lo * sd) * (int64_t)sizeof(double));
  const uint32_t zb = PACK ? zd + (uint32_t)((int64_t)pl * d.spl * (int64_t)sizeof(double)) : zd;
  const int64_t rowb = sd * (int64_t)sizeof(double);
  const int64_t pu = PACK ? 0 : (int64_t)pl * d.spl;  // the plane when it is uniform
  const char *yb = (const char *)(p + pu);
  const char *db = (const char *)dpi;
  char *lb = (char *)(pleft + pu), *rb = (char *)(pright + pu);
#define WM_AT(base, t, off) (uniform_row((base), (int64_t)((t) - lo) * rowb) + (off))
#define WM_LD(base, t) (*(const __attribute__((address_space(1))) double *)WM_AT(base, t, zb))
#define WM_LDD(base, t) (*(const __attribute__((address_space(1))) double *)WM_AT(base, t, zd))
#define WM_ST(base, t) (*(__attribute__((address_space(1))) double *)WM_AT(base, t, zb))
  // windows: yw = y(t-5..t+2), W0 = w0(t-4..t), W1/W2/M0 = (t-3..t),
  // LD = stage-D fraction (t-4..t-2); the cusp fractions of t-2..t, each
  // 0 or 1, are bits 2..0 of lofb
  double yw[8], W0[5], W1[4], W2[4], M0[4], LD[3];
  unsigned lofb = 0u;
#pragma unroll
  for (int a = 0; a < 8; a++) yw[a] = 0.;
#pragma unroll
  for (int a = 0; a < 5; a++) W0[a] = 1.;
#pragma unroll
  for (int a = 0; a < 4; a++) { W1[a] = 1.; W2[a] = 1.; M0[a] = 0.; }
#pragma unroll
  for (int a = 0; a < 3; a++) LD[a] = 0.;
  // before step t = e0-3 the y window holds y(t-6..t+1); points below e0-5 are never read
#pragma unroll
  for (int a = 4; a < 8; a++) yw[a] = WM_LD(yb, e0 - 9 + a);
  // The loads of step t are issued one step ahead (step t-1), so that their
  // latency hides behind a step's arithmetic; indices are clamped to points
  // read anyway (the final step's prefetch and the indicator before e0 are
  // never used).  Every step issues the same memory operations, two loads
  // then two stores: the steps before the first output store their
  // (unused) values at e0, which the step of e0 overwrites, and two such
  // stores follow the first loads.  The wait for a step's loads then leaves
  // the previous step's stores in flight; with stores on some paths only it
  // would have to wait for them too.
  const int tlast = e1 + 3;
  double ny = WM_LD(yb, e0 - 1);
  double ndp = WM_LDD(db, e0);
  WM_ST(lb, e0) = 0.;
  WM_ST(rb, e0) = 0.;
  for (int t = e0 - 3; t <= tlast; t++) {
    const double cy = ny, cdp = ndp;
    {
      const int tn = (t + 1 < tlast) ? t + 1 : tlast;
      ny = WM_LD(yb, tn + 2);
      ndp = WM_LDD(db, (tn - 3 > e0) ? tn - 3 : e0);
    }
#pragma unroll
    for (int a = 0; a < 7; a++) yw[a] = yw[a + 1];
    yw[7] = cy;
    double w[3], m0, lof0;
    stage_c<false>(yw + 5, w, m0, lof0);
#pragma unroll
    for (int a = 0; a < 4; a++) W0[a] = W0[a + 1];
    W0[4] = w[0];
#pragma unroll
    for (int a = 0; a < 3; a++) { W1[a] = W1[a + 1]; W2[a] = W2[a + 1]; M0[a] = M0[a + 1]; }
    W1[3] = w[1];
    W2[3] = w[2];
    M0[3] = m0;
    lofb = ((lofb << 1) | (lof0 != 0.0 ? 1u : 0u)) & 7u;
    // D(t-2)
    LD[0] = LD[1];
    LD[1] = LD[2];
    LD[2] = stage_d(W2[1], W0[0], W0[2], W2[3], (lofb & 4u) ? 1.0 : 0.0);
    // E(t-3)
    double outl = 0., outr = 0.;
    if (t - 3 >= e0) stage_e<false>(yw + 2, M0[0], W0[1], W1[0], W2[0], LD[0], LD[1], LD[2], cdp, outl, outr);
    const int ts = (t - 3 > e0) ? t - 3 : e0;
    WM_ST(lb, ts) = outl;
    WM_ST(rb, ts) = outr;
  }
#undef WM_AT
#undef WM_LD
#undef WM_LDD
#undef WM_ST
}

template <int DIR>
__global__ void __launch_bounds__(256) k_weno_m(const double *__restrict__ p, const double *__restrict__ dpi,
                                                double *__restrict__ pleft, double *__restrict__ pright, Dims d,
                                                int nseg, int seglen, int wide, int nlb) {
  weno_m_body<DIR>(p, dpi, pleft, pright, d, nseg, seglen, wide, nlb, blockIdx.x);
}

// grouped launches (GridView, group_of) of the standard scheme's loops
#define WG_PRELUDE                                     \
  const int g_ = group_of(first, n, blockIdx.x);       \
  const GridView &v = views[g_];                       \
  const unsigned bid = blockIdx.x - first[g_], nblk = first[g_ + 1] - first[g_]; \
  (void)nblk;

__global__ void __launch_bounds__(256) kg_weno_prep(const GridView *__restrict__ views,
                                                    const unsigned *__restrict__ first, int n) {
  WG_PRELUDE
  weno_prep_body(v.pb, v.geo, v.aux, v.d, v.gam, 0, v.ndt_bits, bid);
}

__global__ void __launch_bounds__(256) kg_weno_dpi(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  WG_PRELUDE
  weno_dpi_body<0>(v.pb, v.aux, v.d, 0, bid, nblk);
}

__global__ void __launch_bounds__(256) kg_weno_x(const GridView *__restrict__ views,
                                                 const unsigned *__restrict__ first, int n) {
  WG_PRELUDE
  weno_x_body(v.pb, v.aux + 2 * v.d.spl, v.pl[0], v.pr[0], v.d, 0, bid);
}

template <int DIR>
__global__ void __launch_bounds__(256) kg_weno_m(const GridView *__restrict__ views,
                                                 const unsigned *__restrict__ first, int n) {
  WG_PRELUDE
  weno_m_body<DIR, true>(v.pb, v.aux + (1 + DIR) * v.d.spl, v.pl[DIR - 1], v.pr[DIR - 1], v.d, v.nseg[DIR],
                         v.seglen[DIR], 0, v.nlb[DIR], bid);
}
#undef WG_PRELUDE

void launch_weno_prep(const double *p, const double *geo, double *aux, const Dims &d, double gam, int wide,
                      bool all_dpi, unsigned long long *ndt_reset, hipStream_t s) {
  const int e1 = d.n1 > 1 ? (wide ? d.nb1 : 2) : 0, e2 = d.n2 > 1 ? (wide ? d.nb2 : 2) : 0;
  const int e3 = d.n3 > 1 ? (wide ? d.nb3 : 2) : 0;
  int64_t n = (int64_t)(d.n1 + 2 * e1) * (d.n2 + 2 * e2) * (d.n3 + 2 * e3);
  hipLaunchKernelGGL(k_weno_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, geo, aux, d, gam, wide,
                     ndt_reset);
  if (all_dpi) hipLaunchKernelGGL(k_weno_dpi<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, aux, d, wide);
}

static void weno_march_shape(int dir, const Dims &d, const Tune &tu, int wide, int &nseg, int &seglen, int &nlb) {
  const int s1 = d.n1 > 1 ? (wide ? d.nb1 : 1) : 0, s2 = d.n2 > 1 ? (wide ? d.nb2 : 1) : 0;
  const int s3 = d.n3 > 1 ? (wide ? d.nb3 : 1) : 0;
  const int n = (dir == 1) ? d.n1 : (dir == 2) ? d.n2 : d.n3;
  const int64_t lanes = (dir == 1) ? (int64_t)(d.n2 + 2 * s2) * (d.n3 + 2 * s3)
                                   : (int64_t)(d.n1 + 2 * s1) * ((dir == 2) ? d.n3 + 2 * s3 : d.n2 + 2 * s2);
  // split long lines so that the grid holds enough wavefronts
  const int want = (dir == 1) ? tu.weno_xmarch : (dir == 2) ? tu.weno_nseg : tu.weno_nseg3;
  nseg = want > 0 ? want : ((n + 2 >= 96) ? 2 : 1);
  if (nseg > n + 2) nseg = n + 2;
  seglen = (n + 2 + nseg - 1) / nseg;
  nseg = (n + 2 + seglen - 1) / seglen;  // every segment starts at or before point n
  nlb = (int)((lanes + 255) / 256);
}

void launch_weno(int dir, const double *p, double *aux, double *pleft, double *pright, const Dims &d,
                 const Tune &tu, int wide, bool with_dpi, hipStream_t s) {
  // transverse halo rows (the kernels use the same widths)
  const int s2 = d.n2 > 1 ? (wide ? d.nb2 : 1) : 0, s3 = d.n3 > 1 ? (wide ? d.nb3 : 1) : 0;
  // the dP/P indicators of this direction, over the box of k_weno_prep
  // (unless launch_weno_prep formed every direction's)
  if (with_dpi) {
    const int e1 = d.n1 > 1 ? (wide ? d.nb1 : 2) : 0, e2 = d.n2 > 1 ? (wide ? d.nb2 : 2) : 0;
    const int e3 = d.n3 > 1 ? (wide ? d.nb3 : 2) : 0;
    const int64_t nbox = (int64_t)(d.n1 + 2 * e1) * (d.n2 + 2 * e2) * (d.n3 + 2 * e3);
    const dim3 gb((unsigned)((nbox + 255) / 256));
    if (dir == 1) hipLaunchKernelGGL(k_weno_dpi<1>, gb, dim3(256), 0, s, p, aux, d, wide);
    else if (dir == 2) hipLaunchKernelGGL(k_weno_dpi<2>, gb, dim3(256), 0, s, p, aux, d, wide);
    else hipLaunchKernelGGL(k_weno_dpi<3>, gb, dim3(256), 0, s, p, aux, d, wide);
  }
  const double *dpi = aux + (int64_t)(1 + dir) * d.spl;
  if (dir == 1 && tu.weno_xmarch == 0) {
    const int64_t nw = weno_x_waves(d, s2, s3);
    int wpb = tu.weno_xwaves;
    hipLaunchKernelGGL(k_weno_x, dim3((unsigned)((nw + wpb - 1) / wpb)), dim3(64 * wpb), 0, s, p, dpi, pleft,
                       pright, d, wide);
  } else {
    int nseg, seglen, nlb;
    weno_march_shape(dir, d, tu, wide, nseg, seglen, nlb);
    // units of (lane block, segment), padded to whole groups of 8, times
    // the 8 primitives (NPR = 8)
    const int64_t units = ((int64_t)nlb * nseg + 7) / 8 * 8;
    dim3 grid((unsigned)(units * NPR));
    if (dir == 1)
      hipLaunchKernelGGL(k_weno_m<1>, grid, dim3(256), 0, s, p, dpi, pleft, pright, d, nseg, seglen, wide, nlb);
    else if (dir == 2)
      hipLaunchKernelGGL(k_weno_m<2>, grid, dim3(256), 0, s, p, dpi, pleft, pright, d, nseg, seglen, wide, nlb);
    else
      hipLaunchKernelGGL(k_weno_m<3>, grid, dim3(256), 0, s, p, dpi, pleft, pright, d, nseg, seglen, wide, nlb);
  }
}

void weno_group_shapes(GridView &v, const Tune &tu) {
  v.nseg[1] = v.seglen[1] = v.nlb[1] = 0;
  v.nseg[0] = v.seglen[0] = v.nlb[0] = 0;
  for (int dir = 2; dir <= 3; dir++) {
    weno_march_shape(dir, v.d, tu, 0, v.nseg[dir], v.seglen[dir], v.nlb[dir]);
    // packed lanes (weno_m_body PACK): blocks of 256 over all primitives
    const Dims &d = v.d;
    const int64_t lanes = (int64_t)(d.n1 + 2 * (d.n1 > 1)) * ((dir == 2) ? d.n3 + 2 * (d.n3 > 1) : d.n2 + 2 * (d.n2 > 1));
    v.nlb[dir] = (int)((NPR * lanes + 255) / 256);
  }
}

// grouped: the blocks of the launchers above with the default shapes (four
// wavefronts per block along i; the view's segments along j and k), a
// multiple of 8 per grid for the XCD-ordered classes
unsigned group_blocks_weno(int k, const GridView &v) {
  const Dims &d = v.d;
  const int e1 = d.n1 > 1 ? 2 : 0, e2 = d.n2 > 1 ? 2 : 0, e3 = d.n3 > 1 ? 2 : 0;
  const int64_t nbox = (int64_t)(d.n1 + 2 * e1) * (d.n2 + 2 * e2) * (d.n3 + 2 * e3);
  switch (k) {
    case GK_PREP: return (unsigned)((nbox + 255) / 256);
    case GK_DPI: return (unsigned)((nbox + 255) / 256 + 7) / 8 * 8;
    case GK_WENO1: {
      if (d.n1 <= 1) return 0u;
      const int64_t nw = weno_x_waves(d, d.n2 > 1 ? 1 : 0, d.n3 > 1 ? 1 : 0);
      return (unsigned)((nw + 3) / 4);
    }
    case GK_WENO2: case GK_WENO3: {
      // packed lanes: one block per (lane block, segment) unit
      const int dir = (k == GK_WENO2) ? 2 : 3;
      if (((dir == 2) ? d.n2 : d.n3) <= 1) return 0u;
      return (unsigned)v.nlb[dir] * (unsigned)v.nseg[dir];
    }
    default: return 0u;
  }
}

void launch_group_weno(int k, const GridView *views, const unsigned *first, int n, unsigned total, hipStream_t s) {
  const dim3 g(total), b(256);
  if (k == GK_PREP) hipLaunchKernelGGL(kg_weno_prep, g, b, 0, s, views, first, n);
  if (k == GK_DPI) hipLaunchKernelGGL(kg_weno_dpi, g, b, 0, s, views, first, n);
  if (k == GK_WENO1) hipLaunchKernelGGL(kg_weno_x, g, b, 0, s, views, first, n);
  if (k == GK_WENO2) hipLaunchKernelGGL(kg_weno_m<2>, g, b, 0, s, views, first, n);
  if (k == GK_WENO3) hipLaunchKernelGGL(kg_weno_m<3>, g, b, 0, s, views, first, n);
}

}  // namespace hg
