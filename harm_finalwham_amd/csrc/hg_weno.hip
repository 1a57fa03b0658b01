// WENO5BND interface reconstruction on gfx950 (interpline.c:517
// slope_lim_linetype_c2e with lim = WENO5BND, PRIMTOINTERP, DEFAULT reduction).
//
// One workgroup owns TL pencils along DIR.  The pencil algorithm has data
// dependencies along the line (weights at s+-2, lower-order fractions at
// s+-1), so every stage is evaluated once per point into LDS and the next
// stage reads its neighbours there:
//   A  dP/P jump indicator (interpline.c:2148), shared by the 8 primitives
//   per primitive:
//   B  load the line
//   C  smooth-monotonicity indicator + order-3 weights (smonointerp.c:15,
//      reconstructeno.c:1406, :702)
//   D  weight ratios -> lower-order fraction (reconstructeno.c:1962)
//   E  post/pre-shock reduction, optimised weights, order-3/order-2 blend,
//      stencil evaluation and the ENO5 blend (reconstructeno.c:650, :1617,
//      :2087, :730, :470-490) -> pleft/pright
// Only the pencils and points the fluxes read are computed: transverse
// indices -1..n and points -1..n along the line (weights at -4..n+3).
// All arithmetic is FP64 in the reference's operation order.
#include <hip/hip_runtime.h>

#include <cfloat>

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
#pragma unroll
  for (int s = 0; s < 3; s++) si[s] /= (ss + DBL_MIN);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double t = si[s] + WENO_EPS / 1.;
    t *= t;
    w[s] = 1. / t;
  }
  double sw = sum3(w[0], w[1], w[2]);
#pragma unroll
  for (int s = 0; s < 3; s++) w[s] /= (sw + 0);
  return lof;
}

// order-2 unoptimised + optimised weights (reconstructeno.c:1406 and :1617 with order 2)
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
  si[0] /= (ss + DBL_MIN);
  si[1] /= (ss + DBL_MIN);
#pragma unroll
  for (int s = 0; s < 2; s++) {
    double t = si[s] + WENO_EPS / 1.;
    t *= t;
    w[s] = 1. / t;
  }
  double sw = sum2(w[0], w[1]);
  w[0] /= (sw + 0);
  w[1] /= (sw + 0);
  wl[0] = w[0] * (1. / 4.);
  wl[1] = w[1] * (3. / 4);
  wr[0] = w[0] * (3. / 4.);
  wr[1] = w[1] * (1. / 4.);
  double sl = sum2(wl[0], wl[1]);
  wl[0] /= (sl + DBL_MIN);
  wl[1] /= (sl + DBL_MIN);
  double sr = sum2(wr[0], wr[1]);
  wr[0] /= (sr + DBL_MIN);
  wr[1] /= (sr + DBL_MIN);
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
  double ld1 = (((-3 * yin[-2] + yin[2]) + (-10 * yin[-1] - 6 * yin[1])) + 18 * yin[0]) / 12.;
  double ld2 = (((11 * yin[-2] - yin[2]) + (-20 * yin[-1] + 4 * yin[1])) + 6 * yin[0]) / 12.;
  double ld3 = (((-3 * yin[-2] - yin[2]) + (10 * yin[-1] + 6 * yin[1])) - 12 * yin[0]) / 2.;
  double rd1 = (((-yin[-2] + 3 * yin[2]) + (6 * yin[-1] + 10 * yin[1])) - 18 * yin[0]) / 12.;
  double rd2 = (((-yin[-2] + 11 * yin[2]) + (4 * yin[-1] - 20 * yin[1])) + 6 * yin[0]) / 12.;
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

// LDS arrays per line, each of length LS (points -nb .. n+nb-1 at offset nb)
enum { A_Y, A_W0, A_W1, A_W2, A_M0, A_M1, A_LOF0, A_LOFD, A_PAR, A_UPAR, A_RHO, A_DP, A_NUM };

template <int DIR, int TL>
__global__ void __launch_bounds__(256) k_weno(const double *__restrict__ p, const double *__restrict__ geo,
                                              double *__restrict__ pleft, double *__restrict__ pright, Dims d,
                                              double gam, int LS) {
  extern __shared__ double sm[];
  const int n = (DIR == 1) ? d.n1 : (DIR == 2) ? d.n2 : d.n3;
  const int nb = (DIR == 1) ? d.nb1 : (DIR == 2) ? d.nb2 : d.nb3;
  // transverse extents: the flux faces read one layer beyond the active zones
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  int c1, c2, lo1, lo2;
  if (DIR == 1) { c1 = d.n2 + 2 * s2; lo1 = -s2; c2 = d.n3 + 2 * s3; lo2 = -s3; }
  else if (DIR == 2) { c1 = d.n1 + 2 * s1; lo1 = -s1; c2 = d.n3 + 2 * s3; lo2 = -s3; }
  else { c1 = d.n1 + 2 * s1; lo1 = -s1; c2 = d.n2 + 2 * s2; lo2 = -s2; }
  const int nlines = c1 * c2;
  const int tid = threadIdx.x;
  const int l = tid % TL, sg = tid / TL, nsg = 256 / TL;
  const int line = blockIdx.x * TL + l;
  const bool active = line < nlines;
  int t1 = lo1 + (active ? line % c1 : 0), t2 = lo2 + (active ? line / c1 : 0);
  int i0, j0, k0;
  if (DIR == 1) { i0 = 0; j0 = t1; k0 = t2; }
  else if (DIR == 2) { i0 = t1; j0 = 0; k0 = t2; }
  else { i0 = t1; j0 = t2; k0 = 0; }
  const int64_t z0 = d.idx(i0, j0, k0);
  const int64_t sd = (DIR == 1) ? 1 : (DIR == 2) ? d.sj : d.sk;
  double *A = sm + (int64_t)l * A_NUM * LS;
#define ARR(a, s) A[(a)*LS + (s) + nb]

  // ---- stage A: dP/P indicator (interpline.c:2148 compute_df_line_new) ----
  if (active) {
    for (int s = -2 + sg; s <= n + 1; s += nsg) {
      double pr[NPR];
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) pr[pl] = p[z0 + s * sd + pl * d.spl];
      int gi = (DIR == 1) ? s : i0, gj = (DIR == 2) ? s : j0;
      Geom G;
      {
        const double *b = geo + (int64_t)GL_CENT * GCOMP * d.gplane + d.gidx(gi, gj);
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
      ARR(A_PAR, s) = p1 * (q.ucon[0]);
      ARR(A_UPAR, s) = gam * pr[UU] * (q.ucon[0] * q.ucov[0]) + (gam - 1.0) * pr[UU];
      ARR(A_RHO, s) = pr[RHO];
    }
  }
  __syncthreads();
  if (active) {
    for (int s = -1 + sg; s <= n; s += nsg) {
      double rho = ARR(A_RHO, s);
      ARR(A_DP, s) = (fabs(rho * (ARR(A_PAR, s + 1) - ARR(A_PAR, s - 1))) +
                      fabs(ARR(A_UPAR, s + 1) - ARR(A_UPAR, s - 1))) /
                     (fabs(rho * ARR(A_PAR, s)) + fabs(ARR(A_UPAR, s)) + DBL_MIN);
    }
  }

  for (int pl = 0; pl < NPR; pl++) {
    __syncthreads();
    // ---- stage B: the line ----
    if (active)
      for (int s = -nb + sg; s <= n + nb - 1; s += nsg) ARR(A_Y, s) = p[z0 + s * sd + pl * d.spl];
    __syncthreads();
    // ---- stage C: monotonicity indicator and order-3 weights at -4..n+3 ----
    if (active) {
      for (int s = -4 + sg; s <= n + 3; s += nsg) {
        const double *y = &ARR(A_Y, s);
        double m0, m1;
        if (cusp(y)) { m0 = 0; m1 = 1; }
        else { m0 = mono_eno5(y); m1 = m0; }
        double w[3];
        double lof0 = weights3(m0, m1, y, w);
        if (m0 > 0.0 && m0 < 1.0) {  // reconstructeno.c:702 mono_adjust_unoptimized_weights
          double simpleweight = 1.0 / 3.0;
#pragma unroll
          for (int c = 0; c < 3; c++) w[c] = (1.0 - m0) * w[c] + m0 * simpleweight;
        }
        ARR(A_W0, s) = w[0];
        ARR(A_W1, s) = w[1];
        ARR(A_W2, s) = w[2];
        ARR(A_M0, s) = m0;
        ARR(A_M1, s) = m1;
        ARR(A_LOF0, s) = lof0;
      }
    }
    __syncthreads();
    // ---- stage D: weight ratios at -2..n+1 (reconstructeno.c:1962) ----
    if (active) {
      for (int s = -2 + sg; s <= n + 1; s += nsg) {
        double wl = (ARR(A_W2, s) + WENO_EPS) / (ARR(A_W0, s - 2) + WENO_EPS);
        double wr = (ARR(A_W0, s) + WENO_EPS) / (ARR(A_W2, s + 2) + WENO_EPS);
        double lof = (wmin(wl, wr) - MIN_TR) / (MAX_TR - MIN_TR);
        lof = wmax(0., wmin(lof, 1.0));
        ARR(A_LOFD, s) = wmax(lof, ARR(A_LOF0, s));
      }
    }
    __syncthreads();
    // ---- stage E: reduction, optimised weights, outputs at -1..n ----
    if (active) {
      for (int s = -1 + sg; s <= n; s += nsg) {
        const double *y = &ARR(A_Y, s);
        double m0 = ARR(A_M0, s);
        double lof = ARR(A_LOFD, s), dPi = ARR(A_DP, s);
        double shock = wmin((fabs(dPi) - 0.5 * REDUCE_DP) / (0.5 * REDUCE_DP), 1.);
        lof = wmax(lof, shock * ARR(A_LOFD, s + 1));
        lof = wmax(lof, shock * ARR(A_LOFD, s - 1));
        double outv[2];
        if (m0 == 1) {
          outv[0] = eno5(false, y);
          outv[1] = eno5(true, y);
        } else {
          // compute_optimized_cf_stencil_weights (order 3)
          double w[3] = {ARR(A_W0, s), ARR(A_W1, s), ARR(A_W2, s)};
          double wl[3] = {w[0] * (1. / 16.), w[1] * (5. / 8.), w[2] * (5. / 16.)};
          double wr[3] = {w[0] * (5. / 16.), w[1] * (5. / 8.), w[2] * (1. / 16.)};
          double sl = sum3(wl[0], wl[1], wl[2]), sr = sum3(wr[0], wr[1], wr[2]);
#pragma unroll
          for (int c = 0; c < 3; c++) { wl[c] /= (sl + DBL_MIN); wr[c] /= (sr + DBL_MIN); }
          double l2[2] = {0., 0.}, r2[2] = {0., 0.};
          bool two = (0.0 < lof);
          if (two) {
#pragma unroll
            for (int c = 0; c < 3; c++) { wl[c] *= (1. - lof); wr[c] *= (1. - lof); }
            weights2(m0, y, l2, r2);
            l2[0] *= lof; l2[1] *= lof; r2[0] *= lof; r2[1] *= lof;
          }
#pragma unroll
          for (int side = 0; side < 2; side++) {
            double prev = 0.0;
            if (m0 != 0.0) prev = eno5(side == 1, y);
            double ov = 0.0;
            ov += cvt3(side == 1, side ? wr : wl, y);
            if (two) ov += cvt2(side == 1, side ? r2 : l2, y);
            outv[side] = (1.0 - m0) * ov + m0 * prev;
          }
        }
        int64_t z = z0 + s * sd + pl * d.spl;
        pleft[z] = outv[0];
        pright[z] = outv[1];
      }
    }
  }
#undef ARR
}

void launch_weno(int dir, const double *p, const double *geo, double *pleft, double *pright, const Dims &d,
                 double gam, hipStream_t s) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  if (dir == 1) {
    constexpr int TL = 1;
    int LS = d.m1;
    int64_t nl = (int64_t)(d.n2 + 2 * s2) * (d.n3 + 2 * s3);
    size_t shm = (size_t)TL * A_NUM * LS * sizeof(double);
    hipLaunchKernelGGL((k_weno<1, TL>), dim3((unsigned)((nl + TL - 1) / TL)), dim3(256), shm, s, p, geo, pleft,
                       pright, d, gam, LS);
  } else if (dir == 2) {
    constexpr int TL = 4;
    int LS = d.m2;
    int64_t nl = (int64_t)(d.n1 + 2 * s1) * (d.n3 + 2 * s3);
    size_t shm = (size_t)TL * A_NUM * LS * sizeof(double);
    hipLaunchKernelGGL((k_weno<2, TL>), dim3((unsigned)((nl + TL - 1) / TL)), dim3(256), shm, s, p, geo, pleft,
                       pright, d, gam, LS);
  } else {
    constexpr int TL = 8;
    int LS = d.m3;
    int64_t nl = (int64_t)(d.n1 + 2 * s1) * (d.n2 + 2 * s2);
    size_t shm = (size_t)TL * A_NUM * LS * sizeof(double);
    hipLaunchKernelGGL((k_weno<3, TL>), dim3((unsigned)((nl + TL - 1) / TL)), dim3(256), shm, s, p, geo, pleft,
                       pright, d, gam, LS);
  }
}

}  // namespace hg
