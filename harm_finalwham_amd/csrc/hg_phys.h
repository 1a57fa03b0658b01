// Device point physics of the GRMHD hot path (FP64, vector ALU).
//
// Each function evaluates the same expressions, in the same order, as the
// reference functions it cites, so that results agree bit for bit with the
// reference's FP64 build when compiled with -ffp-contract=off (no FMA
// contraction; FP64 division and sqrt are correctly rounded on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hg_grid.h"

// point physics is shared by the device kernels and the host-side set-up
// (initial data), which must produce the same bits
#define HG_HD __host__ __device__ __forceinline__

namespace hg {

constexpr double DBL_EPS = 2.2204460492503131e-16;

// failure codes, global.h:782-799
constexpr int FAIL_CONVW = 2, FAIL_CONVUTSQ = 3, FAIL_CONVGUESSUTSQ = 4, FAIL_RHONEG = 9,
              FAIL_UNEG = 10, FAIL_RHOUNEG = 11;

struct Geom {
  double gcov[4][4];
  double gcon[4][4];
  double gpert0;  // gcovpert[TT]
  double g;       // sqrt(-g), the EOM factor of every variable (WHICHEOM WITHGDET)
};

struct State {
  double ucon[4], ucov[4], bcon[4], bcov[4];
};

// phys.c:1797 max / phys.c:1804 min
HG_HD double rmax(double a, double b) { return (a > b) ? a : b; }
HG_HD double rmin(double a, double b) { return (a > b) ? b : a; }

// phys.c:766 lower_vec
HG_HD void lower_vec(const double *v, const Geom &G, double *w) {
  const double(*g)[4] = G.gcov;
  w[0] = g[0][0] * v[0] + g[0][1] * v[1] + g[0][2] * v[2] + g[0][3] * v[3];
  w[1] = g[0][1] * v[0] + g[1][1] * v[1] + g[1][2] * v[2] + g[1][3] * v[3];
  w[2] = g[0][2] * v[0] + g[1][2] * v[1] + g[2][2] * v[2] + g[2][3] * v[3];
  w[3] = g[0][3] * v[0] + g[1][3] * v[1] + g[2][3] * v[2] + g[3][3] * v[3];
}

// phys.c:825 raise_vec
HG_HD void raise_vec(const double *v, const Geom &G, double *w) {
  const double(*g)[4] = G.gcon;
  w[0] = g[0][0] * v[0] + g[0][1] * v[1] + g[0][2] * v[2] + g[0][3] * v[3];
  w[1] = g[0][1] * v[0] + g[1][1] * v[1] + g[1][2] * v[2] + g[1][3] * v[3];
  w[2] = g[0][2] * v[0] + g[1][2] * v[1] + g[2][2] * v[2] + g[2][3] * v[3];
  w[3] = g[0][3] * v[0] + g[1][3] * v[1] + g[2][3] * v[2] + g[3][3] * v[3];
}

// phys.c:1069 qsq_calc, phys.c:1040 gamma_calc_fromuconrel with JONCHECKS
HG_HD bool gamma_calc(const double *pr, const Geom &G, double &gamma) {
  const double(*g)[4] = G.gcov;
  double u1 = pr[U1], u2 = pr[U2], u3 = pr[U3];
  double qsq = +g[1][1] * u1 * u1 + g[2][2] * u2 * u2 + g[3][3] * u3 * u3 +
               2. * (g[1][2] * u1 * u2 + g[1][3] * u1 * u3 + g[2][3] * u2 * u3);
  if (qsq < 0.0) {
    if (qsq > -DBL_EPS * 100) qsq = 0.0;
    else return true;
  }
  gamma = sqrt(1. + qsq);
  return false;
}

// phys.c:1006 ucon_calc_rel4vel
HG_HD bool ucon_calc(const double *pr, const Geom &G, double *ucon) {
  double alpha = 1. / sqrt(-G.gcon[0][0]);
  double beta1 = G.gcon[0][1] * alpha * alpha;
  double beta2 = G.gcon[0][2] * alpha * alpha;
  double beta3 = G.gcon[0][3] * alpha * alpha;
  double gamma;
  if (gamma_calc(pr, G, gamma)) return true;
  ucon[0] = gamma / alpha;
  ucon[1] = pr[U1] - ucon[0] * beta1;
  ucon[2] = pr[U2] - ucon[0] * beta2;
  ucon[3] = pr[U3] - ucon[0] * beta3;
  return false;
}

// phys.c:849 get_state (with phys.c:309 bcon_calc)
HG_HD bool get_state(const double *pr, const Geom &G, State &q) {
  if (ucon_calc(pr, G, q.ucon)) return true;
  lower_vec(q.ucon, G, q.ucov);
  q.bcon[0] = pr[B1] * q.ucov[1] + pr[B2] * q.ucov[2] + pr[B3] * q.ucov[3];
#pragma unroll
  for (int j = 1; j <= 3; j++) q.bcon[j] = (pr[B1 - 1 + j] + q.bcon[0] * q.ucon[j]) / q.ucon[0];
  lower_vec(q.bcon, G, q.bcov);
  return false;
}

// phys.c:541 compute_1plusud0
HG_HD double one_plus_ud0(const State &q, const Geom &G) {
  double vcon[4];
#pragma unroll
  for (int j = 1; j < 4; j++) vcon[j] = q.ucon[j] / q.ucon[0];
  double plus1gv00 = G.gpert0;
  double vsq = G.gpert0;
#pragma unroll
  for (int j = 1; j < 4; j++) vsq += 2.0 * G.gcov[0][j] * vcon[j];
#pragma unroll
  for (int j = 1; j < 4; j++)
#pragma unroll
    for (int k = 1; k < 4; k++) vsq += G.gcov[j][k] * vcon[j] * vcon[k];
  double gvtt = G.gcov[0][0];
  double alpha = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) alpha += G.gcov[j][0] * q.ucon[j];
  double uu0 = q.ucon[0];
  return alpha + ((1.0 - gvtt) * plus1gv00 - uu0 * uu0 * vsq * gvtt * gvtt) / (1.0 - gvtt * uu0);
}

// Electromagnetic T^dir_nu: phys.c:178 Mcon_calc, phys.ffde.c:1707 lower_A,
// phys.ffde.c:1483 ffdestresstensor_dir.  The covariant Maxwell tensor does
// not depend on dir, so callers that need several rows compute it once.
struct Maxwell {
  double Mcon[4][4], Mcov[4][4], Msq;
};

HG_HD void maxwell(const double *pr, const State &q, const Geom &G, Maxwell &M) {
  const double(*g)[4] = G.gcov;
#pragma unroll
  for (int j = 0; j < 4; j++) M.Mcon[j][j] = 0.0;
#pragma unroll
  for (int k = 1; k < 4; k++) {
    M.Mcon[k][0] = pr[B1 + k - 1];
    M.Mcon[0][k] = -M.Mcon[k][0];
  }
  // The space-space entries (B^j u^k - B^k u^j) / u^t of phys.c:178, three
  // divisions instead of nine, same bits for finite products B^j u^k (every
  // state a step holds): a diagonal numerator is a product less itself, +0,
  // and +0 / u^t = +0 (u^t > 0).  Entry [k][j] subtracts the same two
  // products the other way round, the exact negative of [j][k]'s numerator
  // unless the two products are equal: then both numerators are the zero of
  // their own subtraction, whose quotient by u^t > 0 is that zero itself.
#pragma unroll
  for (int j = 1; j < 4; j++)
#pragma unroll
    for (int k = j + 1; k < 4; k++) {
      const double a = pr[B1 + j - 1] * q.ucon[k], b = pr[B1 + k - 1] * q.ucon[j];
      const double njk = a - b;
      M.Mcon[j][k] = njk / q.ucon[0];
      M.Mcon[k][j] = (njk == 0.0) ? (b - a) : -M.Mcon[j][k];
    }
  double a01 = 0., a02 = 0., a03 = 0., a12 = 0., a23 = 0., a31 = 0.;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a01 += (g[0][j]) * (g[1][k]) * M.Mcon[j][k];
      a02 += (g[0][j]) * (g[2][k]) * M.Mcon[j][k];
      a03 += (g[0][j]) * (g[3][k]) * M.Mcon[j][k];
      a12 += (g[1][j]) * (g[2][k]) * M.Mcon[j][k];
      a23 += (g[2][j]) * (g[3][k]) * M.Mcon[j][k];
      a31 += (g[3][j]) * (g[1][k]) * M.Mcon[j][k];
    }
  M.Mcov[0][0] = 0.; M.Mcov[1][1] = 0.; M.Mcov[2][2] = 0.; M.Mcov[3][3] = 0.;
  M.Mcov[0][1] = a01; M.Mcov[0][2] = a02; M.Mcov[0][3] = a03;
  M.Mcov[1][2] = a12; M.Mcov[2][3] = a23; M.Mcov[3][1] = a31;
  M.Mcov[1][0] = -a01; M.Mcov[2][0] = -a02; M.Mcov[3][0] = -a03;
  M.Mcov[2][1] = -a12; M.Mcov[3][2] = -a23; M.Mcov[1][3] = -a31;
  double Msq = 0.;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int k = 0; k < 4; k++) Msq += M.Mcov[j][k] * M.Mcon[j][k];
  M.Msq = Msq;
}

// phys.c:442 mhd_calc_norestmass: row dir of T^mu_nu with rest mass removed from T^t_t
HG_HD void mhd_row(const double *pr, int dir, const State &q, const Maxwell &M,
                                        double plus1ud0, double gam, double *mhd) {
  double rho = pr[RHO], u = pr[UU];
  double P = (gam - 1.) * u;  // eos.c:58
  double w = P + rho + u;
  double eta = w, ptot = P;
  mhd[0] = (P + u) * q.ucon[dir] * q.ucov[0] + rho * q.ucon[dir] * plus1ud0 + ptot * ((dir == 0) ? 1. : 0.);
#pragma unroll
  for (int j = 1; j < 4; j++) mhd[j] = eta * q.ucon[dir] * q.ucov[j] + ptot * ((dir == j) ? 1. : 0.);
  double T[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    T[k] = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) T[k] += M.Mcon[dir][i] * M.Mcov[k][i];
  }
  T[dir] -= 0.25 * M.Msq;
#pragma unroll
  for (int j = 0; j < 4; j++) mhd[j] += T[j];
}

// phys.c:16 primtoflux, returntype UEVOLVE (x gdet)
HG_HD void primtoflux(const double *pr, const State &q, const Maxwell &M,
                                          double plus1ud0, int dir, const Geom &G, double gam,
                                          double *flux) {
  double fl[NPR];
  fl[RHO] = pr[RHO] * q.ucon[dir];
  mhd_row(pr, dir, q, M, plus1ud0, gam, &fl[UU]);
  if (dir > 0) {
    // (B^k u^dir - B^dir u^k) / u^t is the Maxwell entry [k][dir] of the
    // same state, the same operations (maxwell above)
    fl[B1] = M.Mcon[1][dir];
    fl[B2] = M.Mcon[2][dir];
    fl[B3] = M.Mcon[3][dir];
  } else {
    fl[B1] = pr[B1];
    fl[B2] = pr[B2];
    fl[B3] = pr[B3];
  }
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) flux[pl] = fl[pl] * G.g;
}

// conserved variables U (phys.c:225 primtoU)
HG_HD bool primtoU(const double *pr, const Geom &G, double gam, double *U) {
  State q;
  if (get_state(pr, G, q)) return true;
  Maxwell M;
  maxwell(pr, q, G, M);
  double p1 = one_plus_ud0(q, G);
  primtoflux(pr, q, M, p1, 0, G, gam, U);
  return false;
}

// vchar.c:213 simplefast
HG_HD bool simplefast(int dir, const Geom &G, const State &q, double cms2,
                                          double &vmin, double &vmax) {
  double Acov[4] = {0., 0., 0., 0.}, Bcov[4] = {1., 0., 0., 0.}, Acon[4], Bcon[4];
  Acov[dir] = 1.;
  raise_vec(Acov, G, Acon);
  raise_vec(Bcov, G, Bcon);
  double Asq = Acon[0] * Acov[0] + Acon[1] * Acov[1] + Acon[2] * Acov[2] + Acon[3] * Acov[3];
  double Bsq = Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3];
  double Au = Acov[0] * q.ucon[0] + Acov[1] * q.ucon[1] + Acov[2] * q.ucon[2] + Acov[3] * q.ucon[3];
  double Bu = Bcov[0] * q.ucon[0] + Bcov[1] * q.ucon[1] + Bcov[2] * q.ucon[2] + Bcov[3] * q.ucon[3];
  double AB = Acon[0] * Bcov[0] + Acon[1] * Bcov[1] + Acon[2] * Bcov[2] + Acon[3] * Bcov[3];
  double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  double A = Bu2 * (1.0 - cms2) - Bsq * cms2;
  double B = 2. * (AuBu * (1.0 - cms2) - AB * cms2);
  double discr = 4.0 * cms2 *
                 ((AB * AB - Asq * Bsq) * cms2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (cms2 - 1.0));
  if ((discr < 0.0) && (discr > -1E-10)) discr = 0.0;
  else if (discr < 0.0) return true;
  discr = sqrt(discr);
  double vm = -(-B + discr) / (2. * A);
  double vp = -(-B - discr) / (2. * A);
  if (!isfinite(vm) || !isfinite(vp)) return true;
  if (vp < vm) { vmax = vm; vmin = vp; }
  else { vmax = vp; vmin = vm; }
  return false;
}

// vchar.c:37 vchar
HG_HD bool vchar(const double *pr, const State &q, int dir, const Geom &G,
                                     double gam, double &vmax, double &vmin) {
  double bsq = q.bcon[0] * q.bcov[0] + q.bcon[1] * q.bcov[1] + q.bcon[2] * q.bcov[2] + q.bcon[3] * q.bcov[3];
  double rho = pr[RHO], u = pr[UU];
  if (u < 0) u = 0;
  if (rho < 0) rho = 0;
  double P = (gam - 1.) * u;
  double WW = rho + u + P, cs2;
  if (WW > 0.0) {
    double pressure = (gam - 1.) * u;  // eos.c:87 cs2_compute_grmhd
    double h = rho + u + pressure;
    cs2 = gam * pressure / h;
  } else {
    WW = cs2 = 0.0;
  }
  if (cs2 >= 1.0) cs2 = 1.0 - 1e-9;
  else if (cs2 <= 0.0) cs2 = 0.0;
  double EF = bsq + WW, va2;
  if (EF > 0.0) va2 = bsq / EF;
  else EF = va2 = 0.0;
  double cms2 = cs2 + va2 - cs2 * va2;
  if (cms2 < 0.0) {
    if (cms2 / (cs2 + va2) > -DBL_EPS * 100.0) cms2 = 0.0;
    else return true;
  }
  if (cms2 > 1.) {
    if (cms2 < 1.0 + DBL_EPS * 100.0) cms2 = 1.0;
    else if (rho < 0.0) cms2 = 1.0;
    else return true;
  }
  return simplefast(dir, G, q, cms2, vmin, vmax);
}

// fixup.c:1859 limit_gamma
HG_HD int limit_gamma(double gammamax, double *pr, const Geom &G) {
  if (gammamax <= 1.0) return 0;
  double gamma;
  if (gamma_calc(pr, G, gamma)) return 1;
  if ((gamma > gammamax && (gamma != 1.0))) {
    double pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
    if (pref < 0.0) return 1;
    double f = sqrt(pref);
    pr[U1] *= f;
    pr[U2] *= f;
    pr[U3] *= f;
    return -1;
  }
  return 0;
}

// fixup.c:2271 inflow_check_rel4vel, radial direction at zone centres;
// side < 0: inner boundary (u^r > 0 is fixed), side > 0: outer (u^r < 0)
HG_HD int inflow_check_r(double *pr, const Geom &G, int side, double gammamax) {
  double ucon[4];
  if (ucon_calc(pr, G, ucon)) return 1;
  if (!((side < 0 && ucon[1] > 0.) || (side > 0 && ucon[1] < 0.))) return 0;
  double gamma;
  if (gamma_calc(pr, G, gamma)) return 1;
  pr[U1] /= gamma;
  pr[U2] /= gamma;
  pr[U3] /= gamma;
  double alpha = 1. / sqrt(-G.gcon[0][0]);
  double betacon = G.gcon[0][1] * alpha * alpha;
  pr[U1] = betacon / alpha;
  double vsq = 0.;
#pragma unroll
  for (int j = 1; j < 4; j++)
#pragma unroll
    for (int k = 1; k < 4; k++) vsq += G.gcov[j][k] * pr[U1 + j - 1] * pr[U1 + k - 1];
  if (vsq < 0.0) {
    if (vsq > -DBL_EPS * 100.0) vsq = 0.0;
    else return 1;
  }
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gammamax * gammamax);
  gamma = 1. / sqrt(1. - vsq);
  pr[U1] *= gamma;
  pr[U2] *= gamma;
  pr[U3] *= gamma;
  return -1;
}

}  // namespace hg
