// Problem set-up: modified Kerr-Schild geometry tables and the magnetised
// Fishbone-Moncrief torus.
//
// The metric is phi independent, so the geometry is a 2-D table built once on
// the host.  The torus initial data is also built on the host (it draws one
// sequential random number per torus zone) except for the ghost fills, which
// use the device kernels of the step.  Every expression keeps the operation
// order of the reference function it cites: with -ffp-contract=off and the
// same libm, the tables and the initial state are the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "harmgpu.h"
#include "hg_host.h"
#include "hg_kernels.h"
#include "hg_phys.h"

using namespace hg;

namespace {

constexpr double SINGSMALL = 1E-14;  // init.fishmon.h:193, polar-axis guard of theta
constexpr double CONNDELTA = 1E-5;   // metric.c:148, step of the numerical connection
constexpr int NLOC = 7;              // host table slots: FACE1 FACE2 FACE3 CORN1 CORN2 CORN3 CENT
constexpr int SLOT_CENT = 6;

struct Mks {
  double a, R0, hslope;
  double startx[4], dx[4];
};

// sin and cos of one angle.  Where the reference evaluates both in one
// function (the metric components, the torus solution) its gcc build computes
// them with a single libm sincos() call, whose results can differ in the last
// bit from separate sin() and cos() calls; the set-up does the same there.
// Lone sin or cos (the coordinate map and its Jacobian) stay separate calls.
inline void sincos_ref(double x, double &s, double &c) { sincos(x, &s, &c); }

// coord.c:1142 set_points (LOGRSINTH)
Mks make_mks(double a, double Rin, double Rout, double R0, double hslope, int n1, int n2, int n3) {
  Mks m;
  m.a = a;
  m.R0 = R0;
  m.hslope = hslope;
  m.startx[0] = 0.;
  m.startx[1] = log(Rin - R0);
  m.startx[2] = 0.;
  m.startx[3] = 0.;
  m.dx[0] = 0.;
  m.dx[1] = log((Rout - R0) / (Rin - R0)) / n1;
  m.dx[2] = 1. / n2;
  m.dx[3] = 1.0 / n3;
  return m;
}

// coord.c:1195 coord: zone-centre offsets of each table slot
void coord(const Mks &m, int i, int j, int k, int slot, double X[4]) {
  static const double off[NLOC][3] = {{0., .5, .5}, {.5, 0., .5}, {.5, .5, 0.}, {.5, 0., 0.},
                                      {0., .5, 0.}, {0., 0., .5}, {.5, .5, .5}};
  X[0] = 0.;
  X[1] = m.startx[1] + (i + off[slot][0]) * m.dx[1];
  X[2] = m.startx[2] + (j + off[slot][1]) * m.dx[2];
  X[3] = m.startx[3] + (k + off[slot][2]) * m.dx[3];
}

// coord.c:389 bl_coord (LOGRSINTH) with the COORDSINGFIX guard (coord.c:571-579)
void bl_coord(const Mks &m, const double X[4], double V[4]) {
  V[0] = 0.;
  V[1] = m.R0 + exp(X[1]);
  if (X[2] < 0.5) V[2] = M_PI * X[2] + ((1. - m.hslope) / 2.) * sin(2. * M_PI * X[2]);
  else V[2] = M_PI - (M_PI * (1.0 - X[2])) + ((1. - m.hslope) / 2.) * (-sin(2. * M_PI * (1.0 - X[2])));
  V[3] = 2.0 * M_PI * X[3];
  if (fabs(V[2]) < SINGSMALL) {
    if (V[2] >= 0) V[2] = SINGSMALL;
    if (V[2] < 0) V[2] = -SINGSMALL;
  }
  if (fabs(M_PI - V[2]) < SINGSMALL) {
    if (V[2] >= M_PI) V[2] = M_PI + SINGSMALL;
    if (V[2] < M_PI) V[2] = M_PI - SINGSMALL;
  }
}

// coord.c:610 dxdxp_analytic (LOGRSINTH): the Jacobian is diagonal
void dxdxp_diag(const Mks &m, const double X[4], const double V[4], double d[4]) {
  d[0] = 1.0;
  d[1] = V[1] - m.R0;
  if (X[2] < 0.5) d[2] = M_PI + (1. - m.hslope) * M_PI * cos(2. * M_PI * X[2]);
  else d[2] = M_PI + (1. - m.hslope) * M_PI * cos(2. * M_PI * (1.0 - X[2]));
  d[3] = 2.0 * M_PI;
}

// metric.c:697 set_gcov_ksmetric (Kerr-Schild g_{mu nu} at r, theta)
void ks_gcov(double a, double r, double th, double g[4][4]) {
  double sth, cth;
  sincos_ref(th, sth, cth);
  double s2 = sth * sth;
  double rho2 = r * r + a * a * cth * cth;
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) g[p][q] = 0.;
  g[0][0] = -1. + 2. * r / rho2;
  g[0][1] = g[1][0] = 2. * r / rho2;
  g[0][3] = g[3][0] = -2. * a * r * s2 / rho2;
  g[1][1] = 1. + 2. * r / rho2;
  g[1][3] = g[3][1] = -a * s2 * (1. + 2. * r / rho2);
  g[2][2] = rho2;
  g[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
}

// metric.c:863 set_gcon_ksmetric (Kerr-Schild g^{mu nu})
void ks_gcon(double a, double r, double th, double g[4][4]) {
  double cth, sth;
  sincos_ref(th, sth, cth);
  double s2 = sth * sth;
  double rho2 = r * r + a * a * cth * cth;
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) g[p][q] = 0.;
  g[0][0] = -(1. + 2. * r / rho2);
  g[0][1] = g[1][0] = 2. * r / rho2;
  g[1][1] = (r * (r - 2.) + a * a) / rho2;
  g[1][3] = g[3][1] = a / rho2;
  g[2][2] = 1. / rho2;
  g[3][3] = 1. / (rho2 * s2);
}

// metric.c:777 set_gcov_blmetric (Boyer-Lindquist g_{mu nu}); only the rows
// the 4-velocity normalisation reads
void bl_gcov(double a, double r, double th, double g[4][4]) {
  double cth, sth;
  sincos_ref(th, sth, cth);
  double s2 = sth * sth, a2 = a * a, r2 = r * r;
  double DDpert = -2. / r + a2 / r2;
  double DD = 1. + DDpert;
  double mupert = +a2 * cth * cth / r2;
  double mu = 1. + mupert;
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) g[p][q] = 0.;
  g[0][0] = -(1. - 2. / (r * mu));
  g[0][3] = g[3][0] = -2. * a * s2 / (r * mu);
  g[1][1] = mu / DD;
  g[2][2] = r2 * mu;
  g[3][3] = r2 * sth * sth * (1. + a2 / r2 + 2. * a2 * s2 / (r2 * r * mu));
}

// metric.c:11 gcov_func(getprim=1, KSCOORDS) with metric.c:1216 gcov2gcovprim:
// g'_{jk} = g_{lm} T^l_j T^m_k summed in (l,m) order; with the diagonal
// Jacobian only (l,m)=(j,k) is non-zero and the zero terms leave it unchanged
void gcov_prime(const Mks &m, const double X[4], double gp[4][4], double *gpert0) {
  double V[4], d[4], g[4][4];
  bl_coord(m, X, V);
  ks_gcov(m.a, V[1], V[2], g);
  dxdxp_diag(m, X, V, d);
  for (int j = 0; j < 4; j++)
    for (int k = 0; k < 4; k++) gp[j][k] = g[j][k] * d[j] * d[k];
  if (gpert0) *gpert0 = gp[0][0] - (-1.);  // metric.c:1278, g'_tt - eta_tt
}

// metric.c:79 gcon_func(getprim=1, KSCOORDS) with metric.c:1286 gcon2gconprim;
// the Gauss-Jordan inverse of the diagonal Jacobian (metric.c:1346) is 1/d
void gcon_prime(const Mks &m, const double X[4], double gp[4][4]) {
  double V[4], d[4], g[4][4], id[4];
  bl_coord(m, X, V);
  ks_gcon(m.a, V[1], V[2], g);
  dxdxp_diag(m, X, V, d);
  for (int j = 0; j < 4; j++) id[j] = 1.0 / d[j];
  for (int j = 0; j < 4; j++)
    for (int k = 0; k < 4; k++) gp[j][k] = id[j] * id[k] * g[j][k];
}

// metric.c:1310 gdet_func: sqrt(-det g) from an LU decomposition with
// implicit-scaling partial pivoting (Crout's method, Numerical Recipes in C,
// 2nd ed., section 2.3, as ludcmp.c implements it).  Returns -1 for a
// singular or wrong-signature metric.
double gdet_lu(const double g[4][4]) {
  double a[4][4], vv[4], d = 1.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) a[i][j] = g[i][j];
  for (int i = 0; i < 4; i++) {
    double big = 0.0, temp;
    for (int j = 0; j < 4; j++)
      if ((temp = fabs(a[i][j])) > big) big = temp;
    if (big == 0.0) return -1.0;
    vv[i] = 1.0 / big;
  }
  int imax = 0;
  for (int j = 0; j < 4; j++) {
    for (int i = 0; i < j; i++) {
      double sum = a[i][j];
      for (int k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
    }
    double big = 0.0, dum;
    for (int i = j; i < 4; i++) {
      double sum = a[i][j];
      for (int k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
      if ((dum = vv[i] * fabs(sum)) >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < 4; k++) {
        dum = a[imax][k];
        a[imax][k] = a[j][k];
        a[j][k] = dum;
      }
      d = -d;
      vv[imax] = vv[j];
    }
    if (a[j][j] == 0.0) a[j][j] = 1.0e-30;
    if (j != 3) {
      dum = 1.0 / (a[j][j]);
      for (int i = j + 1; i < 4; i++) a[i][j] *= dum;
    }
  }
  for (int j = 0; j < 4; j++) d *= a[j][j];
  if (d >= 0) return -1.0;
  return sqrt(-d);
}

// metric.c:1388 conn_func_numerical1 (centred differences of g'_{ij} in
// each X^k, then Gamma^i_{jk} = g'^{il} Gamma_{ljk})
void conn_numerical(const Mks &m, const double X[4], const double gcon[4][4], double conn[4][4][4]) {
  double c[4][4][4], tmp[4][4][4];
  for (int k = 0; k < 4; k++) {
    double Xh[4], Xl[4], gh[4][4], gl[4][4];
    for (int l = 0; l < 4; l++) {
      Xh[l] = X[l];
      Xl[l] = X[l];
    }
    Xh[k] += CONNDELTA;
    Xl[k] -= CONNDELTA;
    gcov_prime(m, Xh, gh, nullptr);
    gcov_prime(m, Xl, gl, nullptr);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) c[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) tmp[i][j][k] = 0.5 * (c[j][i][k] + c[k][i][j] - c[k][j][i]);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) {
        conn[i][j][k] = 0.;
        for (int l = 0; l < 4; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
      }
}

// prime-coordinate geometry at one point, in the device's Geom form
bool point_geom(const Mks &m, const double X[4], Geom &G) {
  gcov_prime(m, X, G.gcov, &G.gpert0);
  gcon_prime(m, X, G.gcon);
  G.g = gdet_lu(G.gcov);
  return G.g > 0.;
}

// ranc.c:30 ranc: 64 multiplicative congruential streams seeded from the C
// library's rand(), indices chained through the previous output
struct Ranc {
  static constexpr int N = 64;
  int aa[N], S[N], n = 0;
  void seed(int iseed) {
    static const int P[N] = {46337, 46327, 46309, 46307, 46301, 46279, 46273, 46271, 46261, 46237, 46229,
                             46219, 46199, 46187, 46183, 46181, 46171, 46153, 46147, 46141, 46133, 46103,
                             46099, 46093, 46091, 46073, 46061, 46051, 46049, 46027, 46021, 45989, 45979,
                             45971, 45959, 45953, 45949, 45943, 45893, 45887, 45869, 45863, 45853, 45841,
                             45833, 45827, 45823, 45821, 45817, 45779, 45767, 45763, 45757, 45751, 45737,
                             45707, 45697, 45691, 45677, 45673, 45667, 45659, 45641, 45631};
    std::memcpy(Pm, P, sizeof P);
    const int MINN = 2 << 8;
    srand((unsigned)iseed);
    n = 0;
    for (int i = 0; i < N; i++) {
      aa[i] = 0;
      S[i] = 0;
      while ((aa[i] = (0x7fff & rand())) < MINN || aa[i] > Pm[i] - MINN) {
      }
      // the second draw tests aa, not S, exactly as ranc.c does
      while ((S[i] = (0x7fff & rand())) < 1 || aa[i] > Pm[i] - 1) {
      }
    }
  }
  double next() {
    n = S[n] & (N - 1);
    S[n] = (S[n] * aa[n]) % Pm[n];
    return (double)S[n] / (double)Pm[n];
  }
  int Pm[N];
};

// init.fishmon.c:657 lfish_calc: specific angular momentum of the torus
// with pressure maximum at r
double lfish_calc(double a, double r) {
  return (((pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2)) *
           ((-2. * a * r * (pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))) / sqrt(2. * a * sqrt(r) + (-3. + r) * r) +
            ((a + (-2. + r) * sqrt(r)) * (pow(r, 3) + pow(a, 2) * (2. + r))) /
                sqrt(1 + (2. * a) / pow(r, 1.5) - 3. / r))) /
          (pow(r, 3) * sqrt(2. * a * sqrt(r) + (-3. + r) * r) * (pow(a, 2) + (-2. + r) * r)));
}

// transforms.c:330 ucon2pr (VELREL4)
void ucon2pr_rel4(const double ucon[4], const Geom &G, double *pr) {
  double alphasq = 1. / (-G.gcon[0][0]);
  double gammaoalpha = ucon[0];
  double beta[4];
  for (int j = 1; j < 4; j++) beta[j] = alphasq * G.gcon[0][j];
  for (int j = 1; j < 4; j++) pr[U1 + j - 1] = ucon[j] + beta[j] * gammaoalpha;
}

// transforms.c:9 bl2met2metp2v for PRIMECOORDS and VELREL4 input: the
// relative 4-velocity goes through u^mu and back
bool rel4_round_trip(double *pr, const Geom &G) {
  double ucon[4];
  if (ucon_calc(pr, G, ucon)) return true;
  ucon2pr_rel4(ucon, G, pr);
  return false;
}

// the per-(i,j) set-up data of one zone column
struct Column {
  Geom G;           // prime geometry at the zone centre
  double X[4], r, th, d[4];
  double floor_rho, floor_u;
  // torus solution (phi independent part)
  bool torus;
  double rho, u, up;
  double bl[4][4];  // Boyer-Lindquist g_{mu nu}, torus zones only
};

}  // namespace

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
extern "C" int hg_mks_geometry_tables(int n1, int n2, int n3, int nb, double a, double Rin, double Rout,
                                      double R0, double hslope, double *gcov, double *gcon, double *gcovpert,
                                      double *gdet, double *conn, double *r_cent, double dx[4]) {
  if (n1 < 1 || n2 < 1 || n3 < 1 || nb < 0) {
    set_error("hg_mks_geometry_tables: bad grid sizes");
    return 1;
  }
  if (!gcov || !gcon || !gcovpert || !gdet || !conn || !r_cent) {
    set_error("hg_mks_geometry_tables: null argument");
    return 1;
  }
  if (!(Rin > R0) || !(Rout > Rin) || !(hslope > 0.) || !(hslope <= 1.)) {
    set_error("hg_mks_geometry_tables: need R0 < Rin < Rout and 0 < hslope <= 1");
    return 1;
  }
  Dims d = make_dims(n1, n2, n3, nb);
  Mks m = make_mks(a, Rin, Rout, R0, hslope, n1, n2, n3);
  // initbase.c FULLLOOPP1 over every location (set_grid.c:47)
  for (int ii = 0; ii < d.g1; ii++)
    for (int jj = 0; jj < d.g2; jj++)
      for (int s = 0; s < NLOC; s++) {
        size_t h = ((size_t)ii * d.g2 + jj) * NLOC + s;
        double X[4];
        coord(m, ii - d.nb1, jj - d.nb2, 0, s, X);
        Geom G;
        if (!point_geom(m, X, G)) {
          set_error("hg_mks_geometry_tables: metric has bad signature at i=%d j=%d", ii - d.nb1, jj - d.nb2);
          return 1;
        }
        for (int p = 0; p < 4; p++)
          for (int q = 0; q < 4; q++) {
            gcov[h * 16 + p * 4 + q] = G.gcov[p][q];
            gcon[h * 16 + p * 4 + q] = G.gcon[p][q];
          }
        gcovpert[h * 4 + 0] = G.gpert0;
        for (int q = 1; q < 4; q++) gcovpert[h * 4 + q] = G.gcov[q][q] - 1.;
        gdet[h] = G.g;
      }
  // set_grid.c:107 connection at zone centres
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++) {
      double X[4], V[4], gc[4][4], cn[4][4][4];
      coord(m, ii - d.nb1, jj - d.nb2, 0, SLOT_CENT, X);
      bl_coord(m, X, V);
      gcon_prime(m, X, gc);
      conn_numerical(m, X, gc, cn);
      size_t h = (size_t)ii * d.m2 + jj;
      std::memcpy(&conn[h * 64], cn, sizeof cn);
      r_cent[h] = V[1];
    }
  if (dx)
    for (int q = 0; q < 4; q++) dx[q] = m.dx[q];
  return 0;
}

extern "C" int hg_build_mks_geometry(hg_ctx *c, double a, double Rin, double Rout, double R0, double hslope) {
  if (!c) {
    set_error("hg_build_mks_geometry: null context");
    return 1;
  }
  const Dims &d = c->d;
  size_t nl = (size_t)d.g1 * d.g2 * NLOC, nc = (size_t)d.m1 * d.m2;
  std::vector<double> gcov(nl * 16), gcon(nl * 16), gpert(nl * 4), gdet(nl), conn(nc * 64), r(nc);
  double dx[4];
  if (hg_mks_geometry_tables(c->cfg.n1, c->cfg.n2, c->cfg.n3, c->cfg.nb, a, Rin, Rout, R0, hslope, gcov.data(),
                             gcon.data(), gpert.data(), gdet.data(), conn.data(), r.data(), dx))
    return 1;
  for (int q = 1; q < 4; q++) c->cfg.dx[q] = dx[q];
  if (upload_geometry_tables(c, gcov.data(), gcon.data(), gpert.data(), gdet.data(), conn.data(), r.data()))
    return 1;
  c->have_mks = true;
  c->mks_a = a;
  c->mks_Rin = Rin;
  c->mks_Rout = Rout;
  c->mks_R0 = R0;
  c->mks_hslope = hslope;
  return 0;
}

// ---------------------------------------------------------------------------
// torus initial data
// ---------------------------------------------------------------------------
static int device_bound(hg_ctx *c, std::vector<double> &P) {
  if (copy_prims(c, c->p, P.data())) return 1;
  launch_bound(c->p, c->geo, c->d, c->cfg.gammamax, c->err, 0);
  hipError_t e = hipStreamSynchronize(0);
  int flag = 0;
  if (e == hipSuccess) e = hipMemcpy(&flag, c->err, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error("hg_init_torus: ghost fill: %s", hipGetErrorString(e));
    return 1;
  }
  if (flag) {
    (void)hipMemset(c->err, 0, sizeof(int));
    set_error("hg_init_torus: ghost fill reported failure mask %d", flag);
    return 1;
  }
  return fetch_prims(c, c->p, P.data());
}

extern "C" int hg_init_torus(hg_ctx *c, double rin, double rmax, double kappa, double beta, double randfact,
                             uint64_t seed, double *dt) {
  if (!c || !dt) {
    set_error("hg_init_torus: null argument");
    return 1;
  }
  if (!c->have_mks) {
    set_error("hg_init_torus: call hg_build_mks_geometry first");
    return 1;
  }
  const Dims &d = c->d;
  const hg_config &f = c->cfg;
  const double a = c->mks_a, gam = f.gam;
  Mks m = make_mks(a, c->mks_Rin, c->mks_Rout, c->mks_R0, c->mks_hslope, f.n1, f.n2, f.n3);

  // per-column data: geometry and the phi-independent torus solution
  // (init.fishmon.c:364 init_dsandvels)
  const double l = lfish_calc(a, rmax);
  const double thin = M_PI / 2.;
  const double sthin = sin(thin), cthin = cos(thin);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  std::vector<Column> col((size_t)d.m1 * d.m2);
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++) {
      Column &C = col[(size_t)ii * d.m2 + jj];
      double V[4];
      coord(m, ii - d.nb1, jj - d.nb2, 0, SLOT_CENT, C.X);
      bl_coord(m, C.X, V);
      dxdxp_diag(m, C.X, V, C.d);
      if (!point_geom(m, C.X, C.G)) {
        set_error("hg_init_torus: bad metric at i=%d j=%d", ii - d.nb1, jj - d.nb2);
        return 1;
      }
      double r = V[1], th = V[2];
      C.r = r;
      C.th = th;
      // fixup.c:1562 set_density_floors (rescaletype 0) and fixup.c:1686 set_atmosphere
      C.floor_rho = f.rhomin * pow(r, -1.5);
      C.floor_u = f.uumin * pow(r, -2.5);
      double sth, cth;
      sincos_ref(th, sth, cth);
      double DD = r * r - 2. * r + a * a;
      double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
      double SS = r * r + a * a * cth * cth;
      double lnh;
      if (r >= rin) {
        lnh = 0.5 * log((1. + sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) / (SS * DD / AA)) -
              0.5 * sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) - 2. * a * r * l / AA -
              (0.5 * log((1. + sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin))) /
                         (SSin * DDin / AAin)) -
               0.5 * sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin)) -
               2. * a * rin * l / AAin);
      } else {
        lnh = 1.;
      }
      C.torus = !(lnh < 0. || r < rin);
      if (C.torus) {
        double hm1 = exp(lnh) - 1.;
        C.rho = pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
        C.u = kappa * pow(C.rho, gam) / (gam - 1.);
        double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
        double up1 = sqrt((-1. + sqrt(1. + 4. * l * l * expm2chi)) / 2.);
        C.up = 2. * a * r * sqrt(1. + up1 * up1) / sqrt(AA * SS * DD) + sqrt(SS / AA) * up1 / sth;
        bl_gcov(a, r, th, C.bl);
      }
    }

  // FULLLOOP (i, then j, then k) over every zone: torus or atmosphere, then
  // the transformation to the evolved velocity (transforms.c:9)
  std::vector<double> P((size_t)d.m1 * d.m2 * d.m3 * NPR);
  auto Z = [&](int ii, int jj, int kk) { return (((size_t)ii * d.m2 + jj) * d.m3 + kk) * NPR; };
  // init.fishmon.c:134 seeds with ranc(7); that call also returns (and the
  // set-up discards) the first draw
  Ranc rng;
  rng.seed((int)seed);
  (void)rng.next();
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++) {
      const Column &C = col[(size_t)ii * d.m2 + jj];
      for (int kk = 0; kk < d.m3; kk++) {
        double *pr = &P[Z(ii, jj, kk)];
        if (!C.torus) {
          // fixup.c:1665 set_atmosphere(whichcond=-1, VELREL4): floor values at rest, no field
          pr[RHO] = f.rhomin * pow(C.r, -1.5);
          pr[UU] = f.uumin * pow(C.r, -2.5);
          pr[U1] = pr[U2] = pr[U3] = 0.0;
          pr[B1] = pr[B2] = pr[B3] = 0;
          if (rel4_round_trip(pr, C.G)) {
            set_error("hg_init_torus: space-like atmosphere velocity");
            return 1;
          }
          continue;
        }
        pr[RHO] = C.rho;
        pr[UU] = C.u * (1. + randfact * (rng.next() - 0.5));
        pr[U1] = 0.;
        pr[U2] = 0.;
        pr[U3] = 1.0 * C.up;
        pr[B1] = pr[B2] = pr[B3] = 0.;
        // phys.c:1192 ucon_calc_4vel with the Boyer-Lindquist metric
        double ucon[4] = {0., pr[U1], pr[U2], pr[U3]};
        const double(*g)[4] = C.bl;
        double AA = g[0][0];
        double BB = 2. * (g[0][1] * ucon[1] + g[0][2] * ucon[2] + g[0][3] * ucon[3]);
        double CCM1 = g[1][1] * ucon[1] * ucon[1] + g[2][2] * ucon[2] * ucon[2] + g[3][3] * ucon[3] * ucon[3] +
                      2. * (g[1][2] * ucon[1] * ucon[2] + g[1][3] * ucon[1] * ucon[3] + g[2][3] * ucon[2] * ucon[3]);
        double CC = CCM1 + 1.0;
        double discr = BB * BB - 4. * AA * CC;
        if (discr < 0.) {
          set_error("hg_init_torus: space-like torus velocity at i=%d j=%d", ii - d.nb1, jj - d.nb2);
          return 1;
        }
        ucon[0] = (-BB - sqrt(discr)) / (2. * AA);
        // transforms.c:109 bltoks, then transforms.c:269 mettometp, summed
        // term by term as the reference's matrix products
        double r = C.r;
        double T[4][4] = {{1, 2. * r / (r * r - 2. * r + a * a), 0, 0},
                          {0, 1, 0, 0},
                          {0, 0, 1, 0},
                          {0, a / (r * r - 2. * r + a * a), 0, 1}};
        double tmp[4];
        for (int j = 0; j < 4; j++) {
          tmp[j] = 0.;
          for (int k = 0; k < 4; k++) tmp[j] += T[j][k] * ucon[k];
        }
        for (int j = 0; j < 4; j++) ucon[j] = tmp[j];
        for (int j = 0; j < 4; j++) {
          tmp[j] = 0.;
          for (int k = 0; k < 4; k++) tmp[j] += ((j == k) ? 1.0 / C.d[j] : 0.) * ucon[k];
        }
        ucon2pr_rel4(tmp, C.G, pr);
      }
    }

  // init.fishmon.c:554 normalize_densities over the active zones
  double rhomax = 0, umax = 0;
  for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
    for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++) {
      const Column &C = col[(size_t)ii * d.m2 + jj];
      for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
        const double *pr = &P[Z(ii, jj, kk)];
        if (pr[RHO] > rhomax) rhomax = pr[RHO];
        if (pr[UU] > umax && C.r > rin) umax = pr[UU];
      }
    }
  if (!(rhomax > 0.)) {
    set_error("hg_init_torus: the torus misses the grid");
    return 1;
  }
  for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
    for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++)
      for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
        double *pr = &P[Z(ii, jj, kk)];
        pr[RHO] /= rhomax;
        pr[UU] /= rhomax;
      }
  umax /= rhomax;
  rhomax = 1.;

  // init.fishmon.c:179 init_atmosphere + round trip, fixup.c:100 fixup
  auto fixup_active = [&]() -> int {
    for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
      for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++) {
        const Column &C = col[(size_t)ii * d.m2 + jj];
        double f0 = C.floor_rho, f1 = C.floor_u;
        if (f0 < 1.e-20) f0 = 1.e-20;
        if (f1 < 1.e-20) f1 = 1.e-20;
        for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
          double *pr = &P[Z(ii, jj, kk)];
          if (f0 > pr[RHO]) pr[RHO] = f0;
          if (f1 > pr[UU]) pr[UU] = f1;
          if (limit_gamma(f.gammamax, pr, C.G) > 0) return 1;
        }
      }
    return 0;
  };
  for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
    for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++) {
      const Column &C = col[(size_t)ii * d.m2 + jj];
      for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
        double *pr = &P[Z(ii, jj, kk)];
        if (pr[RHO] < C.floor_rho) {
          pr[RHO] = C.floor_rho;
          pr[UU] = C.floor_u;
          pr[U1] = pr[U2] = pr[U3] = 0.0;
        }
        if (rel4_round_trip(pr, C.G)) {
          set_error("hg_init_torus: space-like velocity after the atmosphere");
          return 1;
        }
      }
    }
  if (fixup_active()) {
    set_error("hg_init_torus: limit_gamma failed");
    return 1;
  }
  if (device_bound(c, P)) return 1;

  // init.fishmon.c:488 init_vpot (A_phi at CORN3 from the averaged density)
  // and initbase.c:558 vpot2field.  A_phi on the last corner row and column
  // (i = N1+N1BND, j = N2+N2BND) averages zones past the end of the
  // reference's arrays; those values only reach ghost-zone fields, which the
  // ghost fill below overwrites, so the nearest zone stands in here.
  const int c1 = d.m1 + 1, c2 = d.m2 + 1;
  std::vector<double> A3((size_t)c1 * c2);
  auto rho_at = [&](int ii, int jj, int kk) {
    if (ii > d.m1 - 1) ii = d.m1 - 1;
    if (jj > d.m2 - 1) jj = d.m2 - 1;
    return P[Z(ii, jj, kk) + RHO];
  };
  const int sh1 = d.n1 > 1, sh2 = d.n2 > 1, sh3 = d.n3 > 1;
  for (int kk = 0; kk < d.m3; kk++) {
    for (int ii = 0; ii < c1; ii++)
      for (int jj = 0; jj < c2; jj++) {
        double rho_av;
        if (ii == 0 && jj == 0) rho_av = rho_at(ii, jj, kk);
        else if (ii == 0) rho_av = 0.5 * (rho_at(ii, jj, kk) + rho_at(ii, jj - sh2, kk));
        else if (jj == 0) rho_av = 0.5 * (rho_at(ii, jj, kk) + rho_at(ii - sh1, jj, kk));
        else
          rho_av = 0.5 * (0.5 * (rho_at(ii, jj, kk) + rho_at(ii, jj - sh2, kk)) +
                          0.5 * (rho_at(ii - sh1, jj, kk) + rho_at(ii - sh1, jj - sh2, kk)));
        double q = rho_av / rhomax - 0.2;
        double A = 0;
        if (q > 0.) A += q;
        A3[(size_t)ii * c2 + jj] = A;
      }
    for (int ii = 0; ii < d.m1; ii++)
      for (int jj = 0; jj < d.m2; jj++) {
        const Column &C = col[(size_t)ii * d.m2 + jj];
        double *pr = &P[Z(ii, jj, kk)];
        double g = C.G.g;
        auto a3 = [&](int i2, int j2) { return A3[(size_t)i2 * c2 + j2]; };
        int ip = ii + sh1, jp = jj + sh2;
        (void)sh3;
        // AVGCORN_1(F,i,j) = 0.5*(F[i+1][j] + F[i][j]); AVGCORN_2(F,i,j) = 0.5*(F[i][j+1] + F[i][j])
        pr[B1] = +((0.5 * (a3(ip, jp) + a3(ii, jp))) - (0.5 * (a3(ip, jj) + a3(ii, jj)))) / (g * m.dx[2]);
        pr[B1] += -((0.5 * (0. + 0.)) - (0.5 * (0. + 0.))) / (g * m.dx[3]);
        pr[B2] = +((0.5 * (0. + 0.)) - (0.5 * (0. + 0.))) / (g * m.dx[3]);
        pr[B2] += -((0.5 * (a3(ip, jp) + a3(ip, jj))) - (0.5 * (a3(ii, jp) + a3(ii, jj)))) / (g * m.dx[1]);
        pr[B3] = +((0.5 * (0. + 0.)) - (0.5 * (0. + 0.))) / (g * m.dx[1]);
        pr[B3] += -((0.5 * (0. + 0.)) - (0.5 * (0. + 0.))) / (g * m.dx[2]);
      }
  }

  // init.fishmon.c:589 normalize_field: beta_min = (gam-1) umax / (b^2_max / 2)
  double bsq_max = 0.;
  for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
    for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++) {
      const Column &C = col[(size_t)ii * d.m2 + jj];
      for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
        State q;
        if (get_state(&P[Z(ii, jj, kk)], C.G, q)) {
          set_error("hg_init_torus: get_state failed while normalising the field");
          return 1;
        }
        double bsq = q.bcon[0] * q.bcov[0] + q.bcon[1] * q.bcov[1] + q.bcon[2] * q.bcov[2] + q.bcon[3] * q.bcov[3];
        if (bsq > bsq_max) bsq_max = bsq;
      }
    }
  if (!(bsq_max > 0.)) {
    set_error("hg_init_torus: no field inside the torus");
    return 1;
  }
  double beta_act = (gam - 1.) * umax / (0.5 * bsq_max);
  double norm = sqrt(beta_act / beta);
  for (int ii = d.nb1; ii < d.nb1 + d.n1; ii++)
    for (int jj = d.nb2; jj < d.nb2 + d.n2; jj++)
      for (int kk = d.nb3; kk < d.nb3 + d.n3; kk++) {
        double *pr = &P[Z(ii, jj, kk)];
        pr[B1] *= norm;
        pr[B2] *= norm;
        pr[B3] *= norm;
      }

  // initbase.c init(): fixup, ghost fill, set_dt
  if (fixup_active()) {
    set_error("hg_init_torus: limit_gamma failed");
    return 1;
  }
  if (copy_prims(c, c->p, P.data())) return 1;
  c->ucons_stale = true;
  c->unew_in_uf = false;
  launch_bound(c->p, c->geo, c->d, f.gammamax, c->err, 0);
  launch_reset_ndt(c->ndt_bits, 0);
  launch_set_dt(c->p, c->geo, c->d, f.gam, f.cour, f.dx, c->ndt_bits, c->err, 0);
  hipError_t e = hipStreamSynchronize(0);
  int flag = 0;
  unsigned long long bits = 0;
  if (e == hipSuccess) e = hipMemcpy(&flag, c->err, sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&bits, c->ndt_bits, sizeof bits, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error("hg_init_torus: %s", hipGetErrorString(e));
    return 1;
  }
  if (flag) {
    (void)hipMemset(c->err, 0, sizeof(int));
    set_error("hg_init_torus: device reported failure mask %d", flag);
    return 1;
  }
  double ndt;
  std::memcpy(&ndt, &bits, sizeof ndt);
  *dt = ndt;
  return hg_set_dt(c, ndt);
}
