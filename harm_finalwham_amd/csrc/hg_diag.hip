// Device-side diagnostics of the reference's dump_ener()/diag():
//   constotal  (dump_ener.c:596)  sum of U dV over the active zones
//   divbmaxavg (dump_ener.c:482)  max and mean of the flux-CT div B
//   diag_flux  (diag.c:734)       fluxes through the domain boundary, their
//                                 term-by-term split, the split restricted
//                                 to the jet region, and the time integral
// for the global energy region.  diag_flux is the call of advance.c:130
// (ACCURATEDIAG): the last fluxcalc's fluxes with the state it read (after a
// step, the RK midpoint state of the final substep) and, for two of its
// terms, the reference's global p at that point (the state at the start of
// the step), which the step records when hg_diag_record is on.  The per-zone and per-face terms are formed
// in parallel with the reference's arithmetic; every sum is then taken by one
// thread in the reference's loop order, so the sums are bit-identical to the
// reference's sequential double sums.  The div B normalisation of the
// reference is evaluated in x87 extended precision (global.h:505 defines
// THIRD as a long double); here it is FP64, which leaves a few units in the
// last place of difference per zone.
#include <hip/hip_runtime.h>

#include "harmgpu.h"
#include "hg_host.h"
#include "hg_phys.h"

using namespace hg;

namespace {

constexpr int NFLT = 7;              // global.h:885 NUMFLUXTERMS
constexpr double JETBSQORHO = 3.162;  // diag.c:729
// face-term record: 8 flux values, 7 x 8 terms, the jet flag
constexpr int FREC = NPR + NFLT * NPR + 1;

struct DiagView {
  double *ud;      // [8][nact] U dV per zone, ZLOOP order
  double *divb;    // [nact]
  double *face;    // [nface][FREC]
  double *out;     // results: cons 8, divbmax, divbavg, imax (as double), pdot 6x8, terms 6x7x8, jet2 6x7x8
  double *stage;   // [4][6][8] pdot of each RK4 substep (for pcum)
  int nface[6];    // faces per direction (X1UP, X1DN, X2UP, X2DN, X3UP, X3DN), 0 = no flux there
  int face0[6];    // first record of each direction
};

constexpr int OUT_CONS = 0, OUT_DIVBMAX = 8, OUT_DIVBAVG = 9, OUT_DIVBIDX = 10, OUT_PDOT = 11,
              OUT_TERMS = OUT_PDOT + 6 * NPR, OUT_JET2 = OUT_TERMS + 6 * NFLT * NPR,
              OUT_N = OUT_JET2 + 6 * NFLT * NPR;

__device__ __forceinline__ void load_full_geom(const double *geo, int loc, const Dims &d, int i, int j, Geom &G) {
  const double *b = geo + (int64_t)loc * GCOMP * d.gplane + d.gidx(i, j);
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

// global.h:1586-1645 Fg, AVG_1/2, AVG_for1/2/3 on B^i gdet(CENT)
struct DivbCtx {
  const double *p, *gcent, *gcorn3;
  Dims d;
  int s1, s2, s3;
  __device__ double Fg(int i, int j, int k, int pl) const {
    return p[d.idx(i, j, k) + pl * d.spl] * gcent[d.gidx(i, j)];
  }
  __device__ double avg1(int i, int j, int k, int pl) const { return 0.5 * (Fg(i, j, k, pl) + Fg(i - s1, j, k, pl)); }
  __device__ double avg2(int i, int j, int k, int pl) const { return 0.5 * (Fg(i, j, k, pl) + Fg(i, j - s2, k, pl)); }
  __device__ double for1(int i, int j, int k) const { return 0.5 * (avg2(i, j, k, B1) + avg2(i, j, k - s3, B1)); }
  __device__ double for2(int i, int j, int k) const { return 0.5 * (avg1(i, j, k, B2) + avg1(i, j, k - s3, B2)); }
  __device__ double for3(int i, int j, int k) const { return 0.5 * (avg2(i, j, k, B3) + avg2(i - s1, j, k, B3)); }
};

__global__ void __launch_bounds__(256) k_dg_zone(const double *__restrict__ p, const double *__restrict__ geo,
                                                 const double *__restrict__ corn, const double *__restrict__ ucons,
                                                 Dims d, double gam, double dVF, double dx1, double dx2,
                                                 double dx3, DiagView v, int *__restrict__ err) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nact = (int64_t)d.n1 * d.n2 * d.n3;
  if (t >= nact) return;
  // ZLOOP order: i outer, k inner
  int k = (int)(t % d.n3);
  int64_t r = t / d.n3;
  int j = (int)(r % d.n2);
  int i = (int)(r / d.n2);
  int64_t z = d.idx(i, j, k);
  Geom G;
  load_full_geom(geo, GL_CENT, d, i, j, G);
  double pr[NPR], U[NPR];
  if (ucons) {
    // dump_ener.c:611-618: the finite-volume scheme sums its carried unew
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) U[pl] = ucons[z + pl * d.spl];
  } else {
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) pr[pl] = p[z + pl * d.spl];
    if (primtoU(pr, G, gam, U)) atomicOr(err, 64);
  }
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) v.ud[(int64_t)pl * nact + t] = U[pl] * dVF;
  // SETFDIVBFLUXCT: every active zone satisfies DIVBCONDITION
  DivbCtx c{p, geo + ((int64_t)GL_CENT * GCOMP + 33) * d.gplane, corn + 2 * d.gplane, d, d.n1 > 1, d.n2 > 1,
            d.n3 > 1};
  double f1 = c.for1(i, j, k), f1m = c.for1(i - c.s1, j, k);
  double f2 = c.for2(i, j, k), f2m = c.for2(i, j - c.s2, k);
  double f3 = c.for3(i, j, k), f3m = c.for3(i, j, k - c.s3);
  double dmax = rmax(rmax(dx1, dx2), dx3);
  double nrm = dmax / ((1.0 / 3.0) * c.gcorn3[d.gidx(i, j)] * fabs((f1 + f1m) + (f2 + f2m) + (f3 + f3m)) + 1.e-50);
  double dv = ((f1 - f1m) / dx1 + (f2 - f2m) / dx2 + (f3 - f3m) / dx3) * nrm;
  v.divb[t] = fabs(dv);
}

// direction index of diag.c's DIRLOOP: X1UP 0, X1DN 1, X2UP 2, X2DN 3, X3UP 4, X3DN 5
__global__ void __launch_bounds__(256) k_dg_face(const double *__restrict__ p, const double *__restrict__ pglob,
                                                 const double *__restrict__ geo,
                                                 const double *__restrict__ F1, const double *__restrict__ F2,
                                                 const double *__restrict__ F3, const double *__restrict__ e1,
                                                 const double *__restrict__ e2, const double *__restrict__ e3,
                                                 Dims d, double gam, double dx1, double dx2, double dx3, DiagView v,
                                                 int *__restrict__ err) {
  int dir = blockIdx.y;
  int n = v.nface[dir];
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  int i, j, k, fluxdir;
  double surface;
  const double *flux;
  // GENLOOP order over the face (i outer, k inner), doflux = 0 or N
  if (dir <= 1) {
    i = (dir == 0) ? d.n1 : 0;
    j = f / d.n3;
    k = f % d.n3;
    surface = dx2 * dx3;
    flux = F1;
    fluxdir = 1;
  } else if (dir <= 3) {
    j = (dir == 2) ? d.n2 : 0;
    i = f / d.n3;
    k = f % d.n3;
    surface = dx1 * dx3;
    flux = F2;
    fluxdir = 2;
  } else {
    k = (dir == 4) ? d.n3 : 0;
    i = f / d.n2;
    j = f % d.n2;
    surface = dx1 * dx2;
    flux = F3;
    fluxdir = 3;
  }
  int64_t z = d.idx(i, j, k);
  const int64_t P = d.spl;
  const int64_t ip = s1 ? 1 : 0, jp = s2 ? d.sj : 0, kp = s3 ? d.sk : 0;
  // the face's fluxes; the field components as flux_ct left them (fluxct.c)
  double Fl[NPR];
#pragma unroll
  for (int pl = 0; pl < B1; pl++) Fl[pl] = flux[z + pl * P];
  if (fluxdir == 1) {
    Fl[B1] = 0.;
    Fl[B2] = 0.5 * (e3[z] + e3[z + jp]);
    Fl[B3] = -0.5 * (e2[z] + e2[z + kp]);
  } else if (fluxdir == 2) {
    Fl[B1] = -0.5 * (e3[z] + e3[z + ip]);
    Fl[B2] = 0.;
    Fl[B3] = 0.5 * (e1[z] + e1[z + kp]);
  } else {
    Fl[B1] = 0.5 * (e2[z] + e2[z + ip]);
    Fl[B2] = -0.5 * (e1[z] + e1[z + jp]);
    Fl[B3] = 0.;
  }
  double *rec = v.face + (int64_t)(v.face0[dir] + f) * FREC;
  // UtoU(UEVOLVE, UDIAG): divide by the EOM factor (gdet), multiply by gdet
  {
    int gpos = (fluxdir == 1) ? GL_FACE1 : (fluxdir == 2) ? GL_FACE2 : GL_FACE3;
    double gf = geo[((int64_t)gpos * GCOMP + 33) * d.gplane + d.gidx(i, j)];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      double ft = Fl[pl] * surface;
      rec[pl] = (ft / gf) * gf;
    }
  }
  // term-by-term split from the zone-centred state (diag.c:860-990)
  Geom G;
  load_full_geom(geo, GL_CENT, d, i, j, G);
  double pr[NPR];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pr[pl] = p[z + pl * P];
  State q;
  if (get_state(pr, G, q)) {
    atomicOr(err, 64);
    return;
  }
  double pgas = (gam - 1.) * pr[UU];
  double bsq = 0;
  for (int l = 0; l < 4; l++) bsq += (q.bcon[l]) * (q.bcov[l]);
  int condjet2 = 0;
  if (dir == 1) {
    double bsqorho = bsq / pr[RHO];
    condjet2 = (bsqorho > JETBSQORHO);
  } else if (dir == 0) {
    condjet2 = ((q.ucon[1] > 0.0) && (-q.ucov[0] - 1.0 > 0.0));
  }
  double surgdet = (G.g) * surface;
  double *tm = rec + NPR;
  for (int e = 0; e < NFLT * NPR; e++) tm[e] = 0.0;
  const double uf = q.ucon[fluxdir];
  tm[0 * NPR + RHO] = pr[RHO] * uf * surgdet;
  for (int pl = UU; pl <= U3; pl++) {
    int l = pl - UU;
    double dl = (fluxdir == l) ? 1.0 : 0.0;
    tm[0 * NPR + pl] = pgas * uf * (q.ucov[l]) * surgdet;
    tm[1 * NPR + pl] = pglob[z + RHO * P] * uf * (q.ucov[l]) * surgdet;
    tm[2 * NPR + pl] = pglob[z + UU * P] * uf * (q.ucov[l]) * surgdet;
    tm[3 * NPR + pl] = bsq * uf * (q.ucov[l]) * surgdet;
    tm[4 * NPR + pl] = pgas * dl * surgdet;
    tm[5 * NPR + pl] = 0.5 * bsq * dl * surgdet;
    tm[6 * NPR + pl] = -(q.bcon[fluxdir]) * (q.bcov[l]) * surgdet;
  }
  for (int c = 1; c <= 3; c++) {
    tm[0 * NPR + B1 + c - 1] = (q.bcon[c]) * uf * surgdet;
    tm[1 * NPR + B1 + c - 1] = -(q.bcon[fluxdir]) * (q.ucon[c]) * surgdet;
  }
  rec[FREC - 1] = condjet2 ? 1.0 : 0.0;
}

// pdot alone (diag.c:760-1007 for pl < NPR), one thread per sum, into the
// record of RK4 substep `stage`
__global__ void __launch_bounds__(64) k_dg_pdot(DiagView v, int stage) {
  int t = threadIdx.x;
  if (t >= 6 * NPR) return;
  int dir = t / NPR, pl = t % NPR;
  int n = v.nface[dir];
  const double *rec = v.face + (int64_t)v.face0[dir] * FREC;
  double s = 0.0;
  for (int f = 0; f < n; f++) s += rec[(int64_t)f * FREC + pl];
  v.stage[stage * 6 * NPR + t] = s;
}

// every sum by one thread in loop order
__global__ void __launch_bounds__(1024) k_dg_sum(DiagView v, int64_t nact) {
  int t = threadIdx.x;
  if (t < NPR) {
    const double *a = v.ud + (int64_t)t * nact;
    double s = 0.0;
    for (int64_t z = 0; z < nact; z++) s += a[z];
    v.out[OUT_CONS + t] = s;
  } else if (t == NPR) {
    double mx = 0, sum = 0;
    int64_t imx = 0;
    for (int64_t z = 0; z < nact; z++) {
      double dv = v.divb[z];
      if (dv > mx) {
        mx = dv;
        imx = z;
      }
      sum += dv;
    }
    v.out[OUT_DIVBMAX] = mx;
    v.out[OUT_DIVBAVG] = sum / (double)nact;
    v.out[OUT_DIVBIDX] = (double)imx;
  } else {
    // (dir, kind, fl, pl): kind 0 pdot, 1 terms, 2 jet2 terms
    int u = t - (NPR + 1);
    const int per_dir = NPR + 2 * NFLT * NPR;
    if (u >= 6 * per_dir) return;
    int dir = u / per_dir, w = u % per_dir;
    int n = v.nface[dir];
    const double *rec = v.face + (int64_t)v.face0[dir] * FREC;
    double s = 0.0;
    if (w < NPR) {
      for (int f = 0; f < n; f++) s += rec[(int64_t)f * FREC + w];
      v.out[OUT_PDOT + dir * NPR + w] = s;
    } else if (w < NPR + NFLT * NPR) {
      int e = w - NPR;
      for (int f = 0; f < n; f++) s += rec[(int64_t)f * FREC + NPR + e];
      v.out[OUT_TERMS + dir * NFLT * NPR + e] = s;
    } else {
      int e = w - NPR - NFLT * NPR;
      for (int f = 0; f < n; f++)
        if (rec[(int64_t)f * FREC + FREC - 1] != 0.0) s += rec[(int64_t)f * FREC + NPR + e];
      v.out[OUT_JET2 + dir * NFLT * NPR + e] = s;
    }
  }
}

}  // namespace

static DiagView view(hg_ctx *c) {
  const Dims &d = c->d;
  const int64_t nact = (int64_t)d.n1 * d.n2 * d.n3;
  DiagView v;
  // diag.c:1151-1189: a boundary carries flux only where its dimension exists
  v.nface[0] = v.nface[1] = (d.n1 > 1) ? d.n2 * d.n3 : 0;
  v.nface[2] = v.nface[3] = (d.n2 > 1) ? d.n1 * d.n3 : 0;
  v.nface[4] = v.nface[5] = (d.n3 > 1) ? d.n1 * d.n2 : 0;
  int nf = 0;
  for (int k = 0; k < 6; k++) {
    v.face0[k] = nf;
    nf += v.nface[k];
  }
  v.ud = c->dg_buf;
  v.divb = v.ud ? v.ud + (size_t)NPR * nact : nullptr;
  v.face = v.ud ? v.divb + nact : nullptr;
  v.out = v.ud ? v.face + (size_t)nf * FREC : nullptr;
  v.stage = v.ud ? v.out + OUT_N : nullptr;
  return v;
}

static size_t buf_doubles(const Dims &d, const DiagView &v) {
  int nf = v.face0[5] + v.nface[5];
  return (size_t)(NPR + 1) * ((int64_t)d.n1 * d.n2 * d.n3) + (size_t)nf * FREC + OUT_N + 4 * 6 * NPR;
}

namespace hg {

void diag_free(hg_ctx *c) {
  if (c->dg_buf) (void)hipFree(c->dg_buf);
  c->dg_buf = nullptr;
}

int diag_alloc(hg_ctx *c) {
  if (c->dg_buf) return 0;
  DiagView v = view(c);
  hipError_t e = hipMalloc((void **)&c->dg_buf, buf_doubles(c->d, v) * sizeof(double));
  if (e != hipSuccess) {
    c->dg_buf = nullptr;
    set_error("hg_diagnostics: %s", hipGetErrorString(e));
    return 1;
  }
  return 0;
}

// diag_flux's per-face terms (advance.c:130): fluxes of the last fluxcalc,
// prim = the state it read, pglob = the global p of the reference at that
// point (the state at the start of the step)
void launch_diag_faces(hg_ctx *c, const double *prim, const double *pglob, hipStream_t s) {
  DiagView v = view(c);
  const Dims &d = c->d;
  const double *dx = c->cfg.dx;
  int fmax = 1;
  for (int k = 0; k < 6; k++) fmax = v.nface[k] > fmax ? v.nface[k] : fmax;
  hipLaunchKernelGGL(k_dg_face, dim3((unsigned)((fmax + 255) / 256), 6), dim3(256), 0, s, prim, pglob, c->geo,
                     c->F[0], c->F[1], c->F[2], c->emf[0], c->emf[1], c->emf[2], d, c->cfg.gam, dx[1], dx[2], dx[3],
                     v, c->err);
}

// the pdot of RK4 substep `stage` from the faces just recorded
void launch_diag_stage(hg_ctx *c, int stage, hipStream_t s) {
  hipLaunchKernelGGL(k_dg_pdot, dim3(1), dim3(64), 0, s, view(c), stage);
}

}  // namespace hg

extern "C" int hg_diag_record(hg_ctx *c, int on) {
  if (!c) {
    set_error("hg_diag_record: null context");
    return 1;
  }
  if (on && diag_alloc(c)) return 1;
  c->dg_record = on != 0;
  c->dg_faces_valid = false;
  return 0;
}

extern "C" int hg_diagnostics(hg_ctx *c, double Dt, hg_diag *out, void *stream) {
  if (!c || !out) {
    set_error("hg_diagnostics: null argument");
    return 1;
  }
  if (!c->have_geometry) {
    set_error("hg_diagnostics: geometry not set");
    return 1;
  }
  if (diag_alloc(c)) return 1;
  const Dims &d = c->d;
  const int64_t nact = (int64_t)d.n1 * d.n2 * d.n3;
  DiagView v = view(c);
  hipStream_t s = (hipStream_t)stream;
  const double *dx = c->cfg.dx;
  const double dVF = dx[1] * dx[2] * dx[3];  // set_grid.c:21-22
  ensure_conserved(c, s);
  hipLaunchKernelGGL(k_dg_zone, dim3((unsigned)((nact + 255) / 256)), dim3(256), 0, s, c->p, c->geo, c->corn,
                     c->ucons, d, c->cfg.gam, dVF, dx[1], dx[2], dx[3], v, c->err);
  // the faces recorded by the last step, else those of the last fluxcalc on p
  if (!(c->dg_record && c->dg_faces_valid))
    launch_diag_faces(c, c->flux_state ? c->flux_state : c->p, c->p, s);
  hipLaunchKernelGGL(k_dg_sum, dim3(1), dim3(1024), 0, s, v, nact);
  double h[OUT_N], hs[4 * 6 * NPR];
  // RK4 with recording: every substep's pdot enters pcum (advance.c:130
  // calls diag_flux with Dt = Cunew[1] dt on each of them)
  const bool stages = c->cfg.timeorder == 4 && c->dg_record && c->dg_faces_valid;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, v.out, sizeof h, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && stages) e = hipMemcpyAsync(hs, v.stage, sizeof hs, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error("hg_diagnostics: %s", hipGetErrorString(e));
    return 1;
  }
  int em = 0;
  e = hipMemcpy(&em, c->err, sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && (em & 64)) {
    (void)hipMemset(c->err, 0, sizeof(int));
    set_error("hg_diagnostics: state of a zone failed (get_state/primtoU)");
    return 1;
  }
  for (int pl = 0; pl < NPR; pl++) out->cons[pl] = h[OUT_CONS + pl];
  out->divbmax = h[OUT_DIVBMAX];
  out->divbavg = h[OUT_DIVBAVG];
  int64_t im = (int64_t)h[OUT_DIVBIDX];
  out->divbmax_ijk[2] = (int32_t)(im % d.n3);
  out->divbmax_ijk[1] = (int32_t)((im / d.n3) % d.n2);
  out->divbmax_ijk[0] = (int32_t)(im / ((int64_t)d.n3 * d.n2));
  for (int dir = 0; dir < 6; dir++)
    for (int pl = 0; pl < NPR; pl++) {
      out->pdot[dir][pl] = h[OUT_PDOT + dir * NPR + pl];
      for (int fl = 0; fl < NFLT; fl++) {
        out->pdotterms[dir][fl][pl] = h[OUT_TERMS + (dir * NFLT + fl) * NPR + pl];
        out->pdottermsjet2[dir][fl][pl] = h[OUT_JET2 + (dir * NFLT + fl) * NPR + pl];
      }
    }
  // diag.c:1007 pcum += pdot Dt, for the boundaries that carry flux; with
  // RK4 recording, once per substep with Dt = Cunew[t][1] Dt (step_ch.c:464-467)
  static const double cn1[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  for (int t = 0; t < (stages ? 4 : 1); t++) {
    const double Dts = stages ? cn1[t] * Dt : Dt;
    for (int dir = 0; dir < 6; dir++) {
      if (v.nface[dir] == 0) continue;
      for (int pl = 0; pl < NPR; pl++)
        c->pcum[dir][pl] += (stages ? hs[(t * 6 + dir) * NPR + pl] : out->pdot[dir][pl]) * Dts;
    }
  }
  for (int dir = 0; dir < 6; dir++)
    for (int pl = 0; pl < NPR; pl++) out->pcum[dir][pl] = c->pcum[dir][pl];
  return 0;
}
