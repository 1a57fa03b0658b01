// HIP kernels of the per-step GRMHD hot path for gfx950 (FP64 vector ALU).
//
//   k_bound_x1/x2/x3   ghost fill (bounds.fishmon.c:29)
//   k_flux_mc<dir>     MC reconstruction fused with the Riemann flux (flux.c:469)
//   k_flux_para<dir>   PARA (para4) reconstruction fused with the Riemann flux
//   k_flux_lr<dir>     Riemann flux from precomputed interface states (WENO5BND)
//   k_emf              flux-CT corner EMFs (fluxct.c:10)
//   k_update           conserved update + source + Noble 2D + fixup1zone (advance.c:58)
//   k_dt               time-step update (advance.c:231, step_ch.c:243)
//   k_set_dt           initial time step of the set-up (advance.c:957)
#include <hip/hip_runtime.h>

#include "harmgpu.h"
#include "hg_c2p.h"
#include "hg_c2p_io.h"
#include "hg_grid.h"
#include "hg_kernels.h"

namespace hg {

__device__ __forceinline__ void load_geom(const double *__restrict__ full, int loc, int64_t gplane, int64_t gi,
                                          Geom &G) {
  const double *b = full + (int64_t)loc * GCOMP * gplane + gi;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      G.gcov[a][c] = b[(a * 4 + c) * gplane];
      G.gcon[a][c] = b[(16 + a * 4 + c) * gplane];
    }
  G.gpert0 = b[32 * gplane];
  G.g = b[33 * gplane];
}

// Marks every geometry value as redefined.  The metric products that
// maxwell() forms depend on the geometry alone; without this barrier the
// compiler keeps the products of one state alive for the next state that
// uses the same geometry, which costs more registers than recomputing them.
__device__ __forceinline__ void geom_barrier(Geom &G) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      asm volatile("" : "+v"(G.gcov[a][c]));
      asm volatile("" : "+v"(G.gcon[a][c]));
    }
  asm volatile("" : "+v"(G.gpert0));
  asm volatile("" : "+v"(G.g));
}

// Pins n values at this point of the program: everything that computes
// them completes before, and nothing that reads them starts earlier.
__device__ __forceinline__ void value_barrier(double *v, int n) {
#pragma unroll
  for (int a = 0; a < n; a++) asm volatile("" : "+v"(v[a]));
}

// block-wide min of a positive double, folded into *out with an integer
// atomic (positive IEEE doubles order like their bit patterns)
__device__ __forceinline__ void block_min_to(double v, unsigned long long *out) {
  __shared__ double red[32];
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off, 64);
    v = (o < v) ? o : v;
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    double m = red[0];
    for (int w = 1; w < nw; w++) m = (red[w] < m) ? red[w] : m;
    atomicMin(out, (unsigned long long)__double_as_longlong(m));
  }
}

__device__ __forceinline__ void block_sum_to(unsigned long long a, unsigned long long b,
                                             unsigned long long *out_a, unsigned long long *out_b) {
  __shared__ unsigned long long ra[32], rb[32];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { ra[wid] = a; rb[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    unsigned long long sa = 0, sb = 0;
    for (int w = 0; w < nw; w++) { sa += ra[w]; sb += rb[w]; }
    if (sa) atomicAdd(out_a, sa);
    if (sb) atomicAdd(out_b, sb);
  }
}

// ---------------------------------------------------------------------------
// ghost fill, bounds.fishmon.c:29 (EXTRAP 0, POLEDEATH 1, POSDEFMETRIC 0)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bound_x1_body(double *__restrict__ p, const double *__restrict__ geo, const Dims &d,
                                              double gammamax, int *__restrict__ err, unsigned bid) {
  // threads: ((k * n2 + j) * 2 + side) * nb1 + g: the ghost layers of a row
  // on consecutive lanes (contiguous in memory, one source zone per side)
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = 2LL * d.nb1 * d.n2 * d.n3;
  if (t >= total) return;
  int g = (int)(t % d.nb1);
  int64_t r = t / d.nb1;
  int side = (int)(r % 2);
  r /= 2;
  int j = (int)(r % d.n2);
  int k = (int)(r / d.n2);
  int i = side ? d.n1 + g : -1 - g;
  int isrc = side ? d.n1 - 1 : 0;
  double pr[NPR];
  int64_t src = d.idx(isrc, j, k), dst = d.idx(i, j, k);
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pr[pl] = p[src + pl * d.spl];
  Geom G;
  load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
  if (inflow_check_r(pr, G, side ? +1 : -1, gammamax) > 0) atomicOr(err, 1);
  if (limit_gamma(gammamax, pr, G) > 0) atomicOr(err, 1);
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) p[dst + pl * d.spl] = pr[pl];
}

__global__ void k_bound_x1(double *__restrict__ p, const double *__restrict__ geo, Dims d, double gammamax,
                           int *__restrict__ err) {
  bound_x1_body(p, geo, d, gammamax, err, blockIdx.x);
}

__device__ __forceinline__ void bound_x2_body(double *__restrict__ p, const Dims &d, unsigned bid) {
  // threads: ((slot * n3 + k) * m1 + ii), slot in [0, 2*nb2)
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = 2LL * d.nb2 * d.n3 * d.m1;
  if (t >= total) return;
  int ii = t % d.m1;
  int64_t r = t / d.m1;
  int k = r % d.n3;
  int slot = (int)(r / d.n3);
  int i = ii - d.nb1;
  int j, jsrc, jact = -1;
  if (slot < d.nb2) {  // inner: j = -1-slot mirrors row slot
    j = -1 - slot;
    jsrc = slot;
    if (slot == 0) jact = 0;
  } else {
    int s = slot - d.nb2;  // outer: j = n2+s mirrors row n2-1-s
    j = d.n2 + s;
    jsrc = d.n2 - 1 - s;
    if (s == 0) jact = d.n2 - 1;
  }
  int64_t src = d.idx(i, jsrc, k), dst = d.idx(i, j, k);
  double pr[NPR];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pr[pl] = p[src + pl * d.spl];
  pr[U2] *= 0;
  pr[B2] *= 0.;
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) p[dst + pl * d.spl] = pr[pl];
  if (jact >= 0) {  // POLEDEATH also clears the first active row next to the axis
    int64_t a = d.idx(i, jact, k);
    p[a + U2 * d.spl] *= 0;
    p[a + B2 * d.spl] *= 0.;
  }
}

__global__ void k_bound_x2(double *__restrict__ p, Dims d) { bound_x2_body(p, d, blockIdx.x); }

__device__ __forceinline__ void bound_x3_body(double *__restrict__ p, const Dims &d, unsigned bid) {
  // threads: ((slot * m2 + jj) * m1 + ii), slot in [0, 2*nb3)
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = 2LL * d.nb3 * d.m2 * d.m1;
  if (t >= total) return;
  int ii = t % d.m1;
  int64_t r = t / d.m1;
  int jj = r % d.m2;
  int slot = (int)(r / d.m2);
  int i = ii - d.nb1, j = jj - d.nb2;
  int k, ksrc;
  if (slot < d.nb3) { k = -1 - slot; ksrc = d.n3 + k; }
  else { k = d.n3 + (slot - d.nb3); ksrc = k - d.n3; }
  int64_t src = d.idx(i, j, ksrc), dst = d.idx(i, j, k);
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) p[dst + pl * d.spl] = p[src + pl * d.spl];
}

__global__ void k_bound_x3(double *__restrict__ p, Dims d) { bound_x3_body(p, d, blockIdx.x); }

// ---------------------------------------------------------------------------
// Riemann flux at one face from left/right states, fluxcompute.c:12
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool face_flux(int dir, Geom &G, double *p_l, double *p_r, double gam,
                                          int fluxmethod, double *F, double &ctop) {
  // flux.c:629 BDIRCONT
  double bav = 0.5 * (p_l[B1 + dir - 1] + p_r[B1 + dir - 1]);
  p_l[B1 + dir - 1] = bav;
  p_r[B1 + dir - 1] = bav;
  // each side is finished (fluxes, conserved, wave speeds) before the other
  // starts, so only its results stay live
  double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
  double cmax_l, cmin_l, cmax_r, cmin_r;
  {
    State ql;
    if (get_state(p_l, G, ql)) return true;
    Maxwell M;
    maxwell(p_l, ql, G, M);
    double p1 = one_plus_ud0(ql, G);
    primtoflux(p_l, ql, M, p1, dir, G, gam, F_l);
    primtoflux(p_l, ql, M, p1, 0, G, gam, U_l);
    if (vchar(p_l, ql, dir, G, gam, cmax_l, cmin_l)) return true;
  }
  geom_barrier(G);
  {
    State qr;
    if (get_state(p_r, G, qr)) return true;
    Maxwell M;
    maxwell(p_r, qr, G, M);
    double p1 = one_plus_ud0(qr, G);
    primtoflux(p_r, qr, M, p1, dir, G, gam, F_r);
    primtoflux(p_r, qr, M, p1, 0, G, gam, U_r);
    if (vchar(p_r, qr, dir, G, gam, cmax_r, cmin_r)) return true;
  }
  // wavespeeds.c:325 cminmax_calc
  double lmin = rmin(cmin_l, cmin_r);
  double lmax = rmax(cmax_l, cmax_r);
  lmin = fabs(rmax(0., -lmin));
  lmax = fabs(rmax(0., lmax));
  ctop = rmax(lmax, lmin);
  if (fluxmethod == 1 && (lmax + lmin != 0.0)) {
#pragma unroll
    for (int pl = 0; pl < NPR; pl++)
      F[pl] = ((lmax * F_l[pl] + lmin * F_r[pl]) - (lmax * lmin) * (U_r[pl] - U_l[pl])) / (lmax + lmin);
  } else {
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
  }
  return false;
}

// ---------------------------------------------------------------------------
// the other flux methods of fluxcompute.c:100 flux_compute: FORCE, MUSTA and
// HLLLAXF1 (fluxmethod 2, 3, 4).  FORCE and MUSTA invert intermediate
// conserved states at the face, so they live apart from the fast path.
// ---------------------------------------------------------------------------
// fluxcompute.c:57 p2SFUevolve + vchar of one side
__device__ __forceinline__ bool side_flux(int dir, const Geom &G, const double *p, double gam, double *F,
                                          double *U, double &cmax, double &cmin) {
  State q;
  if (get_state(p, G, q)) return true;
  Maxwell M;
  maxwell(p, q, G, M);
  double p1 = one_plus_ud0(q, G);
  primtoflux(p, q, M, p1, dir, G, gam, F);
  if (U) primtoflux(p, q, M, p1, 0, G, gam, U);
  return vchar(p, q, dir, G, gam, cmax, cmin) != 0;
}

// fluxcompute.c:232 hllflux_compute (LaxF when cmax + cmin = 0)
__device__ __forceinline__ void hll_or_laxf(double cmin, double cmax, double ctop, const double *U_l,
                                            const double *U_r, const double *F_l, const double *F_r, double *F) {
  if (cmax + cmin != 0.0) {
    for (int pl = 0; pl < NPR; pl++)
      F[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
  } else {
    for (int pl = 0; pl < NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
  }
}

__device__ __noinline__ bool face_flux_ext(int dir, const Geom &G, double *p_l, double *p_r, double gam,
                                           int fluxmethod, double *F, double &ctop) {
  double bav = 0.5 * (p_l[B1 + dir - 1] + p_r[B1 + dir - 1]);  // flux.c:629 BDIRCONT
  p_l[B1 + dir - 1] = bav;
  p_r[B1 + dir - 1] = bav;
  double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
  double cmax_l, cmin_l, cmax_r, cmin_r;
  if (side_flux(dir, G, p_l, gam, F_l, U_l, cmax_l, cmin_l)) return true;
  if (side_flux(dir, G, p_r, gam, F_r, U_r, cmax_r, cmin_r)) return true;
  double cmin = fabs(rmax(0., -rmin(cmin_l, cmin_r)));
  double cmax = fabs(rmax(0., rmax(cmax_l, cmax_r)));
  ctop = rmax(cmax, cmin);
  int iters;
  if (fluxmethod == 4) {
    // fluxcompute.c:189 HLLLAXF1: HLL and LaxF blended by the jump in u
    double dPoP = fabs((p_r[UU] - p_l[UU]) / (0.5 * (p_r[UU] + p_l[UU])));
    double f_s = (1.0 - 0.0) / (0.3 - 0.0) * (dPoP - 0.0) + 0.0;  // MINDPOP 0, MAXDPOP 0.3
    if (f_s > 1.0) f_s = 1.0;
    if (f_s < 0.0) f_s = 0.0;
    if (cmax + cmin != 0.0) {
      for (int pl = 0; pl < NPR; pl++) {
        double hll = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
        double lax = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
        F[pl] = hll * f_s + lax * (1.0 - f_s);
      }
    } else {
      for (int pl = 0; pl < NPR; pl++) F[pl] = 0.5 * ((F_l[pl] + F_r[pl]) - ctop * (U_r[pl] - U_l[pl]));
    }
    return false;
  }
  if (fluxmethod == 2) {
    // fluxcompute.c:298 forceflux_compute with cforce = ctop: the two-step
    // Lax-Wendroff state inverted (guesses: the mean, then each side),
    // else HLL; GFORCE with MUSTACOEF 1 (fluxcompute.c:95)
    const double cforce = ctop;
    double umid[NPR], pmid[NPR];
    for (int pl = 0; pl < NPR; pl++) umid[pl] = 0.5 * ((U_l[pl] + U_r[pl]) - (F_r[pl] - F_l[pl]) / cforce);
    for (int pl = 0; pl < NPR; pl++) pmid[pl] = 0.5 * (p_l[pl] + p_r[pl]);
    int fl = utoprimgen_fp64(umid, G, gam, pmid, iters);
    if (fl) {
      for (int pl = 0; pl < NPR; pl++) pmid[pl] = p_l[pl];
      fl = utoprimgen_fp64(umid, G, gam, pmid, iters);
      if (fl) {
        for (int pl = 0; pl < NPR; pl++) pmid[pl] = p_r[pl];
        fl = utoprimgen_fp64(umid, G, gam, pmid, iters);
      }
    }
    if (fl) {
      hll_or_laxf(cmin, cmax, ctop, U_l, U_r, F_l, F_r, F);
      return false;
    }
    double fmid[NPR];
    {
      State q;
      if (get_state(pmid, G, q)) return true;
      Maxwell M;
      maxwell(pmid, q, G, M);
      double p1 = one_plus_ud0(q, G);
      primtoflux(pmid, q, M, p1, dir, G, gam, fmid);
    }
    for (int pl = 0; pl < NPR; pl++)
      F[pl] = ((fmid[pl]) + (1.0) * (0.5 * ((F_l[pl] + F_r[pl]) - cforce * (U_r[pl] - U_l[pl])))) / (1.0 + (1.0));
    return false;
  }
  // fluxcompute.c:398 mustaflux_compute -> :538 musta1flux_compute (MUSTAHLL,
  // one single-cell iteration, MUSTAVERSION 0, MUSTACOEF 1): HLL predictor,
  // both sides' U corrected by it and inverted (guess: that side's state),
  // then HLL of the corrected states; any failed inversion falls back to
  // HLL of the original states.  The energy-momentum fluxes are always the
  // original states' HLL (fluxcompute.c:527-536)
  double Fh[NPR];
  hll_or_laxf(cmin, cmax, ctop, U_l, U_r, F_l, F_r, Fh);
  double Ul2[NPR], Ur2[NPR], pl2[NPR], pr2[NPR], Fl2[NPR], Fr2[NPR];
  const double cmusta = ctop;
  for (int pl = 0; pl < NPR; pl++) {
    Ul2[pl] = U_l[pl] - (1.0) * (Fh[pl] - F_l[pl]) / cmusta;
    Ur2[pl] = U_r[pl] - (1.0) * (F_r[pl] - Fh[pl]) / cmusta;
    pl2[pl] = p_l[pl];
    pr2[pl] = p_r[pl];
  }
  bool musta = utoprimgen_fp64(Ul2, G, gam, pl2, iters) == 0;
  if (musta) musta = utoprimgen_fp64(Ur2, G, gam, pr2, iters) == 0;
  if (musta) {
    double cmax_l2, cmin_l2, cmax_r2, cmin_r2;
    if (side_flux(dir, G, pl2, gam, Fl2, nullptr, cmax_l2, cmin_l2)) return true;
    if (side_flux(dir, G, pr2, gam, Fr2, nullptr, cmax_r2, cmin_r2)) return true;
    double cmin2 = fabs(rmax(0., -rmin(cmin_l2, cmin_r2)));
    double cmax2 = fabs(rmax(0., rmax(cmax_l2, cmax_r2)));
    double ctop2 = rmax(cmax2, cmin2);
    hll_or_laxf(cmin2, cmax2, ctop2, Ul2, Ur2, Fl2, Fr2, F);
  } else {
    for (int pl = 0; pl < NPR; pl++) F[pl] = Fh[pl];
  }
  for (int pl = UU; pl <= U3; pl++) F[pl] = Fh[pl];
  return false;
}

// face range of interp_loop_set (initbase.c:1027-1054): faces 0..n along dir,
// one extra layer on each side transverse to it
struct FaceRange {
  int is, js, ks, ni, nj, nk;
};
__host__ __device__ inline FaceRange face_range(const Dims &d, int dir) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  FaceRange f;
  f.is = -s1; f.ni = d.n1 + 2 * s1;
  f.js = -s2; f.nj = d.n2 + 2 * s2;
  f.ks = -s3; f.nk = d.n3 + 2 * s3;
  if (dir == 1) { f.is = 0; f.ni = d.n1 + s1; }
  if (dir == 2) { f.js = 0; f.nj = d.n2 + s2; }
  if (dir == 3) { f.ks = 0; f.nk = d.n3 + s3; }
  return f;
}
// the finite-volume scheme's flux loops (initbase.c:949-993 with avgscheme
// DONOR): faces 0..n along dir, every ghost layer transverse to it
__host__ __device__ inline FaceRange face_range_fv(const Dims &d, int dir) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  FaceRange f;
  f.is = -d.nb1; f.ni = d.m1;
  f.js = -d.nb2; f.nj = d.m2;
  f.ks = -d.nb3; f.nk = d.m3;
  if (dir == 1) { f.is = 0; f.ni = d.n1 + s1; }
  if (dir == 2) { f.js = 0; f.nj = d.n2 + s2; }
  if (dir == 3) { f.ks = 0; f.nk = d.n3 + s3; }
  return f;
}

// global.h:361 MINMOD
__device__ __forceinline__ double minmod(double a, double b) {
  return ((a) * (b) <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? (a) : (b));
}
// interppoint.c:387 slope_lim_3points for the other point limiters: DONOR
// (0), VANL (1), MINM (2, theta 1); MC (3) is mc_slope below
__device__ __forceinline__ double slope3_other(int lim, double yl, double yc, double yr) {
  if (lim == 1) {
    double Dqm = (yc - yl);
    double Dqp = (yr - yc);
    double s = Dqm * Dqp;
    return (s <= 0.) ? 0. : (2.0 * s / (Dqm + Dqp));
  }
  if (lim == 2) {
    double Dqm = 1.0 * (yc - yl);
    double Dqp = 1.0 * (yr - yc);
    double Dqc = 0.5 * (yr - yl);
    double inner = ((Dqm) * (Dqc) <= 0) ? 0.0 : ((fabs(Dqm) < fabs(Dqc)) ? (Dqm) : (Dqc));
    return ((inner) * (Dqp) <= 0) ? 0.0 : ((fabs(inner) < fabs(Dqp)) ? (inner) : (Dqp));
  }
  return 0.0;
}

// interppoint.c:387 slope_lim_3points, MC
__device__ __forceinline__ double mc_slope(double yl, double yc, double yr) {
  double Dqm = 2.0 * (yc - yl);
  double Dqp = 2.0 * (yr - yc);
  double Dqc = 0.5 * (yr - yl);
  double inner = minmod(Dqm, Dqp);
  return minmod(Dqc, inner);
}

// phys.c:1797 max, phys.c:1804 min, phys.c:1778 sign: plain comparisons, so
// NaN and signed-zero operands resolve as in the reference
__device__ __forceinline__ double ref_max(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double ref_min(double a, double b) { return (a > b) ? b : a; }
__device__ __forceinline__ double ref_sign(double a) { return (a > 0.0) ? 1.0 : ((a < 0.0) ? -1.0 : 0.0); }

// interppoint.c:738 para4 with PARA2LIM MC (global.h:426): y[0..4] holds the
// zone's five-point stencil (offsets -2..2); returns the zone's left and
// right interface values
__device__ __forceinline__ void para4(const double y[5], double &lout, double &rout) {
  double dq[3];  // CW1.7 limited slopes at offsets -1, 0, 1
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double Dqm = 2.0 * (y[m + 1] - y[m]);
    double Dqp = 2.0 * (y[m + 2] - y[m + 1]);
    double Dqc = 0.5 * (y[m + 2] - y[m]);
    double s = Dqm * Dqp;
    dq[m] = (s <= 0.) ? 0. : ref_min(fabs(Dqc), ref_min(fabs(Dqm), fabs(Dqp))) * ref_sign(Dqc);
  }
  // CW1.6 with 1/8 in place of 1/6, bounded by the neighbouring values
  double l = 0.5 * (y[2] + y[1]) - (dq[1] - dq[0]) / 8.0;
  double r = 0.5 * (y[3] + y[2]) - (dq[2] - dq[1]) / 8.0;
  l = ref_max(ref_min(y[2], y[1]), l);
  l = ref_min(ref_max(y[2], y[1]), l);
  r = ref_max(ref_min(y[2], y[3]), r);
  r = ref_min(ref_max(y[2], y[3]), r);
  double qa = (r - y[2]) * (y[2] - l);
  double qd = (r - l);
  double qe = 6.0 * (y[2] - 0.5 * (l + r));
  if (qa <= 0.) {
    l = y[2];
    r = y[2];
  } else {
    // steepening at the top or bottom of a gradient; r sees the new l
    if (qd * (qd - qe) < 0.0) l = 3.0 * y[2] - 2.0 * r;
    if (qd * (qd + qe) < 0.0) r = 3.0 * y[2] - 2.0 * l;
  }
  lout = l;
  rout = r;
}

// point-type reconstruction fused with the Riemann flux (flux.c:469
// fluxcalc_standard): MC slopes (interppoint.c:387) or PARA interface values
// (interppoint.c:289), turned into the face states by flux.c:753 getp2interplr
template <int DIR, bool PARA, bool EXT = false>
__device__ __forceinline__ void flux_point_body(const double *__restrict__ p, const double *__restrict__ geo,
                                                double *__restrict__ F, const Dims &d, const FaceRange &fr,
                                                double gam, double cour, double dxdir, int fluxmethod, int slim,
                                                unsigned long long *__restrict__ ndt_bits, int *__restrict__ err) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)fr.ni * fr.nj * fr.nk;
  double dtij = 1.e+50;
  if (t < total) {
    // i fastest, then k: the blocks in flight share one j row of the
    // phi-independent geometry, which then stays in L2
    int i = fr.is + (int)(t % fr.ni);
    int64_t r = t / fr.ni;
    int k = fr.ks + (int)(r % fr.nk);
    int j = fr.js + (int)(r / fr.nk);
    int64_t z = d.idx(i, j, k);
    int64_t s = (DIR == 1) ? 1 : (DIR == 2) ? d.sj : d.sk;
    double p_l[NPR], p_r[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      const double *c = p + z + pl * d.spl;
      if (PARA) {
        // p_l = right value of zone z-s (stencil -3..1), p_r = left value
        // of zone z (stencil -2..2)
        double y[6] = {c[-3 * s], c[-2 * s], c[-s], c[0], c[s], c[2 * s]};
        double lm, rm, l0, r0;
        para4(y, lm, rm);
        para4(y + 1, l0, r0);
        p_l[pl] = rm;
        p_r[pl] = l0;
      } else {
        double ym2 = c[-2 * s], ym1 = c[-s], y0 = c[0], yp1 = c[s];
        if (slim == HG_LIM_MC) {
          p_l[pl] = ym1 + 0.5 * mc_slope(ym2, ym1, y0);
          p_r[pl] = y0 - 0.5 * mc_slope(ym1, y0, yp1);
        } else {
          p_l[pl] = ym1 + 0.5 * slope3_other(slim, ym2, ym1, y0);
          p_r[pl] = y0 - 0.5 * slope3_other(slim, ym1, y0, yp1);
        }
      }
    }
    Geom G;
    load_geom(geo, DIR, d.gplane, d.gidx(i, j), G);
    double Fz[NPR], ctop;
    if (EXT ? face_flux_ext(DIR, G, p_l, p_r, gam, fluxmethod, Fz, ctop)
            : face_flux(DIR, G, p_l, p_r, gam, fluxmethod, Fz, ctop)) {
      atomicOr(err, 2);
    } else {
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) F[z + pl * d.spl] = Fz[pl];
      dtij = cour * dxdir / ctop;
    }
  }
  block_min_to(dtij, ndt_bits);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_flux_mc(const double *__restrict__ p, const double *__restrict__ geo,
                                                 double *__restrict__ F, Dims d, double gam, double cour,
                                                 double dxdir, int fluxmethod, int slim,
                                                 unsigned long long *__restrict__ ndt_bits, int *__restrict__ err) {
  FaceRange fr = face_range(d, DIR);
  flux_point_body<DIR, false>(p, geo, F, d, fr, gam, cour, dxdir, fluxmethod, slim, ndt_bits, err);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_flux_para(const double *__restrict__ p, const double *__restrict__ geo,
                                                   double *__restrict__ F, Dims d, double gam, double cour,
                                                   double dxdir, int fluxmethod, int slim,
                                                   unsigned long long *__restrict__ ndt_bits,
                                                   int *__restrict__ err) {
  FaceRange fr = face_range(d, DIR);
  flux_point_body<DIR, true>(p, geo, F, d, fr, gam, cour, dxdir, fluxmethod, slim, ndt_bits, err);
}

// the finite-volume flux loops (every transverse ghost row)
template <int DIR, bool PARA>
__global__ void __launch_bounds__(256) k_flux_pt_fv(const double *__restrict__ p, const double *__restrict__ geo,
                                                    double *__restrict__ F, Dims d, double gam, double cour,
                                                    double dxdir, int fluxmethod, int slim,
                                                    unsigned long long *__restrict__ ndt_bits,
                                                    int *__restrict__ err) {
  FaceRange fr = face_range_fv(d, DIR);
  flux_point_body<DIR, PARA>(p, geo, F, d, fr, gam, cour, dxdir, fluxmethod, slim, ndt_bits, err);
}

// MC or PARA with the FORCE / MUSTA / HLLLAXF1 fluxes
template <int DIR, bool PARA>
__global__ void __launch_bounds__(256) k_flux_pt_ext(const double *__restrict__ p, const double *__restrict__ geo,
                                                     double *__restrict__ F, Dims d, int fv, double gam,
                                                     double cour, double dxdir, int fluxmethod, int slim,
                                                     unsigned long long *__restrict__ ndt_bits,
                                                     int *__restrict__ err) {
  FaceRange fr = fv ? face_range_fv(d, DIR) : face_range(d, DIR);
  flux_point_body<DIR, PARA, true>(p, geo, F, d, fr, gam, cour, dxdir, fluxmethod, slim, ndt_bits, err);
}

template <int DIR, bool EXT>
__device__ __forceinline__ void flux_lr_body(const double *__restrict__ pleft,
                                                 const double *__restrict__ pright,
                                                 const double *__restrict__ geo, double *__restrict__ F, Dims d,
                                                 FaceRange fr, double gam, double cour, double dxdir,
                                                 int fluxmethod, unsigned long long *__restrict__ ndt_bits,
                                                 int *__restrict__ err, unsigned bid) {
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)fr.ni * fr.nj * fr.nk;
  double dtij = 1.e+50;
  if (t < total) {
    // i fastest, then k: the blocks in flight share one j row of the
    // phi-independent geometry, which then stays in L2
    int i = fr.is + (int)(t % fr.ni);
    int64_t r = t / fr.ni;
    int k = fr.ks + (int)(r % fr.nk);
    int j = fr.js + (int)(r / fr.nk);
    int64_t z = d.idx(i, j, k);
    int64_t s = (DIR == 1) ? 1 : (DIR == 2) ? d.sj : d.sk;
    double p_l[NPR], p_r[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      p_l[pl] = pright[z - s + pl * d.spl];
      p_r[pl] = pleft[z + pl * d.spl];
    }
    Geom G;
    load_geom(geo, DIR, d.gplane, d.gidx(i, j), G);
    double Fz[NPR], ctop;
    if (EXT ? face_flux_ext(DIR, G, p_l, p_r, gam, fluxmethod, Fz, ctop)
            : face_flux(DIR, G, p_l, p_r, gam, fluxmethod, Fz, ctop)) {
      atomicOr(err, 2);
    } else {
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) F[z + pl * d.spl] = Fz[pl];
      dtij = cour * dxdir / ctop;
    }
  }
  block_min_to(dtij, ndt_bits);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_flux_lr(const double *__restrict__ pleft,
                                                 const double *__restrict__ pright,
                                                 const double *__restrict__ geo, double *__restrict__ F, Dims d,
                                                 double gam, double cour, double dxdir, int fluxmethod,
                                                 unsigned long long *__restrict__ ndt_bits, int *__restrict__ err) {
  FaceRange fr = face_range(d, DIR);
  flux_lr_body<DIR, false>(pleft, pright, geo, F, d, fr, gam, cour, dxdir, fluxmethod, ndt_bits, err, blockIdx.x);
}

// the finite-volume flux loops (every transverse ghost row)
template <int DIR>
__global__ void __launch_bounds__(256) k_flux_lr_fv(const double *__restrict__ pleft,
                                                    const double *__restrict__ pright,
                                                    const double *__restrict__ geo, double *__restrict__ F, Dims d,
                                                    double gam, double cour, double dxdir, int fluxmethod,
                                                    unsigned long long *__restrict__ ndt_bits,
                                                    int *__restrict__ err) {
  FaceRange fr = face_range_fv(d, DIR);
  flux_lr_body<DIR, false>(pleft, pright, geo, F, d, fr, gam, cour, dxdir, fluxmethod, ndt_bits, err, blockIdx.x);
}

// WENO5BND states with the FORCE / MUSTA / HLLLAXF1 fluxes
template <int DIR>
__global__ void __launch_bounds__(256) k_flux_lr_ext(const double *__restrict__ pleft,
                                                     const double *__restrict__ pright,
                                                     const double *__restrict__ geo, double *__restrict__ F, Dims d,
                                                     int fv, double gam, double cour, double dxdir,
                                                     int fluxmethod, unsigned long long *__restrict__ ndt_bits,
                                                     int *__restrict__ err) {
  FaceRange fr = fv ? face_range_fv(d, DIR) : face_range(d, DIR);
  flux_lr_body<DIR, true>(pleft, pright, geo, F, d, fr, gam, cour, dxdir, fluxmethod, ndt_bits, err, blockIdx.x);
}

// ---------------------------------------------------------------------------
// flux-CT corner EMFs, fluxct.c:74-218 (FLUXCTTOTH, CORNGDETVERSION 1):
// field fluxes divided by gdet at their faces, averaged to the corner,
// multiplied by gdet at the corner.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void emf_body(const double *__restrict__ F1, const double *__restrict__ F2,
                                         const double *__restrict__ F3, const double *__restrict__ geo,
                                         const double *__restrict__ corn, double *__restrict__ emf1,
                                         double *__restrict__ emf2, double *__restrict__ emf3, const Dims &d,
                                         unsigned bid) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  int ni = d.n1 + s1, nj = d.n2 + s2, nk = d.n3 + s3;
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)ni * nj * nk) return;
  int i = (int)(t % ni);
  int64_t r = t / ni;
  int k = (int)(r % nk);
  int j = (int)(r / nk);
  int64_t z = d.idx(i, j, k);
  int64_t zim = s1 ? z - 1 : z, zjm = s2 ? z - d.sj : z, zkm = s3 ? z - d.sk : z;
  const int64_t P = d.spl;
  int64_t g0 = d.gidx(i, j);
  int64_t gim = s1 ? d.gidx(i - 1, j) : g0, gjm = s2 ? d.gidx(i, j - 1) : g0;
  // gdet at FACE1/2/3 of (i,j) and of the neighbours the averages read
  const double *gF1 = geo + ((int64_t)GL_FACE1 * GCOMP + 33) * d.gplane;
  const double *gF2 = geo + ((int64_t)GL_FACE2 * GCOMP + 33) * d.gplane;
  const double *gF3 = geo + ((int64_t)GL_FACE3 * GCOMP + 33) * d.gplane;
  double c1 = (s2 && s3) ? 0.25 : 0.5, c2 = (s1 && s3) ? 0.25 : 0.5, c3 = (s1 && s2) ? 0.25 : 0.5;
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
  if (s2 && s3)
    e1 = c1 * (+F2[z + B3 * P] / gF2[g0] + F2[zkm + B3 * P] / gF2[g0] - F3[z + B2 * P] / gF3[g0] -
               F3[zjm + B2 * P] / gF3[gjm]);
  else if (s2)
    e1 = c1 * (+F2[z + B3 * P] / gF2[g0] + F2[zkm + B3 * P] / gF2[g0]);
  else if (s3)
    e1 = c1 * (-F3[z + B2 * P] / gF3[g0] - F3[zjm + B2 * P] / gF3[gjm]);
  if (s1 && s3)
    e2 = c2 * (+F3[z + B1 * P] / gF3[g0] + F3[zim + B1 * P] / gF3[gim] - F1[z + B3 * P] / gF1[g0] -
               F1[zkm + B3 * P] / gF1[g0]);
  else if (s3)
    e2 = c2 * (+F3[z + B1 * P] / gF3[g0] + F3[zim + B1 * P] / gF3[gim]);
  else if (s1)
    e2 = c2 * (-F1[z + B3 * P] / gF1[g0] - F1[zkm + B3 * P] / gF1[g0]);
  if (s1 && s2)
    e3 = c3 * (+F1[z + B2 * P] / gF1[g0] + F1[zjm + B2 * P] / gF1[gjm] - F2[z + B1 * P] / gF2[g0] -
               F2[zim + B1 * P] / gF2[gim]);
  else if (s1)
    e3 = c3 * (+F1[z + B2 * P] / gF1[g0] + F1[zjm + B2 * P] / gF1[gjm]);
  else if (s2)
    e3 = c3 * (-F2[z + B1 * P] / gF2[g0] - F2[zim + B1 * P] / gF2[gim]);
  emf1[z] = e1 * corn[0 * d.gplane + g0];
  emf2[z] = e2 * corn[1 * d.gplane + g0];
  emf3[z] = e3 * corn[2 * d.gplane + g0];
}

__global__ void k_emf(const double *__restrict__ F1, const double *__restrict__ F2, const double *__restrict__ F3,
                      const double *__restrict__ geo, const double *__restrict__ corn, double *__restrict__ emf1,
                      double *__restrict__ emf2, double *__restrict__ emf3, Dims d) {
  emf_body(F1, F2, F3, geo, corn, emf1, emf2, emf3, d, blockIdx.x);
}

// ---------------------------------------------------------------------------
// per-zone update, advance.c:143-213, in two kernels:
//   k_update: U(pi) + dt * (flux divergence + source) -> Uf
//   k_invert: Noble 2D inversion of Uf with guess pb, fixup1zone on the
//             final substep -> pf
// (one kernel holding both needs more registers than either, and the Uf
// round trip through HBM costs less than the occupancy it loses)
// ---------------------------------------------------------------------------
template <bool UCUM, bool FV>
__device__ __forceinline__ void update_body(const double *__restrict__ pi, const double *__restrict__ pb,
                                            double *__restrict__ ubuf, const double *__restrict__ F1,
                                            const double *__restrict__ F2, const double *__restrict__ F3,
                                            const double *__restrict__ emf1, const double *__restrict__ emf2,
                                            const double *__restrict__ emf3, const double *__restrict__ geo,
                                            const double *__restrict__ conn, const double *__restrict__ dtp,
                                            const RkStage &rk, const Dims &d, double gam, double idx1, double idx2,
                                            double idx3, int *__restrict__ err, unsigned bid) {
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)d.n1 * d.n2 * d.n3;
  if (t < total) {
    // i fastest, then k (geometry of one j row stays in L2)
    int i = (int)(t % d.n1);
    int64_t r = t / d.n1;
    int k = (int)(r % d.n3);
    int j = (int)(r / d.n3);
    int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
    int64_t z = d.idx(i, j, k);
    const int64_t P = d.spl;
    Geom G;
    load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
    double Ui[NPR], dUg[NPR], prb[NPR];
    bool bad = false;
    if (FV) {
      // advance.c:321-334: Ui is the conserved state carried from the last
      // step's ucum (or set from the primitives at the start)
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) Ui[pl] = rk.ucons[z + pl * P];
    } else {
      double pri[NPR];
#pragma unroll
      for (int pl = 0; pl < NPR; pl++) pri[pl] = pi[z + pl * P];
      bad = primtoU(pri, G, gam, Ui);
    }
    // the three phases (U of p_i, source of p_b, flux divergence) are kept
    // apart: interleaved, their live values exceed the register file
    value_barrier(Ui, NPR);
    geom_barrier(G);
    asm volatile("" ::: "memory");
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) prb[pl] = pb[z + pl * P];
    // phys.c:578 source -> phys.c:623 source_conn
    {
      State q;
      bad |= get_state(prb, G, q);
      Maxwell M;
      maxwell(prb, q, G, M);
      double p1 = one_plus_ud0(q, G);
      const double *cn = conn + d.cidx(i, j);
      double dUc[4] = {0., 0., 0., 0.};
      // dU_l = sum_jj sum_kk T^jj_kk Gamma^kk_{l jj}: one row of T at a time;
      // every dU_l still accumulates in (jj, kk) order
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        double mrow[4];
        mhd_row(prb, jj, q, M, p1, gam, mrow);
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
          for (int kk = 0; kk < 4; kk++) dUc[l] += mrow[kk] * cn[(kk * 16 + l * 4 + jj) * d.cplane];
      }
#pragma unroll
      for (int l = 0; l < 4; l++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) dUc[l] += -prb[RHO] * q.ucon[jj] * cn[(0 * 16 + l * 4 + jj) * d.cplane];
      dUg[RHO] = 0.0 + G.g * 0;
#pragma unroll
      for (int l = 0; l < 4; l++) dUg[UU + l] = 0.0 + G.g * dUc[l];
      dUg[B1] = 0.0 + G.g * 0;
      dUg[B2] = 0.0 + G.g * 0;
      dUg[B3] = 0.0 + G.g * 0;
    }
    value_barrier(dUg, NPR);
    geom_barrier(G);
    asm volatile("" ::: "memory");
    // advance.c:741 flux2dUavg, field components from the flux-CT EMFs (fluxct.c:720-781)
    double Uf[NPR];
    const double dt = dtp[0];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      double dU1 = 0, dU2 = 0, dU3 = 0;
      if (pl < B1) {
        if (s1) dU1 = (-(F1[z + 1 + pl * P] - F1[z + pl * P]) * idx1);
        if (s2) dU2 = (-(F2[z + d.sj + pl * P] - F2[z + pl * P]) * idx2);
        if (s3) dU3 = (-(F3[z + d.sk + pl * P] - F3[z + pl * P]) * idx3);
      } else {
        // F1[B1]=0, F1[B2]=0.5(e3+e3(j+1)), F1[B3]=-0.5(e2+e2(k+1)); likewise F2, F3
        int64_t jp = s2 ? d.sj : 0, kp = s3 ? d.sk : 0, ip = s1 ? 1 : 0;
        if (s1) {
          double a, b;
          if (pl == B1) { a = 0.; b = 0.; }
          else if (pl == B2) { a = 0.5 * (emf3[z + 1] + emf3[z + 1 + jp]); b = 0.5 * (emf3[z] + emf3[z + jp]); }
          else { a = -0.5 * (emf2[z + 1] + emf2[z + 1 + kp]); b = -0.5 * (emf2[z] + emf2[z + kp]); }
          dU1 = (-(a - b) * idx1);
        }
        if (s2) {
          double a, b;
          if (pl == B1) { a = -0.5 * (emf3[z + d.sj] + emf3[z + d.sj + ip]); b = -0.5 * (emf3[z] + emf3[z + ip]); }
          else if (pl == B2) { a = 0.; b = 0.; }
          else { a = 0.5 * (emf1[z + d.sj] + emf1[z + d.sj + kp]); b = 0.5 * (emf1[z] + emf1[z + kp]); }
          dU2 = (-(a - b) * idx2);
        }
        if (s3) {
          double a, b;
          if (pl == B1) { a = 0.5 * (emf2[z + d.sk] + emf2[z + d.sk + ip]); b = 0.5 * (emf2[z] + emf2[z + ip]); }
          else if (pl == B2) { a = -0.5 * (emf1[z + d.sk] + emf1[z + d.sk + jp]); b = -0.5 * (emf1[z] + emf1[z + jp]); }
          else { a = 0.; b = 0.; }
          dU3 = (-(a - b) * idx3);
        }
      }
      double dUr = dU1 + dU2 + dU3;
      // advance.c:922 dUtoU with CUf = {1, 0, cuf2}: uf carries no memory
      // across substeps (CUf[1] = 0 in both schemes)
      if (UCUM) {
        // general substep: CUf = {cuf0, cuf1, cuf2}; uf is the previous
        // substep's Uf (ubuf), zeroed on the first substep (advance.c:141)
        const double ufp = rk.first ? 0.0 : ubuf[z + pl * P];
        Uf[pl] = rk.cuf0 * Ui[pl] + rk.cuf1 * ufp + rk.cuf2 * dt * (dUr + dUg[pl]);
      } else {
        Uf[pl] = 1.0 * Ui[pl] + 0.0 * 0.0 + rk.cuf2 * dt * (dUr + dUg[pl]);
      }
      if (UCUM) {
        // ucum, zeroed on the first substep (advance.c:141), collects
        // Cunew = {cn0, cn1, cn2}; the last substep inverts ucum, which the
        // finite-volume scheme also carries into the next step
        double uc = rk.first ? 0.0 : rk.ucum[z + pl * P];
        uc += rk.cn0 * Ui[pl] + rk.cn1 * dt * (dUr + dUg[pl]) + rk.cn2 * Uf[pl];
        if (rk.last) {
          Uf[pl] = uc;
          if (FV) rk.ucons[z + pl * P] = uc;
        } else {
          rk.ucum[z + pl * P] = uc;
        }
      }
      // midpoint scheme: ucum equals Uf on the final substep (Cunew = {0,0,1})
    }
    if (bad) atomicOr(err, 4);
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) ubuf[z + pl * P] = Uf[pl];
  }
}

template <bool UCUM, bool FV>
__global__ void __launch_bounds__(256) k_update(const double *__restrict__ pi, const double *__restrict__ pb,
                                                double *__restrict__ ubuf, const double *__restrict__ F1,
                                                const double *__restrict__ F2, const double *__restrict__ F3,
                                                const double *__restrict__ emf1, const double *__restrict__ emf2,
                                                const double *__restrict__ emf3, const double *__restrict__ geo,
                                                const double *__restrict__ conn, const double *__restrict__ dtp,
                                                RkStage rk, Dims d, double gam, double idx1, double idx2,
                                                double idx3, int *__restrict__ err) {
  update_body<UCUM, FV>(pi, pb, ubuf, F1, F2, F3, emf1, emf2, emf3, geo, conn, dtp, rk, d, gam, idx1, idx2, idx3, err,
                        blockIdx.x);
}

// The inversion of one zone in the step's mode: the reference's FP64
// iteration (INV_FP64), or the mixed-precision path of the batched inversion
// (FP32 Newton from the plain start, or from the learned start for
// INV_LEARNED, then the polish to the converged root and double-double
// primitives) with the converged-root reference inversion (HG_C2P_REFINED)
// from the plain guess for any zone that path does not settle
constexpr int INV_FP64 = 0, INV_MIXED = 1, INV_LEARNED = 2;
template <int MODE>
__device__ __forceinline__ int invert_zone(const double *Uf, const Geom &G, double gam, double *pr, float ylearn,
                                           int &iters) {
  if (MODE == INV_FP64) return utoprimgen_fp64(Uf, G, gam, pr, iters);
  {
    double Ug[NPR], p2[NPR], W_last, alpha;
#pragma unroll
    for (int i = 0; i < NPR; i++) p2[i] = pr[i];
    for (int i = B1; i <= B3; i++) p2[i] = Uf[i] / G.g;  // as the set-up's Ug
    C2PInv c;
    int code = c2p_setup_zone(Uf, G, gam, p2, c, W_last, alpha, Ug);
    if (code == 0) {
      const double igr = (gam - 1.0) / gam;
      double x0, x1;
      int it = 0, flag;
      bool used;
      if (MODE == INV_LEARNED)
        it = learned_start(c, igr, W_last, ylearn, x0, x1, used);
      else
        newton_start(c, W_last, x0, x1);
      if (mixed_zone(c, igr, x0, x1, p2, it, flag)) {
#pragma unroll
        for (int i = 0; i < NPR; i++) pr[i] = p2[i];
        iters = it;
        return flag;
      }
    }
  }
  bool pol;
  return utoprimgen_refined(Uf, G, gam, pr, iters, pol);
}

template <bool FV, int MODE>
__device__ __forceinline__ void invert_body(const double *__restrict__ ubuf, const double *__restrict__ pb,
                                            double *__restrict__ pf, const double *__restrict__ geo,
                                            const double *__restrict__ floors, int finalstep, const Dims &d,
                                            double gam, double gammamax, unsigned long long *__restrict__ counters,
                                            int8_t *__restrict__ pflag, double *__restrict__ ucons,
                                            const float *__restrict__ ylearn, unsigned *__restrict__ fx_reset,
                                            int *__restrict__ err, unsigned bid) {
  // the post_fixup lists of this substep start empty (launch_post_fixup
  // then skips its own reset)
  if (fx_reset && bid == 0 && threadIdx.x < 3) fx_reset[threadIdx.x] = 0u;
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)d.n1 * d.n2 * d.n3;
  unsigned long long nfail = 0, nit = 0;
  if (t < total) {
    int i = (int)(t % d.n1);
    int64_t r = t / d.n1;
    int k = (int)(r % d.n3);
    int j = (int)(r / d.n3);
    int64_t z = d.idx(i, j, k);
    const int64_t P = d.spl;
    Geom G;
    load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
    double Uf[NPR], pr[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      Uf[pl] = ubuf[z + pl * P];
      pr[pl] = pb[z + pl * P];
    }
    bool bad = false;
    int iters = 0;
    int flag = invert_zone<MODE>(Uf, G, gam, pr, MODE == INV_LEARNED ? ylearn[t] : 0.f, iters);
    pflag[z] = (int8_t)flag;
    nit = (unsigned long long)iters;
    nfail = flag ? 1ull : 0ull;
    if (finalstep) {
      // fixup.c:341 fixup1zone with set_density_floors rescaletype 0
      double pr0[NPR];
      if (FV)
#pragma unroll
        for (int pl = 0; pl < NPR; pl++) pr0[pl] = pr[pl];
      bool changed = false;
      double f0 = floors[d.cidx(i, j)], f1 = floors[d.cplane + d.cidx(i, j)];
      if (f0 < 1.e-20) f0 = 1.e-20;
      if (f1 < 1.e-20) f1 = 1.e-20;
      if (f0 > pr[RHO]) { pr[RHO] = f0; changed = true; }
      if (f1 > pr[UU]) { pr[UU] = f1; changed = true; }
      int lg = limit_gamma(gammamax, pr, G);
      if (lg > 0) bad = true;
      if (lg == -1) changed = true;
      if (FV && changed) {
        // fixup.c:443 diag_fixup -> fixup.c:240-248: the carried conserved
        // state takes the change the floors made, each U passed through
        // UtoU (UDIAG -> UEVOLVE: divided, then multiplied by gdet), once
        // per energy region (three with DOJETDIAG, global.h:436)
        double U0[NPR], U1[NPR];
        primtoU(pr0, G, gam, U0);
        primtoU(pr, G, gam, U1);
#pragma unroll
        for (int pl = 0; pl < NPR; pl++) {
          const double du = (U1[pl] / G.g) * G.g - (U0[pl] / G.g) * G.g;
          double u = ucons[z + pl * P];
          for (int region = 0; region < NENERREGIONS; region++) u = u + du;
          ucons[z + pl * P] = u;
        }
      }
    }
    if (bad) atomicOr(err, 4);
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) pf[z + pl * P] = pr[pl];
  }
  block_sum_to(nfail, nit, &counters[0], &counters[1]);
}

template <bool FV, int MODE>
__global__ void __launch_bounds__(256) k_invert(const double *__restrict__ ubuf, const double *__restrict__ pb,
                                                double *__restrict__ pf, const double *__restrict__ geo,
                                                const double *__restrict__ floors, int finalstep, Dims d,
                                                double gam, double gammamax,
                                                unsigned long long *__restrict__ counters,
                                                int8_t *__restrict__ pflag, double *__restrict__ ucons,
                                                const float *__restrict__ ylearn, unsigned *__restrict__ fx_reset,
                                                int *__restrict__ err) {
  invert_body<FV, MODE>(ubuf, pb, pf, geo, floors, finalstep, d, gam, gammamax, counters, pflag, ucons, ylearn,
                        fx_reset, err, blockIdx.x);
}

// advance.c:231 ndt = defcon / (1/ndt1 + 1/ndt2 + 1/ndt3); on the final
// substep step_ch.c:243-247 limits the growth and stores the next dt
__device__ __forceinline__ void dt_body(unsigned long long *ndt_bits, double *dtp, double defcon, double safe) {
  double n1 = __longlong_as_double((long long)ndt_bits[0]);
  double n2 = __longlong_as_double((long long)ndt_bits[1]);
  double n3 = __longlong_as_double((long long)ndt_bits[2]);
  double ndt = defcon * 1. / (1. / n1 + 1. / n2 + 1. / n3);
  double dt = dtp[0];
  if (ndt > safe * dt) ndt = safe * dt;
  dtp[0] = ndt;
}

__global__ void k_dt(unsigned long long *ndt_bits, double *dtp, double defcon, double safe) {
  dt_body(ndt_bits, dtp, defcon, safe);
}

// ---------------------------------------------------------------------------
// grouped launches of the kernels above (GridView, group_of)
// ---------------------------------------------------------------------------
#define HG_GROUP_PRELUDE                               \
  const int g_ = group_of(first, n, blockIdx.x);       \
  const GridView &v = views[g_];                       \
  const unsigned bid = blockIdx.x - first[g_];

__global__ void __launch_bounds__(128) kg_bound_x1(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  bound_x1_body(v.pf, v.geo, v.d, v.gammamax, v.err, bid);
}

__global__ void __launch_bounds__(256) kg_bound_x2(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  bound_x2_body(v.pf, v.d, bid);
}

__global__ void __launch_bounds__(256) kg_bound_x3(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  bound_x3_body(v.pf, v.d, bid);
}

template <int DIR>
__global__ void __launch_bounds__(256) kg_flux_lr(const GridView *__restrict__ views,
                                                  const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  flux_lr_body<DIR, false>(v.pl[DIR - 1], v.pr[DIR - 1], v.geo, v.F[DIR - 1], v.d, face_range(v.d, DIR), v.gam,
                           v.cour, v.dx[DIR], v.fluxmethod, v.ndt_bits + DIR - 1, v.err, bid);
}

__global__ void __launch_bounds__(256) kg_emf(const GridView *__restrict__ views, const unsigned *__restrict__ first,
                                              int n) {
  HG_GROUP_PRELUDE
  emf_body(v.F[0], v.F[1], v.F[2], v.geo, v.corn, v.emf[0], v.emf[1], v.emf[2], v.d, bid);
}

template <bool UCUM>
__global__ void __launch_bounds__(256) kg_update(const GridView *__restrict__ views,
                                                 const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  update_body<UCUM, false>(v.pi, v.pb, v.uf, v.F[0], v.F[1], v.F[2], v.emf[0], v.emf[1], v.emf[2], v.geo, v.conn,
                            v.dt, v.rk, v.d, v.gam, v.idx[1], v.idx[2], v.idx[3], v.err, bid);
}

template <int MODE>
__global__ void __launch_bounds__(256) kg_invert(const GridView *__restrict__ views,
                                                 const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  invert_body<false, MODE>(v.uf, v.pb, v.pf, v.geo, v.floors, v.finalstep, v.d, v.gam, v.gammamax, v.counters,
                           v.pflag, nullptr, MODE == INV_LEARNED ? v.ly : nullptr, v.fx.count, v.err, bid);
}

__global__ void kg_dt(const GridView *__restrict__ views, const unsigned *__restrict__ first, int n) {
  HG_GROUP_PRELUDE
  (void)bid;
  if (threadIdx.x == 0) dt_body(v.ndt_bits, v.dt, v.defcon, v.safe);
}
#undef HG_GROUP_PRELUDE

__global__ void k_reset_ndt(unsigned long long *ndt_bits) {
  if (threadIdx.x < 3) ndt_bits[threadIdx.x] = (unsigned long long)__double_as_longlong(1.e+50);
}

// ---------------------------------------------------------------------------
// the inversions of a substep as a batch (hg_dump_inversions): for every
// active zone t = i + n1 (j + n2 k), the conserved state it inverted, the
// zone-centre metric and the guess, in hg_utoprim_batch's layout
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gather_inversions(const double *__restrict__ ubuf,
                                                           const double *__restrict__ pb,
                                                           const double *__restrict__ geo, Dims d,
                                                           double *__restrict__ U, double *__restrict__ gcov,
                                                           double *__restrict__ gcon, double *__restrict__ gdet,
                                                           double *__restrict__ gcovpert,
                                                           double *__restrict__ guess) {
  const int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % d.n1);
  const int64_t r = t / d.n1;
  const int j = (int)(r % d.n2), k = (int)(r / d.n2);
  const int64_t z = d.idx(i, j, k);
  const int64_t P = d.spl;
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) {
    U[pl * n + t] = ubuf[z + pl * P];
    guess[pl * n + t] = pb[z + pl * P];
  }
  const double *b = geo + (int64_t)GL_CENT * GCOMP * d.gplane + d.gidx(i, j);
#pragma unroll
  for (int c = 0; c < 16; c++) {
    gcov[c * n + t] = b[c * d.gplane];
    gcon[c * n + t] = b[(16 + c) * d.gplane];
  }
  gcovpert[t] = b[32 * d.gplane];
  gdet[t] = b[33 * d.gplane];
}

void launch_gather_inversions(const double *ubuf, const double *pb, const double *geo, const Dims &d, double *U,
                              double *gcov, double *gcon, double *gdet, double *gcovpert, double *guess,
                              hipStream_t s) {
  const int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  hipLaunchKernelGGL(k_gather_inversions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ubuf, pb, geo, d, U,
                     gcov, gcon, gdet, gcovpert, guess);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_bound(double *p, const double *geo, const Dims &d, double gammamax, int *err, hipStream_t s) {
  if (d.n1 > 1) {
    int64_t n = 2LL * d.nb1 * d.n2 * d.n3;
    hipLaunchKernelGGL(k_bound_x1, dim3(nblocks(n, 128)), dim3(128), 0, s, p, geo, d, gammamax, err);
  }
  if (d.n2 > 1) {
    int64_t n = 2LL * d.nb2 * d.n3 * d.m1;
    hipLaunchKernelGGL(k_bound_x2, dim3(nblocks(n, 256)), dim3(256), 0, s, p, d);
  }
  if (d.n3 > 1) {
    int64_t n = 2LL * d.nb3 * d.m2 * d.m1;
    hipLaunchKernelGGL(k_bound_x3, dim3(nblocks(n, 256)), dim3(256), 0, s, p, d);
  }
}

void launch_flux_point(int dir, int lim, bool fv, const double *p, const double *geo, double *F, const Dims &d,
                       double gam, double cour, double dxdir, int fluxmethod, unsigned long long *ndt_bits,
                       int *err, const Tune &tu, hipStream_t s) {
  FaceRange fr = fv ? face_range_fv(d, dir) : face_range(d, dir);
  int64_t n = (int64_t)fr.ni * fr.nj * fr.nk;
  dim3 g(nblocks(n, tu.bs_flux)), b(tu.bs_flux);
  const bool para = lim == HG_LIM_PARA;
  if (fluxmethod >= 2) {
    if (para) {
      if (dir == 1) hipLaunchKernelGGL((k_flux_pt_ext<1, true>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 2) hipLaunchKernelGGL((k_flux_pt_ext<2, true>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 3) hipLaunchKernelGGL((k_flux_pt_ext<3, true>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    } else {
      if (dir == 1) hipLaunchKernelGGL((k_flux_pt_ext<1, false>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 2) hipLaunchKernelGGL((k_flux_pt_ext<2, false>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 3) hipLaunchKernelGGL((k_flux_pt_ext<3, false>), g, b, 0, s, p, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    }
    return;
  }
  if (fv) {
    if (para) {
      if (dir == 1) hipLaunchKernelGGL((k_flux_pt_fv<1, true>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 2) hipLaunchKernelGGL((k_flux_pt_fv<2, true>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 3) hipLaunchKernelGGL((k_flux_pt_fv<3, true>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    } else {
      if (dir == 1) hipLaunchKernelGGL((k_flux_pt_fv<1, false>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 2) hipLaunchKernelGGL((k_flux_pt_fv<2, false>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
      if (dir == 3) hipLaunchKernelGGL((k_flux_pt_fv<3, false>), g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    }
    return;
  }
  if (para) {
    if (dir == 1) hipLaunchKernelGGL(k_flux_para<1>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    if (dir == 2) hipLaunchKernelGGL(k_flux_para<2>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    if (dir == 3) hipLaunchKernelGGL(k_flux_para<3>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
    return;
  }
  if (dir == 1) hipLaunchKernelGGL(k_flux_mc<1>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
  if (dir == 2) hipLaunchKernelGGL(k_flux_mc<2>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
  if (dir == 3) hipLaunchKernelGGL(k_flux_mc<3>, g, b, 0, s, p, geo, F, d, gam, cour, dxdir, fluxmethod, lim, ndt_bits, err);
}

// initbase.c:18 init_conservatives (primtoU of every active zone) into the
// carried conserved state of the finite-volume scheme
__global__ void __launch_bounds__(256) k_cons_from_prims(const double *__restrict__ p, const double *__restrict__ geo,
                                                         double *__restrict__ U, Dims d, double gam,
                                                         int *__restrict__ err) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)d.n1 * d.n2 * d.n3;
  if (t >= total) return;
  int i = (int)(t % d.n1);
  int64_t r = t / d.n1;
  int k = (int)(r % d.n3);
  int j = (int)(r / d.n3);
  int64_t z = d.idx(i, j, k);
  Geom G;
  load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
  double pr[NPR], Uz[NPR];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pr[pl] = p[z + pl * d.spl];
  if (primtoU(pr, G, gam, Uz)) atomicOr(err, 4);
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) U[z + pl * d.spl] = Uz[pl];
}

void launch_cons_from_prims(const double *p, const double *geo, double *U, const Dims &d, double gam, int *err,
                            hipStream_t s) {
  int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  hipLaunchKernelGGL(k_cons_from_prims, dim3(nblocks(n, 256)), dim3(256), 0, s, p, geo, U, d, gam, err);
}

void launch_flux_lr(int dir, const double *pleft, const double *pright, const double *geo, double *F,
                    const Dims &d, double gam, double cour, double dxdir, int fluxmethod,
                    unsigned long long *ndt_bits, int *err, const Tune &tu, bool fv, hipStream_t s) {
  FaceRange fr = fv ? face_range_fv(d, dir) : face_range(d, dir);
  int64_t n = (int64_t)fr.ni * fr.nj * fr.nk;
  dim3 g(nblocks(n, tu.bs_flux)), b(tu.bs_flux);
  if (fluxmethod >= 2) {
    if (dir == 1) hipLaunchKernelGGL(k_flux_lr_ext<1>, g, b, 0, s, pleft, pright, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    if (dir == 2) hipLaunchKernelGGL(k_flux_lr_ext<2>, g, b, 0, s, pleft, pright, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    if (dir == 3) hipLaunchKernelGGL(k_flux_lr_ext<3>, g, b, 0, s, pleft, pright, geo, F, d, (int)fv, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    return;
  }
  if (fv) {
    if (dir == 1) hipLaunchKernelGGL(k_flux_lr_fv<1>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    if (dir == 2) hipLaunchKernelGGL(k_flux_lr_fv<2>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    if (dir == 3) hipLaunchKernelGGL(k_flux_lr_fv<3>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
    return;
  }
  if (dir == 1) hipLaunchKernelGGL(k_flux_lr<1>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
  if (dir == 2) hipLaunchKernelGGL(k_flux_lr<2>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
  if (dir == 3) hipLaunchKernelGGL(k_flux_lr<3>, g, b, 0, s, pleft, pright, geo, F, d, gam, cour, dxdir, fluxmethod, ndt_bits, err);
}

void launch_emf(const double *F1, const double *F2, const double *F3, const double *geo, const double *corn,
                double *e1, double *e2, double *e3, const Dims &d, const Tune &tu, hipStream_t s) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  int64_t n = (int64_t)(d.n1 + s1) * (d.n2 + s2) * (d.n3 + s3);
  hipLaunchKernelGGL(k_emf, dim3(nblocks(n, tu.bs_emf)), dim3(tu.bs_emf), 0, s, F1, F2, F3, geo, corn, e1, e2, e3, d);
}

// fluxct.c:720-781: the field fluxes rebuilt from the corner EMFs, in place
// (F_i[B_i] = 0, the transverse components the averages of two EMFs), over
// the faces i = 0..n1, j = 0..n2, k = 0..n3 of the dimensions present
__global__ void k_fluxct_fields(double *__restrict__ F1, double *__restrict__ F2, double *__restrict__ F3,
                                const double *__restrict__ e1, const double *__restrict__ e2,
                                const double *__restrict__ e3, Dims d) {
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const int ni = d.n1 + s1, nj = d.n2 + s2, nk = d.n3 + s3;
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)ni * nj * nk) return;
  const int i = (int)(t % ni);
  const int64_t r = t / ni;
  const int k = (int)(r % nk), j = (int)(r / nk);
  const int64_t z = d.idx(i, j, k), P = d.spl;
  const int64_t zip = z + s1, zjp = z + (s2 ? d.sj : 0), zkp = z + (s3 ? d.sk : 0);
  if (s1 && j < d.n2 && k < d.n3) {
    F1[z + B1 * P] = 0.;
    F1[z + B2 * P] = 0.5 * (e3[z] + e3[zjp]);
    F1[z + B3 * P] = -0.5 * (e2[z] + e2[zkp]);
  }
  if (s2 && i < d.n1 && k < d.n3) {
    F2[z + B1 * P] = -0.5 * (e3[z] + e3[zip]);
    F2[z + B2 * P] = 0.;
    F2[z + B3 * P] = 0.5 * (e1[z] + e1[zkp]);
  }
  if (s3 && i < d.n1 && j < d.n2) {
    F3[z + B1 * P] = 0.5 * (e2[z] + e2[zip]);
    F3[z + B2 * P] = -0.5 * (e1[z] + e1[zjp]);
    F3[z + B3 * P] = 0.;
  }
}

void launch_fluxct_fields(double *F1, double *F2, double *F3, const double *e1, const double *e2, const double *e3,
                          const Dims &d, hipStream_t s) {
  int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  int64_t n = (int64_t)(d.n1 + s1) * (d.n2 + s2) * (d.n3 + s3);
  hipLaunchKernelGGL(k_fluxct_fields, dim3(nblocks(n, 256)), dim3(256), 0, s, F1, F2, F3, e1, e2, e3, d);
}

void launch_update(const double *pi, const double *pb, double *ubuf, const double *F1, const double *F2,
                   const double *F3, const double *e1, const double *e2, const double *e3, const double *geo,
                   const double *conn, const double *dtp, const RkStage &rk, const Dims &d, double gam,
                   const double *dx, int *err, const Tune &tu, hipStream_t s) {
  int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  dim3 g(nblocks(n, tu.bs_update)), b(tu.bs_update);
  const double i1 = 1.0 / dx[1], i2 = 1.0 / dx[2], i3 = 1.0 / dx[3];
  if (rk.ucons)
    hipLaunchKernelGGL((k_update<true, true>), g, b, 0, s, pi, pb, ubuf, F1, F2, F3, e1, e2, e3, geo, conn, dtp, rk, d,
                       gam, i1, i2, i3, err);
  else if (rk.ucum)
    hipLaunchKernelGGL((k_update<true, false>), g, b, 0, s, pi, pb, ubuf, F1, F2, F3, e1, e2, e3, geo, conn, dtp, rk,
                       d, gam, i1, i2, i3, err);
  else
    hipLaunchKernelGGL((k_update<false, false>), g, b, 0, s, pi, pb, ubuf, F1, F2, F3, e1, e2, e3, geo, conn, dtp, rk,
                       d, gam, i1, i2, i3, err);
}

template <int MODE>
static void launch_invert_mode(const double *ubuf, const double *pb, double *pf, const double *geo,
                               const double *floors, const RkStage &rk, int finalstep, const Dims &d, double gam,
                               double gammamax, unsigned long long *counters, int8_t *pflag, const float *ylearn,
                               unsigned *fx_reset, int *err, const Tune &tu, hipStream_t s) {
  int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  dim3 gi(nblocks(n, tu.bs_invert)), bi(tu.bs_invert);
  if (rk.ucons && finalstep)
    hipLaunchKernelGGL((k_invert<true, MODE>), gi, bi, 0, s, ubuf, pb, pf, geo, floors, finalstep, d, gam, gammamax,
                       counters, pflag, rk.ucons, ylearn, fx_reset, err);
  else
    hipLaunchKernelGGL((k_invert<false, MODE>), gi, bi, 0, s, ubuf, pb, pf, geo, floors, finalstep, d, gam, gammamax,
                       counters, pflag, nullptr, ylearn, fx_reset, err);
}

void launch_invert(const double *ubuf, const double *pb, double *pf, const double *geo, const double *floors,
                   const RkStage &rk, int finalstep, const Dims &d, double gam, double gammamax,
                   unsigned long long *counters, int8_t *pflag, int mode, const float *ylearn, unsigned *fx_reset,
                   int *err, const Tune &tu, hipStream_t s) {
  if (mode == INV_MIXED)
    launch_invert_mode<INV_MIXED>(ubuf, pb, pf, geo, floors, rk, finalstep, d, gam, gammamax, counters, pflag,
                                  ylearn, fx_reset, err, tu, s);
  else if (mode == INV_LEARNED)
    launch_invert_mode<INV_LEARNED>(ubuf, pb, pf, geo, floors, rk, finalstep, d, gam, gammamax, counters, pflag,
                                    ylearn, fx_reset, err, tu, s);
  else
    launch_invert_mode<INV_FP64>(ubuf, pb, pf, geo, floors, rk, finalstep, d, gam, gammamax, counters, pflag,
                                 ylearn, fx_reset, err, tu, s);
}

// the learned start's features of the inversions k_invert will do (its zone
// order; the guess pb with the field of ubuf, as the set-up sees it)
__device__ __forceinline__ void step_features_body(const double *__restrict__ ubuf, const double *__restrict__ pb,
                                                   const double *__restrict__ geo, const Dims &d, double gam,
                                                   float *__restrict__ X, unsigned bid) {
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)d.n1 * d.n2 * d.n3) return;
  int i = (int)(t % d.n1);
  int64_t r = t / d.n1;
  int k = (int)(r % d.n3);
  int j = (int)(r / d.n3);
  int64_t z = d.idx(i, j, k);
  const int64_t P = d.spl;
  Geom G;
  load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
  double Uf[NPR], pr[NPR], Ug[NPR], W_last = 1., alpha = 1.;
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) {
    Uf[pl] = ubuf[z + pl * P];
    pr[pl] = pb[z + pl * P];
  }
  for (int pl = B1; pl <= B3; pl++) pr[pl] = Uf[pl] / G.g;
  C2PInv c;
  float x[C2P_NFEAT];
  if (c2p_setup_zone(Uf, G, gam, pr, c, W_last, alpha, Ug) == 0) {
    c2p_features(c, W_last, alpha, gam, x);
  } else {
#pragma unroll
    for (int q = 0; q < C2P_NFEAT; q++) x[q] = 0.f;
  }
#pragma unroll
  for (int q = 0; q < C2P_NFEAT; q++) X[t * C2P_NFEAT + q] = x[q];
}

__global__ void __launch_bounds__(256) k_step_features(const double *__restrict__ ubuf,
                                                       const double *__restrict__ pb,
                                                       const double *__restrict__ geo, Dims d, double gam,
                                                       float *__restrict__ X) {
  step_features_body(ubuf, pb, geo, d, gam, X, blockIdx.x);
}

__global__ void __launch_bounds__(256) kg_step_features(const GridView *__restrict__ views,
                                                        const unsigned *__restrict__ first, int n) {
  const int g_ = group_of(first, n, blockIdx.x);
  const GridView &v = views[g_];
  step_features_body(v.uf, v.pb, v.geo, v.d, v.gam, v.lx, blockIdx.x - first[g_]);
}

void launch_step_features(const double *ubuf, const double *pb, const double *geo, const Dims &d, double gam,
                          float *X, hipStream_t s) {
  int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  hipLaunchKernelGGL(k_step_features, dim3(nblocks(n, 256)), dim3(256), 0, s, ubuf, pb, geo, d, gam, X);
}

void launch_dt(unsigned long long *ndt_bits, double *dtp, double defcon, double safe, hipStream_t s) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, s, ndt_bits, dtp, defcon, safe);
}

void launch_reset_ndt(unsigned long long *ndt_bits, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_ndt, dim3(1), dim3(64), 0, s, ndt_bits);
}

// grouped launches: the block counts of the single-grid launchers above at
// their default block sizes
unsigned group_blocks(int k, const GridView &v) {
  const Dims &d = v.d;
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const int64_t act = (int64_t)d.n1 * d.n2 * d.n3;
  switch (k) {
    case GK_BOUND1: return s1 ? nblocks(2LL * d.nb1 * d.n2 * d.n3, 128) : 0u;
    case GK_BOUND2: return s2 ? nblocks(2LL * d.nb2 * d.n3 * d.m1, 256) : 0u;
    case GK_BOUND3: return s3 ? nblocks(2LL * d.nb3 * d.m2 * d.m1, 256) : 0u;
    case GK_FLUX1: case GK_FLUX2: case GK_FLUX3: {
      const int dir = 1 + k - GK_FLUX1;
      if (!((dir == 1) ? s1 : (dir == 2) ? s2 : s3)) return 0u;
      const FaceRange fr = face_range(d, dir);
      return nblocks((int64_t)fr.ni * fr.nj * fr.nk, 256);
    }
    case GK_EMF: return nblocks((int64_t)(d.n1 + s1) * (d.n2 + s2) * (d.n3 + s3), 256);
    case GK_UPDATE: case GK_UPDATEU: case GK_INVERT: case GK_FEATURES: return nblocks(act, 256);
    case GK_DT: return 1u;
    case GK_FXCHECK: case GK_FXRESOLVE: case GK_FXAVG: case GK_FXAPPLY: return group_blocks_fixup(k, v);
    default: return group_blocks_weno(k, v);
  }
}

void launch_group(int k, const GridView *views, const unsigned *first, int n, unsigned total, int mode,
                  hipStream_t s) {
  if (total == 0) return;
  const dim3 g(total);
  switch (k) {
    case GK_BOUND1: hipLaunchKernelGGL(kg_bound_x1, g, dim3(128), 0, s, views, first, n); return;
    case GK_BOUND2: hipLaunchKernelGGL(kg_bound_x2, g, dim3(256), 0, s, views, first, n); return;
    case GK_BOUND3: hipLaunchKernelGGL(kg_bound_x3, g, dim3(256), 0, s, views, first, n); return;
    case GK_FLUX1: hipLaunchKernelGGL(kg_flux_lr<1>, g, dim3(256), 0, s, views, first, n); return;
    case GK_FLUX2: hipLaunchKernelGGL(kg_flux_lr<2>, g, dim3(256), 0, s, views, first, n); return;
    case GK_FLUX3: hipLaunchKernelGGL(kg_flux_lr<3>, g, dim3(256), 0, s, views, first, n); return;
    case GK_EMF: hipLaunchKernelGGL(kg_emf, g, dim3(256), 0, s, views, first, n); return;
    case GK_UPDATE: hipLaunchKernelGGL(kg_update<false>, g, dim3(256), 0, s, views, first, n); return;
    case GK_UPDATEU: hipLaunchKernelGGL(kg_update<true>, g, dim3(256), 0, s, views, first, n); return;
    case GK_INVERT:
      if (mode == INV_MIXED) hipLaunchKernelGGL(kg_invert<INV_MIXED>, g, dim3(256), 0, s, views, first, n);
      else if (mode == INV_LEARNED) hipLaunchKernelGGL(kg_invert<INV_LEARNED>, g, dim3(256), 0, s, views, first, n);
      else hipLaunchKernelGGL(kg_invert<INV_FP64>, g, dim3(256), 0, s, views, first, n);
      return;
    case GK_FEATURES: hipLaunchKernelGGL(kg_step_features, g, dim3(256), 0, s, views, first, n); return;
    case GK_DT: hipLaunchKernelGGL(kg_dt, g, dim3(64), 0, s, views, first, n); return;
    case GK_FXCHECK: case GK_FXRESOLVE: case GK_FXAVG: case GK_FXAPPLY:
      launch_group_fixup(k, views, first, n, total, s);
      return;
    default: launch_group_weno(k, views, first, n, total, s);
  }
}


}  // namespace hg

// ---------------------------------------------------------------------------
// initial time step of the set-up
// ---------------------------------------------------------------------------
namespace hg {

// advance.c:957 set_dt over every zone including ghosts
__global__ void k_set_dt(const double *__restrict__ p, const double *__restrict__ geo, Dims d, double gam,
                         double cour, double dx1, double dx2, double dx3, unsigned long long *__restrict__ out,
                         int *__restrict__ err) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double ndt = 1.e+50;
  if (t < d.spl) {
    int ii = (int)(t % d.m1);
    int64_t r = t / d.m1;
    int jj = (int)(r % d.m2), kk = (int)(r / d.m2);
    int i = ii - d.nb1, j = jj - d.nb2;
    double pr[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) pr[pl] = p[t + pl * d.spl];
    Geom G;
    load_geom(geo, GL_CENT, d.gplane, d.gidx(i, j), G);
    State q;
    (void)kk;
    if (get_state(pr, G, q)) {
      atomicOr(err, 8);
    } else {
      double dxs[4] = {0., dx1, dx2, dx3};
      int nn[4] = {0, d.n1, d.n2, d.n3};
      for (int dir = 1; dir <= 3; dir++) {
        if (nn[dir] <= 1) continue;
        double cmax, cmin;
        if (vchar(pr, q, dir, G, gam, cmax, cmin)) { atomicOr(err, 8); continue; }
        double a = fabs(cmax), b = fabs(cmin);
        double dtij = cour * dxs[dir] / ((a > b) ? a : b);
        if (dtij < ndt) ndt = dtij;
      }
    }
  }
  block_min_to(ndt, out);
}

void launch_set_dt(const double *p, const double *geo, const Dims &d, double gam, double cour, const double *dx,
                   unsigned long long *out, int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt, dim3(nblocks(d.spl, 256)), dim3(256), 0, s, p, geo, d, gam, cour, dx[1], dx[2],
                     dx[3], out, err);
}

}  // namespace hg
