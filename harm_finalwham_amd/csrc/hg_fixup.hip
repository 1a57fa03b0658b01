// post_fixup (fixup.c:37) on the device: the neighbour check of good
// inversions (fixup_checksolution, fixup.c:460) and the neighbour average of
// failed ones (fixup_utoprim, fixup.c:669), between the two bound_prim calls
// of a substep (step_ch.c:132-157).
//
// The reference runs both loops in ZLOOP order (i outer, k inner) and every
// flag it sets is seen by the zones after it.  Here:
//   k_fx_check    one thread per active zone: the zone's gamma; a zone with
//                 gamma >= 2 (a candidate) stores, per neighbour, whether the
//                 gamma ratio and the u ratio would vote, and joins a list;
//                 any other zone whose inversion failed joins a second list
//   k_fx_resolve  one workgroup: the candidates' flags as the sequential loop
//                 leaves them.  A candidate's flag depends on the flags of its
//                 three earlier neighbours (i,j-1), (i-1,j), (i-1,j-1) as the
//                 loop has left them, and on the inversion codes of the rest;
//                 Jacobi sweeps over the candidates reach that unique fixed
//                 point (the dependencies follow the loop order, so they
//                 cannot cycle) after at most (list length + 1) sweeps
//   k_fx_average  one thread per entry of the two lists whose flag is
//                 nonzero (every zone the average can change is in one of
//                 them): the new primitives and flag into a list.  A neighbour earlier in the
//                 loop order is seen with the flag the loop gave it (FIXED
//                 for every code the accounting resets), a later one with its
//                 checked flag, a ghost with its bound_pflag copy; values are
//                 read from the primitives as they were before any fix
//                 (ORDERINDEPENDENT), which no kernel writes until
//   k_fx_apply    writes the list back.
// Flags are int8 in the layout of one primitive; ghost flags are not stored:
// bound_pflag (boundsint.c:6) for outflow x1, polar x2, periodic x3 is the
// index map (clamp i, mirror j, wrap k) onto the active zones' inversion codes.
#include <hip/hip_runtime.h>

#include "harmgpu.h"
#include "hg_host.h"
#include "hg_kernels.h"
#include "hg_phys.h"

using namespace hg;

namespace {

constexpr int FX_FIXED = -1;  // global.h:782 UTOPRIMFAILFIXED
constexpr int FX_GAMMAPERC = 12, FX_UPERC = 13, FX_U2AVG1 = 14, FX_U2AVG2 = 15;
constexpr int FX_RHONEG = 9, FX_UNEG = 10, FX_RHOUNEG = 11;
constexpr double GAMMAPERCDIFFMAX = 2.0;  // init.fishmon.h:178
constexpr double UPERCDIFFMAX = 10.0;     // init.fishmon.h:179

__device__ __forceinline__ int64_t zidx(const Dims &d, int i, int j, int k) { return d.idx(i, j, k); }

// bound_pflag: the active zone whose code a (possibly ghost) zone holds
__device__ __forceinline__ int64_t flag_src(const Dims &d, int i, int j, int k) {
  if (d.n1 > 1) i = (i < 0) ? 0 : (i >= d.n1 ? d.n1 - 1 : i);
  if (d.n2 > 1) j = (j < 0) ? -1 - j : (j >= d.n2 ? 2 * d.n2 - 1 - j : j);
  if (d.n3 > 1) k = (k < 0) ? k + d.n3 : (k >= d.n3 ? k - d.n3 : k);
  return d.idx(i, j, k);
}

__device__ __forceinline__ bool is_active(const Dims &d, int i, int j, int k) {
  return i >= 0 && i < d.n1 && j >= 0 && j < d.n2 && k >= 0 && k < d.n3;
}

// gamma_calc (phys.c:1023) from the spatial metric at CENT of (i,j)
__device__ __forceinline__ bool zone_gamma(const double *p, const double *geo, const Dims &d, int i, int j,
                                           int k, double &gamma) {
  Geom G;
  const double *b = geo + (int64_t)GL_CENT * GCOMP * d.gplane + d.gidx(i, j);
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int c = 1; c < 4; c++) G.gcov[a][c] = b[(a * 4 + c) * d.gplane];
  double pr[NPR];
  int64_t z = d.idx(i, j, k);
  pr[U1] = p[z + U1 * d.spl];
  pr[U2] = p[z + U2 * d.spl];
  pr[U3] = p[z + U3 * d.spl];
  return gamma_calc(pr, G, gamma);
}

// the neighbours of fixup_checksolution's votes, in its order (fixup.c:523-540);
// entry 6 repeats entry 4
__device__ __forceinline__ void vote_nbr(int l, int i, int j, int s1, int s2, int &ni, int &nj) {
  const int di[8] = {0, 0, 1, -1, 1, 1, 1, -1};
  const int dj[8] = {1, -1, 0, 0, 1, -1, 1, -1};
  ni = i + di[l] * s1;
  nj = j + dj[l] * s2;
}

// flag of neighbour (ni,nj,k) as the check of zone (i,j,k) sees it:
// active zones earlier in the loop -> their checked flag so far (cur);
// later active zones -> inversion code; ghosts -> bound_pflag copy
__device__ __forceinline__ int chk_flag(const Dims &d, const int8_t *flag0, const int8_t *cur, int i, int j,
                                        int ni, int nj, int k) {
  if (!is_active(d, ni, nj, k)) return flag0[flag_src(d, ni, nj, k)];
  bool earlier = (ni < i) || (ni == i && nj < j);
  return earlier ? cur[d.idx(ni, nj, k)] : flag0[d.idx(ni, nj, k)];
}

__device__ __forceinline__ int vote_result(int f0, unsigned bits, const int *nflag) {
  if (nflag[0] > 0) return f0;  // percdiff[0] < 0: no votes (fixup.c:550)
  int v0 = 0, v1 = 0;
#pragma unroll
  for (int l = 0; l < 8; l++) {
    bool good = nflag[l] <= 0;
    v0 += good && ((bits >> l) & 1u);
    v1 += good && ((bits >> (8 + l)) & 1u);
  }
  int f = f0;
  if (v0 > 8 * 0.5) f = FX_GAMMAPERC;
  if (v1 > 8 * 0.5) f = FX_UPERC;
  return f;
}

__device__ __forceinline__ void fx_check_body(const double *__restrict__ p, const double *__restrict__ geo,
                                              const Dims &d, const FxBufs &b, int *__restrict__ err, unsigned bid) {
  int64_t t = bid * (int64_t)blockDim.x + threadIdx.x;
  int64_t total = (int64_t)d.n1 * d.n2 * d.n3;
  if (t >= total) return;
  int i = (int)(t % d.n1);
  int64_t r = t / d.n1;
  int k = (int)(r % d.n3);
  int j = (int)(r / d.n3);
  int64_t z = d.idx(i, j, k);
  const int8_t f0 = b.flag0[z];
  b.flag1[z] = f0;
  double g;
  if (zone_gamma(p, geo, d, i, j, k, g)) {
    atomicOr(err, 8);
    return;
  }
  if (!(g >= 2.0)) {
    if (f0 != 0) b.fail_z[atomicAdd(&b.count[2], 1u)] = (int)z;
    return;
  }
  const int s1 = d.n1 > 1, s2 = d.n2 > 1;
  const double uu = p[z + UU * d.spl];
  unsigned bits = 0;
#pragma unroll
  for (int l = 0; l < 8; l++) {
    int ni, nj;
    vote_nbr(l, i, j, s1, s2, ni, nj);
    double gn;
    if (zone_gamma(p, geo, d, ni, nj, k, gn)) {
      atomicOr(err, 8);
      return;
    }
    double rg = gn / g;
    double ru = p[d.idx(ni, nj, k) + UU * d.spl] / uu;
    if ((fabs(rg) > GAMMAPERCDIFFMAX) || (fabs(rg) < 1.0 / GAMMAPERCDIFFMAX)) bits |= 1u << l;
    if ((fabs(ru) > UPERCDIFFMAX) || (fabs(ru) < 1.0 / UPERCDIFFMAX)) bits |= 1u << (8 + l);
  }
  unsigned slot = atomicAdd(&b.count[0], 1u);
  b.cand_z[slot] = (int)z;
  b.cand_bits[slot] = (uint16_t)bits;
}

__global__ void __launch_bounds__(256) k_fx_check(const double *__restrict__ p, const double *__restrict__ geo,
                                                  Dims d, FxBufs b, int *__restrict__ err) {
  fx_check_body(p, geo, d, b, err, blockIdx.x);
}

// zone (i,j,k) of a padded index
__device__ __forceinline__ void unidx(const Dims &d, int64_t z, int &i, int &j, int &k) {
  i = (int)(z % d.m1) - d.nb1;
  int64_t r = z / d.m1;
  j = (int)(r % d.m2) - d.nb2;
  k = (int)(r / d.m2) - d.nb3;
}

__device__ __forceinline__ void fx_resolve_body(const Dims &d, const FxBufs &b, int *__restrict__ err) {
  __shared__ int changed;
  const unsigned n = b.count[0];
  if (n == 0) return;
  const int s1 = d.n1 > 1, s2 = d.n2 > 1;
  for (unsigned sweep = 0;; sweep++) {
    for (unsigned c = threadIdx.x; c < n; c += blockDim.x) {
      int i, j, k;
      int64_t z = b.cand_z[c];
      unidx(d, z, i, j, k);
      int nflag[8];
#pragma unroll
      for (int l = 0; l < 8; l++) {
        int ni, nj;
        vote_nbr(l, i, j, s1, s2, ni, nj);
        nflag[l] = chk_flag(d, b.flag0, b.flag1, i, j, ni, nj, k);
      }
      b.cand_new[c] = (int8_t)vote_result(b.flag0[z], b.cand_bits[c], nflag);
    }
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    for (unsigned c = threadIdx.x; c < n; c += blockDim.x) {
      int64_t z = b.cand_z[c];
      if (b.flag1[z] != b.cand_new[c]) {
        b.flag1[z] = b.cand_new[c];
        changed = 1;
      }
    }
    __syncthreads();
    int ch = changed;
    __syncthreads();
    if (!ch) break;
    if (sweep > n + 1) {  // cannot happen for an acyclic dependency; guard the loop
      if (threadIdx.x == 0) atomicOr(err, 16);
      break;
    }
  }
}

__global__ void __launch_bounds__(1024) k_fx_resolve(Dims d, FxBufs b, int *__restrict__ err) {
  fx_resolve_body(d, b, err);
}

// fixup_utoprim accounting (fixup.c:1345-1420): does the loop reset a zone
// with this code to FIXED?  (-1 no, 1 yes, 2 unknown code)
__device__ __forceinline__ int fx_counted(int f, int finalstep) {
  if (f == 0) return -1;
  if (f == 1 || f == 4 || f >= 101 || f == 2 || f == 5 || f == 7 || f == 8 || f == 6 || f == 3) return 1;
  if (f == FX_RHONEG || f == FX_UNEG || f == FX_U2AVG1 || f == FX_U2AVG2 || f == FX_RHOUNEG)
    return finalstep ? 1 : -1;
  if (f == FX_GAMMAPERC || f == FX_UPERC) return 1;
  return 2;
}

__device__ __forceinline__ int fx_final(int f, int finalstep) {
  return (f != 0 && fx_counted(f, finalstep) == 1) ? FX_FIXED : f;
}

// the flag of neighbour (ni,nj,nk) as fixup_utoprim's loop at (i,j,k) sees it
__device__ __forceinline__ int avg_flag(const Dims &d, const int8_t *flag0, const int8_t *flag1, int finalstep,
                                        int i, int j, int k, int ni, int nj, int nk) {
  if (!is_active(d, ni, nj, nk)) return flag0[flag_src(d, ni, nj, nk)];
  int f = flag1[d.idx(ni, nj, nk)];
  bool earlier = (ni < i) || (ni == i && (nj < j || (nj == j && nk < k)));
  return earlier ? fx_final(f, finalstep) : f;
}

// one zone of the average (z: padded index of an active zone with flag f != 0)
__device__ __forceinline__ void fx_average_zone(const double *__restrict__ p, const double *__restrict__ geo,
                                                const Dims &d, const FxBufs &b, int64_t z, int f, int finalstep,
                                                double gammamax, double gam, double *__restrict__ ucons,
                                                int *__restrict__ err) {
  int i, j, k;
  unidx(d, z, i, j, k);
  const int64_t P = d.spl;
  const int cnt = fx_counted(f, finalstep);
  if (cnt == 2) {
    atomicOr(err, 32);
    return;
  }
  double pv[NPR];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) pv[pl] = p[z + pl * P];
  // STEPOVERNEG* == -1: negative densities are left as they are
  const bool fixed = (f == FX_UNEG || f == FX_RHONEG || f == FX_RHOUNEG);
  Geom G;
  {
    const double *gb = geo + (int64_t)GL_CENT * GCOMP * d.gplane + d.gidx(i, j);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        G.gcov[a][c] = gb[(a * 4 + c) * d.gplane];
        G.gcon[a][c] = gb[(16 + a * 4 + c) * d.gplane];
      }
    G.gpert0 = gb[32 * d.gplane];
    G.g = gb[33 * d.gplane];
  }
  if (!fixed) {
    double gamma;
    if (gamma_calc(pv, G, gamma)) {
      atomicOr(err, 8);
      return;
    }
    if (gamma > gammamax) gamma = gammamax;
    const bool u2avg = (f == FX_U2AVG1 || f == FX_U2AVG2);
    const int startpl = u2avg ? UU : RHO, endpl = u2avg ? UU : U3;
    const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
    const int rnx = s1 ? 3 : 1, rny = s2 ? 3 : 1;
    const int npairs = ((s1 ? 3 : 1) * (s2 ? 3 : 1) * (s3 ? 3 : 1) - 1) / 2;
    double sum0[5] = {0., 0., 0., 0., 0.}, sum1[5] = {0., 0., 0., 0., 0.};
    double lastmin = 1E50;
    int numavg0 = 0, numavg1 = 0;
    for (int qq = 0; qq < npairs; qq++) {
      int ii = (qq % rnx) - s1, jj = ((qq % (rnx * rny)) / rnx) - s2, kk = (qq / (rnx * rny)) - s3;
      int64_t za = d.idx(i + ii, j + jj, k + kk), zb = d.idx(i - ii, j - jj, k - kk);
      int thisnotfail = avg_flag(d, b.flag0, b.flag1, finalstep, i, j, k, i + ii, j + jj, k + kk) <= 0;
      int thatnotfail = avg_flag(d, b.flag0, b.flag1, finalstep, i, j, k, i - ii, j - jj, k - kk) <= 0;
      numavg1 += (p[za + UU * P] > 0.0);
      numavg1 += (p[zb + UU * P] > 0.0);
      numavg0 += thisnotfail + thatnotfail;
      if (!u2avg) {
        double *s = (qq % 2) ? sum1 : sum0;
        for (int pl = RHO; pl <= U3; pl++) s[pl] += p[za + pl * P] * thisnotfail + p[zb + pl * P] * thatnotfail;
      } else {
        double ft = p[za + UU * P];
        if (ft >= 0.0) lastmin = (lastmin > ft) ? ft : lastmin;
        ft = p[zb + UU * P];
        if (ft >= 0.0) lastmin = (lastmin > ft) ? ft : lastmin;
      }
    }
    int numavg;
    if (!u2avg) {
      if (numavg0 != 0)
        for (int pl = RHO; pl <= U3; pl++) pv[pl] = (sum0[pl] + sum1[pl]) / ((double)(numavg0));
      numavg = numavg0;
    } else {
      if (numavg1 != 0) pv[UU] = lastmin;
      numavg = numavg1;
    }
    if (f == FX_U2AVG2) numavg++;
    if (numavg == 0) {
      // NOGOODAVERAGE (fixup.c:1288) with AVG4_1, AVG4_2 (fixup.c:579-580)
      int ip = i + s1, im = i - s1, jp = j + s2, jm = j - s2;
      for (int pl = startpl; pl <= endpl; pl++) {
        double a41 = 0.25 * ((p[d.idx(i, jp, k) + pl * P] + p[d.idx(i, jm, k) + pl * P]) +
                             (p[d.idx(im, j, k) + pl * P] + p[d.idx(ip, j, k) + pl * P]));
        double a42 = 0.25 * ((p[d.idx(ip, jp, k) + pl * P] + p[d.idx(ip, jm, k) + pl * P]) +
                             (p[d.idx(im, jp, k) + pl * P] + p[d.idx(im, jm, k) + pl * P]));
        pv[pl] = 0.5 * (a41 + a42);
      }
    }
    if (limit_gamma(gamma, pv, G) >= 1) {
      atomicOr(err, 8);
      return;
    }
  }
  if (ucons && finalstep && cnt == 1) {
    // fixup.c:1428 diag_fixup -> fixup.c:240-248 (finite-volume scheme):
    // the carried conserved state of this zone follows the fix; U of the
    // pre-fix and fixed states each passed through UtoU (UDIAG -> UEVOLVE),
    // once per energy region (the zone is correctable in each of them)
    double pr0[NPR], U0[NPR], U1[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) pr0[pl] = p[z + pl * P];
    primtoU(pr0, G, gam, U0);
    primtoU(pv, G, gam, U1);
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) {
      const double du = (U1[pl] / G.g) * G.g - (U0[pl] / G.g) * G.g;
      double u = ucons[z + pl * P];
      for (int region = 0; region < NENERREGIONS; region++) u = u + du;
      ucons[z + pl * P] = u;
    }
  }
  unsigned slot = atomicAdd(&b.count[1], 1u);
  b.fix_z[slot] = (int)z;
  for (int pl = RHO; pl <= U3; pl++) b.fix_v[(int64_t)pl * b.cap + slot] = pv[pl];
  b.fix_flag[slot] = (int8_t)fx_final(f, finalstep);
}

// over the candidates (their flags as k_fx_resolve left them) and the other
// failed inversions; a few blocks, since the lists are short next to the
// grid and the average takes the whole register file of a SIMD lane
__device__ __forceinline__ void fx_average_body(const double *__restrict__ p, const double *__restrict__ geo,
                                                const Dims &d, const FxBufs &b, int finalstep, double gammamax,
                                                double gam, double *__restrict__ ucons, int *__restrict__ err,
                                                unsigned bid, unsigned nblk) {
  const unsigned nc = b.count[0], n = nc + b.count[2];
  for (unsigned s = bid * blockDim.x + threadIdx.x; s < n; s += nblk * blockDim.x) {
    const int64_t z = (s < nc) ? b.cand_z[s] : b.fail_z[s - nc];
    const int f = b.flag1[z];
    if (f != 0) fx_average_zone(p, geo, d, b, z, f, finalstep, gammamax, gam, ucons, err);
  }
}

__global__ void __launch_bounds__(256) k_fx_average(const double *__restrict__ p, const double *__restrict__ geo,
                                                    Dims d, FxBufs b, int finalstep, double gammamax, double gam,
                                                    double *__restrict__ ucons, int *__restrict__ err) {
  fx_average_body(p, geo, d, b, finalstep, gammamax, gam, ucons, err, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void fx_apply_body(double *__restrict__ p, const Dims &d, const FxBufs &b, unsigned bid,
                                              unsigned nblk) {
  const unsigned n = b.count[1];
  for (unsigned s = bid * blockDim.x + threadIdx.x; s < n; s += nblk * blockDim.x) {
    int64_t z = b.fix_z[s];
    for (int pl = RHO; pl <= U3; pl++) p[z + pl * d.spl] = b.fix_v[(int64_t)pl * b.cap + s];
    b.flag0[z] = b.fix_flag[s];
  }
}

__global__ void __launch_bounds__(256) k_fx_apply(double *__restrict__ p, Dims d, FxBufs b) {
  fx_apply_body(p, d, b, blockIdx.x, gridDim.x);
}

// grouped launches (one per class over several grids; the standard scheme,
// so no carried conserved state)
#define FX_GROUP_PRELUDE                              \
  const int g_ = group_of(first, n, blockIdx.x);      \
  const GridView &v = views[g_];                      \
  const unsigned bid = blockIdx.x - first[g_], nblk = first[g_ + 1] - first[g_]; \
  (void)bid;                                          \
  (void)nblk;

__global__ void __launch_bounds__(256) kg_fx_check(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  FX_GROUP_PRELUDE
  fx_check_body(v.pf, v.geo, v.d, v.fx, v.err, bid);
}

__global__ void __launch_bounds__(1024) kg_fx_resolve(const GridView *__restrict__ views,
                                                      const unsigned *__restrict__ first, int n) {
  FX_GROUP_PRELUDE
  fx_resolve_body(v.d, v.fx, v.err);
}

__global__ void __launch_bounds__(256) kg_fx_average(const GridView *__restrict__ views,
                                                     const unsigned *__restrict__ first, int n) {
  FX_GROUP_PRELUDE
  fx_average_body(v.pf, v.geo, v.d, v.fx, v.finalstep, v.gammamax, v.gam, nullptr, v.err, bid, nblk);
}

__global__ void __launch_bounds__(256) kg_fx_apply(const GridView *__restrict__ views,
                                                   const unsigned *__restrict__ first, int n) {
  FX_GROUP_PRELUDE
  fx_apply_body(v.pf, v.d, v.fx, bid, nblk);
}
#undef FX_GROUP_PRELUDE

__global__ void k_fx_reset(unsigned *count) {
  if (threadIdx.x < 3) count[threadIdx.x] = 0u;
}

}  // namespace

namespace hg {

int fixup_alloc(hg_ctx *c) {
  const Dims &d = c->d;
  int64_t act = (int64_t)d.n1 * d.n2 * d.n3;
  if (act > 0x7fffffff || d.spl > 0x7fffffff) {
    set_error("hg_ctx_create: grid too large for the fixup lists");
    return 1;
  }
  hipError_t e = hipSuccess;
  auto get = [&](void **ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, bytes);
  };
  get((void **)&c->pflag, (size_t)d.spl);
  get((void **)&c->pchk, (size_t)d.spl);
  get((void **)&c->fx_cand_z, (size_t)act * sizeof(int));
  get((void **)&c->fx_cand_bits, (size_t)act * sizeof(uint16_t));
  get((void **)&c->fx_cand_new, (size_t)act);
  get((void **)&c->fx_fix_z, (size_t)act * sizeof(int));
  get((void **)&c->fx_fix_v, (size_t)act * 5 * sizeof(double));
  get((void **)&c->fx_fix_flag, (size_t)act);
  get((void **)&c->fx_fail_z, (size_t)act * sizeof(int));
  get((void **)&c->fx_count, 3 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(c->pflag, 0, (size_t)d.spl);
  if (e != hipSuccess) {
    set_error("hg_ctx_create: %s", hipGetErrorString(e));
    return 1;
  }
  return 0;
}

void fixup_free(hg_ctx *c) {
  void *b[] = {c->pflag, c->pchk, c->fx_cand_z, c->fx_cand_bits, c->fx_cand_new, c->fx_fix_z, c->fx_fix_v,
               c->fx_fix_flag, c->fx_fail_z, c->fx_count};
  for (void *p : b)
    if (p) (void)hipFree(p);
}

FxBufs fixup_bufs(hg_ctx *c) {
  FxBufs b;
  b.flag0 = c->pflag;
  b.flag1 = c->pchk;
  b.cand_z = c->fx_cand_z;
  b.cand_bits = c->fx_cand_bits;
  b.cand_new = c->fx_cand_new;
  b.fix_z = c->fx_fix_z;
  b.fix_v = c->fx_fix_v;
  b.fix_flag = c->fx_fix_flag;
  b.fail_z = c->fx_fail_z;
  b.count = c->fx_count;
  b.cap = (int)((int64_t)c->d.n1 * c->d.n2 * c->d.n3);
  return b;
}

// fixup_checksolution + fixup_utoprim on the bounded primitives p, with the
// inversion codes in c->pflag; kernels only (capturable)
void launch_post_fixup(hg_ctx *c, double *p, int finalstep, hipStream_t s, bool reset) {
  const Dims &d = c->d;
  FxBufs b = fixup_bufs(c);
  int64_t n = (int64_t)d.n1 * d.n2 * d.n3;
  unsigned nb = (unsigned)((n + 255) / 256);
  if (reset) hipLaunchKernelGGL(k_fx_reset, dim3(1), dim3(64), 0, s, b.count);
  hipLaunchKernelGGL(k_fx_check, dim3(nb), dim3(256), 0, s, p, c->geo, d, b, c->err);
  hipLaunchKernelGGL(k_fx_resolve, dim3(1), dim3(1024), 0, s, d, b, c->err);
  hipLaunchKernelGGL(k_fx_average, dim3(256), dim3(256), 0, s, p, c->geo, d, b, finalstep, c->cfg.gammamax,
                     c->cfg.gam, c->ucons, c->err);
  hipLaunchKernelGGL(k_fx_apply, dim3(256), dim3(256), 0, s, p, d, b);
}

// grouped: the list kernels take 8 blocks per grid (their lists are short
// next to the grid; the single-grid launch takes 256)
unsigned group_blocks_fixup(int k, const GridView &v) {
  const int64_t n = (int64_t)v.d.n1 * v.d.n2 * v.d.n3;
  if (k == GK_FXCHECK) return (unsigned)((n + 255) / 256);
  if (k == GK_FXRESOLVE) return 1u;
  if (k == GK_FXAVG || k == GK_FXAPPLY) return 8u;
  return 0u;
}

void launch_group_fixup(int k, const GridView *views, const unsigned *first, int n, unsigned total,
                        hipStream_t s) {
  if (k == GK_FXCHECK) hipLaunchKernelGGL(kg_fx_check, dim3(total), dim3(256), 0, s, views, first, n);
  if (k == GK_FXRESOLVE) hipLaunchKernelGGL(kg_fx_resolve, dim3(total), dim3(1024), 0, s, views, first, n);
  if (k == GK_FXAVG) hipLaunchKernelGGL(kg_fx_average, dim3(total), dim3(256), 0, s, views, first, n);
  if (k == GK_FXAPPLY) hipLaunchKernelGGL(kg_fx_apply, dim3(total), dim3(256), 0, s, views, first, n);
}

}  // namespace hg
