// Launchers of the device kernels (host side of hg_kernels.hip / hg_weno.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hg_grid.h"

namespace hg {

// launch shapes that do not change results (every zone's arithmetic is
// independent of the block it runs in; the block reductions are a minimum and
// integer sums): block sizes (64, 128 or 256 threads), wavefronts per block of
// the i-direction reconstruction, segments per line of the j/k-direction
// reconstruction (0: chosen from the line length)
struct Tune {
  int bs_flux = 256, bs_emf = 256, bs_update = 256, bs_invert = 256;
  int weno_xwaves = 4, weno_nseg = 0, weno_nseg3 = 0;  // segments per line along j, along k
  int weno_xmarch = 0;  // along i: 0 = one lane per point (shuffles), n > 0 = marching, n segments
  int concurrent = 1;   // 1: the three directions' fluxes on three streams, 0: one after another
};

void launch_bound(double *p, const double *geo, const Dims &d, double gammamax, int *err, hipStream_t s);
// point-type reconstruction (lim: HG_LIM_PARA, HG_LIM_MC or another
// three-point limiter) fused with the flux, over the faces of the standard
// (fv = false) or the finite-volume scheme's loops
void launch_flux_point(int dir, int lim, bool fv, const double *p, const double *geo, double *F, const Dims &d,
                       double gam, double cour, double dxdir, int fluxmethod, unsigned long long *ndt_bits,
                       int *err, const Tune &tu, hipStream_t s);
// primtoU of every active zone of p into U (same layout)
void launch_cons_from_prims(const double *p, const double *geo, double *U, const Dims &d, double gam, int *err,
                            hipStream_t s);
void launch_flux_lr(int dir, const double *pleft, const double *pright, const double *geo, double *F,
                    const Dims &d, double gam, double cour, double dxdir, int fluxmethod,
                    unsigned long long *ndt_bits, int *err, const Tune &tu, bool fv, hipStream_t s);
void launch_fluxct_fields(double *F1, double *F2, double *F3, const double *e1, const double *e2, const double *e3,
                          const Dims &d, hipStream_t s);
void launch_emf(const double *F1, const double *F2, const double *F3, const double *geo, const double *corn,
                double *e1, double *e2, double *e3, const Dims &d, const Tune &tu, hipStream_t s);
// per-zone update (-> ubuf) then inversion + fixup (-> pf); ubuf has the
// layout of the primitives
// coefficients of one Runge-Kutta substep (step_ch.c:442 setup_rktimestep):
// without ucum (the midpoint scheme's fast path) Uf = Ui + cuf2 dt dU and the
// last substep inverts Uf; with ucum (every other case) Uf = cuf0 Ui +
// cuf1 uf + cuf2 dt dU with uf the previous substep's Uf, and ucum +=
// cn0 Ui + cn1 dt dU + cn2 Uf, both zeroed first on the first substep; the
// last substep inverts ucum
struct RkStage {
  double cuf2, cn0, cn1, cn2;
  double *ucum;
  int first, last;
  double *ucons;  // finite-volume scheme: the carried conserved state (Ui),
                  // replaced by ucum on the last substep; null otherwise
  double cuf0 = 1.0, cuf1 = 0.0;
};
void launch_update(const double *pi, const double *pb, double *ubuf, const double *F1, const double *F2,
                   const double *F3, const double *e1, const double *e2, const double *e3, const double *geo,
                   const double *conn, const double *dtp, const RkStage &rk, const Dims &d, double gam,
                   const double *dx, int *err, const Tune &tu, hipStream_t s);
// the inversion of ubuf with the guess pb, fixup1zone on the final substep
// -> pf; mode 0: the reference's FP64 iteration, 1: mixed precision to the
// converged root, 2: the same from the learned start (ylearn[zone], the
// network's log(W / W_guess) in k_invert's zone order); fx_reset (may be
// null): post_fixup's three list counters, cleared by the kernel
void launch_invert(const double *ubuf, const double *pb, double *pf, const double *geo, const double *floors,
                   const RkStage &rk, int finalstep, const Dims &d, double gam, double gammamax,
                   unsigned long long *counters, int8_t *pflag, int mode, const float *ylearn, unsigned *fx_reset,
                   int *err, const Tune &tu, hipStream_t s);
// the learned start's features of the inversions launch_invert will do
// (8 floats per zone, k_invert's zone order)
void launch_step_features(const double *ubuf, const double *pb, const double *geo, const Dims &d, double gam,
                          float *X, hipStream_t s);
void launch_dt(unsigned long long *ndt_bits, double *dtp, double defcon, double safe, hipStream_t s);
void launch_reset_ndt(unsigned long long *ndt_bits, hipStream_t s);


// WENO5BND (hg_weno.hip): the per-zone dP/P inputs (launch_weno_prep), then
// per direction the dP/P indicator and the interface states along dir for
// every zone (launch_weno; aux holds 5 x spl doubles: the two inputs and the
// three directions' indicators); pleft/pright have the layout of the
// primitives
// wide: the finite-volume flux loops, which span every transverse ghost row
// launch_weno_prep: also every direction's indicator when all_dpi, and the
// substep's ndt slots reset when ndt_reset is given; launch_weno: the
// direction's indicator first when with_dpi
void launch_weno_prep(const double *p, const double *geo, double *aux, const Dims &d, double gam, int wide,
                      bool all_dpi, unsigned long long *ndt_reset, hipStream_t s);
void launch_weno(int dir, const double *p, double *aux, double *pleft, double *pright, const Dims &d,
                 const Tune &tu, int wide, bool with_dpi, hipStream_t s);

// post_fixup's buffers of one grid (hg_fixup.hip)
struct FxBufs {
  int8_t *flag0;       // inversion codes (active zones)
  int8_t *flag1;       // codes after the neighbour check
  int *cand_z;         // candidate zone indices
  uint16_t *cand_bits; // per neighbour l: bit l gamma vote, bit 8+l u vote
  int8_t *cand_new;    // Jacobi scratch
  int *fix_z;          // zones the average changed
  double *fix_v;       // their RHO..U3 (5 x cap, structure of arrays)
  int8_t *fix_flag;    // their final flag
  int *fail_z;         // failed inversions that are not candidates
  unsigned *count;     // [candidates, fixes, failed non-candidates]
  int cap;
};

// ---------------------------------------------------------------------------
// Grouped launches: one launch of a kernel class covers the blocks of several
// grids.  Grid g owns blocks [first[g], first[g+1]) of the launch (none when
// the class has no work on it) and reads its arguments from views[g]; inside,
// the kernel runs the same per-zone code as its single-grid launch, with the
// grid's own block numbering.  The classes whose block order is laid out for
// the XCDs (work-groups go to the 8 XCDs round robin) get a multiple of 8
// blocks per grid, so that block b % 8 is the same in both numberings.
// ---------------------------------------------------------------------------
struct GridView {
  Dims d;
  const double *pi, *pb;  // this substep's primitives: Ui from pi, fluxes and guess from pb
  double *pf;             // its result, bounded and fixed up in place
  const double *geo, *corn, *conn, *floors;
  double *F[3], *emf[3], *pl[3], *pr[3], *aux, *uf, *dt;
  float *lx, *ly;         // the learned start's features and predictions
  unsigned long long *ndt_bits, *counters;
  int8_t *pflag;
  FxBufs fx;
  int *err;
  RkStage rk;
  double gam, gammamax, cour, defcon, safe;
  double dx[4], idx[4];   // idx = 1 / dx, as launch_update forms it
  int finalstep, fluxmethod;
  int nseg[4], seglen[4], nlb[4];  // reconstruction line segments per direction
};

__device__ __forceinline__ int group_of(const unsigned *__restrict__ first, int n, unsigned b) {
  int lo = 0, hi = n - 1;  // the last grid g with first[g] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

enum GroupClass {
  GK_PREP, GK_DPI, GK_WENO1, GK_WENO2, GK_WENO3, GK_FLUX1, GK_FLUX2, GK_FLUX3, GK_EMF, GK_UPDATE, GK_INVERT,
  GK_DT, GK_BOUND1, GK_BOUND2, GK_BOUND3, GK_FXCHECK, GK_FXRESOLVE, GK_FXAVG, GK_FXAPPLY, GK_FEATURES, GK_MLP,
  GK_UPDATEU,  // the update with ucum (every scheme but the midpoint one)
  GK_NCLASS
};
// blocks of class k on one grid (0: no work there), and one grouped launch of
// class k over n grids (total = first[n] blocks; mode: the inversion's, 0
// FP64, 1 mixed, 2 mixed from the learned start; GK_MLP goes through
// launch_group_mlp, which needs the network)
unsigned group_blocks(int k, const GridView &v);
void launch_group(int k, const GridView *views, const unsigned *first, int n, unsigned total, int mode,
                  hipStream_t s);
// the per-TU parts of the two above
unsigned group_blocks_weno(int k, const GridView &v);
void launch_group_weno(int k, const GridView *views, const unsigned *first, int n, unsigned total, hipStream_t s);
unsigned group_blocks_fixup(int k, const GridView &v);
void launch_group_fixup(int k, const GridView *views, const unsigned *first, int n, unsigned total, hipStream_t s);
// the reconstruction's line segments of a view (v.d set), as launch_weno
// forms them for the standard scheme with the tune's segment requests
void weno_group_shapes(GridView &v, const Tune &tu);

// initial time step over every zone (advance.c:957 set_dt), folded into *out
void launch_set_dt(const double *p, const double *geo, const Dims &d, double gam, double cour, const double *dx,
                   unsigned long long *out, int *err, hipStream_t s);
// the conserved states, zone-centre metric and guesses of a substep's
// inversions as a batch in hg_utoprim_batch's layout
void launch_gather_inversions(const double *ubuf, const double *pb, const double *geo, const Dims &d, double *U,
                              double *gcov, double *gcon, double *gdet, double *gcovpert, double *guess,
                              hipStream_t s);
}  // namespace hg
