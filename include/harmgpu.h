/*
 * harmgpu.h -- C ABI of the MI355X implementation of the per-step GRMHD
 * hot path of pseudotensor/harm_finalwham.
 *
 * The entry points replace, on device-resident state, the functions that
 * advance.c / step_ch.c call in the reference:
 *
 *   hg_bound_prim     bound_prim()        step_ch.c:568 -> bounds.fishmon.c:29
 *   hg_fluxcalc       fluxcalc()          flux.c:7 (which ends in flux_ct(), fluxct.c:10)
 *   hg_step           step_ch()           step_ch.c:16 -> step_ch_simplempi step_ch.c:35
 *                     (RK2 midpoint: two advance() calls, advance.c:58 -- flux
 *                     update + source + Utoprimgen + fixup1zone per zone --
 *                     each followed by bound_prim(), then the dt update)
 *   hg_utoprim_batch  Utoprimgen()        utoprimgen.c:23 with UTOPRIM2DFINAL
 *                     (utoprim_2d_final.c:34), on a batch of independent zones
 *
 * and the problem set-up of init.fishmon.c / initbase.c init()
 * (hg_mks_geometry_tables, hg_build_mks_geometry, hg_init_torus).
 *
 * Conventions follow the reference: NPR=8 primitives
 * [rho, u, U1, U2, U3, B1, B2, B3] (relative 4-velocity, WHICHVEL VELREL4),
 * conserved U = gdet*[rho u^t, T^t_t + rho u^t, T^t_i, B^i]
 * (REMOVERESTMASSFROMUU 2), failure codes of global.h:782-799.
 *
 * Every function returns 0 on success and >=1 on error (the MYFUN
 * convention of global.h:1700); hg_last_error() describes the last error.
 * Soft per-zone inversion failures are not errors: they are reported as
 * failure codes (hg_utoprim_batch) or counted (hg_step, hg_get_counters).
 *
 * Host arrays use the reference's layouts:
 *   primitives  p[i][j][k][pl], i in [0,n1+2nb), ... (ghost zones included,
 *               nb ghost layers on each side of every direction of size > 1)
 *   geometry    per (i,j) with i in [0,n1+2nb+1), j in [0,n2+2nb+1) and
 *               7 locations (FACE1, FACE2, FACE3, CORN1, CORN2, CORN3, CENT):
 *               gcov[..][7][4][4], gcon[..][7][4][4], gcovpert[..][7][4], gdet[..][7]
 *               conn[i][j][4][4][4] over (n1+2nb) x (n2+2nb), r_cent likewise.
 * Device arrays passed to hg_utoprim_batch are structure-of-arrays (see there).
 * Streams are hipStream_t passed as void* (NULL = default stream).
 */
#ifndef HARMGPU_H
#define HARMGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_NPR 8

/* reconstruction and flux choices (reference values of lim/fluxmethod, global.h:247-314) */
#define HG_LIM_DONOR 0 /* no slope (interppoint.c:433) */
#define HG_LIM_VANL 1  /* van Leer (interppoint.c:399) */
#define HG_LIM_MINM 2  /* minmod of the one-sided and central slopes (interppoint.c:416) */
#define HG_LIM_MC 3
#define HG_LIM_PARA 4 /* para4 (interppoint.c:738) with PARA2LIM MC, WHICHPARA PARA4 */
#define HG_LIM_WENO5BND 17
#define HG_FLUX_LAXF 0
#define HG_FLUX_HLL 1
#define HG_FLUX_FORCE 2    /* GFORCE with the Lax-Wendroff state inverted at the face (fluxcompute.c:298) */
#define HG_FLUX_MUSTA 3    /* single-cell MUSTA on HLL (fluxcompute.c:398, 538) */
#define HG_FLUX_HLLLAXF1 4 /* HLL / LaxF blended by the jump in u (fluxcompute.c:189) */
/* DOENOFLUX (global.h:188-191) */
#define HG_NOENOFLUX 0
#define HG_ENOFINITEVOLUME 3

typedef struct hg_ctx hg_ctx;

typedef struct {
  int n1, n2, n3;      /* active zones */
  int nb;              /* ghost layers per side (>= 6 for WENO5BND, initbase.c:792) */
  int lim;             /* HG_LIM_DONOR, _VANL, _MINM, _MC, _PARA or _WENO5BND */
  int fluxmethod;      /* HG_FLUX_LAXF, HG_FLUX_HLL, HG_FLUX_FORCE, HG_FLUX_MUSTA or HG_FLUX_HLLLAXF1 */
  double gam;          /* adiabatic index */
  double cour;         /* Courant factor */
  double safe;         /* max dt growth per step (SAFE, step_ch.c:245) */
  double defcon;       /* dt prefactor (advance.c:231) */
  double dx[4];        /* coordinate spacings dx[1..3] */
  double rhomin, uumin;/* floor coefficients RHOMIN, UUMIN (initbase.c:459-460) */
  double gammamax;     /* GAMMAMAX (initbase.c:471) */
  int timeorder;       /* TIMEORDER (step_ch.c:442 setup_rktimestep): 1 Euler, 2 (or 0) midpoint,
                          3 TVD RK3, 4 classical RK4 */
  int doenoflux;       /* HG_NOENOFLUX (advance.c:58 advance_standard) or HG_ENOFINITEVOLUME
                          (advance.c:245 advance_finitevolume with the torus file's avgscheme
                          DONOR: the conserved state is carried from step to step and the
                          flux loops span every transverse ghost layer) */
} hg_config;

/* Context with device-resident state for one grid.  The caller owns the
   context; nothing in the step allocates. */
int hg_ctx_create(const hg_config *cfg, hg_ctx **out);
int hg_ctx_destroy(hg_ctx *ctx);

/* Geometry tables in the host layout described above (reference set_grid.c:7
   output restricted to the locations the hot path reads). */
int hg_upload_geometry(hg_ctx *ctx, const double *gcov, const double *gcon,
                       const double *gcovpert, const double *gdet, const double *conn,
                       const double *r_cent);

/* Geometry of the modified Kerr-Schild grid in the host layout of
   hg_upload_geometry (set_grid.c:7 with coord.c:389 LOGRSINTH, metric.c:697
   KS metric, metric.c:1388 numerical connection): spin a, radial range
   [Rin, Rout] with R0, theta concentration hslope.  Host only, needs no GPU.
   dx (may be NULL) receives the coordinate spacings (coord.c:1142). */
int hg_mks_geometry_tables(int n1, int n2, int n3, int nb, double a, double Rin, double Rout,
                           double R0, double hslope, double *gcov, double *gcon,
                           double *gcovpert, double *gdet, double *conn, double *r_cent,
                           double dx[4]);

/* hg_mks_geometry_tables for the context's grid, uploaded; also sets the
   context's dx[1..3] to the grid spacings (read them with hg_get_config). */
int hg_build_mks_geometry(hg_ctx *ctx, double a, double Rin, double Rout, double R0,
                          double hslope);

/* Copy of the context's configuration. */
int hg_get_config(const hg_ctx *ctx, hg_config *out);

/* Fishbone-Moncrief torus initial data (init.fishmon.c:237 init_primitives and
   initbase.c:5 init) on the grid built by hg_build_mks_geometry: r_in, r_pmax,
   kappa, beta_min, relative amplitude of the uniform random perturbation of u
   and its seed.  Returns the initial dt (advance.c:957 set_dt) in *dt. */
int hg_init_torus(hg_ctx *ctx, double rin, double rmax, double kappa, double beta,
                  double randfact, uint64_t seed, double *dt);

int hg_set_prims(hg_ctx *ctx, const double *p_host);
int hg_get_prims(hg_ctx *ctx, double *p_host);

/* The conserved state the finite-volume scheme carries between steps (unew
   of the reference, restart.c:244), in the layout of hg_set_prims (ghost
   entries ignored on input, zero on output).  hg_set_prims and
   hg_init_torus reset it to U(p) (initbase.c:147 init_conservatives) at the
   next step; hg_set_conserved overrides that, as a restart does
   (initbase.c:155-160).  Errors unless doenoflux is HG_ENOFINITEVOLUME. */
int hg_set_conserved(hg_ctx *ctx, const double *U_host);
int hg_get_conserved(hg_ctx *ctx, double *U_host);

/* bound_prim (step_ch.c:568) on the current primitives. */
int hg_bound_prim(hg_ctx *ctx, void *stream);

/* fluxcalc + flux_ct (flux.c:7, fluxct.c:10) on the current primitives;
   F1..F3 (host, may be NULL) receive the fluxes in the p layout, ndt[3] the
   per-direction time-step limits. */
int hg_fluxcalc(hg_ctx *ctx, double *F1, double *F2, double *F3, double ndt[3], void *stream);

/* One full RK2 step (step_ch.c:35 with TIMEORDER 2).  *dt in: step size,
   out: the next step size.  nfail/iters (may be NULL) receive the number of
   failed inversions and the total Newton iterations of the step. */
int hg_step(hg_ctx *ctx, double *dt, int64_t *nfail, int64_t *iters, void *stream);

/* One substep of the reference's step loop with the caller's coefficients
   (step_ch.c:97-166: advance (advance.c:16), bound_prim, post_fixup,
   bound_prim), for a caller that keeps its own setup_rktimestep.  pi, pb, pf
   name the context's state arrays: 0 the primitives (hg_set_prims), 1 and 2
   the two intermediate states (pk[0], pk[1]).  CUf[0..2] and Cunew[0..2] as
   in dUtoU (advance.c:922); timeorder 0 zeroes uf and ucum, the last
   (numtimeorders - 1) inverts ucum and applies fixup1zone.  dt is the step's
   dt; *ndt (may be NULL) receives advance's ndt (advance.c:231).  The
   context's timeorder does not matter here. */
int hg_advance_substep(hg_ctx *ctx, int pi, int pb, int pf, const double CUf[4], const double Cunew[3],
                       int timeorder, int numtimeorders, double dt, double *ndt, void *stream);

/* Same as hg_step but leaves dt on the device and does not synchronise:
   n steps are enqueued on `stream`; hg_get_dt reads the device dt. */
int hg_step_async(hg_ctx *ctx, int nsteps, void *stream);
int hg_set_dt(hg_ctx *ctx, double dt);
int hg_get_dt(hg_ctx *ctx, double *dt);
int hg_get_counters(hg_ctx *ctx, int64_t *nfail, int64_t *iters);

/* post_fixup (fixup.c:37) with the bound_prim calls around it
   (step_ch.c:132-157), on the context's primitives: bound_prim, the
   neighbour check of good inversions (fixup_checksolution, fixup.c:460),
   the neighbour average of failed ones (fixup_utoprim, fixup.c:669),
   bound_prim.  hg_step runs it in every substep.  codes (host, may be NULL:
   the codes of the last inversion) are failure codes of the active zones,
   [i][j][k] over n1 x n2 x n3; finalstep as in the reference (1 on the last
   RK substep).  hg_get_pflag returns the active zones' codes after the
   last post_fixup or inversion (FIXED = -1 for repaired zones). */
int hg_post_fixup(hg_ctx *ctx, const int32_t *codes, int finalstep, void *stream);
int hg_get_pflag(hg_ctx *ctx, int32_t *codes);

/* Diagnostics of the global energy region on the context's state, computed
   on the device (dump_ener.c / diag.c of the reference):
     cons[pl]        constotal (dump_ener.c:596): sum over the active zones of
                     U[pl] dx1 dx2 dx3 (U in the evolved form, with gdet)
     divbmax/avg     divbmaxavg (dump_ener.c:482): max and mean over the
                     active zones of the normalised flux-CT |div B|, and the
                     (i,j,k) of the first maximum
     pdot[dir][pl]   diag_flux (diag.c:734) as advance.c:130 calls it: flux
                     through each boundary (dir = X1UP, X1DN, X2UP, X2DN,
                     X3UP, X3DN; zero where the dimension is flat) from the
                     last fluxcalc (after hg_step: its final substep, on the
                     RK midpoint state), field fluxes as flux-CT leaves them
     pdotterms       its term-by-term split from the state that fluxcalc
                     read, [dir][term 0..6][pl], and pdottermsjet2 the same
                     over the jet boundary zones (b^2/rho > 3.162 at X1DN,
                     outgoing unbound at X1UP)
     pcum[dir][pl]   the context's running sum of pdot Dt (diag.c:1002).
                     The reference adds Cunew[1] dt per substep, which is 0
                     in both substeps of its RK2 midpoint scheme, so its pcum
                     stays 0; pass Dt = 0 to follow it.
   Sums are taken in the reference's loop order (bit-identical to its
   double sums); div B carries a few ulp, the reference normalising it in
   x87 extended precision. */
typedef struct {
  double cons[8];
  double divbmax, divbavg;
  int32_t divbmax_ijk[3];
  double pdot[6][8];
  double pdotterms[6][7][8];
  double pdottermsjet2[6][7][8];
  double pcum[6][8];
} hg_diag;
int hg_diagnostics(hg_ctx *ctx, double Dt, hg_diag *out, void *stream);
/* Two of diag_flux's terms read the state at the start of the step, which a
   step overwrites.  hg_diag_record(ctx, 1) makes every step record the
   boundary terms where the reference computes them, so that hg_diagnostics
   after hg_step reproduces diag_flux exactly; without it hg_diagnostics
   takes those two terms from the current state.  With RK4 and recording,
   hg_diagnostics adds every substep's pdot to pcum with Dt = Cunew[t][1] Dt
   (pass the step's dt as Dt), as the reference's per-substep diag_flux does;
   otherwise pcum += pdot Dt.  The finite-volume scheme's conserved totals
   are those of its carried state (dump_ener.c:611). */
int hg_diag_record(hg_ctx *ctx, int on);

/* Restart dumps of the reference (restart.c:206 restart_write, text output):
   a header line, then one line per active zone (k outer, i fastest) with
   p[0..7] and unew[0..7] at 15 significant digits.  hg_rdump_parse reads
   one on the host (no GPU): the active zones' primitives into p_active and
   the conserved state into unew (both [i][j][k][8] over n1 x n2 x n3, may be
   NULL), and the header fields the hot path uses into *info (may be NULL);
   the header's defcoord is skipped (the reference prints that double with
   "%d").  hg_rdump_read puts the primitives into a context and runs
   bound_prim; a context running the finite-volume scheme also takes unew
   as its carried conserved state.  hg_rdump_write (restart.c:206
   restart_write) writes header (the text of the reference's header line,
   restart.c:641, which the caller formats from its global state) and then
   p and unew of every active zone in the reference's format; unew is the
   last step's ucum (the carried state for the finite-volume scheme), so
   the standard scheme needs a step since the primitives were set. */
typedef struct {
  int n[3];
  double t, tf;
  long nstep;
  double a, gam, cour;
  double DT[4];            /* DTd, DTavg, DTener, DTi */
  long DTr;
  double DTdebug;
  long counts[5];          /* dump, image, rdump, avg, debug counters */
  double dt;
  int lim, timeorder, fluxmethod, fluxb, utoprimversion, failed;
  double R0, Rin, Rout, hslope;
  int bctype[6];           /* X1UP, X1DN, X2UP, X2DN, X3UP, X3DN */
  long realnstep;
  double defcon, safe, rhomin, uumin, gammamax, gammafail;
} hg_rdump_info;
int hg_rdump_parse(const char *path, int n1, int n2, int n3, double *p_active, double *unew, hg_rdump_info *info);
int hg_rdump_read(hg_ctx *ctx, const char *path, hg_rdump_info *info, double *unew, void *stream);
int hg_rdump_write(hg_ctx *ctx, const char *path, const char *header, void *stream);

/* Batched Noble-2D inversion on device memory (utoprimgen.c:23 for each of
   n independent zones).  All arrays are structure-of-arrays over n zones:
     U[8][n] evolved conserved variables (with gdet), gcov[16][n], gcon[16][n]
     (symmetric: only the upper triangles, rows 4a+b with a <= b, are read),
     gdet[n], gcovpert[n] (the tt component; not read), pr[8][n] guess in /
     solution out (the inversion reads the guess's u and u~^i; it writes the
     field, and rho, u, u~^i where the reference assigns them: codes 0, 9, 10,
     11), pflag[n] failure code out, iters[n] Newton steps out (may be NULL);
     n < 2^32.
   mode:
     HG_C2P_FP64     the reference arithmetic: same iterates, codes, counts.
     HG_C2P_REFINED  HG_C2P_FP64, then for each solved zone the root of the
                     same FP64 system converged (Newton steps with the
                     residuals in double-double arithmetic) and the primitives
                     formed from it in double-double: the reference's codes
                     and counts (plus the polish steps), values within ~1e-13
                     of the exact root.  No counterpart in the reference (its
                     EXTRA_NEWT_ITER, u2p_defs.h:77, continues the FP64
                     iteration, which cannot go below its rounding noise).
     HG_C2P_MIXED    the reference's FP64 set-up and start, the Newton
                     iteration in FP32 to |dW/W| <= 2e-6 (or to its rounding
                     floor), then the polish of HG_C2P_REFINED and the
                     primitives from the converged W with the reference's
                     FP64 formulas (within ~1e-12 of HG_C2P_REFINED's); a
                     persistent work-queue kernel.  Zones whose FP32 phase or
                     polish does not converge are inverted by HG_C2P_REFINED
                     from the guess.
                     Allocates its scratch (4 n + 24 bytes) stream-ordered on
                     the given stream. */
#define HG_C2P_FP64 0
#define HG_C2P_MIXED 1
#define HG_C2P_REFINED 2
int hg_utoprim_batch(int64_t n, double gam, const double *U, const double *gcov,
                     const double *gcon, const double *gdet, const double *gcovpert,
                     double *pr, int32_t *pflag, int32_t *iters, int mode, void *stream);

/* hg_utoprim_batch without any allocation.  No counterpart in the reference:
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  Same
   arguments, kernels and results as hg_utoprim_batch.  HG_C2P_MIXED keeps its
   work-queue counters and compacted fallback list in a static device array
   of the library (HG_C2P_NOALLOC_MAX_N zones: 4 * 2^24 + 24 bytes, held on
   every GPU the library runs on until the process ends); larger n is
   rejected.  Every call on one GPU shares that GPU's array: calls that
   overlap in time on one GPU (other streams, other threads) race and must
   not be issued. */
#define HG_C2P_NOALLOC_MAX_N ((int64_t)1 << 24)
int hg_utoprim_batch_noalloc(int64_t n, double gam, const double *U, const double *gcov,
                             const double *gcon, const double *gdet, const double *gcovpert,
                             double *pr, int32_t *pflag, int32_t *iters, int mode, void *stream);

/* The inversions of the last step's final substep (advance.c:193-201,
   Utoprimgen of Uf with the guess pb) as a batch for hg_utoprim_batch: for
   each active zone t = i + n1 (j + n2 k), U[8][n] the conserved state it
   inverted, gcov/gcon[16][n], gdet[n] and gcovpert[n] at the zone centre,
   guess[8][n] the substep's pb; n = n1 n2 n3, device memory.  The midpoint
   scheme (timeorder 2) without the finite-volume step, after hg_step or
   hg_step_async.  No counterpart in the reference (the learned start's
   training and evaluation data from the step itself). */
int hg_dump_inversions(hg_ctx *ctx, double *U, double *gcov, double *gcon, double *gdet, double *gcovpert,
                       double *guess, void *stream);

/* Device-side generator of the batched-inversion benchmark distribution
   (see hg_utoprim_batch for the layouts): KS metric of spin a at r, theta,
   primitives, their conserved variables and a perturbed guess. */
int hg_make_c2p_batch(int64_t n, double gam, double a, uint64_t seed, double *U, double *gcov,
                      double *gcon, double *gdet, double *gcovpert, double *p_true,
                      double *p_guess, void *stream);

/* Number of zones of the last HG_C2P_MIXED batch on the current device that
   were inverted by the HG_C2P_REFINED path (the FP32 phase or the polish did
   not converge).  No counterpart in the reference. */
int hg_c2p_fallbacks(int64_t *nfallback);

/* Learned initial guess of the Noble-2D Newton iteration, no counterpart in
   the reference (whose guess is the previous primitives, utoprim_2d_final.c
   Utoprim_new_body).  A perceptron 8 -> 64 -> 64 -> 1 (tanh) maps invariants
   of the conserved state and the plain guess to y = log(W / W_guess); the
   hidden layers run on the matrix cores with BF16 weights (each input as a
   hi + lo BF16 pair) and FP32 accumulation, the FP32 master weights are
   trained on the GPU (Adam) from inversions this library solved.
   Parameter block (FP32, hg_mlp_nparams() = 4817 entries):
     W1[8][64] b1[64] W2[64][64] b2[64] W3[64] b3 mu[8] sd[8]
   where the network sees (x - mu) / sd.  Features x per zone:
     log(-Q.n / D), log1p(Qt^2 / (Q.n)^2), log1p(B^2 / D), (Q.B)^2/(B^2 Qt^2),
     log10(D), alpha, log(W_guess / D), gam
   with D = alpha rho u^t and W_guess the plain guess's W.
   All arrays are device memory; X is [n][8] float, y [n] float, the
   inversion arrays are laid out as for hg_utoprim_batch. */
typedef struct hg_mlp hg_mlp;
int hg_mlp_create(uint64_t seed, hg_mlp **out);
int hg_mlp_destroy(hg_mlp *m);
int hg_mlp_nparams(void);
int hg_mlp_get_params(const hg_mlp *m, float *host);
int hg_mlp_set_params(hg_mlp *m, const float *host);
/* features X[n][8] of n inversions from their plain guess p_guess[8][n];
   with solved primitives p_solved[8][n] (may be NULL) also the training
   target y[n] = log(W_solved / W_guess), which the network predicts; ok[n]
   (may be NULL) is 1 where the set-up succeeded and every value is finite */
int hg_mlp_features(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                    const double *gdet, const double *p_guess, const double *p_solved, float *X, float *y,
                    int32_t *ok, void *stream);
int hg_mlp_forward(const hg_mlp *m, int64_t n, const float *X, float *y, void *stream);
/* gradient of mean((y - Y)^2) with respect to the parameter block (host
   copy in grad_host, may be NULL; normalisation entries are 0), the loss in
   *loss; the weights are not changed */
int hg_mlp_gradient(hg_mlp *m, int64_t n, const float *X, const float *Y, float *grad_host, float *loss,
                    void *stream);
/* `epochs` passes of Adam over minibatches of `batch` (>= 64) samples in order */
int hg_mlp_train(hg_mlp *m, int64_t n, const float *X, const float *Y, int epochs, int batch, float lr,
                 float *loss, void *stream);
/* hg_utoprim_batch's HG_C2P_MIXED inversion with a learned start.  Where
   the network predicts |y| > 0.25 (the plain guess off by more than ~28% in
   W), the Newton step of the FP64 system is evaluated at W_guess exp(y); when
   it is finite and at most half of W it is taken as the first step and the
   FP32 iteration continues from there, else the FP32 iteration starts from
   the plain start (the reference's, from the guess).  The polish and
   double-double primitives of HG_C2P_REFINED follow.  A zone whose FP32
   phase or polish does not converge is inverted by HG_C2P_REFINED from its
   plain guess.  The values therefore agree with HG_C2P_REFINED to ~1e-12
   (the same converged root), not bit for bit.  iters counts Newton steps and
   the learned start's evaluation.  counts (may be NULL) receives [zones solved from the learned
   start, zones inverted by the fallback].  Allocates its scratch (4 n + 24
   bytes) stream-ordered on the given stream. */
int hg_utoprim_batch_learned(const hg_mlp *m, int64_t n, double gam, const double *U, const double *gcov,
                             const double *gcon, const double *gdet, const double *gcovpert, double *pr,
                             int32_t *pflag, int32_t *iters, int64_t counts[2], void *stream);

/* The inversion the step uses for every zone (advance.c:193-201, Utoprimgen
   of Uf with the guess pb), no counterpart in the reference:
     HG_C2P_FP64     the reference's FP64 iteration, bit for bit (default);
     HG_C2P_MIXED    hg_utoprim_batch's HG_C2P_MIXED path: FP32 Newton from
                     the reference's start, the polish to the converged root
                     of the same FP64 system and double-double primitives;
                     a zone it does not settle is inverted by HG_C2P_REFINED
                     from its guess in the same kernel;
     HG_C2P_LEARNED  HG_C2P_MIXED from the learned start of network m
                     (hg_utoprim_batch_learned's rule), the network
                     evaluated on the matrix cores before each substep's
                     inversion; m must outlive its use by the context.
   The failure codes are the reference's; the primitives of solved zones
   are the converged root (within ~1e-13 of it), not the reference's last
   iterate (up to ~1e-9 apart at Lorentz factors near 50).  Allocates the
   learned mode's buffers (36 bytes per zone) here, not in the step. */
#define HG_C2P_LEARNED 3
int hg_set_inversion(hg_ctx *ctx, int mode, const hg_mlp *m);

/* Several grids advanced by one HIP graph, no counterpart in the reference
   (which runs one grid per process): the graph holds nsteps RK2 steps of
   every context, one concurrent branch per context, and each launch advances
   all of them on the device (dt stays per context on the device).  Every
   grid evolves exactly as it would alone.  The contexts must outlive the
   graph and must not be stepped otherwise while a launch is in flight. */
typedef struct hg_graph hg_graph;
int hg_graph_create(hg_ctx *const *ctxs, int n, int nsteps, hg_graph **out);
/* The same with grouped launches, no counterpart in the reference: each
   kernel of a substep is one launch over the blocks of every context (each
   block reads its grid's arguments from a table on the device), so a step
   of all the grids takes ~40 launches however many grids there are, with
   every grid's arithmetic that of its own run, bit for bit.  For the
   standard scheme with WENO5BND and the LAXF or HLL flux, every context
   with the same timeorder and the same inversion mode (HG_C2P_FP64,
   HG_C2P_MIXED, or HG_C2P_LEARNED with one network for all), and no
   diagnostics recording; other contexts are rejected with an error
   (hg_graph_create takes them).  Launch shapes are
   the defaults, with each context's segment counts along j and k where its
   tune sets them; otherwise the library picks one count for all, enough
   wavefronts to fill the device (twice its wavefront slots). */
int hg_graph_create_grouped(hg_ctx *const *ctxs, int n, int nsteps, hg_graph **out);
int hg_graph_launch(hg_graph *g, void *stream);
int hg_graph_destroy(hg_graph *g);

/* Launch shapes of the step (no effect on results): block sizes of the flux,
   EMF, update and inversion kernels (64, 128 or 256), wavefronts per block of
   the i-direction reconstruction (1, 2 or 4), segments per line of the j
   and of the k reconstruction (0 = by line length), and the scheme along i
   (weno_xmarch 0: one lane per point with wavefront shuffles; n > 0: lanes
   across j and k marching along i in n segments).  hg_autotune times nsteps steps per
   candidate (coordinate descent) from the current state, restores the state
   (primitives, intermediate states, the finite-volume scheme's carried
   conserved state, unew, failure codes), dt and counters, keeps the fastest
   shapes and returns them in *best (may be NULL).  Diagnostics recording is
   paused during the trials.  No counterpart in the reference. */
typedef struct {
  int bs_flux, bs_emf, bs_update, bs_invert;
  int weno_xwaves, weno_nseg, weno_nseg3;
  int weno_xmarch;
  int concurrent; /* 1: the three directions' fluxes on three streams, 0: in turn */
} hg_tune;
int hg_get_tune(const hg_ctx *ctx, hg_tune *t);
int hg_set_tune(hg_ctx *ctx, const hg_tune *t);
int hg_autotune(hg_ctx *ctx, int nsteps, hg_tune *best, void *stream);

const char *hg_last_error(void);

/* Device self-check of the shared-reciprocal FP64 division the WENO5BND
   kernels use (a test hook, no counterpart in the reference): n generated
   operand sets from seed; result[0], [1], [2] count the mismatches against
   the hardware's own x / d of three numerators sharing one divisor, of 1 / d,
   and of four signed numerators of the divisor 12 (a zero quotient compared
   without its sign); result[3] counts the sets inside the shortcut's
   ranges.  n at most 2^32. */
int hg_check_division(int64_t n, uint64_t seed, int64_t result[4]);

/* Kernel-class timing, no counterpart in the reference (its timing is the
   wall clock of the whole step, diag.c): while a profile is open, every
   launch of hg_step/hg_step_async/hg_fluxcalc is bracketed by HIP events on
   its stream.  hg_profile(ctx, 1) opens a profile (clearing the previous
   one), hg_profile(ctx, 0) closes it; hg_profile_read waits for the events
   and returns, per class, the summed milliseconds and the interval count. */
#define HG_PROF_BOUND 0
#define HG_PROF_RECON1 1 /* WENO5BND interface states, direction 1 (2, 3 follow) */
#define HG_PROF_FLUX1 4  /* Riemann flux, direction 1 (MC: fused reconstruction) */
#define HG_PROF_EMF 7
#define HG_PROF_UPDATE 8 /* update + source + inversion + fixup1zone (+ dt) */
#define HG_PROF_FIXUP 9  /* post_fixup: neighbour check and average of failed zones */
#define HG_PROF_PREP 10  /* WENO5BND: the per-zone dP/P inputs the three directions share */
#define HG_PROF_NCLASS 11
int hg_profile(hg_ctx *ctx, int enable);
int hg_profile_read(hg_ctx *ctx, double ms[HG_PROF_NCLASS], int64_t launches[HG_PROF_NCLASS]);
const char *hg_profile_class_name(int k);

#ifdef __cplusplus
}
#endif
#endif
