/*
 * harm_oracle.h -- CPU restatement of the per-step GRMHD hot path of
 * pseudotensor/harm_finalwham (reconstruction, Riemann fluxes, flux-CT,
 * conserved update, Noble 2D con2prim, fixup1zone, ghost fill).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (harm_finalwham_amd/) never links, calls or includes it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * FP64 and written in the reference's operation order; the library is built
 * with -O2 -ffp-contract=off so that no FMA contraction changes rounding.
 *
 * Pinning: every function here is checked bit for bit against output of the
 * reference's own code, compiled from its sources into oracle/_ref and run
 * on the same inputs (tests/golden/make_golden.py writes those fixtures;
 * the CPU tests under tests/ check them, and planted mistakes in
 * test_oracle_golden.py, test_fixup_golden.py and test_rdump.py show that
 * the checks can fail).  Parity unpinned: none.
 *
 * Conventions (same as the reference, global.h:916-923):
 *   primitives  p = [rho, u, U1, U2, U3, B1, B2, B3]  (U^i relative 4-velocity)
 *   conserved   U = gdet * [rho u^t, T^t_t + rho u^t, T^t_i, B^i]
 *                (REMOVERESTMASSFROMUU==2, WHICHEOM==WITHGDET)
 */
#ifndef HARM_ORACLE_H
#define HARM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NPR 8
#define OR_RHO 0
#define OR_UU 1
#define OR_U1 2
#define OR_U2 3
#define OR_U3 4
#define OR_B1 5
#define OR_B2 6
#define OR_B3 7

/* geometry locations, global.h:946-953 */
#define OR_FACE1 0
#define OR_FACE2 1
#define OR_FACE3 2
#define OR_CORN1 3
#define OR_CORN2 4
#define OR_CORN3 5
#define OR_CORNT 6
#define OR_CENT 7
#define OR_NPG 8

/* failure codes written by the inversion, global.h:782-799 */
#define OR_FAIL_NONE 0
#define OR_FAIL_CONVW 2
#define OR_FAIL_CONVUTSQ 3
#define OR_FAIL_CONVGUESSUTSQ 4
#define OR_FAIL_CONVRET 101
#define OR_FAIL_RHONEG 9
#define OR_FAIL_UNEG 10
#define OR_FAIL_RHOUNEG 11

typedef struct {
  double gcov[4][4];
  double gcon[4][4];
  double gcovpert[4];
  double g; /* sqrt(-det g); also the EOM factor for every variable */
} or_geom;

typedef struct {
  double ucon[4], ucov[4], bcon[4], bcov[4];
} or_state;

/* ---- point physics (phys.c, eos.c, vchar.c) ---- */
int or_get_state(const double *pr, const or_geom *g, or_state *q);
void or_primtoflux(const double *pr, const or_state *q, int dir, const or_geom *g, double gam,
                   double *flux);
void or_primtoU(const double *pr, const or_geom *g, double gam, double *U);
int or_vchar(const double *pr, const or_state *q, int dir, const or_geom *g, double gam,
             double *vmax, double *vmin);
void or_source(const double *pr, const or_geom *g, const double conn[4][4][4], double gam,
               double *dU);

/* ---- con2prim (utoprimgen.c:23 + utoprim_2d_final.c) ---- */
/* U in evolved form (with gdet); pr in: guess, out: solution. */
int or_utoprimgen(const double *U, const or_geom *g, double gam, double *pr,
                  int *pflag, int *iters);

/* ---- fixup1zone (fixup.c:341) ---- */
/* returns 1 when a floor or the gamma limit changed pr */
int or_fixup1zone(double *pr, const or_geom *g, double rhofloor, double ufloor, double gammamax);
/* fixup.c:181 diag_fixup's update of the finite-volume conserved state */
void or_diag_fixup_fv(const double *pr0, const double *pr, const or_geom *g, double gam, double *u);
int or_limit_gamma(double gammamax, double *pr, const or_geom *g);

/* ---- batch helpers exported for tests ---- */
void or_c2p_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                  const double *gcovpert, const double *U, double gam, double *pr_inout,
                  int32_t *pflag, int32_t *iters);
/* the converged root of the reference's FP64 Noble 2D system in long double
   (or_c2p_root.c) */
void or_c2p_converged_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                            const double *gcovpert, const double *U, double gam, double *pr_inout,
                            int32_t *pflag, int start, const double *start_prims, double *Wv, double *resid);
void or_primtoU_batch(int64_t n, const double *gcov, const double *gcon, const double *gdet,
                      const double *gcovpert, const double *pr, double gam, double *U);
int or_inflow_check_r(double *pr, const or_geom *g, int side, double gammamax);
void or_floors(int64_t n, const double *r, double rhomin, double uumin, double *out);

/* ---- grid-level step (advance.c, flux.c, fluxct.c, step_ch.c, bounds.fishmon.c) ---- */
typedef struct {
  int n1, n2, n3, nb;       /* active zones per direction and ghost width */
  int lim;                  /* 0 = three-point limiter (slim), 1 = WENO5BND, 2 = PARA (para4) */
  int flux;                 /* 0 = LAXF, 1 = HLL, 2 = FORCE, 3 = MUSTA, 4 = HLLLAXF1 */
  double gam, cour, dx[4];
  double rhomin, uumin, gammamax;
  double defcon, safe;
  /* geometry per (i,j) with ghosts, phi independent:
     gcov [(n1+2nb+1)*(n2+2nb+1)][8][16], gcon same, gcovpert [..][8][4], gdet [..][8]
     conn [(n1+2nb)*(n2+2nb)][64], rfloor [(n1+2nb)*(n2+2nb)][2] (rho, u floors at CENT) */
  const double *gcov, *gcon, *gcovpert, *gdet, *conn, *floors;
  int timeorder;            /* 1 Euler, 2 (or 0) midpoint RK2, 3 TVD RK3, 4 RK4 (step_ch.c:442) */
  int fv;                   /* 1 = ENOFINITEVOLUME with avgscheme DONOR (advance.c:245) */
  int slim;                 /* lim 0: the three-point limiter, DONOR 0, VANL 1, MINM 2 or MC (else) */
} or_grid;

/* one step on p (full array with ghosts, layout [i][j][k][pl] with
   dims (n1+2nb)x(n2+2nb)x(n3+2nb3)); returns the next dt in *dt_io.
   Finite-volume grids (fv) use or_step_fv. */
int or_step(const or_grid *G, double *p, double *dt_io, int64_t *nfail, int64_t *iters);
/* one finite-volume step: ucons (layout of p, active zones used) is the
   conserved state carried between steps, replaced by the step's ucum */
/* or_step that also returns the step's unew (the last substep's ucum, zero
   outside the active zones), as restart_write (restart.c:240) dumps it */
int or_step_unew(const or_grid *G, double *p, double *unew, double *dt_io, int64_t *nfail, int64_t *iters);
int or_step_fv(const or_grid *G, double *p, double *ucons, double *dt_io, int64_t *nfail, int64_t *iters);
/* initbase.c:18 init_conservatives: U(p) of the active zones into ucons */
void or_cons_from_prims(const or_grid *G, const double *p, double *ucons);
/* fluxcalc + flux_ct on p for diagnostics; F arrays same layout as p */
int or_fluxcalc(const or_grid *G, const double *p, double *F1, double *F2, double *F3,
                double *ndt);
int or_bound_prim(const or_grid *G, double *p);
/* post_fixup (fixup.c:37): ghost fill of the failure flags, the
   check of good solutions against their neighbours and the neighbour
   average of failed zones, on p (bounded) and pflag (full array of int,
   layout of p without the pl index: inversion codes in the active zones) */
int or_post_fixup(const or_grid *G, double *p, int *pflag, int finalstep);
/* the same on a finite-volume grid: fixed zones also move ucons (fixup.c:1428) */
int or_post_fixup_fv(const or_grid *G, double *p, int *pflag, int finalstep, double *ucons);
void or_bound_pflag(const or_grid *G, int *pflag);
/* interface states of the reconstruction along dir for the whole grid */
int or_reconstruct(const or_grid *G, const double *p, int dir, double *pleft, double *pright);

#ifdef __cplusplus
}
#endif
#endif
