/*
 * Golden-vector harness around the reference (pseudotensor/harm_finalwham),
 * linked against the reference's own object files (see build_ref.sh).
 *
 * This is test infrastructure only: it writes .npy files that pin the
 * project's oracle and its GPU kernels.  It never ships in the product.
 *
 * Modes
 *   torus <outdir> <scheme 0|1|2> <nsteps>
 *       Fishbone-Moncrief torus in MKS (LOGRSINTH) coordinates with the
 *       overrides of the problem's init_global/init_grid marked below, then
 *       writes geometry tables, the initial primitives, one call of
 *       fluxcalc() (which ends in flux_ct()), the slope_lim() interface
 *       states per direction, and the primitives after step 1 and after
 *       <nsteps> steps of step_ch().  Schemes: 0 MC + LaxF + RK2,
 *       1 WENO5BND + HLL + RK2, 2 PARA + HLL + RK4 (the torus file's own
 *       lim, fluxmethod and TIMEORDER, init.fishmon.c:110-114, with
 *       DOENOFLUX NOENOFLUX), 3 the same with its DOENOFLUX ENOFINITEVOLUME
 *       (and its avgscheme DONOR, init.fishmon.c:97): the torus file's own
 *       scheme; 4 MC + LaxF + TVD RK3, 5 MC + LaxF + Euler (the other
 *       TIMEORDERs of step_ch.c:442); 6 MC + FORCE, 7 PARA + MUSTA,
 *       8 WENO5BND + HLLLAXF1, each with RK2 (the other fluxmethods of
 *       fluxcompute.c:100); 9 WENO5BND + HLL + RK2 with the finite-volume
 *       step; 10 VANL + LaxF, 11 MINM + HLL, 12 DONOR + LaxF (RK2; the
 *       other three-point limiters).  Every mode's <weno> argument below is
 *       this scheme number.
 *   c2p <infile> <outfile> <gamma>
 *       Batched inversion: each input record holds a metric (gcov, gcon,
 *       gdet, gcovpert), the true primitives and a guess.  U is built with
 *       the reference's get_state()+primtoU() and inverted with its
 *       Utoprimgen() (UTOPRIM2DFINAL).
 *   time <weno 0|1> <nsteps>
 *       Times <nsteps> steps of the torus (IO excluded) and prints
 *       zone-updates/s; used as the reference CPU baseline.
 *   fixup <infile> <outdir> <weno 0|1> <finalstep>
 *       bound_prim + post_fixup + bound_prim on a given state and given
 *       inversion codes (see mode_fixup).
 *   diag <outdir> <weno 0|1> <nsteps>
 *       constotal, divbmaxavg and diag_flux after <nsteps> steps (mode_diag).
 *   long <outdir> <scheme> <nsteps>
 *       The dt and failure sequences over <nsteps> steps, then constotal()
 *       and divbmaxavg() of the final state and the final state itself
 *       (mode_long): the acceptance run of the full-size grids.
 *
 * Environment overrides of the set-up (absent: the values below):
 *   HR_A    black-hole spin a;   HR_RIN  inner radius of the grid, Rin.
 *
 * The initialisation below follows init() in initbase.c:5-176 call for call;
 * the only additions are the parameter overrides marked OVERRIDE.
 */
#include "defs.h"
#include <sys/stat.h>
#include <sys/time.h>

extern int pre_init(int argc, char *argv[]);
extern int pre_init_specific_init(void);
extern int init_defglobal(void);
extern int init_defconsts(void);
extern int init_global(void);
extern int init_defgrid(void);
extern int init_grid(void);
extern int init_grid_post_set_grid(void);
extern int post_par_set(void);
extern int post_init(void);
extern int post_init_specific_init(void);
extern int init_primitives(FTYPE p[][N2M][N3M][NPR]);
extern int set_dt(FTYPE (*prim)[N2M][N3M][NPR], SFTYPE *dt);
extern int init_conservatives(FTYPE p[][N2M][N3M][NPR], FTYPE Utemp[][N2M][N3M][NPR], FTYPE U[][N2M][N3M][NPR]);
extern int step_ch(void);
extern int fluxcalc(int stage, FTYPE pr[][N2M][N3M][NPR], FTYPE F1[][N2M][N3M][NPR],
                    FTYPE F2[][N2M][N3M][NPR], FTYPE F3[][N2M][N3M][NPR], FTYPE CUf,
                    FTYPE *ndt1, FTYPE *ndt2, FTYPE *ndt3);
extern void slope_lim(int numprims, int dir, int idel, int jdel, int kdel,
                      FTYPE (*primreal)[N2M][N3M][NPR], FTYPE (*p2interp)[N2M][N3M][NPR2INTERP],
                      FTYPE (*dq)[N2M][N3M][NPR2INTERP], FTYPE (*pleft)[N2M][N3M][NPR2INTERP],
                      FTYPE (*pright)[N2M][N3M][NPR2INTERP]);
extern int Utoprimgen(int evolvetype, int inputtype, FTYPE *U, struct of_geom *ptrgeom, FTYPE *pr);

/* ------------------------------------------------------------------ */
/* minimal .npy writer (format version 1.0, little-endian)            */
/* ------------------------------------------------------------------ */
static void write_npy(const char *dir, const char *name, const char *descr,
                      int ndim, const long *shape, const void *data, size_t elsize)
{
  char path[1024], hdr[512], shp[256];
  size_t n = 1, len, pad;
  int d;
  FILE *f;
  unsigned short hl;
  shp[0] = 0;
  for (d = 0; d < ndim; d++) {
    char tmp[32];
    n *= (size_t)shape[d];
    snprintf(tmp, sizeof tmp, "%ld%s", shape[d], (ndim == 1) ? "," : (d < ndim - 1 ? ", " : ""));
    strcat(shp, tmp);
  }
  snprintf(hdr, sizeof hdr, "{'descr': '%s', 'fortran_order': False, 'shape': (%s), }", descr, shp);
  len = strlen(hdr);
  pad = 64 - ((10 + len + 1) % 64);
  if (pad == 64) pad = 0;
  memset(hdr + len, ' ', pad);
  hdr[len + pad] = '\n';
  hdr[len + pad + 1] = 0;
  hl = (unsigned short)(len + pad + 1);
  snprintf(path, sizeof path, "%s/%s.npy", dir, name);
  f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  fwrite("\x93NUMPY\x01\x00", 1, 8, f);
  fwrite(&hl, 2, 1, f);
  fwrite(hdr, 1, hl, f);
  fwrite(data, elsize, n, f);
  fclose(f);
}

/* copy of the full primitive array including ghost zones: [N1M][N2M][N3M][NPR] */
static void dump_prims(const char *dir, const char *name, FTYPE (*pr)[N2M][N3M][NPR])
{
  long shape[4] = {N1M, N2M, N3M, NPR};
  /* pr points at the first active zone; the allocation starts N?BND earlier */
  FTYPE *base = &pr[-N1BND][-N2BND][-N3BND][0];
  write_npy(dir, name, "<f8", 4, shape, base, sizeof(FTYPE));
}

static void dump_flux(const char *dir, const char *name, FTYPE (*F)[N2M][N3M][NPR])
{
  long shape[4] = {N1M, N2M, N3M, NPR};
  FTYPE *base = &F[-N1BND][-N2BND][-N3BND][0];
  write_npy(dir, name, "<f8", 4, shape, base, sizeof(FTYPE));
}

/* ------------------------------------------------------------------ */
static void harness_init(int scheme)
{
  int use_weno = (scheme == 1);
  char *fake_argv[] = {"harness", NULL};

  fprintf(stderr, "harness init: N=%dx%dx%d MAXBND=%d scheme=%d\n", N1, N2, N3, MAXBND, scheme);
  pre_init(1, fake_argv);
  pre_init_specific_init();

  init_defglobal();
  init_defconsts();
  init_global();
  /* OVERRIDE: the hot-path configurations (init.fishmon.c:96-118 picks PARA/RK4/FV) */
  lim = (scheme == 2 || scheme == 3 || scheme == 7) ? PARA : (use_weno || scheme == 8 || scheme == 9) ? WENO5BND
      : (scheme == 10) ? VANL : (scheme == 11) ? MINM : (scheme == 12) ? DONOR : MC;
  fluxmethod = (scheme == 6) ? FORCEFLUX : (scheme == 7) ? MUSTAFLUX : (scheme == 8) ? HLLLAXF1FLUX
             : (scheme == 0 || scheme == 4 || scheme == 5 || scheme == 10 || scheme == 12) ? LAXFFLUX : HLLFLUX;
  FLUXB = FLUXCTTOTH;
  TIMEORDER = (scheme == 2 || scheme == 3) ? 4 : (scheme == 4) ? 3 : (scheme == 5) ? 1 : 2;
  DOENOFLUX = (scheme == 3 || scheme == 9) ? ENOFINITEVOLUME : NOENOFLUX;
  cour = 0.8;
  rescaletype = 0; /* floors rho>=RHOMIN r^-1.5, u>=UUMIN r^-2.5 (fixup.c:1562) */
  DODIAGS = 0;

  init_defgrid();
  init_grid();
  /* OVERRIDE: log-r / sin-theta MKS grid, r in [1.3,40], hslope 0.3 (init.fishmon.c:54-76 uses defcoord 9) */
  a = getenv("HR_A") ? atof(getenv("HR_A")) : 0.9375;
  R0 = 0.0;
  Rin = getenv("HR_RIN") ? atof(getenv("HR_RIN")) : 1.3;
  Rout = 40.0;
  hslope = 0.3;
  defcoord = LOGRSINTH;

  set_coord_parms();
  set_grid();
  init_grid_post_set_grid();
  post_par_set();

  if (init_primitives(p) >= 1) { fprintf(stderr, "init_primitives failed\n"); exit(3); }

#if (FIXUPAFTERINIT)
  if (fixup(STAGEM1, p, 0) >= 1) { fprintf(stderr, "fixup failed\n"); exit(3); }
#endif
  if (bound_prim(STAGEM1, p) >= 1) { fprintf(stderr, "bound_prim failed\n"); exit(3); }
  if (pre_fixup(STAGEM1, p) >= 1) { fprintf(stderr, "pre_fixup failed\n"); exit(3); }
  set_dt(p, &dt);

  post_init();
  post_init_specific_init();
  /* OVERRIDE: Noble 2D inversion (post_init_specific_init sets UTOPRIMJONNONRELCOMPAT) */
  UTOPRIMVERSION = UTOPRIM2DFINAL;

  if (DOENOFLUXMEMORY) init_conservatives(p, ulast, uinitial);
  else init_conservatives(p, ulast, unew);
  fprintf(stderr, "harness init done: dt=%.17g\n", dt);
  /* dualfprintf()/trifprintf() (mpi_fileio.c:1978-2036) hand one va_list to
     vfprintf twice when myid==0; the second pass reads consumed arguments and
     crashes on "%s" formats in the inversion report (utoprimgen.c:420).  The
     hot path never reads myid (only diagnostics do), so the second, stderr,
     pass is switched off here; the log files still get every message. */
  myid = 1;
}

static long count_failed(void)
{
  int i, j, k;
  long n = 0;
  ZLOOP { if (pflag[i][j][k][FLAGUTOPRIMFAIL] != 0) n++; }
  return n;
}

/* geometry at k=0 for all (i,j) including ghosts; it is phi-independent */
static void dump_geometry(const char *dir)
{
  int i, j, k, loc, a1, b1, c1;
  long ni = N1M + SHIFT1, nj = N2M + SHIFT2;
  long shp_g[5] = {ni, nj, NPG, NDIM, NDIM};
  long shp_p[4] = {ni, nj, NPG, NDIM};
  long shp_d[3] = {ni, nj, NPG};
  long shp_c[5] = {N1M, N2M, NDIM, NDIM, NDIM};
  long shp_x[4] = {N1M, N2M, 2, NDIM}; /* X and V at CENT */
  FTYPE *gco = malloc(sizeof(FTYPE) * ni * nj * NPG * 16);
  FTYPE *gcn = malloc(sizeof(FTYPE) * ni * nj * NPG * 16);
  FTYPE *gpt = malloc(sizeof(FTYPE) * ni * nj * NPG * 4);
  FTYPE *gdt = malloc(sizeof(FTYPE) * ni * nj * NPG);
  FTYPE *cnn = malloc(sizeof(FTYPE) * N1M * N2M * 64);
  FTYPE *xv = malloc(sizeof(FTYPE) * N1M * N2M * 8);
  double maxkdiff = 0.0;
  long idx;

  for (i = -N1BND; i < N1 + N1BND + SHIFT1; i++)
    for (j = -N2BND; j < N2 + N2BND + SHIFT2; j++)
      for (loc = 0; loc < NPG; loc++) {
        idx = ((long)(i + N1BND) * nj + (j + N2BND)) * NPG + loc;
        for (a1 = 0; a1 < 4; a1++) {
          for (b1 = 0; b1 < 4; b1++) {
            gco[idx * 16 + a1 * 4 + b1] = gcov[i][j][0][loc][a1][b1];
            gcn[idx * 16 + a1 * 4 + b1] = gcon[i][j][0][loc][a1][b1];
          }
          gpt[idx * 4 + a1] = gcovpert[i][j][0][loc][a1];
        }
        gdt[idx] = gdet[i][j][0][loc];
        /* check phi independence of the stored tables */
        for (k = 1; k < N3; k++) {
          double d = fabs(gdet[i][j][k][loc] - gdet[i][j][0][loc]);
          for (a1 = 0; a1 < 4; a1++)
            for (b1 = 0; b1 < 4; b1++)
              d += fabs(gcon[i][j][k][loc][a1][b1] - gcon[i][j][0][loc][a1][b1]);
          if (d > maxkdiff) maxkdiff = d;
        }
      }
  for (i = -N1BND; i < N1 + N1BND; i++)
    for (j = -N2BND; j < N2 + N2BND; j++) {
      FTYPE X[NDIM], V[NDIM];
      idx = (long)(i + N1BND) * N2M + (j + N2BND);
      for (a1 = 0; a1 < 4; a1++)
        for (b1 = 0; b1 < 4; b1++)
          for (c1 = 0; c1 < 4; c1++)
            cnn[idx * 64 + a1 * 16 + b1 * 4 + c1] = conn[i][j][0][a1][b1][c1];
      coord(i, j, 0, CENT, X);
      bl_coord(X, V);
      for (a1 = 0; a1 < 4; a1++) { xv[idx * 8 + a1] = X[a1]; xv[idx * 8 + 4 + a1] = V[a1]; }
      for (k = 1; k < N3; k++)
        for (a1 = 0; a1 < 4; a1++)
          for (b1 = 0; b1 < 4; b1++)
            for (c1 = 0; c1 < 4; c1++) {
              double d = fabs(conn[i][j][k][a1][b1][c1] - conn[i][j][0][a1][b1][c1]);
              if (d > maxkdiff) maxkdiff = d;
            }
    }
  write_npy(dir, "gcov", "<f8", 5, shp_g, gco, sizeof(FTYPE));
  write_npy(dir, "gcon", "<f8", 5, shp_g, gcn, sizeof(FTYPE));
  write_npy(dir, "gcovpert", "<f8", 4, shp_p, gpt, sizeof(FTYPE));
  write_npy(dir, "gdet", "<f8", 3, shp_d, gdt, sizeof(FTYPE));
  write_npy(dir, "conn", "<f8", 5, shp_c, cnn, sizeof(FTYPE));
  write_npy(dir, "xv_cent", "<f8", 4, shp_x, xv, sizeof(FTYPE));
  fprintf(stderr, "geometry: max |table(k)-table(0)| = %g\n", maxkdiff);
  free(gco); free(gcn); free(gpt); free(gdt); free(cnn); free(xv);
}

static int mode_torus(const char *outdir, int use_weno, int nsteps)
{
  FTYPE ndt1, ndt2, ndt3;
  FTYPE *dts;
  long *fails;
  long shp1[1];
  int dir, n;
  FTYPE params[16];
  long shp_par[1] = {16};
  FTYPE (*pl_save)[N2M][N3M][NPR] = malloc(sizeof(FTYPE) * N1M * N2M * N3M * NPR);
  FTYPE (*pr_save)[N2M][N3M][NPR] = malloc(sizeof(FTYPE) * N1M * N2M * N3M * NPR);

  mkdir(outdir, 0755);
  harness_init(use_weno);

  params[0] = N1; params[1] = N2; params[2] = N3; params[3] = MAXBND;
  params[4] = a; params[5] = gam; params[6] = Rin; params[7] = Rout;
  params[8] = hslope; params[9] = R0; params[10] = cour; params[11] = dt;
  params[12] = dx[1]; params[13] = dx[2]; params[14] = dx[3]; params[15] = startx[1];
  write_npy(outdir, "params", "<f8", 1, shp_par, params, sizeof(FTYPE));
  {
    FTYPE sx[4] = {startx[0], startx[1], startx[2], startx[3]};
    long s4[1] = {4};
    write_npy(outdir, "startx", "<f8", 1, s4, sx, sizeof(FTYPE));
  }

  dump_geometry(outdir);
  dump_prims(outdir, "p_init", p);

  /* per-direction interface states of slope_lim() on the initial data */
  for (dir = 1; dir <= 3; dir++) {
    int idel = (dir == 1), jdel = (dir == 2), kdel = (dir == 3);
    char nm[64];
    FTYPE (*dqd)[N2M][N3M][NPR2INTERP] = (dir == 1) ? dq1 : (dir == 2) ? dq2 : dq3;
    if ((dir == 1 && N1 == 1) || (dir == 2 && N2 == 1) || (dir == 3 && N3 == 1)) continue;
    fprintf(stderr, "slope_lim dir %d\n", dir);
    slope_lim(NPR2INTERP, dir, idel, jdel, kdel, p, p, dqd, pleft, pright);
    if (lim >= WENO3 || lim == PARA) {
      snprintf(nm, sizeof nm, "pleft_dir%d", dir);
      dump_flux(outdir, nm, (FTYPE (*)[N2M][N3M][NPR])pleft);
      snprintf(nm, sizeof nm, "pright_dir%d", dir);
      dump_flux(outdir, nm, (FTYPE (*)[N2M][N3M][NPR])pright);
    } else {
      snprintf(nm, sizeof nm, "dq_dir%d", dir);
      dump_flux(outdir, nm, (FTYPE (*)[N2M][N3M][NPR])dqd);
    }
  }
  (void)pl_save; (void)pr_save;

  /* one fluxcalc (+flux_ct) on the initial data, first-substep coefficient CUf[2]=0.5 */
  {
    int i, j, k, pl;
    /* F1..F3 are only allocated for directions with more than one zone (set_arrays.c:58-66) */
    FULLLOOP PLOOP(pl) {
      if (N1 > 1) F1[i][j][k][pl] = 0;
      if (N2 > 1) F2[i][j][k][pl] = 0;
      if (N3 > 1) F3[i][j][k][pl] = 0;
    }
    fprintf(stderr, "fluxcalc on initial data\n");
    fluxcalc(STAGEM1, p, F1, F2, F3, 0.5, &ndt1, &ndt2, &ndt3);
    fprintf(stderr, "fluxcalc done\n");
    if (N1 > 1) dump_flux(outdir, "F1_init", F1);
    if (N2 > 1) dump_flux(outdir, "F2_init", F2);
    if (N3 > 1) dump_flux(outdir, "F3_init", F3);
    {
      FTYPE nd[3] = {ndt1, ndt2, ndt3};
      long s3[1] = {3};
      write_npy(outdir, "ndt_init", "<f8", 1, s3, nd, sizeof(FTYPE));
    }
  }

  dts = malloc(sizeof(FTYPE) * (nsteps + 1));
  fails = malloc(sizeof(long) * (nsteps + 1));
  dts[0] = dt;
  fails[0] = 0;
  for (n = 1; n <= nsteps; n++) {
    nstroke = 0;
    step_ch();
    nstep++;
    dts[n] = dt; /* dt to be used by step n+1 */
    fails[n] = count_failed();
    if (n == 1) dump_prims(outdir, "p_step1", p);
    fprintf(stderr, "step %d t=%.17g dt=%.17g failed=%ld iters/zone=%g\n", n, t, dt, fails[n],
            (double)nstroke / (2.0 * N1 * N2 * N3));
  }
  dump_prims(outdir, "p_final", p);
  shp1[0] = nsteps + 1;
  write_npy(outdir, "dt_seq", "<f8", 1, shp1, dts, sizeof(FTYPE));
  write_npy(outdir, "fail_seq", "<i8", 1, shp1, fails, sizeof(long));
  return 0;
}

static int mode_time(int use_weno, int nsteps)
{
  struct timeval t0, t1;
  double sec;
  int n;
  harness_init(use_weno);
  gettimeofday(&t0, NULL);
  for (n = 0; n < nsteps; n++) {
    nstroke = 0;
    step_ch();
    nstep++;
  }
  gettimeofday(&t1, NULL);
  sec = (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec);
  printf("{\"zones\": %ld, \"steps\": %d, \"seconds\": %.6f, \"zone_updates_per_s\": %.6e}\n",
         (long)N1 * N2 * N3, nsteps, sec, (double)N1 * N2 * N3 * nsteps / sec);
  return 0;
}

/*
 * c2p record layout (doubles): gcov[16] gcon[16] gdet gcovpert[4] p_true[8] p_guess[8]  = 53
 * output record (doubles):     U[8] p_out[8] pflag iters                                = 18
 */
#define C2P_IN 53
#define C2P_OUT 18
static int mode_c2p(const char *infile, const char *outfile, double gamma_eos)
{
  FILE *fi = fopen(infile, "rb"), *fo = fopen(outfile, "wb");
  double rec[C2P_IN], out[C2P_OUT];
  FTYPE gco[NDIM][NDIM], gcn[NDIM][NDIM], gpert[NDIM];
  struct of_geom geom;
  struct of_state q;
  FTYPE U[NPR], pr[NPR];
  long n = 0;
  int a1, b1, pl;

  if (!fi || !fo) { fprintf(stderr, "c2p: cannot open files\n"); return 2; }
  harness_init(0);
  gam = gamma_eos;
  while (fread(rec, sizeof(double), C2P_IN, fi) == C2P_IN) {
    for (a1 = 0; a1 < 4; a1++) {
      for (b1 = 0; b1 < 4; b1++) { gco[a1][b1] = rec[a1 * 4 + b1]; gcn[a1][b1] = rec[16 + a1 * 4 + b1]; }
      gpert[a1] = rec[33 + a1];
    }
    geom.gcov = gco;
    geom.gcon = gcn;
    geom.gcovpert = gpert;
    geom.g = rec[32];
    PLOOP(pl) geom.e[pl] = geom.g;
    geom.i = 0; geom.j = 0; geom.k = 0; geom.p = CENT;
    PLOOP(pl) pr[pl] = rec[37 + pl];
    get_state(pr, &geom, &q);
    primtoU(UEVOLVE, pr, &q, &geom, U);
    PLOOP(pl) pr[pl] = rec[45 + pl];
    Utoprimgen(EVOLVEUTOPRIM, UEVOLVE, U, &geom, pr);
    PLOOP(pl) { out[pl] = U[pl]; out[8 + pl] = pr[pl]; }
    out[16] = pflag[0][0][0][FLAGUTOPRIMFAIL];
    out[17] = lntries;
    fwrite(out, sizeof(double), C2P_OUT, fo);
    n++;
  }
  fclose(fi);
  fclose(fo);
  fprintf(stderr, "c2p: %ld records\n", n);
  return 0;
}

/*
 * fixup mode: the primitives p[N1M][N2M][N3M][NPR] (ghosts included) and the
 * inversion codes of the active zones (int32 [N1][N2][N3]) are read from
 * <infile>; then bound_prim, post_fixup (fixup.c:37) and bound_prim run as
 * in step_ch.c:132-157, and p and the full flag array are written.
 */
extern int post_fixup(int stageit, FTYPE (*pv)[N2M][N3M][NPR], FTYPE (*pbackup)[N2M][N3M][NPR], int finalstep);
static int mode_fixup(const char *infile, const char *outdir, int use_weno, int finalstep)
{
  FILE *fi = fopen(infile, "rb");
  int i, j, k, pl;
  size_t np = (size_t)N1M * N2M * N3M * NPR;
  int *fl = malloc(sizeof(int) * N1 * N2 * N3);
  int *flout = malloc(sizeof(int) * N1M * N2M * N3M);
  long shp_f[3] = {N1M, N2M, N3M};
  if (!fi) { fprintf(stderr, "fixup: cannot open %s\n", infile); return 2; }
  mkdir(outdir, 0755);
  harness_init(use_weno);
  if (fread(&p[-N1BND][-N2BND][-N3BND][0], sizeof(FTYPE), np, fi) != np) { fprintf(stderr, "fixup: short p\n"); return 2; }
  if (fread(fl, sizeof(int), (size_t)N1 * N2 * N3, fi) != (size_t)N1 * N2 * N3) { fprintf(stderr, "fixup: short flags\n"); return 2; }
  fclose(fi);
  FULLLOOP pflag[i][j][k][FLAGUTOPRIMFAIL] = 0;
  ZLOOP pflag[i][j][k][FLAGUTOPRIMFAIL] = fl[((long)i * N2 + j) * N3 + k];
  if (bound_prim(STAGEM1, p) >= 1) { fprintf(stderr, "bound_prim failed\n"); return 3; }
  dump_prims(outdir, "p_in_bounded", p);
  if (post_fixup(STAGEM1, p, p, finalstep) >= 1) { fprintf(stderr, "post_fixup failed\n"); return 3; }
  if (bound_prim(STAGEM1, p) >= 1) { fprintf(stderr, "bound_prim failed\n"); return 3; }
  dump_prims(outdir, "p_out", p);
  FULLLOOP flout[((long)(i + N1BND) * N2M + (j + N2BND)) * N3M + (k + N3BND)] = pflag[i][j][k][FLAGUTOPRIMFAIL];
  write_npy(outdir, "pflag_out", "<i4", 3, shp_f, flout, sizeof(int));
  (void)pl;
  free(fl);
  free(flout);
  return 0;
}

/*
 * diag mode: the torus over <nsteps> RK2 steps (step_ch() calls diag_flux()
 * after each step, step_ch.c:221), then constotal() and divbmaxavg() of the
 * final state; writes them with the global region's pdot, pdotterms,
 * pdottermsjet2 and pcum (diag.c:734) and the final state.
 */
extern int constotal(int enerregion, SFTYPE *vars);
extern void divbmaxavg(FTYPE p[][N2M][N3M][NPR], SFTYPE *ptrdivbmax, SFTYPE *ptrdivbavg);
static int mode_diag(const char *outdir, int use_weno, int nsteps)
{
  SFTYPE vars[NPR], dmax, davg, dv[2];
  long s1[1] = {NPR}, s2[1] = {2}, sp[2] = {COMPDIM * 2, NPR}, st[3] = {COMPDIM * 2, NUMFLUXTERMS, NPR};
  long shp1[1];
  FTYPE *dts = malloc(sizeof(FTYPE) * (nsteps + 1));
  int n;
  mkdir(outdir, 0755);
  harness_init(use_weno);
  dts[0] = dt;
  for (n = 1; n <= nsteps; n++) {
    step_ch();
    nstep++;
    dts[n] = dt;
  }
  if (constotal(0, vars) >= 1) { fprintf(stderr, "constotal failed\n"); return 3; }
  divbmaxavg(p, &dmax, &davg);
  dv[0] = dmax;
  dv[1] = davg;
  write_npy(outdir, "cons", "<f8", 1, s1, vars, sizeof(SFTYPE));
  write_npy(outdir, "divb", "<f8", 1, s2, dv, sizeof(SFTYPE));
  write_npy(outdir, "pdot", "<f8", 2, sp, &pdotreg[0][0][0], sizeof(SFTYPE));
  write_npy(outdir, "pdotterms", "<f8", 3, st, &pdottermsreg[0][0][0][0], sizeof(SFTYPE));
  write_npy(outdir, "pdottermsjet2", "<f8", 3, st, &pdottermsjet2[0][0][0], sizeof(SFTYPE));
  write_npy(outdir, "pcum", "<f8", 2, sp, &pcumreg[0][0][0], sizeof(SFTYPE));
  shp1[0] = nsteps + 1;
  write_npy(outdir, "dt_seq", "<f8", 1, shp1, dts, sizeof(FTYPE));
  dump_prims(outdir, "p_final", p);
  free(dts);
  return 0;
}

/*
 * long mode: <nsteps> steps of the torus in the given scheme, recording dt
 * and the number of failed inversions after every step (as mode_torus),
 * then constotal() and divbmaxavg() of the final state (as mode_diag) and
 * the final state.
 */
static int mode_long(const char *outdir, int scheme, int nsteps)
{
  SFTYPE vars[NPR], dmax, davg, dv[2];
  long s1[1] = {NPR}, s2[1] = {2}, shp1[1];
  FTYPE *dts = malloc(sizeof(FTYPE) * (nsteps + 1));
  long *fails = malloc(sizeof(long) * (nsteps + 1));
  int n;
  mkdir(outdir, 0755);
  harness_init(scheme);
  dump_prims(outdir, "p_init", p);
  dts[0] = dt;
  fails[0] = 0;
  for (n = 1; n <= nsteps; n++) {
    nstroke = 0;
    step_ch();
    nstep++;
    dts[n] = dt;
    fails[n] = count_failed();
    fprintf(stderr, "step %d t=%.17g dt=%.17g failed=%ld\n", n, t, dt, fails[n]);
  }
  if (constotal(0, vars) >= 1) { fprintf(stderr, "constotal failed\n"); return 3; }
  divbmaxavg(p, &dmax, &davg);
  dv[0] = dmax;
  dv[1] = davg;
  write_npy(outdir, "cons", "<f8", 1, s1, vars, sizeof(SFTYPE));
  write_npy(outdir, "divb", "<f8", 1, s2, dv, sizeof(SFTYPE));
  shp1[0] = nsteps + 1;
  write_npy(outdir, "dt_seq", "<f8", 1, shp1, dts, sizeof(FTYPE));
  write_npy(outdir, "fail_seq", "<i8", 1, shp1, fails, sizeof(long));
  dump_prims(outdir, "p_final", p);
  free(dts);
  free(fails);
  return 0;
}

/*
 * rdump mode: the torus over <nsteps> RK2 steps, then restart_write(0)
 * (restart.c:206), which writes dumps/rdump-0 in the working directory, and
 * the state as .npy (p with ghosts, unew of the active zones).
 *
 * rdread mode: restart_read(0) (restart.c:262) of dumps/rdump-0 in the
 * working directory after the usual set-up; writes p, unew and the header
 * values the hot path uses.
 */
extern int restart_write(long dump_cnt);
extern int restart_read(long dump_cnt);
static void dump_unew(const char *dir, const char *name)
{
  long shape[4] = {N1, N2, N3, NPR};
  FTYPE *buf = malloc(sizeof(FTYPE) * N1 * N2 * N3 * NPR);
  int i, j, k, pl;
  ZLOOP PLOOP(pl) buf[(((long)i * N2 + j) * N3 + k) * NPR + pl] = unew[i][j][k][pl];
  write_npy(dir, name, "<f8", 4, shape, buf, sizeof(FTYPE));
  free(buf);
}
static void dump_header_values(const char *dir)
{
  FTYPE h[16];
  long s[1] = {16};
  h[0] = t; h[1] = tf; h[2] = (FTYPE)nstep; h[3] = a; h[4] = gam; h[5] = cour; h[6] = dt;
  h[7] = lim; h[8] = fluxmethod; h[9] = R0; h[10] = Rin; h[11] = Rout; h[12] = hslope;
  h[13] = RHOMIN; h[14] = UUMIN; h[15] = GAMMAMAX;
  write_npy(dir, "header", "<f8", 1, s, h, sizeof(FTYPE));
}
static int mode_rdump(const char *outdir, int use_weno, int nsteps)
{
  int n;
  mkdir(outdir, 0755);
  harness_init(use_weno);
  for (n = 1; n <= nsteps; n++) {
    step_ch();
    nstep++;
    t += 0.0;
  }
  mkdir("dumps", 0755);
  if (restart_write(0) >= 1) { fprintf(stderr, "restart_write failed\n"); return 3; }
  dump_prims(outdir, "p", p);
  dump_unew(outdir, "unew");
  dump_header_values(outdir);
  return 0;
}
static int mode_rdread(const char *outdir, int use_weno)
{
  int i, j, k, pl;
  mkdir(outdir, 0755);
  harness_init(use_weno);
  FULLLOOP PLOOP(pl) p[i][j][k][pl] = 0.0;
  if (restart_read(0) >= 1) { fprintf(stderr, "restart_read failed\n"); return 3; }
  dump_prims(outdir, "p", p);
  dump_unew(outdir, "unew");
  dump_header_values(outdir);
  return 0;
}

int main(int argc, char **argv)
{
  if (argc >= 5 && !strcmp(argv[1], "torus")) return mode_torus(argv[2], atoi(argv[3]), atoi(argv[4]));
  if (argc >= 5 && !strcmp(argv[1], "c2p")) return mode_c2p(argv[2], argv[3], atof(argv[4]));
  if (argc >= 4 && !strcmp(argv[1], "time")) return mode_time(atoi(argv[2]), atoi(argv[3]));
  if (argc >= 6 && !strcmp(argv[1], "fixup")) return mode_fixup(argv[2], argv[3], atoi(argv[4]), atoi(argv[5]));
  if (argc >= 5 && !strcmp(argv[1], "diag")) return mode_diag(argv[2], atoi(argv[3]), atoi(argv[4]));
  if (argc >= 5 && !strcmp(argv[1], "rdump")) return mode_rdump(argv[2], atoi(argv[3]), atoi(argv[4]));
  if (argc >= 4 && !strcmp(argv[1], "rdread")) return mode_rdread(argv[2], atoi(argv[3]));
  if (argc >= 5 && !strcmp(argv[1], "long")) return mode_long(argv[2], atoi(argv[3]), atoi(argv[4]));
  fprintf(stderr, "usage: %s torus <outdir> <weno> <nsteps> | c2p <in> <out> <gamma> | time <weno> <nsteps> | fixup <in> <outdir> <weno> <finalstep>\n", argv[0]);
  return 1;
}
