// Host side of the C ABI (include/harmgpu.h): context, device memory,
// layout conversion and the RK2 driver that plays the role of step_ch.c.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "harmgpu.h"
#include "hg_grid.h"
#include "hg_host.h"
#include "hg_kernels.h"

using namespace hg;

static thread_local std::string g_last_error;

void hg::set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

extern "C" const char *hg_last_error(void) { return g_last_error.c_str(); }

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      hg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

static int dalloc(double **p, size_t n) {
  HIPCHECK(hipMalloc((void **)p, n * sizeof(double)));
  HIPCHECK(hipMemset(*p, 0, n * sizeof(double)));
  return 0;
}

extern "C" int hg_ctx_create(const hg_config *cfg, hg_ctx **out) {
  if (!cfg || !out) { set_error("hg_ctx_create: null argument"); return 1; }
  *out = nullptr;
  if (cfg->n1 < 1 || cfg->n2 < 1 || cfg->n3 < 1) { set_error("hg_ctx_create: grid sizes must be >= 1"); return 1; }
  if (cfg->lim != HG_LIM_DONOR && cfg->lim != HG_LIM_VANL && cfg->lim != HG_LIM_MINM && cfg->lim != HG_LIM_MC &&
      cfg->lim != HG_LIM_PARA && cfg->lim != HG_LIM_WENO5BND) {
    set_error("hg_ctx_create: unsupported lim %d", cfg->lim);
    return 1;
  }
  if (cfg->timeorder < 0 || cfg->timeorder > 4) {
    set_error("hg_ctx_create: unsupported timeorder %d (1 to 4)", cfg->timeorder);
    return 1;
  }
  if (cfg->doenoflux != HG_NOENOFLUX && cfg->doenoflux != HG_ENOFINITEVOLUME) {
    set_error("hg_ctx_create: unsupported doenoflux %d", cfg->doenoflux);
    return 1;
  }
  if (cfg->fluxmethod < HG_FLUX_LAXF || cfg->fluxmethod > HG_FLUX_HLLLAXF1) {
    set_error("hg_ctx_create: unsupported fluxmethod %d", cfg->fluxmethod);
    return 1;
  }
  int need = (cfg->lim == HG_LIM_WENO5BND) ? 6 : 3;
  if (cfg->nb < need) { set_error("hg_ctx_create: nb=%d < %d ghost layers needed", cfg->nb, need); return 1; }
  if (cfg->n3 > 1 && cfg->n3 < cfg->nb) { set_error("hg_ctx_create: periodic n3=%d < nb=%d", cfg->n3, cfg->nb); return 1; }
  if (cfg->n2 > 1 && cfg->n2 < cfg->nb) { set_error("hg_ctx_create: n2=%d < nb=%d", cfg->n2, cfg->nb); return 1; }
  {
    // the reconstruction kernels address a primitive plane with 32-bit byte offsets
    const Dims dd = make_dims(cfg->n1, cfg->n2, cfg->n3, cfg->nb);
    if ((uint64_t)dd.spl * sizeof(double) > (uint64_t)UINT32_MAX) {
      set_error("hg_ctx_create: %lld zones with ghosts exceed the 2^29 per primitive plane the kernels address",
                (long long)dd.spl);
      return 1;
    }
  }
  hg_ctx *c = new hg_ctx();
  c->cfg = *cfg;
  c->d = make_dims(cfg->n1, cfg->n2, cfg->n3, cfg->nb);
  const Dims &d = c->d;
  size_t ns = (size_t)d.spl * NPR, nz = (size_t)d.spl;
  int rc = 0;
  rc |= dalloc(&c->p, ns);
  rc |= dalloc(&c->pk, ns);
  rc |= dalloc(&c->uf, ns);
  // pk2 and ucum also serve hg_advance_substep with any coefficients
  const bool fv = cfg->doenoflux == HG_ENOFINITEVOLUME;
  rc |= dalloc(&c->pk2, ns);
  rc |= dalloc(&c->ucum, ns);
  if (fv) rc |= dalloc(&c->ucons, ns);
  for (int k = 0; k < 3; k++) rc |= dalloc(&c->F[k], ns);
  for (int k = 0; k < 3; k++) rc |= dalloc(&c->emf[k], nz);
  if (cfg->lim == HG_LIM_WENO5BND) {
    rc |= dalloc(&c->pl, ns);
    rc |= dalloc(&c->pr, ns);
    if (d.n2 > 1) { rc |= dalloc(&c->pl2, ns); rc |= dalloc(&c->pr2, ns); }
    if (d.n3 > 1) { rc |= dalloc(&c->pl3, ns); rc |= dalloc(&c->pr3, ns); }
    rc |= dalloc(&c->aux, 5 * nz);
  }
  rc |= dalloc(&c->geo, (size_t)4 * GCOMP * d.gplane);
  rc |= dalloc(&c->corn, (size_t)3 * d.gplane);
  rc |= dalloc(&c->conn, (size_t)64 * d.cplane);
  rc |= dalloc(&c->floors, (size_t)2 * d.cplane);
  rc |= dalloc(&c->dt, 1);
  if (!rc) {
    hipError_t e = hipMalloc((void **)&c->ndt_bits, 3 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&c->counters, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&c->err, sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->counters, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->err, 0, sizeof(int));
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
      e = hipStreamCreateWithFlags(&c->xs[k], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e != hipSuccess) { set_error("hg_ctx_create: %s", hipGetErrorString(e)); rc = 1; }
  }
  if (!rc) rc = fixup_alloc(c);
  if (rc) { hg_ctx_destroy(c); return 1; }
  *out = c;
  return 0;
}

extern "C" int hg_ctx_destroy(hg_ctx *c) {
  if (!c) return 0;
  double *bufs[] = {c->p,      c->pk,     c->pk2,    c->ucum,  c->ucons, c->uf,  c->F[0],  c->F[1],
                    c->F[2],   c->emf[0], c->emf[1], c->emf[2], c->pl,    c->pr,  c->pl2,   c->pr2,
                    c->pl3,    c->pr3,    c->aux,    c->geo,    c->corn,  c->conn, c->floors, c->dt};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (c->ndt_bits) (void)hipFree(c->ndt_bits);
  if (c->counters) (void)hipFree(c->counters);
  if (c->err) (void)hipFree(c->err);
  for (int k = 0; k < 2; k++) {
    if (c->xs[k]) (void)hipStreamDestroy(c->xs[k]);
    if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->lx) (void)hipFree(c->lx);
  if (c->ly) (void)hipFree(c->ly);
  fixup_free(c);
  diag_free(c);
  for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
  delete c;
  return 0;
}

// geometry in the host layout of harmgpu.h -> device tables of hg_grid.h
int hg::upload_geometry_tables(hg_ctx *c, const double *gcov, const double *gcon, const double *gcovpert,
                               const double *gdet, const double *conn, const double *r_cent) {
  const Dims &d = c->d;
  // host layout: [i][j][loc7] with loc order FACE1 FACE2 FACE3 CORN1 CORN2 CORN3 CENT
  static const int full_src[4] = {6, 0, 1, 2};  // CENT, FACE1, FACE2, FACE3
  std::vector<double> geo((size_t)4 * GCOMP * d.gplane), corn((size_t)3 * d.gplane);
  for (int ii = 0; ii < d.g1; ii++)
    for (int jj = 0; jj < d.g2; jj++) {
      size_t h = ((size_t)ii * d.g2 + jj) * 7;
      size_t g = (size_t)ii + (size_t)jj * d.g1;
      for (int l = 0; l < 4; l++) {
        size_t hs = h + full_src[l];
        double *b = &geo[(size_t)l * GCOMP * d.gplane + g];
        for (int a = 0; a < 16; a++) {
          b[a * d.gplane] = gcov[hs * 16 + a];
          b[(16 + a) * d.gplane] = gcon[hs * 16 + a];
        }
        b[32 * d.gplane] = gcovpert[hs * 4 + 0];
        b[33 * d.gplane] = gdet[hs];
      }
      for (int l = 0; l < 3; l++) corn[(size_t)l * d.gplane + g] = gdet[h + 3 + l];
    }
  std::vector<double> cn((size_t)64 * d.cplane), fl((size_t)2 * d.cplane);
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++) {
      size_t h = (size_t)ii * d.m2 + jj;
      size_t g = (size_t)ii + (size_t)jj * d.m1;
      for (int a = 0; a < 64; a++) cn[(size_t)a * d.cplane + g] = conn[h * 64 + a];
      double r = r_cent[h];
      // fixup.c:1562-1568 set_density_floors, rescaletype 0
      fl[g] = c->cfg.rhomin * pow(r, -1.5);
      fl[d.cplane + g] = c->cfg.uumin * pow(r, -2.5);
    }
  HIPCHECK(hipMemcpy(c->geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->corn, corn.data(), corn.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->conn, cn.data(), cn.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->floors, fl.data(), fl.size() * sizeof(double), hipMemcpyHostToDevice));
  c->have_geometry = true;
  return 0;
}

extern "C" int hg_upload_geometry(hg_ctx *c, const double *gcov, const double *gcon, const double *gcovpert,
                                  const double *gdet, const double *conn, const double *r_cent) {
  if (!c || !gcov || !gcon || !gcovpert || !gdet || !conn || !r_cent) {
    set_error("hg_upload_geometry: null argument");
    return 1;
  }
  return upload_geometry_tables(c, gcov, gcon, gcovpert, gdet, conn, r_cent);
}

// AoS host [i][j][k][pl] <-> SoA device [pl][k][j][i]
int hg::copy_prims(hg_ctx *c, double *dev, const double *host) {
  const Dims &d = c->d;
  std::vector<double> t((size_t)d.spl * NPR);
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++)
      for (int kk = 0; kk < d.m3; kk++) {
        size_t h = (((size_t)ii * d.m2 + jj) * d.m3 + kk) * NPR;
        size_t g = (size_t)ii + (size_t)jj * d.sj + (size_t)kk * d.sk;
        for (int pl = 0; pl < NPR; pl++) t[g + pl * d.spl] = host[h + pl];
      }
  HIPCHECK(hipMemcpy(dev, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int hg::fetch_prims(hg_ctx *c, const double *dev, double *host) {
  const Dims &d = c->d;
  std::vector<double> t((size_t)d.spl * NPR);
  HIPCHECK(hipMemcpy(t.data(), dev, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int ii = 0; ii < d.m1; ii++)
    for (int jj = 0; jj < d.m2; jj++)
      for (int kk = 0; kk < d.m3; kk++) {
        size_t h = (((size_t)ii * d.m2 + jj) * d.m3 + kk) * NPR;
        size_t g = (size_t)ii + (size_t)jj * d.sj + (size_t)kk * d.sk;
        for (int pl = 0; pl < NPR; pl++) host[h + pl] = t[g + pl * d.spl];
      }
  return 0;
}

extern "C" int hg_set_prims(hg_ctx *c, const double *p) {
  if (!c || !p) { set_error("hg_set_prims: null argument"); return 1; }
  c->ucons_stale = true;
  c->unew_in_uf = false;
  return copy_prims(c, c->p, p);
}

static int need_fv(hg_ctx *c, const char *who) {
  if (!c) { set_error("%s: null argument", who); return 1; }
  if (!c->ucons) { set_error("%s: the context does not run the finite-volume scheme", who); return 1; }
  return 0;
}

extern "C" int hg_set_conserved(hg_ctx *c, const double *U) {
  if (need_fv(c, "hg_set_conserved")) return 1;
  if (!U) { set_error("hg_set_conserved: null argument"); return 1; }
  const Dims &d = c->d;
  std::vector<double> act((size_t)d.m1 * d.m2 * d.m3 * NPR, 0.0);
  for (int i = 0; i < d.n1; i++)
    for (int j = 0; j < d.n2; j++)
      for (int k = 0; k < d.n3; k++) {
        size_t h = ((((size_t)(i + d.nb1)) * d.m2 + (j + d.nb2)) * d.m3 + (k + d.nb3)) * NPR;
        for (int pl = 0; pl < NPR; pl++) act[h + pl] = U[h + pl];
      }
  if (copy_prims(c, c->ucons, act.data())) return 1;
  c->ucons_stale = false;
  return 0;
}

extern "C" int hg_get_conserved(hg_ctx *c, double *U) {
  if (need_fv(c, "hg_get_conserved")) return 1;
  if (!U) { set_error("hg_get_conserved: null argument"); return 1; }
  if (c->ucons_stale) {
    if (!c->have_geometry) { set_error("hg_get_conserved: geometry not set"); return 1; }
    hg::ensure_conserved(c, nullptr);
  }
  HIPCHECK(hipDeviceSynchronize());
  return fetch_prims(c, c->ucons, U);
}

extern "C" int hg_get_prims(hg_ctx *c, double *p) {
  if (!c || !p) { set_error("hg_get_prims: null argument"); return 1; }
  HIPCHECK(hipDeviceSynchronize());
  return fetch_prims(c, c->p, p);
}

static int check_err(hg_ctx *c) {
  int e = 0;
  HIPCHECK(hipMemcpy(&e, c->err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    HIPCHECK(hipMemset(c->err, 0, sizeof(int)));
    set_error("device reported failure mask %d (1: ghost fill, 2: face state, 4: zone update, 8: post_fixup "
              "gamma, 16: post_fixup check did not settle, 32: unknown failure code)", e);
    return 1;
  }
  return 0;
}

static int need_geometry(hg_ctx *c, const char *who) {
  if (!c) { set_error("%s: null context", who); return 1; }
  if (!c->have_geometry) { set_error("%s: geometry not set", who); return 1; }
  return 0;
}

extern "C" int hg_bound_prim(hg_ctx *c, void *stream) {
  if (need_geometry(c, "hg_bound_prim")) return 1;
  hipStream_t s = (hipStream_t)stream;
  launch_bound(c->p, c->geo, c->d, c->cfg.gammamax, c->err, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  return check_err(c);
}

// the failure codes of the active zones, host order [i][j][k] <-> device int8
static int codes_to_device(hg_ctx *c, const int32_t *h) {
  const Dims &d = c->d;
  std::vector<int8_t> buf((size_t)d.spl, 0);
  for (int i = 0; i < d.n1; i++)
    for (int j = 0; j < d.n2; j++)
      for (int k = 0; k < d.n3; k++) {
        int32_t v = h[((size_t)i * d.n2 + j) * d.n3 + k];
        if (v < -128 || v > 127) {
          set_error("hg_post_fixup: failure code %d out of range", (int)v);
          return 1;
        }
        buf[d.idx(i, j, k)] = (int8_t)v;
      }
  HIPCHECK(hipMemcpy(c->pflag, buf.data(), buf.size(), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int hg_get_pflag(hg_ctx *c, int32_t *h) {
  if (!c || !h) { set_error("hg_get_pflag: null argument"); return 1; }
  const Dims &d = c->d;
  std::vector<int8_t> buf((size_t)d.spl);
  HIPCHECK(hipMemcpy(buf.data(), c->pflag, buf.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < d.n1; i++)
    for (int j = 0; j < d.n2; j++)
      for (int k = 0; k < d.n3; k++) h[((size_t)i * d.n2 + j) * d.n3 + k] = buf[d.idx(i, j, k)];
  return 0;
}

extern "C" int hg_post_fixup(hg_ctx *c, const int32_t *codes, int finalstep, void *stream) {
  if (need_geometry(c, "hg_post_fixup")) return 1;
  hipStream_t s = (hipStream_t)stream;
  HIPCHECK(hipStreamSynchronize(s));
  if (codes && codes_to_device(c, codes)) return 1;
  ensure_conserved(c, s);
  launch_bound(c->p, c->geo, c->d, c->cfg.gammamax, c->err, s);
  launch_post_fixup(c, c->p, finalstep ? 1 : 0, s);
  launch_bound(c->p, c->geo, c->d, c->cfg.gammamax, c->err, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  return check_err(c);
}

// ---------------------------------------------------------------------------
// per-kernel-class timing with HIP events on the launch stream.  While a
// profile is open, prof_mark(c, k, s) closes the running interval (if any)
// and opens one of class k (k < 0: none); the events are read back and summed
// by hg_profile_read.
// ---------------------------------------------------------------------------
static void prof_mark(hg_ctx *c, int cls, hipStream_t s) {
  if (!c->prof_on) return;
  if (c->prof_open >= 0) {
    hipEvent_t stop;
    if (c->prof_used < c->prof_events.size()) stop = c->prof_events[c->prof_used];
    else {
      if (hipEventCreate(&stop) != hipSuccess) { c->prof_on = false; return; }
      c->prof_events.push_back(stop);
    }
    (void)hipEventRecord(stop, s);
    c->prof_intervals.push_back({c->prof_open, c->prof_used - 1, c->prof_used});
    c->prof_used++;
    c->prof_open = -1;
  }
  if (cls < 0) return;
  hipEvent_t start;
  if (c->prof_used < c->prof_events.size()) start = c->prof_events[c->prof_used];
  else {
    if (hipEventCreate(&start) != hipSuccess) { c->prof_on = false; return; }
    c->prof_events.push_back(start);
  }
  (void)hipEventRecord(start, s);
  c->prof_used++;
  c->prof_open = cls;
}

extern "C" int hg_profile(hg_ctx *c, int enable) {
  if (!c) { set_error("hg_profile: null context"); return 1; }
  c->prof_on = enable != 0;
  c->prof_open = -1;
  if (enable) {
    c->prof_used = 0;
    c->prof_intervals.clear();
    // enough events for ~150 steps, created here rather than while timing
    while (c->prof_events.size() < 4096) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      c->prof_events.push_back(e);
    }
  }
  return 0;
}

extern "C" int hg_profile_read(hg_ctx *c, double ms[HG_PROF_NCLASS], int64_t launches[HG_PROF_NCLASS]) {
  if (!c || !ms) { set_error("hg_profile_read: null argument"); return 1; }
  for (int k = 0; k < HG_PROF_NCLASS; k++) {
    ms[k] = 0.;
    if (launches) launches[k] = 0;
  }
  for (const auto &iv : c->prof_intervals) {
    HIPCHECK(hipEventSynchronize(c->prof_events[iv.stop]));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, c->prof_events[iv.start], c->prof_events[iv.stop]));
    ms[iv.cls] += t;
    if (launches) launches[iv.cls]++;
  }
  return 0;
}

extern "C" const char *hg_profile_class_name(int k) {
  static const char *names[HG_PROF_NCLASS] = {"bound", "recon1", "recon2", "recon3", "flux1", "flux2",
                                              "flux3", "emf",    "update", "fixup",  "prep"};
  return (k >= 0 && k < HG_PROF_NCLASS) ? names[k] : "";
}

// fluxcalc (flux.c:7) on primitives pb: per-direction fluxes and ndt bits,
// then the flux-CT EMFs (fluxct.c:10)
static void enqueue_fluxcalc(hg_ctx *c, const double *pb, hipStream_t s) {
  const Dims &d = c->d;
  const hg_config &f = c->cfg;
  c->flux_state = pb;
  c->dg_faces_valid = false;
  int nn[4] = {0, d.n1, d.n2, d.n3};
  // the directions are independent (own interface states, fluxes and ndt
  // slot): j and k on the side streams, i on s, joined before the EMFs.
  // A profile runs them one after another so that every interval holds one
  // kernel class; so does any stream capture, hg_graph_create's or a
  // caller's: the fork records the context's own events onto its own side
  // streams, which a capture would pull into the graph, and that path is
  // kept out of captures (a graph runs its grids side by side instead).
  bool conc = c->tune.concurrent && !c->prof_on && !c->capturing;
  if (conc) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) conc = false;
  }
  // Run one after another, the three directions' dP/P indicators take one
  // launch with the shared inputs' (fewer launches for the small grids of a
  // graph); a profile keeps them with their directions, as the concurrent
  // step runs them.  The prep kernel also resets the ndt slots.
  const bool weno = f.lim == HG_LIM_WENO5BND;
  const bool all_dpi = !conc && !c->prof_on;
  if (weno) {
    // the dP/P inputs are shared by the three directions
    prof_mark(c, HG_PROF_PREP, s);
    launch_weno_prep(pb, c->geo, c->aux, d, f.gam, f.doenoflux == HG_ENOFINITEVOLUME, all_dpi, c->ndt_bits, s);
    prof_mark(c, -1, s);
  } else {
    launch_reset_ndt(c->ndt_bits, s);
  }
  hipStream_t ds[4] = {nullptr, s, s, s};
  if (conc) {
    (void)hipEventRecord(c->ev_fork, s);
    for (int dir = 2; dir <= 3; dir++)
      if (nn[dir] > 1) {
        ds[dir] = c->xs[dir - 2];
        (void)hipStreamWaitEvent(ds[dir], c->ev_fork, 0);
      }
  }
  double *pls[4] = {nullptr, c->pl, c->pl2, c->pl3}, *prs[4] = {nullptr, c->pr, c->pr2, c->pr3};
  // concurrent: the side streams' work first; a profile keeps the order 1, 2, 3
  const int order[2][3] = {{1, 2, 3}, {3, 2, 1}};
  for (int o = 0; o < 3; o++) {
    const int dir = order[conc ? 1 : 0][o];
    if (nn[dir] <= 1) continue;
    hipStream_t sd = ds[dir];
    if (f.lim != HG_LIM_WENO5BND) {
      prof_mark(c, HG_PROF_FLUX1 + dir - 1, sd);
      launch_flux_point(dir, f.lim, f.doenoflux == HG_ENOFINITEVOLUME, pb, c->geo, c->F[dir - 1], d,
                        f.gam, f.cour, f.dx[dir], f.fluxmethod, c->ndt_bits + dir - 1, c->err, c->tune, sd);
      prof_mark(c, -1, sd);
    } else {
      prof_mark(c, HG_PROF_RECON1 + dir - 1, sd);
      launch_weno(dir, pb, c->aux, pls[dir], prs[dir], d, c->tune, f.doenoflux == HG_ENOFINITEVOLUME, !all_dpi, sd);
      prof_mark(c, HG_PROF_FLUX1 + dir - 1, sd);
      launch_flux_lr(dir, pls[dir], prs[dir], c->geo, c->F[dir - 1], d, f.gam, f.cour, f.dx[dir], f.fluxmethod,
                     c->ndt_bits + dir - 1, c->err, c->tune, f.doenoflux == HG_ENOFINITEVOLUME, sd);
      prof_mark(c, -1, sd);
    }
  }
  if (conc)
    for (int dir = 2; dir <= 3; dir++)
      if (nn[dir] > 1) {
        (void)hipEventRecord(c->ev_join[dir - 2], ds[dir]);
        (void)hipStreamWaitEvent(s, c->ev_join[dir - 2], 0);
      }
  prof_mark(c, HG_PROF_EMF, s);
  launch_emf(c->F[0], c->F[1], c->F[2], c->geo, c->corn, c->emf[0], c->emf[1], c->emf[2], d, c->tune, s);
  prof_mark(c, -1, s);
}

extern "C" int hg_fluxcalc(hg_ctx *c, double *F1, double *F2, double *F3, double ndt[3], void *stream) {
  if (need_geometry(c, "hg_fluxcalc")) return 1;
  hipStream_t s = (hipStream_t)stream;
  enqueue_fluxcalc(c, c->p, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  if (check_err(c)) return 1;
  if (ndt) {
    unsigned long long b[3];
    HIPCHECK(hipMemcpy(b, c->ndt_bits, sizeof b, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) memcpy(&ndt[k], &b[k], sizeof(double));
  }
  const Dims &d = c->d;
  // flux_ct's rewrite of the field fluxes (fluxct.c:720-781), on the device;
  // the step itself rebuilds them from the EMFs inside k_update
  launch_fluxct_fields(c->F[0], c->F[1], c->F[2], c->emf[0], c->emf[1], c->emf[2], d, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  std::vector<double> h[3];
  for (int k = 0; k < 3; k++) {
    h[k].resize((size_t)d.spl * NPR);
    HIPCHECK(hipMemcpy(h[k].data(), c->F[k], h[k].size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  double *outs[3] = {F1, F2, F3};
  for (int k = 0; k < 3; k++) {
    if (!outs[k]) continue;
    for (int ii = 0; ii < d.m1; ii++)
      for (int jj = 0; jj < d.m2; jj++)
        for (int kk = 0; kk < d.m3; kk++) {
          size_t hh = (((size_t)ii * d.m2 + jj) * d.m3 + kk) * NPR;
          size_t g = (size_t)ii + (size_t)jj * d.sj + (size_t)kk * d.sk;
          for (int pl = 0; pl < NPR; pl++) outs[k][hh + pl] = h[k][g + pl * d.spl];
        }
  }
  return 0;
}

// one substep of step_ch.c:94-166: advance, bound_prim, post_fixup, bound_prim
static void enqueue_substep(hg_ctx *c, double *pi, double *pb, double *pf, const RkStage &rk, int stage,
                            int finalstep, hipStream_t s) {
  const hg_config &f = c->cfg;
  enqueue_fluxcalc(c, pb, s);
  const bool rk4 = f.timeorder == 4;
  if (c->dg_record && (finalstep || rk4)) {
    // diag_flux of advance.c:130, where the reference's p is still pi; the
    // midpoint scheme's first substep has Dt = 0, RK4 needs every substep
    launch_diag_faces(c, pb, pi, s);
    if (rk4) launch_diag_stage(c, stage, s);
    c->dg_faces_valid = finalstep != 0;
  }
  prof_mark(c, HG_PROF_UPDATE, s);
  launch_update(pi, pb, c->uf, c->F[0], c->F[1], c->F[2], c->emf[0], c->emf[1], c->emf[2], c->geo, c->conn, c->dt,
                rk, c->d, f.gam, f.dx, c->err, c->tune, s);
  if (c->c2p_mode == HG_C2P_LEARNED) {
    // the network's start for every inversion of this substep
    launch_step_features(c->uf, pb, c->geo, c->d, f.gam, c->lx, s);
    launch_mlp_forward(c->mlp, (int64_t)c->d.n1 * c->d.n2 * c->d.n3, c->lx, c->ly, s);
  }
  launch_invert(c->uf, pb, pf, c->geo, c->floors, rk, finalstep, c->d, f.gam, f.gammamax, c->counters, c->pflag,
                c->c2p_mode == HG_C2P_MIXED ? 1 : c->c2p_mode == HG_C2P_LEARNED ? 2 : 0, c->ly, c->fx_count,
                c->err, c->tune, s);
  if (finalstep) launch_dt(c->ndt_bits, c->dt, f.defcon, f.safe, s);
  // step_ch.c:127 and :166 call bound_prim twice per substep; the second call
  // is not a repeat of the first: POLEDEATH has zeroed U2 in the rows next to
  // the axis, which the radial inflow check then sees in the corner ghosts
  prof_mark(c, HG_PROF_BOUND, s);
  launch_bound(pf, c->geo, c->d, f.gammamax, c->err, s);
  prof_mark(c, HG_PROF_FIXUP, s);
  launch_post_fixup(c, pf, finalstep, s, false);
  prof_mark(c, HG_PROF_BOUND, s);
  launch_bound(pf, c->geo, c->d, f.gammamax, c->err, s);
  prof_mark(c, -1, s);
}

// the finite-volume scheme's conserved state from p when p was set since
// (initbase.c:147 init_conservatives); kernels only
void hg::ensure_conserved(hg_ctx *c, hipStream_t s) {
  if (!c->ucons || !c->ucons_stale) return;
  launch_cons_from_prims(c->p, c->geo, c->ucons, c->d, c->cfg.gam, c->err, s);
  c->ucons_stale = false;
}

// the substeps of step_ch.c:442 setup_rktimestep for the context's
// TIMEORDER: each one's pb, pf and coefficients (pi is p on every substep)
int hg::step_substeps(hg_ctx *c, double *pb[4], double *pf[4], RkStage rk[4]) {
  const int to = c->cfg.timeorder;
  if (to == 1 || to == 3) {
    // TIMEORDER 1 Euler: CUf {1, 0, 1}, Cunew {0, 0, 1}, pb = pf = p.
    // TIMEORDER 3 TVD RK3: CUf {1, 0, 1}, {3/4, 1/4, 1/4}, {1/3, 2/3, 2/3},
    // Cunew {0, 0, 0}, {0, 0, 0}, {0, 0, 1}; pb = p, pk, pk2; pf = pk, pk2, p
    const int n = to;
    const double cuf[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const double cn3[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    double *b3[3] = {c->p, c->pk, c->pk2}, *f3[3] = {c->pk, c->pk2, c->p};
    for (int t = 0; t < n; t++) {
      const double *cn = (n == 1) ? cn3[2] : cn3[t];
      rk[t] = RkStage{cuf[t][2], cn[0], cn[1], cn[2], c->ucum, t == 0, t == n - 1, c->ucons, cuf[t][0], cuf[t][1]};
      pb[t] = (n == 1) ? c->p : b3[t];
      pf[t] = (n == 1) ? c->p : f3[t];
    }
    return n;
  }
  if (to != 4) {
    // TIMEORDER 2 midpoint: CUf[.][2] = 0.5, 1; Cunew = {0, 0, 0}, {0, 0, 1};
    // pb = p, pk; pf = pk, p.  The standard scheme inverts Uf, which equals
    // that ucum; the finite-volume one keeps ucum itself
    rk[0] = RkStage{0.5, 0.0, 0.0, 0.0, c->ucons ? c->ucum : nullptr, 1, 0, c->ucons};
    rk[1] = RkStage{1.0, 0.0, 0.0, 1.0, c->ucons ? c->ucum : nullptr, 0, 1, c->ucons};
    pb[0] = c->p;
    pf[0] = c->pk;
    pb[1] = c->pk;
    pf[1] = c->p;
    return 2;
  }
  // TIMEORDER 4: CUf[.][2] = 1/2, 1/2, 1, 1; Cunew = {1, 1/6, 0}, {0, 1/3, 0},
  // {0, 1/3, 0}, {0, 1/6, 0}; pb = p, pk, pk2, pk; pf = pk, pk2, pk, p
  const double cuf2[4] = {0.5, 0.5, 1.0, 1.0};
  const double cn[4][3] = {{1.0, 1.0 / 6.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 6.0, 0.0}};
  double *b4[4] = {c->p, c->pk, c->pk2, c->pk}, *f4[4] = {c->pk, c->pk2, c->pk, c->p};
  for (int t = 0; t < 4; t++) {
    rk[t] = RkStage{cuf2[t], cn[t][0], cn[t][1], cn[t][2], c->ucum, t == 0, t == 3, c->ucons};
    pb[t] = b4[t];
    pf[t] = f4[t];
  }
  return 4;
}

// step_ch.c:35 step_ch_simplempi over the substeps of step_ch.c:442
// setup_rktimestep
void hg::enqueue_step(hg_ctx *c, hipStream_t s) {
  ensure_conserved(c, s);
  c->unew_in_uf = true;
  double *pb[4], *pf[4];
  RkStage rk[4];
  const int n = step_substeps(c, pb, pf, rk);
  for (int t = 0; t < n; t++) enqueue_substep(c, c->p, pb[t], pf[t], rk[t], t, t == n - 1, s);
}

extern "C" int hg_set_dt(hg_ctx *c, double dt) {
  if (!c) { set_error("hg_set_dt: null context"); return 1; }
  HIPCHECK(hipMemcpy(c->dt, &dt, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int hg_get_dt(hg_ctx *c, double *dt) {
  if (!c || !dt) { set_error("hg_get_dt: null argument"); return 1; }
  HIPCHECK(hipMemcpy(dt, c->dt, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int hg_get_counters(hg_ctx *c, int64_t *nfail, int64_t *iters) {
  if (!c) { set_error("hg_get_counters: null context"); return 1; }
  unsigned long long h[2];
  HIPCHECK(hipMemcpy(h, c->counters, sizeof h, hipMemcpyDeviceToHost));
  if (nfail) *nfail = (int64_t)h[0];
  if (iters) *iters = (int64_t)h[1];
  HIPCHECK(hipMemset(c->counters, 0, sizeof h));
  return 0;
}

extern "C" int hg_step_async(hg_ctx *c, int nsteps, void *stream) {
  if (need_geometry(c, "hg_step_async")) return 1;
  if (nsteps < 0) { set_error("hg_step_async: nsteps < 0"); return 1; }
  hipStream_t s = (hipStream_t)stream;
  for (int n = 0; n < nsteps; n++) enqueue_step(c, s);
  HIPCHECK(hipGetLastError());
  return 0;
}

// one pass of step_ch.c:97-166 with the caller's substep: advance
// (advance.c:16) from pi/pb into pf, bound_prim, post_fixup, bound_prim
extern "C" int hg_advance_substep(hg_ctx *c, int pi, int pb, int pf, const double CUf[4], const double Cunew[3],
                                  int timeorder, int numtimeorders, double dt, double *ndt, void *stream) {
  if (need_geometry(c, "hg_advance_substep")) return 1;
  if (!CUf || !Cunew || pi < 0 || pi > 2 || pb < 0 || pb > 2 || pf < 0 || pf > 2 || numtimeorders < 1 ||
      numtimeorders > 4 || timeorder < 0 || timeorder >= numtimeorders) {
    set_error("hg_advance_substep: bad arguments");
    return 1;
  }
  double *slot[3] = {c->p, c->pk, c->pk2};
  hipStream_t s = (hipStream_t)stream;
  HIPCHECK(hipStreamSynchronize(s));
  if (hg_set_dt(c, dt)) return 1;
  ensure_conserved(c, s);
  const int last = timeorder == numtimeorders - 1;
  if (last) c->unew_in_uf = true;
  RkStage rk{CUf[2], Cunew[0], Cunew[1], Cunew[2], c->ucum, timeorder == 0, last, c->ucons, CUf[0], CUf[1]};
  enqueue_substep(c, slot[pi], slot[pb], slot[pf], rk, timeorder, last, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  if (check_err(c)) return 1;
  if (ndt) {
    // advance.c:231
    unsigned long long b[3];
    double n[3];
    HIPCHECK(hipMemcpy(b, c->ndt_bits, sizeof b, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) memcpy(&n[k], &b[k], sizeof(double));
    *ndt = c->cfg.defcon * 1. / (1. / n[0] + 1. / n[1] + 1. / n[2]);
  }
  return 0;
}

extern "C" int hg_step(hg_ctx *c, double *dt, int64_t *nfail, int64_t *iters, void *stream) {
  if (need_geometry(c, "hg_step")) return 1;
  if (!dt) { set_error("hg_step: null dt"); return 1; }
  hipStream_t s = (hipStream_t)stream;
  // work the caller queued on s (hg_step_async) still reads dt and adds to
  // the counters: let it finish before they are set
  HIPCHECK(hipStreamSynchronize(s));
  if (hg_set_dt(c, *dt)) return 1;
  HIPCHECK(hipMemset(c->counters, 0, 2 * sizeof(unsigned long long)));
  enqueue_step(c, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  if (check_err(c)) return 1;
  if (hg_get_dt(c, dt)) return 1;
  return hg_get_counters(c, nfail, iters);
}

// the final substep's inversions of the last step (advance.c:193-201): the
// midpoint scheme inverts Uf (kept in uf) with the guess pb = pk
extern "C" int hg_dump_inversions(hg_ctx *c, double *U, double *gcov, double *gcon, double *gdet, double *gcovpert,
                                  double *guess, void *stream) {
  if (need_geometry(c, "hg_dump_inversions")) return 1;
  if (!U || !gcov || !gcon || !gdet || !gcovpert || !guess) {
    set_error("hg_dump_inversions: null argument");
    return 1;
  }
  if (c->cfg.timeorder != 2 || c->ucons || !c->unew_in_uf) {
    set_error("hg_dump_inversions: needs a step of the midpoint (TIMEORDER 2) standard scheme since the "
              "primitives were set");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  launch_gather_inversions(c->uf, c->pk, c->geo, c->d, U, gcov, gcon, gdet, gcovpert, guess, s);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int hg_get_config(const hg_ctx *c, hg_config *out) {
  if (!c || !out) { set_error("hg_get_config: null argument"); return 1; }
  *out = c->cfg;
  return 0;
}

// the inversion the step uses (advance.c:193-201); the learned mode's
// feature and prediction buffers are allocated here, never in the step
extern "C" int hg_set_inversion(hg_ctx *c, int mode, const hg_mlp *m) {
  if (!c) { set_error("hg_set_inversion: null context"); return 1; }
  if (mode != HG_C2P_FP64 && mode != HG_C2P_MIXED && mode != HG_C2P_LEARNED) {
    set_error("hg_set_inversion: unknown mode %d", mode);
    return 1;
  }
  if (mode == HG_C2P_LEARNED && !m) { set_error("hg_set_inversion: HG_C2P_LEARNED needs a network"); return 1; }
  if (mode == HG_C2P_LEARNED && !(c->lx && c->ly)) {
    // both buffers or neither: a failed allocation leaves the context as it was
    const size_t n = (size_t)c->d.n1 * c->d.n2 * c->d.n3;
    float *lx = nullptr, *ly = nullptr;
    hipError_t e = hipMalloc(&lx, n * C2P_NFEAT * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&ly, n * sizeof(float));
    if (e != hipSuccess) {
      if (lx) (void)hipFree(lx);
      if (ly) (void)hipFree(ly);
      set_error("hg_set_inversion: %s", hipGetErrorString(e));
      return 1;
    }
    if (c->lx) (void)hipFree(c->lx);
    if (c->ly) (void)hipFree(c->ly);
    c->lx = lx;
    c->ly = ly;
  }
  c->c2p_mode = mode;
  c->mlp = mode == HG_C2P_LEARNED ? m : nullptr;
  return 0;
}
