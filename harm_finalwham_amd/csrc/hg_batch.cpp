// Many grids in one HIP graph, and the launch-shape autotuner.
//
// hg_graph_create captures nsteps RK2 steps of every grid into one graph:
// each grid gets its own branch (a stream forked from the capture stream and
// joined back), so the GPU runs the grids' kernels side by side and one
// hipGraphLaunch advances all of them.  dt stays on the device per grid, so
// replays need no host work.  The step enqueues kernels only (no memset or
// copy nodes).
//
// hg_graph_create_grouped captures the same steps as one launch per kernel of
// a substep over the blocks of all the grids (GridView tables on the device,
// hg_kernels.h): the graph then holds ~20 launches per substep instead of
// ~20 per substep and grid, which the runtime would mostly run one at a time.
//
// hg_autotune times the step under different launch shapes (block sizes,
// reconstruction wavefronts per block and line segments) and keeps the
// fastest; the shapes do not change any result bit.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "harmgpu.h"
#include "hg_host.h"

using namespace hg;

struct hg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *tables = nullptr;  // grouped: the views and block tables the launches read
};

#define HIPTRY(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess && rc == 0) {                                             \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      rc = 1;                                                                      \
    }                                                                              \
  } while (0)

extern "C" int hg_graph_create(hg_ctx *const *ctxs, int n, int nsteps, hg_graph **out) {
  if (!ctxs || !out || n < 1 || nsteps < 1) {
    set_error("hg_graph_create: need n >= 1 contexts, nsteps >= 1 and an output");
    return 1;
  }
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || !ctxs[i]->have_geometry) {
      set_error("hg_graph_create: context %d missing or without geometry", i);
      return 1;
    }
    if (ctxs[i]->prof_on) {
      set_error("hg_graph_create: context %d has a profile open", i);
      return 1;
    }
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) {
        set_error("hg_graph_create: context %d appears twice", i);
        return 1;
      }
  }
  int rc = 0;
  hipStream_t cap = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipStream_t> st(n, nullptr);
  std::vector<hipEvent_t> join(n, nullptr);
  HIPTRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  HIPTRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  for (int i = 0; i < n && !rc; i++) {
    HIPTRY(hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking));
    HIPTRY(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
  }
  hg_graph *g = nullptr;
  // a pending conserved state of the finite-volume scheme is derived now,
  // outside the graph, so that replays do not repeat it
  for (int i = 0; i < n && !rc; i++) {
    ensure_conserved(ctxs[i], nullptr);
    HIPTRY(hipDeviceSynchronize());
  }
  if (!rc) {
    g = new hg_graph();
    HIPTRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    if (!rc) {
      HIPTRY(hipEventRecord(fork, cap));
      for (int i = 0; i < n && !rc; i++) {
        HIPTRY(hipStreamWaitEvent(st[i], fork, 0));
        ctxs[i]->capturing = true;
        for (int k = 0; k < nsteps; k++) enqueue_step(ctxs[i], st[i]);
        ctxs[i]->capturing = false;
        HIPTRY(hipGetLastError());
        HIPTRY(hipEventRecord(join[i], st[i]));
        HIPTRY(hipStreamWaitEvent(cap, join[i], 0));
      }
      // always end the capture, also after an error, so the stream is usable
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(cap, &graph);
      if (e != hipSuccess && !rc) {
        set_error("hg_graph_create: capture failed: %s", hipGetErrorString(e));
        rc = 1;
      }
      g->graph = graph;
      if (!rc) HIPTRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    }
  }
  for (int i = 0; i < n; i++) {
    if (st[i]) (void)hipStreamDestroy(st[i]);
    if (join[i]) (void)hipEventDestroy(join[i]);
  }
  if (fork) (void)hipEventDestroy(fork);
  if (cap) (void)hipStreamDestroy(cap);
  if (rc) {
    if (g) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
    }
    return 1;
  }
  *out = g;
  return 0;
}

// ---------------------------------------------------------------------------
// grouped graph
// ---------------------------------------------------------------------------
namespace {

// the views of the substeps of one step of a context (enqueue_step's pb,
// pf and coefficients; pi = p); returns their number
int make_views(hg_ctx *c, const Tune &tu, GridView v[4]) {
  const hg_config &f = c->cfg;
  double *pbs[4], *pfs[4];
  RkStage rks[4];
  const int ns = step_substeps(c, pbs, pfs, rks);
  for (int sub = 0; sub < ns; sub++) {
    GridView &w = v[sub];
    memset(&w, 0, sizeof w);
    w.d = c->d;
    w.pi = c->p;
    w.pb = pbs[sub];
    w.pf = pfs[sub];
    w.geo = c->geo;
    w.corn = c->corn;
    w.conn = c->conn;
    w.floors = c->floors;
    double *pls[3] = {c->pl, c->pl2, c->pl3}, *prs[3] = {c->pr, c->pr2, c->pr3};
    for (int k = 0; k < 3; k++) {
      w.F[k] = c->F[k];
      w.emf[k] = c->emf[k];
      w.pl[k] = pls[k];
      w.pr[k] = prs[k];
    }
    w.aux = c->aux;
    w.uf = c->uf;
    w.lx = c->lx;
    w.ly = c->ly;
    w.dt = c->dt;
    w.ndt_bits = c->ndt_bits;
    w.counters = c->counters;
    w.pflag = c->pflag;
    w.fx = fixup_bufs(c);
    w.err = c->err;
    w.rk = rks[sub];
    w.gam = f.gam;
    w.gammamax = f.gammamax;
    w.cour = f.cour;
    w.defcon = f.defcon;
    w.safe = f.safe;
    for (int k = 1; k < 4; k++) {
      w.dx[k] = f.dx[k];
      w.idx[k] = 1.0 / f.dx[k];
    }
    w.finalstep = sub == ns - 1;
    w.fluxmethod = f.fluxmethod;
    weno_group_shapes(w, tu);
  }
  return ns;
}

// wavefronts of the grouped marching reconstruction along dir (2 or 3) of
// one grid with the segments of a view (lanes packed over the primitives)
int64_t march_waves(const GridView &v, int dir) {
  const Dims &d = v.d;
  if (((dir == 2) ? d.n2 : d.n3) <= 1) return 0;
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const int64_t lanes = (int64_t)(d.n1 + 2 * s1) * ((dir == 2) ? d.n3 + 2 * s3 : d.n2 + 2 * s2);
  return (NPR * lanes + 63) / 64 * v.nseg[dir];
}

// The segments along j and k of a grouped launch: the contexts' own where
// their tune sets them, else one count for all, the fewest that give the
// marching launch two wavefronts per wavefront slot of the device (four
// per SIMD).  Batches of small grids have short lines and few lanes, and
// their launches fill the device only with more segments than a grid that
// runs alone would take.
void grouped_tunes(hg_ctx *const *ctxs, int n, std::vector<Tune> &tunes) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const int64_t want = 2LL * cus * 4 * 4;
  tunes.resize(n);
  for (int i = 0; i < n; i++) tunes[i] = ctxs[i]->tune;
  for (int dir = 2; dir <= 3; dir++) {
    int pick = 64;
    for (int seg = 1; seg <= 64; seg++) {
      int64_t waves = 0;
      for (int i = 0; i < n; i++) {
        Tune t = ctxs[i]->tune;
        int &req = (dir == 2) ? t.weno_nseg : t.weno_nseg3;
        if (req == 0) req = seg;
        GridView v;
        memset(&v, 0, sizeof v);
        v.d = ctxs[i]->d;
        weno_group_shapes(v, t);
        waves += march_waves(v, dir);
      }
      if (waves >= want) {
        pick = seg;
        break;
      }
    }
    for (int i = 0; i < n; i++) {
      int &req = (dir == 2) ? tunes[i].weno_nseg : tunes[i].weno_nseg3;
      if (req == 0) req = pick;
    }
  }
}

const char *grouped_unsupported(const hg_ctx *c, const hg_ctx *c0) {
  const hg_config &f = c->cfg;
  if (f.lim != HG_LIM_WENO5BND) return "reconstruction other than WENO5BND";
  if (f.doenoflux == HG_ENOFINITEVOLUME) return "the finite-volume scheme";
  if (f.timeorder != c0->cfg.timeorder) return "a timeorder unlike the first context's";
  if (f.fluxmethod != HG_FLUX_LAXF && f.fluxmethod != HG_FLUX_HLL) return "a flux other than LAXF or HLL";
  if (c->c2p_mode != HG_C2P_FP64 && c->c2p_mode != HG_C2P_MIXED && c->c2p_mode != HG_C2P_LEARNED)
    return "an inversion mode other than FP64, MIXED or LEARNED";
  if (c->c2p_mode != c0->c2p_mode) return "an inversion mode unlike the first context's";
  if (c->c2p_mode == HG_C2P_LEARNED && (!c->mlp || c->mlp != c0->mlp || !c->lx || !c->ly))
    return "a learned start with a network unlike the first context's";
  if (c->dg_record) return "diagnostics recording on";
  if (c->ucons) return "a carried conserved state";
  // the packed reconstruction lanes address the 8 planes with 32-bit offsets
  if (c->d.spl >= ((int64_t)1 << 26)) return "a primitive plane of 2^26 zones or more";
  return nullptr;
}

}  // namespace

extern "C" int hg_graph_create_grouped(hg_ctx *const *ctxs, int n, int nsteps, hg_graph **out) {
  if (!ctxs || !out || n < 1 || nsteps < 1) {
    set_error("hg_graph_create_grouped: need n >= 1 contexts, nsteps >= 1 and an output");
    return 1;
  }
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || !ctxs[i]->have_geometry) {
      set_error("hg_graph_create_grouped: context %d missing or without geometry", i);
      return 1;
    }
    if (ctxs[i]->prof_on) {
      set_error("hg_graph_create_grouped: context %d has a profile open", i);
      return 1;
    }
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) {
        set_error("hg_graph_create_grouped: context %d appears twice", i);
        return 1;
      }
    if (const char *why = grouped_unsupported(ctxs[i], ctxs[0])) {
      set_error("hg_graph_create_grouped: context %d has %s (hg_graph_create takes it)", i, why);
      return 1;
    }
  }
  // host tables: ns x n views (substep-major), then per class n + 1 block
  // offsets
  std::vector<Tune> tunes;
  grouped_tunes(ctxs, n, tunes);
  std::vector<GridView> views;
  int ns = 0;
  for (int i = 0; i < n; i++) {
    GridView v[4];
    const int k = make_views(ctxs[i], tunes[i], v);
    if (i == 0) {
      ns = k;
      views.resize((size_t)ns * n);
    }
    for (int sub = 0; sub < ns; sub++) views[(size_t)sub * n + i] = v[sub];
  }
  const bool ucum = views[0].rk.ucum != nullptr;
  std::vector<unsigned> first((size_t)GK_NCLASS * (n + 1));
  unsigned total[GK_NCLASS];
  for (int k = 0; k < GK_NCLASS; k++) {
    uint64_t acc = 0;
    for (int i = 0; i < n; i++) {
      first[(size_t)k * (n + 1) + i] = (unsigned)acc;
      acc += (k == GK_MLP) ? group_blocks_mlp(views[i]) : group_blocks(k, views[i]);
    }
    if (acc > 0x7fffffffull) {
      set_error("hg_graph_create_grouped: %llu blocks in one launch exceed the grid limit",
                (unsigned long long)acc);
      return 1;
    }
    first[(size_t)k * (n + 1) + n] = (unsigned)acc;
    total[k] = (unsigned)acc;
  }
  const size_t vbytes = views.size() * sizeof(GridView), vpad = (vbytes + 255) / 256 * 256;
  const size_t bytes = vpad + first.size() * sizeof(unsigned);
  int rc = 0;
  hg_graph *g = new hg_graph();
  HIPTRY(hipMalloc(&g->tables, bytes));
  if (!rc) HIPTRY(hipMemcpy(g->tables, views.data(), vbytes, hipMemcpyHostToDevice));
  if (!rc) HIPTRY(hipMemcpy((char *)g->tables + vpad, first.data(), first.size() * sizeof(unsigned),
                            hipMemcpyHostToDevice));
  const GridView *dv = (const GridView *)g->tables;
  const unsigned *dfirst = (const unsigned *)((char *)g->tables + vpad);
  const int cm = ctxs[0]->c2p_mode;
  const int mode = cm == HG_C2P_MIXED ? 1 : cm == HG_C2P_LEARNED ? 2 : 0;
  hipStream_t cap = nullptr;
  if (!rc) HIPTRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  if (!rc) HIPTRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  if (!rc) {
    auto L = [&](int k, int sub) {
      const GridView *vs = dv + (size_t)sub * n;
      const unsigned *fs = dfirst + (size_t)k * (n + 1);
      if (k == GK_MLP) launch_group_mlp(ctxs[0]->mlp, vs, fs, n, total[k], cap);
      else launch_group(k, vs, fs, n, total[k], mode, cap);
    };
    // one substep of step_ch.c:94-166 on every grid: fluxcalc (directions in
    // turn, every direction's dP/P indicator in one launch), update, the
    // learned start's features and network, inversion, dt on the final
    // substep, bound_prim, post_fixup, bound_prim
    const int order[] = {GK_PREP, GK_DPI, GK_WENO1, GK_FLUX1, GK_WENO2, GK_FLUX2, GK_WENO3, GK_FLUX3,
                         GK_EMF, ucum ? GK_UPDATEU : GK_UPDATE, GK_FEATURES, GK_MLP, GK_INVERT, GK_DT, GK_BOUND1,
                         GK_BOUND2, GK_BOUND3, GK_FXCHECK, GK_FXRESOLVE, GK_FXAVG, GK_FXAPPLY, GK_BOUND1,
                         GK_BOUND2, GK_BOUND3};
    for (int step = 0; step < nsteps; step++)
      for (int sub = 0; sub < ns; sub++)
        for (int k : order) {
          if (k == GK_DT && sub != ns - 1) continue;
          if ((k == GK_FEATURES || k == GK_MLP) && mode != 2) continue;
          L(k, sub);
        }
    HIPTRY(hipGetLastError());
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &graph);
    if (e != hipSuccess && !rc) {
      set_error("hg_graph_create_grouped: capture failed: %s", hipGetErrorString(e));
      rc = 1;
    }
    g->graph = graph;
    if (!rc) HIPTRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  }
  if (cap) (void)hipStreamDestroy(cap);
  if (rc) {
    (void)hg_graph_destroy(g);
    return 1;
  }
  // the host-side record of each context after the captured steps, as
  // enqueue_step leaves it
  for (int i = 0; i < n; i++) {
    ctxs[i]->unew_in_uf = true;
    ctxs[i]->flux_state = views[(size_t)(ns - 1) * n + i].pb;
    ctxs[i]->dg_faces_valid = false;
  }
  *out = g;
  return 0;
}

extern "C" int hg_graph_launch(hg_graph *g, void *stream) {
  if (!g || !g->exec) {
    set_error("hg_graph_launch: no graph");
    return 1;
  }
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error("hg_graph_launch: %s", hipGetErrorString(e));
    return 1;
  }
  return 0;
}

extern "C" int hg_graph_destroy(hg_graph *g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->tables) (void)hipFree(g->tables);
  delete g;
  return 0;
}

// ---------------------------------------------------------------------------
// launch shapes
// ---------------------------------------------------------------------------
static bool valid_bs(int b) { return b == 64 || b == 128 || b == 256; }

static int check_tune(const hg_tune *t) {
  if (!valid_bs(t->bs_flux) || !valid_bs(t->bs_emf) || !valid_bs(t->bs_update) || !valid_bs(t->bs_invert)) {
    set_error("hg_set_tune: block sizes must be 64, 128 or 256");
    return 1;
  }
  if (t->weno_xwaves != 1 && t->weno_xwaves != 2 && t->weno_xwaves != 4) {
    set_error("hg_set_tune: weno_xwaves must be 1, 2 or 4");
    return 1;
  }
  if (t->concurrent != 0 && t->concurrent != 1) {
    set_error("hg_set_tune: concurrent must be 0 or 1");
    return 1;
  }
  if (t->weno_nseg < 0 || t->weno_nseg > 64 || t->weno_nseg3 < 0 || t->weno_nseg3 > 64 || t->weno_xmarch < 0 ||
      t->weno_xmarch > 64) {
    set_error("hg_set_tune: weno_nseg, weno_nseg3 and weno_xmarch must be in [0, 64]");
    return 1;
  }
  return 0;
}

static void to_c(const Tune &a, hg_tune *t) {
  t->bs_flux = a.bs_flux;
  t->bs_emf = a.bs_emf;
  t->bs_update = a.bs_update;
  t->bs_invert = a.bs_invert;
  t->weno_xwaves = a.weno_xwaves;
  t->weno_nseg = a.weno_nseg;
  t->weno_nseg3 = a.weno_nseg3;
  t->weno_xmarch = a.weno_xmarch;
  t->concurrent = a.concurrent;
}

static void from_c(const hg_tune *t, Tune &a) {
  a.bs_flux = t->bs_flux;
  a.bs_emf = t->bs_emf;
  a.bs_update = t->bs_update;
  a.bs_invert = t->bs_invert;
  a.weno_xwaves = t->weno_xwaves;
  a.weno_nseg = t->weno_nseg;
  a.weno_nseg3 = t->weno_nseg3;
  a.weno_xmarch = t->weno_xmarch;
  a.concurrent = t->concurrent;
}

extern "C" int hg_get_tune(const hg_ctx *c, hg_tune *t) {
  if (!c || !t) {
    set_error("hg_get_tune: null argument");
    return 1;
  }
  to_c(c->tune, t);
  return 0;
}

extern "C" int hg_set_tune(hg_ctx *c, const hg_tune *t) {
  if (!c || !t) {
    set_error("hg_set_tune: null argument");
    return 1;
  }
  if (check_tune(t)) return 1;
  from_c(t, c->tune);
  return 0;
}

// The state a step reads or writes, saved before tuning and restored after:
// the primitives and intermediate states, the conserved arrays (the
// finite-volume scheme's carried state, ucum, and uf, which holds unew for
// hg_rdump_write after a step), the failure codes, dt and the counters, with
// the host flags that describe them.  Trials start from the saved p, dt and
// carried state, so that every trial times the same work.
struct Snapshot {
  std::vector<std::pair<void *, size_t>> arrays;  // device array, bytes
  std::vector<size_t> offs;
  char *buf = nullptr;
  bool ucons_stale = true, unew_in_uf = false, dg_faces_valid = false;
  const double *flux_state = nullptr;
  double dt = 0.;
  unsigned long long cnt[2] = {0, 0};
  size_t carried = 0;                              // arrays[0..carried): restored before every trial
};

static int snapshot_take(hg_ctx *c, Snapshot &sn, hipStream_t s) {
  int rc = 0;
  const size_t st = (size_t)c->d.spl * NPR * sizeof(double);
  sn.arrays.push_back({c->p, st});
  if (c->ucons) sn.arrays.push_back({c->ucons, st});
  sn.carried = sn.arrays.size();
  for (double *a : {c->pk, c->pk2, c->ucum, c->uf})
    if (a) sn.arrays.push_back({a, st});
  sn.arrays.push_back({c->pflag, (size_t)c->d.spl});
  size_t tot = 0;
  for (auto &a : sn.arrays) {
    sn.offs.push_back(tot);
    tot += (a.second + 255) & ~(size_t)255;
  }
  HIPTRY(hipMalloc((void **)&sn.buf, tot));
  if (rc) { sn.buf = nullptr; return rc; }
  HIPTRY(hipStreamSynchronize(s));
  for (size_t i = 0; i < sn.arrays.size(); i++)
    HIPTRY(hipMemcpy(sn.buf + sn.offs[i], sn.arrays[i].first, sn.arrays[i].second, hipMemcpyDeviceToDevice));
  HIPTRY(hipMemcpy(&sn.dt, c->dt, sizeof(double), hipMemcpyDeviceToHost));
  HIPTRY(hipMemcpy(sn.cnt, c->counters, sizeof sn.cnt, hipMemcpyDeviceToHost));
  sn.ucons_stale = c->ucons_stale;
  sn.unew_in_uf = c->unew_in_uf;
  sn.dg_faces_valid = c->dg_faces_valid;
  sn.flux_state = c->flux_state;
  return rc;
}

// restore the first n saved arrays (asynchronously on s), dt and the carried
// state's flag
static int snapshot_restore(hg_ctx *c, const Snapshot &sn, size_t n, hipStream_t s) {
  int rc = 0;
  for (size_t i = 0; i < n; i++)
    HIPTRY(hipMemcpyAsync(sn.arrays[i].first, sn.buf + sn.offs[i], sn.arrays[i].second, hipMemcpyDeviceToDevice, s));
  HIPTRY(hipMemcpyAsync(c->dt, &sn.dt, sizeof(double), hipMemcpyHostToDevice, s));
  c->ucons_stale = sn.ucons_stale;
  return rc;
}

// time nsteps steps from the saved state
static int time_steps(hg_ctx *c, const Snapshot &sn, int nsteps, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                      float &ms) {
  int rc = snapshot_restore(c, sn, sn.carried, s);
  HIPTRY(hipEventRecord(e0, s));
  for (int k = 0; k < nsteps; k++) enqueue_step(c, s);
  HIPTRY(hipGetLastError());
  HIPTRY(hipEventRecord(e1, s));
  HIPTRY(hipEventSynchronize(e1));
  HIPTRY(hipEventElapsedTime(&ms, e0, e1));
  return rc;
}

extern "C" int hg_autotune(hg_ctx *c, int nsteps, hg_tune *best, void *stream) {
  if (!c || !c->have_geometry) {
    set_error("hg_autotune: context missing or without geometry");
    return 1;
  }
  if (nsteps < 1) nsteps = 1;
  hipStream_t s = (hipStream_t)stream;
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Snapshot sn;
  // no diagnostics recording during the trials: the recorded boundary terms
  // stay those of the caller's last step
  const bool rec = c->dg_record;
  c->dg_record = false;
  HIPTRY(hipEventCreate(&e0));
  HIPTRY(hipEventCreate(&e1));
  if (!rc) rc = snapshot_take(c, sn, s);
  Tune cur = c->tune;
  if (!rc) {
    // one warm-up, then coordinate descent over the shapes, one at a time
    float ms = 0.f;
    rc = time_steps(c, sn, nsteps, s, e0, e1, ms);
    float best_ms = 0.f;
    if (!rc) rc = time_steps(c, sn, nsteps, s, e0, e1, best_ms);
    struct Axis {
      int Tune::*field;
      std::vector<int> values;
    };
    std::vector<Axis> axes = {{&Tune::bs_flux, {64, 128, 256}},   {&Tune::bs_update, {64, 128, 256}},
                              {&Tune::bs_invert, {64, 128, 256}}, {&Tune::bs_emf, {128, 256}},
                              {&Tune::weno_xwaves, {1, 2, 4}},    {&Tune::weno_nseg, {1, 2, 3, 4, 6}},
                              {&Tune::weno_xmarch, {0, 2, 4}}, {&Tune::weno_nseg3, {1, 2, 3, 4}}};
    if (c->cfg.lim == HG_LIM_MC) axes.resize(4);
    else if (c->d.n3 == 1) axes.pop_back();  // no reconstruction along k
    for (const Axis &ax : axes) {
      if (rc) break;
      for (int v : ax.values) {
        if (cur.*(ax.field) == v) continue;
        Tune trial = cur;
        trial.*(ax.field) = v;
        c->tune = trial;
        rc = time_steps(c, sn, nsteps, s, e0, e1, ms);
        if (rc) break;
        // a shape must win by 1% to replace the current one: timing noise
        // of a few steps is about that size
        if (ms < 0.99f * best_ms) {
          best_ms = ms;
          cur = trial;
        }
      }
    }
    c->tune = cur;
  }
  // the state, dt, counters and flags as they were
  if (sn.buf) {
    int r2 = snapshot_restore(c, sn, sn.arrays.size(), s);
    if (!rc) rc = r2;
    HIPTRY(hipStreamSynchronize(s));
    HIPTRY(hipMemcpy(c->counters, sn.cnt, sizeof sn.cnt, hipMemcpyHostToDevice));
    c->unew_in_uf = sn.unew_in_uf;
    c->dg_faces_valid = sn.dg_faces_valid;
    c->flux_state = sn.flux_state;
    (void)hipFree(sn.buf);
  }
  c->dg_record = rec;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc && best) to_c(c->tune, best);
  return rc;
}
