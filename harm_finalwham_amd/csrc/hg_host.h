// Internal host-side definitions shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <vector>

#include "harmgpu.h"
#include "hg_grid.h"
#include "hg_kernels.h"

struct hg_ctx {
  hg_config cfg;
  hg::Dims d;
  double *p = nullptr, *pk = nullptr;        // primitives and the first intermediate state
  double *pk2 = nullptr, *ucum = nullptr;    // RK4 only: second intermediate state, ucum
  // finite-volume scheme: carried conserved state; stale until derived from
  // p after hg_set_prims / hg_init_torus
  double *ucons = nullptr;
  bool ucons_stale = true;
  // the standard scheme's unew (the last substep's ucum) sits in uf after a
  // step until primitives are set again
  bool unew_in_uf = false;
  double *F[3] = {nullptr, nullptr, nullptr};
  double *emf[3] = {nullptr, nullptr, nullptr};
  double *pl = nullptr, *pr = nullptr;       // WENO5BND interface states along i
  double *pl2 = nullptr, *pr2 = nullptr;     // ... along j
  double *pl3 = nullptr, *pr3 = nullptr;     // ... along k
  // the j and k reconstruction + flux run on these streams beside the i one
  hipStream_t xs[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  bool capturing = false;                    // inside hg_graph_create: no side streams
  double *aux = nullptr;                     // WENO5BND per-zone dP/P inputs
  // the step's inversion (hg_set_inversion): mode, and for the learned start
  // the network with its per-zone features and predictions
  int c2p_mode = HG_C2P_FP64;
  const hg_mlp *mlp = nullptr;
  float *lx = nullptr, *ly = nullptr;
  double *uf = nullptr;                      // updated conserved variables before the inversion
  double *geo = nullptr, *corn = nullptr, *conn = nullptr, *floors = nullptr;
  double *dt = nullptr;
  unsigned long long *ndt_bits = nullptr, *counters = nullptr;
  int *err = nullptr;
  // post_fixup: inversion codes and checked codes (int8, layout of one
  // primitive), candidate and fix lists (hg_fixup.hip)
  int8_t *pflag = nullptr, *pchk = nullptr;
  int *fx_cand_z = nullptr, *fx_fix_z = nullptr;
  uint16_t *fx_cand_bits = nullptr;
  int8_t *fx_cand_new = nullptr, *fx_fix_flag = nullptr;
  double *fx_fix_v = nullptr;
  int *fx_fail_z = nullptr;
  unsigned *fx_count = nullptr;
  // diagnostics (hg_diag.hip): scratch on first use, time-integrated
  // boundary fluxes (diag.c pcum) of the global region
  double *dg_buf = nullptr;
  double pcum[6][8] = {};
  const double *flux_state = nullptr;        // the primitives the last fluxcalc read
  bool dg_record = false, dg_faces_valid = false;
  bool have_geometry = false;
  hg::Tune tune;                             // launch shapes (hg_autotune)
  // modified Kerr-Schild grid parameters when the geometry was built by
  // hg_build_mks_geometry (the torus set-up needs the coordinates)
  bool have_mks = false;
  double mks_a = 0, mks_Rin = 0, mks_Rout = 0, mks_R0 = 0, mks_hslope = 0;
  // kernel-class timing (hg_profile): event pool and recorded intervals
  struct ProfInterval {
    int cls;
    size_t start, stop;
  };
  bool prof_on = false;
  int prof_open = -1;
  size_t prof_used = 0;
  std::vector<hipEvent_t> prof_events;
  std::vector<ProfInterval> prof_intervals;
};

namespace hg {
void set_error(const char *fmt, ...);
int upload_geometry_tables(hg_ctx *c, const double *gcov, const double *gcon, const double *gcovpert,
                           const double *gdet, const double *conn, const double *r_cent);
int copy_prims(hg_ctx *c, double *dev, const double *host);
// one step of the context on stream s (kernels only: capturable)
void enqueue_step(hg_ctx *c, hipStream_t s);
// the substeps of one step of the context's scheme: their number (at most 4)
// and each one's pb, pf and coefficients (pi is p on every substep)
int step_substeps(hg_ctx *c, double *pb[4], double *pf[4], RkStage rk[4]);
// finite-volume scheme: derive the carried conserved state from p if p was
// set since (kernels only)
void ensure_conserved(hg_ctx *c, hipStream_t s);
int fetch_prims(hg_ctx *c, const double *dev, double *host);
// returns 1 (with the error set) when the work counter cannot be set up
// post_fixup buffers (hg_fixup.hip)
int fixup_alloc(hg_ctx *c);
void fixup_free(hg_ctx *c);
void diag_free(hg_ctx *c);
int diag_alloc(hg_ctx *c);
void launch_diag_faces(hg_ctx *c, const double *prim, const double *pglob, hipStream_t s);
void launch_diag_stage(hg_ctx *c, int stage, hipStream_t s);
// fixup_checksolution + fixup_utoprim on bounded primitives p (kernels only)
// reset false: the list counters were cleared already (by the substep's
// inversion kernel)
void launch_post_fixup(hg_ctx *c, double *p, int finalstep, hipStream_t s, bool reset = true);
FxBufs fixup_bufs(hg_ctx *c);
// grouped launches of the network's forward pass (hg_mlp.hip): every grid's
// features lx -> predictions ly through one network
unsigned group_blocks_mlp(const GridView &v);
void launch_group_mlp(const hg_mlp *m, const GridView *views, const unsigned *first, int n, unsigned total,
                      hipStream_t s);
// batched inversion (hg_c2p_batch.hip).  The mixed path's scratch holds
// C2P_NCTR counters (work queue, list length, list work queue) and a list of
// n 32-bit zone indices: c2p_scratch_bytes(n) bytes.
constexpr int C2P_WORK = 0, C2P_LISTLEN = 1, C2P_LISTWORK = 2, C2P_NCTR = 3;
size_t c2p_scratch_bytes(int64_t n);
int launch_c2p_batch(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                     const double *gdet, double *pr, int *pflag, int *iters, bool refined, hipStream_t s);
int launch_c2p_batch_mixed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                           const double *gdet, double *pr, int *pflag, int *iters, void *scratch, hipStream_t s);
void launch_c2p_reset(void *scratch, hipStream_t s);
// the zones a mixed or learned pass listed, inverted by HG_C2P_REFINED
int launch_c2p_listed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                      const double *gdet, double *pr, int *pflag, int *iters, void *scratch, hipStream_t s);
void c2p_note_fallbacks(long long nfb);
// the network's predictions y[n] from features X[n][8] (hg_mlp.hip; kernels only)
void launch_mlp_forward(const hg_mlp *m, int64_t n, const float *X, float *y, hipStream_t s);
constexpr int C2P_NFEAT = 8;  // features per inversion of the learned start
}  // namespace hg
