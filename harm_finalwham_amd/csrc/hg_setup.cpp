// Problem set-up on the host: MKS/KS geometry tables and torus initial data.
#include <hip/hip_runtime.h>

#include "harmgpu.h"
#include "hg_host.h"

using namespace hg;

extern "C" int hg_build_mks_geometry(hg_ctx *c, double a, double Rin, double Rout, double R0, double hslope) {
  (void)c; (void)a; (void)Rin; (void)Rout; (void)R0; (void)hslope;
  set_error("hg_build_mks_geometry: not available in this build");
  return 1;
}

extern "C" int hg_init_torus(hg_ctx *c, double rin, double rmax, double kappa, double beta, double randfact,
                             uint64_t seed, double *dt) {
  (void)c; (void)rin; (void)rmax; (void)kappa; (void)beta; (void)randfact; (void)seed; (void)dt;
  set_error("hg_init_torus: not available in this build");
  return 1;
}

extern "C" int hg_make_c2p_batch(int64_t n, double gam, double a, uint64_t seed, double *U, double *gcov,
                                 double *gcon, double *gdet, double *gcovpert, double *p_true, double *p_guess,
                                 void *stream) {
  (void)n; (void)gam; (void)a; (void)seed; (void)U; (void)gcov; (void)gcon; (void)gdet; (void)gcovpert;
  (void)p_true; (void)p_guess; (void)stream;
  set_error("hg_make_c2p_batch: not available in this build");
  return 1;
}

namespace hg {
void launch_c2p_batch_mixed(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                            const double *gdet, const double *gcovpert, double *pr, int *pflag, int *iters,
                            hipStream_t s) {
  (void)n; (void)gam; (void)U; (void)gcov; (void)gcon; (void)gdet; (void)gcovpert; (void)pr; (void)pflag;
  (void)iters; (void)s;
}
}  // namespace hg
