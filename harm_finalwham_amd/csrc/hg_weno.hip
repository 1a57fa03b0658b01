// WENO5BND line reconstruction (interpline.c:517 slope_lim_linetype_c2e).
#include <hip/hip_runtime.h>

#include "hg_kernels.h"

namespace hg {
void launch_weno(int dir, const double *p, const double *geo, double *pleft, double *pright, const Dims &d,
                 double gam, hipStream_t s) {
  (void)dir; (void)p; (void)geo; (void)pleft; (void)pright; (void)d; (void)gam; (void)s;
}
}  // namespace hg
