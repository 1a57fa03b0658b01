/* placeholder, replaced by the WENO5BND restatement */
#include "or_internal.h"
void or_weno5bnd_line(const or_grid *G, const or_dims *d, const double *p, int dir, int i, int j,
                      int k, double *pleft, double *pright)
{
  (void)G; (void)d; (void)p; (void)dir; (void)i; (void)j; (void)k; (void)pleft; (void)pright;
}
