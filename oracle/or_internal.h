/* Internal helpers of the CPU oracle (test infrastructure only). */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H
#include "harm_oracle.h"
#include <stddef.h>

/* geometry location slots in the tables handed to the oracle */
#define OL_FACE1 0
#define OL_FACE2 1
#define OL_FACE3 2
#define OL_CORN1 3
#define OL_CORN2 4
#define OL_CORN3 5
#define OL_CENT 6
#define OL_NLOC 7

typedef struct {
  int n1, n2, n3;
  int nb1, nb2, nb3;    /* ghost width per direction (0 for a direction of size 1) */
  int m1, m2, m3;       /* allocated extent per direction n + 2 nb */
  int g1, g2;           /* geometry table extent (m + 1 in each of the first two dirs) */
} or_dims;

static inline or_dims or_get_dims(const or_grid *G)
{
  or_dims d;
  d.n1 = G->n1; d.n2 = G->n2; d.n3 = G->n3;
  d.nb1 = (G->n1 > 1) ? G->nb : 0;
  d.nb2 = (G->n2 > 1) ? G->nb : 0;
  d.nb3 = (G->n3 > 1) ? G->nb : 0;
  d.m1 = d.n1 + 2 * d.nb1;
  d.m2 = d.n2 + 2 * d.nb2;
  d.m3 = d.n3 + 2 * d.nb3;
  d.g1 = d.m1 + 1;
  d.g2 = d.m2 + 1;
  return d;
}

/* zone index of (i,j,k) with i in [-nb1, n1+nb1) etc. */
static inline ptrdiff_t or_zidx(const or_dims *d, int i, int j, int k)
{
  return (((ptrdiff_t)(i + d->nb1) * d->m2 + (j + d->nb2)) * d->m3 + (k + d->nb3));
}

/* fetch the geometry of location `loc` at (i,j) (phi independent) */
static inline void or_load_geom(const or_grid *G, const or_dims *d, int i, int j, int loc, or_geom *g)
{
  ptrdiff_t gi = ((ptrdiff_t)(i + d->nb1) * d->g2 + (j + d->nb2)) * OL_NLOC + loc;
  int a, b;
  for (a = 0; a < 4; a++) {
    for (b = 0; b < 4; b++) {
      g->gcov[a][b] = G->gcov[gi * 16 + a * 4 + b];
      g->gcon[a][b] = G->gcon[gi * 16 + a * 4 + b];
    }
    g->gcovpert[a] = G->gcovpert[gi * 4 + a];
  }
  g->g = G->gdet[gi];
}

static inline double or_gdet(const or_grid *G, const or_dims *d, int i, int j, int loc)
{
  return G->gdet[((ptrdiff_t)(i + d->nb1) * d->g2 + (j + d->nb2)) * OL_NLOC + loc];
}

/* phys.c:1023 gamma_calc (VELREL4, JONCHECKS): 0 ok, 1 failure */
int or_gamma_calc(const double *pr, const or_geom *g, double *gamma);

/* or_weno.c: WENO5BND interface states along one pencil (interpline.c:517) */
void or_weno5bnd_line(const or_grid *G, const or_dims *d, const double *p, int dir, int i, int j,
                      int k, double *pleft, double *pright);

#endif
