// Device data layout of one grid.
//
// Primitives, fluxes and conserved scratch are structure-of-arrays with the
// radial index fastest: a[pl][k][j][i], i in [0,m1) = active [nb1, nb1+n1)
// plus nb1 ghost layers each side (likewise j, k).  Directional sweeps map
// threads to i so every load and store of a wavefront is contiguous.
//
// Geometry is phi independent and stored once per (i,j) on the (m1+1) x
// (m2+1) table, structure-of-arrays per location and component:
//   full[(loc*GCOMP + c)][j][i]  for loc in {CENT, FACE1, FACE2, FACE3},
//        c = gcov[16], gcon[16], gcovpert_tt, gdet        (GCOMP = 34)
//   corn[c][j][i]                gdet at CORN1, CORN2, CORN3
//   conn[c][j][i] (m1 x m2)      Gamma^a_bc at CENT (64)
//   floor[c][j][i] (m1 x m2)     RHOMIN r^-1.5, UUMIN r^-2.5 at CENT
#pragma once
#include <stdint.h>

namespace hg {

// primitive indices (global.h:916-923)
constexpr int RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7, NPR = 8;
// energy regions of the diagnostics (global.h:436, DOJETDIAG 1 in the torus
// file): the global region and two jet regions
constexpr int NENERREGIONS = 3;

constexpr int GCOMP = 34;
constexpr int GL_CENT = 0, GL_FACE1 = 1, GL_FACE2 = 2, GL_FACE3 = 3;

struct Dims {
  int n1, n2, n3;
  int nb1, nb2, nb3;
  int m1, m2, m3;
  int g1, g2;            // geometry table extents (m1+1, m2+1)
  int64_t sj, sk, spl;   // strides of j, k, pl in a[pl][k][j][i]
  int64_t gplane;        // g1*g2
  int64_t cplane;        // m1*m2
  __host__ __device__ int64_t idx(int i, int j, int k) const {  // i,j,k relative to first active zone
    return (int64_t)(i + nb1) + (int64_t)(j + nb2) * sj + (int64_t)(k + nb3) * sk;
  }
  __host__ __device__ int64_t gidx(int i, int j) const { return (int64_t)(i + nb1) + (int64_t)(j + nb2) * g1; }
  __host__ __device__ int64_t cidx(int i, int j) const { return (int64_t)(i + nb1) + (int64_t)(j + nb2) * m1; }
};

inline Dims make_dims(int n1, int n2, int n3, int nb) {
  Dims d;
  d.n1 = n1; d.n2 = n2; d.n3 = n3;
  d.nb1 = (n1 > 1) ? nb : 0;
  d.nb2 = (n2 > 1) ? nb : 0;
  d.nb3 = (n3 > 1) ? nb : 0;
  d.m1 = n1 + 2 * d.nb1;
  d.m2 = n2 + 2 * d.nb2;
  d.m3 = n3 + 2 * d.nb3;
  d.g1 = d.m1 + 1;
  d.g2 = d.m2 + 1;
  d.sj = d.m1;
  d.sk = (int64_t)d.m1 * d.m2;
  d.spl = d.sk * d.m3;
  d.gplane = (int64_t)d.g1 * d.g2;
  d.cplane = (int64_t)d.m1 * d.m2;
  return d;
}

}  // namespace hg
