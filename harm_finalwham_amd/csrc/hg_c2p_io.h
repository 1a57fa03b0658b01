// Device helpers shared by the batched inversion kernels (hg_c2p_batch.hip,
// hg_mlp.hip): loading and storing one zone of a structure-of-arrays batch,
// the compacted list of zones left to the reference path, and the
// mixed-precision inversion of one zone.
#pragma once
#include <hip/hip_runtime.h>

#include "hg_c2p.h"
#include "hg_host.h"

namespace hg {

// one zone of a batch: the symmetric metric from its upper triangle (the
// reference's metric tables are symmetric), gdet, U, the guess's u and u~^i
__device__ __forceinline__ void load_batch_zone(int64_t n, int64_t t, const double *__restrict__ U,
                                                const double *__restrict__ gcov, const double *__restrict__ gcon,
                                                const double *__restrict__ gdet, const double *__restrict__ pr,
                                                Geom &G, double *u, double *p) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = a; c < 4; c++) {
      G.gcov[a][c] = G.gcov[c][a] = gcov[(a * 4 + c) * n + t];
      G.gcon[a][c] = G.gcon[c][a] = gcon[(a * 4 + c) * n + t];
    }
  G.gpert0 = 0.;
  G.g = gdet[t];
#pragma unroll
  for (int pl = 0; pl < NPR; pl++) u[pl] = U[pl * n + t];
  p[RHO] = 0.;
#pragma unroll
  for (int pl = UU; pl <= U3; pl++) p[pl] = pr[pl * n + t];
#pragma unroll
  for (int pl = B1; pl <= B3; pl++) p[pl] = 0.;
}

// the solution: the field always, rho .. u~^3 where the inversion assigned
// them (elsewhere the guess stays in place, as the reference leaves it)
__device__ __forceinline__ void store_batch_zone(int64_t n, int64_t t, double *__restrict__ pr, const double *p,
                                                 bool assigned) {
  if (assigned)
#pragma unroll
    for (int pl = RHO; pl <= U3; pl++) pr[pl * n + t] = p[pl];
#pragma unroll
  for (int pl = B1; pl <= B3; pl++) pr[pl * n + t] = p[pl];
}

// the reference assigns rho .. u~^3 for codes 0 and the sign failures
// 9, 10, 11 (utoprim_2d_final.c:418-431), not for the others
__device__ __forceinline__ bool code_assigns(int code) {
  return code == 0 || code == FAIL_RHONEG || code == FAIL_UNEG || code == FAIL_RHOUNEG;
}

// lane l of the wavefront appends zone t to the list when fb is set (one
// atomic per wavefront)
__device__ __forceinline__ void list_append(bool fb, int64_t t, unsigned long long *ctr, uint32_t *list) {
  const int lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(fb);
  if (!m) return;
  const int first = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == first) base = atomicAdd(&ctr[C2P_LISTLEN], (unsigned long long)__popcll(m));
  base = __shfl(base, first, 64);
  if (fb) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)t;
}

// the FP32 bulk, polish and double-double primitives of one zone whose
// set-up succeeded; false: the zone goes to the list
__device__ __forceinline__ bool mixed_zone(const C2PInv &c, double igr, double x0, double x1, double *p, int &it,
                                           int &flag) {
  if (!newton_f32(c, igr, x0, x1, it)) return false;
  if (!c2p_polish(c, igr, x0, x1, POLISH_MAX_F64, it)) return false;
  double out[NPR];
  bool assign;
  flag = c2p_finish_dd(c, x0, igr, out, assign);
  if (assign)
#pragma unroll
    for (int i = 0; i < B1; i++) p[i] = out[i];
  return true;
}

}  // namespace hg
