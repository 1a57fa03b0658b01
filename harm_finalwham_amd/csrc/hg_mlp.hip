// Learned Newton initial guess for the Noble 2D inversion.
//
// A multilayer perceptron 8 -> 64 -> 64 -> 1 (tanh) maps dimensionless
// invariants of the conserved state, the metric and the plain guess to the
// correction log(W / W_guess), W being the enthalpy-like unknown of the 2D
// scheme and W_guess the reference's start (from the previous primitives).
// Weights live in FP32 (training) with BF16 copies for the matrix cores: the
// two hidden layers run on gfx950 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulate), 16 inversions per tile.
//
//   k_mlp_features  FP64 set-up of the inversion -> 8 features per zone, and
//                   the training target log(W_solved / W_guess)
//   k_mlp_forward   features -> prediction (MFMA; each wavefront keeps the
//                   network in registers and runs over 16-sample tiles)
//   k_mlp_backward  forward + backward of a 64-sample chunk on the vector
//                   units (FP32), per-block weight gradients added to grad[]
//   k_mlp_adam      Adam update of the FP32 weights
//   k_mlp_refresh   BF16 copies of the hidden-layer weights
//   k_c2p_learned   the mixed-precision inversion (FP32 Newton, polish,
//                   double-double primitives) from the learned start where
//                   the network predicts a large correction and the start's
//                   first FP64 Newton step passes a size check, else from
//                   the plain start; a zone that does not converge joins
//                   the list that the reference path (HG_C2P_REFINED)
//                   inverts from the plain guess
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "harmgpu.h"
#include "hg_c2p.h"
#include "hg_c2p_io.h"
#include "hg_host.h"

using namespace hg;

namespace {

constexpr int NIN = 8, NH = 64;
// parameter block (FP32): W1[8][64] b1[64] W2[64][64] b2[64] W3[64] b3 mu[8] sd[8]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, OFF_MU = OFF_B3 + 1, OFF_SD = OFF_MU + NIN,
              NPARAM = OFF_SD + NIN;
// trainable parameters are [0, OFF_MU)
constexpr int NTRAIN = OFF_MU;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// W = (rho + gam u) (1 + u~^2) of primitives p
__device__ __forceinline__ double W_of_prims(const Geom &G, const double *p, double gam) {
  double utsq = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) utsq += G.gcov[i][j] * p[U1 + i - 1] * p[U1 + j - 1];
  return (p[RHO] + gam * p[UU]) * (1.0 + utsq);
}

// features X[n][8] from the plain guess; with solved primitives, the
// training target y = log(W_solved / W_guess) in FP64 rounded once to FP32
__global__ void __launch_bounds__(256) k_mlp_features(int64_t n, double gam, const double *__restrict__ U,
                                                      const double *__restrict__ gcov,
                                                      const double *__restrict__ gcon,
                                                      const double *__restrict__ gdet,
                                                      const double *__restrict__ guess,
                                                      const double *__restrict__ solved, float *__restrict__ X,
                                                      float *__restrict__ y, int *__restrict__ ok) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  Geom G;
  double u[NPR], p[NPR], Ug[NPR];
  load_batch_zone(n, t, U, gcov, gcon, gdet, guess, G, u, p);
  C2PInv c;
  double W_last, alpha;
  int code = c2p_setup_zone(u, G, gam, p, c, W_last, alpha, Ug);
  float x[NIN];
  c2p_features(c, W_last, alpha, gam, x);
  bool good = (code == 0);
#pragma unroll
  for (int k = 0; k < NIN; k++) {
    good = good && isfinite(x[k]);
    X[t * NIN + k] = x[k];
  }
  if (solved) {
    double ps[NPR];
#pragma unroll
    for (int pl = 0; pl < NPR; pl++) ps[pl] = solved[pl * n + t];
    double yt = log(W_of_prims(G, ps, gam) / W_last);
    good = good && isfinite(yt);
    if (y) y[t] = (float)yt;
  }
  if (ok) ok[t] = good ? 1 : 0;
}

// ---------------------------------------------------------------------------
// forward pass on the matrix cores: one wavefront, 16 samples per MFMA tile.
// v_mfma_f32_16x16x32_bf16 lane layout (lane l, q = l / 16, r = l % 16):
//   A (16 x 32):  A[r][8q + j], j = 0..7
//   B (32 x 16):  B[8q + j][r]
//   C (16 x 16):  C[4q + i][r], i = 0..3
// The network input x enters as a pair of BF16 numbers hi = bf16(x),
// lo = bf16(x - hi) against the same BF16 weights in both halves of K, so
// the product keeps ~16 bits of the input (the weights are BF16, the
// accumulation FP32): layer 1 has K = 32 = [x_hi(8) x_lo(8) 0(16)].  The
// hidden activations enter layer 2 as BF16 (K = 64).
// ---------------------------------------------------------------------------
struct MlpDev {
  const float *w;       // FP32 parameters (biases, W3, normalisation)
  const __bf16 *w1b;    // BF16 W1 B-fragments: [nt 4][lane 64][j 8] of the K = 32 layer-1 matrix
  const __bf16 *w2b;    // BF16 W2 B-fragments: [ks 2][nt 4][lane 64][j 8]
};

// the network as one wavefront holds it: the B-fragments of both hidden
// layers (lane l = 16 q + r holds rows 8 q .. 8 q + 7 of column 16 nt + r),
// and the biases and output weights of the lane's columns 16 nt + r
struct MlpRegs {
  bf16x8 w1[4], w2[8];
  float b1[4], b2[4], w3[4], b3;
};

__device__ __forceinline__ void mlp_load(const MlpDev &m, MlpRegs &R) {
  const int l = threadIdx.x & 63, r = l & 15;
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    R.w1[nt] = *(const bf16x8 *)(m.w1b + (nt * 64 + l) * 8);
    R.b1[nt] = m.w[OFF_B1 + 16 * nt + r];
    R.b2[nt] = m.w[OFF_B2 + 16 * nt + r];
    R.w3[nt] = m.w[OFF_W3 + 16 * nt + r];
  }
#pragma unroll
  for (int f = 0; f < 8; f++) R.w2[f] = *(const bf16x8 *)(m.w2b + (f * 64 + l) * 8);
  R.b3 = m.w[OFF_B3];
}

__device__ __forceinline__ __bf16 bf_hi(float v) { return (__bf16)v; }

// tanh as 1 - 2 / (exp(2x) + 1) with the hardware exponential and
// reciprocal (absolute error ~1e-7; saturates to +-1 without overflow)
__device__ __forceinline__ float act(float x) {
  float e = __expf(2.0f * x);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ __bf16 bf_lo(float v) { return (__bf16)(v - (float)(__bf16)v); }

// 16 samples: xin[16][8] (normalised) -> y[16]; h[16][64] float LDS scratch
// for the layer-1 activations (the C layout of layer 1 -- sample 4 q + i in
// lane 16 q + r -- is not the A layout of layer 2, sample r in lane 16 q + r);
// layer 3 is summed per lane over its columns, then across the 16 lanes of
// each sample group
__device__ __forceinline__ void forward16(const MlpRegs &R, const float (*xin)[NIN], float (*h)[NH], float *y) {
  const int l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  // layer 1
  bf16x8 a;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    float v = (q < 2) ? xin[r][j] : 0.0f;
    a[j] = (q == 0) ? bf_hi(v) : bf_lo(v);  // q >= 2: bf16(0 - 0) = 0
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, R.w1[nt], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) h[4 * q + i][16 * nt + r] = act(acc[i] + R.b1[nt]);
  }
  __syncthreads();
  // layer 2: K = 64 in two steps, BF16 inputs
  f32x4 acc2[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) acc2[nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    const int kb = 32 * ks + 8 * q;
    bf16x8 a2;
#pragma unroll
    for (int j = 0; j < 8; j++) a2[j] = bf_hi(h[r][kb + j]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      acc2[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, R.w2[4 * ks + nt], acc2[nt], 0, 0, 0);
  }
  // layer 3: lane (q, r) sums its four columns for samples 4 q + i
  float part[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float sacc = 0.f;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) sacc += act(acc2[nt][i] + R.b2[nt]) * R.w3[nt];
    part[i] = sacc;
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1)
#pragma unroll
    for (int i = 0; i < 4; i++) part[i] += __shfl_xor(part[i], off, 64);
  if (r == 0)
#pragma unroll
    for (int i = 0; i < 4; i++) y[4 * q + i] = R.b3 + part[i];
  __syncthreads();
}

// one wavefront per block holds the network in registers and runs over
// tiles of 16 samples (grid-stride)
__device__ __forceinline__ void mlp_forward_body(int64_t n, const float *__restrict__ X, float *__restrict__ y,
                                                 const MlpDev &m, unsigned bid, unsigned nblk) {
  __shared__ float xin[16][NIN];
  __shared__ float h[16][NH];
  __shared__ float yo[16];
  const int l = threadIdx.x;
  MlpRegs R;
  mlp_load(m, R);
  // lane 8 s + k (s < 8) normalises feature k of sample s and of s + 8
  const int k = l & 7;
  const float mk = m.w[OFF_MU + k], sk = m.w[OFF_SD + k];
  for (int64_t base = (int64_t)bid * 16; base < n; base += (int64_t)nblk * 16) {
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int s = half * 8 + (l >> 3);
      const int64_t t = base + s;
      xin[s][k] = (t < n) ? (X[t * NIN + k] - mk) / sk : 0.f;
    }
    __syncthreads();
    forward16(R, xin, h, yo);
    if (l < 16 && base + l < n) y[base + l] = yo[l];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64) k_mlp_forward(int64_t n, const float *__restrict__ X, float *__restrict__ y,
                                                    MlpDev m) {
  mlp_forward_body(n, X, y, m, blockIdx.x, gridDim.x);
}

// grouped: every grid's step features through the one network
__global__ void __launch_bounds__(64) kg_mlp_forward(const GridView *__restrict__ views,
                                                     const unsigned *__restrict__ first, int ng, MlpDev m) {
  const int g_ = group_of(first, ng, blockIdx.x);
  const GridView &v = views[g_];
  const int64_t n = (int64_t)v.d.n1 * v.d.n2 * v.d.n3;
  mlp_forward_body(n, v.lx, v.ly, m, blockIdx.x - first[g_], first[g_ + 1] - first[g_]);
}

// ---------------------------------------------------------------------------
// backward pass, 64 samples per block of one wavefront, FP32 on the vector
// units.  The forward is recomputed from the same BF16-rounded weights and
// layer inputs the matrix-core forward uses (hi + lo input, BF16 hidden
// activations; the roundings are treated as the identity when
// differentiating: the gradient is that of the FP32 master weights).  Activations are staged in LDS (row stride 65 floats, so both
// per-sample rows and per-unit columns are free of bank conflicts):
//   A  thread = sample: x, h1, h2, dy
//   B  thread = unit j: dW3[j], dz2[s][j] in place of h2, db2[j]
//   C  thread = sample: dz1[s][k] = (W2 dz2)[k] (1 - h1^2)
//   D  thread = unit k: dW2[k][:], dW1[:][k], db1[k]
// Each block adds its sums into grad[] once per parameter.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bfr(float x) { return (float)(__bf16)x; }
// the value a layer input carries into the matrix cores: hi + lo
__device__ __forceinline__ float eff(float v) { return (float)bf_hi(v) + (float)bf_lo(v); }

constexpr int LS = NH + 1;  // LDS row stride

__global__ void __launch_bounds__(64) k_mlp_backward(int64_t n, const float *__restrict__ X,
                                                     const float *__restrict__ Y, const float *__restrict__ w,
                                                     float inv_n, float loss_scale, float *__restrict__ grad,
                                                     float *__restrict__ loss) {
  __shared__ float xs[64][NIN + 1];
  __shared__ float h1s[64][LS];
  __shared__ float h2s[64][LS];  // h2, then dz2
  __shared__ float d1s[64][LS];  // dz1
  __shared__ float dys[64];
  const int l = threadIdx.x;
  const int64_t t = (int64_t)blockIdx.x * 64 + l;
  const bool live = t < n;
  // A: forward of sample l
  {
    float x[NIN];
#pragma unroll
    for (int i = 0; i < NIN; i++) {
      x[i] = live ? eff((X[t * NIN + i] - w[OFF_MU + i]) / w[OFF_SD + i]) : 0.f;
      xs[l][i] = x[i];
    }
    for (int j = 0; j < NH; j++) {
      float s = 0.f;
#pragma unroll
      for (int i = 0; i < NIN; i++) s += x[i] * bfr(w[OFF_W1 + i * NH + j]);
      h1s[l][j] = act(s + w[OFF_B1 + j]);
    }
    float yp = w[OFF_B3];
    for (int j = 0; j < NH; j++) {
      float s = 0.f;
      for (int k = 0; k < NH; k++) s += bfr(h1s[l][k]) * bfr(w[OFF_W2 + k * NH + j]);
      float h2 = act(s + w[OFF_B2 + j]);
      h2s[l][j] = h2;
      yp += h2 * w[OFF_W3 + j];
    }
    float r = live ? yp - Y[t] : 0.f;
    dys[l] = 2.f * r * inv_n;
    float lsum = r * r * loss_scale;
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
    if (l == 0) atomicAdd(loss, lsum);
  }
  __syncthreads();
  // B: unit j = l of layer 3 / layer 2 output
  {
    const int j = l;
    const float w3 = w[OFF_W3 + j];
    float gw3 = 0.f, gb2 = 0.f;
    for (int s = 0; s < 64; s++) {
      float h2 = h2s[s][j], dy = dys[s];
      gw3 += dy * h2;
      float dz2 = dy * w3 * (1.f - h2 * h2);
      h2s[s][j] = dz2;
      gb2 += dz2;
    }
    atomicAdd(&grad[OFF_W3 + j], gw3);
    atomicAdd(&grad[OFF_B2 + j], gb2);
    if (l == 0) {
      float gb3 = 0.f;
      for (int s = 0; s < 64; s++) gb3 += dys[s];
      atomicAdd(&grad[OFF_B3], gb3);
    }
  }
  __syncthreads();
  // C: sample l, back through W2 and the first tanh
  for (int k = 0; k < NH; k++) {
    float dh = 0.f;
    for (int j = 0; j < NH; j++) dh += bfr(w[OFF_W2 + k * NH + j]) * h2s[l][j];
    float h1 = h1s[l][k];
    d1s[l][k] = dh * (1.f - h1 * h1);
  }
  __syncthreads();
  // D: unit k = l of the first hidden layer
  {
    const int k = l;
    float g2[NH];
#pragma unroll
    for (int j = 0; j < NH; j++) g2[j] = 0.f;
    float g1[NIN];
#pragma unroll
    for (int i = 0; i < NIN; i++) g1[i] = 0.f;
    float gb1 = 0.f;
    for (int s = 0; s < 64; s++) {
      float hk = bfr(h1s[s][k]);
#pragma unroll
      for (int j = 0; j < NH; j++) g2[j] += hk * h2s[s][j];
      float dz1 = d1s[s][k];
      gb1 += dz1;
#pragma unroll
      for (int i = 0; i < NIN; i++) g1[i] += xs[s][i] * dz1;
    }
#pragma unroll
    for (int j = 0; j < NH; j++) atomicAdd(&grad[OFF_W2 + k * NH + j], g2[j]);
#pragma unroll
    for (int i = 0; i < NIN; i++) atomicAdd(&grad[OFF_W1 + i * NH + k], g1[i]);
    atomicAdd(&grad[OFF_B1 + k], gb1);
  }
}

// BF16 B-fragments of the hidden layers for lane l = 16 q + r, element j:
// layer 1 [nt][l][j] = W1x[8 q + j][16 nt + r] with W1x the K = 32 matrix
// whose rows 0..7 and 8..15 are both W1 (the input's hi and lo halves) and
// rows 16..31 zero; layer 2 [ks][nt][l][j] = W2[32 ks + 8 q + j][16 nt + r]
__device__ __forceinline__ void refresh_bf16(int e, const float *w, __bf16 *w1b, __bf16 *w2b) {
  const int j = e & 7, l = (e >> 3) & 63, q = l >> 4, r = l & 15;
  if (e < 4 * 64 * 8) {
    const int nt = e >> 9, row = 8 * q + j;
    w1b[e] = (row < 2 * NIN) ? (__bf16)w[OFF_W1 + (row % NIN) * NH + 16 * nt + r] : (__bf16)0.0f;
  }
  if (e < 8 * 64 * 8) {
    const int f = e >> 9, ks = f >> 2, nt = f & 3;
    w2b[e] = (__bf16)w[OFF_W2 + (32 * ks + 8 * q + j) * NH + 16 * nt + r];
  }
}

__global__ void k_mlp_adam(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float b1t, float b2t) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < NTRAIN) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[e];
    m1[e] = b1 * m1[e] + (1.f - b1) * gi;
    m2[e] = b2 * m2[e] + (1.f - b2) * gi * gi;
    float mh = m1[e] / (1.f - b1t), vh = m2[e] / (1.f - b2t);
    w[e] -= lr * mh / (sqrtf(vh) + eps);
    g[e] = 0.f;
  }
}

__global__ void k_mlp_refresh(const float *__restrict__ w, __bf16 *__restrict__ w1b, __bf16 *__restrict__ w2b) {
  refresh_bf16(blockIdx.x * blockDim.x + threadIdx.x, w, w1b, w2b);
}

// ---------------------------------------------------------------------------
// inversion with the learned start.  The residual of the Newton system is
// evaluated in FP64 at the plain start (the reference's W from the previous
// primitives) and at the learned one, W_guess exp(y); the iteration starts
// from the smaller.  From there it is the mixed-precision inversion of
// HG_C2P_MIXED (FP32 Newton, FP64 and double-double polish, double-double
// primitives); a zone that does not converge joins the list that
// HG_C2P_REFINED inverts from the plain guess.  counts[0] counts the zones
// solved from the learned start.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_c2p_learned(int64_t n, double gam, const double *__restrict__ U,
                                                    const double *__restrict__ gcov,
                                                    const double *__restrict__ gcon,
                                                    const double *__restrict__ gdet, double *__restrict__ pr,
                                                    int *__restrict__ pflag, int *__restrict__ iters, MlpDev m,
                                                    unsigned long long *__restrict__ ctr,
                                                    uint32_t *__restrict__ list,
                                                    unsigned long long *__restrict__ counts) {
  __shared__ float xin[64][NIN];
  __shared__ float h[16][NH];
  __shared__ float yo[64];
  const int l = threadIdx.x;
  const int64_t t = (int64_t)blockIdx.x * 64 + l;
  const bool live = t < n;
  Geom G;
  double u[NPR], p[NPR], Ug[NPR];
  C2PInv c;
  double W_last = 1., alpha = 1.;
  int code = 1;
  if (live) {
    load_batch_zone(n, t, U, gcov, gcon, gdet, pr, G, u, p);
    code = c2p_setup_zone(u, G, gam, p, c, W_last, alpha, Ug);
    float x[NIN];
    c2p_features(c, W_last, alpha, gam, x);
#pragma unroll
    for (int k = 0; k < NIN; k++) xin[l][k] = (x[k] - m.w[OFF_MU + k]) / m.w[OFF_SD + k];
  } else {
#pragma unroll
    for (int k = 0; k < NIN; k++) xin[l][k] = 0.f;
  }
  __syncthreads();
  MlpRegs R;
  mlp_load(m, R);
  for (int g16 = 0; g16 < 4; g16++) {
    forward16(R, (const float(*)[NIN]) & xin[16 * g16][0], h, &yo[16 * g16]);
    __syncthreads();
  }
  const double igr = (gam - 1.0) / gam;
  bool fb = false, learned = false;
  if (live) {
#pragma unroll
    for (int i = B1; i <= B3; i++) p[i] = Ug[i];
    int it = 0, flag = code;
    if (code == 0) {
      double x0, x1;
      bool cand;
      it = learned_start(c, igr, W_last, yo[l], x0, x1, cand);
      fb = !mixed_zone(c, igr, x0, x1, p, it, flag);
      learned = cand && !fb;
    }
    if (!fb) {
      store_batch_zone(n, t, pr, p, code_assigns(flag));
      pflag[t] = flag;
      if (iters) iters[t] = it;
    }
  }
  list_append(fb, t, ctr, list);
  const unsigned long long ml = __ballot(learned);
  if (l == 0 && ml) atomicAdd(&counts[0], (unsigned long long)__popcll(ml));
}

}  // namespace

struct hg_mlp {
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
  __bf16 *w1b = nullptr, *w2b = nullptr;
  unsigned long long *counts = nullptr;  // [learned starts, fallbacks]
  int step = 0;
};

static MlpDev dev_view(const hg_mlp *m) { return MlpDev{m->w, m->w1b, m->w2b}; }

// the forward kernel's one-wavefront blocks: one per 16-sample tile up to
// 8192 (32 per CU), each then loops over its share of the tiles
static unsigned forward_blocks(int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  return (unsigned)(tiles < 8192 ? tiles : 8192);
}

#define MLPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

extern "C" int hg_mlp_destroy(hg_mlp *m) {
  if (!m) return 0;
  void *b[] = {m->w, m->g, m->m1, m->m2, m->loss, m->w1b, m->w2b, m->counts};
  for (void *p : b)
    if (p) (void)hipFree(p);
  delete m;
  return 0;
}

extern "C" int hg_mlp_create(uint64_t seed, hg_mlp **out) {
  if (!out) { set_error("hg_mlp_create: null output"); return 1; }
  *out = nullptr;
  hg_mlp *m = new hg_mlp();
  std::vector<float> w(NPARAM, 0.f);
  // Glorot-uniform hidden weights from a 64-bit LCG; unit normalisation
  uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto uni = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * 0x1.0p-53;
  };
  double a1 = sqrt(6.0 / (NIN + NH)), a2 = sqrt(6.0 / (NH + NH)), a3 = sqrt(6.0 / (NH + 1));
  for (int e = 0; e < NIN * NH; e++) w[OFF_W1 + e] = (float)(a1 * (2 * uni() - 1));
  for (int e = 0; e < NH * NH; e++) w[OFF_W2 + e] = (float)(a2 * (2 * uni() - 1));
  for (int e = 0; e < NH; e++) w[OFF_W3 + e] = (float)(a3 * (2 * uni() - 1));
  for (int k = 0; k < NIN; k++) w[OFF_SD + k] = 1.f;
  hipError_t e = hipMalloc((void **)&m->w, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->g, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->m1, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->m2, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->loss, sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->w1b, 32 * NH * sizeof(__bf16));
  if (e == hipSuccess) e = hipMalloc((void **)&m->w2b, NH * NH * sizeof(__bf16));
  if (e == hipSuccess) e = hipMalloc((void **)&m->counts, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemcpy(m->w, w.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->g, 0, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->m1, 0, NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->m2, 0, NPARAM * sizeof(float));
  if (e != hipSuccess) {
    set_error("hg_mlp_create: %s", hipGetErrorString(e));
    hg_mlp_destroy(m);
    return 1;
  }
  hipLaunchKernelGGL(k_mlp_refresh, dim3((NH * NH + 255) / 256), dim3(256), 0, 0, m->w, m->w1b, m->w2b);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error("hg_mlp_create: %s", hipGetErrorString(e));
    hg_mlp_destroy(m);
    return 1;
  }
  *out = m;
  return 0;
}

extern "C" int hg_mlp_nparams(void) { return NPARAM; }

extern "C" int hg_mlp_get_params(const hg_mlp *m, float *host) {
  if (!m || !host) { set_error("hg_mlp_get_params: null argument"); return 1; }
  MLPCHECK(hipMemcpy(host, m->w, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int hg_mlp_set_params(hg_mlp *m, const float *host) {
  if (!m || !host) { set_error("hg_mlp_set_params: null argument"); return 1; }
  for (int k = 0; k < NIN; k++)
    if (!(host[OFF_SD + k] > 0.f)) { set_error("hg_mlp_set_params: feature scales must be > 0"); return 1; }
  MLPCHECK(hipMemcpy(m->w, host, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_refresh, dim3((NH * NH + 255) / 256), dim3(256), 0, 0, m->w, m->w1b, m->w2b);
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipDeviceSynchronize());
  return 0;
}

extern "C" int hg_mlp_features(int64_t n, double gam, const double *U, const double *gcov, const double *gcon,
                               const double *gdet, const double *p_guess, const double *p_solved, float *X,
                               float *y, int32_t *ok, void *stream) {
  if (n < 0 || (n > 0 && (!U || !gcov || !gcon || !gdet || !p_guess || !X || (y && !p_solved)))) {
    set_error("hg_mlp_features: bad arguments");
    return 1;
  }
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, gam, U, gcov, gcon,
                     gdet, p_guess, p_solved, X, y, ok);
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipStreamSynchronize(s));
  return 0;
}

unsigned hg::group_blocks_mlp(const GridView &v) {
  return forward_blocks((int64_t)v.d.n1 * v.d.n2 * v.d.n3);
}

void hg::launch_group_mlp(const hg_mlp *m, const GridView *views, const unsigned *first, int n, unsigned total,
                          hipStream_t s) {
  if (total == 0) return;
  hipLaunchKernelGGL(kg_mlp_forward, dim3(total), dim3(64), 0, s, views, first, n, dev_view(m));
}

void hg::launch_mlp_forward(const hg_mlp *m, int64_t n, const float *X, float *y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3(forward_blocks(n)), dim3(64), 0, s, n, X, y, dev_view(m));
}

extern "C" int hg_mlp_forward(const hg_mlp *m, int64_t n, const float *X, float *y, void *stream) {
  if (!m || n < 0 || (n > 0 && (!X || !y))) { set_error("hg_mlp_forward: bad arguments"); return 1; }
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_forward, dim3(forward_blocks(n)), dim3(64), 0, s, n, X, y, dev_view(m));
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipStreamSynchronize(s));
  return 0;
}

// gradients of the mean squared error over n samples (no update); grad_host
// (may be NULL) receives them, *loss the error
extern "C" int hg_mlp_gradient(hg_mlp *m, int64_t n, const float *X, const float *Y, float *grad_host,
                               float *loss, void *stream) {
  if (!m || n <= 0 || !X || !Y) { set_error("hg_mlp_gradient: bad arguments"); return 1; }
  hipStream_t s = (hipStream_t)stream;
  MLPCHECK(hipMemsetAsync(m->g, 0, NPARAM * sizeof(float), s));
  MLPCHECK(hipMemsetAsync(m->loss, 0, sizeof(float), s));
  hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, X, Y, m->w,
                     (float)(1.0 / (double)n), (float)(1.0 / (double)n), m->g, m->loss);
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipStreamSynchronize(s));
  if (grad_host) MLPCHECK(hipMemcpy(grad_host, m->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) MLPCHECK(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  MLPCHECK(hipMemset(m->g, 0, NPARAM * sizeof(float)));
  return 0;
}

// epochs of Adam over the n samples in minibatches of `batch` (in order);
// *loss receives the mean minibatch loss of the last epoch
extern "C" int hg_mlp_train(hg_mlp *m, int64_t n, const float *X, const float *Y, int epochs, int batch,
                            float lr, float *loss, void *stream) {
  if (!m || n <= 0 || !X || !Y || epochs < 0 || batch < 64) {
    set_error("hg_mlp_train: bad arguments (batch >= 64)");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t nb = (n + batch - 1) / batch;
  float last = 0.f;
  for (int ep = 0; ep < epochs; ep++) {
    MLPCHECK(hipMemsetAsync(m->loss, 0, sizeof(float), s));
    for (int64_t b = 0; b < nb; b++) {
      int64_t off = b * batch, cnt = (off + batch <= n) ? batch : n - off;
      hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)((cnt + 63) / 64)), dim3(64), 0, s, cnt,
                         X + off * NIN, Y + off, m->w, (float)(1.0 / (double)cnt),
                         (float)(1.0 / (double)cnt / (double)nb), m->g, m->loss);
      m->step++;
      float b1t = powf(0.9f, (float)m->step), b2t = powf(0.999f, (float)m->step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((NTRAIN + 255) / 256), dim3(256), 0, s, m->w, m->g, m->m1, m->m2, lr,
                         b1t, b2t);
      hipLaunchKernelGGL(k_mlp_refresh, dim3((NH * NH + 255) / 256), dim3(256), 0, s, m->w, m->w1b, m->w2b);
    }
    MLPCHECK(hipGetLastError());
    MLPCHECK(hipMemcpyAsync(&last, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    MLPCHECK(hipStreamSynchronize(s));
  }
  if (loss) *loss = last;
  return 0;
}

extern "C" int hg_utoprim_batch_learned(const hg_mlp *m, int64_t n, double gam, const double *U,
                                        const double *gcov, const double *gcon, const double *gdet,
                                        const double *gcovpert, double *pr, int32_t *pflag, int32_t *iters,
                                        int64_t *counts, void *stream) {
  (void)gcovpert;
  if (!m || n < 0 || n > (int64_t)0xFFFFFFFFll) { set_error("hg_utoprim_batch_learned: bad arguments"); return 1; }
  if (n == 0) {
    if (counts) counts[0] = counts[1] = 0;
    return 0;
  }
  if (!U || !gcov || !gcon || !gdet || !pr || !pflag) {
    set_error("hg_utoprim_batch_learned: null argument");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  // scratch of this call only (see hg_utoprim_batch)
  void *scr = nullptr;
  MLPCHECK(hipMallocAsync(&scr, c2p_scratch_bytes(n), s));
  unsigned long long *ctr = (unsigned long long *)scr;
  launch_c2p_reset(scr, s);
  hipError_t e = hipMemsetAsync(m->counts, 0, 2 * sizeof(unsigned long long), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_c2p_learned, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, gam, U, gcov, gcon, gdet,
                       pr, pflag, iters, dev_view(m), ctr, (uint32_t *)(ctr + C2P_NCTR), m->counts);
    if (launch_c2p_listed(n, gam, U, gcov, gcon, gdet, pr, pflag, iters, scr, s)) e = hipErrorUnknown;
  }
  unsigned long long h[2] = {0, 0}, len = 0;
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, m->counts, sizeof h, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&len, ctr + C2P_LISTLEN, sizeof len, hipMemcpyDeviceToHost, s);
  hipError_t e2 = hipFreeAsync(scr, s);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error("hg_utoprim_batch_learned: %s", hipGetErrorString(e));
    return 1;
  }
  if (counts) {
    counts[0] = (int64_t)h[0];
    counts[1] = (int64_t)len;
  }
  return 0;
}
