// mg_sur.hip -- the learned screening surrogate: a nin-64-64-1 tanh network over per-structure descriptor
// features, trained with mean-squared error and Adam, entirely in these kernels.
//
// Features (k_sur_features): for structure s and species q, the mean over the atoms of species q of the
// same-species block of their descriptors (G2 block ind2[0], then G4 block ind4[0][0]: for one species pair
// the ng2 + ng4 functions of the basis), giving nspc * (ng2 + ng4) inputs; zero for a species absent from s.
//
// Network (flat FP32 parameter vector, the order of the reference's per-layer weight blocks, src/nmlp.c:591-608):
//   W0[64][nin] b0[64] W1[64][64] b1[64] w2[64] b2
//   x' = (x - mu) / sig,  h1 = tanh(W0 x' + b0),  h2 = tanh(W1 h1 + b1),  y' = w2 . h2 + b2,  y = ymu + ysig y'
// Loss over a mini-batch of B samples: L = (1/B) sum_b (y'_b - t'_b)^2 with t' = (t - ymu) / ysig.
// Gradients (the chain of the reference's training derivative DE_ANN_PARA, src/nmlp.c:292-354, for this loss):
//   e = 2 (y' - t') / B;  g_b2 = e;  g_w2 = e h2;  d2 = e w2 (1 - h2^2);  g_W1 = d2 h1^T;  g_b1 = d2;
//   d1 = (W1^T d2)(1 - h1^2);  g_W0 = d1 x'^T;  g_b0 = d1.
//
// One wave per block; lane t owns hidden unit t of both layers.  Weights sit in LDS with odd row strides so
// that both the row-per-lane and the column-per-lane sweeps are free of bank conflicts; each block runs a fixed
// chunk of samples and accumulates its parameter gradients in registers, then writes one partial per block.
// The reduction over blocks (fixed order, so results do not depend on timing) is fused with the Adam update.
#include <hip/hip_runtime.h>

#include "mg_internal.h"

#define SUR_H 64
#define SUR_MAXIN 128
#define SUR_S0 (SUR_MAXIN + 1)  // LDS row stride of W0
#define SUR_S1 (SUR_H + 1)      // LDS row stride of W1

namespace {

__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct SurLds {
  float W0[SUR_H * SUR_S0];
  float W1[SUR_H * SUR_S1];
  float x[SUR_MAXIN];
  float h1[SUR_H];
  float d2[SUR_H];
};

__device__ void load_weights(SurLds& L, const float* __restrict__ p, int nin, int lane) {
  for (int e = lane; e < SUR_H * nin; e += 64) L.W0[(e / nin) * SUR_S0 + e % nin] = p[e];
  const float* W1 = p + SUR_H * nin + SUR_H;
  for (int e = lane; e < SUR_H * SUR_H; e += 64) L.W1[(e >> 6) * SUR_S1 + (e & 63)] = W1[e];
}

// forward of one sample; returns y' (wave-uniform) and leaves h1 in LDS and h1/h2 of unit `lane` in registers
__device__ float forward(SurLds& L, const float* __restrict__ p, int nin, int lane, const float* __restrict__ xs,
                         const float* __restrict__ mu, const float* __restrict__ sig, float& h1t, float& h2t) {
  for (int k = lane; k < nin; k += 64) L.x[k] = (xs[k] - mu[k]) / sig[k];
  __syncthreads();
  const float* b0 = p + SUR_H * nin;
  const float* b1 = b0 + SUR_H + SUR_H * SUR_H;
  const float* w2 = b1 + SUR_H;
  const float b2 = w2[SUR_H];
  float a = b0[lane];
  for (int k = 0; k < nin; k++) a += L.W0[lane * SUR_S0 + k] * L.x[k];
  h1t = tanhf(a);
  L.h1[lane] = h1t;
  __syncthreads();
  float c = b1[lane];
#pragma unroll 8
  for (int k = 0; k < SUR_H; k++) c += L.W1[lane * SUR_S1 + k] * L.h1[k];
  h2t = tanhf(c);
  return wsum(w2[lane] * h2t) + b2;
}

}  // namespace

// per-structure features: block per structure, threads over (species, function)
__global__ __launch_bounds__(128) void k_sur_features(int S, const mg_struct_dev* __restrict__ str,
                                                      const int* __restrict__ spc, const float* __restrict__ G, int D,
                                                      int nspc, int ng2, int ng4, int off2, int off4,
                                                      float* __restrict__ feat) {
  const int s = blockIdx.x;
  if (s >= S) return;
  const int a0 = str[s].atom0, n = str[s].natom;
  const int nf = ng2 + ng4;
  for (int f = threadIdx.x; f < nspc * nf; f += blockDim.x) {
    const int q = f / nf, c = f % nf;
    const int col = c < ng2 ? off2 + c : off4 + (c - ng2);
    float sum = 0.0f;
    int cnt = 0;
    for (int i = a0; i < a0 + n; i++)
      if (spc[i] == q) { sum += G[(size_t)i * D + col]; cnt++; }
    feat[(size_t)s * nspc * nf + f] = cnt ? sum / (float)cnt : 0.0f;
  }
}

// fused forward + backward over a mini-batch (indices idx[0..B)), chunk samples per block;
// part[blk][P] receives the block's gradient sum, part_loss[blk] its squared-error sum
__global__ __launch_bounds__(64) void k_sur_grad(int B, int chunk, int nin, const int* __restrict__ idx,
                                                 const float* __restrict__ X, const float* __restrict__ Y,
                                                 const float* __restrict__ p, const float* __restrict__ norm,
                                                 float* __restrict__ part, float* __restrict__ part_loss) {
  __shared__ SurLds L;
  const int lane = threadIdx.x;
  const int P = SUR_H * nin + SUR_H * SUR_H + 3 * SUR_H + 1;  // W0 b0 W1 b1 w2 b2
  const float* mu = norm;
  const float* sig = norm + nin;
  const float ymu = norm[2 * nin], ysig = norm[2 * nin + 1];
  load_weights(L, p, nin, lane);
  const float* w2 = p + SUR_H * nin + SUR_H + SUR_H * SUR_H + SUR_H;
  float gW0[SUR_MAXIN], gW1[SUR_H];
#pragma unroll
  for (int k = 0; k < SUR_MAXIN; k++) gW0[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < SUR_H; k++) gW1[k] = 0.0f;
  float gb0 = 0.f, gb1 = 0.f, gw2 = 0.f, gb2 = 0.f, loss = 0.f;
  const float inv = 2.0f / (float)B;
  const int b_lo = blockIdx.x * chunk, b_hi = min(B, b_lo + chunk);
  for (int b = b_lo; b < b_hi; b++) {
    const int r = idx ? idx[b] : b;
    float h1t, h2t;
    __syncthreads();
    const float y = forward(L, p, nin, lane, X + (size_t)r * nin, mu, sig, h1t, h2t);
    const float t = (Y[r] - ymu) / ysig;
    const float diff = y - t;
    loss += diff * diff;
    const float e = inv * diff;
    gb2 += e;
    gw2 += e * h2t;
    const float d2 = e * w2[lane] * (1.0f - h2t * h2t);
    L.d2[lane] = d2;
    gb1 += d2;
#pragma unroll
    for (int k = 0; k < SUR_H; k++) gW1[k] += d2 * L.h1[k];
    __syncthreads();
    float s1 = 0.0f;
#pragma unroll 8
    for (int u = 0; u < SUR_H; u++) s1 += L.W1[u * SUR_S1 + lane] * L.d2[u];
    const float d1 = s1 * (1.0f - h1t * h1t);
    gb0 += d1;
#pragma unroll
    for (int k = 0; k < SUR_MAXIN; k++)
      if (k < nin) gW0[k] += d1 * L.x[k];
  }
  float* o = part + (size_t)blockIdx.x * P;
#pragma unroll
  for (int k = 0; k < SUR_MAXIN; k++)
    if (k < nin) o[lane * nin + k] = gW0[k];
  o[SUR_H * nin + lane] = gb0;
  float* o1 = o + SUR_H * nin + SUR_H;
#pragma unroll
  for (int k = 0; k < SUR_H; k++) o1[lane * SUR_H + k] = gW1[k];
  o1[SUR_H * SUR_H + lane] = gb1;
  o1[SUR_H * SUR_H + SUR_H + lane] = gw2;
  // e, and so gb2 and the loss, are the same on every lane
  if (lane == 0) {
    o1[SUR_H * SUR_H + 2 * SUR_H] = gb2;
    part_loss[blockIdx.x] = loss;
  }
}

// sum of the block partials in block order; Adam step when lr > 0 (state m, v; step t >= 1), else the summed
// gradient goes to grad
__global__ __launch_bounds__(256) void k_sur_adam(int P, int nblk, const float* __restrict__ part,
                                                  float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                  float* __restrict__ grad, float lr, float b1, float b2, float eps,
                                                  float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; b++) g += part[(size_t)b * P + k];
  if (grad) grad[k] = g;
  if (lr > 0.0f) {
    const float mk = b1 * m[k] + (1.0f - b1) * g;
    const float vk = b2 * v[k] + (1.0f - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    // c1 = 1 - b1^t, c2 = 1 - b2^t (bias corrections)
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  }
}

__global__ __launch_bounds__(64) void k_sur_predict(int n, int nin, const float* __restrict__ X,
                                                    const float* __restrict__ p, const float* __restrict__ norm,
                                                    float* __restrict__ out) {
  __shared__ SurLds L;
  const int lane = threadIdx.x;
  load_weights(L, p, nin, lane);
  const float ymu = norm[2 * nin], ysig = norm[2 * nin + 1];
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    float h1t, h2t;
    __syncthreads();
    const float y = forward(L, p, nin, lane, X + (size_t)r * nin, norm, norm + nin, h1t, h2t);
    if (lane == 0) out[r] = ymu + ysig * y;
  }
}

extern "C" {

int mgk_sur_features(int S, const mg_struct_dev* str, const int* spc, const float* G, int D, int nspc, int ng2,
                     int ng4, int off2, int off4, float* feat, hipStream_t st) {
  if (S == 0) return 0;
  hipLaunchKernelGGL(k_sur_features, dim3(S), dim3(128), 0, st, S, str, spc, G, D, nspc, ng2, ng4, off2, off4, feat);
  return mgk_status();
}

int mgk_sur_grad(int B, int chunk, int nin, const int* idx, const float* X, const float* Y, const float* p,
                 const float* norm, float* part, float* part_loss, hipStream_t st) {
  if (nin < 1 || nin > SUR_MAXIN || chunk < 1 || B < 1) return -1;
  hipLaunchKernelGGL(k_sur_grad, dim3((B + chunk - 1) / chunk), dim3(64), 0, st, B, chunk, nin, idx, X, Y, p, norm,
                     part, part_loss);
  return mgk_status();
}

int mgk_sur_adam(int P, int nblk, const float* part, float* p, float* m, float* v, float* grad, float lr, float b1,
                 float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_sur_adam, dim3((P + 255) / 256), dim3(256), 0, st, P, nblk, part, p, m, v, grad, lr, b1, b2,
                     eps, c1, c2);
  return mgk_status();
}

int mgk_sur_predict(int n, int nin, const float* X, const float* p, const float* norm, float* out, hipStream_t st) {
  if (n == 0) return 0;
  if (nin < 1 || nin > SUR_MAXIN) return -1;
  const int grid = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(k_sur_predict, dim3(grid), dim3(64), 0, st, n, nin, X, p, norm, out);
  return mgk_status();
}

}  // extern "C"
