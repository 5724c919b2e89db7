// mg_sur.hip -- the learned screening surrogate: a nin-64-64-1 tanh network over per-structure descriptor
// features, trained with mean-squared error and Adam, entirely in these kernels.
//
// Features (k_sur_features): for structure s and species q, the mean over the atoms of species q of their whole
// descriptor row (all D functions), giving nspc * D inputs; zero for a species absent from s.
//
// Network (flat FP32 parameter vector, the order of the reference's per-layer weight blocks, src/nmlp.c:591-608):
//   W0[64][nin] b0[64] W1[64][64] b1[64] w2[64] b2
//   x' = (x - mu) / sig,  h1 = tanh(W0 x' + b0),  h2 = tanh(W1 h1 + b1),  y' = w2 . h2 + b2,  y = ymu + ysig y'
// Loss over a mini-batch of B samples: L = (1/B) sum_b (y'_b - t'_b)^2 with t' = (t - ymu) / ysig.
// Gradients (the chain of the reference's training derivative DE_ANN_PARA, src/nmlp.c:292-354, for this loss):
//   e = 2 (y' - t') / B;  g_b2 = e;  g_w2 = e h2;  d2 = e w2 (1 - h2^2);  g_W1 = d2 h1^T;  g_b1 = d2;
//   d1 = (W1^T d2)(1 - h1^2);  g_W0 = d1 x'^T;  g_b0 = d1.
//
// A block is SUR_W waves, one sample per wave, lane t owning hidden unit t of both layers.  W0 streams through
// LDS in tiles of SUR_T columns shared by the block's samples, W1 sits in LDS whole (odd row strides keep the
// row-per-lane and column-per-lane sweeps free of bank conflicts).  After the backward sweep the block writes
// the sum of its samples' gradients as outer products from LDS (coalesced over the input index), adding into its
// own partial on later rounds; the reduction over blocks (fixed order, so results do not depend on timing) is
// fused with the Adam update.
#include <hip/hip_runtime.h>

#include "mg_internal.h"

#define SUR_H 64
#define SUR_MAXIN 1024
#define SUR_W 4                 // samples (waves) per block
#define SUR_T 128               // W0 columns per LDS tile
#define SUR_ST (SUR_T + 1)      // LDS row stride of a W0 tile
#define SUR_S1 (SUR_H + 1)      // LDS row stride of W1
#define SUR_THREADS (64 * SUR_W)

namespace {

__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct SurLds {
  float W0t[SUR_H * SUR_ST];
  float W1[SUR_H * SUR_S1];
  float x[SUR_W][SUR_MAXIN];
  float h1[SUR_W][SUR_H];
  float d1[SUR_W][SUR_H];
  float d2[SUR_W][SUR_H];
  float h2e[SUR_W][SUR_H];
  float e[SUR_W];
  float loss[SUR_W];
};

// forward of the block's samples (wave w: sample row `row`, or none when row < 0, which computes on x' = 0);
// called by every thread of the block.  Leaves x' and h1 in LDS, h1/h2 of unit `lane` in registers and returns y'
// (wave-uniform).
__device__ float forward(SurLds& L, const float* __restrict__ p, int nin, const float* __restrict__ X, int row,
                         const float* __restrict__ mu, const float* __restrict__ sig, int w, int lane, int tid,
                         float& h1t, float& h2t) {
  for (int k = lane; k < nin; k += 64) L.x[w][k] = row >= 0 ? (X[(size_t)row * nin + k] - mu[k]) / sig[k] : 0.0f;
  const float* b0 = p + SUR_H * nin;
  const float* W1 = b0 + SUR_H;
  const float* b1 = W1 + SUR_H * SUR_H;
  const float* w2 = b1 + SUR_H;
  for (int e = tid; e < SUR_H * SUR_H; e += SUR_THREADS) L.W1[(e >> 6) * SUR_S1 + (e & 63)] = W1[e];
  float a = b0[lane];
  for (int t0 = 0; t0 < nin; t0 += SUR_T) {
    const int nt = min(SUR_T, nin - t0);
    __syncthreads();  // the previous tile is consumed (and, the first time, x' is written)
    for (int e = tid; e < SUR_H * SUR_T; e += SUR_THREADS) {
      const int u = e / SUR_T, j = e % SUR_T;
      if (j < nt) L.W0t[u * SUR_ST + j] = p[(size_t)u * nin + t0 + j];
    }
    __syncthreads();
    for (int j = 0; j < nt; j++) a += L.W0t[lane * SUR_ST + j] * L.x[w][t0 + j];
  }
  h1t = tanhf(a);
  L.h1[w][lane] = h1t;
  __syncthreads();  // W1 staged by the whole block, h1 written
  float c = b1[lane];
#pragma unroll 8
  for (int k = 0; k < SUR_H; k++) c += L.W1[lane * SUR_S1 + k] * L.h1[w][k];
  h2t = tanhf(c);
  return wsum(w2[lane] * h2t) + w2[SUR_H];
}

}  // namespace

// per-structure features: block per structure, threads over (species, function)
__global__ __launch_bounds__(256) void k_sur_features(int S, const mg_struct_dev* __restrict__ str,
                                                      const int* __restrict__ spc, const float* __restrict__ G, int D,
                                                      int nspc, float* __restrict__ feat) {
  const int s = blockIdx.x;
  if (s >= S) return;
  const int a0 = str[s].atom0, n = str[s].natom;
  for (int f = threadIdx.x; f < nspc * D; f += blockDim.x) {
    const int q = f / D, c = f - q * D;
    float sum = 0.0f;
    int cnt = 0;
    for (int i = a0; i < a0 + n; i++)
      if (spc[i] == q) { sum += G[(size_t)i * D + c]; cnt++; }
    feat[(size_t)s * nspc * D + f] = cnt ? sum / (float)cnt : 0.0f;
  }
}

// fused forward + backward over B samples (indices idx[0..B), or 0..B when idx is null): block b takes samples
// SUR_W at a time from b*SUR_W with stride gridDim.x*SUR_W; part[b][P] receives the sum of its samples'
// gradients, part_loss[b] their squared errors.  Every block has at least one round (gridDim.x*SUR_W < B+SUR_W).
__global__ __launch_bounds__(SUR_THREADS) void k_sur_grad(int B, int nin, const int* __restrict__ idx,
                                                          const float* __restrict__ X, const float* __restrict__ Y,
                                                          const float* __restrict__ p, const float* __restrict__ norm,
                                                          float* __restrict__ part, float* __restrict__ part_loss) {
  __shared__ SurLds L;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int P = SUR_H * nin + SUR_H * SUR_H + 3 * SUR_H + 1;  // W0 b0 W1 b1 w2 b2
  const float* mu = norm;
  const float* sig = norm + nin;
  const float ymu = norm[2 * nin], ysig = norm[2 * nin + 1];
  const float* w2 = p + SUR_H * nin + SUR_H + SUR_H * SUR_H + SUR_H;
  const float inv = 2.0f / (float)B;
  float* o = part + (size_t)blockIdx.x * P;
  float* o1 = o + SUR_H * nin + SUR_H;  // W1 block
  int round = 0;
  for (int r0 = blockIdx.x * SUR_W; r0 < B; r0 += gridDim.x * SUR_W, round++) {
    const int b = r0 + w;
    const int row = b < B ? (idx ? idx[b] : b) : -1;
    float h1t, h2t;
    __syncthreads();  // the previous round's LDS contents are consumed
    const float y = forward(L, p, nin, X, row, mu, sig, w, lane, tid, h1t, h2t);
    const float diff = row >= 0 ? y - (Y[row] - ymu) / ysig : 0.0f;
    const float e = inv * diff;
    const float d2 = e * w2[lane] * (1.0f - h2t * h2t);
    L.d2[w][lane] = d2;
    L.h2e[w][lane] = e * h2t;
    if (lane == 0) {
      L.e[w] = e;
      L.loss[w] = diff * diff;
    }
    __syncthreads();
    float s1 = 0.0f;
#pragma unroll 8
    for (int u = 0; u < SUR_H; u++) s1 += L.W1[u * SUR_S1 + lane] * L.d2[w][u];
    L.d1[w][lane] = s1 * (1.0f - h1t * h1t);
    __syncthreads();
    // the block's gradient, summed over its samples, as outer products
    const bool add = round > 0;
    for (int e2 = tid; e2 < SUR_H * nin; e2 += SUR_THREADS) {
      const int u = e2 / nin, k = e2 - u * nin;
      float g = 0.0f;
#pragma unroll
      for (int q = 0; q < SUR_W; q++) g += L.d1[q][u] * L.x[q][k];
      o[e2] = add ? o[e2] + g : g;
    }
    for (int e2 = tid; e2 < SUR_H * SUR_H; e2 += SUR_THREADS) {
      const int u = e2 >> 6, k = e2 & 63;
      float g = 0.0f;
#pragma unroll
      for (int q = 0; q < SUR_W; q++) g += L.d2[q][u] * L.h1[q][k];
      o1[e2] = add ? o1[e2] + g : g;
    }
    if (tid < SUR_H) {
      float gb0 = 0.0f, gb1 = 0.0f, gw2 = 0.0f;
#pragma unroll
      for (int q = 0; q < SUR_W; q++) {
        gb0 += L.d1[q][tid];
        gb1 += L.d2[q][tid];
        gw2 += L.h2e[q][tid];
      }
      o[SUR_H * nin + tid] = add ? o[SUR_H * nin + tid] + gb0 : gb0;
      o1[SUR_H * SUR_H + tid] = add ? o1[SUR_H * SUR_H + tid] + gb1 : gb1;
      o1[SUR_H * SUR_H + SUR_H + tid] = add ? o1[SUR_H * SUR_H + SUR_H + tid] + gw2 : gw2;
    }
    if (tid == 0) {
      float gb2 = 0.0f, ls = 0.0f;
#pragma unroll
      for (int q = 0; q < SUR_W; q++) {
        gb2 += L.e[q];
        ls += L.loss[q];
      }
      o1[SUR_H * SUR_H + 2 * SUR_H] = add ? o1[SUR_H * SUR_H + 2 * SUR_H] + gb2 : gb2;
      part_loss[blockIdx.x] = add ? part_loss[blockIdx.x] + ls : ls;
    }
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

__global__ __launch_bounds__(SUR_THREADS) void k_sur_predict(int n, int nin, const float* __restrict__ X,
                                                             const float* __restrict__ p,
                                                             const float* __restrict__ norm, float* __restrict__ out) {
  __shared__ SurLds L;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const float ymu = norm[2 * nin], ysig = norm[2 * nin + 1];
  for (int r0 = blockIdx.x * SUR_W; r0 < n; r0 += gridDim.x * SUR_W) {
    const int row = r0 + w < n ? r0 + w : -1;
    float h1t, h2t;
    __syncthreads();
    const float y = forward(L, p, nin, X, row, norm, norm + nin, w, lane, tid, h1t, h2t);
    if (lane == 0 && row >= 0) out[row] = ymu + ysig * y;
  }
}

extern "C" {

int mgk_sur_features(int S, const mg_struct_dev* str, const int* spc, const float* G, int D, int nspc, float* feat,
                     hipStream_t st) {
  if (S == 0) return 0;
  hipLaunchKernelGGL(k_sur_features, dim3(S), dim3(256), 0, st, S, str, spc, G, D, nspc, feat);
  return mgk_status();
}

// nblk blocks, each with at least one round of SUR_W samples: 1 <= nblk <= ceil(B / SUR_W)
int mgk_sur_grad(int B, int nblk, int nin, const int* idx, const float* X, const float* Y, const float* p,
                 const float* norm, float* part, float* part_loss, hipStream_t st) {
  if (nin < 1 || nin > SUR_MAXIN || B < 1 || nblk < 1 || nblk > (B + SUR_W - 1) / SUR_W) return -1;
  hipLaunchKernelGGL(k_sur_grad, dim3(nblk), dim3(SUR_THREADS), 0, st, B, nin, idx, X, Y, p, norm, part, part_loss);
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
  const int grid = (n + SUR_W - 1) / SUR_W < 512 ? (n + SUR_W - 1) / SUR_W : 512;
  hipLaunchKernelGGL(k_sur_predict, dim3(grid), dim3(SUR_THREADS), 0, st, n, nin, X, p, norm, out);
  return mgk_status();
}

}  // extern "C"
