// mg_mlp.hip -- the MAISE network (D-H1-H2-1, tanh) on gfx950 matrix cores.
//
// Built without the VGPR-form MFMA option of mg_kernels.hip: these kernels keep their accumulators in AGPRs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_internal.h"
#include "mg_dev.h"

extern "C" int mgk_status(void);

// ---------------------------------------------------------------------------------------------- MLP on MFMA
// One wave per tile of up to 32 atoms of one species (atoms reached through a species-sorted permutation).
// Layers are GEMMs with hidden units as rows and atoms as columns: out[u][atom] = W[u][k] x[k][atom], on
// v_mfma_f32_32x32x2_f32 (FP32) or v_mfma_f32_32x32x16_bf16 (BF16 inputs, FP32 accumulation).
// Accumulator layout (32x32 blocks): lane l, register r holds row 8(r/4) + 4(l/32) + r%4, column l%32.  Each
// lane therefore keeps all hidden units of one atom spread over its registers, and those registers are fed
// straight back as the next layer's B operand; the host packs the weight fragments (A operand) in the same
// k order, so no data movement is needed between layers.  Forward tanh layers, E_i = W2.h2 + b2, and the
// backward chain dE/dG = W0'^T[(W1^T[W2 (1-h2^2)]) (1-h1^2)] (src/nmlp.c:367-431 and the chain of 433-478).
__device__ __forceinline__ int acc_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// PREC 0: FP32 MFMA (K=2).  PREC 1: BF16 MFMA (K=16), one pass.  PREC 2: compensated BF16 with split operands
// a = a_hi + a_lo, b = b_hi + b_lo (each a BF16 value) and three passes a_lo b_hi + a_hi b_lo + a_hi b_hi into the
// FP32 accumulator: single-pass BF16 rounds the weights to 2^-9 relative, which alone moves energies by ~1e-3
// eV/atom on the 51-10-10-1 workload; the split keeps the error near FP32 (~1e-6 eV/atom).
template <int PREC>
__device__ __forceinline__ f32x16 mfma_step(const float* __restrict__ afrag, const __bf16* __restrict__ afrag_bf,
                                            const __bf16* __restrict__ afrag_lo, int idx, const float* bvals,
                                            f32x16 acc) {
  if (PREC == 0) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(afrag[idx], bvals[0], acc, 0, 0, 0);
  } else {
    const bf16x8 a = *(const bf16x8*)(afrag_bf + 8 * (size_t)idx);
    bf16x8 b;
#pragma unroll
    for (int v = 0; v < 8; v++) b[v] = (__bf16)bvals[v];
    if (PREC == 2) {
      const bf16x8 al = *(const bf16x8*)(afrag_lo + 8 * (size_t)idx);
      bf16x8 bl;
#pragma unroll
      for (int v = 0; v < 8; v++) bl[v] = (__bf16)(bvals[v] - (float)b[v]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl, acc, 0, 0, 0);
    }
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
}

// Fragment offsets (in A-operand slots of 64 lanes; BF16 slots hold 8 values) of one species, see mg_model.c.  The
// input layer (W0' x and W0'^T delta1) and the hidden layer (W1 h1 and W1^T delta2) may use different sets: FP32
// (one value per slot, K = 2 per step) or BF16 (eight values, K = 16).
struct MlpFrag {
  int D, H1, H2, hb1, hb2, db, k1;   // k1: layer-1 steps of the input layer's set
  int o_l1, o_g0;                     // slot offsets of the input-layer A operands (input layer's set)
  int o_l2, o_b1;                     // slot offsets of the hidden-layer A operands (hidden layer's set)
  int sf, sb;                         // per-species strides of the FP32 and BF16 sets (slots)
  int stride_vec;                     // per-species stride of b0[hb1*32], b1[hb2*32], w2[hb2*32], b2, xoff[D]
};

// PIN: precision of the input layer, PHID: of the hidden layer (0 FP32, 1 one-pass BF16, 2 compensated BF16)
template <int PIN, int PHID, int HB1, int HB2>
__global__ __launch_bounds__(64) void k_mlp_mfma(const int4* __restrict__ tiles, const int* __restrict__ perm,
                                                 const float* __restrict__ G, const float* __restrict__ frag,
                                                 const __bf16* __restrict__ frag_bf, const __bf16* __restrict__ frag_lo,
                                                 const float* __restrict__ vecs, MlpFrag f, float* __restrict__ Ei,
                                                 float* __restrict__ dEdG) {
  constexpr bool BIN = PIN != 0, BHID = PHID != 0;
  constexpr int KS = BIN ? 8 : 1;     // B values per lane per step, input layer
  constexpr int KSH = BHID ? 8 : 1;   // hidden layer
  const int4 tl = tiles[blockIdx.x];
  const int sp = tl.x, start = tl.y, count = tl.z;
  const int lane = threadIdx.x, col = lane & 31, h = lane >> 5;
  const int atom = col < count ? perm[start + col] : -1;
  extern __shared__ float gsm[];  // [32][Kmax + 1] staged inputs
  const int D = f.D;
  const float* fa = frag + (size_t)sp * f.sf * 64;
  const __bf16* fb = frag_bf + (size_t)sp * f.sb * 64 * 8;
  const __bf16* fl = frag_lo + (size_t)sp * f.sb * 64 * 8;
  const float* vv = vecs + (size_t)sp * f.stride_vec;
  const float* b0 = vv;
  const float* b1 = b0 + HB1 * 32;
  const float* w2 = b1 + HB2 * 32;
  const float* xo = w2 + HB2 * 32 + 1;
  const float b2 = w2[HB2 * 32];

  // ---- layer 1: a1[u][atom] = b0 + W0' x
  f32x16 a1[HB1];
#pragma unroll
  for (int rb = 0; rb < HB1; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) a1[rb][r] = b0[rb * 32 + acc_row(r, h)];
  // the tile's inputs x - xoff staged in LDS first ([32][Kmax + 1], zero past D and for missing atoms): 32 rows of
  // coalesced loads all in flight at once, instead of two dependent memory round trips per K step
  const int k1 = f.k1;
  const int Kmax = BIN ? 16 * k1 : 2 * k1;
  const int RS = Kmax + 1;
  for (int k0 = 0; k0 < Kmax; k0 += WAVE) {
    const int k = k0 + lane, kc = k < D ? k : D - 1;
    const float xk = xo[kc];
    float vals[32];
#pragma unroll
    for (int c = 0; c < 32; c++) {
      const int ac = __builtin_amdgcn_readlane(atom, c);  // tile atom c, scalar
      const float keep = (ac >= 0 && k < D) ? 1.0f : 0.0f;  // a product, not a select: the loads stay unconditional
      vals[c] = (G[(size_t)(ac >= 0 ? ac : 0) * D + kc] - xk) * keep;
    }
    if (k < Kmax)
#pragma unroll
      for (int c = 0; c < 32; c++) gsm[c * RS + k] = vals[c];
  }
  wave_sync();
  const float* gr = gsm + col * RS;
  // four steps at a time (steps past k1 multiply the zero padding): their weight loads are in flight together
  for (int st0 = 0; st0 < k1; st0 += 4) {
    float xs[4][KS];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int su = st0 + u < k1 ? st0 + u : k1 - 1;
#pragma unroll
      for (int v = 0; v < KS; v++) xs[u][v] = st0 + u < k1 ? gr[BIN ? 16 * su + 8 * h + v : 2 * su + h] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int st = st0 + u < k1 ? st0 + u : k1 - 1;
#pragma unroll
      for (int rb = 0; rb < HB1; rb++)
        a1[rb] = mfma_step<PIN>(fa, fb, fl, f.o_l1 * 64 + (rb * k1 + st) * 64 + lane, xs[u], a1[rb]);
    }
  }
#pragma unroll
  for (int rb = 0; rb < HB1; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) a1[rb][r] = tanh_f(a1[rb][r]);

  // ---- layer 2: a2 = tanh(b1 + W1 h1); energy; delta2 = W2 (1 - h2^2)
  constexpr int S2 = BHID ? 2 : 16;  // hidden-layer steps per 32-unit block
  f32x16 a2[HB2];
#pragma unroll
  for (int rb = 0; rb < HB2; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) a2[rb][r] = b1[rb * 32 + acc_row(r, h)];
#pragma unroll
  for (int rb1 = 0; rb1 < HB1; rb1++)
#pragma unroll
    for (int q = 0; q < S2; q++) {
      float xb[KSH];
#pragma unroll
      for (int v = 0; v < KSH; v++) xb[v] = a1[rb1][BHID ? 8 * q + v : q];
      const int st = rb1 * S2 + q;
#pragma unroll
      for (int rb = 0; rb < HB2; rb++)
        a2[rb] = mfma_step<PHID>(fa, fb, fl, f.o_l2 * 64 + (rb * HB1 * S2 + st) * 64 + lane, xb, a2[rb]);
    }
  float e = 0.0f;
#pragma unroll
  for (int rb = 0; rb < HB2; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float w = w2[rb * 32 + acc_row(r, h)];
      const float t = tanh_f(a2[rb][r]);
      e += w * t;
      a2[rb][r] = w * (1.0f - t * t);  // delta2
    }
  e += __shfl_xor(e, 32, WAVE);
  if (atom >= 0 && h == 0) Ei[atom] = e + b2;

  // ---- delta1 = (W1^T delta2) (1 - h1^2)
  f32x16 d1[HB1];
#pragma unroll
  for (int rb = 0; rb < HB1; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) d1[rb][r] = 0.0f;
#pragma unroll
  for (int rb2 = 0; rb2 < HB2; rb2++)
#pragma unroll
    for (int q = 0; q < S2; q++) {
      float xb[KSH];
#pragma unroll
      for (int v = 0; v < KSH; v++) xb[v] = a2[rb2][BHID ? 8 * q + v : q];
      const int st = rb2 * S2 + q;
#pragma unroll
      for (int rb = 0; rb < HB1; rb++)
        d1[rb] = mfma_step<PHID>(fa, fb, fl, f.o_b1 * 64 + (rb * HB2 * S2 + st) * 64 + lane, xb, d1[rb]);
    }
#pragma unroll
  for (int rb = 0; rb < HB1; rb++)
#pragma unroll
    for (int r = 0; r < 16; r++) d1[rb][r] *= 1.0f - a1[rb][r] * a1[rb][r];

  // ---- dE/dG = W0'^T delta1, 32 features per block (input layer's precision)
  constexpr int SG = BIN ? 2 : 16;
  float* out = atom >= 0 ? dEdG + (size_t)atom * D : nullptr;
  for (int rbd = 0; rbd < f.db; rbd++) {
    f32x16 gacc;
#pragma unroll
    for (int r = 0; r < 16; r++) gacc[r] = 0.0f;
#pragma unroll
    for (int rb1 = 0; rb1 < HB1; rb1++)
#pragma unroll
      for (int q = 0; q < SG; q++) {
        float xb[KS];
#pragma unroll
        for (int v = 0; v < KS; v++) xb[v] = d1[rb1][BIN ? 8 * q + v : q];
        gacc = mfma_step<PIN>(fa, fb, fl, f.o_g0 * 64 + (rbd * HB1 * SG + rb1 * SG + q) * 64 + lane, xb, gacc);
      }
    if (out)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int k = rbd * 32 + acc_row(r, h);
        if (k < D) out[k] = gacc[r];
      }
  }
}

// The A operand of one MFMA step (one FP32 value, or eight BF16 values and their low halves), loaded ahead of the
// step that uses it, and the step itself (mfma_step with the load taken out).
template <int PREC> struct AFrag { float a; };
template <> struct AFrag<1> { bf16x8 a; };
template <> struct AFrag<2> { bf16x8 a, lo; };

template <int PREC>
__device__ __forceinline__ AFrag<PREC> afrag_load(const float* __restrict__ afrag, const __bf16* __restrict__ afrag_bf,
                                                  const __bf16* __restrict__ afrag_lo, int idx) {
  AFrag<PREC> r;
  if constexpr (PREC == 0) {
    r.a = afrag[idx];
  } else {
    r.a = *(const bf16x8*)(afrag_bf + 8 * (size_t)idx);
    if constexpr (PREC == 2) r.lo = *(const bf16x8*)(afrag_lo + 8 * (size_t)idx);
  }
  return r;
}

template <int PREC>
__device__ __forceinline__ f32x16 mfma_frag(const AFrag<PREC>& A, const float* bvals, f32x16 acc) {
  if constexpr (PREC == 0) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(A.a, bvals[0], acc, 0, 0, 0);
  } else {
    bf16x8 b;
#pragma unroll
    for (int v = 0; v < 8; v++) b[v] = (__bf16)bvals[v];
    if constexpr (PREC == 2) {
      bf16x8 bl;
#pragma unroll
      for (int v = 0; v < 8; v++) bl[v] = (__bf16)(bvals[v] - (float)b[v]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.lo, b, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a, bl, acc, 0, 0, 0);
    }
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a, b, acc, 0, 0, 0);
  }
}

// floats of the two-wave kernel's exchange regions: layer 1's (always), delta2's and the energy partials (a two-block
// second layer), delta1's (a two-block first layer).  Only what the layout uses is allocated, so a one-block network
// on a wide input keeps more tiles resident per CU.
static inline int mlp2_exchange_floats(int hb1, int hb2) {
  return 2 * 16 * 64 + (hb2 == 2 ? 2 * 16 * 64 + 2 * 64 : 0) + (hb1 == 2 ? 2 * 16 * 64 : 0);
}

// ---- two waves per tile.  The single-wave kernel above leaves one wave per SIMD on most of the chip (a tile is 32
// atoms), so each wave's load latencies and its chain of matrix-core and tanh work run alone.  Here the two waves
// of a workgroup share a tile: wave w owns hidden block w of a 64-wide layer (HB == 2), half of the input layer's
// K steps when that layer is one block wide (HB1 == 1; the two partial sums are added in the same order by both
// waves), and every second output block of dE/dG.  Activations, deltas and partial sums cross between the waves
// through LDS (regions of [2 waves][16 registers][64 lanes] floats, one per exchange, so no region is rewritten
// while the other wave may still read it).  Same fragments, same arithmetic per block as the single-wave kernel.
template <int PIN, int PHID, int HB1, int HB2>
__global__ __launch_bounds__(128) void k_mlp_mfma2(const int4* __restrict__ tiles, const int* __restrict__ perm,
                                                   const float* __restrict__ G, const float* __restrict__ frag,
                                                   const __bf16* __restrict__ frag_bf, const __bf16* __restrict__ frag_lo,
                                                   const float* __restrict__ vecs, MlpFrag f, float* __restrict__ Ei,
                                                   float* __restrict__ dEdG) {
  constexpr bool BIN = PIN != 0, BHID = PHID != 0;
  constexpr int KS = BIN ? 8 : 1;
  constexpr int KSH = BHID ? 8 : 1;
  constexpr int S2 = BHID ? 2 : 16;  // hidden-layer steps per 32-unit block
  constexpr int SG = BIN ? 2 : 16;   // dE/dG steps per 32-unit block of d1
  const int4 tl = tiles[blockIdx.x];
  const int sp = tl.x, start = tl.y, count = tl.z;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  const int atom = col < count ? perm[start + col] : -1;
  extern __shared__ float gsm[];
  const int D = f.D;
  const float* fa = frag + (size_t)sp * f.sf * 64;
  const __bf16* fb = frag_bf + (size_t)sp * f.sb * 64 * 8;
  const __bf16* fl = frag_lo + (size_t)sp * f.sb * 64 * 8;
  const float* vv = vecs + (size_t)sp * f.stride_vec;
  const float* b0 = vv;
  const float* b1 = b0 + HB1 * 32;
  const float* w2 = b1 + HB2 * 32;
  const float* xo = w2 + HB2 * 32 + 1;
  const float b2 = w2[HB2 * 32];
  const int k1 = f.k1;
  const int Kmax = BIN ? 16 * k1 : 2 * k1;
  const int RS = Kmax + 1;
  // exchange regions after the staged inputs, only those the block layout uses (mlp2_exchange_floats)
  float* xa = gsm + 32 * RS;                             // h1 (HB1 == 2) or the layer-1 partial sums (HB1 == 1)
  float* xb2 = xa + 2 * 16 * 64;                         // delta2 (HB2 == 2)
  float* eb = xb2 + 2 * 16 * 64;                         // the energy partials (HB2 == 2)
  float* xc = HB2 == 2 ? eb + 2 * 64 : xa + 2 * 16 * 64;  // d1 (HB1 == 2)

  // the tile's inputs x - xoff in LDS, wave w staging tile atoms 16 w .. 16 w + 15
  for (int k0 = 0; k0 < Kmax; k0 += WAVE) {
    const int k = k0 + lane, kc = k < D ? k : D - 1;
    const float xk = xo[kc];
    float vals[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const int ac = __builtin_amdgcn_readlane(atom, 16 * w + c);
      const float keep = (ac >= 0 && k < D) ? 1.0f : 0.0f;
      vals[c] = (G[(size_t)(ac >= 0 ? ac : 0) * D + kc] - xk) * keep;
    }
    if (k < Kmax)
#pragma unroll
      for (int c = 0; c < 16; c++) gsm[(16 * w + c) * RS + k] = vals[c];
  }
  __syncthreads();
  const float* gr = gsm + col * RS;

  // ---- layer 1: a1 = tanh(b0 + W0' x), HB1 blocks, all held by both waves afterwards.  The A operands (weight
  // fragments) are loaded a group of U steps ahead of the steps that use them, and each later layer's fragments
  // are loaded before the exchange that precedes it, so their latency passes under the barrier.
  constexpr int U = BIN ? 4 : 8;
  constexpr int N2 = HB1 * S2, N3 = HB2 * S2, N4 = HB1 * SG;  // steps of layer 2, delta1 and a dE/dG block
  const int rb1w = HB1 == 2 ? w : 0, rb2w = HB2 == 2 ? w : 0;  // the blocks this wave computes
  f32x16 a1[HB1];
  AFrag<PHID> f2[N2];
  {
    f32x16 acc;
    const bool bias = HB1 == 2 || w == 0;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = bias ? b0[rb1w * 32 + acc_row(r, h)] : 0.0f;
    const int half = (k1 + 1) >> 1;
    const int sa = HB1 == 2 ? 0 : w * half;
    const int se = HB1 == 2 ? k1 : (w == 0 ? half : k1);
    const int i0 = f.o_l1 * 64 + rb1w * k1 * 64 + lane;
    AFrag<PIN> fc[U];
#pragma unroll
    for (int u = 0; u < U; u++) fc[u] = afrag_load<PIN>(fa, fb, fl, i0 + max(min(sa + u, se - 1), 0) * 64);
    for (int st0 = sa; st0 < se; st0 += U) {
      AFrag<PIN> fn[U];  // the next group (past the end: the last step again, unused)
#pragma unroll
      for (int u = 0; u < U; u++) fn[u] = afrag_load<PIN>(fa, fb, fl, i0 + min(st0 + U + u, se - 1) * 64);
      float xs[U][KS];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int su = st0 + u < se ? st0 + u : se - 1;
#pragma unroll
        for (int v = 0; v < KS; v++) xs[u][v] = st0 + u < se ? gr[BIN ? 16 * su + 8 * h + v : 2 * su + h] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; u++) acc = mfma_frag<PIN>(fc[u], xs[u], acc);
#pragma unroll
      for (int u = 0; u < U; u++) fc[u] = fn[u];
    }
#pragma unroll
    for (int s = 0; s < N2; s++) f2[s] = afrag_load<PHID>(fa, fb, fl, f.o_l2 * 64 + (rb2w * N2 + s) * 64 + lane);
    if (HB1 == 2) {
#pragma unroll
      for (int r = 0; r < 16; r++) xa[(w * 16 + r) * 64 + lane] = tanh_f(acc[r]);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < HB1; q++)
#pragma unroll
        for (int r = 0; r < 16; r++) a1[q][r] = xa[(q * 16 + r) * 64 + lane];
    } else {
#pragma unroll
      for (int r = 0; r < 16; r++) xa[(w * 16 + r) * 64 + lane] = acc[r];
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 16; r++) a1[0][r] = tanh_f(xa[r * 64 + lane] + xa[(16 + r) * 64 + lane]);
    }
  }

  // ---- layer 2: a2 = tanh(b1 + W1 h1); energy; delta2 = W2 (1 - h2^2): wave w's block when HB2 == 2, else both
  // waves compute the single block
  f32x16 a2[HB2];
  AFrag<PHID> f3[N3];
  {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = b1[rb2w * 32 + acc_row(r, h)];
#pragma unroll
    for (int rb1 = 0; rb1 < HB1; rb1++)
#pragma unroll
      for (int q = 0; q < S2; q++) {
        float xb[KSH];
#pragma unroll
        for (int v = 0; v < KSH; v++) xb[v] = a1[rb1][BHID ? 8 * q + v : q];
        acc = mfma_frag<PHID>(f2[rb1 * S2 + q], xb, acc);
      }
#pragma unroll
    for (int s = 0; s < N3; s++) f3[s] = afrag_load<PHID>(fa, fb, fl, f.o_b1 * 64 + (rb1w * N3 + s) * 64 + lane);
    float e = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float wr = w2[rb2w * 32 + acc_row(r, h)];
      const float t = tanh_f(acc[r]);
      e += wr * t;
      acc[r] = wr * (1.0f - t * t);  // delta2
    }
    e += __shfl_xor(e, 32, WAVE);
    if (HB2 == 2) {
#pragma unroll
      for (int r = 0; r < 16; r++) xb2[(w * 16 + r) * 64 + lane] = acc[r];
      eb[w * 64 + lane] = e;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < HB2; q++)
#pragma unroll
        for (int r = 0; r < 16; r++) a2[q][r] = xb2[(q * 16 + r) * 64 + lane];
      e = eb[lane] + eb[64 + lane];
    } else {
      a2[0] = acc;
    }
    if (w == 0 && atom >= 0 && h == 0) Ei[atom] = e + b2;
  }

  // ---- delta1 = (W1^T delta2) (1 - h1^2): wave w's block when HB1 == 2, else both waves
  f32x16 d1[HB1];
  AFrag<PIN> f4[N4];
  {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
    for (int rb2 = 0; rb2 < HB2; rb2++)
#pragma unroll
      for (int q = 0; q < S2; q++) {
        float xb[KSH];
#pragma unroll
        for (int v = 0; v < KSH; v++) xb[v] = a2[rb2][BHID ? 8 * q + v : q];
        acc = mfma_frag<PHID>(f3[rb2 * S2 + q], xb, acc);
      }
    // the first dE/dG block's fragments (block w; a wave without a block loads block 0, unused)
    const int rbd0 = w < f.db ? w : 0;
#pragma unroll
    for (int s = 0; s < N4; s++) f4[s] = afrag_load<PIN>(fa, fb, fl, f.o_g0 * 64 + (rbd0 * N4 + s) * 64 + lane);
    if (HB1 == 2) {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float hv = w == 0 ? a1[0][r] : a1[HB1 - 1][r];
        xc[(w * 16 + r) * 64 + lane] = acc[r] * (1.0f - hv * hv);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < HB1; q++)
#pragma unroll
        for (int r = 0; r < 16; r++) d1[q][r] = xc[(q * 16 + r) * 64 + lane];
    } else {
#pragma unroll
      for (int r = 0; r < 16; r++) d1[0][r] = acc[r] * (1.0f - a1[0][r] * a1[0][r]);
    }
  }

  // ---- dE/dG = W0'^T delta1, 32 features per block, blocks alternating between the waves (the next block's
  // fragments loaded before this block's steps)
  float* out = atom >= 0 ? dEdG + (size_t)atom * D : nullptr;
  for (int rbd = w; rbd < f.db; rbd += 2) {
    AFrag<PIN> fn[N4];
    const int rn = rbd + 2 < f.db ? rbd + 2 : rbd;
#pragma unroll
    for (int s = 0; s < N4; s++) fn[s] = afrag_load<PIN>(fa, fb, fl, f.o_g0 * 64 + (rn * N4 + s) * 64 + lane);
    f32x16 gacc;
#pragma unroll
    for (int r = 0; r < 16; r++) gacc[r] = 0.0f;
#pragma unroll
    for (int rb1 = 0; rb1 < HB1; rb1++)
#pragma unroll
      for (int q = 0; q < SG; q++) {
        float xb[KS];
#pragma unroll
        for (int v = 0; v < KS; v++) xb[v] = d1[rb1][BIN ? 8 * q + v : q];
        gacc = mfma_frag<PIN>(f4[rb1 * SG + q], xb, gacc);
      }
    if (out)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int k = rbd * 32 + acc_row(r, h);
        if (k < D) out[k] = gacc[r];
      }
#pragma unroll
    for (int s = 0; s < N4; s++) f4[s] = fn[s];
  }
}

// prec: 0 FP32 throughout; 1 (MG_PREC_BF16) one-pass BF16 in the hidden layer (W1 h1 and W1^T delta2), FP32 input
// layer; 2 (MG_PREC_BF16X3) compensated BF16 in every layer
extern "C" int mgk_mlp_mfma(int ntiles, const void* tiles, const int* perm, const float* G, const float* frag,
                            const void* frag_bf, const void* frag_lo, const float* vecs, const int* fdesc, int prec,
                            float* Ei, float* dEdG, hipStream_t st) {
  if (ntiles == 0) return 0;
  const bool bin = prec == 2, bhid = prec != 0;
  MlpFrag f;
  f.D = fdesc[0]; f.H1 = fdesc[1]; f.H2 = fdesc[2]; f.hb1 = fdesc[3]; f.hb2 = fdesc[4]; f.db = fdesc[5];
  f.k1 = bin ? fdesc[6] : fdesc[7];
  const int* oi = bin ? fdesc + 8 : fdesc + 14;   // input layer's set
  const int* oh = bhid ? fdesc + 8 : fdesc + 14;  // hidden layer's set
  f.o_l1 = oi[0]; f.o_g0 = oi[3];
  f.o_l2 = oh[1]; f.o_b1 = oh[2];
  f.sb = fdesc[12]; f.sf = fdesc[18];
  f.stride_vec = fdesc[13];
  const int4* t = (const int4*)tiles;
  const __bf16* fb = (const __bf16*)frag_bf;
  const __bf16* fl = (const __bf16*)frag_lo;
  const size_t sm = (size_t)32 * ((bin ? 16 : 2) * f.k1 + 1) * sizeof(float);  // staged inputs of a tile
  if (sm > 64 * 1024) return -2;
  // two waves per tile where the exchange regions fit beside the staged inputs (every shipped width does)
  const size_t sm2 = sm + (size_t)mlp2_exchange_floats(f.hb1, f.hb2) * sizeof(float);
  const bool two = sm2 <= 64 * 1024;
#define MLP_L(PI, PH, H1B, H2B)                                                                                      \
  do {                                                                                                               \
    if (two)                                                                                                         \
      hipLaunchKernelGGL((k_mlp_mfma2<PI, PH, H1B, H2B>), dim3(ntiles), dim3(128), sm2, st, t, perm, G, frag, fb, fl, \
                         vecs, f, Ei, dEdG);                                                                         \
    else                                                                                                             \
      hipLaunchKernelGGL((k_mlp_mfma<PI, PH, H1B, H2B>), dim3(ntiles), dim3(64), sm, st, t, perm, G, frag, fb, fl,   \
                         vecs, f, Ei, dEdG);                                                                         \
  } while (0)
#define MLP_P(PI, PH)                                     \
  do {                                                    \
    if (f.hb1 == 1 && f.hb2 == 1) MLP_L(PI, PH, 1, 1);    \
    else if (f.hb1 == 1) MLP_L(PI, PH, 1, 2);             \
    else if (f.hb2 == 1) MLP_L(PI, PH, 2, 1);             \
    else MLP_L(PI, PH, 2, 2);                             \
  } while (0)
  if (prec == 2) MLP_P(2, 2);
  else if (prec == 1) MLP_P(0, 1);
  else MLP_P(0, 0);
#undef MLP_P
#undef MLP_L
  return mgk_status();
}
