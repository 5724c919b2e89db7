// mg_kernels.hip -- CDNA4 (gfx950) kernels of the MAISE neural-network potential engine.
//
// Pipeline for one evaluation (all structures of a context at once):
//   k_prepare      CAST_CELL fractional->Cartesian, JAR wrap in FP64 (src/cmod.c:1112-1121, src/cell.c:446-508),
//                  cell-list bin of every atom
//   k_bin_fill / k_bin_sort    counting sort of atoms into bins (deterministic order inside a bin)
//   k_nbr<false>   per-atom neighbour count per species over the bin stencil, FP64 test r < Rc^2 in the
//                  reference's expression order (src/cell.c:109-123) so the neighbour SETS are bit-exact
//   k_nbr<true>    fill: 64-bit key (species, j, image) + FP32 (d, r) per edge
//   k_row_sort     one wave per row: bitonic sort of the row by key (canonical, bin-independent order)
//   k_rev          reverse-edge index by binary search in the neighbour's row
//   k_desc<NE>     one wave per atom: G2 over neighbours, G4 over neighbour pairs of each species class,
//                  neighbour tile staged in LDS, (eta, zeta, lambda) grid accumulators in VGPRs,
//                  reduce-scatter across the wave (src/nprs.c:504-566)
//   k_mlp_fp32     per-atom D-H1-H2-1 forward and dE/dG backward (src/nmlp.c:367-431 + chain of 444-478)
//   k_force<NE>    one wave per atom: recomputes pair/triplet geometry, contracts dG/dr with dE/dG in
//                  registers, forces on neighbour slots via LDS atomics, per-atom virial (no Fn/Sn storage)
//   k_gather       F_a = own term + sum over reverse edges (deterministic)
//   k_reduce       per-structure FP64 energy and virial sums (deterministic tree)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "mg_internal.h"

#define WAVE 64

typedef unsigned long long u64;

__device__ __forceinline__ int floordiv_i(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

__device__ __forceinline__ unsigned pack_img(int a, int b, int c) {
  return (unsigned)(a + MG_IMG_OFF) | ((unsigned)(b + MG_IMG_OFF) << 10) | ((unsigned)(c + MG_IMG_OFF) << 20);
}
__device__ __forceinline__ unsigned neg_img(unsigned p) {
  int a = (int)(p & 1023u) - MG_IMG_OFF, b = (int)((p >> 10) & 1023u) - MG_IMG_OFF, c = (int)((p >> 20) & 1023u) - MG_IMG_OFF;
  return pack_img(-a, -b, -c);
}
__device__ __forceinline__ u64 make_key(int s, int j, unsigned img) {
  return ((u64)(unsigned)s << 60) | ((u64)(unsigned)j << 30) | (u64)img;
}
__device__ __forceinline__ int key_j(u64 k) { return (int)((k >> 30) & 0x3fffffffull); }

// ---------------------------------------------------------------------------------------------- prepare
__global__ void k_prepare(int N, const double* __restrict__ xin, const int* __restrict__ atom_str,
                          const mg_struct_dev* __restrict__ str, double* __restrict__ x, int* __restrict__ bin,
                          int* __restrict__ bin_count) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const mg_struct_dev* st = &str[atom_str[i]];
  double X[3] = {xin[3 * i], xin[3 * i + 1], xin[3 * i + 2]};
  if (st->xt == 0) {  // fractional input -> Cartesian, x_q = sum_q1 X_q1 L_q1q
    double c[3];
    for (int q = 0; q < 3; q++) {
      c[q] = 0.0;
      for (int q1 = 0; q1 < 3; q1++) c[q] += X[q1] * st->L[3 * q1 + q];
    }
    X[0] = c[0]; X[1] = c[1]; X[2] = c[2];
  }
  int b[3] = {0, 0, 0};
  if (st->nd > 0) {
    double f[3];
    for (int q = 0; q < 3; q++) {
      f[q] = 0.0;
      for (int j = 0; j < 3; j++) f[q] += X[j] * st->R[3 * q + j];
    }
    for (int q = 0; q < 3; q++) {
      // the reference wraps with repeated +-1.0 (rounding at each step); far-away atoms are first brought
      // close with floor so a corrupted position cannot stall the kernel
      if (f[q] < -1048576.0 || f[q] >= 1048576.0) f[q] -= floor(f[q]);
      while (f[q] < 0.0) f[q] += 1.0;
      while (f[q] >= 1.0) f[q] -= 1.0;
    }
    for (int q = 0; q < 3; q++) {
      X[q] = 0.0;
      for (int j = 0; j < 3; j++) X[q] += f[j] * st->L[3 * j + q];
    }
    for (int q = 0; q < 3; q++) {
      int v = (int)(f[q] * (double)st->nb[q]);
      b[q] = v < 0 ? 0 : (v >= st->nb[q] ? st->nb[q] - 1 : v);
    }
  }
  x[3 * i] = X[0]; x[3 * i + 1] = X[1]; x[3 * i + 2] = X[2];
  int id = st->bin0 + (b[0] * st->nb[1] + b[1]) * st->nb[2] + b[2];
  bin[i] = id;
  atomicAdd(&bin_count[id], 1);
}

__global__ void k_bin_fill(int N, const int* __restrict__ bin, const int* __restrict__ bin_start,
                           int* __restrict__ bin_cur, int* __restrict__ bin_atoms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b = bin[i];
  int p = bin_start[b] + atomicAdd(&bin_cur[b], 1);
  bin_atoms[p] = i;
}

__global__ void k_bin_sort(int nbins, const int* __restrict__ bin_start, int* __restrict__ bin_atoms) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int s = bin_start[b], e = bin_start[b + 1];
  for (int p = s + 1; p < e; p++) {
    int v = bin_atoms[p], q = p - 1;
    while (q >= s && bin_atoms[q] > v) { bin_atoms[q + 1] = bin_atoms[q]; q--; }
    bin_atoms[q + 1] = v;
  }
}

// ---------------------------------------------------------------------------------------------- neighbours
// One thread per atom.  Stencil over bins within K_q in each direction; the image of a candidate follows
// from the periodic wrap of the bin coordinate.  The FP64 distance is formed exactly as LIST does.
template <bool FILL>
__global__ void k_nbr(int N, const double* __restrict__ x, const int* __restrict__ spc, const int* __restrict__ atom_str,
                      const mg_struct_dev* __restrict__ str, const int* __restrict__ bin, const int* __restrict__ bin_start,
                      const int* __restrict__ bin_atoms, int* __restrict__ cnt, int* __restrict__ rowcnt,
                      int* __restrict__ maxrow, const int* __restrict__ row_off, u64* __restrict__ key_u,
                      float4* __restrict__ vec_u) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const mg_struct_dev* st = &str[atom_str[i]];
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  const double RC = st->rc2;
  const int nb0 = st->nb[0], nb1 = st->nb[1], nb2 = st->nb[2];
  int local = bin[i] - st->bin0;
  const int b2 = local % nb2, b1 = (local / nb2) % nb1, b0 = local / (nb1 * nb2);
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0;  // per-species counts
  int pos = FILL ? row_off[i] : 0;
  for (int d0 = -st->K[0]; d0 <= st->K[0]; d0++) {
    int w0 = b0 + d0, ic0 = floordiv_i(w0, nb0);
    w0 -= ic0 * nb0;
    for (int d1 = -st->K[1]; d1 <= st->K[1]; d1++) {
      int w1 = b1 + d1, ic1 = floordiv_i(w1, nb1);
      w1 -= ic1 * nb1;
      for (int d2 = -st->K[2]; d2 <= st->K[2]; d2++) {
        int w2 = b2 + d2, ic2 = floordiv_i(w2, nb2);
        w2 -= ic2 * nb2;
        double sh[3];
        for (int q = 0; q < 3; q++) {
          double v = 0.0;
          v += (double)ic0 * st->L[q];
          v += (double)ic1 * st->L[3 + q];
          v += (double)ic2 * st->L[6 + q];
          sh[q] = v;
        }
        const bool zero_img = (ic0 == 0 && ic1 == 0 && ic2 == 0);
        int w = st->bin0 + (w0 * nb1 + w1) * nb2 + w2;
        int s = bin_start[w], e = bin_start[w + 1];
        for (int p = s; p < e; p++) {
          int j = bin_atoms[p];
          if (j == i && zero_img) continue;
          double a0 = sh[0]; a0 += x[3 * j] - xi0;
          double a1 = sh[1]; a1 += x[3 * j + 1] - xi1;
          double a2 = sh[2]; a2 += x[3 * j + 2] - xi2;
          double r = 0.0;
          r += a0 * a0;
          r += a1 * a1;
          r += a2 * a2;
          if (r < RC) {
            int sj = spc[j];
            if (FILL) {
              key_u[pos] = make_key(sj, j, pack_img(ic0, ic1, ic2));
              vec_u[pos] = make_float4((float)a0, (float)a1, (float)a2, (float)sqrt(r));
              pos++;
            } else {
              c0 += sj == 0; c1 += sj == 1; c2 += sj == 2; c3 += sj == 3;
            }
          }
        }
      }
    }
  }
  if (!FILL) {
    cnt[4 * i] = c0; cnt[4 * i + 1] = c1; cnt[4 * i + 2] = c2; cnt[4 * i + 3] = c3;
    int t = c0 + c1 + c2 + c3;
    rowcnt[i] = t;
    atomicMax(maxrow, t);
  }
}

// One wave (block of 64) per row: bitonic sort of the keys in LDS, then gather (key, vec) into place.
__global__ __launch_bounds__(64) void k_row_sort(int N, const int* __restrict__ row_off, const u64* __restrict__ key_u,
                                                 const float4* __restrict__ vec_u, u64* __restrict__ key,
                                                 float4* __restrict__ vec, int cap) {
  extern __shared__ unsigned char smem_sort[];
  u64* sk = (u64*)smem_sort;
  int* si = (int*)(sk + cap);
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int off = row_off[i], M = row_off[i + 1] - off;
  if (M > cap) {  // long row: one lane sorts in place (rare: only when the LDS tile is too small)
    if (lane == 0) {
      for (int t = 0; t < M; t++) { key[off + t] = key_u[off + t]; vec[off + t] = vec_u[off + t]; }
      for (int p = 1; p < M; p++) {
        u64 kv = key[off + p];
        float4 vv = vec[off + p];
        int q = p - 1;
        while (q >= 0 && key[off + q] > kv) { key[off + q + 1] = key[off + q]; vec[off + q + 1] = vec[off + q]; q--; }
        key[off + q + 1] = kv;
        vec[off + q + 1] = vv;
      }
    }
    return;
  }
  int n2 = 1;
  while (n2 < M) n2 <<= 1;
  for (int t = lane; t < n2; t += WAVE) {
    sk[t] = t < M ? key_u[off + t] : ~0ull;
    si[t] = t;
  }
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < n2; t += WAVE) {
        int p = t ^ j;
        if (p > t) {
          bool asc = (t & k) == 0;
          u64 a = sk[t], b = sk[p];
          if ((a > b) == asc) {
            sk[t] = b; sk[p] = a;
            int ia = si[t]; si[t] = si[p]; si[p] = ia;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = lane; t < M; t += WAVE) {
    key[off + t] = sk[t];
    vec[off + t] = vec_u[off + si[t]];
  }
}

// Reverse edge of e = (i -> j, img): the entry (j -> i, -img) of row j, found by binary search on keys.
__global__ void k_rev(int N, const int* __restrict__ spc, const int* __restrict__ row_off, const u64* __restrict__ key,
                      int* __restrict__ rev, int* __restrict__ err) {
  int i = blockIdx.x;
  int off = row_off[i], M = row_off[i + 1] - off;
  int si = spc[i];
  for (int t = threadIdx.x; t < M; t += blockDim.x) {
    int e = off + t;
    u64 k = key[e];
    int j = key_j(k);
    u64 target = make_key(si, i, neg_img((unsigned)(k & 0x3fffffffull)));
    int lo = row_off[j], hi = row_off[j + 1] - 1, found = -1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      u64 v = key[mid];
      if (v == target) { found = mid; break; }
      if (v < target) lo = mid + 1; else hi = mid - 1;
    }
    if (found < 0) { atomicOr(err, 1); found = e; }
    rev[e] = found;
  }
}

// ---------------------------------------------------------------------------------------------- wave helpers
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k, WAVE);
  return v;
}

// Reduce-scatter of 64 per-lane accumulators: afterwards lane l holds the wave sum of accumulator l.
template <int A>
__device__ __forceinline__ void rs_step(float* v, int lane, int k) {
  const bool hi = (lane & k) != 0;
#pragma unroll
  for (int p = 0; p < A / 2; p++) {
    float send = hi ? v[p] : v[p + A / 2];
    float keep = hi ? v[p + A / 2] : v[p];
    v[p] = keep + __shfl_xor(send, k, WAVE);
  }
}
__device__ __forceinline__ float reduce_scatter64(float* v, int lane) {
  rs_step<64>(v, lane, 32);
  rs_step<32>(v, lane, 16);
  rs_step<16>(v, lane, 8);
  rs_step<8>(v, lane, 4);
  rs_step<4>(v, lane, 2);
  rs_step<2>(v, lane, 1);
  return v[0];
}

__device__ __forceinline__ float ipowf(float u, int n) {
  float r = 1.0f;
  while (n) {
    if (n & 1) r *= u;
    u *= u;
    n >>= 1;
  }
  return r;
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
// cos(pi * t) and sin(pi * t) for t in [0, 1]: the hardware takes revolutions, cos(2 pi y)
__device__ __forceinline__ float cospi_(float t) { return __builtin_amdgcn_cosf(0.5f * t); }
__device__ __forceinline__ float sinpi_(float t) { return __builtin_amdgcn_sinf(0.5f * t); }

// Per-neighbour LDS record (64 B): (dx, dy, dz, r), (1/r, fc, dfc, -), A[0..7] = exp(-eta4(i,j)_e r^2)
#define NREC 4

struct RowInfo {
  int off, M, si;
  int m0, m1, m2, m3;  // neighbours of each species (rows are sorted by species)
  __device__ __forceinline__ int m(int s) const { return s == 0 ? m0 : (s == 1 ? m1 : (s == 2 ? m2 : m3)); }
  __device__ __forceinline__ int o(int s) const {
    return s == 0 ? 0 : (s == 1 ? m0 : (s == 2 ? m0 + m1 : m0 + m1 + m2));
  }
};

__device__ __forceinline__ void row_info(RowInfo& ri, int i, const int* __restrict__ spc, const int* __restrict__ row_off,
                                         const int* __restrict__ cnt, int S) {
  ri.off = row_off[i];
  ri.M = row_off[i + 1] - ri.off;
  ri.si = spc[i];
  ri.m0 = cnt[4 * i];
  ri.m1 = S > 1 ? cnt[4 * i + 1] : 0;
  ri.m2 = S > 2 ? cnt[4 * i + 2] : 0;
  ri.m3 = S > 3 ? cnt[4 * i + 3] : 0;
}

// stage the row of atom i into LDS
template <int NE>
__device__ __forceinline__ void stage_row(float4* rec, const RowInfo& ri, const float4* __restrict__ vec,
                                          const mg_basis_dev* __restrict__ bp, int lane, int S) {
  for (int t = lane; t < ri.M; t += WAVE) {
    float4 v = vec[ri.off + t];
    int s = 0;
#pragma unroll
    for (int q = 1; q < MG_MAXSPC; q++) s += (q < S && t >= ri.o(q)) ? 1 : 0;
    int p = ri.si * MG_MAXSPC + s;
    float rc = bp->rc[p];
    float tt = v.w / rc;
    float fc = v.w < rc ? 0.5f * (cospi_(tt) + 1.0f) : 0.0f;
    float dfc = v.w < rc ? -0.5f * 3.14159265358979f / rc * sinpi_(tt) : 0.0f;
    rec[NREC * t + 0] = v;
    rec[NREC * t + 1] = make_float4(1.0f / v.w, fc, dfc, 0.0f);
    float A[8];
    float r2 = v.w * v.w;
#pragma unroll
    for (int e = 0; e < 8; e++) A[e] = e < NE ? fast_exp2(-(float)MG_LOG2E * bp->eta4[p][e] * r2) : 0.0f;
    rec[NREC * t + 2] = make_float4(A[0], A[1], A[2], A[3]);
    rec[NREC * t + 3] = make_float4(A[4], A[5], A[6], A[7]);
  }
}

// ---- G4 function patterns on the (eta, zeta, lambda) grid: bitmask of the used (zeta, lambda) cells of each eta
// (cell q = z*2 + l).  PAT 0 is generic (dense grid, runtime integer zeta); PAT 1 is the 43-function table of the
// shipped v0/v1 models, PAT 2 the 22-function table of the shipped vp models (both zeta = {1,2,4,16},
// lambda = {-1,+1}).  Unused cells cost nothing: the masks prune them at compile time.
template <int PAT> struct Pat;
template <> struct Pat<0> {
  static constexpr int NE = 8;
  static constexpr bool ZSTD = false;
  static constexpr unsigned mask(int e) { return 0xFFu; }
};
template <> struct Pat<1> {
  static constexpr int NE = 7;
  static constexpr bool ZSTD = true;
  static constexpr unsigned mask(int e) { return e < 3 ? 0x0Fu : (e < 6 ? 0xFFu : 0x7Fu); }
};
template <> struct Pat<2> {
  static constexpr int NE = 6;
  static constexpr bool ZSTD = true;
  static constexpr unsigned mask(int e) { return e < 2 ? 0x0Fu : (e == 2 ? 0x0Au : 0xAAu); }
};

// P_zl(c) = 2^(1-zeta) (1 + lambda c)^zeta and dP/dc, cell q = z*2 + l
template <int PAT>
__device__ __forceinline__ void angular(const mg_basis_dev* __restrict__ bp, float cs, float* P, float* Pd) {
#pragma unroll
  for (int l = 0; l < MG_MAXL; l++) {
    const float lam = bp->lam[l];
    const float u = 1.0f + lam * cs;
    if (Pat<PAT>::ZSTD) {
      const float u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u16 = u8 * u8;
      P[0 + l] = bp->pref[0] * u;
      P[2 + l] = bp->pref[1] * u2;
      P[4 + l] = bp->pref[2] * u4;
      P[6 + l] = bp->pref[3] * u16;
      if (Pd) {
        Pd[0 + l] = bp->pref[0] * lam;
        Pd[2 + l] = bp->pref[1] * 2.0f * lam * u;
        Pd[4 + l] = bp->pref[2] * 4.0f * lam * u2 * u;
        Pd[6 + l] = bp->pref[3] * 16.0f * lam * (u8 * u4) * (u2 * u);
      }
    } else {
#pragma unroll
      for (int z = 0; z < MG_MAXZ; z++) {
        int zt = bp->zeta[z];
        float pm1 = ipowf(u, zt - 1);
        P[z * MG_MAXL + l] = bp->pref[z] * pm1 * u;
        if (Pd) Pd[z * MG_MAXL + l] = bp->pref[z] * (float)zt * lam * pm1;
      }
    }
  }
}

// Unordered neighbour pairs {j, k} of one species class with r_jk < rc.  Enumeration: triangle (same
// species) in circulant order (j, j + d mod m), d = 1 .. m/2, each pair once; rectangle (a x b) with the longer
// side as the lane-fast index.  Passing pairs are compacted into an LDS list (ballot + lane prefix count) and
// processed with all 64 lanes busy; about half of the enumerated pairs fail the r_jk test in bulk metals.
#define PCH 512
template <typename F>
__device__ __forceinline__ void for_pairs(const float4* rec, unsigned* plist, bool tri, int oa, int ma, int ob, int mb,
                                          float rc2, int lane, F&& body) {
  int mf, ms;  // fast / slow extents
  if (tri) {
    if (ma < 2) return;
    mf = ma; ms = ma >> 1;
  } else {
    if (ma == 0 || mb == 0) return;
    mf = ma >= mb ? ma : mb; ms = ma >= mb ? mb : ma;
  }
  const bool even = (ma & 1) == 0;
  int f = lane, sl = tri ? 1 : 0;
  while (f >= mf) { f -= mf; sl++; }
  bool live = tri ? sl <= ms : sl < ms;
  int cnt = 0;
  for (;;) {
    bool ok = false;
    unsigned pk = 0;
    if (live) {
      int jj, kk;
      bool use = true;
      if (tri) {
        use = !(even && sl == ms && f >= ms);
        int k = f + sl;
        if (k >= ma) k -= ma;
        jj = oa + f; kk = oa + k;
      } else if (ma >= mb) {
        jj = oa + f; kk = ob + sl;
      } else {
        jj = oa + sl; kk = ob + f;
      }
      if (use) {
        float4 a = rec[NREC * jj], b = rec[NREC * kk];
        float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        ok = cx * cx + cy * cy + cz * cz < rc2;
        pk = (unsigned)jj | ((unsigned)kk << 16);
      }
      f += WAVE;
      while (f >= mf) { f -= mf; sl++; }
      live = tri ? sl <= ms : sl < ms;
    }
    const unsigned long long bm = __ballot(ok);
    const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
    if (ok) plist[cnt + below] = pk;
    cnt += __popcll(bm);
    const bool more = __ballot(live) != 0ull;
    if (cnt > PCH - WAVE || !more) {
      __syncthreads();
      for (int t = lane; t < cnt; t += WAVE) {
        unsigned p = plist[t];
        body((int)(p & 0xffffu), (int)(p >> 16));
      }
      __syncthreads();
      cnt = 0;
      if (!more) break;
    }
  }
}

// ---------------------------------------------------------------------------------------------- descriptors
template <int PAT>
__global__ __launch_bounds__(64) void k_desc(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
                                             const int* __restrict__ cnt, const float4* __restrict__ vec,
                                             const mg_basis_dev* __restrict__ bp, float* __restrict__ G,
                                             float* __restrict__ erep, float urep, int cap) {
  constexpr int NE = Pat<PAT>::NE;
  extern __shared__ float4 smem_d[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int D = bp->D;
  RowInfo ri;
  row_info(ri, i, spc, row_off, cnt, S);
  float4* rec = smem_d;
  unsigned* plist = (unsigned*)(smem_d + NREC * cap);
  stage_row<NE>(rec, ri, vec, bp, lane, S);
  __syncthreads();
  float* Gi = G + (size_t)i * D;

  // ---- G2 (+ UREP energy) per neighbour species block
  float eu = 0.0f;
  for (int s = 0; s < S; s++) {
    const int blk = bp->ind2[(s - ri.si + S) % S];
    if (ri.m(s) == 0) {
      if (lane < bp->ng2) Gi[blk + lane] = 0.0f;
      continue;
    }
    const int p = ri.si * MG_MAXSPC + s;
    float acc[MG_MAXG2];
#pragma unroll
    for (int c = 0; c < MG_MAXG2; c++) acc[c] = 0.0f;
    const float Rrep = bp->urep_rm[ri.si] + bp->urep_rm[s];
    for (int t = ri.o(s) + lane; t < ri.o(s) + ri.m(s); t += WAVE) {
      float4 v = rec[NREC * t];
      float fc = rec[NREC * t + 1].y;
      float r = v.w;
#pragma unroll
      for (int c = 0; c < MG_MAXG2; c++)
        if (c < bp->ng2) {
          float dr = r - bp->rs2[p][c];
          acc[c] += fc * fast_exp2(-(float)MG_LOG2E * bp->eta2[p][c] * dr * dr);
        }
      if (urep > 0.0f && r < Rrep) eu += urep * 0.5f * (1.0f + cospi_(r / Rrep)) * Rrep / r;
    }
#pragma unroll
    for (int c = 0; c < MG_MAXG2; c++)
      if (c < bp->ng2) {
        float v = wave_sum(acc[c]);
        if (lane == 0) Gi[blk + c] = v;
      }
  }
  if (erep) {
    float v = wave_sum(eu);
    if (lane == 0) erep[i] = v;
  }

  // ---- G4 per species class (sa <= sb)
  for (int sa = 0; sa < S; sa++)
    for (int sb = sa; sb < S; sb++) {
      const int ds1 = (sa - ri.si + S) % S, ds2 = (sb - ri.si + S) % S;
      const int blk = bp->ind4[ds1][ds2];
      const int pjk = sa * MG_MAXSPC + sb;
      const float rc = bp->rc[pjk];
      const float rinv = 1.0f / rc;
      float ejk[NE];
#pragma unroll
      for (int e = 0; e < NE; e++) ejk[e] = -(float)MG_LOG2E * bp->eta4[pjk][e];
      float V[MG_GRID];
#pragma unroll
      for (int q = 0; q < MG_GRID; q++) V[q] = 0.0f;
      auto body = [&](int j, int k) {
        float4 a = rec[NREC * j], b = rec[NREC * k];
        float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        float r3s = cx * cx + cy * cy + cz * cz;
        float4 ia = rec[NREC * j + 1], ib = rec[NREC * k + 1];
        float r3 = __builtin_sqrtf(r3s);
        float fc3 = 0.5f * (cospi_(r3 * rinv) + 1.0f);
        float cs = (a.x * b.x + a.y * b.y + a.z * b.z) * ia.x * ib.x;
        float F = 2.0f * ia.y * ib.y * fc3;
        float4 Aj0 = rec[NREC * j + 2], Aj1 = rec[NREC * j + 3], Ak0 = rec[NREC * k + 2], Ak1 = rec[NREC * k + 3];
        float Aj[8] = {Aj0.x, Aj0.y, Aj0.z, Aj0.w, Aj1.x, Aj1.y, Aj1.z, Aj1.w};
        float Ak[8] = {Ak0.x, Ak0.y, Ak0.z, Ak0.w, Ak1.x, Ak1.y, Ak1.z, Ak1.w};
        float P[MG_MAXZ * MG_MAXL];
        angular<PAT>(bp, cs, P, nullptr);
#pragma unroll
        for (int e = 0; e < NE; e++) {
          float B = F * Aj[e] * Ak[e] * fast_exp2(ejk[e] * r3s);
#pragma unroll
          for (int q = 0; q < MG_MAXZ * MG_MAXL; q++)
            if (Pat<PAT>::mask(e) & (1u << q)) V[e * MG_MAXZ * MG_MAXL + q] += B * P[q];
        }
      };
      for_pairs(rec, plist, sa == sb, ri.o(sa), ri.m(sa), ri.o(sb), ri.m(sb), rc * rc, lane, body);
      float mine = reduce_scatter64(V, lane);  // lane l: grid cell l
      int cell = lane < bp->ng4 ? bp->g4cell[lane] : 0;
      float v = __shfl(mine, cell, WAVE);
      if (lane < bp->ng4) Gi[blk + lane] = v;
    }
}

// ---------------------------------------------------------------------------------------------- MLP (FP32 VALU)
__device__ __forceinline__ float tanh_f(float x) {
  // tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2|x|): no overflow, full FP32 accuracy
  float ax = fabsf(x);
  float e = __expf(-2.0f * ax);
  float t = (1.0f - e) / (1.0f + e);
  return copysignf(t, x);
}

// one thread per atom; hidden activations in LDS [unit][thread]
__global__ __launch_bounds__(64) void k_mlp_fp32(int N, const int* __restrict__ spc, const float* __restrict__ G,
                                                 const float* __restrict__ W, int wstride, int D, int H1, int H2,
                                                 float* __restrict__ Ei, float* __restrict__ dEdG) {
  extern __shared__ float smem_m[];
  const int t = threadIdx.x;
  const int i = blockIdx.x * 64 + t;
  float* h1 = smem_m;            // [H1][64]
  float* h2 = smem_m + H1 * 64;  // [H2][64]
  if (i >= N) return;
  const float* w = W + (size_t)spc[i] * wstride;
  const float* W0 = w;
  const float* b0 = W0 + H1 * D;
  const float* W1 = b0 + H1;
  const float* b1 = W1 + H2 * H1;
  const float* W2 = b1 + H2;
  const float* b2 = W2 + H2;
  const float* xo = b2 + 1;
  const float* g = G + (size_t)i * D;
  for (int o = 0; o < H1; o++) {
    float acc = b0[o];
    const float* wr = W0 + o * D;
    for (int n = 0; n < D; n++) acc += (g[n] - xo[n]) * wr[n];
    h1[o * 64 + t] = tanh_f(acc);
  }
  float e = b2[0];
  for (int o = 0; o < H2; o++) {
    float acc = b1[o];
    const float* wr = W1 + o * H1;
    for (int n = 0; n < H1; n++) acc += h1[n * 64 + t] * wr[n];
    float h = tanh_f(acc);
    e += W2[o] * h;
    h2[o * 64 + t] = W2[o] * (1.0f - h * h);  // delta2
  }
  Ei[i] = e;
  for (int n = 0; n < H1; n++) {
    float s = 0.0f;
    for (int o = 0; o < H2; o++) s += h2[o * 64 + t] * W1[o * H1 + n];
    float h = h1[n * 64 + t];
    h1[n * 64 + t] = s * (1.0f - h * h);  // delta1
  }
  float* out = dEdG + (size_t)i * D;
  for (int l = 0; l < D; l++) {
    float s = 0.0f;
    for (int n = 0; n < H1; n++) s += h1[n * 64 + t] * W0[n * D + l];
    out[l] = s;
  }
}

// ---------------------------------------------------------------------------------------------- forces
// Per atom i: gradient of T_i = sum_n g_n G_n(i) with respect to its neighbour vectors.  Pair terms
// act along d_ij; a triplet (j, k) with a = d_ij, b = d_ik, c = b - a, cos = a.b/(|a||b|):
//   T = F * sum_e S_e B_e, F = f1 f2 f3, B_e = 2 exp(-(e_ij r1^2 + e_ik r2^2 + e_jk r3^2)),
//   S_e = sum_(z,l) g(e,z,l) P_zl(cos), S'_e with P'_zl,
//   dT/dr1 = -2 r1 F sum_e e_ij S_e B_e + f1' f2 f3 sum_e S_e B_e (r2, r3 likewise), dT/dcos = F sum_e S'_e B_e,
//   dT/da = dT/dr1 a/r1 - dT/dr3 c/r3 + dT/dcos (b/(r1 r2) - cos a/r1^2), dT/db symmetric.
// Forces: F_j -= dT/da, F_k -= dT/db, F_i += dT/da + dT/db; virial U -= dT/da (x) a + dT/db (x) b.
template <int PAT>
__global__ __launch_bounds__(64) void k_force(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
                                              const int* __restrict__ cnt, const float4* __restrict__ vec,
                                              const mg_basis_dev* __restrict__ bp, const float* __restrict__ dEdG,
                                              float4* __restrict__ fself, float4* __restrict__ fe,
                                              float* __restrict__ vir, float urep, int cap) {
  constexpr int NE = Pat<PAT>::NE;
  extern __shared__ float4 smem_f[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int D = bp->D;
  RowInfo ri;
  row_info(ri, i, spc, row_off, cnt, S);
  float4* rec = smem_f;
  float* fnb = (float*)(smem_f + NREC * cap);     // [cap][4]
  unsigned* plist = (unsigned*)(fnb + 4 * cap);  // [PCH]
  stage_row<NE>(rec, ri, vec, bp, lane, S);
  for (int t = lane; t < ri.M; t += WAVE) {
    fnb[4 * t] = 0.0f; fnb[4 * t + 1] = 0.0f; fnb[4 * t + 2] = 0.0f; fnb[4 * t + 3] = 0.0f;
  }
  __syncthreads();
  const float* gi = dEdG + (size_t)i * D;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  float u0 = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, u4 = 0.f, u5 = 0.f;

  // ---- pair terms (G2) and UREP
  for (int s = 0; s < S; s++) {
    if (ri.m(s) == 0) continue;
    const int p = ri.si * MG_MAXSPC + s;
    const int blk = bp->ind2[(s - ri.si + S) % S];
    const float Rrep = bp->urep_rm[ri.si] + bp->urep_rm[s];
    for (int t = ri.o(s) + lane; t < ri.o(s) + ri.m(s); t += WAVE) {
      float4 v = rec[NREC * t];
      float4 w = rec[NREC * t + 1];
      float r = v.w, fc = w.y, dfc = w.z;
      float dT = 0.0f;
#pragma unroll
      for (int c = 0; c < MG_MAXG2; c++)
        if (c < bp->ng2) {
          float eta = bp->eta2[p][c];
          float dr = r - bp->rs2[p][c];
          float ex = fast_exp2(-(float)MG_LOG2E * eta * dr * dr);
          dT += gi[blk + c] * ex * (-2.0f * eta * dr * fc + dfc);
        }
      float sc = dT * w.x;
      float gx = sc * v.x, gy = sc * v.y, gz = sc * v.z;
      fnb[4 * t] -= gx; fnb[4 * t + 1] -= gy; fnb[4 * t + 2] -= gz;
      fx += gx; fy += gy; fz += gz;
      u0 -= gx * v.x; u1 -= gy * v.y; u2 -= gz * v.z;
      u3 -= 0.5f * (gx * v.y + gy * v.x); u4 -= 0.5f * (gy * v.z + gz * v.y); u5 -= 0.5f * (gz * v.x + gx * v.z);
      if (urep > 0.0f && r < Rrep) {
        // per ordered pair, acting on atom i only, as CELL_UREP (src/cmod.c:95-116)
        float FC = 0.5f * (1.0f + cospi_(r / Rrep));
        float dFC = -0.5f * sinpi_(r / Rrep) * 3.14159265358979f / Rrep;
        float tt = -2.0f * urep * Rrep / (r * r) * (-FC / r + dFC);
        fx -= tt * v.x; fy -= tt * v.y; fz -= tt * v.z;
        u0 += 0.5f * tt * v.x * v.x; u1 += 0.5f * tt * v.y * v.y; u2 += 0.5f * tt * v.z * v.z;
        u3 += 0.5f * tt * v.x * v.y; u4 += 0.5f * tt * v.y * v.z; u5 += 0.5f * tt * v.z * v.x;
      }
    }
  }

  // ---- triplet terms (G4) per species class
  for (int sa = 0; sa < S; sa++)
    for (int sb = sa; sb < S; sb++) {
      const int ds1 = (sa - ri.si + S) % S, ds2 = (sb - ri.si + S) % S;
      const int blk = bp->ind4[ds1][ds2];
      const int pjk = sa * MG_MAXSPC + sb, pij = ri.si * MG_MAXSPC + sa, pik = ri.si * MG_MAXSPC + sb;
      const float rc = bp->rc[pjk];
      const float rinv = 1.0f / rc;
      const float dfk = -0.5f * 3.14159265358979f * rinv;
      float ejk[NE], eij[NE], eik[NE], e3[NE];
#pragma unroll
      for (int e = 0; e < NE; e++) {
        ejk[e] = -(float)MG_LOG2E * bp->eta4[pjk][e];
        e3[e] = bp->eta4[pjk][e];
        eij[e] = bp->eta4[pij][e];
        eik[e] = bp->eta4[pik][e];
      }
      float gt[NE * MG_MAXZ * MG_MAXL];
#pragma unroll
      for (int q = 0; q < NE * MG_MAXZ * MG_MAXL; q++) {
        if (Pat<PAT>::mask(q / (MG_MAXZ * MG_MAXL)) & (1u << (q % (MG_MAXZ * MG_MAXL)))) {
          int n = bp->g4map[q];
          gt[q] = n >= 0 ? gi[blk + n] : 0.0f;
        } else {
          gt[q] = 0.0f;
        }
      }
      auto body = [&](int j, int k) {
        float4 a = rec[NREC * j], b = rec[NREC * k];
        float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        float r3s = cx * cx + cy * cy + cz * cz;
        float4 ia = rec[NREC * j + 1], ib = rec[NREC * k + 1];
        float r3 = __builtin_sqrtf(r3s);
        float i3 = 1.0f / r3;
        float f3 = 0.5f * (cospi_(r3 * rinv) + 1.0f);
        float d3 = dfk * sinpi_(r3 * rinv);
        float i1 = ia.x, i2 = ib.x, f1 = ia.y, f2 = ib.y, d1 = ia.z, d2 = ib.z;
        float cs = (a.x * b.x + a.y * b.y + a.z * b.z) * i1 * i2;
        float4 Aj0 = rec[NREC * j + 2], Aj1 = rec[NREC * j + 3], Ak0 = rec[NREC * k + 2], Ak1 = rec[NREC * k + 3];
        float Aj[8] = {Aj0.x, Aj0.y, Aj0.z, Aj0.w, Aj1.x, Aj1.y, Aj1.z, Aj1.w};
        float Ak[8] = {Ak0.x, Ak0.y, Ak0.z, Ak0.w, Ak1.x, Ak1.y, Ak1.z, Ak1.w};
        float P[MG_MAXZ * MG_MAXL], Pd[MG_MAXZ * MG_MAXL];
        angular<PAT>(bp, cs, P, Pd);
        float X = 0.f, Y1 = 0.f, Y2 = 0.f, Y3 = 0.f, Zc = 0.f;
#pragma unroll
        for (int e = 0; e < NE; e++) {
          float B = 2.0f * Aj[e] * Ak[e] * fast_exp2(ejk[e] * r3s);
          float Se = 0.f, Sd = 0.f;
#pragma unroll
          for (int q = 0; q < MG_MAXZ * MG_MAXL; q++)
            if (Pat<PAT>::mask(e) & (1u << q)) {
              Se += gt[e * MG_MAXZ * MG_MAXL + q] * P[q];
              Sd += gt[e * MG_MAXZ * MG_MAXL + q] * Pd[q];
            }
          float SB = Se * B;
          X += SB;
          Y1 += eij[e] * SB;
          Y2 += eik[e] * SB;
          Y3 += e3[e] * SB;
          Zc += Sd * B;
        }
        float r1 = a.w, r2 = b.w;
        float Fp = f1 * f2 * f3;
        float dT1 = -2.0f * r1 * Fp * Y1 + d1 * f2 * f3 * X;
        float dT2 = -2.0f * r2 * Fp * Y2 + f1 * d2 * f3 * X;
        float dT3 = -2.0f * r3 * Fp * Y3 + f1 * f2 * d3 * X;
        float dTc = Fp * Zc;
        float i12 = i1 * i2;
        float ca = dT1 * i1 - dTc * cs * i1 * i1;  // coefficient of a in dT/da
        float cb = dTc * i12;                      // coefficient of b in dT/da (and of a in dT/db)
        float c3c = dT3 * i3;                      // coefficient of c
        float cbb = dT2 * i2 - dTc * cs * i2 * i2; // coefficient of b in dT/db
        float gax = ca * a.x + cb * b.x - c3c * cx, gay = ca * a.y + cb * b.y - c3c * cy, gaz = ca * a.z + cb * b.z - c3c * cz;
        float gbx = cbb * b.x + cb * a.x + c3c * cx, gby = cbb * b.y + cb * a.y + c3c * cy, gbz = cbb * b.z + cb * a.z + c3c * cz;
        atomicAdd(&fnb[4 * j], -gax); atomicAdd(&fnb[4 * j + 1], -gay); atomicAdd(&fnb[4 * j + 2], -gaz);
        atomicAdd(&fnb[4 * k], -gbx); atomicAdd(&fnb[4 * k + 1], -gby); atomicAdd(&fnb[4 * k + 2], -gbz);
        fx += gax + gbx; fy += gay + gby; fz += gaz + gbz;
        u0 -= gax * a.x + gbx * b.x;
        u1 -= gay * a.y + gby * b.y;
        u2 -= gaz * a.z + gbz * b.z;
        u3 -= 0.5f * (gax * a.y + gay * a.x + gbx * b.y + gby * b.x);
        u4 -= 0.5f * (gay * a.z + gaz * a.y + gby * b.z + gbz * b.y);
        u5 -= 0.5f * (gaz * a.x + gax * a.z + gbz * b.x + gbx * b.z);
      };
      for_pairs(rec, plist, sa == sb, ri.o(sa), ri.m(sa), ri.o(sb), ri.m(sb), rc * rc, lane, body);
    }
  fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
  u0 = wave_sum(u0); u1 = wave_sum(u1); u2 = wave_sum(u2);
  u3 = wave_sum(u3); u4 = wave_sum(u4); u5 = wave_sum(u5);
  if (lane == 0) {
    fself[i] = make_float4(fx, fy, fz, 0.0f);
    float* vv = vir + 6 * (size_t)i;
    vv[0] = u0; vv[1] = u1; vv[2] = u2; vv[3] = u3; vv[4] = u4; vv[5] = u5;
  }
  __syncthreads();
  for (int t = lane; t < ri.M; t += WAVE) fe[ri.off + t] = make_float4(fnb[4 * t], fnb[4 * t + 1], fnb[4 * t + 2], 0.0f);
}

__global__ void k_gather(int N, const int* __restrict__ row_off, const int* __restrict__ rev,
                         const float4* __restrict__ fself, const float4* __restrict__ fe, double* __restrict__ F) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  float4 s = fself[a];
  double fx = s.x, fy = s.y, fz = s.z;
  for (int e = row_off[a]; e < row_off[a + 1]; e++) {
    float4 v = fe[rev[e]];
    fx += v.x; fy += v.y; fz += v.z;
  }
  F[3 * a] = fx; F[3 * a + 1] = fy; F[3 * a + 2] = fz;
}

// one block of 256 per structure; FP64, fixed summation order
__global__ __launch_bounds__(256) void k_reduce(int nstr, const mg_struct_dev* __restrict__ str, const float* __restrict__ Ei,
                                                const float* __restrict__ erep, const float* __restrict__ vir, int have_vir,
                                                double* __restrict__ Eout, double* __restrict__ Uout) {
  __shared__ double sh[7][256];
  const int s = blockIdx.x, t = threadIdx.x;
  const int a0 = str[s].atom0, n = str[s].natom;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int a = a0 + t; a < a0 + n; a += 256) {
    acc[0] += (double)Ei[a] + (erep ? (double)erep[a] : 0.0);
    if (have_vir)
      for (int q = 0; q < 6; q++) acc[1 + q] += (double)vir[6 * (size_t)a + q];
  }
  for (int q = 0; q < 7; q++) sh[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 7; q++) sh[q][t] += sh[q][t + w];
    __syncthreads();
  }
  if (t == 0) {
    Eout[s] = sh[0][0];
    for (int q = 0; q < 6; q++) Uout[6 * s + q] = sh[1 + q][0];
  }
}

// ---------------------------------------------------------------------------------------------- launchers
extern "C" {

int mgk_scan(const int* in, int* out, int n, void* tmp, size_t* tmp_bytes, hipStream_t st) {
  // exclusive scan of n values into out[0..n]; out[n] = total.  tmp == NULL queries the scratch size.
  size_t need = 0;
  hipcub::DeviceScan::InclusiveSum(nullptr, need, in, out + 1, n, st);
  if (!tmp) { *tmp_bytes = need; return 0; }
  if (*tmp_bytes < need) return -1;
  if (hipMemsetAsync(out, 0, sizeof(int), st) != hipSuccess) return -1;
  if (n > 0 && hipcub::DeviceScan::InclusiveSum(tmp, need, in, out + 1, n, st) != hipSuccess) return -1;
  return 0;
}

int mgk_prepare(int N, const double* xin, const int* atom_str, const mg_struct_dev* str, double* x, int* bin,
                int* bin_count, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_prepare, dim3((N + 255) / 256), dim3(256), 0, st, N, xin, atom_str, str, x, bin, bin_count);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_bins(int N, int nbins, const int* bin, const int* bin_start, int* bin_cur, int* bin_atoms, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_bin_fill, dim3((N + 255) / 256), dim3(256), 0, st, N, bin, bin_start, bin_cur, bin_atoms);
  hipLaunchKernelGGL(k_bin_sort, dim3((nbins + 255) / 256), dim3(256), 0, st, nbins, bin_start, bin_atoms);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_nbr_count(int N, const double* x, const int* spc, const int* atom_str, const mg_struct_dev* str, const int* bin,
                  const int* bin_start, const int* bin_atoms, int* cnt, int* rowcnt, int* maxrow, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_nbr<false>, dim3((N + 127) / 128), dim3(128), 0, st, N, x, spc, atom_str, str, bin, bin_start,
                     bin_atoms, cnt, rowcnt, maxrow, (const int*)nullptr, (u64*)nullptr, (float4*)nullptr);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_nbr_fill(int N, const double* x, const int* spc, const int* atom_str, const mg_struct_dev* str, const int* bin,
                 const int* bin_start, const int* bin_atoms, const int* row_off, void* key_u, void* vec_u,
                 void* key, void* vec, int* rev, int* err, int maxrow, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_nbr<true>, dim3((N + 127) / 128), dim3(128), 0, st, N, x, spc, atom_str, str, bin, bin_start,
                     bin_atoms, (int*)nullptr, (int*)nullptr, (int*)nullptr, row_off, (u64*)key_u, (float4*)vec_u);
  int cap = 64;
  while (cap < maxrow && cap < 1024) cap <<= 1;
  size_t sm = (size_t)cap * (sizeof(u64) + sizeof(int));
  hipLaunchKernelGGL(k_row_sort, dim3(N), dim3(64), sm, st, N, row_off, (const u64*)key_u, (const float4*)vec_u,
                     (u64*)key, (float4*)vec, cap);
  hipLaunchKernelGGL(k_rev, dim3(N), dim3(64), 0, st, N, spc, row_off, (const u64*)key, rev, err);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_desc(int N, int S, int pat, const int* spc, const int* row_off, const int* cnt, const void* vec,
             const mg_basis_dev* bp, float* G, float* erep, float urep, int maxrow, hipStream_t st) {
  if (N == 0) return 0;
  int cap = maxrow < 1 ? 1 : maxrow;
  size_t sm = (size_t)cap * NREC * sizeof(float4) + PCH * sizeof(unsigned);
  if (pat == 1)
    hipLaunchKernelGGL(k_desc<1>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, G, erep, urep, cap);
  else if (pat == 2)
    hipLaunchKernelGGL(k_desc<2>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, G, erep, urep, cap);
  else
    hipLaunchKernelGGL(k_desc<0>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, G, erep, urep, cap);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_mlp(int N, const int* spc, const float* G, const float* W, int wstride, int D, int H1, int H2, float* Ei,
            float* dEdG, hipStream_t st) {
  if (N == 0) return 0;
  size_t sm = (size_t)(H1 + H2) * 64 * sizeof(float);
  hipLaunchKernelGGL(k_mlp_fp32, dim3((N + 63) / 64), dim3(64), sm, st, N, spc, G, W, wstride, D, H1, H2, Ei, dEdG);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_force(int N, int S, int pat, const int* spc, const int* row_off, const int* cnt, const void* vec,
              const mg_basis_dev* bp, const float* dEdG, void* fself, void* fe, float* vir, float urep, int maxrow,
              hipStream_t st) {
  if (N == 0) return 0;
  int cap = maxrow < 1 ? 1 : maxrow;
  size_t sm = (size_t)cap * (NREC + 1) * sizeof(float4) + PCH * sizeof(unsigned);
  if (pat == 1)
    hipLaunchKernelGGL(k_force<1>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, dEdG,
                       (float4*)fself, (float4*)fe, vir, urep, cap);
  else if (pat == 2)
    hipLaunchKernelGGL(k_force<2>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, dEdG,
                       (float4*)fself, (float4*)fe, vir, urep, cap);
  else
    hipLaunchKernelGGL(k_force<0>, dim3(N), dim3(64), sm, st, N, S, spc, row_off, cnt, (const float4*)vec, bp, dEdG,
                       (float4*)fself, (float4*)fe, vir, urep, cap);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_gather(int N, const int* row_off, const int* rev, const void* fself, const void* fe, double* F, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_gather, dim3((N + 255) / 256), dim3(256), 0, st, N, row_off, rev, (const float4*)fself,
                     (const float4*)fe, F);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mgk_reduce(int nstr, const mg_struct_dev* str, const float* Ei, const float* erep, const float* vir, int have_vir,
               double* E, double* U, hipStream_t st) {
  if (nstr == 0) return 0;
  hipLaunchKernelGGL(k_reduce, dim3(nstr), dim3(256), 0, st, nstr, str, Ei, erep, vir, have_vir, E, U);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
