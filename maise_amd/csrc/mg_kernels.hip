// mg_kernels.hip -- CDNA4 (gfx950) kernels of the MAISE neural-network potential engine.
//
// Pipeline for one evaluation (all structures of a context at once):
//   k_prepare      CAST_CELL fractional->Cartesian, JAR wrap in FP64 (src/cmod.c:1112-1121, src/cell.c:446-508),
//                  cell-list bin of every atom
//   k_bin_fill     counting sort of atoms into bins (order inside a bin is free: rows are sorted later)
//   k_scan_one     bin offsets (one block, LDS-staged passes)
//   k_nbr<0>       one wave per atom: neighbour count per species over the bin stencil, FP64 test r < Rc^2 in
//                  the reference's expression order (src/cell.c:109-123) so the neighbour SETS are bit-exact
//   k_nbr<1>       fill: 64-bit key (species, j, image) + FP32 (d, r) per edge, row sorted by key in LDS, and
//                  the row's reverse-lookup hash table
//   k_nbr<2>       skin candidates (MD, FIRE): the fill's search at Rc + skin, keys only, when the lists are stale
//   k_nbr_skin     every step with skin lists: LIST's FP64 test on the candidates only, same rows as the fill
//   k_desc<NE>     one wave per atom: G2 over neighbours, G4 over neighbour pairs of each species class,
//                  neighbour tile staged in LDS (record planes), (eta, zeta, lambda) grid accumulators in VGPRs,
//                  reduce-scatter across the wave (src/nprs.c:504-566); with forces, the reverse-edge index of
//                  the row's owned pairs through the partners' hash tables
//   k_mlp_fp32     per-atom D-H1-H2-1 forward and dE/dG backward (src/nmlp.c:367-431 + chain of 444-478)
//   k_mlp_mfma2    the same on MFMA for hidden layers up to 64 wide, two waves per tile (mg_mlp.hip)
//   k_force<NE>    one wave per atom: recomputes pair/triplet geometry, contracts dG/dr with dE/dG in
//                  registers, forces on neighbour slots in LDS without atomics, per-atom virial (no Fn/Sn
//                  storage); each neighbour's force goes to the reverse edge's slot in the neighbour's row
//   k_force_mh/_mm the same with the triplet contraction over eta on the matrix cores (split FP16 in FP32, split
//                  BF16 in the BF16 precisions)
//   k_assemble     one launch: F_a = own term + sum over the atom's own row (deterministic), and the
//                  per-structure FP64 energy and virial sums (chunk sums, then a fixed tree by a structure's last
//                  chunk block)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mg_internal.h"
#include "mg_dev.h"


typedef unsigned long long u64;

// floor(a / b) for |a| < 2^21, 0 < b <= 2^21, with rb = 1.0f / b (correctly rounded): a + 0.5 is exact in
// float, (a + 0.5) / b lies at least 0.5 / b from an integer, and the two roundings move it by under
// (|a| + 0.5) 2^-23 / b, less than that while |a| < 2^21
__device__ __forceinline__ int floordiv_f(int a, float rb) { return (int)floorf(((float)a + 0.5f) * rb); }

// Wave64 scans and reductions through DPP moves in the VALU (no LDS round trips).  Inclusive scan: row shifts by 1, 2,
// 4, 8 inside rows of 16 lanes, then row 0's last lane into row 1 and row 2's into row 3 (row_bcast:15) and row 1's
// into rows 2 and 3 (row_bcast:31).  Sum / min / max: xor 1 and xor 2 inside quads, rotations by 4 and 8 inside rows
// (every lane then holds its row's value), the two broadcasts, and lane 63 read into a scalar register (the same
// value in every lane).  Lanes a move leaves without a source take 0 (scan, sum) or keep their own value (min, max).
template <int CTRL, int RM>
__device__ __forceinline__ int dpp0(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, RM, 0xf, false); }
template <int CTRL, int RM>
__device__ __forceinline__ float dppf0(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, RM, 0xf, false));
}
template <int CTRL, int RM>
__device__ __forceinline__ float dppf_self(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, CTRL, RM, 0xf, false));
}
__device__ __forceinline__ int wave_scan_incl(int v) {
  v += dpp0<0x111, 0xf>(v);
  v += dpp0<0x112, 0xf>(v);
  v += dpp0<0x114, 0xf>(v);
  v += dpp0<0x118, 0xf>(v);
  v += dpp0<0x142, 0xa>(v);
  v += dpp0<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ float lane63f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dppf0<0xB1, 0xf>(v);
  v += dppf0<0x4E, 0xf>(v);
  v += dppf0<0x124, 0xf>(v);
  v += dppf0<0x128, 0xf>(v);
  v += dppf0<0x142, 0xa>(v);
  v += dppf0<0x143, 0xc>(v);
  return lane63f(v);
}
__device__ __forceinline__ float wave_min(float v) {
  v = fminf(v, dppf_self<0xB1, 0xf>(v));
  v = fminf(v, dppf_self<0x4E, 0xf>(v));
  v = fminf(v, dppf_self<0x124, 0xf>(v));
  v = fminf(v, dppf_self<0x128, 0xf>(v));
  v = fminf(v, dppf_self<0x142, 0xa>(v));
  v = fminf(v, dppf_self<0x143, 0xc>(v));
  return lane63f(v);
}
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dppf_self<0xB1, 0xf>(v));
  v = fmaxf(v, dppf_self<0x4E, 0xf>(v));
  v = fmaxf(v, dppf_self<0x124, 0xf>(v));
  v = fmaxf(v, dppf_self<0x128, 0xf>(v));
  v = fmaxf(v, dppf_self<0x142, 0xa>(v));
  v = fmaxf(v, dppf_self<0x143, 0xc>(v));
  return lane63f(v);
}

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
// Home slot of neighbour (atom j, packed image img) in a row's reverse-lookup hash table of hmask + 1 slots
__device__ __forceinline__ int rev_slot(int j, unsigned img, int hmask) {
  unsigned h = (unsigned)j * 0x9E3779B1u ^ img * 0x85EBCA77u;
  h ^= h >> 15;
  return (int)(h & (unsigned)hmask);
}
// Insert val (1..65535) at the first empty 16-bit slot from h on (linear probing) in an LDS table of 16-bit slots
// packed two per word.  A compare-and-swap that fails means another lane filled the other half of the word (or this
// slot): the slot is read again.  The table is at most half full, so an empty slot lies within hslots / 2 probes, and
// each failed swap is another lane's insertion: the bound below is never reached.
__device__ __forceinline__ void htab_insert(unsigned* tab, int h, int hmask, unsigned val) {
  for (int it = 0; it < 2 * (hmask + 1) + 2 * WAVE; it++) {
    unsigned* w = tab + (h >> 1);
    const int sh = (h & 1) * 16;
    const unsigned old = __atomic_load_n(w, __ATOMIC_RELAXED);
    if ((old >> sh) & 0xffffu) {
      h = (h + 1) & hmask;
      continue;
    }
    if (atomicCAS(w, old, old | (val << sh)) == old) return;
  }
}

// ---- scans, clears and the capacity check (graph-capturable: kernels only)
#define SCAN_TILE 1024  // values per block: 256 threads x 4

// exclusive scan of one tile; single != 0: the whole array is one tile and out[0] = 0 is written here
__global__ __launch_bounds__(256) void k_scan_tile(const int* __restrict__ in, int* __restrict__ out, int n,
                                                   int* __restrict__ bsum, int single) {
  __shared__ int ts[256];
  const int t = threadIdx.x, base = blockIdx.x * SCAN_TILE + 4 * t;
  int v[4], run = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) { v[q] = base + q < n ? in[base + q] : 0; run += v[q]; }
  ts[t] = run;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // Hillis-Steele inclusive scan of the thread sums
    int a = t >= o ? ts[t - o] : 0;
    __syncthreads();
    ts[t] += a;
    __syncthreads();
  }
  int acc = ts[t] - run;  // exclusive prefix of this thread within the tile
#pragma unroll
  for (int q = 0; q < 4; q++) {
    acc += v[q];
    if (base + q < n) out[base + q + 1] = acc;  // inclusive value at q = exclusive value at q + 1
  }
  if (t == 255 && bsum) bsum[blockIdx.x] = ts[255];
  if (single && blockIdx.x == 0 && t == 0) out[0] = 0;
}

// exclusive scan of the nb tile sums in place (one block, chunks of 256 with a carry); bsum[nb] = total
__global__ __launch_bounds__(256) void k_scan_top(int* __restrict__ bsum, int nb) {
  __shared__ int ts[256];
  __shared__ int carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nb; c0 += 256) {
    const int v = c0 + t < nb ? bsum[c0 + t] : 0;
    ts[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      int a = t >= o ? ts[t - o] : 0;
      __syncthreads();
      ts[t] += a;
      __syncthreads();
    }
    if (c0 + t < nb) bsum[c0 + t] = carry + ts[t] - v;
    __syncthreads();
    if (t == 0) carry += ts[255];
    __syncthreads();
  }
  if (t == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ bsum) {
  const int b = blockIdx.x, add = bsum[b];
  for (int q = threadIdx.x; q < SCAN_TILE; q += 256) {
    const int k = b * SCAN_TILE + q;
    if (k < n) out[k + 1] += add;
  }
  if (b == 0 && threadIdx.x == 0) out[0] = 0;
}

// exclusive scan of up to 1024 * SCAN1_PT values in one block.  Each pass covers 1024 x SCAN1_SEG values: they are
// loaded coalesced into LDS (all loads of a thread issued before any is used: one memory round trip per pass), thread t
// then sums its run of SCAN1_SEG consecutive values from LDS, a shuffle scan of the 1024 run totals (64 per wave, 16
// waves through LDS) gives each run its offset, the prefix sums go back to LDS and out with coalesced stores, and
// the pass total carries into the next pass.  LDS index k + k/32 spreads a run's strided reads over the banks.
#define SCAN1_PT 64
#define SCAN1_SEG 8
__global__ __launch_bounds__(1024) void k_scan_one(const int* __restrict__ in, int* __restrict__ out, int n) {
  __shared__ int buf[1024 * SCAN1_SEG + 1024 * SCAN1_SEG / 32];
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) { carry_s = 0; out[0] = 0; }
  for (int c0 = 0; c0 < n; c0 += 1024 * SCAN1_SEG) {
    int v[SCAN1_SEG];
#pragma unroll
    for (int q = 0; q < SCAN1_SEG; q++) {
      const int k = c0 + q * 1024 + t;
      v[q] = k < n ? in[k] : 0;
    }
#pragma unroll
    for (int q = 0; q < SCAN1_SEG; q++) {
      const int k = q * 1024 + t;
      buf[k + (k >> 5)] = v[q];
    }
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int q = 0; q < SCAN1_SEG; q++) {
      const int k = t * SCAN1_SEG + q;
      run += buf[k + (k >> 5)];
      v[q] = run;  // inclusive within the run
    }
    const int x = wave_scan_incl(run);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int off = carry_s + x - run;
    for (int q = 0; q < w; q++) off += wsum[q];
#pragma unroll
    for (int q = 0; q < SCAN1_SEG; q++) {
      const int k = t * SCAN1_SEG + q;
      buf[k + (k >> 5)] = off + v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SCAN1_SEG; q++) {
      const int k = q * 1024 + t;
      if (c0 + k < n) out[c0 + k + 1] = buf[k + (k >> 5)];
    }
    if (t == 1023) carry_s = off + run;
    __syncthreads();
  }
}

__global__ void k_clear3(int* __restrict__ a, int na, int* __restrict__ b, int nb, int* __restrict__ c, int nc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < na) a[k] = 0;
  if (k < nb) b[k] = 0;
  if (k < nc) c[k] = 0;
}

// misc[0] = longest row of this evaluation; raises the overflow flag misc[3] when the edges or the longest row
// exceed the capacities the edge kernels were sized for, and keeps the largest totals seen in misc[4], misc[5]
__global__ void k_check_caps(int N, const int* __restrict__ row_off, int* __restrict__ misc, int cap_edges,
                             int cap_row) {
  if (threadIdx.x != 0) return;
  const int ne = row_off[N], mr = misc[0];
  misc[6] = ne;
  if (ne > cap_edges || mr > cap_row) misc[3] = 1;
  if (ne > misc[4]) misc[4] = ne;
  if (mr > misc[5]) misc[5] = mr;
}

// ---------------------------------------------------------------------------------------------- prepare
// Wraps positions and counts atoms per bin (bin_count is zero on entry: the bin fill of the previous evaluation
// counted it back down).  Also clears the first nclear flags of misc for this evaluation.
// With skin lists (xb non-null): the image shift m_i that carries the wrapped position at the lists' build, xb_i,
// next to the current one, m_i = rint((xb_i - x_i) . R) per lattice direction, goes to mimg[i] (packed); when the
// atom moved more than sqrt(lim2) from there (minimum image) the lists are marked stale (misc[8] = 0).
__global__ void k_prepare(int N, const double* __restrict__ xin, const int* __restrict__ atom_str,
                          const mg_struct_dev* __restrict__ str, double* __restrict__ x, int* __restrict__ bin,
                          int* __restrict__ bin_count, int* __restrict__ misc, int nclear,
                          const double* __restrict__ xb, unsigned* __restrict__ mimg, double lim2) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nclear) misc[i] = 0;
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
  if (xb) {
    const double d[3] = {xb[3 * i] - X[0], xb[3 * i + 1] - X[1], xb[3 * i + 2] - X[2]};
    int mq[3] = {0, 0, 0};
    if (st->nd > 0)
      for (int q = 0; q < 3; q++) mq[q] = (int)rint(d[0] * st->R[3 * q] + d[1] * st->R[3 * q + 1] + d[2] * st->R[3 * q + 2]);
    double u2 = 0.0;
    for (int q = 0; q < 3; q++) {
      const double u = -d[q] + mq[0] * st->L[q] + mq[1] * st->L[3 + q] + mq[2] * st->L[6 + q];
      u2 += u * u;
    }
    // a shift beyond the packing range cannot be carried: mark the lists stale as for a long move
    const bool far = mq[0] < -256 || mq[0] > 255 || mq[1] < -256 || mq[1] > 255 || mq[2] < -256 || mq[2] > 255;
    if (u2 > lim2 || far) misc[8] = 0;
    mimg[i] = far ? pack_img(0, 0, 0) : pack_img(mq[0], mq[1], mq[2]);
  }
  int id = st->bin0 + (b[0] * st->nb[1] + b[1]) * st->nb[2] + b[2];
  bin[i] = id;
  bin[N + i] = b[0] | (b[1] << 10) | (b[2] << 20);  // the bin's coordinates (nb <= 1024 per direction)
  atomicAdd(&bin_count[id], 1);
}

// Atoms in bin order, each with its wrapped position and species, so a neighbour candidate is one 32-byte load.
struct __align__(16) BinRec {
  double x, y, z;
  int s, j;
};

// Places atoms in their bins, taking slots by counting bin_count back down to zero (ready for the next evaluation).
__global__ void k_bin_fill(int N, const int* __restrict__ bin, const int* __restrict__ bin_start,
                           int* __restrict__ bin_count, const double* __restrict__ x, const int* __restrict__ spc,
                           BinRec* __restrict__ brec) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b = bin[i];
  int p = bin_start[b] + atomicSub(&bin_count[b], 1) - 1;
  BinRec r;
  r.x = x[3 * i]; r.y = x[3 * i + 1]; r.z = x[3 * i + 2];
  r.s = spc[i]; r.j = i;
  brec[p] = r;
}

// a lane's double in every lane (scalar read; lane must be wave-uniform)
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// ---------------------------------------------------------------------------------------------- block order, wave sync
// The 8 XCDs of the chip take workgroups round robin (block b runs on XCD b mod 8), each with its own L2.  Kernels
// that read the rows of an atom's neighbours remap blocks so every XCD works through one contiguous range of
// atoms: neighbouring atoms then share their neighbours' rows in that XCD's L2.
#define MG_NXCD 8
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  const int per = nblk / MG_NXCD, rem = nblk % MG_NXCD;  // XCD x takes per (+1 for x < rem) blocks
  const int x = b % MG_NXCD, k = b / MG_NXCD;
  return x * per + (x < rem ? x : rem) + k;
}


// ---------------------------------------------------------------------------------------------- neighbours
// One wave per atom.  The stencil of (2K0+1)(2K1+1)(2K2+1) bins around the atom's bin (each with the image
// shift its periodic wrap implies) is processed in chunks of 64 entries: lane l builds entry l (bin start,
// atom count, FP64 image shift), a wave prefix scan over the counts flattens the chunk into one candidate
// range, and the lanes then test candidates c = lane, lane + 64, ... (entry found by binary search in LDS).
// The FP64 distance uses LIST's expression order (src/cell.c:115-123), so the accepted set is bit-exact.
// FILL: accepted entries are appended to an LDS row (ballot + lane prefix), the row is sorted by key
// (species, j, image) with a rank sort and written once: canonical order, independent of the binning.
// Rows are (row_off[i], rowcnt[i]).  The count pass (!FILL) writes rowcnt and the per-species counts, and a scan
// then gives row_off, so FILL writes CSR rows.  With padded, FILL runs without the count pass: row i lies at
// i * cap_row and the fill writes row_off[i], rowcnt[i] and the species counts itself (every consumer addresses rows
// through row_off and rowcnt).  A row longer than cap_row raises the overflow flag misc[3] and is not written;
// misc[5] keeps the longest row seen, to size the next capacities.
#define NBR_CHUNK 64
// per-wave LDS of the stencil chunk tables (shifts, bin starts, images, prefix), a multiple of 16 bytes
#define NBR_LDS_CHUNK ((NBR_CHUNK * (3 * 8 + 4 + 4 + 4) + 4 + 15) / 16 * 16)
#define NBR_WPB 4  // most atoms (waves) per block; XCD-aware block order, so neighbouring atoms share bins in one L2
// MODE 0 counts, 1 fills; MODE 2 fills skin candidates: the search at the radius
// Rc + skin with the wider stencil Ks, rows of keys only (species, j, image at this build) at i * cap_row in key[]
// with their lengths in rowcnt[], run only when the lists are stale (misc[8] == 0); its longest row goes to misc[7].
template <int MODE>
__global__ __launch_bounds__(64 * NBR_WPB) void k_nbr(int N, const double* __restrict__ x, const int* __restrict__ spc,
                                            const int* __restrict__ atom_str, const mg_struct_dev* __restrict__ str,
                                            const int* __restrict__ bin, const int* __restrict__ bin_start,
                                            const BinRec* __restrict__ brec, int* __restrict__ cnt,
                                            int* __restrict__ rowcnt, int* __restrict__ misc,
                                            int* __restrict__ row_off, u64* __restrict__ key,
                                            float4* __restrict__ vec, int* __restrict__ rev,
                                            unsigned short* __restrict__ rhash, int hslots, int cap, int padded,
                                            int cap_row, int wpb, double skin) {
#pragma clang fp contract(off)
  constexpr bool FILL = MODE != 0;
  constexpr bool CAND = MODE == 2;
  if (FILL && !padded && misc[3]) return;  // rows exceed the capacities this launch was sized for
  if (CAND && misc[8]) return;             // the skin lists are current
  extern __shared__ __align__(16) unsigned char smem_n[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = xcd_block(blockIdx.x, gridDim.x) * wpb + wv;
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  // the wave's LDS: the stencil chunk tables (NBR_LDS_CHUNK bytes), then with FILL the row's keys, vectors and hash
  unsigned char* wl = smem_n + (size_t)wv * (NBR_LDS_CHUNK + 24 * (size_t)cap + 2 * (size_t)hslots);
  double (*e_sh)[3] = (double (*)[3])wl;                          // [NBR_CHUNK] image shifts
  int* e_start = (int*)(wl + NBR_CHUNK * 3 * sizeof(double));    // [NBR_CHUNK] first record of each stencil bin
  unsigned* e_img = (unsigned*)(e_start + NBR_CHUNK);             // [NBR_CHUNK] packed images
  int* e_pre = (int*)(e_img + NBR_CHUNK);                         // [NBR_CHUNK + 1] candidate prefix
  u64* rk = (u64*)(wl + NBR_LDS_CHUNK);                           // [cap] keys of the row (FILL)
  float4* rv = (float4*)(rk + cap);                                // [cap] vectors
  unsigned* htab = (unsigned*)(rv + cap);                          // [hslots / 2] words: 16-bit hash slots (FILL)
  const mg_struct_dev* st = &str[atom_str[i]];
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  double RC = st->rc2;
  if (CAND) {
    const double rs = sqrt(st->rc2) + skin;
    RC = rs * rs * (1.0 + 1e-12);
  }
  const int nb0 = st->nb[0], nb1 = st->nb[1], nb2 = st->nb[2];
  const int K0 = CAND ? st->Ks[0] : st->K[0], K1 = CAND ? st->Ks[1] : st->K[1], K2 = CAND ? st->Ks[2] : st->K[2];
  const int w1n = 2 * K1 + 1, w2n = 2 * K2 + 1;
  const int nsten = (2 * K0 + 1) * w1n * w2n;
  if (nsten >= (1 << 21)) {  // a cell thinner than rc/64: beyond the index arithmetic below; no row, error raised
    if (lane == 0) {
      misc[1] = 1;
      if (CAND) {
        rowcnt[i] = 0;
      } else if (!FILL || padded) {
        cnt[4 * i] = cnt[4 * i + 1] = cnt[4 * i + 2] = cnt[4 * i + 3] = 0;
        rowcnt[i] = 0;
        if (FILL) row_off[i] = i * cap_row;
      }
    }
    return;
  }
  const float rw1 = 1.0f / (float)w1n, rw2 = 1.0f / (float)w2n;
  float rfn0, rfn1, rfn2;  // 1.0f / nb_q
  const int bq = bin[N + i];
  const int b0 = bq & 1023, b1 = (bq >> 10) & 1023, b2 = bq >> 20;
  const unsigned img0 = pack_img(0, 0, 0);
  // stencil pruning: a lower bound on the distance from atom i to every point of a stencil bin.  Along lattice
  // direction q the bin is the slab of fractional coordinates [(b_q + d_q)/nb_q, (b_q + d_q + 1)/nb_q); its gap to
  // the atom's fractional coordinate times the plane spacing 1/|R_q| bounds the distance.  The largest of the three
  // bounds holds for any cell; for mutually orthogonal lattice rows the slabs are orthogonal and the bounds add in
  // quadrature.  Bins whose bound exceeds rc (with a relative margin far above rounding) hold no neighbour.
  // per-direction constants computed once, lane q for direction q, then broadcast: plane spacing 1/|R_q|, 1/nb_q,
  // and the float reciprocals for the stencil index arithmetic
  double fi[3], sp[3], rnb[3];
  {
    double spl = 0.0, rnl = 0.0;
    float rfl = 0.0f;
    if (lane < 3) {
      const double* Rq = st->R + 3 * lane;
      spl = 1.0 / sqrt(Rq[0] * Rq[0] + Rq[1] * Rq[1] + Rq[2] * Rq[2]);
      const int nq = lane == 0 ? nb0 : (lane == 1 ? nb1 : nb2);
      rnl = 1.0 / (double)nq;
      rfl = 1.0f / (float)nq;
    }
    for (int q = 0; q < 3; q++) {
      fi[q] = xi0 * st->R[3 * q] + xi1 * st->R[3 * q + 1] + xi2 * st->R[3 * q + 2];
      sp[q] = readlane_d(spl, q);
      rnb[q] = readlane_d(rnl, q);
    }
    rfn0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, rfl), 0));
    rfn1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, rfl), 1));
    rfn2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, rfl), 2));
  }
  bool orth = true;
  for (int q = 0; q < 3; q++) {
    const double* u = st->L + 3 * q;
    const double* v = st->L + 3 * ((q + 1) % 3);
    const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (uv * uv > 1e-24 * uu * vv) orth = false;
  }
  const double rc2m = RC * (1.0 + 1e-9);
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0;  // per-species counts (wave-uniform)
  int n = 0;                           // accepted so far (wave-uniform)
  for (int base = 0; base < nsten; base += NBR_CHUNK) {
    const int t = base + lane;
    int count = 0;
    if (t < nsten) {
      // t = (d0 * w1n + d1) * w2n + d2 (offsets by K): quotients through float reciprocals, exact for t < 2^21
      const int t2 = floordiv_f(t, rw2), t1 = floordiv_f(t2, rw1);
      const int d2 = t - t2 * w2n - K2, d1 = t2 - t1 * w1n - K1, d0 = t1 - K0;
      const int dd[3] = {d0, d1, d2}, bb[3] = {b0, b1, b2};
      double lb[3];
      for (int q = 0; q < 3; q++) {
        // (b + d) / nb by the reciprocal: within an ulp of the quotient, far inside the bound's margin
        double g = 0.0;
        if (dd[q] > 0) g = (double)(bb[q] + dd[q]) * rnb[q] - fi[q];
        else if (dd[q] < 0) g = fi[q] - (double)(bb[q] + dd[q] + 1) * rnb[q];
        lb[q] = g > 0.0 ? g * sp[q] : 0.0;
      }
      const double lb2 = orth ? lb[0] * lb[0] + lb[1] * lb[1] + lb[2] * lb[2]
                              : fmax(lb[0] * lb[0], fmax(lb[1] * lb[1], lb[2] * lb[2]));
      const bool far = lb2 > rc2m;
      int w0 = b0 + d0, ic0 = floordiv_f(w0, rfn0);
      w0 -= ic0 * nb0;
      int w1 = b1 + d1, ic1 = floordiv_f(w1, rfn1);
      w1 -= ic1 * nb1;
      int w2 = b2 + d2, ic2 = floordiv_f(w2, rfn2);
      w2 -= ic2 * nb2;
      const int w = st->bin0 + (w0 * nb1 + w1) * nb2 + w2;
      e_start[lane] = bin_start[w];
      count = far ? 0 : bin_start[w + 1] - bin_start[w];
      e_img[lane] = pack_img(ic0, ic1, ic2);
      for (int q = 0; q < 3; q++) {
        double v = 0.0;
        v += (double)ic0 * st->L[q];
        v += (double)ic1 * st->L[3 + q];
        v += (double)ic2 * st->L[6 + q];
        e_sh[lane][q] = v;
      }
    }
    const int inc = wave_scan_incl(count);  // inclusive scan of the counts across the wave
    e_pre[lane + 1] = inc;
    if (lane == 0) e_pre[0] = 0;
    const int total = __builtin_amdgcn_readlane(inc, WAVE - 1);
    wave_sync();
    for (int cb = 0; cb < total; cb += WAVE) {
      const int c = cb + lane;
      bool ok = false;
      int j = 0, sj = 0;
      unsigned img = 0;
      double a0 = 0, a1 = 0, a2 = 0, r = 0;
      if (c < total) {
        int lo = 0, hi = NBR_CHUNK - 1;  // last entry e with e_pre[e] <= c
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (e_pre[mid] <= c) lo = mid; else hi = mid - 1;
        }
        const BinRec br = brec[e_start[lo] + (c - e_pre[lo])];
        j = br.j;
        img = e_img[lo];
        if (!(j == i && img == img0)) {
          a0 = e_sh[lo][0]; a0 += br.x - xi0;
          a1 = e_sh[lo][1]; a1 += br.y - xi1;
          a2 = e_sh[lo][2]; a2 += br.z - xi2;
          r = 0.0;
          r += a0 * a0;
          r += a1 * a1;
          r += a2 * a2;
          ok = r < RC;
          sj = br.s;
        }
      }
      const u64 bm = __ballot(ok);
      if (FILL) {
        const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
        if (ok && n + below < cap) {
          rk[n + below] = make_key(sj, j, img);
          if (!CAND) rv[n + below] = make_float4((float)a0, (float)a1, (float)a2, (float)sqrt(r));
        }
      }
      if (!CAND && (!FILL || padded)) {
        c0 += __popcll(__ballot(ok && sj == 0));
        c1 += __popcll(__ballot(ok && sj == 1));
        c2 += __popcll(__ballot(ok && sj == 2));
        c3 += __popcll(__ballot(ok && sj == 3));
      }
      n += __popcll(bm);
    }
    wave_sync();
  }
  if (CAND) {
    if (lane == 0) {
      rowcnt[i] = n < cap_row ? n : cap_row;
      if (n > __atomic_load_n(misc + 7, __ATOMIC_RELAXED)) atomicMax(misc + 7, n);
      if (n > cap_row) misc[3] = 1;
    }
    if (n > cap_row) return;
  } else if (!FILL || padded) {
    if (lane == 0) {
      cnt[4 * i] = c0; cnt[4 * i + 1] = c1; cnt[4 * i + 2] = c2; cnt[4 * i + 3] = c3;
      rowcnt[i] = n;
      // a plain read first: almost every row is shorter than the running maximum, so few atomics contend
      if (n > __atomic_load_n(misc, __ATOMIC_RELAXED)) atomicMax(misc, n);
      if (FILL) {
        row_off[i] = i * cap_row;
        if (n > __atomic_load_n(misc + 5, __ATOMIC_RELAXED)) atomicMax(misc + 5, n);
        if (n > cap_row) misc[3] = 1;
      }
    }
    if (!FILL || n > cap_row) return;  // count pass done, or overflow: nothing of this row is written
  }
  const int M = n < cap ? n : cap;  // n <= cap by construction (cap >= longest row)
  wave_sync();
  if (CAND) {  // candidate rows: the keys in sorted order, nothing else
    const int off = i * cap_row;
    for (int t0 = 0; t0 < M; t0 += WAVE) {
      const int ta = t0 + lane;
      const u64 ka = ta < M ? rk[ta] : ~0ull;
      int ra = 0;
#pragma unroll 4
      for (int u = 0; u < M; u++) ra += rk[u] < ka ? 1 : 0;
      if (ta < M) key[off + ra] = ka;
    }
    return;
  }
  // rank sort: the keys of a row are distinct, so the number of smaller keys is an entry's place in the sorted row;
  // every lane counts against broadcast reads of the whole row and writes its entries to their places
  const int off = padded ? i * cap_row : row_off[i];
  // the row's reverse-lookup table: (neighbour atom, image) -> sorted position + 1, open addressing with linear
  // probing in hslots >= 2 * cap slots (at most half full, so every probe sequence ends at an empty slot).  Only the
  // entries another row will look up go in: those whose reverse edge owns the pair, i.e. j < i, or j == i with
  // the larger packed image.
  const int hmask = hslots - 1;
  for (int q = lane; q < hslots / 2; q += WAVE) htab[q] = 0u;
  wave_sync();
  for (int t0 = 0; t0 < M; t0 += 2 * WAVE) {
    const int ta = t0 + lane, tb = t0 + WAVE + lane;
    const u64 ka = ta < M ? rk[ta] : ~0ull, kb = tb < M ? rk[tb] : ~0ull;
    int ra = 0, rb = 0;
#pragma unroll 4
    for (int u = 0; u < M; u++) {
      const u64 ku = rk[u];
      ra += ku < ka ? 1 : 0;
      rb += ku < kb ? 1 : 0;
    }
    if (ta < M) {
      key[off + ra] = ka; vec[off + ra] = rv[ta]; rev[off + ta] = -1;
      const int ja = key_j(ka);
      const unsigned ia = (unsigned)(ka & 0x3fffffffull);
      if (ja < i || (ja == i && neg_img(ia) < ia)) htab_insert(htab, rev_slot(ja, ia, hmask), hmask, (unsigned)(ra + 1));
    }
    if (tb < M) {
      key[off + rb] = kb; vec[off + rb] = rv[tb]; rev[off + tb] = -1;
      const int jb = key_j(kb);
      const unsigned ib = (unsigned)(kb & 0x3fffffffull);
      if (jb < i || (jb == i && neg_img(ib) < ib)) htab_insert(htab, rev_slot(jb, ib, hmask), hmask, (unsigned)(rb + 1));
    }
  }
  wave_sync();
  unsigned* ht = (unsigned*)(rhash + (size_t)i * hslots);  // hslots is even: the row's table is word aligned
  for (int q = lane; q < hslots / 2; q += WAVE) ht[q] = htab[q];
}

// Skin lists, every step: row i of the padded CSR (what k_nbr<1> writes with padded = 1: keys, vectors, rev = -1,
// species counts, row lengths and offsets, the reverse-lookup hash table, misc[0]/misc[5]/overflow) from the row's
// skin candidates alone.  A candidate (species, j, image c at the lists' build) lies at image c + m_j - m_i now
// (m from k_prepare); its vector and distance come from the same FP64 expressions as the fill's (shift
// sum_q c_q L_q accumulated in that order, then + (x_j - x_i); r = a0^2 + a1^2 + a2^2 accumulated in that order),
// so the accepted set and the vectors are bit-identical to a fresh fill.  The candidate row is sorted by key and
// every candidate of one (species, j) moves by the same image shift, which keeps the lexicographic order of the
// packed images: the accepted entries come out in the fill's sorted order without a sort.  Marks the lists
// current (misc[8] = 1) for the next step: the stale mark of this step was consumed by k_nbr<2> and k_skin_commit,
// which ran before this kernel.
__global__ __launch_bounds__(64 * NBR_WPB) void k_nbr_skin(int N, const double* __restrict__ x,
                                                  const int* __restrict__ atom_str,
                                                  const mg_struct_dev* __restrict__ str,
                                                  const u64* __restrict__ ckey, const int* __restrict__ crow,
                                                  int ccap, const unsigned* __restrict__ mimg,
                                                  int* __restrict__ cnt, int* __restrict__ rowcnt,
                                                  int* __restrict__ misc, int* __restrict__ row_off,
                                                  u64* __restrict__ key, float4* __restrict__ vec,
                                                  int* __restrict__ rev, unsigned short* __restrict__ rhash,
                                                  int hslots, int cap, int cap_row, int wpb) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char smem_s[];
  if (blockIdx.x == 0 && threadIdx.x == 0) misc[8] = 1;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = xcd_block(blockIdx.x, gridDim.x) * wpb + wv;
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  unsigned char* wl = smem_s + (size_t)wv * (24 * (size_t)cap + 2 * (size_t)hslots);
  u64* rk = (u64*)wl;                        // [cap] keys of the row
  float4* rv = (float4*)(rk + cap);          // [cap] vectors
  unsigned* htab = (unsigned*)(rv + cap);    // [hslots / 2] words of 16-bit hash slots
  const mg_struct_dev* st = &str[atom_str[i]];
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  const double RC = st->rc2;
  const unsigned mi = mimg[i];
  const int mi0 = (int)(mi & 1023u) - MG_IMG_OFF, mi1 = (int)((mi >> 10) & 1023u) - MG_IMG_OFF,
            mi2 = (int)((mi >> 20) & 1023u) - MG_IMG_OFF;
  const int nc = crow[i];
  const u64* ck = ckey + (size_t)i * ccap;
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  int n = 0;
  // SKIN_ILP chunks of 64 candidates at a time: their key, image-shift and position loads are issued together
  // (one round of memory latency per group instead of one per chunk); the chunks are then tested and appended in
  // candidate order
  constexpr int SKIN_ILP = 4;
  for (int cb = 0; cb < nc; cb += SKIN_ILP * WAVE) {
    u64 kq[SKIN_ILP];
    unsigned mq[SKIN_ILP];
    double xq[SKIN_ILP][3];
#pragma unroll
    for (int q = 0; q < SKIN_ILP; q++) {
      const int c = cb + q * WAVE + lane;
      kq[q] = c < nc ? ck[c] : 0ull;
    }
#pragma unroll
    for (int q = 0; q < SKIN_ILP; q++) {
      const int c = cb + q * WAVE + lane;
      const int j = key_j(kq[q]);
      mq[q] = c < nc ? mimg[j] : 0u;
      xq[q][0] = c < nc ? x[3 * j] : 0.0;
      xq[q][1] = c < nc ? x[3 * j + 1] : 0.0;
      xq[q][2] = c < nc ? x[3 * j + 2] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < SKIN_ILP; q++) {
      const int c = cb + q * WAVE + lane;
      if (cb + q * WAVE >= nc) break;  // wave-uniform
      bool ok = false;
      int j = 0, sj = 0;
      unsigned img = 0;
      double a0 = 0, a1 = 0, a2 = 0, r = 0;
      if (c < nc) {
        const u64 k = kq[q];
        sj = (int)(k >> 60);
        j = key_j(k);
        // image now = image at the build + m_j - m_i (packed fields carry MG_IMG_OFF each)
        const unsigned ib = (unsigned)(k & 0x3fffffffull), mj = mq[q];
        const int ic0 = (int)(ib & 1023u) - MG_IMG_OFF + ((int)(mj & 1023u) - MG_IMG_OFF) - mi0;
        const int ic1 = (int)((ib >> 10) & 1023u) - MG_IMG_OFF + ((int)((mj >> 10) & 1023u) - MG_IMG_OFF) - mi1;
        const int ic2 = (int)((ib >> 20) & 1023u) - MG_IMG_OFF + ((int)((mj >> 20) & 1023u) - MG_IMG_OFF) - mi2;
        img = pack_img(ic0, ic1, ic2);
        double sh[3];
        for (int d = 0; d < 3; d++) {
          double v = 0.0;
          v += (double)ic0 * st->L[d];
          v += (double)ic1 * st->L[3 + d];
          v += (double)ic2 * st->L[6 + d];
          sh[d] = v;
        }
        a0 = sh[0]; a0 += xq[q][0] - xi0;
        a1 = sh[1]; a1 += xq[q][1] - xi1;
        a2 = sh[2]; a2 += xq[q][2] - xi2;
        r = 0.0;
        r += a0 * a0;
        r += a1 * a1;
        r += a2 * a2;
        ok = r < RC;
      }
      const u64 bm = __ballot(ok);
      const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
      if (ok && n + below < cap) {
        rk[n + below] = make_key(sj, j, img);
        rv[n + below] = make_float4((float)a0, (float)a1, (float)a2, (float)sqrt(r));
      }
      c0 += __popcll(__ballot(ok && sj == 0));
      c1 += __popcll(__ballot(ok && sj == 1));
      c2 += __popcll(__ballot(ok && sj == 2));
      c3 += __popcll(__ballot(ok && sj == 3));
      n += __popcll(bm);
    }
  }
  if (lane == 0) {
    cnt[4 * i] = c0; cnt[4 * i + 1] = c1; cnt[4 * i + 2] = c2; cnt[4 * i + 3] = c3;
    rowcnt[i] = n;
    if (n > __atomic_load_n(misc, __ATOMIC_RELAXED)) atomicMax(misc, n);
    row_off[i] = i * cap_row;
    if (n > __atomic_load_n(misc + 5, __ATOMIC_RELAXED)) atomicMax(misc + 5, n);
    if (n > cap_row) misc[3] = 1;
  }
  if (n > cap_row) return;  // overflow: nothing of this row is written
  const int M = n < cap ? n : cap;
  const int off = i * cap_row;
  const int hmask = hslots - 1;
  for (int q = lane; q < hslots / 2; q += WAVE) htab[q] = 0u;
  wave_sync();
  for (int ta = lane; ta < M; ta += WAVE) {
    const u64 ka = rk[ta];
    key[off + ta] = ka; vec[off + ta] = rv[ta]; rev[off + ta] = -1;
    const int ja = key_j(ka);
    const unsigned ia = (unsigned)(ka & 0x3fffffffull);
    if (ja < i || (ja == i && neg_img(ia) < ia)) htab_insert(htab, rev_slot(ja, ia, hmask), hmask, (unsigned)(ta + 1));
  }
  wave_sync();
  unsigned* ht = (unsigned*)(rhash + (size_t)i * hslots);
  for (int q = lane; q < hslots / 2; q += WAVE) ht[q] = htab[q];
}

// After a skin-candidate build (lists stale, misc[8] == 0): the build positions are the current wrapped ones, and
// every image shift is zero; misc[16] counts the builds over the context's life.
__global__ void k_skin_commit(int N, const double* __restrict__ x, double* __restrict__ xb, unsigned* __restrict__ mimg,
                              int* __restrict__ misc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || misc[8]) return;
  xb[3 * i] = x[3 * i]; xb[3 * i + 1] = x[3 * i + 1]; xb[3 * i + 2] = x[3 * i + 2];
  mimg[i] = pack_img(0, 0, 0);
  if (i == 0) misc[16] += 1;
}

// Reverse edge of e = (i -> j, img): the entry (j -> i, -img) of row j.  Edges come in mutual pairs, rev[e] = e' and
// rev[e'] = e, so only the edge that owns its pair (j > i, or j == i with the smaller packed image) looks the other
// up (in the descriptor kernel) and writes both.  The lookup probes row j's hash table (built with the row by the
// fill): a slot value v > 0 names position v - 1, which is checked against the key (one table load and one key load
// per probe, about 1.3 probes at the table's load of at most one half, against log2(row) + 1 dependent key loads for a
// search of the sorted row).  Every load stays inside row j's table and row (positions above rowcnt[j] are
// rejected); an empty slot, or no match, is a missing partner (error flag, rev = self).
__device__ __forceinline__ int rev_find(const unsigned short* __restrict__ tab, const u64* __restrict__ key, int base,
                                        int n, u64 want, int h, int hmask) {
  for (int pr = 0; pr <= hmask; pr++) {
    const int v = tab[h];
    if (v == 0 || v > n) return -1;
    if (key[base + v - 1] == want) return v - 1;
    h = (h + 1) & hmask;
  }
  return -1;
}

// ---------------------------------------------------------------------------------------------- wave helpers

// Reduce-scatter of 64 per-lane accumulators: afterwards lane l holds the wave sum of accumulator l.  Each stage k
// halves the accumulators: a lane keeps the half its bit k selects and adds the partner's (lane ^ k) copy of it.  The
// exchanges stay in the VALU: stages 32 and 16 by the lane-half swaps (v_permlane32/16_swap: a swap of the two
// halves of a pair (v[p], v[p + A/2]) leaves v[p] of both lanes in one register and v[p + A/2] of both in the other,
// so one add finishes a pair), stage 8 by a row rotation, stage 4 by row shifts either way under bank masks, stages 2
// and 1 by quad permutes.  Every move runs with all lanes active (a move never reads a lane a branch has switched off).  The sums are those of the xor exchange, term for term.
// the swap in inline assembly: this compiler's builtin for it returns the first result twice; two wait states
// between the instruction that writes an operand and the swap
__device__ __forceinline__ float fpair_sum32(float a, float b) {
  unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(ua), "+v"(ub));
  return __builtin_bit_cast(float, ua) + __builtin_bit_cast(float, ub);
}
__device__ __forceinline__ float fpair_sum16(float a, float b) {
  unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(ua), "+v"(ub));
  return __builtin_bit_cast(float, ua) + __builtin_bit_cast(float, ub);
}
template <int A, int K>
__device__ __forceinline__ void rs_step(float* v, int lane) {
  const bool hi = (lane & K) != 0;
#pragma unroll
  for (int p = 0; p < A / 2; p++) {
    if (K == 32) {
      v[p] = fpair_sum32(v[p], v[p + A / 2]);
    } else if (K == 16) {
      v[p] = fpair_sum16(v[p], v[p + A / 2]);
    } else {
      const float send = hi ? v[p] : v[p + A / 2];
      const float keep = hi ? v[p + A / 2] : v[p];
      float recv;
      if (K == 8) {
        recv = dppf0<0x128, 0xf>(send);  // row_ror:8 = lane ^ 8
      } else if (K == 4) {
        // lane ^ 4 by two moves with bank masks (banks = 4-lane groups of a row), no select: row_shl:4 (lane + 4)
        // into banks 0 and 2, then row_shr:4 (lane - 4) into banks 1 and 3; both with every lane active
        const int sb = __builtin_bit_cast(int, send);
        const int up = __builtin_amdgcn_update_dpp(0, sb, 0x104, 0xf, 0x5, false);
        recv = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(up, sb, 0x114, 0xf, 0xa, false));
      } else if (K == 2) {
        recv = dppf0<0x4E, 0xf>(send);   // quad_perm [2,3,0,1]
      } else {
        recv = dppf0<0xB1, 0xf>(send);   // quad_perm [1,0,3,2]
      }
      v[p] = keep + recv;
    }
  }
}
__device__ __forceinline__ float reduce_scatter64(float* v, int lane) {
  rs_step<64, 32>(v, lane);
  rs_step<32, 16>(v, lane);
  rs_step<16, 8>(v, lane);
  rs_step<8, 4>(v, lane);
  rs_step<4, 2>(v, lane);
  rs_step<2, 1>(v, lane);
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
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// a = hi + lo in FP16 for two values at once (packed RNE conversions)
__device__ __forceinline__ void split_f16x2(float a, float b, unsigned& hi, unsigned& lo) {
  const f16x2 h = __builtin_convertvector((f32x2){a, b}, f16x2);
  const f32x2 hf = __builtin_convertvector(h, f32x2);
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a - hf[0], b - hf[1]}, f16x2));
}
// the same for one value in the low half, the high half zero (an odd count of etas)
__device__ __forceinline__ void split_f16x1(float a, unsigned& hi, unsigned& lo) {
  const _Float16 h = (_Float16)a;
  hi = (unsigned)__builtin_bit_cast(unsigned short, h);
  lo = (unsigned)__builtin_bit_cast(unsigned short, (_Float16)(a - (float)h));
}
// a = hi + lo for two values at once: hi rounds a to BF16 (one packed conversion), lo rounds the exact residual;
// the BF16 halves of hi are widened back by shifts
__device__ __forceinline__ void split_bf16x2(float a, float b, unsigned& hi, unsigned& lo) {
  const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
  const float ha = __builtin_bit_cast(float, h << 16), hb = __builtin_bit_cast(float, h & 0xffff0000u);
  hi = h;
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a - ha, b - hb}, bf16x2));
}
__device__ __forceinline__ void split_bf16x1(float a, unsigned& hi, unsigned& lo) {
  const __bf16 h = (__bf16)a;
  hi = (unsigned)__builtin_bit_cast(unsigned short, h);
  lo = (unsigned)__builtin_bit_cast(unsigned short, (__bf16)(a - (float)h));
}
// cos(pi * t) and sin(pi * t) for t in [0, 1]: the hardware takes revolutions, cos(2 pi y)
__device__ __forceinline__ float cospi_(float t) { return __builtin_amdgcn_cosf(0.5f * t); }
__device__ __forceinline__ float sinpi_(float t) { return __builtin_amdgcn_sinf(0.5f * t); }


struct RowInfo {
  int off, M, si;
  int m0, m1, m2, m3;  // neighbours of each species (rows are sorted by species)
  __device__ __forceinline__ int m(int s) const { return s == 0 ? m0 : (s == 1 ? m1 : (s == 2 ? m2 : m3)); }
  __device__ __forceinline__ int o(int s) const {
    return s == 0 ? 0 : (s == 1 ? m0 : (s == 2 ? m0 + m1 : m0 + m1 + m2));
  }
};

__device__ __forceinline__ void row_info(RowInfo& ri, int i, const int* __restrict__ spc, const int* __restrict__ row_off,
                                         const int* __restrict__ rowcnt, const int* __restrict__ cnt, int S) {
  ri.off = row_off[i];
  ri.M = rowcnt[i];
  ri.si = spc[i];
  ri.m0 = cnt[4 * i];
  ri.m1 = S > 1 ? cnt[4 * i + 1] : 0;
  ri.m2 = S > 2 ? cnt[4 * i + 2] : 0;
  ri.m3 = S > 3 ? cnt[4 * i + 3] : 0;
}

// Stage the row of atom i into LDS.  Record per neighbour: (dx, dy, dz, r), (1/r, fc, fc'/r, r^2) and, unless
// every species pair shares the G4 etas (UNI), A[0..7] = exp(-eta4(i,j)_e r^2).  With shared etas a triplet's
// radial factor is one exp(-eta_e (r_ij^2 + r_ik^2 + r_jk^2)) and the A tables are not needed.  The records lie
// in planes of cap float4 (record q of slot t at rec[q cap + t]), so consecutive slots are consecutive 16-byte words
// and a wave's reads of consecutive slots (the pairs of one diagonal) meet no bank conflict.  DESC: the descriptor
// kernel's form (dx, dy, dz, r^2), (dx/r, dy/r, dz/r, fc): a triplet's cosine is one dot product of unit vectors
// (that kernel reads r itself from the edge vectors).
template <int NE, bool UNI, bool DESC = false>
__device__ __forceinline__ void stage_row(float4* rec, const RowInfo& ri, const float4* __restrict__ vec,
                                          const mg_basis_dev* __restrict__ bp, int lane, int S, int cap) {
  const int pl = cap;  // plane stride
  for (int t = lane; t < ri.M; t += WAVE) {
    float4 v = vec[ri.off + t];
    int s = 0;
#pragma unroll
    for (int q = 1; q < MG_MAXSPC; q++) s += (q < S && t >= ri.o(q)) ? 1 : 0;
    int p = ri.si * MG_MAXSPC + s;
    float rc = bp->rc[p];
    float tt = v.w * __builtin_amdgcn_rcpf(rc);  // r / rc within 1.5 ulp: no IEEE division sequence per neighbour
    float fc = v.w < rc ? 0.5f * (cospi_(tt) + 1.0f) : 0.0f;
    float r2 = v.w * v.w;
    const float ir = __builtin_amdgcn_rcpf(v.w);
    if (DESC) {
      rec[t] = make_float4(v.x, v.y, v.z, r2);
      rec[pl + t] = make_float4(v.x * ir, v.y * ir, v.z * ir, fc);
    } else {
      float dfc = v.w < rc ? -0.5f * 3.14159265358979f / rc * sinpi_(tt) : 0.0f;
      rec[t] = v;
      rec[pl + t] = make_float4(ir, fc, dfc * ir, r2);
    }
    if (!UNI) {
      float A[8];
#pragma unroll
      for (int e = 0; e < 8; e++) A[e] = e < NE ? fast_exp2(-(float)MG_LOG2E * bp->eta4[p][e] * r2) : 0.0f;
      rec[2 * pl + t] = make_float4(A[0], A[1], A[2], A[3]);
      rec[3 * pl + t] = make_float4(A[4], A[5], A[6], A[7]);
    }
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
// (1 + lambda c)^zeta and its derivative in c, cell q = z*2 + l.  The prefactors 2^(1 - zeta) are applied once
// outside the pair loops (to the reduced descriptors, and folded into the dE/dG grid in the force kernel).
// HALF: the same on w = u / 2, i.e. P = w^zeta = 2^(1 - zeta) P / 2 and P' = 2^(1 - zeta) P' / 2, the zeta prefactor
// taken into the powers (one scale for every cell)
template <int PAT, bool HALF = false>
__device__ __forceinline__ void angular(const mg_basis_dev* __restrict__ bp, float cs, float* P, float* Pd) {
#pragma unroll
  for (int l = 0; l < MG_MAXL; l++) {
    // the two shipped tables have lambda = {-1, +1} (checked when the model is loaded): constants here
    const float lam = Pat<PAT>::ZSTD ? (l == 0 ? -1.0f : 1.0f) : bp->lam[l];
    const float u = HALF ? __builtin_fmaf(0.5f * lam, cs, 0.5f) : 1.0f + lam * cs;
    const float dk = HALF ? 0.5f : 1.0f;
    if (Pat<PAT>::ZSTD) {
      const float u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u16 = u8 * u8;
      P[0 + l] = u;
      P[2 + l] = u2;
      P[4 + l] = u4;
      P[6 + l] = u16;
      if (Pd) {
        Pd[0 + l] = dk * lam;
        Pd[2 + l] = (dk * 2.0f * lam) * u;
        Pd[4 + l] = (dk * 4.0f * lam) * u2 * u;
        Pd[6 + l] = (dk * 16.0f * lam) * (u8 * u4) * (u2 * u);
      }
    } else {
#pragma unroll
      for (int z = 0; z < MG_MAXZ; z++) {
        int zt = bp->zeta[z];
        float pm1 = ipowf(u, zt - 1);
        P[z * MG_MAXL + l] = pm1 * u;
        if (Pd) Pd[z * MG_MAXL + l] = dk * (float)zt * lam * pm1;
      }
    }
  }
}

// Unordered neighbour pairs {j, k} of one species class with r_jk < rc.  Enumeration: triangle (same
// species) in circulant order (j, j + d mod m), d = 1 .. m/2, each pair once; rectangle (a x b) with the longer
// side as the lane-fast index.  Passing pairs are compacted into an LDS list (ballot + lane prefix count) and
// processed with all 64 lanes busy; about half of the enumerated pairs fail the r_jk test in bulk metals.
#define PRING 128  // per-wave ring of surviving pairs (a power of two >= 2 * WAVE)
// Pair words: j | k << 11 | tag << 22 (slots < 2048, tag < 1024).  tag numbers the group of the pair, a set of pairs in
// which no slot repeats on either side, so the force kernel can accumulate one group at a time without atomics.
// The fields are taken out with v_bfe_u32 written out: the compiler turns (p & 0x7ff) * 32 + base into a shift, a
// mask and an add, where a field extract and one shift-add suffice.
template <int OFF, int W>
__device__ __forceinline__ unsigned ubfe_v(unsigned x) {
  unsigned r;
  asm("v_bfe_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "i"(OFF), "i"(W));
  return r;
}
#define PK_J(p) ((int)ubfe_v<0, 11>(p))
#define PK_K(p) ((int)ubfe_v<11, 11>(p))
#define PK_TAG(p) ((int)((p) >> 22))
// j and k renamed between the reads of two record planes, so that each plane's address is one shift-add of the slot
// (shared, the shifted slot costs a shift and one add per plane)
__device__ __forceinline__ void fresh2(int& j, int& k) { asm("" : "+v"(j), "+v"(k)); }
// lane / n for 0 <= lane < 64, 0 < n <= 2048: lane + 0.5 is exact, and the hardware reciprocal (1 ulp) and the
// product move (lane + 0.5) / n by less than 64 * 1.5 * 2^-23 = 1.2e-5, under the 0.5 / n >= 2.4e-4 that separates
// it from an integer
__device__ __forceinline__ int lane_div(int lane, int n) {
  return (int)(((float)lane + 0.5f) * __builtin_amdgcn_rcpf((float)n));
}

// Enumerate the neighbour pairs (j, k) of one species class with r_jk < rc, compacting the survivors into a ring in
// LDS and handing them to body(j, k, tag, valid) in batches of 64 as soon as 64 are waiting (valid = false on the
// lanes past the end of the class's last batch), so every batch but the class's last is full.  The pairs are laid out
// on circulant diagonals, each a conflict-free group (tag = the diagonal's q, q + 1 for one species):
//   same species (m slots): diagonal q = 0..m/2-1 holds the pairs (f, f + q + 1 mod m), f = 0..m-1 (the last
//     diagonal of an even m only f < m/2), so every unordered pair appears once;
//   two species (long side a >= short side b): diagonal q = 0..a-1 holds (u, u + q mod a) for the short side's slots
//     u = 0..b-1 (a <= 1024 is required: a longer side raises the error flag, bit 1).
// Candidate t = q * n + f (n = m, or b) is taken 64 at a time in order, so a batch of 64 survivors spans only a few
// diagonals; each lane steps its own (q, f) by 64 candidates with one add and one wrap (no division per candidate),
// and the wraps are unsigned minima (x mod n = min(x, x - n) for x < 2n).  Only the class's last step checks the
// candidate range.  With gout, the survivors are also written there in order (the pair list of the class, which the
// force kernel reads instead of enumerating again); returns their number.  The batches are the consecutive runs of 64
// of that list, the same batches the stored-list reader forms.  R: float4 from one slot's position record to the
// next (1 for the records' planes).
template <int MODE>
struct PairMode : std::integral_constant<int, MODE> {};

template <int R, typename F>
__device__ __forceinline__ int for_pairs(const float4* rec, unsigned* ring, bool tri, int oa, int ma, int ob, int mb,
                                         float rc2, int lane, unsigned* __restrict__ gout, int* __restrict__ err,
                                         F&& body) {
  int wp = 0, rp = 0;  // ring write / read positions: survivors found so far, survivors handed to body
  // mode 0: one species; 1: two species, long side a; 2: two species, long side b
  auto run = [&](auto MODEC, int n, int nl, int ol, int os) {
    constexpr int mode = decltype(MODEC)::value;
    // n: slots per diagonal; nl: the wrap of the partner slot (m, or the long side); ol, os: block offsets of the
    // partner (long) side and of the lane's (short) side
    const int ntot = mode == 0 ? ((n >> 1) - 1) * n + ((n & 1) ? n : n >> 1) : nl * n;
    // wave-uniform step of (q, f) per 64 candidates (df pinned in a scalar register: left as 64 - dq n, the compiler
    // spends two vector instructions per step on f + df)
    const int dq = WAVE / n;
    int df = WAVE - dq * n;
    asm("" : "+s"(df));
    // lane state: f, the short-side slot, and qq = q (+ 1 for one species), the partner's offset along the diagonal;
    // the partner slot is p = f + qq mod nl, and qq is also the pair's tag (a diagonal number, distinct per diagonal)
    int qq = lane_div(lane, n);
    unsigned f = (unsigned)(lane - qq * n);
    if (mode == 0) qq += 1;
    const float4* recf = rec + R * os;  // the short side's records
    const float4* recp = rec + R * ol;  // the partner side's records
    // the pair word j | k << 11 | tag << 22 with block slots: j is the short side's slot except in mode 1
    const unsigned pkc = mode == 1 ? (unsigned)ol | ((unsigned)os << 11) : (unsigned)os | ((unsigned)ol << 11);
    // one step of 64 candidates; GUARD: the class's last, partial step, where lanes past the last candidate skip
    // the reads (elsewhere f < n and f + qq < 2 nl hold on every lane, so both slots lie in the class's blocks)
    auto step = [&](auto GUARD, int base) {
      constexpr bool guard = decltype(GUARD)::value;
      const unsigned t = f + (unsigned)qq;
      const unsigned p = min(t, t - (unsigned)nl);
      bool ok = false;
      if (!guard || lane < ntot - base) {
        const float4 a = recf[R * f], b = recp[R * p];
        const float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        ok = cx * cx + cy * cy + cz * cz < rc2;
      }
      const unsigned pk = (mode == 1 ? (p | (f << 11)) : (f | (p << 11))) + (((unsigned)qq << 22) + pkc);
      const unsigned f1 = f + (unsigned)df;
      const bool c = f1 >= (unsigned)n;
      f = c ? f1 - (unsigned)n : f1;
      qq += dq + (c ? 1 : 0);
      const unsigned long long bm = __ballot(ok);
      // wp + the survivors on lower lanes
      const unsigned at = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, (unsigned)wp));
      if (ok) ring[at & (PRING - 1)] = pk;
      wp += __popcll(bm);
      const bool last = guard || base + WAVE == ntot;
      if (wp - rp >= WAVE || last && wp > rp) {
        // at most 127 survivors wait here (fewer than 64 before this step's 64), so the ring never overwrites an
        // unread entry; the second sync orders the batch's ring reads before the next step's writes
        wave_sync();
        while (wp - rp >= WAVE || (last && wp > rp)) {
          const unsigned tt = (unsigned)(rp + lane);
          const bool valid = tt < (unsigned)wp;
          // read on every lane: past wp the ring holds older pairs of this row (or stale words), whose slots the
          // body masks; only the valid lanes write
          const unsigned pw = ring[tt & (PRING - 1)];
          // a 32-bit byte offset from the row's base (rows hold far fewer than 2^30 pairs): a scalar-based store
          if (gout && valid) *(unsigned*)((char*)gout + (tt << 2)) = pw;
          body(PK_J(pw), PK_K(pw), PK_TAG(pw), valid);
          rp += WAVE;
        }
        wave_sync();
      }
    };
    int base = 0;
    for (; base + WAVE <= ntot; base += WAVE) step(std::false_type{}, base);
    if (base < ntot) step(std::true_type{}, base);
  };
  if (tri) {
    if (ma < 2) return 0;
    run(PairMode<0>{}, ma, ma, oa, oa);
    return wp;
  }
  if (ma == 0 || mb == 0) return 0;
  if ((ma >= mb ? ma : mb) > 1024) {  // tags would not fit 10 bits
    if (lane == 0) atomicOr(err, 2);
    return 0;
  }
  if (ma >= mb) run(PairMode<1>{}, mb, ma, oa, ob);  // lanes over species b's slots, partners in a
  else run(PairMode<2>{}, ma, mb, ob, oa);            // lanes over species a's slots, partners in b
  return wp;
}

// The stored pair list of one class (np pairs from pl), handed to body in batches of 64 like for_pairs; the next
// batch's words are loaded before the current batch runs.  The loads are unconditional (index clamped to the last
// pair), so the only wait for one is where its batch starts.
template <typename F>
__device__ __forceinline__ void for_stored_pairs(const unsigned* __restrict__ pl, int np, int lane, F&& body) {
  if (np <= 0) return;
  // a lane's pair as a 32-bit byte offset from the class's (scalar) list base, clamped to the last pair
  const char* base = (const char*)pl;
  const unsigned endb = 4u * (unsigned)np, lastb = endb - 4u;
  unsigned ob = 4u * (unsigned)lane;
  unsigned pn = *(const unsigned*)(base + (ob < lastb ? ob : lastb));
  for (int tb = 0; tb < np; tb += WAVE) {
    const unsigned p = pn;
    const bool valid = ob < endb;
    ob += 4u * WAVE;
    pn = *(const unsigned*)(base + (ob < lastb ? ob : lastb));
    body(PK_J(p), PK_K(p), PK_TAG(p), valid);
  }
}

// fj[j] += gj, fk[k] += gk over one batch of pairs: the pairs of one tag group touch distinct slots on each side, and
// the two sides accumulate into separate arrays, so each group is one plain read-modify-write of both; the groups of
// a batch (a few) go one after the other, ordered by the LDS executing one wave's accesses in program order.
__device__ __forceinline__ void group_accum(float4* fj, float4* fk, int j, float jx, float jy, float jz, int k, float kx,
                                            float ky, float kz, int tag, bool valid) {
  unsigned long long rem = __ballot(valid);
  while (rem) {
    const int cur = __builtin_amdgcn_readlane(tag, (int)__builtin_ctzll(rem));  // scalar read, no LDS round trip
    const bool same = tag == cur;
    const unsigned long long hit = __ballot(same);  // rem holds valid lanes only, so this clears what is done
    if (valid && same) {
      float4 v = fj[j], w = fk[k];
      v.x += jx; v.y += jy; v.z += jz;
      w.x += kx; w.y += ky; w.z += kz;
      fj[j] = v;
      fk[k] = w;
    }
    rem &= ~hit;
  }
}

// ---------------------------------------------------------------------------------------------- descriptors
template <int PAT, bool UNI>
__global__ __launch_bounds__(256) void k_desc(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
                                                   const int* __restrict__ rowcnt, const int* __restrict__ cnt, const float4* __restrict__ vec,
                                                   const mg_basis_dev* __restrict__ bp, float* __restrict__ G,
                                                   float* __restrict__ erep, float urep, int cap,
                                                   unsigned* __restrict__ gpl, int* __restrict__ npl, size_t pstride,
                                                   const u64* __restrict__ key, const unsigned short* __restrict__ rhash,
                                                   int hslots, int* __restrict__ rev, int* __restrict__ err,
                                                   const int* __restrict__ ovf) {
  constexpr int NE = Pat<PAT>::NE;
  constexpr int R = UNI ? 2 : 4;
  extern __shared__ float4 smem_d[];
  if (*ovf) return;  // rows may exceed the LDS tile of cap entries
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: per-atom data in SGPRs
  const int i = blockIdx.x * (blockDim.x >> 6) + wv;  // one wave per atom, 1-4 waves per block
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const int D = bp->D;
  RowInfo ri;
  row_info(ri, i, spc, row_off, rowcnt, cnt, S);
  float4* rec = smem_d + (size_t)wv * (R * cap + PRING / 4);
  unsigned* plist = (unsigned*)(rec + R * cap);
  stage_row<NE, UNI, true>(rec, ri, vec, bp, lane, S, cap);
  const float4* rec1 = rec + cap;  // (d / r, fc) plane
  const float4* rec2 = rec + 2 * cap;  // A tables (without UNI)
  const float4* rec3 = rec + 3 * cap;
  // reverse edges (for the force kernel): the edges of this row that own their pair look the partner up in the
  // partner row's hash table and write both rev entries; the probes' load latency hides under the other waves'
  // descriptor arithmetic
  if (rev) {
    const int hmask = hslots - 1;
    for (int t = lane; t < ri.M; t += WAVE) {
      const int e = ri.off + t;
      const u64 k = key[e];
      const int j = key_j(k);
      const unsigned im = (unsigned)(k & 0x3fffffffull);
      if (j > i || (j == i && im < neg_img(im))) {
        const unsigned nim = neg_img(im);
        const int base = row_off[j], n = rowcnt[j];
        const int p = rev_find(rhash + (size_t)j * hslots, key, base, n, make_key(ri.si, i, nim),
                               rev_slot(i, nim, hmask), hmask);
        if (p >= 0) { rev[e] = base + p; rev[base + p] = e; }
        else { atomicOr(err, 1); rev[e] = e; }
      }
    }
  }
  wave_sync();
  float* Gi = G + (size_t)i * D;
  const float* vr = (const float*)(vec + ri.off) + 3;  // vr[4 t] = r of neighbour t (the G2 sums read it here)
  // The shipped G4 tables leave the grid cells of the etas past NE free (8 cells for the 43-function table, 16 for
  // the 22-function one): there the G2 sums of neighbour species s ride along in the reduce-scatter of class (s, s)
  // instead of one wave reduction per G2 function.
  constexpr int G2OFF = NE * MG_MAXZ * MG_MAXL;
  const bool g2grid = PAT != 0 && bp->ng2 <= MG_GRID - G2OFF;
  float eu = 0.0f;  // UREP energy of the atom (CELL_UREP, src/cmod.c:70-86)

  // ---- G2 (+ UREP energy) per neighbour species block (generic grid)
  for (int s = 0; s < (g2grid ? 0 : S); s++) {
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
      const float r = vr[4 * t];
      const float fc = rec1[t].w;
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

  // ---- G4 per species class (sa <= sb); with gpl, the surviving pairs of each class are stored for the force
  // kernel (class after class from gpl + i * pstride; npl holds each class's count)
  unsigned* gout = gpl ? gpl + (size_t)i * pstride : nullptr;
  int cls = 0;
  for (int sa = 0; sa < S; sa++)
    for (int sb = sa; sb < S; sb++, cls++) {
      const int ds1 = (sa - ri.si + S) % S, ds2 = (sb - ri.si + S) % S;
      const int blk = bp->ind4[ds1][ds2];
      const int pjk = sa * MG_MAXSPC + sb;
      const float rc = bp->rc[pjk];
      const float rinv = __builtin_amdgcn_rcpf(rc), rinvh = 0.5f * rinv;  // hardware reciprocal (1 ulp)
      float ejk[NE];
#pragma unroll
      for (int e = 0; e < NE; e++) ejk[e] = -(float)MG_LOG2E * bp->eta4[pjk][e];
      float V[MG_GRID];
#pragma unroll
      for (int q = 0; q < MG_GRID; q++) V[q] = 0.0f;
      const bool g2here = g2grid && sa == sb;
      if (g2here) {
        // G2 of neighbour species sa (and its UREP energy) into the free cells G2OFF + c
        const int p = ri.si * MG_MAXSPC + sa;
        const float Rrep = bp->urep_rm[ri.si] + bp->urep_rm[sa];
        for (int t = ri.o(sa) + lane; t < ri.o(sa) + ri.m(sa); t += WAVE) {
          const float r = vr[4 * t];
          const float fc = rec1[t].w;
#pragma unroll
          for (int c = 0; c < MG_GRID - G2OFF; c++)
            if (c < MG_MAXG2 && c < bp->ng2) {
              const float dr = r - bp->rs2[p][c];
              V[G2OFF + c] += fc * fast_exp2(-(float)MG_LOG2E * bp->eta2[p][c] * dr * dr);
            }
          if (urep > 0.0f && r < Rrep) eu += urep * 0.5f * (1.0f + cospi_(r / Rrep)) * Rrep / r;
        }
      }
      auto body = [&](int j, int k, int, bool valid) {
        if (!valid) return;
        float4 a = rec[j], b = rec[k];  // (d, r^2)
        fresh2(j, k);
        float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        float r3s = cx * cx + cy * cy + cz * cz;
        float4 ua = rec1[j], ub = rec1[k];  // (d / r, fc)
        float r3 = __builtin_amdgcn_sqrtf(r3s);  // hardware square root (1 ulp): no IEEE fix-up sequence
        // fc(r_jk), clamped to [0, 1] by the instruction (the hardware cosine may step past -1 by an ulp)
        float fc3 = __builtin_amdgcn_fmed3f(__builtin_fmaf(__builtin_amdgcn_cosf(r3 * rinvh), 0.5f, 0.5f), 0.0f, 1.0f);
        float cs = ua.x * ub.x + ua.y * ub.y + ua.z * ub.z;
        float F = ua.w * ub.w * fc3;  // fc fc fc; the factor 2 of the G4 definition goes on with the zeta prefactors
        float P[MG_MAXZ * MG_MAXL];
        angular<PAT>(bp, cs, P, nullptr);
        float Bq[NE];
        if (UNI) {
          // F exp2(c rs) = exp2(c rs + log2 F): one logarithm in place of NE products (F = 0 gives -inf, then 0)
          const float rs = a.w + b.w + r3s;
          const float lF = __builtin_amdgcn_logf(F);
#pragma unroll
          for (int e = 0; e < NE; e++) Bq[e] = fast_exp2(__builtin_fmaf(ejk[e], rs, lF));
        } else {
          float4 Aj0 = rec2[j], Aj1 = rec3[j], Ak0 = rec2[k], Ak1 = rec3[k];
          float Aj[8] = {Aj0.x, Aj0.y, Aj0.z, Aj0.w, Aj1.x, Aj1.y, Aj1.z, Aj1.w};
          float Ak[8] = {Ak0.x, Ak0.y, Ak0.z, Ak0.w, Ak1.x, Ak1.y, Ak1.z, Ak1.w};
#pragma unroll
          for (int e = 0; e < NE; e++) Bq[e] = F * Aj[e] * Ak[e] * fast_exp2(ejk[e] * r3s);
        }
#pragma unroll
        for (int e = 0; e < NE; e++)
#pragma unroll
          for (int q = 0; q < MG_MAXZ * MG_MAXL; q++)
            if (Pat<PAT>::mask(e) & (1u << q)) V[e * MG_MAXZ * MG_MAXL + q] += Bq[e] * P[q];
      };
      const int np = for_pairs<1>(rec, plist, sa == sb, ri.o(sa), ri.m(sa), ri.o(sb), ri.m(sb), rc * rc, lane, gout, err, body);
      if (gout) {
        gout += np;
        if (lane == 0) npl[(size_t)i * MG_NCLS + cls] = np;
      }
      // lane l: grid cell l = (e*MAXZ + z)*MAXL + lambda; the zeta prefactor 2^(1 - zeta) goes on here
      const float raw = reduce_scatter64(V, lane);
      float mine = raw * (2.0f * bp->pref[(lane / MG_MAXL) % MG_MAXZ]);
      int cell = lane < bp->ng4 ? bp->g4cell[lane] : 0;
      float v = __shfl(mine, cell, WAVE);
      if (lane < bp->ng4) Gi[blk + lane] = v;
      if (g2here && lane >= G2OFF && lane - G2OFF < bp->ng2) Gi[bp->ind2[ds1] + lane - G2OFF] = raw;
    }
  if (erep) {
    // no UREP term (urep = 0, the usual case) leaves eu = 0 on every lane: no reduction needed
    const float v = urep > 0.0f ? wave_sum(eu) : 0.0f;
    if (lane == 0) erep[i] = v;
  }
}

// ---------------------------------------------------------------------------------------------- MLP (FP32 VALU)

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
//   T = F * sum_e S_e B_e, F = f1 f2 f3, B_e = exp(-(e_ij r1^2 + e_ik r2^2 + e_jk r3^2)),
//   S_e = sum_(z,l) g(e,z,l) P_zl(cos), S'_e with P'_zl (g carries 2^(1 - zeta) and the factor 2 of G4),
//   dT/dr1 = -2 r1 F sum_e e_ij S_e B_e + f1' f2 f3 sum_e S_e B_e (r2, r3 likewise), dT/dcos = F sum_e S'_e B_e,
//   dT/da = dT/dr1 a/r1 - dT/dr3 c/r3 + dT/dcos (b/(r1 r2) - cos a/r1^2), dT/db symmetric.
// Forces: F_j -= dT/da, F_k -= dT/db, F_i += dT/da + dT/db; virial U -= dT/da (x) a + dT/db (x) b.
// The dE/dG grid g(e,z,l) of the current species block is wave-uniform; it is staged through LDS once per class
// and kept in registers for the class's triplets.
template <int PAT, bool UNI, int MMA, bool ST>
__device__ __forceinline__ void force_row(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
                                                    const int* __restrict__ rowcnt, const int* __restrict__ cnt, const float4* __restrict__ vec,
                                                    const mg_basis_dev* __restrict__ bp, const float* __restrict__ dEdG,
                                                    float4* __restrict__ fself, float4* __restrict__ fe,
                                                    float* __restrict__ vir, float urep, int cap,
                                                    const unsigned* __restrict__ gpl, const int* __restrict__ npl,
                                                    size_t pstride, const int* __restrict__ rev, int* __restrict__ err,
                                                    const int* __restrict__ ovf) {
  constexpr int NE = Pat<PAT>::NE;
  constexpr int R = UNI ? 2 : 4;
  constexpr int ZL = MG_MAXZ * MG_MAXL;
  constexpr bool MM = UNI && MMA != 0;  // contraction on the matrix cores (one exp per eta needs shared etas)
  constexpr bool MH = UNI && MMA == 2;  // FP16 hi/lo operands (FP32 precision) instead of BF16
  extern __shared__ float4 smem_f[];
  if (*ovf) return;  // rows may exceed the LDS tile of cap entries
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: per-atom data in SGPRs
  const int i = blockIdx.x * (blockDim.x >> 6) + wv;  // one wave per atom, 1-4 waves per block
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const int D = bp->D;
  RowInfo ri;
  row_info(ri, i, spc, row_off, rowcnt, cnt, S);
  // the reverse slots of the first two batches of the row, loaded now so the final scatter does not wait on them
  const int rq0 = lane < ri.M ? rev[ri.off + lane] : -1;
  const int rq1 = lane + WAVE < ri.M ? rev[ri.off + lane + WAVE] : -1;
  float4* rec = smem_f + (size_t)wv * ((R + 2) * cap + MG_GRID / 4 + PRING / 4);
  float4* gts4 = rec + R * cap;                  // [MG_GRID / 4]
  float4* fnb4 = gts4 + MG_GRID / 4;             // [cap] force on each neighbour slot (pair terms, triplet j side)
  float4* fnk4 = fnb4 + cap;                     // [cap] triplet k side
  float* fnb = (float*)fnb4;
  unsigned* plist = (unsigned*)(fnk4 + cap);     // [PRING]
  stage_row<NE, UNI>(rec, ri, vec, bp, lane, S, cap);
  const float4* rec1 = rec + cap;  // (1/r, fc, fc'/r, r^2) plane
  const float4* rec2 = rec + 2 * cap;  // A tables (without UNI)
  const float4* rec3 = rec + 3 * cap;
  for (int t = lane; t < ri.M; t += WAVE) {
    fnb4[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    fnk4[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  wave_sync();
  // MH: lb = 2 min_j r_ij^2, a lower bound of r_ij^2 + r_ik^2 + r_jk^2 over the row's triplets (wave-uniform)
  float rlb = 0.0f;
  if (MH) {
    float mn = 3.0e38f;
    for (int t = lane; t < ri.M; t += WAVE) mn = fminf(mn, rec1[t].w);
    mn = wave_min(mn);
    rlb = ri.M >= 2 ? 2.0f * mn : 0.0f;
  }
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
      float4 v = rec[t];
      float4 w = rec1[t];
      // sc = dT/dr / r from fc/r and fc'/r
      float r = v.w, fcr = w.y * w.x, dfr = w.z;
      float sc = 0.0f;
#pragma unroll
      for (int c = 0; c < MG_MAXG2; c++)
        if (c < bp->ng2) {
          float eta = bp->eta2[p][c];
          float dr = r - bp->rs2[p][c];
          float ex = fast_exp2(-(float)MG_LOG2E * eta * dr * dr);
          sc += gi[blk + c] * ex * (-2.0f * eta * dr * fcr + dfr);
        }
      fnb[4 * t] -= sc * v.x; fnb[4 * t + 1] -= sc * v.y; fnb[4 * t + 2] -= sc * v.z;
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

  // ---- triplet terms (G4) per species class; the pairs come from the descriptor kernel's stored list when there
  // is one, else they are enumerated here as there
  const unsigned* pl = gpl ? gpl + (size_t)i * pstride : nullptr;
  int cls = 0;
  for (int sa = 0; sa < S; sa++)
    for (int sb = sa; sb < S; sb++, cls++) {
      const int ds1 = (sa - ri.si + S) % S, ds2 = (sb - ri.si + S) % S;
      const int blk = bp->ind4[ds1][ds2];
      const int pjk = sa * MG_MAXSPC + sb, pij = ri.si * MG_MAXSPC + sa, pik = ri.si * MG_MAXSPC + sb;
      const float rc = bp->rc[pjk];
      const float rinv = __builtin_amdgcn_rcpf(rc), rinvh = 0.5f * rinv;  // hardware reciprocal (1 ulp)
      const float dfk = -0.5f * 3.14159265358979f * rinv;
      // the class's etas (wave-uniform)
      float e3[NE], n3[NE], e1[NE], e2[NE];
#pragma unroll
      for (int e = 0; e < NE; e++) {
        e3[e] = bp->eta4[pjk][e];
        n3[e] = -(float)MG_LOG2E * e3[e];
        e1[e] = UNI ? e3[e] : bp->eta4[pij][e];
        e2[e] = UNI ? e3[e] : bp->eta4[pik][e];
      }
      {
        // grid cell lane, with the zeta prefactor 2^(1 - zeta) and the factor 2 of the G4 definition folded in
        int n = bp->g4map[lane];
        ((float*)gts4)[lane] = n >= 0 ? gi[blk + n] * (2.0f * bp->pref[(lane / MG_MAXL) % MG_MAXZ]) : 0.0f;
      }
      wave_sync();
      // MM (shared etas): the contraction over eta runs on the matrix cores, beside the VALU work of the other waves
      // (MH: split FP16 operands, the FP32 precision; otherwise split BF16, the BF16 precisions; per-pair etas keep
      // the VALU contraction below).  Per batch of 64 triplets, one v_mfma_f32_32x32x16 (three with the split
      // operands) computes, for each triplet t,
      //   C_q(t) = sum_e g_eq B_e(t)  and  D_q(t) = -2 sum_e eta_e g_eq B_e(t),   B_e = exp(-eta_e (r1^2 + r2^2 + r3^2)),
      // so that X = sum_e B_e S_e = sum_q C_q P_q, -2 Y = -2 sum_e eta_e B_e S_e = sum_q D_q P_q and
      // Z = sum_e B_e S'_e = sum_q C_q P'_q.  B operand: lane t gives its B_0..7 as k = 8(t/32) + e of column t%32.
      // A operand: block diagonal, so that the rows a lane receives (8(r/4) + 4(t/32) + r%4, register r) depend on
      // its own k half only; register r of every lane then holds output r (C_0..7, D_0..7) of its own triplet.
      // Both operands are split a = hi + lo and the three products lo.hi + hi.lo + hi.hi accumulate in FP32: with
      // BF16 halves the relative error is ~2^-17 of sum_e |g_eq B_e| (forces measured within ~3e-5 eV/A of FP64),
      // with FP16 halves ~2^-22 (forces at the FP32 VALU contraction's level, a few 1e-6 eV/A).
      bf16x8 ahi, alo;
      f16x8 hhi, hlo;
      float oscale = 1.0f;
      if (MH) {
        // FP16 form: A = 2 g / pref (the zeta prefactor goes into the powers of w = u/2, so every cell's
        // coefficient keeps its own size).  B_e = exp(-eta_e rsum) is taken as exp(-eta_e (rsum - lb)) 2^14 with
        // lb = 2 min_j r_ij^2 <= rsum, so every B lies in (0, 2^14] and the triplets that dominate a row's sum
        // (rsum near lb) stay far above the FP16 subnormals whatever the eta; the factor exp(-eta_e lb) moves
        // into A_e.  A is then scaled by one power of two to below 2^14, so the hi/lo split of both operands stays
        // clear of FP16 overflow, and only terms smaller than 2^-28 of the largest lose their low half.  Unused
        // zeta slots (pref = 0) have no coefficient.
        const int m = lane & 31, h = lane >> 5;
        const int r = 4 * (m >> 3) + (m & 3);
        const int qq = r < ZL ? r : r - ZL;
        const bool own = ((m >> 2) & 1) == h;
        const float* gts = (const float*)gts4;
        const float pz = bp->pref[(qq / MG_MAXL) % MG_MAXZ];
        const float ip = pz > 0.0f ? 2.0f * __builtin_amdgcn_rcpf(pz) : 0.0f;  // pz = 2^(1 - zeta): exact for integer zeta
        float a[8], amax = 0.f;
#pragma unroll
        for (int v = 0; v < 8; v++) {
          a[v] = 0.f;
          if (own && v < NE)
            a[v] = (r < ZL ? gts[v * ZL + qq] : -2.0f * e3[v < NE ? v : 0] * gts[v * ZL + qq]) * ip *
                   fast_exp2(n3[v < NE ? v : 0] * rlb);
          amax = fmaxf(amax, fabsf(a[v]));
        }
        amax = wave_max(amax);
        int ex = 0;
        if (amax > 0.f) (void)frexpf(amax, &ex);
        oscale = ldexpf(1.0f, ex - 28);
        unsigned hh[4], hl[4];
#pragma unroll
        for (int v = 0; v < 4; v++) split_f16x2(ldexpf(a[2 * v], 14 - ex), ldexpf(a[2 * v + 1], 14 - ex), hh[v], hl[v]);
        hhi = __builtin_bit_cast(f16x8, (u32x4){hh[0], hh[1], hh[2], hh[3]});
        hlo = __builtin_bit_cast(f16x8, (u32x4){hl[0], hl[1], hl[2], hl[3]});
      }
      if (MM && !MH) {
        const int m = lane & 31, h = lane >> 5;
        const int r = 4 * (m >> 3) + (m & 3);
        const bool own = ((m >> 2) & 1) == h;
        const float* gts = (const float*)gts4;
#pragma unroll
        for (int v = 0; v < 8; v++) {
          float a = 0.f;
          if (own && v < NE) a = r < ZL ? gts[v * ZL + r] : -2.0f * e3[v < NE ? v : 0] * gts[v * ZL + r - ZL];
          ahi[v] = (__bf16)a;
          alo[v] = (__bf16)(a - (float)ahi[v]);
        }
      }
      // the matrix-core sums come out scaled by 1/oscale; every term of dT is linear in f3 or d3, so the scale is
      // taken there (a power of two: exact) instead of on X, Y and Z
      const float h3 = 0.5f * oscale, dfo = dfk * oscale;
      // the grid of this species block, loop-invariant: kept in registers for the whole class (VALU contraction)
      float gv[NE * ZL];
#pragma unroll
      for (int e = 0; e < NE; e++) {
        const float4 g0 = (Pat<PAT>::mask(e) & 0x0Fu) ? gts4[e * 2] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 g1 = (Pat<PAT>::mask(e) & 0xF0u) ? gts4[e * 2 + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        gv[e * ZL + 0] = g0.x; gv[e * ZL + 1] = g0.y; gv[e * ZL + 2] = g0.z; gv[e * ZL + 3] = g0.w;
        gv[e * ZL + 4] = g1.x; gv[e * ZL + 5] = g1.y; gv[e * ZL + 6] = g1.z; gv[e * ZL + 7] = g1.w;
      }
      auto body = [&](int j, int k, int tag, bool valid) {
        // a harmless slot pair for the lanes past the batch (their gradient is discarded below); the stored list's
        // reader already gives them the row's last pair
        if (!ST && !valid) { j = 0; k = 1; }
        float4 a = rec[j], b = rec[k];
        fresh2(j, k);
        float cx = b.x - a.x, cy = b.y - a.y, cz = b.z - a.z;
        float r3s = cx * cx + cy * cy + cz * cz;
        float4 ia = rec1[j], ib = rec1[k];
        float r3 = __builtin_amdgcn_sqrtf(r3s);  // hardware square root (1 ulp): no IEEE fix-up sequence
        float i3 = __builtin_amdgcn_rcpf(r3);
        const float th = r3 * rinvh;  // cos(pi r / rc) = cos(2 pi th) in the hardware's revolutions
        float f3 = __builtin_fmaf(__builtin_amdgcn_cosf(th), h3, h3);
        float d3 = dfo * __builtin_amdgcn_sinf(th);
        float i1 = ia.x, i2 = ib.x, f1 = ia.y, f2 = ib.y, d1 = ia.z, d2 = ib.z;
        float cs = (a.x * b.x + a.y * b.y + a.z * b.z) * i1 * i2;
        float P[ZL], Pd[ZL];
        angular<PAT, MH>(bp, cs, P, Pd);
        float X = 0.f, Y1 = 0.f, Y2 = 0.f, Y3 = 0.f, Zc = 0.f;
        float Aj[8], Ak[8];
        if (!UNI) {
          float4 Aj0 = rec2[j], Aj1 = rec3[j], Ak0 = rec2[k], Ak1 = rec3[k];
          Aj[0] = Aj0.x; Aj[1] = Aj0.y; Aj[2] = Aj0.z; Aj[3] = Aj0.w; Aj[4] = Aj1.x; Aj[5] = Aj1.y; Aj[6] = Aj1.z; Aj[7] = Aj1.w;
          Ak[0] = Ak0.x; Ak[1] = Ak0.y; Ak[2] = Ak0.z; Ak[3] = Ak0.w; Ak[4] = Ak1.x; Ak[5] = Ak1.y; Ak[6] = Ak1.z; Ak[7] = Ak1.w;
        }
        const float rsum = ia.w + ib.w + r3s;
        if (MM) {
          float B[8];
#pragma unroll
          for (int e = 0; e < 8; e++)
            B[e] = e < NE ? fast_exp2(MH ? __builtin_fmaf(n3[e < NE ? e : 0], rsum - rlb, 14.0f)
                                         : n3[e < NE ? e : 0] * rsum)
                          : 0.f;
          unsigned h[4], l[4];
          f32x16 acc = {};
          if (MH) {
#pragma unroll
            for (int v = 0; v < 4; v++)
              if (2 * v + 1 < NE) split_f16x2(B[2 * v], B[2 * v + 1], h[v], l[v]);
              else if (2 * v < NE) split_f16x1(B[2 * v], h[v], l[v]);
              else h[v] = l[v] = 0u;
            const f16x8 bh = __builtin_bit_cast(f16x8, (u32x4){h[0], h[1], h[2], h[3]});
            const f16x8 bl = __builtin_bit_cast(f16x8, (u32x4){l[0], l[1], l[2], l[3]});
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(hlo, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(hhi, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(hhi, bh, acc, 0, 0, 0);
          } else {
#pragma unroll
            for (int v = 0; v < 4; v++)
              if (2 * v + 1 < NE) split_bf16x2(B[2 * v], B[2 * v + 1], h[v], l[v]);
              else if (2 * v < NE) split_bf16x1(B[2 * v], h[v], l[v]);
              else h[v] = l[v] = 0u;
            const bf16x8 bh = __builtin_bit_cast(bf16x8, (u32x4){h[0], h[1], h[2], h[3]});
            const bf16x8 bl = __builtin_bit_cast(bf16x8, (u32x4){l[0], l[1], l[2], l[3]});
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bh, acc, 0, 0, 0);
          }
#pragma unroll
          for (int q = 0; q < ZL; q++) {
            X += acc[q] * P[q];
            Y1 += acc[ZL + q] * P[q];
            if (!Pat<PAT>::ZSTD) Zc += acc[q] * Pd[q];
          }
          if (Pat<PAT>::ZSTD) {
            // zeta = 1, 2, 4, 16 on u_l (cell 2z + l): sum_z C_zl dP_zl/dc = dk lam_l (C_0l + 2u (C_1l + 2u^2 (C_2l
            // + 4 u^12 C_3l))), with dk = 1/2 for the halved powers; lam = -1, +1
            float zl[2];
#pragma unroll
            for (int l = 0; l < 2; l++) {
              const float u = P[l], u2 = P[2 + l], u12 = P[4 + l] * (P[4 + l] * P[4 + l]);
              const float s2 = __builtin_fmaf(acc[6 + l] * u12, 4.0f, acc[4 + l]);
              const float s1 = __builtin_fmaf(u2 * s2, 2.0f, acc[2 + l]);
              zl[l] = __builtin_fmaf(u * s1, 2.0f, acc[l]);
            }
            Zc = (MH ? 0.5f : 1.0f) * (zl[1] - zl[0]);
          }
        }
#pragma unroll
        for (int e = 0; e < (MM ? 0 : NE); e++) {
          const float eta3 = e3[e];
          float B = UNI ? fast_exp2(n3[e] * rsum) : Aj[e] * Ak[e] * fast_exp2(n3[e] * r3s);
          float Se = 0.f, Sd = 0.f;
#pragma unroll
          for (int q = 0; q < ZL; q++)
            if (Pat<PAT>::mask(e) & (1u << q)) { Se += gv[e * ZL + q] * P[q]; Sd += gv[e * ZL + q] * Pd[q]; }
          float SB = Se * B;
          X += SB;
          if (UNI) {
            Y1 += eta3 * SB;
          } else {
            Y1 += e1[e] * SB;
            Y2 += e2[e] * SB;
            Y3 += eta3 * SB;
          }
          Zc += Sd * B;
        }
        if (UNI) { Y2 = Y1; Y3 = Y1; }
        // dT/dr_n / r_n = -2 Fp Y_n + (fc_n'/r_n) (the other two cutoffs) X: the r_n of d(r_n^2)/dr_n cancels the
        // 1/r_n of the direction (d1, d2 are fc'/r from the records; d3 / r3 here)
        const float f12 = f1 * f2, Fp = f12 * f3;
        constexpr float m2 = MM ? 1.0f : -2.0f;  // the matrix-core D sums carry the -2 already
        const float FY1 = m2 * Fp * Y1;
        const float FY2 = UNI ? FY1 : m2 * Fp * Y2, FY3 = UNI ? FY1 : m2 * Fp * Y3;
        const float c1 = __builtin_fmaf(d1, f2 * f3 * X, FY1);
        const float c2 = __builtin_fmaf(d2, f1 * f3 * X, FY2);
        const float c3c = __builtin_fmaf(d3 * i3, f12 * X, FY3);  // coefficient of c
        float dTc = Fp * Zc, dcs = dTc * cs;
        float ca = c1 - dcs * (i1 * i1);   // coefficient of a in dT/da
        float cb = dTc * (i1 * i2);        // coefficient of b in dT/da (and of a in dT/db)
        float cbb = c2 - dcs * (i2 * i2);  // coefficient of b in dT/db
        float gax = ca * a.x + cb * b.x - c3c * cx, gay = ca * a.y + cb * b.y - c3c * cy, gaz = ca * a.z + cb * b.z - c3c * cz;
        float gbx = cbb * b.x + cb * a.x + c3c * cx, gby = cbb * b.y + cb * a.y + c3c * cy, gbz = cbb * b.z + cb * a.z + c3c * cz;
        group_accum(fnb4, fnk4, j, -gax, -gay, -gaz, k, -gbx, -gby, -gbz, tag, valid);
      };
      // ST: the pairs come from the stored list (the usual case; compiled apart from the enumeration, whose live
      // state would otherwise share the register budget of the triplet arithmetic)
      if (ST) {
        const int np = npl[(size_t)i * MG_NCLS + cls];
        for_stored_pairs(pl, np, lane, body);
        pl += np;
      } else {
        for_pairs<1>(rec, plist, sa == sb, ri.o(sa), ri.m(sa), ri.o(sb), ri.m(sb), rc * rc, lane, nullptr, err, body);
      }
      wave_sync();
    }
  // every term of T_i depends on the neighbour vectors d_e only, and fnb[e] = -dT_i/dd_e summed over them: so
  // F_i = -sum_e fnb[e] and the virial U = -sum_e dT/dd_e (x) d_e = sum_e fnb[e] (x) d_e (symmetric).  The force on
  // neighbour j of edge e = (i -> j) goes to the slot of the reverse edge rev[e] = (j -> i) in row j, so each atom's
  // received forces lie in its own row (a missing reverse edge raises the error flag)
  wave_sync();
  for (int t = lane; t < ri.M; t += WAVE) {
    const float4 g1 = fnb4[t], g2 = fnk4[t];
    const float gx = g1.x + g2.x, gy = g1.y + g2.y, gz = g1.z + g2.z;
    const float4 v = rec[t];
    fx -= gx; fy -= gy; fz -= gz;
    u0 += gx * v.x; u1 += gy * v.y; u2 += gz * v.z;
    u3 += 0.5f * (gx * v.y + gy * v.x); u4 += 0.5f * (gy * v.z + gz * v.y); u5 += 0.5f * (gz * v.x + gx * v.z);
    const int q = t < WAVE ? rq0 : (t < 2 * WAVE ? rq1 : rev[ri.off + t]);
    if (q >= 0) fe[q] = make_float4(gx, gy, gz, 0.0f);
    else atomicOr(err, 1);
  }
  fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
  u0 = wave_sum(u0); u1 = wave_sum(u1); u2 = wave_sum(u2);
  u3 = wave_sum(u3); u4 = wave_sum(u4); u5 = wave_sum(u5);
  if (lane == 0) {
    fself[i] = make_float4(fx, fy, fz, 0.0f);
    float* vv = vir + 6 * (size_t)i;
    vv[0] = u0; vv[1] = u1; vv[2] = u2; vv[3] = u3; vv[4] = u4; vv[5] = u5;
  }
}

// one wave per atom: k_force contracts on the VALU in FP32, k_force_mm on the matrix cores (compensated BF16)
template <int PAT, bool UNI, bool ST>
__global__ __launch_bounds__(256) void k_force(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
    const int* __restrict__ rowcnt, const int* __restrict__ cnt, const float4* __restrict__ vec,
    const mg_basis_dev* __restrict__ bp, const float* __restrict__ dEdG, float4* __restrict__ fself,
    float4* __restrict__ fe, float* __restrict__ vir, float urep, int cap, const unsigned* __restrict__ gpl,
    const int* __restrict__ npl, size_t pstride, const int* __restrict__ rev, int* __restrict__ err,
    const int* __restrict__ ovf) {
  force_row<PAT, UNI, 0, ST>(N, S, spc, row_off, rowcnt, cnt, vec, bp, dEdG, fself, fe, vir, urep, cap, gpl, npl, pstride, rev, err, ovf);
}
template <int PAT, bool UNI, bool ST>
__global__ __launch_bounds__(256) void k_force_mm(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
    const int* __restrict__ rowcnt, const int* __restrict__ cnt, const float4* __restrict__ vec,
    const mg_basis_dev* __restrict__ bp, const float* __restrict__ dEdG, float4* __restrict__ fself,
    float4* __restrict__ fe, float* __restrict__ vir, float urep, int cap, const unsigned* __restrict__ gpl,
    const int* __restrict__ npl, size_t pstride, const int* __restrict__ rev, int* __restrict__ err,
    const int* __restrict__ ovf) {
  force_row<PAT, UNI, 1, ST>(N, S, spc, row_off, rowcnt, cnt, vec, bp, dEdG, fself, fe, vir, urep, cap, gpl, npl, pstride, rev, err, ovf);
}
// five waves per SIMD: 96 VGPRs (from 101), the few spilled values are saved once outside the triplet loop
// (-5% kernel time in same-box timing)
template <int PAT, bool UNI, bool ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_force_mh(int N, int S, const int* __restrict__ spc, const int* __restrict__ row_off,
    const int* __restrict__ rowcnt, const int* __restrict__ cnt, const float4* __restrict__ vec,
    const mg_basis_dev* __restrict__ bp, const float* __restrict__ dEdG, float4* __restrict__ fself,
    float4* __restrict__ fe, float* __restrict__ vir, float urep, int cap, const unsigned* __restrict__ gpl,
    const int* __restrict__ npl, size_t pstride, const int* __restrict__ rev, int* __restrict__ err,
    const int* __restrict__ ovf) {
  force_row<PAT, UNI, 2, ST>(N, S, spc, row_off, rowcnt, cnt, vec, bp, dEdG, fself, fe, vir, urep, cap, gpl, npl, pstride, rev, err, ovf);
}

// Force on atom a: its own term plus what its neighbours' centres sent it, which the force kernel wrote into a's row
// (slot e of row a holds the force from the centre at the far end of e): four lanes per atom stream the row.
// One launch for the end of an evaluation: blocks [0, gb) gather the forces (four lanes per atom, XCD-aware block
// order) and blocks [gb, gb + nchunk) sum energy and virial over one reduction chunk of MG_RCHUNK atoms each, in FP64
// in a fixed order.  A structure of one chunk gets its sums from that block directly; for a structure of several,
// each block writes its partial and counts itself in the structure's counter (cnt[chunk0], zero between launches),
// and the block that completes the count sums the partials in a fixed order (lane l takes chunks l, l + 64, ...,
// then a fixed butterfly) and clears the counter.  The partials are released and acquired at device scope around
// the counter, so that block sees every partial whatever XCD wrote it.
__device__ __forceinline__ double dwave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
#define ASM_L 8  // lanes per atom in the force gather
__global__ __launch_bounds__(256) void k_assemble(int N, int gb, const int* __restrict__ row_off,
                                                  const int* __restrict__ rowcnt, const float4* __restrict__ fself,
                                                  const float4* __restrict__ fe, double* __restrict__ F,
                                                  const int* __restrict__ ovf, const int* __restrict__ chunk_str,
                                                  int* __restrict__ cnt, const mg_struct_dev* __restrict__ str,
                                                  const float* __restrict__ Ei, const float* __restrict__ erep,
                                                  const float* __restrict__ vir, int have_vir,
                                                  double* __restrict__ part, double* __restrict__ Eout,
                                                  double* __restrict__ Uout) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if ((int)blockIdx.x < gb) {
    if (*ovf) return;
    // ASM_L lanes per atom: each load of the row's group covers ASM_L consecutive 16-byte edge forces (full lines)
    const int g = xcd_block(blockIdx.x, gb) * blockDim.x + t;
    const int a = g / ASM_L, sub = g % ASM_L;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (a < N) {
      const int e1 = row_off[a] + rowcnt[a];
      for (int e = row_off[a] + sub; e < e1; e += ASM_L) {
        const float4 v = fe[e];
        fx += v.x; fy += v.y; fz += v.z;
      }
    }
#pragma unroll
    for (int m = 1; m < ASM_L; m <<= 1) {
      fx += __shfl_xor(fx, m, ASM_L); fy += __shfl_xor(fy, m, ASM_L); fz += __shfl_xor(fz, m, ASM_L);
    }
    if (a < N && sub == 0) {
      float4 s = fself[a];
      F[3 * a] = fx + s.x; F[3 * a + 1] = fy + s.y; F[3 * a + 2] = fz + s.z;
    }
    return;
  }
  __shared__ double sh[4][7];
  __shared__ int last;
  const int ch = blockIdx.x - gb;
  const int s = chunk_str[ch];
  const mg_struct_dev* st = &str[s];
  const int a0 = st->atom0 + (ch - st->chunk0) * MG_RCHUNK, nc = st->nchunk;
  int a1 = a0 + MG_RCHUNK;
  if (a1 > st->atom0 + st->natom) a1 = st->atom0 + st->natom;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int a = a0 + t; a < a1; a += 256) {
    acc[0] += (double)Ei[a] + (erep ? (double)erep[a] : 0.0);
    if (have_vir)
      for (int q = 0; q < 6; q++) acc[1 + q] += (double)vir[6 * (size_t)a + q];
  }
#pragma unroll
  for (int q = 0; q < 7; q++) acc[q] = dwave_sum(acc[q]);
  if (lane == 0)
    for (int q = 0; q < 7; q++) sh[wv][q] = acc[q];
  __syncthreads();
  if (nc == 1) {  // the structure's sums are this block's
    if (t < 7) {
      const double v = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
      if (t == 0) Eout[s] = v;
      else Uout[6 * s + t - 1] = v;
    }
    return;
  }
  if (t < 7) {
    part[7 * (size_t)ch + t] = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
    __threadfence();  // each writer releases its partial at device scope
  }
  __syncthreads();
  if (t == 0) last = atomicAdd(cnt + st->chunk0, 1) == nc - 1;
  __syncthreads();
  if (!last || wv != 0) return;
  __threadfence();  // acquire the structure's partials
  double r[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int c = lane; c < nc; c += WAVE)
    for (int q = 0; q < 7; q++) r[q] += part[7 * (size_t)(st->chunk0 + c) + q];
#pragma unroll
  for (int q = 0; q < 7; q++) r[q] = dwave_sum(r[q]);
  if (lane == 0) {
    Eout[s] = r[0];
    for (int q = 0; q < 6; q++) Uout[6 * s + q] = r[1 + q];
    cnt[st->chunk0] = 0;
  }
}

// ---------------------------------------------------------------------------------------------- launchers
static thread_local hipError_t g_launch_err = hipSuccess;

extern "C" int mgk_status(void) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return 0;
  g_launch_err = e;
  return -1;
}

extern "C" const char* mgk_last_error(void) { return hipGetErrorString(g_launch_err); }

extern "C" {

int mgk_scan(const int* in, int* out, int n, void* tmp, size_t* tmp_bytes, hipStream_t st) {
  // exclusive scan of n values into out[0..n]; out[n] = total.  tmp == NULL queries the scratch size
  // (one int per SCAN_TILE block).  Three plain launches and no memset, so the scan can sit in a HIP graph.
  const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  const size_t need = (size_t)(nb + 1) * sizeof(int);
  if (!tmp) { *tmp_bytes = need; return 0; }
  if (*tmp_bytes < need) return -1;
  int* bsum = (int*)tmp;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(256), 0, st, in, out, n, (int*)nullptr, 1);
    return mgk_status();
  }
  if (n <= 1024 * SCAN1_PT) {  // one launch
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, st, in, out, n);
    return mgk_status();
  }
  hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(256), 0, st, in, out, n, bsum, 0);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, bsum, nb);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, st, out, n, bsum);
  return mgk_status();
}

int mgk_clear(int* a, int na, int* b, int nb, int* c, int nc, hipStream_t st) {
  const int n = na > nb ? (na > nc ? na : nc) : (nb > nc ? nb : nc);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_clear3, dim3((n + 255) / 256), dim3(256), 0, st, a, na, b, nb, c, nc);
  return mgk_status();
}

int mgk_check_caps(int N, const int* row_off, int* misc, int cap_edges, int cap_row, hipStream_t st) {
  hipLaunchKernelGGL(k_check_caps, dim3(1), dim3(64), 0, st, N, row_off, misc, cap_edges, cap_row);
  return mgk_status();
}

int mgk_prepare(int N, const double* xin, const int* atom_str, const mg_struct_dev* str, double* x, int* bin,
                int* bin_count, int* misc, int nclear, const double* xb, unsigned* mimg, double lim2, hipStream_t st) {
  const int n = N > nclear ? N : nclear;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_prepare, dim3((n + 255) / 256), dim3(256), 0, st, N, xin, atom_str, str, x, bin, bin_count, misc,
                     nclear, xb, mimg, lim2);
  return mgk_status();
}

int mgk_bins(int N, int nbins, const int* bin, const int* bin_start, int* bin_count, const double* x, const int* spc,
             void* brec, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_bin_fill, dim3((N + 255) / 256), dim3(256), 0, st, N, bin, bin_start, bin_count, x, spc,
                     (BinRec*)brec);
  return mgk_status();
}

int mgk_nbr_count(int N, const double* x, const int* spc, const int* atom_str, const mg_struct_dev* str, const int* bin,
                  const int* bin_start, const void* brec, int* cnt, int* rowcnt, int* misc, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_nbr<0>, dim3((N + NBR_WPB - 1) / NBR_WPB), dim3(64 * NBR_WPB), (size_t)NBR_WPB * NBR_LDS_CHUNK,
                     st, N, x, spc, atom_str, str,
                     bin, bin_start, (const BinRec*)brec, cnt, rowcnt, misc, (int*)nullptr, (u64*)nullptr,
                     (float4*)nullptr, (int*)nullptr, (unsigned short*)nullptr, 0, 0, 0, 0, NBR_WPB, 0.0);
  return mgk_status();
}

// Fill and sort the rows (LDS tile of cap entries: a power of two >= the longest row) with their reverse-lookup hash
// tables; the descriptor kernel then builds the reverse-edge index.  padded: no count pass, row i at i * maxrow.  Rows longer than
// 2048 neighbours are rejected (-2).
// slots of a row's reverse-lookup hash table: twice the row's LDS tile (a power of two >= the longest row, >= 64)
int mgk_rev_slots(int maxrow) {
  int cap = 64;
  while (cap < maxrow) cap <<= 1;
  return 2 * cap;
}

int mgk_nbr_fill(int N, const double* x, const int* spc, const int* atom_str, const mg_struct_dev* str, const int* bin,
                 const int* bin_start, const void* brec, int* row_off, int* rowcnt, int* cnt, void* key, void* vec,
                 int* rev, uint16_t* rhash, int* misc, int maxrow, int padded, int wpb_req, hipStream_t st) {
  if (N == 0) return 0;
  int cap = 64;
  while (cap < maxrow) cap <<= 1;
  if (cap > 2048) return -2;
  // per wave: the stencil chunk tables, keys + vectors (3 u64 per slot) and the 2 * cap 16-bit slots of the hash
  // table; fewer waves per block when long rows would not fit 64 KiB (a row of 2,048 fits one wave: 59,664 bytes)
  const int hslots = mgk_rev_slots(maxrow);
  const size_t per_wave = NBR_LDS_CHUNK + (size_t)cap * (sizeof(u64) + sizeof(float4)) + (size_t)hslots * sizeof(uint16_t);
  const size_t budget = 64 * 1024;
  // wpb_req 1, 2 or 4 asks for that many waves per block (fewer when they would not fit); 0 starts from NBR_WPB
  int wpb = (wpb_req == 1 || wpb_req == 2 || wpb_req == 4) ? wpb_req : NBR_WPB;
  while (wpb > 1 && (size_t)wpb * per_wave > budget) wpb >>= 1;
  const size_t sm = (size_t)wpb * per_wave;
  if (sm > budget &&
      hipFuncSetAttribute((const void*)&k_nbr<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_nbr<1>, dim3((N + wpb - 1) / wpb), dim3(64 * wpb), sm, st, N, x, spc, atom_str,
                     str, bin, bin_start, (const BinRec*)brec,
                     cnt, rowcnt, misc, row_off, (u64*)key, (float4*)vec, rev, (unsigned short*)rhash, hslots, cap,
                     padded, maxrow, wpb, 0.0);
  return mgk_status();
}

// Skin candidates at Rc + skin (rows of ccap keys; the stencil half widths Ks of each structure cover that radius),
// only when the lists are stale.  ccap <= 2048.
int mgk_nbr_cand(int N, const double* x, const int* spc, const int* atom_str, const mg_struct_dev* str, const int* bin,
                 const int* bin_start, const void* brec, int* crow, void* ckey, int* misc, int ccap, double skin,
                 hipStream_t st) {
  if (N == 0) return 0;
  int cap = 64;
  while (cap < ccap) cap <<= 1;
  if (cap > 2048) return -2;
  const size_t per_wave = NBR_LDS_CHUNK + (size_t)cap * (sizeof(u64) + sizeof(float4));
  int wpb = NBR_WPB;
  while (wpb > 1 && (size_t)wpb * per_wave > 64 * 1024) wpb >>= 1;
  const size_t sm = (size_t)wpb * per_wave;
  if (sm > 64 * 1024 &&
      hipFuncSetAttribute((const void*)&k_nbr<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_nbr<2>, dim3((N + wpb - 1) / wpb), dim3(64 * wpb), sm, st, N, x, spc, atom_str, str, bin,
                     bin_start, (const BinRec*)brec, (int*)nullptr, crow, misc, (int*)nullptr, (u64*)ckey,
                     (float4*)nullptr, (int*)nullptr, (unsigned short*)nullptr, 0, cap, 1, ccap, wpb, skin);
  return mgk_status();
}

int mgk_skin_commit(int N, const double* x, double* xb, unsigned* mimg, int* misc, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_skin_commit, dim3((N + 255) / 256), dim3(256), 0, st, N, x, xb, mimg, misc);
  return mgk_status();
}

// The padded rows (row i at i * maxrow, as mgk_nbr_fill with padded = 1) from the skin candidates.
int mgk_nbr_skin(int N, const double* x, const int* atom_str, const mg_struct_dev* str, const void* ckey,
                 const int* crow, int ccap, const unsigned* mimg, int* row_off, int* rowcnt, int* cnt, void* key,
                 void* vec, int* rev, uint16_t* rhash, int* misc, int maxrow, int wpb_req, hipStream_t st) {
  if (N == 0) return 0;
  int cap = 64;
  while (cap < maxrow) cap <<= 1;
  if (cap > 2048) return -2;
  const int hslots = mgk_rev_slots(maxrow);
  const size_t per_wave = (size_t)cap * (sizeof(u64) + sizeof(float4)) + (size_t)hslots * sizeof(uint16_t);
  int wpb = (wpb_req == 1 || wpb_req == 2 || wpb_req == 4) ? wpb_req : NBR_WPB;
  while (wpb > 1 && (size_t)wpb * per_wave > 64 * 1024) wpb >>= 1;
  const size_t sm = (size_t)wpb * per_wave;
  if (sm > 64 * 1024 &&
      hipFuncSetAttribute((const void*)&k_nbr_skin, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_nbr_skin, dim3((N + wpb - 1) / wpb), dim3(64 * wpb), sm, st, N, x, atom_str, str,
                     (const u64*)ckey, crow, ccap, mimg, cnt, rowcnt, misc, row_off, (u64*)key, (float4*)vec, rev,
                     (unsigned short*)rhash, hslots, cap, maxrow, wpb);
  return mgk_status();
}

// One wave per atom; wpb waves (atoms) per block.  wpb = 0 picks the largest of 4, 2, 1 whose LDS fits 64 KiB;
// a single wave may use up to the 160 KiB of a gfx950 CU.
#define MG_LDS_BLOCK (64 * 1024)
#define MG_LDS_CU (160 * 1024)
static int pick_wpb(int wpb, size_t sm_wave) {
  if (wpb == 1 || wpb == 2 || wpb == 4) return (size_t)wpb * sm_wave <= MG_LDS_CU ? wpb : -1;
  for (int w = 4; w >= 1; w >>= 1)
    if ((size_t)w * sm_wave <= MG_LDS_BLOCK) return w;
  return sm_wave <= MG_LDS_CU ? 1 : -1;
}
#define LAUNCH_PU(KER, PAT, UNI, SM, ...)                                                                            \
  do {                                                                                                               \
    if ((size_t)(SM) * wpb_ > MG_LDS_BLOCK &&                                                                        \
        hipFuncSetAttribute((const void*)&KER<PAT, UNI>, hipFuncAttributeMaxDynamicSharedMemorySize,                 \
                            (int)((SM) * wpb_)) != hipSuccess)                                                       \
      return -1;                                                                                                     \
    hipLaunchKernelGGL((KER<PAT, UNI>), dim3((N + wpb_ - 1) / wpb_), dim3(64 * wpb_), (SM) * wpb_, st, __VA_ARGS__); \
  } while (0)
#define DISPATCH_PU(KER, SM, ...)                                  \
  do {                                                             \
    const int wpb_ = pick_wpb(wpb, SM);                            \
    if (wpb_ < 0) return -3;                                       \
    const int pat_ = variant & 15;                                 \
    const bool uni_ = (variant >> 4) & 1;                          \
    if (pat_ == 1) { if (uni_) LAUNCH_PU(KER, 1, true, SM, __VA_ARGS__); else LAUNCH_PU(KER, 1, false, SM, __VA_ARGS__); } \
    else if (pat_ == 2) { if (uni_) LAUNCH_PU(KER, 2, true, SM, __VA_ARGS__); else LAUNCH_PU(KER, 2, false, SM, __VA_ARGS__); } \
    else { if (uni_) LAUNCH_PU(KER, 0, true, SM, __VA_ARGS__); else LAUNCH_PU(KER, 0, false, SM, __VA_ARGS__); } \
  } while (0)

// the force kernels also take ST: stored pair lists (gpl != NULL) or their own enumeration
#define LAUNCH_PUS(KER, PAT, UNI, ST, SM, ...)                                                                        \
  do {                                                                                                               \
    if ((size_t)(SM) * wpb_ > MG_LDS_BLOCK &&                                                                        \
        hipFuncSetAttribute((const void*)&KER<PAT, UNI, ST>, hipFuncAttributeMaxDynamicSharedMemorySize,             \
                            (int)((SM) * wpb_)) != hipSuccess)                                                       \
      return -1;                                                                                                     \
    hipLaunchKernelGGL((KER<PAT, UNI, ST>), dim3((N + wpb_ - 1) / wpb_), dim3(64 * wpb_), (SM) * wpb_, st, __VA_ARGS__); \
  } while (0)
#define LAUNCH_PU2(KER, PAT, UNI, SM, ...)                                        \
  do {                                                                            \
    if (st_) LAUNCH_PUS(KER, PAT, UNI, true, SM, __VA_ARGS__);                    \
    else LAUNCH_PUS(KER, PAT, UNI, false, SM, __VA_ARGS__);                       \
  } while (0)
#define DISPATCH_PUS(KER, SM, STORED, ...)                         \
  do {                                                             \
    const int wpb_ = pick_wpb(wpb, SM);                            \
    if (wpb_ < 0) return -3;                                       \
    const int pat_ = variant & 15;                                 \
    const bool uni_ = (variant >> 4) & 1;                          \
    const bool st_ = (STORED);                                     \
    if (pat_ == 1) { if (uni_) LAUNCH_PU2(KER, 1, true, SM, __VA_ARGS__); else LAUNCH_PU2(KER, 1, false, SM, __VA_ARGS__); } \
    else if (pat_ == 2) { if (uni_) LAUNCH_PU2(KER, 2, true, SM, __VA_ARGS__); else LAUNCH_PU2(KER, 2, false, SM, __VA_ARGS__); } \
    else { if (uni_) LAUNCH_PU2(KER, 0, true, SM, __VA_ARGS__); else LAUNCH_PU2(KER, 0, false, SM, __VA_ARGS__); } \
  } while (0)

// variant = G4 pattern (bits 0-3) | shared-eta flag (bit 4) | matrix-core force contraction (bit 5, force only)
int mgk_desc(int N, int S, int variant, const int* spc, const int* row_off, const int* rowcnt, const int* cnt,
             const void* vec,
             const mg_basis_dev* bp, float* G, float* erep, float urep, int maxrow, int wpb, unsigned* gpl, int* npl,
             size_t pstride, const void* key, const uint16_t* rhash, int* rev, int* err, const int* ovf,
             hipStream_t st) {
  const int hslots = mgk_rev_slots(maxrow);
  if (N == 0) return 0;
  const int cap = maxrow < 1 ? 1 : maxrow;
  const int R = ((variant >> 4) & 1) ? 2 : 4;
  size_t sm = (size_t)cap * R * sizeof(float4) + PRING * sizeof(unsigned);
  DISPATCH_PU(k_desc, sm, N, S, spc, row_off, rowcnt, cnt, (const float4*)vec, bp, G, erep, urep, cap, gpl, npl, pstride,
              (const u64*)key, (const unsigned short*)rhash, hslots, rev, err, ovf);
  return mgk_status();
}

int mgk_mlp(int N, const int* spc, const float* G, const float* W, int wstride, int D, int H1, int H2, float* Ei,
            float* dEdG, hipStream_t st) {
  if (N == 0) return 0;
  size_t sm = (size_t)(H1 + H2) * 64 * sizeof(float);
  hipLaunchKernelGGL(k_mlp_fp32, dim3((N + 63) / 64), dim3(64), sm, st, N, spc, G, W, wstride, D, H1, H2, Ei, dEdG);
  return mgk_status();
}

int mgk_force(int N, int S, int variant, const int* spc, const int* row_off, const int* rowcnt, const int* cnt,
              const void* vec,
              const mg_basis_dev* bp, const float* dEdG, void* fself, void* fe, float* vir, float urep, int maxrow,
              int wpb, const unsigned* gpl, const int* npl, size_t pstride, const int* rev, int* err, const int* ovf,
              hipStream_t st) {
  if (N == 0) return 0;
  const int cap = maxrow < 1 ? 1 : maxrow;
  const int R = ((variant >> 4) & 1) ? 2 : 4;
  size_t sm = (size_t)cap * (R + 2) * sizeof(float4) + MG_GRID * sizeof(float) + PRING * sizeof(unsigned);
  if ((variant >> 6) & 1) {
    DISPATCH_PUS(k_force_mh, sm, gpl != nullptr, N, S, spc, row_off, rowcnt, cnt, (const float4*)vec, bp, dEdG, (float4*)fself,
                (float4*)fe, vir, urep, cap, gpl, npl, pstride, rev, err, ovf);
    return mgk_status();
  }
  if ((variant >> 5) & 1) {
    DISPATCH_PUS(k_force_mm, sm, gpl != nullptr, N, S, spc, row_off, rowcnt, cnt, (const float4*)vec, bp, dEdG, (float4*)fself,
                (float4*)fe, vir, urep, cap, gpl, npl, pstride, rev, err, ovf);
    return mgk_status();
  }
  DISPATCH_PUS(k_force, sm, gpl != nullptr, N, S, spc, row_off, rowcnt, cnt, (const float4*)vec, bp, dEdG, (float4*)fself, (float4*)fe, vir,
              urep, cap, gpl, npl, pstride, rev, err, ovf);
  return mgk_status();
}

int mgk_assemble(int N, int forces, const int* row_off, const int* rowcnt, const void* fself, const void* fe, double* F,
                 const int* ovf, int nstr, int nchunk, const int* chunk_str, int* chunk_cnt, const mg_struct_dev* str,
                 const float* Ei, const float* erep, const float* vir, double* part, double* E, double* U,
                 hipStream_t st) {
  if (nstr == 0 || nchunk == 0) return 0;
  const int gb = forces && N > 0 ? (ASM_L * N + 255) / 256 : 0;
  hipLaunchKernelGGL(k_assemble, dim3(gb + nchunk), dim3(256), 0, st, N, gb, row_off, rowcnt, (const float4*)fself,
                     (const float4*)fe, F, ovf, chunk_str, chunk_cnt, str, Ei, erep, vir, forces, part, E, U);
  return mgk_status();
}

}  // extern "C"
