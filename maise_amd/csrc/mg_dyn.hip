// mg_dyn.hip -- device-resident dynamics: NVT molecular dynamics (velocity Verlet with a Nose-Hoover
// thermostat) and batched FIRE relaxation.  State (positions, velocities, thermostat, FIRE scalars) never
// leaves the device between steps; the host only sequences launches on the context stream.
//
// Reference callers replaced: Dynamics/Maxwell/CELL_KIN (src/cmod.c:12-23,224-280,283-549) and the relaxation
// driver CELL_RELX/CELL_MIN (src/cmod.c:901-977, src/cmin.c:177-301).  The reference integrates NVT with one
// force call per step (x += v dt + f dt^2/2, v += f dt, f = F/m - xi v, xi += dt (2K/(3N kT) - 1)/tau^2) and
// relaxes with GSL BFGS2; this build uses velocity Verlet with the same thermostat equation of motion
// (half-step splitting) and FIRE (Bitzek et al., PRL 97, 170201 (2006)).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_internal.h"

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// Maxwell-Boltzmann velocities: v_q = sqrt(kT/m) * N(0,1) by Box-Muller on a counter-based hash
__global__ void k_md_maxwell(int N, unsigned long long seed, const double* __restrict__ mass, double kT,
                             double* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double sd = sqrt(kT / mass[i]);
  for (int q = 0; q < 3; q++) {
    uint64_t h1 = mix64(seed ^ mix64(2ull * (3ull * i + q)));
    uint64_t h2 = mix64(seed ^ mix64(2ull * (3ull * i + q) + 1ull));
    double r = sqrt(-2.0 * log(u01(h1)));
    v[3 * i + q] = sd * r * cos(6.283185307179586 * u01(h2));
  }
}

// one block: remove the centre-of-mass velocity and rescale to exactly the target temperature
__global__ __launch_bounds__(1024) void k_md_fix(int N, const double* __restrict__ mass, double* __restrict__ v,
                                                 double kT, double ndof) {
  __shared__ double sh[4][1024];
  const int t = threadIdx.x;
  double p[4] = {0, 0, 0, 0};
  for (int i = t; i < N; i += 1024) {
    p[0] += mass[i] * v[3 * i]; p[1] += mass[i] * v[3 * i + 1]; p[2] += mass[i] * v[3 * i + 2]; p[3] += mass[i];
  }
  for (int k = 0; k < 4; k++) sh[k][t] = p[k];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) for (int k = 0; k < 4; k++) sh[k][t] += sh[k][t + w];
    __syncthreads();
  }
  const double vx = sh[0][0] / sh[3][0], vy = sh[1][0] / sh[3][0], vz = sh[2][0] / sh[3][0];
  __syncthreads();
  double k2 = 0.0;
  for (int i = t; i < N; i += 1024) {
    v[3 * i] -= vx; v[3 * i + 1] -= vy; v[3 * i + 2] -= vz;
    k2 += mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  }
  sh[0][t] = k2;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) sh[0][t] += sh[0][t + w];
    __syncthreads();
  }
  const double s = sh[0][0] > 0.0 ? sqrt(ndof * kT / sh[0][0]) : 0.0;
  for (int i = t; i < N; i += 1024) { v[3 * i] *= s; v[3 * i + 1] *= s; v[3 * i + 2] *= s; }
}

// one block: kinetic energy K, half-step thermostat update xi += dt/2 (2K/(g kT) - 1)/tau^2 (skipped when
// tau <= 0), velocity scale s = exp(-xi dt/2).  dyn = {xi, K, s}; trace[slot] = K when slot >= 0.
__global__ __launch_bounds__(1024) void k_md_thermo(int N, const double* __restrict__ mass, const double* __restrict__ v,
                                                    double* __restrict__ dyn, double dt, double kT, double ndof,
                                                    double tau, double* __restrict__ trace, int slot) {
  __shared__ double sh[1024];
  const int t = threadIdx.x;
  double k2 = 0.0;
  for (int i = t; i < N; i += 1024)
    k2 += mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  sh[t] = k2;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) {
    double K = 0.5 * sh[0];
    double xi = dyn[0];
    if (tau > 0.0) xi += 0.5 * dt * (2.0 * K / (ndof * kT) - 1.0) / (tau * tau);
    dyn[0] = xi;
    dyn[1] = K;
    dyn[2] = tau > 0.0 ? exp(-0.5 * dt * xi) : 1.0;
    if (trace && slot >= 0) trace[slot] = K;
  }
}

// v <- v s + dt/2 F/m ; x_next <- lambda (x + dt v), lambda = dyn[4] with the barostat on (the lattice was scaled
// by the same factor when lambda was set), else 1
__global__ void k_md_kick_drift(int N, double dt, const double* __restrict__ mass, const double* __restrict__ F,
                                double* __restrict__ v, const double* __restrict__ x, double* __restrict__ xin,
                                const double* __restrict__ dyn, int baro) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double s = dyn[2], h = 0.5 * dt / mass[i], lam = baro ? dyn[4] : 1.0;
  for (int q = 0; q < 3; q++) {
    double vv = v[3 * i + q] * s + h * F[3 * i + q];
    v[3 * i + q] = vv;
    xin[3 * i + q] = baro ? lam * (x[3 * i + q] + dt * vv) : x[3 * i + q] + dt * vv;
  }
}

// Berendsen barostat of the reference's NPT and isobaric runs (src/cmod.c:371, 408-411, 457-468): after a force
// evaluation, lambda = (1 - (ktt dt / 3)(p - p_int) eV2GPa)^(1/3) with the internal pressure
// p_int = (U_xx + U_yy + U_zz) / (3 V) (CELL_PRS, src/cmod.c:552-561; p, p_int in eV/A^3, ktt in 1/(GPa fs)).  The
// lattice rows scale by lambda now (reciprocal rows by 1/lambda, the bin stencil follows) and every position at the
// next drift (the reference rescales only X[0..2], src/cmod.c:462-465; all atoms scale here).  The argument is
// kept >= 1/8 (lambda >= 1/2 per step).  dyn[4] = lambda, dyn[5] = p_int, dyn[6] = new volume.
__global__ void k_md_baro(mg_struct_dev* __restrict__ str, const double* __restrict__ U, double* __restrict__ dyn,
                          double p, double ktt_dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  mg_struct_dev& d = str[0];
  const double* L = d.L;
  const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                     L[2] * (L[3] * L[7] - L[4] * L[6]);
  const double V = fabs(det);
  const double pint = (U[0] + U[1] + U[2]) / (3.0 * V);
  const double arg = fmax(1.0 - (ktt_dt / 3.0) * (p - pint) * 160.2176565, 0.125);
  const double lam = cbrt(arg);
  for (int k = 0; k < 9; k++) { d.L[k] *= lam; d.R[k] /= lam; }
  const double rc = sqrt(d.rc2);
  for (int q = 0; q < 3; q++) {
    const double h = rc * sqrt(d.R[3 * q] * d.R[3 * q] + d.R[3 * q + 1] * d.R[3 * q + 1] + d.R[3 * q + 2] * d.R[3 * q + 2]) *
                     (1.0 + 1e-6);
    const int K = (int)ceil(h * d.nb[q]);
    d.K[q] = K < 1 ? 1 : K;
  }
  dyn[4] = lam;
  dyn[5] = pint;
  dyn[6] = V * lam * lam * lam;
}

// v <- (v + dt/2 F/m) exp(-xi dt/2)
__global__ void k_md_kick_scale(int N, double dt, const double* __restrict__ mass, const double* __restrict__ F,
                                double* __restrict__ v, const double* __restrict__ dyn, double tau) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double s = tau > 0.0 ? exp(-0.5 * dt * dyn[0]) : 1.0, h = 0.5 * dt / mass[i];
  for (int q = 0; q < 3; q++) v[3 * i + q] = (v[3 * i + q] + h * F[3 * i + q]) * s;
}

// end of an MD step replayed from a graph: the step index lives in dyn[3]; trace[s] = E, trace[nsteps + s] = K and,
// with the barostat, trace[2 nsteps + s] = volume, trace[3 nsteps + s] = internal pressure
__global__ void k_md_record(const double* __restrict__ E, double* __restrict__ dyn, double* __restrict__ trace,
                            int nsteps, int baro) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int s = (int)dyn[3];
  if (s >= 0 && s < nsteps) {
    trace[s] = E[0];
    trace[nsteps + s] = dyn[1];
    if (baro) {
      trace[2 * nsteps + s] = dyn[6];
      trace[3 * nsteps + s] = dyn[5];
    }
  }
  dyn[3] = (double)(s + 1);
}

__global__ void k_copy_d(const double* __restrict__ src, double* __restrict__ dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// FIRE, one block per structure.  state[s] = {dt, alpha, npos, done, steps}.  Unit masses, as in the FIRE
// paper's formulation for relaxation; the step dr = dt v is capped at maxstep (A) per structure.
#define FIRE_NMIN 5
#define FIRE_FINC 1.1
#define FIRE_FDEC 0.5
#define FIRE_ASTART 0.1
#define FIRE_FA 0.99
__global__ __launch_bounds__(256) void k_fire(int S, const mg_struct_dev* __restrict__ str, const double* __restrict__ F,
                                              double* __restrict__ v, const double* __restrict__ x,
                                              double* __restrict__ xin, double* __restrict__ state, double fmax,
                                              double dtmax, double maxstep) {
  __shared__ double sh[4][256];
  const int s = blockIdx.x, t = threadIdx.x;
  const int a0 = str[s].atom0, n = str[s].natom;
  double* st = state + 5 * s;
  // a converged structure keeps its positions as they are (re-wrapping them every iteration would move them by
  // rounding, so results would depend on how long the structure stays in the batch)
  if (st[3] != 0.0) {
    for (int i = a0 + t; i < a0 + n; i += 256)
      for (int q = 0; q < 3; q++) v[3 * i + q] = 0.0;
    return;
  }
  double vf = 0, ff = 0, vv = 0, fm = 0;
  for (int i = a0 + t; i < a0 + n; i += 256) {
    double f2 = 0;
    for (int q = 0; q < 3; q++) {
      double f = F[3 * i + q], u = v[3 * i + q];
      vf += f * u; f2 += f * f; vv += u * u;
    }
    ff += f2;
    fm = f2 > fm ? f2 : fm;
  }
  sh[0][t] = vf; sh[1][t] = ff; sh[2][t] = vv; sh[3][t] = fm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] += sh[0][t + w]; sh[1][t] += sh[1][t + w]; sh[2][t] += sh[2][t + w];
      sh[3][t] = sh[3][t] > sh[3][t + w] ? sh[3][t] : sh[3][t + w];
    }
    __syncthreads();
  }
  vf = sh[0][0]; ff = sh[1][0]; vv = sh[2][0]; fm = sh[3][0];
  __syncthreads();
  if (fm < fmax * fmax) {
    // converged now: store the wrapped Cartesian positions once (inputs may still be fractional at step 0)
    for (int i = a0 + t; i < a0 + n; i += 256)
      for (int q = 0; q < 3; q++) { xin[3 * i + q] = x[3 * i + q]; v[3 * i + q] = 0.0; }
    if (t == 0) st[3] = 1.0;
    return;
  }
  double dt = st[0], alpha = st[1], npos = st[2];
  bool reset = false;
  double c1 = 1.0, c2 = 0.0;
  if (vf > 0.0) {
    c1 = 1.0 - alpha;
    c2 = ff > 0.0 ? alpha * sqrt(vv / ff) : 0.0;
    if (npos > FIRE_NMIN) {
      dt = fmin(dt * FIRE_FINC, dtmax);
      alpha *= FIRE_FA;
    }
    npos += 1.0;
  } else {
    reset = true;
    alpha = FIRE_ASTART;
    dt *= FIRE_FDEC;
    npos = 0.0;
  }
  double dr2 = 0.0;
  for (int i = a0 + t; i < a0 + n; i += 256)
    for (int q = 0; q < 3; q++) {
      double u = reset ? 0.0 : c1 * v[3 * i + q] + c2 * F[3 * i + q];
      u += dt * F[3 * i + q];
      v[3 * i + q] = u;
      dr2 += (dt * u) * (dt * u);
    }
  sh[0][t] = dr2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[0][t] += sh[0][t + w];
    __syncthreads();
  }
  const double nrm = sqrt(sh[0][0]);
  const double sc = nrm > maxstep ? maxstep / nrm : 1.0;
  for (int i = a0 + t; i < a0 + n; i += 256)
    for (int q = 0; q < 3; q++) xin[3 * i + q] = x[3 * i + q] + sc * dt * v[3 * i + q];
  if (t == 0) {
    st[0] = dt; st[1] = alpha; st[2] = npos; st[4] += 1.0;
  }
}

// FIRE over positions and cell (the reference's full-cell relaxation, RLXT 3: enthalpy H = E + pV minimised over
// atoms and lattice, src/cmin.c:90-166, there with GSL BFGS2).  The cell degree of freedom is the strain e of the
// current lattice, scaled by the atom count N (generalized coordinate Y = N e, generalized force
// G = (U - p V 1) / N with U the virial, so cell and atom steps are of comparable size).  Each step deforms the
// lattice and the positions by 1 + dY / N, then the lattice, its reciprocal and the neighbour-search stencil of the
// structure are updated on the device (bin counts stay as planned).  Converged when every |F| < fmax and every
// |U - p V 1| / V < smax.  Cell velocities: 9 per structure.
__global__ __launch_bounds__(256) void k_fire_cell(int S, mg_struct_dev* __restrict__ str, const double* __restrict__ F,
                                                   const double* __restrict__ U, double* __restrict__ v,
                                                   double* __restrict__ cellv, const double* __restrict__ x,
                                                   double* __restrict__ xin, double* __restrict__ state, double p,
                                                   double fmax, double smax, double dtmax, double maxstep) {
  __shared__ double sh[5][256];
  __shared__ double Dm[9];
  const int s = blockIdx.x, t = threadIdx.x;
  const int a0 = str[s].atom0, n = str[s].natom;
  const bool cell = str[s].nd == 3;
  double* st = state + 5 * s;
  double* cv = cellv + 9 * s;
  if (st[3] != 0.0) {
    for (int i = a0 + t; i < a0 + n; i += 256)
      for (int q = 0; q < 3; q++) v[3 * i + q] = 0.0;
    return;
  }
  // generalized cell force G (symmetric), from the structure's virial xx yy zz xy yz zx
  double L[9];
  for (int k = 0; k < 9; k++) L[k] = str[s].L[k];
  const double V = fabs(L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                        L[2] * (L[3] * L[7] - L[4] * L[6]));
  const double* u = U + 6 * s;
  double G[9] = {u[0] - p * V, u[3], u[5], u[3], u[1] - p * V, u[4], u[5], u[4], u[2] - p * V};
  double smx = 0.0;
  for (int k = 0; k < 9; k++) {
    const double sk = fabs(G[k]) / V;
    smx = sk > smx ? sk : smx;
    G[k] = cell ? G[k] / (double)n : 0.0;
  }
  if (!cell) smx = 0.0;
  double vf = 0, ff = 0, vv = 0, fm = 0;
  for (int i = a0 + t; i < a0 + n; i += 256) {
    double f2 = 0;
    for (int q = 0; q < 3; q++) {
      double f = F[3 * i + q], w = v[3 * i + q];
      vf += f * w; f2 += f * f; vv += w * w;
    }
    ff += f2;
    fm = f2 > fm ? f2 : fm;
  }
  if (t == 0)
    for (int k = 0; k < 9; k++) { vf += cv[k] * G[k]; ff += G[k] * G[k]; vv += cv[k] * cv[k]; }
  sh[0][t] = vf; sh[1][t] = ff; sh[2][t] = vv; sh[3][t] = fm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] += sh[0][t + w]; sh[1][t] += sh[1][t + w]; sh[2][t] += sh[2][t + w];
      sh[3][t] = sh[3][t] > sh[3][t + w] ? sh[3][t] : sh[3][t + w];
    }
    __syncthreads();
  }
  vf = sh[0][0]; ff = sh[1][0]; vv = sh[2][0]; fm = sh[3][0];
  __syncthreads();
  if (fm < fmax * fmax && smx < smax) {
    for (int i = a0 + t; i < a0 + n; i += 256)
      for (int q = 0; q < 3; q++) { xin[3 * i + q] = x[3 * i + q]; v[3 * i + q] = 0.0; }
    if (t == 0) {
      st[3] = 1.0;
      for (int k = 0; k < 9; k++) cv[k] = 0.0;
    }
    return;
  }
  double dt = st[0], alpha = st[1], npos = st[2];
  bool reset = false;
  double c1 = 1.0, c2 = 0.0;
  if (vf > 0.0) {
    c1 = 1.0 - alpha;
    c2 = ff > 0.0 ? alpha * sqrt(vv / ff) : 0.0;
    if (npos > FIRE_NMIN) {
      dt = fmin(dt * FIRE_FINC, dtmax);
      alpha *= FIRE_FA;
    }
    npos += 1.0;
  } else {
    reset = true;
    alpha = FIRE_ASTART;
    dt *= FIRE_FDEC;
    npos = 0.0;
  }
  double dr2 = 0.0;
  for (int i = a0 + t; i < a0 + n; i += 256)
    for (int q = 0; q < 3; q++) {
      double w = reset ? 0.0 : c1 * v[3 * i + q] + c2 * F[3 * i + q];
      w += dt * F[3 * i + q];
      v[3 * i + q] = w;
      dr2 += (dt * w) * (dt * w);
    }
  double cvn[9];
  for (int k = 0; k < 9; k++) {
    double w = reset ? 0.0 : c1 * cv[k] + c2 * G[k];
    cvn[k] = w + dt * G[k];
  }
  if (t == 0)
    for (int k = 0; k < 9; k++) dr2 += (dt * cvn[k]) * (dt * cvn[k]);
  sh[0][t] = dr2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[0][t] += sh[0][t + w];
    __syncthreads();
  }
  const double nrm = sqrt(sh[0][0]);
  const double sc = nrm > maxstep ? maxstep / nrm : 1.0;
  if (t == 0) {
    // deformation D = 1 + dY / N (symmetric, so no rotation)
    for (int k = 0; k < 9; k++) Dm[k] = (k % 4 == 0 ? 1.0 : 0.0) + (cell ? sc * dt * cvn[k] / (double)n : 0.0);
  }
  __syncthreads();
  for (int i = a0 + t; i < a0 + n; i += 256) {
    const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    for (int q = 0; q < 3; q++)
      xin[3 * i + q] = x0 * Dm[q] + x1 * Dm[3 + q] + x2 * Dm[6 + q] + sc * dt * v[3 * i + q];
  }
  if (t == 0) {
    for (int k = 0; k < 9; k++) cv[k] = cvn[k];
    st[0] = dt; st[1] = alpha; st[2] = npos; st[4] += 1.0;
    if (cell) {
      mg_struct_dev& d = str[s];
      double Ln[9];
      for (int q = 0; q < 3; q++)
        for (int b = 0; b < 3; b++) Ln[3 * q + b] = L[3 * q] * Dm[b] + L[3 * q + 1] * Dm[3 + b] + L[3 * q + 2] * Dm[6 + b];
      // reciprocal rows (L1 x L2, L2 x L0, L0 x L1) / det
      const double* a = Ln; const double* bb = Ln + 3; const double* c = Ln + 6;
      double R[9];
      R[0] = bb[1] * c[2] - bb[2] * c[1]; R[1] = bb[2] * c[0] - bb[0] * c[2]; R[2] = bb[0] * c[1] - bb[1] * c[0];
      R[3] = c[1] * a[2] - c[2] * a[1];   R[4] = c[2] * a[0] - c[0] * a[2];   R[5] = c[0] * a[1] - c[1] * a[0];
      R[6] = a[1] * bb[2] - a[2] * bb[1]; R[7] = a[2] * bb[0] - a[0] * bb[2]; R[8] = a[0] * bb[1] - a[1] * bb[0];
      const double det = a[0] * R[0] + a[1] * R[1] + a[2] * R[2];
      for (int k = 0; k < 9; k++) { d.L[k] = Ln[k]; d.R[k] = R[k] / det; }
      const double rc = sqrt(d.rc2);
      for (int q = 0; q < 3; q++) {
        const double h = rc * sqrt(d.R[3 * q] * d.R[3 * q] + d.R[3 * q + 1] * d.R[3 * q + 1] + d.R[3 * q + 2] * d.R[3 * q + 2]) *
                         (1.0 + 1e-6);
        const int K = (int)ceil(h * d.nb[q]);
        d.K[q] = K < 1 ? 1 : K;
      }
    }
  }
}

__global__ void k_fire_init(int S, double* __restrict__ state, double dt0) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double* st = state + 5 * s;
  st[0] = dt0; st[1] = FIRE_ASTART; st[2] = 0.0; st[3] = 0.0; st[4] = 0.0;
}

__global__ void k_fire_count(int S, const double* __restrict__ state, int* __restrict__ out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S && state[5 * s + 3] != 0.0) atomicAdd(out, 1);
}

extern "C" {

int mgk_md_maxwell(int N, unsigned long long seed, const double* mass, double kT, double* v, hipStream_t st) {
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_md_maxwell, dim3((N + 255) / 256), dim3(256), 0, st, N, seed, mass, kT, v);
  return mgk_status();
}

int mgk_md_fix(int N, const double* mass, double* v, double kT, double ndof, hipStream_t st) {
  hipLaunchKernelGGL(k_md_fix, dim3(1), dim3(1024), 0, st, N, mass, v, kT, ndof);
  return mgk_status();
}

int mgk_md_thermo(int N, const double* mass, const double* v, double* dyn, double dt, double kT, double ndof,
                  double tau, double* trace, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_md_thermo, dim3(1), dim3(1024), 0, st, N, mass, v, dyn, dt, kT, ndof, tau, trace, slot);
  return mgk_status();
}

int mgk_md_kick_drift(int N, double dt, const double* mass, const double* F, double* v, const double* x, double* xin,
                      const double* dyn, int baro, hipStream_t st) {
  hipLaunchKernelGGL(k_md_kick_drift, dim3((N + 255) / 256), dim3(256), 0, st, N, dt, mass, F, v, x, xin, dyn, baro);
  return mgk_status();
}

int mgk_md_baro(mg_struct_dev* str, const double* U, double* dyn, double p, double ktt_dt, hipStream_t st) {
  hipLaunchKernelGGL(k_md_baro, dim3(1), dim3(64), 0, st, str, U, dyn, p, ktt_dt);
  return mgk_status();
}

int mgk_md_kick_scale(int N, double dt, const double* mass, const double* F, double* v, const double* dyn, double tau,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_md_kick_scale, dim3((N + 255) / 256), dim3(256), 0, st, N, dt, mass, F, v, dyn, tau);
  return mgk_status();
}

int mgk_md_record(const double* E, double* dyn, double* trace, int nsteps, int baro, hipStream_t st) {
  hipLaunchKernelGGL(k_md_record, dim3(1), dim3(64), 0, st, E, dyn, trace, nsteps, baro);
  return mgk_status();
}

int mgk_copy_d(const double* src, double* dst, int n, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_copy_d, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n);
  return mgk_status();
}

int mgk_fire_init(int S, double* state, double dt0, hipStream_t st) {
  hipLaunchKernelGGL(k_fire_init, dim3((S + 255) / 256), dim3(256), 0, st, S, state, dt0);
  return mgk_status();
}

int mgk_fire(int S, const mg_struct_dev* str, const double* F, double* v, const double* x, double* xin, double* state,
             double fmax, double dtmax, double maxstep, hipStream_t st) {
  hipLaunchKernelGGL(k_fire, dim3(S), dim3(256), 0, st, S, str, F, v, x, xin, state, fmax, dtmax, maxstep);
  return mgk_status();
}

int mgk_fire_cell(int S, mg_struct_dev* str, const double* F, const double* U, double* v, double* cellv,
                  const double* x, double* xin, double* state, double p, double fmax, double smax, double dtmax,
                  double maxstep, hipStream_t st) {
  hipLaunchKernelGGL(k_fire_cell, dim3(S), dim3(256), 0, st, S, str, F, U, v, cellv, x, xin, state, p, fmax, smax,
                     dtmax, maxstep);
  return mgk_status();
}

int mgk_fire_count(int S, const double* state, int* out, hipStream_t st) {
  hipLaunchKernelGGL(k_fire_count, dim3((S + 255) / 256), dim3(256), 0, st, S, state, out);
  return mgk_status();
}

}  // extern "C"
