// mg_dyn.hip -- device-resident dynamics: NVT molecular dynamics (velocity Verlet with a Nose-Hoover
// thermostat) and batched FIRE and L-BFGS relaxation.  State (positions, velocities, thermostat, optimiser state)
// never leaves the device between steps; the host only sequences launches on the context stream.
//
// Reference callers replaced: Dynamics/Maxwell/CELL_KIN (src/cmod.c:12-23,224-280,283-549) and the relaxation
// driver CELL_RELX/CELL_MIN (src/cmod.c:901-977, src/cmin.c:177-301).  The reference integrates NVT with one
// force call per step (x += v dt + f dt^2/2, v += f dt, f = F/m - xi v, xi += dt (2K/(3N kT) - 1)/tau^2) and
// relaxes with GSL BFGS2; this build uses velocity Verlet with the same thermostat equation of motion
// (half-step splitting), FIRE (Bitzek et al., PRL 97, 170201 (2006)) and limited-memory BFGS.
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

// ---------------------------------------------------------------------------------------------- L-BFGS
// Limited-memory BFGS (Nocedal, Math. Comp. 35 (1980) 773; two-loop recursion as Nocedal & Wright, Numerical
// Optimization, Alg. 7.4), one block of 64 or 256 threads per structure, a thread per atom, every sum in FP64 in a
// fixed order (a fixed DPP tree per wave, then the waves in order), so a run does not depend on how it is launched.
// state[s] = {gamma, head, count, done, steps}: gamma = s.y / y.y of the newest pair (the initial inverse Hessian
// of the recursion; h0 before the first pair), head = the ring slot the next pair goes to, count = stored pairs.
// History ring: hs/hy[slot * hstride + 3 i + q] (hstride = 3 N), rho[s * m + slot] = 1 / (y.s).  dr holds the
// previous step (it doubles as the recursion's work vector), f0 the forces at the previous positions.
// A pair with y.s <= 1e-10 |y||s| (no positive curvature along the step) is not stored; a direction that is not a
// descent direction (p.F <= 0) drops the memory and steps along h0 F.  The step is scaled so no atom moves more than
// maxstep (A).
#define LB_MAXM 32
// Wave sums and maxima through DPP moves in the VALU rather than LDS permutes: xor 1 and xor 2 inside quads, rotations
// by 4 and 8 inside rows of 16 (every lane then holds its row's value), row 0 into row 1 and row 2 into row 3
// (row_bcast:15), rows 0-1 into rows 2-3 (row_bcast:31); lane 63 holds the wave's value, read into a scalar register
// and so the same in every lane.  Lanes outside the row mask keep 0 (sum) or their own value (max).
template <int CTRL, int RM>
__device__ __forceinline__ double lb_dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lb_lane63(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lb_wsum(double v) {
  v += lb_dpp<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += lb_dpp<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v += lb_dpp<0x124, 0xf>(v);  // row_ror:4
  v += lb_dpp<0x128, 0xf>(v);  // row_ror:8
  v += lb_dpp<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += lb_dpp<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return lb_lane63(v);
}
__device__ __forceinline__ double lb_wmax(double v) {  // v >= 0 on every lane (squared norms)
  v = fmax(v, lb_dpp<0xB1, 0xf>(v));
  v = fmax(v, lb_dpp<0x4E, 0xf>(v));
  v = fmax(v, lb_dpp<0x124, 0xf>(v));
  v = fmax(v, lb_dpp<0x128, 0xf>(v));
  v = fmax(v, lb_dpp<0x142, 0xa>(v));
  v = fmax(v, lb_dpp<0x143, 0xc>(v));
  return lb_lane63(v);
}
// block sum / max: red[] holds one partial per wave; the leading barrier keeps the previous call's readers ahead of
// this call's writers
__device__ __forceinline__ double lb_bsum(double v, double* red, int t, int nw) {
  v = lb_wsum(v);
  if (nw == 1) return v;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < nw; w++) r += red[w];
  return r;
}
__device__ __forceinline__ double lb_bmax(double v, double* red, int t, int nw) {
  v = lb_wmax(v);
  if (nw == 1) return v;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < nw; w++) r = fmax(r, red[w]);
  return r;
}

__global__ __launch_bounds__(256) void k_lbfgs(int S, const mg_struct_dev* __restrict__ str, const double* __restrict__ F,
                                               double* __restrict__ dr, double* __restrict__ f0, double* __restrict__ hs,
                                               double* __restrict__ hy, double* __restrict__ rho, int m, size_t hstride,
                                               const double* __restrict__ x, double* __restrict__ xin,
                                               double* __restrict__ state, double ftol, double maxstep, double h0) {
  __shared__ double red[4];
  __shared__ double a_sh[LB_MAXM], rho_sh[LB_MAXM];
  const int s = blockIdx.x, t = threadIdx.x, nt = blockDim.x, nw = nt >> 6;
  const int a0 = str[s].atom0, n = str[s].natom;
  double* st = state + 5 * s;
  if (st[3] != 0.0) {  // converged: positions kept as they are (see k_fire)
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) dr[3 * i + q] = 0.0;
    return;
  }
  double fm = 0.0;
  for (int i = a0 + t; i < a0 + n; i += nt) {
    const double* f = F + 3 * i;
    fm = fmax(fm, f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  }
  fm = lb_bmax(fm, red, t, nw);
  if (fm < ftol * ftol) {
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) { xin[3 * i + q] = x[3 * i + q]; dr[3 * i + q] = 0.0; }
    if (t == 0) st[3] = 1.0;
    return;
  }
  double gam = st[0];
  int head = (int)st[1], cnt = (int)st[2];
  const double steps = st[4];
  double* rs = rho + (size_t)s * m;
  for (int k = t; k < m; k += nt) rho_sh[k] = rs[k];
  // the last step and the change of the forces along it become the newest pair (positive curvature only)
  if (steps > 0.0) {
    double ys = 0.0, yy = 0.0, ss = 0.0;
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) {
        const double y = f0[3 * i + q] - F[3 * i + q], d = dr[3 * i + q];
        ys += y * d; yy += y * y; ss += d * d;
      }
    ys = lb_bsum(ys, red, t, nw);
    yy = lb_bsum(yy, red, t, nw);
    ss = lb_bsum(ss, red, t, nw);
    if (yy > 0.0 && ys > 1e-10 * sqrt(yy * ss)) {
      double* ps = hs + (size_t)head * hstride;
      double* py = hy + (size_t)head * hstride;
      for (int i = a0 + t; i < a0 + n; i += nt)
        for (int q = 0; q < 3; q++) {
          ps[3 * i + q] = dr[3 * i + q];
          py[3 * i + q] = f0[3 * i + q] - F[3 * i + q];
        }
      __syncthreads();  // rho_sh: loaded above by some threads, written here by one
      if (t == 0) rho_sh[head] = 1.0 / ys;
      head = head + 1 == m ? 0 : head + 1;
      cnt = cnt < m ? cnt + 1 : m;
      gam = ys / yy;
    }
  }
  __syncthreads();
  // two-loop recursion on q = F, the negative gradient, so it ends with the search direction p = H F (kept in dr;
  // each thread touches only its own atoms' components)
  for (int i = a0 + t; i < a0 + n; i += nt)
    for (int q = 0; q < 3; q++) dr[3 * i + q] = F[3 * i + q];
  for (int k = 0; k < cnt; k++) {
    int slot = head - 1 - k;
    if (slot < 0) slot += m;
    const double* ps = hs + (size_t)slot * hstride;
    const double* py = hy + (size_t)slot * hstride;
    double d = 0.0;
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) d += ps[3 * i + q] * dr[3 * i + q];
    const double a = rho_sh[slot] * lb_bsum(d, red, t, nw);
    if (t == 0) a_sh[k] = a;
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) dr[3 * i + q] -= a * py[3 * i + q];
  }
  const double g = cnt > 0 ? gam : h0;
  for (int i = a0 + t; i < a0 + n; i += nt)
    for (int q = 0; q < 3; q++) dr[3 * i + q] *= g;
  __syncthreads();  // a_sh
  for (int k = cnt - 1; k >= 0; k--) {
    int slot = head - 1 - k;
    if (slot < 0) slot += m;
    const double* ps = hs + (size_t)slot * hstride;
    const double* py = hy + (size_t)slot * hstride;
    double d = 0.0;
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) d += py[3 * i + q] * dr[3 * i + q];
    const double c = a_sh[k] - rho_sh[slot] * lb_bsum(d, red, t, nw);
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) dr[3 * i + q] += c * ps[3 * i + q];
  }
  // dr = p = H F; not a descent direction (p.F <= 0): steepest descent, memory dropped
  double pf = 0.0;
  for (int i = a0 + t; i < a0 + n; i += nt)
    for (int q = 0; q < 3; q++) pf += dr[3 * i + q] * F[3 * i + q];
  pf = lb_bsum(pf, red, t, nw);
  const bool sd = !(pf > 0.0);
  if (sd) {
    cnt = 0;
    gam = h0;
    for (int i = a0 + t; i < a0 + n; i += nt)
      for (int q = 0; q < 3; q++) dr[3 * i + q] = h0 * F[3 * i + q];
  }
  double l2 = 0.0;
  for (int i = a0 + t; i < a0 + n; i += nt) {
    const double* d = dr + 3 * i;
    l2 = fmax(l2, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }
  l2 = lb_bmax(l2, red, t, nw);
  const double sc = l2 > maxstep * maxstep ? maxstep / sqrt(l2) : 1.0;
  for (int i = a0 + t; i < a0 + n; i += nt)
    for (int q = 0; q < 3; q++) {
      const double d = sc * dr[3 * i + q];
      dr[3 * i + q] = d;
      xin[3 * i + q] = x[3 * i + q] + d;
      f0[3 * i + q] = F[3 * i + q];
    }
  __syncthreads();  // rho_sh
  for (int k = t; k < m; k += nt) rs[k] = rho_sh[k];
  if (t == 0) {
    st[0] = gam; st[1] = (double)head; st[2] = (double)cnt; st[4] = steps + 1.0;
  }
}

// The same step for batches whose structures have at most one atom per thread (n <= blockDim.x): the thread's
// forces, last step, previous forces and recursion vector stay in registers, and each ring slot's step and force
// change are loaded one slot ahead of the sum that needs them, so the recursion's chain of block sums is not also a
// chain of memory round trips.  The same recursion and sums as k_lbfgs (the compiler may contract a product and a
// sum into one FMA differently, so the two kernels can differ in the last bits; each is deterministic).
__device__ __forceinline__ void lb_load3(const double* __restrict__ p, bool on, double* v) {
  v[0] = on ? p[0] : 0.0; v[1] = on ? p[1] : 0.0; v[2] = on ? p[2] : 0.0;
}
__global__ __launch_bounds__(256) void k_lbfgs1(int S, const mg_struct_dev* __restrict__ str, const double* __restrict__ F,
                                                double* __restrict__ dr, double* __restrict__ f0, double* __restrict__ hs,
                                                double* __restrict__ hy, double* __restrict__ rho, int m, size_t hstride,
                                                const double* __restrict__ x, double* __restrict__ xin,
                                                double* __restrict__ state, double ftol, double maxstep, double h0) {
  __shared__ double red[4];
  __shared__ double a_sh[LB_MAXM], rho_sh[LB_MAXM];
  const int s = blockIdx.x, t = threadIdx.x, nt = blockDim.x, nw = nt >> 6;
  const int a0 = str[s].atom0, n = str[s].natom;
  const bool on = t < n;
  const int i = a0 + (on ? t : 0);
  double* st = state + 5 * s;
  if (st[3] != 0.0) {
    if (on) { dr[3 * i] = 0.0; dr[3 * i + 1] = 0.0; dr[3 * i + 2] = 0.0; }
    return;
  }
  double f[3], d0[3], fo[3];
  lb_load3(F + 3 * i, on, f);
  lb_load3(dr + 3 * i, on, d0);
  lb_load3(f0 + 3 * i, on, fo);
  const double fm = lb_bmax(f[0] * f[0] + f[1] * f[1] + f[2] * f[2], red, t, nw);
  if (fm < ftol * ftol) {
    if (on) {
      for (int q = 0; q < 3; q++) { xin[3 * i + q] = x[3 * i + q]; dr[3 * i + q] = 0.0; }
    }
    if (t == 0) st[3] = 1.0;
    return;
  }
  double gam = st[0];
  int head = (int)st[1], cnt = (int)st[2];
  const double steps = st[4];
  double* rs = rho + (size_t)s * m;
  for (int k = t; k < m; k += nt) rho_sh[k] = rs[k];
  if (steps > 0.0) {
    double y[3] = {fo[0] - f[0], fo[1] - f[1], fo[2] - f[2]};
    double ys = 0.0, yy = 0.0, ss = 0.0;
    for (int q = 0; q < 3; q++) { ys += y[q] * d0[q]; yy += y[q] * y[q]; ss += d0[q] * d0[q]; }
    ys = lb_bsum(ys, red, t, nw);
    yy = lb_bsum(yy, red, t, nw);
    ss = lb_bsum(ss, red, t, nw);
    if (yy > 0.0 && ys > 1e-10 * sqrt(yy * ss)) {
      if (on) {
        double* ps = hs + (size_t)head * hstride + 3 * i;
        double* py = hy + (size_t)head * hstride + 3 * i;
        for (int q = 0; q < 3; q++) { ps[q] = d0[q]; py[q] = y[q]; }
      }
      __syncthreads();
      if (t == 0) rho_sh[head] = 1.0 / ys;
      head = head + 1 == m ? 0 : head + 1;
      cnt = cnt < m ? cnt + 1 : m;
      gam = ys / yy;
    }
  }
  __syncthreads();
  // two-loop recursion on q = F, newest pair first; the next slot's values are in flight during each block sum
  double q3[3] = {f[0], f[1], f[2]};
  double sv[3], yv[3], sn[3], yn[3];
  {
    int slot = head - 1;
    if (slot < 0) slot += m;
    lb_load3(hs + (size_t)slot * hstride + 3 * i, on && cnt > 0, sn);
    lb_load3(hy + (size_t)slot * hstride + 3 * i, on && cnt > 0, yn);
  }
  for (int k = 0; k < cnt; k++) {
    int slot = head - 1 - k;
    if (slot < 0) slot += m;
    for (int q = 0; q < 3; q++) { sv[q] = sn[q]; yv[q] = yn[q]; }
    int nslot = slot - 1;
    if (nslot < 0) nslot += m;
    lb_load3(hs + (size_t)nslot * hstride + 3 * i, on && k + 1 < cnt, sn);
    lb_load3(hy + (size_t)nslot * hstride + 3 * i, on && k + 1 < cnt, yn);
    double dd = 0.0;
    for (int q = 0; q < 3; q++) dd += sv[q] * q3[q];
    const double a = rho_sh[slot] * lb_bsum(dd, red, t, nw);
    if (t == 0) a_sh[k] = a;
    for (int q = 0; q < 3; q++) q3[q] -= a * yv[q];
  }
  const double g = cnt > 0 ? gam : h0;
  for (int q = 0; q < 3; q++) q3[q] *= g;
  __syncthreads();  // a_sh
  {
    int slot = head - cnt;
    if (slot < 0) slot += m;
    lb_load3(hs + (size_t)slot * hstride + 3 * i, on && cnt > 0, sn);
    lb_load3(hy + (size_t)slot * hstride + 3 * i, on && cnt > 0, yn);
  }
  for (int k = cnt - 1; k >= 0; k--) {
    int slot = head - 1 - k;
    if (slot < 0) slot += m;
    for (int q = 0; q < 3; q++) { sv[q] = sn[q]; yv[q] = yn[q]; }
    const int nslot = slot + 1 == m ? 0 : slot + 1;
    lb_load3(hs + (size_t)nslot * hstride + 3 * i, on && k > 0, sn);
    lb_load3(hy + (size_t)nslot * hstride + 3 * i, on && k > 0, yn);
    double dd = 0.0;
    for (int q = 0; q < 3; q++) dd += yv[q] * q3[q];
    const double c = a_sh[k] - rho_sh[slot] * lb_bsum(dd, red, t, nw);
    for (int q = 0; q < 3; q++) q3[q] += c * sv[q];
  }
  double pf = 0.0;
  for (int q = 0; q < 3; q++) pf += q3[q] * f[q];
  pf = lb_bsum(pf, red, t, nw);
  if (!(pf > 0.0)) {
    cnt = 0;
    gam = h0;
    for (int q = 0; q < 3; q++) q3[q] = h0 * f[q];
  }
  const double l2 = lb_bmax(q3[0] * q3[0] + q3[1] * q3[1] + q3[2] * q3[2], red, t, nw);
  const double sc = l2 > maxstep * maxstep ? maxstep / sqrt(l2) : 1.0;
  if (on)
    for (int q = 0; q < 3; q++) {
      const double d = sc * q3[q];
      dr[3 * i + q] = d;
      xin[3 * i + q] = x[3 * i + q] + d;
      f0[3 * i + q] = f[q];
    }
  __syncthreads();  // rho_sh
  for (int k = t; k < m; k += nt) rs[k] = rho_sh[k];
  if (t == 0) {
    st[0] = gam; st[1] = (double)head; st[2] = (double)cnt; st[4] = steps + 1.0;
  }
}

__global__ void k_lbfgs_init(int S, double* __restrict__ state, double h0) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double* st = state + 5 * s;
  st[0] = h0; st[1] = 0.0; st[2] = 0.0; st[3] = 0.0; st[4] = 0.0;
}

// out[k * out_stride + r * width + w] = in[k * in_stride + src[r] * width + w] for k < nslab, r < nrows (the
// optimiser's per-atom and per-structure arrays carried over when a batch is re-packed)
__global__ void k_gather_rows(int nrows, const int* __restrict__ src, int width, int nslab, size_t in_stride,
                              size_t out_stride, const double* __restrict__ in, double* __restrict__ out) {
  const size_t per = (size_t)nrows * width;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per * nslab) return;
  const size_t k = idx / per, e = idx - k * per;
  const int r = (int)(e / width), w = (int)(e - (size_t)r * width);
  out[k * out_stride + (size_t)r * width + w] = in[k * in_stride + (size_t)src[r] * width + w];
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

int mgk_lbfgs_init(int S, double* state, double h0, hipStream_t st) {
  hipLaunchKernelGGL(k_lbfgs_init, dim3((S + 255) / 256), dim3(256), 0, st, S, state, h0);
  return mgk_status();
}

// block of one wave when every structure has at most 64 atoms (no barriers in the sums), else four waves; the
// register form when no structure has more atoms than the block has threads
int mgk_lbfgs(int S, int maxn, const mg_struct_dev* str, const double* F, double* dr, double* f0, double* hs, double* hy,
              double* rho, int m, size_t hstride, const double* x, double* xin, double* state, double fmax,
              double maxstep, double h0, hipStream_t st) {
  if (m < 1 || m > LB_MAXM) return -1;
  const int nt = maxn <= 64 ? 64 : 256;
  if (maxn <= nt)  // one atom per thread: registers and prefetched ring slots
    hipLaunchKernelGGL(k_lbfgs1, dim3(S), dim3(nt), 0, st, S, str, F, dr, f0, hs, hy, rho, m, hstride, x, xin, state,
                       fmax, maxstep, h0);
  else
    hipLaunchKernelGGL(k_lbfgs, dim3(S), dim3(nt), 0, st, S, str, F, dr, f0, hs, hy, rho, m, hstride, x, xin, state,
                       fmax, maxstep, h0);
  return mgk_status();
}

int mgk_gather_rows(int nrows, const int* src, int width, int nslab, size_t in_stride, size_t out_stride,
                    const double* in, double* out, hipStream_t st) {
  const size_t tot = (size_t)nrows * width * nslab;
  if (tot == 0) return 0;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nrows, src, width, nslab,
                     in_stride, out_stride, in, out);
  return mgk_status();
}

}  // extern "C"
