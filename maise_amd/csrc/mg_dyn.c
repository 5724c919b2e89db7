/*
 * mg_dyn.c -- host sequencing of the device-resident dynamics (mg_dyn.hip): NVT molecular dynamics and
 * batched FIRE relaxation.  Positions, velocities, forces and integrator state stay on the device; each step
 * is a fixed chain of launches on the context stream.
 *
 * Units: A, fs, eV.  Masses m = atomic mass (amu) * 103.6427 eV fs^2/A^2 and kB = 1/11604 eV/K, the
 * reference's constants (src/util.c:11,984 via CAST_CELL src/cmod.c:1128-1133; src/cmod.c:5).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mg_host.h"

static const double MG_AMU[96] = { /* atomic masses (amu), src/util.c:11 */
  0.0,1.008,4.002602,6.94,9.0121831,10.81,12.011,14.007,15.999,18.99840316,20.1797,22.98976928,24.305,26.9815384,
  28.085,30.973762,32.06,35.45,39.948,39.0983,40.078,44.955908,47.867,50.9415,51.9961,54.938043,55.845,58.933194,
  58.6934,63.546,65.38,69.723,72.63,74.921595,78.971,79.904,83.798,85.4678,87.62,88.90584,91.224,92.90637,95.95,98,
  101.07,102.90549,106.42,107.8682,112.414,114.818,118.71,121.76,127.6,126.90447,131.293,132.9054519,137.327,
  138.90547,140.116,140.90766,144.242,145,150.36,151.964,157.25,158.925354,162.5,164.930328,167.259,168.934218,
  173.045,174.9668,178.49,180.94788,183.84,186.207,190.23,192.217,195.084,196.96657,200.592,204.38,207.2,208.9804,
  209,210,222,223,226,227,232.0377,231.03588,238.02891,237,244,243};
#define MG_MASS_UNIT 103.6427
#define MG_KB (1.0 / 11604.0)

int mgk_md_maxwell(int N, unsigned long long seed, const double *mass, double kT, double *v, hipStream_t st);
int mgk_md_fix(int N, const double *mass, double *v, double kT, double ndof, hipStream_t st);
int mgk_md_thermo(int N, const double *mass, const double *v, double *dyn, double dt, double kT, double ndof, double tau,
                  double *trace, int slot, hipStream_t st);
int mgk_md_kick_drift(int N, double dt, const double *mass, const double *F, double *v, const double *x, double *xin,
                      const double *dyn, hipStream_t st);
int mgk_md_kick_scale(int N, double dt, const double *mass, const double *F, double *v, const double *dyn, double tau,
                      hipStream_t st);
int mgk_copy_d(const double *src, double *dst, int n, hipStream_t st);
int mgk_fire_init(int S, double *state, double dt0, hipStream_t st);
int mgk_fire(int S, const mg_struct_dev *str, const double *F, double *v, const double *x, double *xin, double *state,
             double fmax, double dtmax, double maxstep, hipStream_t st);
int mgk_fire_count(int S, const double *state, int *out, hipStream_t st);

#define KCHK(x) do { if ((x) != 0) return mg_fail(MG_ERR_HIP, "kernel launch failed: %s", #x); } while (0)

static int ensure_dyn(mg_ctx *c) {
  int r;
  if (c->cap_dyn_atoms < (size_t)c->N || !c->d_v) {
    size_t n = c->N + 16;
    if ((r = mg_dev_alloc((void **)&c->d_v, n * 3 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_mass, n * sizeof(double))))
      return r;
    c->cap_dyn_atoms = n;
  }
  if (c->cap_dyn_str < (size_t)c->S || !c->d_fire) {
    size_t n = c->S + 4;
    if ((r = mg_dev_alloc((void **)&c->d_fire, n * 5 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_dyn, 16 * sizeof(double))))
      return r;
    c->cap_dyn_str = n;
  }
  /* per-atom masses from the species of the current structures */
  int *spc = (int *)malloc(sizeof(int) * (c->N > 0 ? c->N : 1));
  double *mass = (double *)malloc(sizeof(double) * (c->N > 0 ? c->N : 1));
  HIPCHK(hipMemcpy(spc, c->d_spc, sizeof(int) * c->N, hipMemcpyDeviceToHost));
  for (int i = 0; i < c->N; i++) {
    int z = c->m->spcz[spc[i]];
    mass[i] = (z > 0 && z < 96 ? MG_AMU[z] : 1.0) * MG_MASS_UNIT;
  }
  hipError_t e = hipMemcpy(c->d_mass, mass, sizeof(double) * c->N, hipMemcpyHostToDevice);
  free(spc);
  free(mass);
  if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "mass upload: %s", hipGetErrorString(e));
  return MG_OK;
}

int mg_md_init(mg_ctx *c, double T, unsigned long long seed) {
  if (!c || c->S != 1) return mg_fail(MG_ERR_ARG, "mg_md_init: needs a context holding one structure");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if ((r = ensure_dyn(c))) return r;
  const double kT = MG_KB * T;
  KCHK(mgk_md_maxwell(c->N, seed, c->d_mass, kT, c->d_v, c->st));
  KCHK(mgk_md_fix(c->N, c->d_mass, c->d_v, kT, 3.0 * c->N, c->st));
  HIPCHK(hipMemsetAsync(c->d_dyn, 0, 16 * sizeof(double), c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return MG_OK;
}

int mg_md_set_velocities(mg_ctx *c, const double *v) {
  if (!c || !v) return mg_fail(MG_ERR_ARG, "mg_md_set_velocities: bad argument");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if ((r = ensure_dyn(c))) return r;
  HIPCHK(hipMemcpy(c->d_v, v, sizeof(double) * 3 * c->N, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_dyn, 0, 16 * sizeof(double)));
  return MG_OK;
}

int mg_md_get_velocities(mg_ctx *c, double *v) {
  if (!c || !v || !c->d_v) return mg_fail(MG_ERR_ARG, "mg_md_get_velocities: no velocities");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(v, c->d_v, sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_md_run(mg_ctx *c, int nsteps, double dt, double T, double tau, double *epot, double *ekin) {
  if (!c || c->S != 1 || nsteps < 0 || dt <= 0.0) return mg_fail(MG_ERR_ARG, "mg_md_run: bad argument");
  if (!c->d_v) return mg_fail(MG_ERR_ARG, "mg_md_run: call mg_md_init or mg_md_set_velocities first");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if ((size_t)(2 * nsteps + 2) > c->cap_trace) {
    if ((r = mg_dev_alloc((void **)&c->d_trace, sizeof(double) * (2 * nsteps + 2)))) return r;
    c->cap_trace = 2 * nsteps + 2;
  }
  const double kT = MG_KB * T, ndof = 3.0 * c->N;
  /* forces at the starting positions */
  if ((r = mg_eval(c, MG_EVAL_FORCES))) return r;
  for (int s = 0; s < nsteps; s++) {
    KCHK(mgk_md_thermo(c->N, c->d_mass, c->d_v, c->d_dyn, dt, kT, ndof, tau, NULL, -1, c->st));
    KCHK(mgk_md_kick_drift(c->N, dt, c->d_mass, c->d_F, c->d_v, c->d_x, c->d_xin, c->d_dyn, c->st));
    /* positions advanced: x_in is Cartesian from now on */
    if (c->h_str[0].xt == 0) {
      c->h_str[0].xt = 1;
      HIPCHK(hipMemcpyAsync(c->d_str, c->h_str, sizeof(mg_struct_dev), hipMemcpyHostToDevice, c->st));
    }
    if ((r = mg_eval(c, MG_EVAL_FORCES))) return r;
    KCHK(mgk_md_kick_scale(c->N, dt, c->d_mass, c->d_F, c->d_v, c->d_dyn, tau, c->st));
    KCHK(mgk_md_thermo(c->N, c->d_mass, c->d_v, c->d_dyn, dt, kT, ndof, tau, c->d_trace + nsteps, s, c->st));
    KCHK(mgk_copy_d(c->d_E, c->d_trace + s, 1, c->st));
  }
  if (nsteps > 0 && epot) HIPCHK(hipMemcpy(epot, c->d_trace, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  if (nsteps > 0 && ekin) HIPCHK(hipMemcpy(ekin, c->d_trace + nsteps, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  HIPCHK(hipStreamSynchronize(c->st));
  return MG_OK;
}

int mg_relax(mg_ctx *c, int max_steps, double fmax, double dt0, double dtmax, double maxstep, int *steps, int *done) {
  if (!c || c->S < 1 || max_steps < 0 || fmax <= 0.0) return mg_fail(MG_ERR_ARG, "mg_relax: bad argument");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if ((r = ensure_dyn(c))) return r;
  HIPCHK(hipMemsetAsync(c->d_v, 0, sizeof(double) * 3 * c->N, c->st));
  KCHK(mgk_fire_init(c->S, c->d_fire, dt0, c->st));
  int ndone = 0, it = 0;
  for (it = 0; it < max_steps; it++) {
    if ((r = mg_eval(c, MG_EVAL_FORCES))) return r;
    KCHK(mgk_fire(c->S, c->d_str, c->d_F, c->d_v, c->d_x, c->d_xin, c->d_fire, fmax, dtmax, maxstep, c->st));
    if (it == 0) {
      int any_frac = 0;
      for (int s = 0; s < c->S; s++) if (c->h_str[s].xt == 0) { c->h_str[s].xt = 1; any_frac = 1; }
      if (any_frac) HIPCHK(hipMemcpyAsync(c->d_str, c->h_str, sizeof(mg_struct_dev) * c->S, hipMemcpyHostToDevice, c->st));
    }
    if ((it % 8) == 7) {
      HIPCHK(hipMemsetAsync(c->d_misc + 2, 0, sizeof(int), c->st));
      KCHK(mgk_fire_count(c->S, c->d_fire, c->d_misc + 2, c->st));
      HIPCHK(hipMemcpyAsync(&ndone, c->d_misc + 2, sizeof(int), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      if (ndone == c->S) break;
    }
  }
  /* energies and forces at the final positions */
  if ((r = mg_eval(c, MG_EVAL_FORCES))) return r;
  if (steps || done) {
    double *state = (double *)malloc(sizeof(double) * 5 * c->S);
    hipError_t e = hipMemcpy(state, c->d_fire, sizeof(double) * 5 * c->S, hipMemcpyDeviceToHost);
    for (int s = 0; s < c->S && e == hipSuccess; s++) {
      if (steps) steps[s] = (int)state[5 * s + 4];
      if (done) done[s] = state[5 * s + 3] != 0.0;
    }
    free(state);
    if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "mg_relax: %s", hipGetErrorString(e));
  }
  return MG_OK;
}
