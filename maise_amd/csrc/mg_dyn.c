/*
 * mg_dyn.c -- host sequencing of the device-resident dynamics (mg_dyn.hip): NVT molecular dynamics and
 * batched FIRE and L-BFGS relaxation.  Positions, velocities, forces and integrator state stay on the device; each step
 * is a fixed chain of launches on the context stream.
 *
 * Units: A, fs, eV.  Masses m = atomic mass (amu) * 103.6427 eV fs^2/A^2 and kB = 1/11604 eV/K, the
 * reference's constants (src/util.c:11,984 via CAST_CELL src/cmod.c:1128-1133; src/cmod.c:5).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mg_host.h"

const double MG_AMU_TABLE[96] = { /* atomic masses (amu), src/util.c:11 */
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
                      const double *dyn, int baro, hipStream_t st);
int mgk_md_baro(mg_struct_dev *str, const double *U, double *dyn, double p, double ktt_dt, hipStream_t st);
int mgk_md_kick_scale(int N, double dt, const double *mass, const double *F, double *v, const double *dyn, double tau,
                      hipStream_t st);
int mgk_copy_d(const double *src, double *dst, int n, hipStream_t st);
int mgk_md_record(const double *E, double *dyn, double *trace, int nsteps, int baro, hipStream_t st);
int mgk_fire_init(int S, double *state, double dt0, hipStream_t st);
int mgk_fire(int S, const mg_struct_dev *str, const double *F, double *v, const double *x, double *xin, double *state,
             double fmax, double dtmax, double maxstep, hipStream_t st);
int mgk_fire_count(int S, const double *state, int *out, hipStream_t st);
int mgk_fire_cell(int S, mg_struct_dev *str, const double *F, const double *U, double *v, double *cellv,
                  const double *x, double *xin, double *state, double p, double fmax, double smax, double dtmax,
                  double maxstep, hipStream_t st);
int mgk_lbfgs_init(int S, double *state, double h0, hipStream_t st);
int mgk_lbfgs(int S, int maxn, const mg_struct_dev *str, const double *F, double *dr, double *f0, double *hs, double *hy,
              double *rho, int m, size_t hstride, const double *x, double *xin, double *state, double fmax,
              double maxstep, double h0, hipStream_t st);
int mgk_gather_rows(int nrows, const int *src, int width, int nslab, size_t in_stride, size_t out_stride,
                    const double *in, double *out, hipStream_t st);

#define KCHK(x) do { if ((x) != 0) return mg_fail(MG_ERR_HIP, "kernel launch failed: %s (%s)", #x, mgk_last_error()); } while (0)

typedef int (*step_fn)(mg_ctx *c, void *arg);
static int graph_capture(mg_ctx *c, step_fn head, step_fn tail, void *arg, hipGraphExec_t *exec);
static int snapshot(mg_ctx *c, int restore);
#define MG_GRAPH_CHECK 8

static int ensure_dyn(mg_ctx *c) {
  int r;
  if (c->cap_dyn_atoms < (size_t)c->N || !c->d_v) {
    size_t n = c->N + 16;
    if ((r = mg_dev_alloc((void **)&c->d_v, n * 3 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_mass, n * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_x, n * 3 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_v, n * 3 * sizeof(double))))
      return r;
    c->cap_dyn_atoms = n;
  }
  if (c->cap_dyn_str < (size_t)c->S || !c->d_fire) {
    size_t n = c->S + 4;
    if ((r = mg_dev_alloc((void **)&c->d_fire, n * 5 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_fire, n * 5 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_cellv, n * 9 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_cellv, n * 9 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_str, n * sizeof(mg_struct_dev))) ||
        (r = mg_dev_alloc((void **)&c->d_dyn, 16 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_dyn, 16 * sizeof(double))))
      return r;
    c->cap_dyn_str = n;
  }
  /* per-atom masses from the species of the current structures */
  int *spc = (int *)malloc(sizeof(int) * (c->N > 0 ? c->N : 1));
  double *mass = (double *)malloc(sizeof(double) * (c->N > 0 ? c->N : 1));
  HIPCHK(hipMemcpy(spc, c->d_spc, sizeof(int) * c->N, hipMemcpyDeviceToHost));
  for (int i = 0; i < c->N; i++) {
    int z = c->m->spcz[spc[i]];
    mass[i] = (z > 0 && z < 96 ? MG_AMU_TABLE[z] : 1.0) * MG_MASS_UNIT;
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
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
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
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  int r;
  if ((r = ensure_dyn(c))) return r;
  HIPCHK(hipMemcpy(c->d_v, v, sizeof(double) * 3 * c->N, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_dyn, 0, 16 * sizeof(double)));
  return MG_OK;
}

int mg_md_get_velocities(mg_ctx *c, double *v) {
  if (!c || !v || !c->d_v) return mg_fail(MG_ERR_ARG, "mg_md_get_velocities: no velocities");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  HIPCHK(hipMemcpy(v, c->d_v, sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost));
  return MG_OK;
}

typedef struct {
  double dt, kT, ndof, tau;
  int nsteps;
  int baro;          /* Berendsen barostat on */
  double p, ktt_dt;  /* target pressure (eV/A^3), coupling x dt (1/GPa) */
} md_arg;

/* velocity Verlet with the reference's Nose-Hoover xi equation, split around the force evaluation */
static int md_head(mg_ctx *c, void *arg) {
  const md_arg *a = (const md_arg *)arg;
  KCHK(mgk_md_thermo(c->N, c->d_mass, c->d_v, c->d_dyn, a->dt, a->kT, a->ndof, a->tau, NULL, -1, c->st));
  KCHK(mgk_md_kick_drift(c->N, a->dt, c->d_mass, c->d_F, c->d_v, c->d_x, c->d_xin, c->d_dyn, a->baro, c->st));
  return MG_OK;
}

static int md_tail(mg_ctx *c, void *arg) {
  const md_arg *a = (const md_arg *)arg;
  KCHK(mgk_md_kick_scale(c->N, a->dt, c->d_mass, c->d_F, c->d_v, c->d_dyn, a->tau, c->st));
  KCHK(mgk_md_thermo(c->N, c->d_mass, c->d_v, c->d_dyn, a->dt, a->kT, a->ndof, a->tau, NULL, -1, c->st));
  if (a->baro) KCHK(mgk_md_baro(c->d_str, c->d_U, c->d_dyn, a->p, a->ktt_dt, c->st));
  KCHK(mgk_md_record(c->d_E, c->d_dyn, c->d_trace, a->nsteps, a->baro, c->st));  /* step counter dyn[3] */
  return MG_OK;
}

/* one step without a graph, with a synchronised evaluation */
static int md_step_plain(mg_ctx *c, md_arg *a) {
  int r;
  if ((r = md_head(c, a))) return r;
  if (c->h_str[0].xt == 0) {  /* positions advanced: x_in is Cartesian from now on */
    c->h_str[0].xt = 1;
    HIPCHK(hipMemcpyAsync(c->d_str, c->h_str, sizeof(mg_struct_dev), hipMemcpyHostToDevice, c->st));
  }
  if ((r = mgh_eval(c, MG_EVAL_FORCES, 0))) return r;
  return md_tail(c, a);
}

/* NVT (tau > 0) or NVE steps, optionally with the Berendsen barostat (NPT / isobaric).  With graphs enabled the
 * first step runs synchronised (sizing the buffers), then one step is captured and replayed MG_GRAPH_CHECK steps at
 * a time; after each block the host reads the overflow flag and, when a row or
 * the edge total outgrew the capture, restores the checkpoint taken before the block, evaluates synchronised and
 * captures again, so the trajectory does not depend on the buffer sizes. */
static int md_impl(mg_ctx *c, int nsteps, double dt, double T, double tau, int baro, double p, double ktt, double *epot,
                   double *ekin, double *vol, double *pint) {
  if (!c || c->S != 1 || nsteps < 0 || dt <= 0.0) return mg_fail(MG_ERR_ARG, "mg_md_run: bad argument");
  if (!c->d_v) return mg_fail(MG_ERR_ARG, "mg_md_run: call mg_md_init or mg_md_set_velocities first");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  int r;
  const size_t ntr = (size_t)(baro ? 4 : 2) * nsteps + 2;
  if (ntr > c->cap_trace) {
    if ((r = mg_dev_alloc((void **)&c->d_trace, sizeof(double) * ntr))) return r;
    c->cap_trace = ntr;
  }
  md_arg a = {dt, MG_KB * T, 3.0 * c->N, tau, nsteps, baro, p, ktt * dt};
  c->skin_on = !baro && c->skin > 0.0;  /* a fixed cell: skin lists hold */
  c->skin_need = 0;
  if (baro) {  /* lambda = 1 until the first evaluation has set it */
    const double one = 1.0;
    HIPCHK(hipMemcpy(c->d_dyn + 4, &one, sizeof(double), hipMemcpyHostToDevice));
  }
  /* forces at the starting positions; step counter to 0 */
  if ((r = mgh_eval(c, MG_EVAL_FORCES, 0))) return r;
  HIPCHK(hipMemsetAsync(c->d_dyn + 3, 0, sizeof(double), c->st));
  hipGraphExec_t gx = NULL;
  int s = 0, s_ck = 0, need_e = 0, need_r = 0, flags[8];
#define MCHK(x) do { if ((r = (x)) != MG_OK) goto out; } while (0)
#define MHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { r = mg_fail(MG_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); goto out; } } while (0)
  while (s < nsteps) {
    if (!c->graph_enable || !gx) {
      MCHK(md_step_plain(c, &a));
      s++;
      if (c->graph_enable && s < nsteps) {
        MCHK(mgh_autotune_once(c));  /* evaluations at the same positions: the state does not change */
        MCHK(mgh_graph_caps(c, c->nedges > need_e ? c->nedges : need_e, c->maxrow > need_r ? c->maxrow : need_r));
        MCHK(graph_capture(c, md_head, md_tail, &a, &gx));
        MCHK(snapshot(c, 0));
        s_ck = s;
        if (mgk_clear(c->d_misc + 3, 3, NULL, 0, NULL, 0, c->st)) { r = mg_fail(MG_ERR_HIP, "misc clear failed"); goto out; }
      }
      continue;
    }
    const int nrun = nsteps - s < MG_GRAPH_CHECK ? nsteps - s : MG_GRAPH_CHECK;
    for (int k = 0; k < nrun; k++) MHIP(hipGraphLaunch(gx, c->st));
    MHIP(hipMemcpyAsync(flags, c->d_misc, sizeof flags, hipMemcpyDeviceToHost, c->st));
    MHIP(hipStreamSynchronize(c->st));
    if (flags[3]) {
      /* restore the checkpoint; the next pass evaluates it synchronised and captures with larger buffers */
      if (flags[7] > c->skin_need) c->skin_need = flags[7];
      MCHK(snapshot(c, 1));
      s = s_ck;
      need_e = flags[4];
      need_r = flags[5];
      hipGraphExecDestroy(gx);
      gx = NULL;
      MCHK(mgh_eval(c, MG_EVAL_FORCES, 0));  /* forces at the restored positions */
      continue;
    }
    if (flags[1]) { r = mg_fail(MG_ERR_NEIGHBOR, "neighbour list error: a reverse edge is missing, or a cell is too thin for the cutoff"); goto out; }
    s += nrun;
    c->nedges = -1;  /* padded rows: summed when asked for */
    c->maxrow = flags[0];
    MCHK(snapshot(c, 0));
    s_ck = s;
    if (mgk_clear(c->d_misc + 3, 3, NULL, 0, NULL, 0, c->st)) { r = mg_fail(MG_ERR_HIP, "misc clear failed"); goto out; }
  }
  if (nsteps > 0 && epot) MHIP(hipMemcpy(epot, c->d_trace, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  if (nsteps > 0 && ekin) MHIP(hipMemcpy(ekin, c->d_trace + nsteps, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  if (baro && nsteps > 0 && vol) MHIP(hipMemcpy(vol, c->d_trace + 2 * nsteps, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  if (baro && nsteps > 0 && pint) MHIP(hipMemcpy(pint, c->d_trace + 3 * nsteps, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  MHIP(hipStreamSynchronize(c->st));
  if (baro) {  /* the lattice moved on the device: bring the host tables up to date */
    MHIP(hipMemcpy(c->h_str, c->d_str, sizeof(mg_struct_dev), hipMemcpyDeviceToHost));
    memcpy(c->h_lat, c->h_str[0].L, 9 * sizeof(double));
  }
  r = MG_OK;
out:
  if (gx) hipGraphExecDestroy(gx);
  c->skin_on = 0;
  return r;
#undef MCHK
#undef MHIP
}

int mg_md_run(mg_ctx *c, int nsteps, double dt, double T, double tau, double *epot, double *ekin) {
  return md_impl(c, nsteps, dt, T, tau, 0, 0.0, 0.0, epot, ekin, NULL, NULL);
}

int mg_md_run_npt(mg_ctx *c, int nsteps, double dt, double T, double tau, double p_gpa, double ktt, double *epot,
                  double *ekin, double *vol, double *pint_gpa) {
  const double ev2gpa = 160.2176565; /* eV/A^3 in GPa, the reference's eV2GPa */
  if (!c || ktt < 0.0 || (c->S == 1 && c->h_str[0].nd != 3)) return mg_fail(MG_ERR_ARG, "mg_md_run_npt: bad argument (periodic cell and ktt >= 0 needed)");
  int r = md_impl(c, nsteps, dt, T, tau, 1, p_gpa / ev2gpa, ktt, epot, ekin, vol, pint_gpa);
  if (r == MG_OK && pint_gpa)
    for (int s = 0; s < nsteps; s++) pint_gpa[s] *= ev2gpa;
  return r;
}

/* Batched FIRE.  Every 8 iterations the converged structures are counted on the device; once the atoms of the
 * structures still moving fall to half of those being evaluated, the finished ones are copied out and the batch
 * is re-packed to the unconverged structures (positions, velocities and FIRE state carried over), so the tail
 * of the relaxation no longer evaluates structures that have stopped.  At the end the full batch is restored
 * with the final positions and evaluated once more, so mg_get / mg_get_batch see every structure. */
/* ---- graph mode: one iteration (evaluation chain + integrator step) captured once and replayed; the
 * edge buffers are sized from the last synchronised evaluation with headroom, the kernels raise an overflow
 * flag instead of overrunning them, and every MG_GRAPH_CHECK replays the host checks the flags and, on
 * overflow, restores the state saved at the previous check, grows the buffers and re-captures. */
static int graph_capture(mg_ctx *c, step_fn head, step_fn tail, void *arg, hipGraphExec_t *exec) {
  hipGraph_t g = NULL;
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int r = head ? head(c, arg) : MG_OK;
  if (r == MG_OK) r = mgh_eval_chain(c, 1, MG_CHAIN_GRAPH);
  if (r == MG_OK && tail) r = tail(c, arg);
  hipError_t e = hipStreamEndCapture(c->st, &g);
  if (r != MG_OK) {
    if (e == hipSuccess && g) hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(exec, g, NULL, NULL, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  return MG_OK;
}

/* save / restore positions, velocities, per-structure FIRE state and thermostat scalars */
static int snapshot(mg_ctx *c, int restore) {
  const size_t nx = sizeof(double) * 3 * c->N, nf = sizeof(double) * 5 * c->S;
  hipMemcpyKind k = hipMemcpyDeviceToDevice;
  if (!restore) {
    HIPCHK(hipMemcpyAsync(c->d_snap_x, c->d_xin, nx, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_snap_v, c->d_v, nx, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_snap_fire, c->d_fire, nf, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_snap_dyn, c->d_dyn, 16 * sizeof(double), k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_snap_cellv, c->d_cellv, sizeof(double) * 9 * c->S, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_snap_str, c->d_str, sizeof(mg_struct_dev) * c->S, k, c->st));
    if (c->lb_on) {
      const size_t nr = sizeof(double) * c->lb_m * c->lb_stride;
      HIPCHK(hipMemcpyAsync(c->d_snap_lb_s, c->d_lb_s, nr, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_snap_lb_y, c->d_lb_y, nr, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_snap_lb_f0, c->d_lb_f0, nx, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_snap_lb_rho, c->d_lb_rho, sizeof(double) * c->lb_m * c->S, k, c->st));
    }
  } else {
    HIPCHK(hipMemcpyAsync(c->d_xin, c->d_snap_x, nx, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_v, c->d_snap_v, nx, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_fire, c->d_snap_fire, nf, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_dyn, c->d_snap_dyn, 16 * sizeof(double), k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_cellv, c->d_snap_cellv, sizeof(double) * 9 * c->S, k, c->st));
    HIPCHK(hipMemcpyAsync(c->d_str, c->d_snap_str, sizeof(mg_struct_dev) * c->S, k, c->st));
    if (c->lb_on) {
      const size_t nr = sizeof(double) * c->lb_m * c->lb_stride;
      HIPCHK(hipMemcpyAsync(c->d_lb_s, c->d_snap_lb_s, nr, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_lb_y, c->d_snap_lb_y, nr, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_lb_f0, c->d_snap_lb_f0, nx, k, c->st));
      HIPCHK(hipMemcpyAsync(c->d_lb_rho, c->d_snap_lb_rho, sizeof(double) * c->lb_m * c->S, k, c->st));
    }
  }
  return MG_OK;
}

/* L-BFGS buffers for the current batch with m pairs per structure (laid out for 3N; compaction only shrinks N) */
static int ensure_lbfgs(mg_ctx *c, int m) {
  int r;
  const size_t ring = (size_t)m * 3 * (size_t)c->N + 16;
  if (c->cap_lb_ring < ring || !c->d_lb_s) {
    if ((r = mg_dev_alloc((void **)&c->d_lb_s, ring * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_lb_y, ring * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_lb_tmp, ring * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_lb_s, ring * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_lb_y, ring * sizeof(double))))
      return r;
    c->cap_lb_ring = ring;
  }
  if (c->cap_lb_atoms < (size_t)c->N + 16 || !c->d_lb_f0) {
    const size_t n = (size_t)c->N + 16;
    if ((r = mg_dev_alloc((void **)&c->d_lb_f0, n * 3 * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_lb_f0, n * 3 * sizeof(double))))
      return r;
    c->cap_lb_atoms = n;
  }
  if (c->cap_lb_str < (size_t)m * (c->S + 4) || !c->d_lb_rho) {
    const size_t n = (size_t)m * (c->S + 4);
    if ((r = mg_dev_alloc((void **)&c->d_lb_rho, n * sizeof(double))) ||
        (r = mg_dev_alloc((void **)&c->d_snap_lb_rho, n * sizeof(double))))
      return r;
    c->cap_lb_str = n;
  }
  /* index maps of a re-pack: N + S ints at most (the map buffer is re-sized with the ring) */
  if ((r = mg_dev_alloc((void **)&c->d_lb_map, ((size_t)c->N + (size_t)c->S + 16) * sizeof(int)))) return r;
  c->lb_m = m;
  c->lb_stride = 3 * (size_t)c->N;
  return MG_OK;
}

/* After a re-pack: carry the L-BFGS rings, previous forces and 1/(y.s) of the kept structures into the new layout.
 * amap[i'] = the old index of new atom i' (N1 of them), smap[k'] = the old slot of new structure k' (S1). */
static int lbfgs_repack(mg_ctx *c, const int *amap, int N1, const int *smap, int S1, size_t old_stride) {
  const int m = c->lb_m;
  int *d_amap = c->d_lb_map, *d_smap = c->d_lb_map + N1;
  HIPCHK(hipMemcpyAsync(d_amap, amap, sizeof(int) * N1, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(d_smap, smap, sizeof(int) * S1, hipMemcpyHostToDevice, c->st));
  const size_t new_stride = 3 * (size_t)N1;
  double *t;
  KCHK(mgk_gather_rows(N1, d_amap, 3, m, old_stride, new_stride, c->d_lb_s, c->d_lb_tmp, c->st));
  t = c->d_lb_s; c->d_lb_s = c->d_lb_tmp; c->d_lb_tmp = t;
  KCHK(mgk_gather_rows(N1, d_amap, 3, m, old_stride, new_stride, c->d_lb_y, c->d_lb_tmp, c->st));
  t = c->d_lb_y; c->d_lb_y = c->d_lb_tmp; c->d_lb_tmp = t;
  KCHK(mgk_gather_rows(N1, d_amap, 3, 1, 0, 0, c->d_lb_f0, c->d_lb_tmp, c->st));
  HIPCHK(hipMemcpyAsync(c->d_lb_f0, c->d_lb_tmp, sizeof(double) * 3 * N1, hipMemcpyDeviceToDevice, c->st));
  KCHK(mgk_gather_rows(S1, d_smap, m, 1, 0, 0, c->d_lb_rho, c->d_lb_tmp, c->st));
  HIPCHK(hipMemcpyAsync(c->d_lb_rho, c->d_lb_tmp, sizeof(double) * m * S1, hipMemcpyDeviceToDevice, c->st));
  /* the maps are read by the gathers above before the host arrays go: wait here */
  HIPCHK(hipStreamSynchronize(c->st));
  c->lb_stride = new_stride;
  return MG_OK;
}

typedef struct {
  double fmax, dtmax, maxstep;
  int cell;            /* also relax the lattice (enthalpy at pressure p) */
  double p, smax;      /* eV/A^3 */
  int lbfgs;           /* L-BFGS (positions only) instead of FIRE */
  double h0;           /* its initial inverse Hessian (A^2/eV) */
  int maxn;            /* most atoms in one structure of the batch (block size of the L-BFGS kernel) */
} fire_arg;

static int fire_tail(mg_ctx *c, void *arg) {
  const fire_arg *a = (const fire_arg *)arg;
  if (a->lbfgs) {
    const int e = mgk_lbfgs(c->S, a->maxn, c->d_str, c->d_F, c->d_v, c->d_lb_f0, c->d_lb_s, c->d_lb_y, c->d_lb_rho,
                            c->lb_m, c->lb_stride, c->d_x, c->d_xin, c->d_fire, a->fmax, a->maxstep, a->h0, c->st);
    return e ? mg_fail(MG_ERR_HIP, "kernel launch failed: L-BFGS step (%s)", mgk_last_error()) : MG_OK;
  }
  const int e = a->cell ? mgk_fire_cell(c->S, c->d_str, c->d_F, c->d_U, c->d_v, c->d_cellv, c->d_x, c->d_xin,
                                        c->d_fire, a->p, a->fmax, a->smax, a->dtmax, a->maxstep, c->st)
                        : mgk_fire(c->S, c->d_str, c->d_F, c->d_v, c->d_x, c->d_xin, c->d_fire, a->fmax, a->dtmax,
                                   a->maxstep, c->st);
  return e ? mg_fail(MG_ERR_HIP, "kernel launch failed: FIRE step (%s)", mgk_last_error()) : MG_OK;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* one record of the relaxation log (grown on demand; a failed allocation only stops the log) */
static void rlog_add(mg_ctx *c, int it, int S, int N, double sec) {
  if (c->rlog_n == c->rlog_cap) {
    const int cap = c->rlog_cap ? 2 * c->rlog_cap : 256;
    double *p = (double *)realloc(c->rlog, sizeof(double) * 4 * (size_t)cap);
    if (!p) return;
    c->rlog = p;
    c->rlog_cap = cap;
  }
  double *r = c->rlog + 4 * (size_t)c->rlog_n++;
  r[0] = it; r[1] = S; r[2] = N; r[3] = sec;
}

int mg_relax_log(mg_ctx *c, int max, double *rec) {
  if (!c || max < 0 || (max > 0 && !rec)) return mg_fail(MG_ERR_ARG, "mg_relax_log: bad argument");
  const int n = c->rlog_n < max ? c->rlog_n : max;
  if (n) memcpy(rec, c->rlog, sizeof(double) * 4 * (size_t)n);
  return c->rlog_n;
}

typedef struct {
  int S0, N0;
  int *off0, *nd0, *spc0;
  double *lat0, *x0, *st0;
  int *act;   /* active slot -> original structure */
  int S;      /* active structures */
} relax_book;

static void book_free(relax_book *b) {
  free(b->off0); free(b->nd0); free(b->spc0); free(b->lat0); free(b->x0); free(b->st0); free(b->act);
}

/* copy the current device state of the active structures into the book; keep (re-pack) those with keep[k] */
static int compact(mg_ctx *c, relax_book *b, int repack) {
  const int S = c->S, N = c->N;
  double *x = (double *)malloc(sizeof(double) * 3 * N), *v = (double *)malloc(sizeof(double) * 3 * N);
  double *st = (double *)malloc(sizeof(double) * 5 * S), *cv = (double *)malloc(sizeof(double) * 9 * S);
  mg_struct_dev *sd = (mg_struct_dev *)malloc(sizeof(mg_struct_dev) * S);
  int r = MG_OK;
  if (hipMemcpyAsync(x, c->d_xin, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(v, c->d_v, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(st, c->d_fire, sizeof(double) * 5 * S, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(cv, c->d_cellv, sizeof(double) * 9 * S, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(sd, c->d_str, sizeof(mg_struct_dev) * S, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess) {
    r = mg_fail(MG_ERR_HIP, "mg_relax: state download failed");
    goto out;
  }
  for (int k = 0; k < S; k++) {
    const int o = b->act[k], n = b->off0[o + 1] - b->off0[o];
    memcpy(b->x0 + 3 * b->off0[o], x + 3 * c->h_atom_off[k], sizeof(double) * 3 * n);
    memcpy(b->st0 + 5 * o, st + 5 * k, sizeof(double) * 5);
    memcpy(b->lat0 + 9 * o, sd[k].L, sizeof(double) * 9);  /* the cell relaxation moves the lattices */
  }
  if (repack) {
    int S1 = 0, N1 = 0;
    for (int k = 0; k < S; k++)
      if (st[5 * k + 3] == 0.0) { b->act[S1++] = b->act[k]; }
    int *off = (int *)malloc(sizeof(int) * (S1 + 1)), *nd = (int *)malloc(sizeof(int) * S1);
    int *xt = (int *)malloc(sizeof(int) * S1), *spc = (int *)malloc(sizeof(int) * b->N0);
    double *lat = (double *)malloc(sizeof(double) * 9 * S1), *xn = (double *)malloc(sizeof(double) * 3 * b->N0);
    double *vn = (double *)malloc(sizeof(double) * 3 * b->N0), *sn = (double *)malloc(sizeof(double) * 5 * S1);
    double *cvn = (double *)malloc(sizeof(double) * 9 * (S1 > 0 ? S1 : 1));
    /* L-BFGS: old atom of each new atom, old slot of each new structure */
    int *amap = (int *)malloc(sizeof(int) * (N > 0 ? N : 1)), *smap = (int *)malloc(sizeof(int) * (S > 0 ? S : 1));
    const size_t old_stride = 3 * (size_t)N;
    int kk = 0;
    for (int k = 0; k < S; k++) {
      if (st[5 * k + 3] != 0.0) continue;
      const int o = b->act[kk], n = b->off0[o + 1] - b->off0[o];
      smap[kk] = k;
      for (int a = 0; a < n; a++) amap[N1 + a] = c->h_atom_off[k] + a;
      off[kk] = N1;
      nd[kk] = b->nd0[o];
      xt[kk] = 1;
      memcpy(lat + 9 * kk, b->lat0 + 9 * o, sizeof(double) * 9);
      memcpy(spc + N1, b->spc0 + b->off0[o], sizeof(int) * n);
      memcpy(xn + 3 * N1, x + 3 * c->h_atom_off[k], sizeof(double) * 3 * n);
      memcpy(vn + 3 * N1, v + 3 * c->h_atom_off[k], sizeof(double) * 3 * n);
      memcpy(sn + 5 * kk, st + 5 * k, sizeof(double) * 5);
      memcpy(cvn + 9 * kk, cv + 9 * k, sizeof(double) * 9);
      N1 += n;
      kk++;
    }
    off[S1] = N1;
    r = mgh_upload(c, S1, off, nd, xt, lat, spc, xn);
    if (r == MG_OK && (r = ensure_dyn(c)) == MG_OK) {
      if (hipMemcpy(c->d_v, vn, sizeof(double) * 3 * N1, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(c->d_fire, sn, sizeof(double) * 5 * S1, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(c->d_cellv, cvn, sizeof(double) * 9 * S1, hipMemcpyHostToDevice) != hipSuccess)
        r = mg_fail(MG_ERR_HIP, "mg_relax: state upload failed");
    }
    if (r == MG_OK && c->lb_on) r = lbfgs_repack(c, amap, N1, smap, S1, old_stride);
    b->S = S1;
    free(off); free(nd); free(xt); free(spc); free(lat); free(xn); free(vn); free(sn); free(cvn);
    free(amap); free(smap);
  }
out:
  free(x); free(v); free(st); free(cv); free(sd);
  return r;
}

static int relax_impl(mg_ctx *c, int cell, double p, double smax, int max_steps, double fmax, double dt0, double dtmax,
                      double maxstep, int lbfgs, int mem, int *steps, int *done) {
  if (!c || c->S < 1 || max_steps < 0 || fmax <= 0.0 || (cell && smax <= 0.0) ||
      (lbfgs && (cell || mem < 1 || mem > 32 || dt0 <= 0.0 || maxstep <= 0.0)))
    return mg_fail(MG_ERR_ARG, "mg_relax: bad argument");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if ((r = mgh_settle(c)) != MG_OK) return r;
  if ((r = ensure_dyn(c))) return r;
  relax_book b;
  memset(&b, 0, sizeof b);
  b.S0 = b.S = c->S;
  b.N0 = c->N;
  b.off0 = (int *)malloc(sizeof(int) * (b.S0 + 1));
  b.nd0 = (int *)malloc(sizeof(int) * b.S0);
  b.spc0 = (int *)malloc(sizeof(int) * (b.N0 > 0 ? b.N0 : 1));
  b.lat0 = (double *)malloc(sizeof(double) * 9 * b.S0);
  b.x0 = (double *)malloc(sizeof(double) * 3 * (b.N0 > 0 ? b.N0 : 1));
  b.st0 = (double *)malloc(sizeof(double) * 5 * b.S0);
  b.act = (int *)malloc(sizeof(int) * b.S0);
  memcpy(b.off0, c->h_atom_off, sizeof(int) * (b.S0 + 1));
  memcpy(b.nd0, c->h_nd, sizeof(int) * b.S0);
  memcpy(b.lat0, c->h_lat, sizeof(double) * 9 * b.S0);
  for (int s = 0; s < b.S0; s++) b.act[s] = s;
  if (hipMemcpy(b.spc0, c->d_spc, sizeof(int) * b.N0, hipMemcpyDeviceToHost) != hipSuccess) {
    book_free(&b);
    return mg_fail(MG_ERR_HIP, "mg_relax: species download failed");
  }
  int compacted = 0;
#define RCHK(x) do { if ((r = (x)) != MG_OK) { if (gx) hipGraphExecDestroy(gx); book_free(&b); return r; } } while (0)
#define RKCHK(x) do { if ((x) != 0) { if (gx) hipGraphExecDestroy(gx); book_free(&b); return mg_fail(MG_ERR_HIP, "kernel launch failed: %s (%s)", #x, mgk_last_error()); } } while (0)
#define RHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { if (gx) hipGraphExecDestroy(gx); book_free(&b); return mg_fail(MG_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
  hipGraphExec_t gx = NULL;
  int maxn = 0;
  for (int s = 0; s < c->S; s++)
    if (c->h_atom_off[s + 1] - c->h_atom_off[s] > maxn) maxn = c->h_atom_off[s + 1] - c->h_atom_off[s];
  /* L-BFGS: dt0 is the initial inverse Hessian h0 (A^2/eV) */
  fire_arg fa = {fmax, dtmax, maxstep, cell, p, smax, lbfgs, dt0, maxn};
  RHIP(hipMemsetAsync(c->d_v, 0, sizeof(double) * 3 * c->N, c->st));
  RHIP(hipMemsetAsync(c->d_cellv, 0, sizeof(double) * 9 * c->S, c->st));
  if (lbfgs) {
    RCHK(ensure_lbfgs(c, mem));
    RHIP(hipMemsetAsync(c->d_lb_rho, 0, sizeof(double) * mem * c->S, c->st));
    RKCHK(mgk_lbfgs_init(c->S, c->d_fire, dt0, c->st));
    c->lb_on = 1;
  } else {
    RKCHK(mgk_fire_init(c->S, c->d_fire, dt0, c->st));
  }
  int it = 0, need_e = 0, need_r = 0;
  c->rlog_n = 0;
  c->skin_on = !cell && c->skin > 0.0;  /* positions only: the cells stay fixed, skin lists hold */
  c->skin_need = 0;
  double t_seg = now_s();
  while (it < max_steps) {
    int nrun;
    if (c->graph_enable && gx) {
      nrun = max_steps - it < MG_GRAPH_CHECK ? max_steps - it : MG_GRAPH_CHECK;
      for (int k = 0; k < nrun; k++) RHIP(hipGraphLaunch(gx, c->st));
    } else {
      /* synchronised iterations: all of them without graphs, else one to size the buffers before a capture */
      nrun = c->graph_enable ? 1 : (max_steps - it < MG_GRAPH_CHECK ? max_steps - it : MG_GRAPH_CHECK);
      for (int k = 0; k < nrun; k++) {
        RCHK(mgh_eval(c, MG_EVAL_FORCES, 0));
        RCHK(fire_tail(c, &fa));
        if (it + k == 0) {
          int any_frac = 0;
          for (int s = 0; s < c->S; s++) if (c->h_str[s].xt == 0) { c->h_str[s].xt = 1; any_frac = 1; }
          if (any_frac) RHIP(hipMemcpyAsync(c->d_str, c->h_str, sizeof(mg_struct_dev) * c->S, hipMemcpyHostToDevice, c->st));
        }
      }
      if (c->graph_enable) {
        it += nrun;
        RCHK(mgh_autotune_once(c));  /* before the capacities and the capture, which depend on the shape */
        RCHK(mgh_graph_caps(c, c->nedges > need_e ? c->nedges : need_e, c->maxrow > need_r ? c->maxrow : need_r));
        RCHK(graph_capture(c, NULL, fire_tail, &fa, &gx));
        RCHK(snapshot(c, 0));
        RKCHK(mgk_clear(c->d_misc + 3, 3, NULL, 0, NULL, 0, c->st));
        continue;
      }
    }
    it += nrun;
    int flags[8];
    RKCHK(mgk_clear(c->d_misc + 2, 1, NULL, 0, NULL, 0, c->st));
    RKCHK(mgk_fire_count(c->S, c->d_fire, c->d_misc + 2, c->st));
    RHIP(hipMemcpyAsync(flags, c->d_misc, sizeof flags, hipMemcpyDeviceToHost, c->st));
    RHIP(hipStreamSynchronize(c->st));
    {
      const double t1 = now_s();
      rlog_add(c, it, c->S, c->N, t1 - t_seg);
      t_seg = t1;
    }
    if (flags[3]) {
      /* a row or the edge total outgrew the capture: redo these iterations with larger buffers */
      if (flags[7] > c->skin_need) c->skin_need = flags[7];
      RCHK(snapshot(c, 1));
      it -= nrun;
      need_e = flags[4];
      need_r = flags[5];
      hipGraphExecDestroy(gx);
      gx = NULL;
      continue;
    }
    if (flags[1]) RCHK(mg_fail(MG_ERR_NEIGHBOR, "neighbour list error: a reverse edge is missing, or a cell is too thin for the cutoff"));
    RCHK(snapshot(c, 0));
    RKCHK(mgk_clear(c->d_misc + 3, 3, NULL, 0, NULL, 0, c->st));
    const int ndone = flags[2];
    if (ndone == c->S) break;
    if (ndone > 0 && c->S > 1) {
      /* atoms still moving, from the per-structure done flags */
      double *st = (double *)malloc(sizeof(double) * 5 * c->S);
      RHIP(hipMemcpy(st, c->d_fire, sizeof(double) * 5 * c->S, hipMemcpyDeviceToHost));
      long moving = 0;
      for (int k = 0; k < c->S; k++)
        if (st[5 * k + 3] == 0.0) moving += c->h_atom_off[k + 1] - c->h_atom_off[k];
      free(st);
      if (2 * moving <= c->N) {
        RCHK(compact(c, &b, 1));
        compacted = 1;
        if (gx) hipGraphExecDestroy(gx);  /* the batch changed: capture again */
        gx = NULL;
        need_e = need_r = 0;
      }
    }
  }
  if (gx) hipGraphExecDestroy(gx);
  gx = NULL;
  if (cell && !compacted) {
    /* the lattices moved on the device: bring the host tables up to date */
    RHIP(hipMemcpy(c->h_str, c->d_str, sizeof(mg_struct_dev) * c->S, hipMemcpyDeviceToHost));
    for (int s = 0; s < c->S; s++) memcpy(c->h_lat + 9 * s, c->h_str[s].L, 9 * sizeof(double));
  }
  if (compacted) {
    /* bring the last active structures into the book and restore the full batch with the final positions */
    RCHK(compact(c, &b, 0));
    int *xt = (int *)malloc(sizeof(int) * b.S0);
    for (int s = 0; s < b.S0; s++) xt[s] = 1;
    r = mgh_upload(c, b.S0, b.off0, b.nd0, xt, b.lat0, b.spc0, b.x0);
    free(xt);
    if (r != MG_OK) { book_free(&b); return r; }
    RCHK(ensure_dyn(c));
    RHIP(hipMemcpy(c->d_fire, b.st0, sizeof(double) * 5 * b.S0, hipMemcpyHostToDevice));
  }
  /* energies and forces at the final positions */
  RCHK(mgh_eval(c, MG_EVAL_FORCES, 0));
  if (steps || done) {
    double *state = (double *)malloc(sizeof(double) * 5 * c->S);
    hipError_t e = hipMemcpy(state, c->d_fire, sizeof(double) * 5 * c->S, hipMemcpyDeviceToHost);
    for (int s = 0; s < c->S && e == hipSuccess; s++) {
      if (steps) steps[s] = (int)state[5 * s + 4];
      if (done) done[s] = state[5 * s + 3] != 0.0;
    }
    free(state);
    if (e != hipSuccess) { book_free(&b); return mg_fail(MG_ERR_HIP, "mg_relax: %s", hipGetErrorString(e)); }
  }
  book_free(&b);
#undef RCHK
#undef RKCHK
#undef RHIP
  return MG_OK;
}

int mg_relax(mg_ctx *c, int max_steps, double fmax, double dt0, double dtmax, double maxstep, int *steps, int *done) {
  const int r = relax_impl(c, 0, 0.0, 0.0, max_steps, fmax, dt0, dtmax, maxstep, 0, 0, steps, done);
  if (c) c->skin_on = 0;  /* skin lists serve this run only (every exit path) */
  return r;
}

int mg_relax_lbfgs(mg_ctx *c, int max_steps, double fmax, double maxstep, int memory, double h0, int *steps,
                   int *done) {
  const int r = relax_impl(c, 0, 0.0, 0.0, max_steps, fmax, h0, 0.0, maxstep, 1, memory, steps, done);
  if (c) { c->skin_on = 0; c->lb_on = 0; }  /* this run only (every exit path) */
  return r;
}

int mg_relax_cell(mg_ctx *c, double p_gpa, int max_steps, double fmax, double smax_gpa, double dt0, double dtmax,
                  double maxstep, int *steps, int *done) {
  const double ev2gpa = 160.2176565; /* eV/A^3 in GPa, the reference's eV2GPa */
  return relax_impl(c, 1, p_gpa / ev2gpa, smax_gpa / ev2gpa, max_steps, fmax, dt0, dtmax, maxstep, 0, 0, steps, done);
}
