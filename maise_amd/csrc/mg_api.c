/*
 * mg_api.c -- host side of libmaise_gpu: evaluation contexts, structure upload (CAST_CELL semantics),
 * the kernel chain of one evaluation, result download, and the CALL_MAISE drop-in.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mg_host.h"
#include "../../include/mlib.h"

/* ------------------------------------------------------------------ errors */
static __thread char g_err[512];

int mg_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *mg_last_error(void) { return g_err; }


/* ------------------------------------------------------------------ context */
int mg_dev_alloc(void **p, size_t bytes) {
  if (*p) hipFree(*p);
  *p = NULL;
  if (bytes == 0) bytes = 16;
  HIPCHK(hipMalloc(p, bytes));
  return MG_OK;
}

#define ALLOC(ptr, bytes)                                                                                        \
  do {                                                                                                           \
    if (c->no_alloc)                                                                                             \
      return mg_fail(MG_ERR_ARG, "fixed scratch: a buffer of %zu bytes (%s) does not fit", (size_t)(bytes), #ptr); \
    int r_ = mg_dev_alloc((void **)&(ptr), (bytes));                                                             \
    if (r_) return r_;                                                                                           \
  } while (0)

/* host temporaries of an upload: malloc, or a slice of the context's host work area (fixed-scratch evaluation);
 * *used tracks the slices taken; NULL when the work area is too small */
static void *htmp(mg_ctx *c, size_t *used, size_t bytes) {
  if (!c->h_work) return malloc(bytes > 0 ? bytes : 1);
  const size_t b = (bytes + 15) & ~(size_t)15;
  if (*used + b > c->h_work_bytes) return NULL;
  void *p = c->h_work + *used;
  *used += b;
  return p;
}
static void hfree(mg_ctx *c, void *p) {
  if (!c->h_work) free(p);
}

static int ensure_atoms(mg_ctx *c, size_t N) {
  if (N <= c->cap_atoms) return MG_OK;
  size_t n = N + N / 4 + 16;
  const int D = c->m->D;
  ALLOC(c->d_atom_str, n * sizeof(int));
  ALLOC(c->d_spc, n * sizeof(int));
  ALLOC(c->d_xin, n * 3 * sizeof(double));
  ALLOC(c->d_x, n * 3 * sizeof(double));
  ALLOC(c->d_bin, 2 * n * sizeof(int));  /* bin id, then the packed bin coordinates at N + i */
  ALLOC(c->d_bin_rec, n * 32);
  ALLOC(c->d_cnt, n * 4 * sizeof(int));
  ALLOC(c->d_npl, n * MG_NCLS * sizeof(int));
  ALLOC(c->d_rowcnt, n * sizeof(int));
  ALLOC(c->d_row_off, (n + 1) * sizeof(int));
  ALLOC(c->d_G, n * D * sizeof(float));
  ALLOC(c->d_dEdG, n * D * sizeof(float));
  ALLOC(c->d_Ei, n * sizeof(float));
  ALLOC(c->d_erep, n * sizeof(float));
  ALLOC(c->d_vir, n * 6 * sizeof(float));
  ALLOC(c->d_fself, n * 16);
  ALLOC(c->d_F, n * 3 * sizeof(double));
  c->cap_atoms = n;
  /* scan scratch for the larger of the atom and bin scans */
  return MG_OK;
}

static int ensure_scan_tmp(mg_ctx *c, int n) {
  size_t need = 0;
  if (mgk_scan(NULL, NULL, n, NULL, &need, c->st) != 0) return mg_fail(MG_ERR_HIP, "scan size query failed");
  if (need > c->tmp_bytes) {
    ALLOC(c->d_tmp, need);
    c->tmp_bytes = need;
  }
  return MG_OK;
}

static int ensure_str(mg_ctx *c, size_t S) {
  if (S <= c->cap_str) return MG_OK;
  size_t n = S + 4;
  ALLOC(c->d_str, n * sizeof(mg_struct_dev));
  ALLOC(c->d_E, n * sizeof(double));
  ALLOC(c->d_U, n * 6 * sizeof(double));
  c->h_str = (mg_struct_dev *)realloc(c->h_str, n * sizeof(mg_struct_dev));
  c->h_atom_off = (int *)realloc(c->h_atom_off, (n + 1) * sizeof(int));
  c->h_nd = (int *)realloc(c->h_nd, n * sizeof(int));
  c->h_lat = (double *)realloc(c->h_lat, n * 9 * sizeof(double));
  c->cap_str = n;
  return MG_OK;
}

static int ensure_bins(mg_ctx *c, size_t nb) {
  if (nb <= c->cap_bins) return MG_OK;
  size_t n = nb + nb / 4 + 16;
  ALLOC(c->d_bin_count, (n + 1) * sizeof(int));
  ALLOC(c->d_bin_start, (n + 1) * sizeof(int));
  c->cap_bins = n;
  c->bins_dirty = 1;
  return MG_OK;
}

/* Stored G4 pair lists: every unordered pair of a row of up to cap_row neighbours fits the per-atom stride.  The
 * allocation holds alloc_row-long rows (so the capacities set after a synchronised evaluation reuse it); returns
 * the stride for use_row, or 0 when the lists would exceed the budget or cannot be allocated (the force kernel then
 * enumerates the pairs itself). */
static size_t pair_stride(int cap_row) { return cap_row < 2 ? 1 : (size_t)cap_row * (size_t)(cap_row - 1) / 2; }

static size_t ensure_pairs(mg_ctx *c, int alloc_row, int use_row) {
  if (!c->pair_store || c->N == 0) return 0;
  const size_t need = (size_t)c->N * pair_stride(alloc_row > use_row ? alloc_row : use_row);
  if (need * sizeof(unsigned) > c->pair_budget) return 0;
  if (need > c->cap_gpl) {
    if (c->d_gpl) hipFree(c->d_gpl);
    c->d_gpl = NULL;
    c->cap_gpl = 0;
    if (hipMalloc((void **)&c->d_gpl, need * sizeof(unsigned)) != hipSuccess) {
      (void)hipGetLastError();  /* out of memory here is not an error: the pairs are enumerated twice instead */
      c->d_gpl = NULL;
      return 0;
    }
    c->cap_gpl = need;
  }
  return pair_stride(use_row);
}

/* the reverse-edge hash tables for rows of up to row entries */
static int ensure_rhash(mg_ctx *c, int row) {
  const size_t need = (size_t)c->N * (size_t)mgk_rev_slots(row);
  if (need <= c->cap_rhash) return MG_OK;
  const size_t n = need + need / 4 + 1024;
  ALLOC(c->d_rhash, n * sizeof(uint16_t));
  c->cap_rhash = n;
  return MG_OK;
}

static int ensure_edges(mg_ctx *c, size_t M) {
  if (M <= c->cap_edges) return MG_OK;
  size_t n = M + M / 4 + 64;
  ALLOC(c->d_key, n * 8);
  ALLOC(c->d_vec, n * 16);
  ALLOC(c->d_rev, n * sizeof(int));
  ALLOC(c->d_fe, n * 16);
  c->cap_edges = n;
  return MG_OK;
}

int mg_ctx_create(const mg_model *m, int device, mg_ctx **out) {
  if (!m || !out) return mg_fail(MG_ERR_ARG, "mg_ctx_create: null argument");
  *out = NULL;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return mg_fail(MG_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));
  mg_ctx *c = (mg_ctx *)calloc(1, sizeof(mg_ctx));
  c->graph_enable = 1;
  c->graph_headroom = 0.125;
  c->pair_store = 1;
  {
    size_t fr = 0, tot = 0;
    c->pair_budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 16 : ((size_t)1 << 30);
  }
  c->m = m;
  c->device = device;
  if (hipStreamCreate(&c->st) != hipSuccess) { free(c); return mg_fail(MG_ERR_HIP, "hipStreamCreate failed"); }
  if (hipHostMalloc((void **)&c->h_flags, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    hipStreamDestroy(c->st);
    free(c);
    return mg_fail(MG_ERR_HIP, "hipHostMalloc failed");
  }
  int r;
  if ((r = mg_dev_alloc((void **)&c->d_basis, sizeof(mg_basis_dev))) ||
      (r = mg_dev_alloc((void **)&c->d_mlp, (size_t)m->wstride * m->nspc * sizeof(float))) ||
      (r = mg_dev_alloc((void **)&c->d_misc, 32 * sizeof(int)))) {
    mg_ctx_free(c);
    return r;
  }
  HIPCHK(hipMemset(c->d_misc, 0, 32 * sizeof(int)));  /* misc[16..31] are never cleared by the chains */
  HIPCHK(hipMemcpy(c->d_basis, &m->basis, sizeof(mg_basis_dev), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_mlp, m->mlp, (size_t)m->wstride * m->nspc * sizeof(float), hipMemcpyHostToDevice));
  if (m->mfma_ok) {
    if ((r = mg_dev_alloc((void **)&c->d_frag, m->frag_n * sizeof(float))) ||
        (r = mg_dev_alloc((void **)&c->d_frag_bf, m->frag_bf_n * sizeof(uint16_t))) ||
        (r = mg_dev_alloc((void **)&c->d_frag_lo, m->frag_bf_n * sizeof(uint16_t))) ||
        (r = mg_dev_alloc((void **)&c->d_fvec, m->fvec_n * sizeof(float)))) {
      mg_ctx_free(c);
      return r;
    }
    HIPCHK(hipMemcpy(c->d_frag, m->frag, m->frag_n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_frag_bf, m->frag_bf, m->frag_bf_n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_frag_lo, m->frag_lo, m->frag_bf_n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_fvec, m->fvec, m->fvec_n * sizeof(float), hipMemcpyHostToDevice));
  }
  *out = c;
  return MG_OK;
}

void mg_ctx_free(mg_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  void *ptrs[] = {c->d_basis, c->d_mlp, c->d_str, c->d_atom_str, c->d_spc, c->d_xin, c->d_x, c->d_bin, c->d_bin_count,
                  c->d_bin_start, c->d_bin_rec, c->d_cnt, c->d_rowcnt, c->d_row_off, c->d_misc,
                  c->d_key, c->d_vec, c->d_rev, c->d_rhash, c->d_G, c->d_dEdG, c->d_Ei, c->d_erep,
                  c->d_vir, c->d_fself, c->d_fe, c->d_gpl, c->d_npl, c->d_F, c->d_E, c->d_U, c->d_tmp, c->d_part, c->d_chunk_str, c->d_chunk_cnt, c->d_frag, c->d_frag_bf, c->d_frag_lo, c->d_fvec, c->d_perm, c->d_tiles,
                  c->d_v, c->d_mass, c->d_dyn, c->d_fire, c->d_trace, c->d_snap_x, c->d_snap_v, c->d_snap_fire,
                  c->d_snap_dyn, c->d_cellv, c->d_snap_cellv, c->d_snap_str, c->d_ckey, c->d_crow, c->d_xb,
                  c->d_mimg, c->d_lb_s, c->d_lb_y, c->d_lb_f0, c->d_lb_rho, c->d_lb_tmp, c->d_snap_lb_s,
                  c->d_snap_lb_y, c->d_snap_lb_f0, c->d_snap_lb_rho, c->d_lb_map};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++)
    if (ptrs[k]) hipFree(ptrs[k]);
  if (c->st) hipStreamSynchronize(c->st);
  if (c->tev_ok)
    for (int q = 0; q < MG_TRING; q++)
      for (int k = 0; k < EV_N; k++) hipEventDestroy(c->tev[q][k]);
  if (c->h_flags) hipHostFree(c->h_flags);
  if (c->st) hipStreamDestroy(c->st);
  free(c->h_str); free(c->h_atom_off); free(c->h_nd); free(c->h_lat);
  free(c->rlog);
  free(c);
}

int mg_ctx_set_urep(mg_ctx *c, double urep) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  c->urep = urep;
  return MG_OK;
}

int mg_ctx_set_precision(mg_ctx *c, int prec) {
  if (!c || prec < MG_PREC_FP32 || prec > MG_PREC_BF16X3) return mg_fail(MG_ERR_ARG, "bad precision");
  c->prec = prec;
  return MG_OK;
}

int mg_ctx_set_timing(mg_ctx *c, int on) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  int r;
  if ((r = mgh_settle(c)) != MG_OK) return r;
  HIPCHK(hipSetDevice(c->device));
  if (on && !c->tev_ok) {
    for (int q = 0; q < MG_TRING; q++)
      for (int k = 0; k < EV_N; k++) HIPCHK(hipEventCreate(&c->tev[q][k]));
    c->tev_ok = 1;
  }
  c->timing = on != 0;
  c->tcount = c->tread = c->tn = 0;
  for (int k = 0; k < 6; k++) c->tsum[k] = 0.0;
  return MG_OK;
}

int mg_ctx_set_wpb(mg_ctx *c, int wpb_desc, int wpb_force) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  if ((wpb_desc != 0 && wpb_desc != 1 && wpb_desc != 2 && wpb_desc != 4) ||
      (wpb_force != 0 && wpb_force != 1 && wpb_force != 2 && wpb_force != 4))
    return mg_fail(MG_ERR_ARG, "waves per block must be 0 (automatic), 1, 2 or 4");
  c->wpb_desc = wpb_desc;
  c->wpb_force = wpb_force;
  return MG_OK;
}

int mg_ctx_set_wpb_nbr(mg_ctx *c, int wpb) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  if (wpb != 0 && wpb != 1 && wpb != 2 && wpb != 4)
    return mg_fail(MG_ERR_ARG, "waves per block must be 0 (automatic), 1, 2 or 4");
  c->wpb_nbr = wpb;
  return MG_OK;
}

int mg_ctx_set_autotune(mg_ctx *c, int on) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  c->autotune = on != 0;
  c->tuned = 0;
  return MG_OK;
}

static void plan_bins(mg_struct_dev *s, double rc, double skin, int natom);

int mg_ctx_set_skin(mg_ctx *c, double skin) {
  if (!c || !(skin >= 0.0) || skin > 4.0) return mg_fail(MG_ERR_ARG, "mg_ctx_set_skin: skin must be 0..4 A");
  if (c->no_alloc && skin > 0.0) return mg_fail(MG_ERR_ARG, "mg_ctx_set_skin: not with the fixed scratch");
  c->skin = skin;
  c->ccap_row = 0;
  if (c->S > 0) {  /* the skin stencils of the structures set */
    for (int s = 0; s < c->S; s++) plan_bins(&c->h_str[s], c->m->rc_list, skin, c->h_str[s].natom);
    HIPCHK(hipMemcpy(c->d_str, c->h_str, c->S * sizeof(mg_struct_dev), hipMemcpyHostToDevice));
  }
  return MG_OK;
}

int mg_ctx_skin_builds(mg_ctx *c, int *builds) {
  if (!c || !builds) return mg_fail(MG_ERR_ARG, "mg_ctx_skin_builds: bad argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(builds, c->d_misc + 16, sizeof(int), hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_ctx_get_shape(mg_ctx *c, int *shape) {
  if (!c || !shape) return mg_fail(MG_ERR_ARG, "mg_ctx_get_shape: bad argument");
  shape[0] = c->wpb_nbr;
  shape[1] = c->wpb_desc;
  shape[2] = c->wpb_force;
  shape[3] = c->pair_store;
  return MG_OK;
}

/* Launch-shape autotuner on the structure (or batch) now set, reps timed evaluations per candidate after one
 * untimed: (1) 1, 2 and 4 waves per block for the neighbour fill, the descriptor and the force kernel, each
 * kernel keeping its fastest (the stages are timed separately, so one sweep serves all three); (2) with those,
 * the force kernel reading the descriptor kernel's stored pair lists against enumerating them itself, by the
 * descriptor + force time.  Candidates that do not fit (LDS, pair-list budget) are skipped.  None of these
 * choices changes a result bit: every atom's arithmetic is the same whatever the block it runs in, and both pair
 * sources visit the same pairs in the same order. */
static int autotune_impl(mg_ctx *c, int reps, int *shape) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  if (c->S == 0) return mg_fail(MG_ERR_ARG, "mg_ctx_autotune: no structure set");
  if (reps < 1) reps = 1;
  const int cand[3] = {1, 2, 4};
  const int timing0 = c->timing, wn0 = c->wpb_nbr, wd0 = c->wpb_desc, wf0 = c->wpb_force, ps0 = c->pair_store;
  float best_n = 1e30f, best_d = 1e30f, best_f = 1e30f, t[6];
  int wn = 0, wd = 0, wf = 0, r;
  if ((r = mg_ctx_set_timing(c, 1)) != MG_OK) return r;
  for (int k = 0; k < 3; k++) {
    c->wpb_nbr = c->wpb_desc = c->wpb_force = cand[k];
    if (mgh_eval(c, MG_EVAL_FORCES, 0) != MG_OK) continue;  /* LDS does not fit: skip this shape */
    if ((r = mg_get_timing(c, t)) != MG_OK) return r;       /* drop the warm-up evaluation */
    int ok = 1;
    for (int q = 0; q < reps && ok; q++)
      if (mgh_eval(c, MG_EVAL_FORCES, 0) != MG_OK) ok = 0;
    if (!ok) continue;
    if ((r = mg_get_timing(c, t)) != MG_OK) return r;       /* stage means over the reps */
    if (t[0] < best_n) { best_n = t[0]; wn = cand[k]; }
    if (t[1] < best_d) { best_d = t[1]; wd = cand[k]; }
    if (t[3] < best_f) { best_f = t[3]; wf = cand[k]; }
  }
  if (!wn || !wd || !wf) {
    c->wpb_nbr = wn0;
    c->wpb_desc = wd0;
    c->wpb_force = wf0;
    mg_ctx_set_timing(c, timing0);
    return mg_fail(MG_ERR_HIP, "mg_ctx_autotune: no launch shape ran");
  }
  c->wpb_nbr = wn;
  c->wpb_desc = wd;
  c->wpb_force = wf;
  /* stored pair lists against enumeration, at the chosen shape; switching them re-sizes the capacities (as
   * mg_ctx_set_pair_store does), so the next evaluation stores the lists or not as asked */
  float tp[2] = {1e30f, 1e30f};
  for (int store = 0; store < 2; store++) {
    c->pair_store = store;
    c->gcap_edges = c->gcap_row = 0;
    c->gpstride = 0;
    if (mgh_eval(c, MG_EVAL_FORCES, 0) != MG_OK) continue;
    if ((r = mg_get_timing(c, t)) != MG_OK) return r;
    int ok = 1;
    for (int q = 0; q < reps && ok; q++)
      if (mgh_eval(c, MG_EVAL_FORCES, 0) != MG_OK) ok = 0;
    if (!ok) continue;
    if ((r = mg_get_timing(c, t)) != MG_OK) return r;
    tp[store] = t[1] + t[3];
  }
  c->pair_store = tp[1] < 1e29f && tp[1] <= tp[0] ? 1 : (tp[0] < 1e29f ? 0 : ps0);
  c->gcap_edges = c->gcap_row = 0;  /* the next evaluation sizes the capacities for the chosen setting */
  c->gpstride = 0;
  if ((r = mg_ctx_set_timing(c, timing0)) != MG_OK) return r;
  c->tuned = 1;
  if (shape) mg_ctx_get_shape(c, shape);
  return MG_OK;
}

int mg_ctx_autotune(mg_ctx *c, int reps, int *wpb_out) {
  int shape[4];
  const int r = autotune_impl(c, reps, shape);
  if (r == MG_OK && wpb_out) { wpb_out[0] = shape[1]; wpb_out[1] = shape[2]; }
  return r;
}

int mg_ctx_autotune_shape(mg_ctx *c, int reps, int *shape) { return autotune_impl(c, reps, shape); }

/* called by relaxation and MD before their first capture: tune once per context batch when asked to */
int mgh_autotune_once(mg_ctx *c) {
  if (!c->autotune || c->tuned) return MG_OK;
  return autotune_impl(c, 3, NULL);
}

/* reciprocal rows R = (L1xL2, L2xL0, L0xL1) / (L0.(L1xL2)) with the reference's operation order
 * (src/cell.c:432-444, src/util.c:186-219); compiled without FMA contraction */
void mgh_reciprocal(const double *L, double *R) {
  const double *a = L, *b = L + 3, *c = L + 6;
  R[6] = a[1] * b[2] - a[2] * b[1];
  R[7] = a[2] * b[0] - a[0] * b[2];
  R[8] = a[0] * b[1] - a[1] * b[0];
  R[0] = b[1] * c[2] - b[2] * c[1];
  R[1] = b[2] * c[0] - b[0] * c[2];
  R[2] = b[0] * c[1] - b[1] * c[0];
  R[3] = c[1] * a[2] - c[2] * a[1];
  R[4] = c[2] * a[0] - c[0] * a[2];
  R[5] = c[0] * a[1] - c[1] * a[0];
  double t = 1.0 / (a[0] * b[1] * c[2] + a[1] * b[2] * c[0] + a[2] * b[0] * c[1] - a[0] * b[2] * c[1] -
                    a[1] * b[0] * c[2] - a[2] * b[1] * c[0]);
  for (int i = 0; i < 9; i++) R[i] *= t;
}

static double cell_volume(const double *L) { /* src/cutl.c:178-188 */
  double V = L[0] * L[4] * L[8] + L[2] * L[3] * L[7] + L[1] * L[5] * L[6] - L[0] * L[5] * L[7] - L[1] * L[3] * L[8] -
             L[2] * L[4] * L[6];
  return fabs(V);
}

/* bins and stencil of one structure: bins at least rc/2 thick (perpendicular to each lattice plane), stencil
 * half width K_q >= rc |R_q| nb_q so every pair within rc lies in the stencil (the neighbour kernels skip the
 * stencil bins that lie farther than rc from the atom). */
static void plan_bins(mg_struct_dev *s, double rc, double skin, int natom) {
  const double eps = 1e-6;
  if (s->nd == 0) {
    s->nb[0] = s->nb[1] = s->nb[2] = 1;
    s->K[0] = s->K[1] = s->K[2] = 0;
    s->Ks[0] = s->Ks[1] = s->Ks[2] = 0;
  } else {
    double h[3];
    for (int q = 0; q < 3; q++) {
      double a = 0.0;
      for (int k = 0; k < 3; k++) a += s->R[3 * q + k] * s->R[3 * q + k];
      h[q] = rc * sqrt(a) * (1.0 + eps); /* rc in units of the lattice-plane spacing */
      double nb = floor(2.0 / h[q]);
      s->nb[q] = nb < 1.0 ? 1 : (nb > 1024.0 ? 1024 : (int)nb);
    }
    /* keep the bin count O(natom) for sparse cells */
    while ((double)s->nb[0] * s->nb[1] * s->nb[2] > 4.0 * natom + 64) {
      int q = 0;
      for (int k = 1; k < 3; k++) if (s->nb[k] > s->nb[q]) q = k;
      s->nb[q] = (s->nb[q] + 1) / 2;
    }
    for (int q = 0; q < 3; q++) {
      int K = (int)ceil(h[q] * s->nb[q]);
      s->K[q] = K < 1 ? 1 : K;
      /* the skin-candidate radius rc + skin needs a stencil that much wider */
      int Ks = (int)ceil(h[q] * (rc + skin) / rc * s->nb[q]);
      s->Ks[q] = Ks < s->K[q] ? s->K[q] : Ks;
    }
  }
  s->nbins = s->nb[0] * s->nb[1] * s->nb[2];
}

int mgh_upload(mg_ctx *c, int S, const int *atom_off, const int *nd, const int *xt, const double *lat, const int *spc,
               const double *x) {
  HIPCHK(hipSetDevice(c->device));
  if (c->pending) {  /* the results of an evaluation still running are dropped with the old structures */
    c->pending = 0;
    HIPCHK(hipStreamSynchronize(c->st));
  }
  c->bins_dirty = 1;  /* the bin layout may change: clear the counts once before the next evaluation */
  int N = atom_off[S];
  for (int i = 0; i < N; i++)
    if (spc[i] < 0 || spc[i] >= c->m->nspc) return mg_fail(MG_ERR_ARG, "atom %d: species %d outside the model's %d", i, spc[i], c->m->nspc);
  int r;
  if ((r = ensure_str(c, S)) || (r = ensure_atoms(c, N))) return r;
  size_t hused = 0;
  int *atom_str = (int *)htmp(c, &hused, sizeof(int) * (N > 0 ? N : 1));
  if (!atom_str) return mg_fail(MG_ERR_ARG, "fixed scratch: %d atoms exceed the host work area", N);
  int bin0 = 0, nchunk = 0;
  const double rc = c->m->rc_list;
  for (int s = 0; s < S; s++) {
    mg_struct_dev *d = &c->h_str[s];
    memset(d, 0, sizeof *d);
    for (int k = 0; k < 9; k++) d->L[k] = lat[9 * s + k] * 1.0;
    mgh_reciprocal(d->L, d->R);
    d->rc2 = rc * rc;
    d->atom0 = atom_off[s];
    d->natom = atom_off[s + 1] - atom_off[s];
    d->nd = nd[s];
    d->xt = xt[s];
    if (d->nd != 0 && d->nd != 3) { hfree(c, atom_str); return mg_fail(MG_ERR_ARG, "structure %d: ND must be 0 or 3", s); }
    plan_bins(d, rc, c->skin, d->natom);
    d->bin0 = bin0;
    bin0 += d->nbins;
    d->chunk0 = nchunk;
    d->nchunk = (d->natom + MG_RCHUNK - 1) / MG_RCHUNK;
    nchunk += d->nchunk;
    for (int i = d->atom0; i < d->atom0 + d->natom; i++) atom_str[i] = s;
    c->h_nd[s] = nd[s];
    memcpy(c->h_lat + 9 * s, d->L, 9 * sizeof(double));
    c->h_atom_off[s] = atom_off[s];
  }
  c->h_atom_off[S] = N;
  c->nbins = bin0;
  if ((size_t)nchunk > c->cap_chunk) {
    if (c->no_alloc) return mg_fail(MG_ERR_ARG, "fixed scratch: %d reduction chunks do not fit", nchunk);
    if ((r = mg_dev_alloc((void **)&c->d_chunk_str, (nchunk + 16) * sizeof(int))) ||
        (r = mg_dev_alloc((void **)&c->d_chunk_cnt, (nchunk + 16) * sizeof(int))) ||
        (r = mg_dev_alloc((void **)&c->d_part, (nchunk + 16) * 7 * sizeof(double)))) { hfree(c, atom_str); return r; }
    c->cap_chunk = nchunk + 16;
  }
  {
    int *cs = (int *)htmp(c, &hused, sizeof(int) * (nchunk > 0 ? nchunk : 1));
    if (!cs) return mg_fail(MG_ERR_ARG, "fixed scratch: the host work area is too small");
    for (int s = 0; s < S; s++)
      for (int k = 0; k < c->h_str[s].nchunk; k++) cs[c->h_str[s].chunk0 + k] = s;
    hipError_t e = hipMemcpy(c->d_chunk_str, cs, sizeof(int) * nchunk, hipMemcpyHostToDevice);
    /* the assembly kernel's per-structure counters start at zero (each launch leaves them so) */
    if (e == hipSuccess) e = hipMemset(c->d_chunk_cnt, 0, sizeof(int) * nchunk);
    hfree(c, cs);
    if (e != hipSuccess) { hfree(c, atom_str); return mg_fail(MG_ERR_HIP, "chunk table: %s", hipGetErrorString(e)); }
  }
  c->nchunk = nchunk;
  /* species-sorted permutation and MLP tiles of <= 32 same-species atoms */
  {
    const int S_ = c->m->nspc;
    int *perm = (int *)htmp(c, &hused, sizeof(int) * (N > 0 ? N : 1));
    if (!perm) return mg_fail(MG_ERR_ARG, "fixed scratch: the host work area is too small");
    int cntS[MG_MAXSPC] = {0, 0, 0, 0}, off[MG_MAXSPC + 1];
    for (int i = 0; i < N; i++) cntS[spc[i]]++;
    off[0] = 0;
    for (int q = 0; q < S_; q++) off[q + 1] = off[q] + cntS[q];
    int pos[MG_MAXSPC];
    for (int q = 0; q < S_; q++) pos[q] = off[q];
    for (int i = 0; i < N; i++) perm[pos[spc[i]]++] = i;
    int nt = 0;
    for (int q = 0; q < S_; q++) nt += (cntS[q] + 31) / 32;
    int *tiles = (int *)htmp(c, &hused, sizeof(int) * 4 * (nt > 0 ? nt : 1));
    if (!tiles) return mg_fail(MG_ERR_ARG, "fixed scratch: the host work area is too small");
    nt = 0;
    for (int q = 0; q < S_; q++)
      for (int a = off[q]; a < off[q + 1]; a += 32) {
        tiles[4 * nt] = q; tiles[4 * nt + 1] = a; tiles[4 * nt + 2] = off[q + 1] - a < 32 ? off[q + 1] - a : 32; tiles[4 * nt + 3] = 0;
        nt++;
      }
    if ((size_t)nt > c->cap_tiles || (size_t)N > c->cap_tiles * 32 || !c->d_perm) {
      if (c->no_alloc) return mg_fail(MG_ERR_ARG, "fixed scratch: %d MLP tiles do not fit", nt);
      size_t cap = (size_t)nt + N / 32 + 64;
      if ((r = mg_dev_alloc((void **)&c->d_tiles, cap * 4 * sizeof(int))) ||
          (r = mg_dev_alloc((void **)&c->d_perm, (cap * 32 + 64) * sizeof(int)))) { hfree(c, perm); hfree(c, tiles); hfree(c, atom_str); return r; }
      c->cap_tiles = cap;
    }
    hipError_t e1 = hipMemcpy(c->d_perm, perm, sizeof(int) * N, hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(c->d_tiles, tiles, sizeof(int) * 4 * nt, hipMemcpyHostToDevice);
    hfree(c, perm);
    hfree(c, tiles);
    if (e1 != hipSuccess || e2 != hipSuccess) { hfree(c, atom_str); return mg_fail(MG_ERR_HIP, "tile upload failed"); }
    c->ntiles = nt;
  }
  if ((r = ensure_bins(c, bin0))) { hfree(c, atom_str); return r; }
  int nscan = N > bin0 ? N : bin0;
  if ((r = ensure_scan_tmp(c, nscan))) { hfree(c, atom_str); return r; }
  HIPCHK(hipMemcpyAsync(c->d_str, c->h_str, S * sizeof(mg_struct_dev), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_atom_str, atom_str, N * sizeof(int), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_spc, spc, N * sizeof(int), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_xin, x, 3 * N * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  hfree(c, atom_str);
  c->S = S;
  c->N = N;
  c->have_forces = c->have_g = 0;
  c->gcap_edges = c->gcap_row = 0;  /* capacities are sized again by the next synchronised evaluation */
  c->gpstride = 0;
  c->ccap_row = 0;                   /* and the skin candidates with them */
  return MG_OK;
}

int mg_set_structure(mg_ctx *c, int N, int nd, int xt, const int *spc, const double *lat, const double *x) {
  if (!c || N < 0 || (N > 0 && (!spc || !x)) || !lat) return mg_fail(MG_ERR_ARG, "mg_set_structure: bad argument");
  int off[2] = {0, N};
  c->tuned = 0;  /* a new workload: tune again when asked to */
  return mgh_upload(c, 1, off, &nd, &xt, lat, spc, x);
}

int mg_set_batch(mg_ctx *c, int S, const int *atom_off, const int *nd, const double *lat, const int *spc, const double *x) {
  if (!c || S < 1 || !atom_off || !nd || !lat || !spc || !x) return mg_fail(MG_ERR_ARG, "mg_set_batch: bad argument");
  int *xt = (int *)malloc(sizeof(int) * S);
  for (int s = 0; s < S; s++) xt[s] = 1;
  c->tuned = 0;
  int r = mgh_upload(c, S, atom_off, nd, xt, lat, spc, x);
  free(xt);
  return r;
}

/* ------------------------------------------------------------------ evaluation */
#define MG_ERR_NEIGHBOR_CAP 1  /* internal: capacity-mode evaluation outgrew its buffers */
#define KCHK(x)                                                                                                     \
  do {                                                                                                              \
    int kr_ = (x);                                                                                                  \
    if (kr_ != 0) c->bins_dirty = 1; /* the chain stopped part way: bin counts may be left non-zero */             \
    if (kr_ == -3) return mg_fail(MG_ERR_NEIGHBOR, "neighbour rows too long for the LDS of one CU (%d)", c->maxrow); \
    if (kr_ == -2) return mg_fail(MG_ERR_NEIGHBOR, "an atom has more than 2048 neighbours (%d)", c->maxrow);        \
    if (kr_ != 0) return mg_fail(MG_ERR_HIP, "kernel launch failed: %s (%s)", #x, mgk_last_error());               \
  } while (0)

static size_t graph_edge_cap(const mg_ctx *c, int need_edges) {
  const double h = c->graph_headroom;
  return (size_t)need_edges + (size_t)(h * need_edges) + (h > 0.0 ? 256 : 0);
}

/* edge slots for rows of up to row neighbours at a fixed stride (the layout of evaluations with known capacities) */
static size_t padded_edges(const mg_ctx *c, int row) { return (size_t)c->N * (size_t)(row > 1 ? row : 1); }

int mgh_graph_caps(mg_ctx *c, int need_edges, int need_row) {
  const double h = c->graph_headroom;
  const int row = need_row + (h > 0.0 ? 8 : 0);
  size_t ne = padded_edges(c, row);
  const size_t nc = graph_edge_cap(c, need_edges > 0 ? need_edges : 0);
  if (nc > ne) ne = nc;
  if (ne > 0x7fffffff) return mg_fail(MG_ERR_ARG, "too many neighbour edge slots (%zu)", ne);
  int r;
  if ((r = ensure_edges(c, ne))) return r;
  if ((r = ensure_rhash(c, row))) return r;
  c->gcap_edges = (int)ne;
  c->gcap_row = row;
  c->gpstride = ensure_pairs(c, c->gcap_row, c->gcap_row);
  if (c->skin_on) {
    /* candidate rows: the neighbour rows scaled by the volume of the larger sphere, with headroom, or the longest
     * candidate row an overflowed step has seen */
    const double rc = c->m->rc_list, g = (rc + c->skin) / rc;
    int cc = (int)ceil(need_row * g * g * g * (1.0 + (h > 0.0 ? h : 0.0))) + 16;
    if (c->skin_need + 16 > cc) cc = c->skin_need + 16;
    if (cc > 2048) cc = 2048;
    const size_t nk = (size_t)c->N * (size_t)cc;
    if (nk > c->cap_ckey) {
      ALLOC(c->d_ckey, nk * sizeof(unsigned long long));
      c->cap_ckey = nk;
    }
    if ((size_t)c->N > c->cap_skin_atoms) {
      const size_t n = (size_t)c->N + 64;
      ALLOC(c->d_crow, n * sizeof(int));
      ALLOC(c->d_xb, 3 * n * sizeof(double));
      ALLOC(c->d_mimg, n * sizeof(unsigned));
      c->cap_skin_atoms = n;
    }
    c->ccap_row = cc;
    /* candidate rows built before this point may have another stride (or lie in a reallocated buffer): mark the
     * lists stale so the next chain builds them again */
    if (mgk_clear(c->d_misc + 8, 1, NULL, 0, NULL, 0, c->st)) return mg_fail(MG_ERR_HIP, "misc clear failed");
  }
  return MG_OK;
}

int mg_ctx_set_pair_store(mg_ctx *c, int enable) {
  if (!c) return mg_fail(MG_ERR_ARG, "mg_ctx_set_pair_store: null context");
  c->pair_store = enable != 0;
  c->gcap_edges = c->gcap_row = 0;  /* the next evaluation sizes the capacities (and the pair lists) again */
  c->gpstride = 0;
  return MG_OK;
}

int mg_ctx_set_graph(mg_ctx *c, int enable, double headroom) {
  if (!c || headroom < 0.0 || headroom > 10.0) return mg_fail(MG_ERR_ARG, "mg_ctx_set_graph: bad argument");
  c->graph_enable = enable != 0;
  c->graph_headroom = headroom;
  return MG_OK;
}

/* flags of a finished chain (h_flags = misc[0..7]) */
static int chain_flags(mg_ctx *c, int mode) {
  const int *f = c->h_flags;
  if (f[3]) {
    if (mode != MG_CHAIN_SYNC) return MG_ERR_NEIGHBOR_CAP;  /* outgrew the capacities: repeat in sync mode */
    return mg_fail(MG_ERR_HIP, "edge capacity check failed in a synchronised evaluation");
  }
  if (f[1] & 2)
    return mg_fail(MG_ERR_NEIGHBOR, "neighbour list error: an atom has more than 1024 neighbours of one species beside "
                                    "neighbours of another (the pair enumeration's limit)");
  if (f[1]) return mg_fail(MG_ERR_NEIGHBOR, "neighbour list error: a reverse edge is missing, or a cell is too thin for the cutoff");
  c->nedges = mode == MG_CHAIN_SYNC ? f[6] : -1;  /* padded rows: the total is summed when it is asked for */
  c->maxrow = f[0];
  return MG_OK;
}

/* read the stage events of the oldest unread timed evaluation into the sums */
static int timing_read_one(mg_ctx *c) {
  hipEvent_t *e = c->tev[c->tread % MG_TRING];
  HIPCHK(hipEventSynchronize(e[EV_END]));
  float t[6];
  HIPCHK(hipEventElapsedTime(&t[0], e[EV_START], e[EV_NBR]));
  HIPCHK(hipEventElapsedTime(&t[1], e[EV_NBR], e[EV_DESC]));
  HIPCHK(hipEventElapsedTime(&t[2], e[EV_DESC], e[EV_MLP]));
  HIPCHK(hipEventElapsedTime(&t[3], e[EV_MLP], e[EV_FORCE]));
  HIPCHK(hipEventElapsedTime(&t[4], e[EV_FORCE], e[EV_END]));
  HIPCHK(hipEventElapsedTime(&t[5], e[EV_START], e[EV_END]));
  for (int k = 0; k < 6; k++) c->tsum[k] += t[k];
  c->tn++;
  c->tread++;
  return MG_OK;
}

/* events for the next timed evaluation; the oldest outstanding one is read first when the ring is full */
static int timing_slot(mg_ctx *c, hipEvent_t **ev) {
  if (c->tcount - c->tread >= MG_TRING) {
    int r;
    if ((r = timing_read_one(c)) != MG_OK) return r;
  }
  *ev = c->tev[c->tcount % MG_TRING];
  c->tcount++;
  return MG_OK;
}

int mgh_eval_chain(mg_ctx *c, int forces, int mode) {
  const mg_model *m = c->m;
  const int N = c->N;
  hipStream_t st = c->st;
  const float urep = (float)c->urep;
  const int timing = c->timing && c->tev_ok && mode != MG_CHAIN_GRAPH;
  const int *ovf = c->d_misc + 3;
  hipEvent_t *ev = NULL;
  if (timing) {
    int r;
    if ((r = timing_slot(c, &ev)) != MG_OK) return r;
    HIPCHK(hipEventRecord(ev[EV_START], st));
  }
  /* neighbour list: bins, row counts, row offsets; graph mode keeps the flags of misc[1..5] for the caller */
  /* bin counts are zero between evaluations (the bin fill counts them back down); they are cleared explicitly
   * only after an allocation or an evaluation that stopped part way */
  if (c->bins_dirty) {
    KCHK(mgk_clear(c->d_bin_count, c->nbins + 1, NULL, 0, NULL, 0, st));
    c->bins_dirty = 0;
  }
  /* skin lists: graph-mode chains of fixed-cell runs; any other chain clears misc[0..15], which marks them stale */
  const int skin = c->skin_on && c->ccap_row > 0 && mode != MG_CHAIN_SYNC;
  const double half = 0.5 * c->skin * (1.0 - 1e-9);
  KCHK(mgk_prepare(N, c->d_xin, c->d_atom_str, c->d_str, c->d_x, c->d_bin, c->d_bin_count, c->d_misc,
                   mode == MG_CHAIN_GRAPH ? 1 : 16, skin ? c->d_xb : NULL, c->d_mimg, half * half, st));
  KCHK(mgk_scan(c->d_bin_count, c->d_bin_start, c->nbins, c->d_tmp, &c->tmp_bytes, st));
  KCHK(mgk_bins(N, c->nbins, c->d_bin, c->d_bin_start, c->d_bin_count, c->d_x, c->d_spc, c->d_bin_rec, st));
  int cap_row;
  size_t pstride;  /* stored pair lists (forces only): entries per atom, 0 = none */
  if (mode == MG_CHAIN_SYNC) {
    /* row counts, CSR offsets, one read-back to size the buffers, then the fill at the offsets */
    KCHK(mgk_nbr_count(N, c->d_x, c->d_spc, c->d_atom_str, c->d_str, c->d_bin, c->d_bin_start, c->d_bin_rec, c->d_cnt,
                       c->d_rowcnt, c->d_misc, st));
    KCHK(mgk_scan(c->d_rowcnt, c->d_row_off, N, c->d_tmp, &c->tmp_bytes, st));
    int hm[2];
    HIPCHK(hipMemcpyAsync(&hm[0], c->d_row_off + N, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&hm[1], c->d_misc, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->nedges = hm[0];
    c->maxrow = hm[1];
    int r;
    /* allocate with the capacity-mode headroom now, so sizing the capacities after this evaluation never has to
     * reallocate (and drop) the edge buffers it has just filled */
    {
      size_t ne = graph_edge_cap(c, c->nedges);
      const size_t np = padded_edges(c, c->maxrow + (c->graph_headroom > 0.0 ? 8 : 0));
      if (np > ne) ne = np;
      if ((r = ensure_edges(c, ne))) return r;
    }
    cap_row = c->maxrow;
    if ((r = ensure_rhash(c, cap_row + (c->graph_headroom > 0.0 ? 8 : 0)))) return r;
    pstride = forces ? ensure_pairs(c, cap_row + (c->graph_headroom > 0.0 ? 8 : 0), cap_row) : 0;
    KCHK(mgk_check_caps(N, c->d_row_off, c->d_misc, c->nedges, cap_row, st));
    KCHK(mgk_nbr_fill(N, c->d_x, c->d_spc, c->d_atom_str, c->d_str, c->d_bin, c->d_bin_start, c->d_bin_rec, c->d_row_off,
                      c->d_rowcnt, c->d_cnt, c->d_key, c->d_vec, c->d_rev, c->d_rhash, c->d_misc, cap_row, 0, c->wpb_nbr,
                      st));
  } else if (skin) {
    /* capacities known, skin lists: candidates rebuilt when stale, then LIST's test on the candidates */
    cap_row = c->gcap_row;
    pstride = forces ? c->gpstride : 0;
    KCHK(mgk_nbr_cand(N, c->d_x, c->d_spc, c->d_atom_str, c->d_str, c->d_bin, c->d_bin_start, c->d_bin_rec, c->d_crow,
                      c->d_ckey, c->d_misc, c->ccap_row, c->skin, st));
    KCHK(mgk_skin_commit(N, c->d_x, c->d_xb, c->d_mimg, c->d_misc, st));
    KCHK(mgk_nbr_skin(N, c->d_x, c->d_atom_str, c->d_str, c->d_ckey, c->d_crow, c->ccap_row, c->d_mimg, c->d_row_off,
                      c->d_rowcnt, c->d_cnt, c->d_key, c->d_vec, c->d_rev, c->d_rhash, c->d_misc, cap_row, c->wpb_nbr,
                      st));
  } else {
    /* capacities known: one pass, row i at i * gcap_row (the edge buffers hold N * gcap_row slots) */
    cap_row = c->gcap_row;
    pstride = forces ? c->gpstride : 0;
    KCHK(mgk_nbr_fill(N, c->d_x, c->d_spc, c->d_atom_str, c->d_str, c->d_bin, c->d_bin_start, c->d_bin_rec, c->d_row_off,
                      c->d_rowcnt, c->d_cnt, c->d_key, c->d_vec, c->d_rev, c->d_rhash, c->d_misc, cap_row, 1, c->wpb_nbr,
                      st));
  }
  if (timing) HIPCHK(hipEventRecord(ev[EV_NBR], st));
  /* descriptors, MLP, forces */
  KCHK(mgk_desc(N, m->nspc, m->pat | (m->uni << 4), c->d_spc, c->d_row_off, c->d_rowcnt, c->d_cnt, c->d_vec, c->d_basis, c->d_G, c->d_erep,
                urep, cap_row, c->wpb_desc, pstride ? c->d_gpl : NULL, c->d_npl, pstride, c->d_key, c->d_rhash,
                forces ? c->d_rev : NULL, c->d_misc + 1, ovf, st));
  if (timing) HIPCHK(hipEventRecord(ev[EV_DESC], st));
  if (m->mfma_ok)
    KCHK(mgk_mlp_mfma(c->ntiles, c->d_tiles, c->d_perm, c->d_G, c->d_frag, c->d_frag_bf, c->d_frag_lo, c->d_fvec, m->fdesc,
                      c->prec, c->d_Ei, c->d_dEdG, st));
  else
    KCHK(mgk_mlp(N, c->d_spc, c->d_G, c->d_mlp, m->wstride, m->D, m->H1, m->H2, c->d_Ei, c->d_dEdG, st));
  if (timing) HIPCHK(hipEventRecord(ev[EV_MLP], st));
  if (forces) {
    /* shared etas: the force kernel contracts the triplet terms on the matrix cores, with split FP16 operands in
       FP32 precision (bit 6: needs every radial factor <= 1, so no negative eta) or split BF16 in the BF16
       precisions (bit 5); otherwise (per-pair etas) on the VALU in k_force, whose register budget is not capped */
    int fvar = m->pat | (m->uni << 4);
    if (m->uni && c->prec != MG_PREC_FP32) fvar |= 1 << 5;
    else if (m->uni && m->eta_nonneg) fvar |= 1 << 6;
    KCHK(mgk_force(N, m->nspc, fvar, c->d_spc, c->d_row_off, c->d_rowcnt, c->d_cnt, c->d_vec, c->d_basis, c->d_dEdG, c->d_fself,
                   c->d_fe, c->d_vir, urep, cap_row, c->wpb_force, pstride ? c->d_gpl : NULL, c->d_npl, pstride, c->d_rev,
                   c->d_misc + 1, ovf, st));
  }
  if (timing) HIPCHK(hipEventRecord(ev[EV_FORCE], st));
  /* force gather and the FP64 energy/virial reductions in one launch */
  KCHK(mgk_assemble(N, forces, c->d_row_off, c->d_rowcnt, c->d_fself, c->d_fe, c->d_F, ovf, c->S, c->nchunk,
                    c->d_chunk_str, c->d_chunk_cnt, c->d_str, c->d_Ei, c->d_erep, c->d_vir, c->d_part, c->d_E, c->d_U,
                    st));
  if (timing) HIPCHK(hipEventRecord(ev[EV_END], st));
  if (mode == MG_CHAIN_GRAPH) return MG_OK;
  /* misc[0..7] (row maximum, error, overflow, edge total) to pinned host memory; ASYNC leaves the check to
   * mgh_settle */
  HIPCHK(hipMemcpyAsync(c->h_flags, c->d_misc, 8 * sizeof(int), hipMemcpyDeviceToHost, st));
  if (mode == MG_CHAIN_ASYNC) {
    c->pending = 1;
    c->pending_forces = forces;
    return MG_OK;
  }
  HIPCHK(hipStreamSynchronize(st));
  return chain_flags(c, mode);
}

int mgh_eval(mg_ctx *c, int flags, int async) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  if (c->S == 0) return mg_fail(MG_ERR_ARG, "no structure set");
  HIPCHK(hipSetDevice(c->device));
  if (c->prec != MG_PREC_FP32 && !c->m->mfma_ok) return mg_fail(MG_ERR_UNSUPPORTED, "BF16 MLP needs hidden layers <= 64 wide");
  const int forces = (flags & MG_EVAL_FORCES) != 0;
  /* repeated evaluations of the same atoms reuse the capacities of the last one (no host round trip in the middle
   * of the chain); a step that outgrows them is detected on the device and evaluated again synchronised.  An
   * asynchronous evaluation supersedes a pending one: only the last evaluation's results can be read. */
  int r = MG_ERR_NEIGHBOR_CAP;
  c->pending = 0;
  if (c->graph_enable && c->gcap_edges > 0) {
    r = mgh_eval_chain(c, forces, async ? MG_CHAIN_ASYNC : MG_CHAIN_CAPS);
    if (async && r == MG_OK) {
      c->have_forces = forces;
      c->have_g = 1;
      return MG_OK;
    }
  }
  if (r == MG_ERR_NEIGHBOR_CAP) {
    if ((r = mgh_eval_chain(c, forces, MG_CHAIN_SYNC)) != MG_OK) return r;
    if ((r = mgh_graph_caps(c, c->nedges, c->maxrow)) != MG_OK) return r;
  } else if (r != MG_OK) {
    return r;
  }
  c->have_forces = forces;
  c->have_g = 1;
  return MG_OK;
}

int mgh_settle(mg_ctx *c) {
  if (!c || !c->pending) return MG_OK;
  c->pending = 0;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->st));
  int r = chain_flags(c, MG_CHAIN_CAPS);
  if (r == MG_ERR_NEIGHBOR_CAP) {
    if ((r = mgh_eval_chain(c, c->pending_forces, MG_CHAIN_SYNC)) != MG_OK) return r;
    r = mgh_graph_caps(c, c->nedges, c->maxrow);
  }
  return r;
}

int mg_eval(mg_ctx *c, int flags) { return mgh_eval(c, flags, 1); }

int mg_get(mg_ctx *c, int s, double *E, double *EA, double *F, double *U, double *X) {
  if (!c || s < 0 || s >= c->S) return mg_fail(MG_ERR_ARG, "mg_get: bad structure index");
  if (!c->have_g) return mg_fail(MG_ERR_ARG, "mg_get: the structure has not been evaluated yet");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  const int a0 = c->h_atom_off[s], n = c->h_atom_off[s + 1] - a0;
  if (E) HIPCHK(hipMemcpy(E, c->d_E + s, sizeof(double), hipMemcpyDeviceToHost));
  if (U) {
    if (!c->have_forces) return mg_fail(MG_ERR_ARG, "mg_get: virial requested but the last eval had no forces");
    HIPCHK(hipMemcpy(U, c->d_U + 6 * s, 6 * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (F) {
    if (!c->have_forces) return mg_fail(MG_ERR_ARG, "mg_get: forces requested but the last eval had no forces");
    HIPCHK(hipMemcpy(F, c->d_F + 3 * (size_t)a0, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (X) HIPCHK(hipMemcpy(X, c->d_x + 3 * (size_t)a0, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  if (EA && n > 0) {
    float *tmp = (float *)malloc(sizeof(float) * n);
    hipError_t e = hipMemcpy(tmp, c->d_Ei + a0, sizeof(float) * n, hipMemcpyDeviceToHost);
    for (int i = 0; i < n; i++) EA[i] = tmp[i];
    free(tmp);
    if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "mg_get: %s", hipGetErrorString(e));
  }
  return MG_OK;
}

int mg_get_batch(mg_ctx *c, double *E, double *U, double *F) {
  if (!c) return mg_fail(MG_ERR_ARG, "null context");
  if (!c->have_g) return mg_fail(MG_ERR_ARG, "mg_get_batch: the batch has not been evaluated yet");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  if (E) HIPCHK(hipMemcpy(E, c->d_E, c->S * sizeof(double), hipMemcpyDeviceToHost));
  if ((U || F) && !c->have_forces) return mg_fail(MG_ERR_ARG, "mg_get_batch: last eval had no forces");
  if (U) HIPCHK(hipMemcpy(U, c->d_U, 6 * c->S * sizeof(double), hipMemcpyDeviceToHost));
  if (F) HIPCHK(hipMemcpy(F, c->d_F, 3 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_get_batch_x(mg_ctx *c, double *X) {
  if (!c || !X) return mg_fail(MG_ERR_ARG, "mg_get_batch_x: bad argument");
  if (!c->have_g) return mg_fail(MG_ERR_ARG, "mg_get_batch_x: the batch has not been evaluated yet");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  if (c->N > 0) HIPCHK(hipMemcpy(X, c->d_x, 3 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_get_neighbors(mg_ctx *c, int *nedges, int *row_off, int *j, int *img) {
  if (!c || !nedges) return mg_fail(MG_ERR_ARG, "mg_get_neighbors: bad argument");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  /* rows are (start, count) on the device: CSR after a synchronised evaluation, a fixed stride after one with known
   * capacities; the export is CSR either way */
  const int N = c->N;
  int *st = (int *)malloc(sizeof(int) * (N > 0 ? N : 1)), *ct = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  hipError_t e = hipMemcpy(st, c->d_row_off, sizeof(int) * N, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ct, c->d_rowcnt, sizeof(int) * N, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { free(st); free(ct); return mg_fail(MG_ERR_HIP, "%s", hipGetErrorString(e)); }
  size_t total = 0, extent = 0;
  for (int i = 0; i < N; i++) {
    total += (size_t)ct[i];
    if ((size_t)st[i] + (size_t)ct[i] > extent) extent = (size_t)st[i] + (size_t)ct[i];
  }
  c->nedges = (int)total;
  *nedges = c->nedges;
  if (!row_off && !j && !img) { free(st); free(ct); return MG_OK; }
  unsigned long long *k = (unsigned long long *)malloc(8 * (extent > 0 ? extent : 1));
  if (j || img) e = hipMemcpy(k, c->d_key, 8 * extent, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { free(st); free(ct); free(k); return mg_fail(MG_ERR_HIP, "%s", hipGetErrorString(e)); }
  int t = 0;
  for (int i = 0; i < N; i++) {
    if (row_off) row_off[i] = t;
    for (int q = 0; q < ct[i]; q++, t++) {
      const unsigned long long kk = k[st[i] + q];
      if (j) j[t] = (int)((kk >> 30) & 0x3fffffffull);
      if (img) {
        unsigned p = (unsigned)(kk & 0x3fffffffull);
        img[3 * t] = (int)(p & 1023u) - MG_IMG_OFF;
        img[3 * t + 1] = (int)((p >> 10) & 1023u) - MG_IMG_OFF;
        img[3 * t + 2] = (int)((p >> 20) & 1023u) - MG_IMG_OFF;
      }
    }
  }
  if (row_off) row_off[N] = t;
  free(st); free(ct); free(k);
  return MG_OK;
}

int mg_get_descriptors(mg_ctx *c, float *G) {
  if (!c || !G) return mg_fail(MG_ERR_ARG, "mg_get_descriptors: bad argument");
  if (!c->have_g) return mg_fail(MG_ERR_ARG, "mg_get_descriptors: no evaluation yet");
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  HIPCHK(hipMemcpy(G, c->d_G, sizeof(float) * (size_t)c->N * c->m->D, hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_get_lattice(mg_ctx *c, int s, double *lat) {
  if (!c || !lat || s < 0 || s >= c->S) return mg_fail(MG_ERR_ARG, "mg_get_lattice: bad argument");
  memcpy(lat, c->h_lat + 9 * s, 9 * sizeof(double));
  return MG_OK;
}

int mg_get_timing(mg_ctx *c, float *t) {
  if (!c || !t) return mg_fail(MG_ERR_ARG, "mg_get_timing: bad argument");
  int r;
  if ((r = mgh_settle(c)) != MG_OK) return r;
  while (c->tread < c->tcount)
    if ((r = timing_read_one(c)) != MG_OK) return r;
  for (int k = 0; k < 6; k++) t[k] = c->tn ? (float)(c->tsum[k] / c->tn) : 0.0f;
  for (int k = 0; k < 6; k++) c->tsum[k] = 0.0;
  c->tn = 0;
  return MG_OK;
}

/* ------------------------------------------------------------------ CALL_MAISE drop-in */
typedef struct {
  mg_model *m;
  mg_ctx *c;
} cm_state;

static void cm_die(const char *what) {
  fprintf(stderr, "%s: %s\n", what, mg_last_error());
  exit(1);
}

double CALL_MAISE(ANN *R, PRS *P, PRS *W, LNK *L, Cell *C, int CODE, int N, int NM, int ND, int NP, int XT, int *ATMN,
                  double *LAT, double *X, double *FRC, double *STR) {
  (void)R; (void)P; (void)W; (void)CODE; (void)NP;
  const double eV2GPa = 160.2176565;
  C->p /= eV2GPa; /* every call, as src/cmod.c:1181 */
  cm_state *s = L->B != 0 ? (cm_state *)L->storage_[0] : NULL;
  if (!s) {
    s = (cm_state *)calloc(1, sizeof(cm_state));
    const char *path = getenv("MAISE_MODEL");
    if (!path || !*path) path = "model";
    if (mg_model_load(path, &s->m) != MG_OK) cm_die("ERROR reading model");
    if (mg_ctx_create(s->m, 0, &s->c) != MG_OK) cm_die("ERROR creating GPU context");
    const char *u = getenv("MAISE_UREP");
    mg_ctx_set_urep(s->c, (u && *u) ? atof(u) : 100.0);
    L->storage_[0] = s;
    L->B = 1;
  }
  if (mg_set_structure(s->c, N, ND, XT, ATMN, LAT, X) != MG_OK) cm_die("CALL_MAISE");
  if (mg_eval(s->c, MG_EVAL_FORCES) != MG_OK) cm_die("CALL_MAISE");
  const char *strict = getenv("MAISE_NM_STRICT");
  if (strict && *strict == '1' && s->c->maxrow >= NM) {
    fprintf(stderr, "Increase 'Max number of nearest neighbors' (%d needed)\n", s->c->maxrow + 1);
    exit(1);
  }
  double E, U[6];
  if (mg_get(s->c, 0, &E, NULL, FRC, U, NULL) != MG_OK) cm_die("CALL_MAISE");
  for (int k = 0; k < 3 * N; k++) if (fabs(FRC[k]) < 1e-8) FRC[k] = 0.0;
  for (int q = 0; q < 6; q++) STR[q] = fabs(U[q]) < 1e-8 ? 0.0 : U[q];
  double H = E;
  if (ND == 3) {
    double Lr[9];
    for (int k = 0; k < 9; k++) Lr[k] = LAT[k] * 1.0;
    H += C->p * cell_volume(Lr);
  }
  return H;
}

void maise_gpu_free(LNK *L) {
  if (!L) return;
  cm_state *s = (cm_state *)L->storage_[0];
  if (s) {
    mg_ctx_free(s->c);
    mg_model_free(s->m);
    free(s);
  }
  L->storage_[0] = NULL;
  L->B = 0;
}

/* ------------------------------------------------------------------ fixed-scratch evaluation */
/* An evaluation that allocates nothing: the device buffers of a throw-away context are slices of the library's
 * static device array (mgs_scratch), the host temporaries come from a static host area, the chain runs on the null
 * stream in its synchronised form, and any buffer the structure would need beyond the scratch fails with an error
 * instead of being allocated.  Calls that overlap on one device (or on the host area) race. */
#define MG_STATIC_MAX_ATOMS 32768
#define MG_STATIC_MAX_ROW 256    /* longest neighbour row the reverse-lookup tables are laid out for */
static unsigned char g_static_hwork[16 * MG_STATIC_MAX_ATOMS + 4096];

typedef struct { unsigned char *base; size_t used, size; } mg_carve;
static void *carve(mg_carve *a, size_t bytes) {
  const size_t off = (a->used + 255) & ~(size_t)255;
  if (off + bytes > a->size) return NULL;
  a->used = off + bytes;
  return a->base + off;
}

int mg_eval_static(const mg_model *m, int N, int nd, int xt, const int *spc, const double *lat, const double *x,
                   double urep, double *E, double *EA, double *F, double *U) {
  if (!m || N < 1 || !spc || !lat || !x) return mg_fail(MG_ERR_ARG, "mg_eval_static: bad argument");
  if (N > MG_STATIC_MAX_ATOMS)
    return mg_fail(MG_ERR_ARG, "mg_eval_static: %d atoms exceed the fixed-scratch limit of %d", N, MG_STATIC_MAX_ATOMS);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  void *base = NULL;
  size_t bytes = 0;
  if (mgs_scratch(dev, &base, &bytes) != 0)
    return mg_fail(MG_ERR_HIP, "mg_eval_static: the scratch array of device %d could not be looked up", dev);
  mg_ctx cc;
  memset(&cc, 0, sizeof cc);
  mg_ctx *c = &cc;
  int flags[16];
  mg_struct_dev h_str[1];
  int h_atom_off[2], h_nd[1];
  double h_lat[9];
  c->m = m;
  c->device = dev;
  c->st = 0;  /* the null stream */
  c->urep = urep;
  c->prec = MG_PREC_FP32;
  c->no_alloc = 1;
  c->pair_store = 0;  /* the force kernel enumerates the pairs itself: no O(N M^2) pair lists in the scratch */
  c->graph_enable = 0;
  c->graph_headroom = 0.0;
  c->h_flags = flags;
  c->h_str = h_str;
  c->h_atom_off = h_atom_off;
  c->h_nd = h_nd;
  c->h_lat = h_lat;
  c->cap_str = 1;
  c->h_work = g_static_hwork;
  c->h_work_bytes = sizeof g_static_hwork;
  mg_carve a = {(unsigned char *)base, 0, bytes};
#define CARVE(ptr, nb)                                                                                              \
  do {                                                                                                              \
    void *p_ = carve(&a, (nb));                                                                                     \
    if (!p_) return mg_fail(MG_ERR_ARG, "mg_eval_static: %d atoms do not fit the %zu-byte fixed scratch", N, bytes); \
    *(void **)&(ptr) = p_;                                                                                          \
  } while (0)
  const size_t n = (size_t)N, D = (size_t)m->D;
  /* model tables, copied in by every call (the scratch keeps whatever the last call left) */
  CARVE(c->d_basis, sizeof(mg_basis_dev));
  CARVE(c->d_mlp, (size_t)m->wstride * m->nspc * sizeof(float));
  if (m->mfma_ok) {
    CARVE(c->d_frag, m->frag_n * sizeof(float));
    CARVE(c->d_frag_bf, m->frag_bf_n * sizeof(uint16_t));
    CARVE(c->d_frag_lo, m->frag_bf_n * sizeof(uint16_t));
    CARVE(c->d_fvec, m->fvec_n * sizeof(float));
  }
  CARVE(c->d_misc, 32 * sizeof(int));
  CARVE(c->d_str, sizeof(mg_struct_dev));
  CARVE(c->d_E, sizeof(double));
  CARVE(c->d_U, 6 * sizeof(double));
  /* per atom (the buffers ensure_atoms would allocate) */
  CARVE(c->d_atom_str, n * sizeof(int));
  CARVE(c->d_spc, n * sizeof(int));
  CARVE(c->d_xin, n * 3 * sizeof(double));
  CARVE(c->d_x, n * 3 * sizeof(double));
  CARVE(c->d_bin, 2 * n * sizeof(int));
  CARVE(c->d_bin_rec, n * 32);
  CARVE(c->d_cnt, n * 4 * sizeof(int));
  CARVE(c->d_npl, n * MG_NCLS * sizeof(int));
  CARVE(c->d_rowcnt, n * sizeof(int));
  CARVE(c->d_row_off, (n + 1) * sizeof(int));
  CARVE(c->d_G, n * D * sizeof(float));
  CARVE(c->d_dEdG, n * D * sizeof(float));
  CARVE(c->d_Ei, n * sizeof(float));
  CARVE(c->d_erep, n * sizeof(float));
  CARVE(c->d_vir, n * 6 * sizeof(float));
  CARVE(c->d_fself, n * 16);
  CARVE(c->d_F, n * 3 * sizeof(double));
  c->cap_atoms = n;
  /* reduction chunks, MLP tiles, bins (at most 4 N + 64 for one structure) and the scan workspace */
  const size_t nchunk = (n + MG_RCHUNK - 1) / MG_RCHUNK + 16;
  CARVE(c->d_chunk_str, nchunk * sizeof(int));
  CARVE(c->d_chunk_cnt, nchunk * sizeof(int));
  CARVE(c->d_part, nchunk * 7 * sizeof(double));
  c->cap_chunk = nchunk;
  const size_t ntile = n / 32 + MG_MAXSPC + 64;
  CARVE(c->d_tiles, ntile * 4 * sizeof(int));
  CARVE(c->d_perm, (ntile * 32 + 64) * sizeof(int));
  c->cap_tiles = ntile;
  const size_t nbins = 4 * n + 64 + 16;
  CARVE(c->d_bin_count, (nbins + 1) * sizeof(int));
  CARVE(c->d_bin_start, (nbins + 1) * sizeof(int));
  c->cap_bins = nbins;
  {
    size_t need = 0;
    if (mgk_scan(NULL, NULL, (int)(nbins > n ? nbins : n), NULL, &need, 0) != 0)
      return mg_fail(MG_ERR_HIP, "mg_eval_static: scan size query failed");
    CARVE(c->d_tmp, need);
    c->tmp_bytes = need;
  }
  /* reverse-lookup tables for rows of up to MG_STATIC_MAX_ROW, then the edge buffers take the rest (44 bytes per edge:
   * key, vector, reverse index, force) */
  const size_t slots = (size_t)mgk_rev_slots(MG_STATIC_MAX_ROW);
  CARVE(c->d_rhash, n * slots * sizeof(uint16_t));
  c->cap_rhash = n * slots;
  {
    const size_t left = a.size > a.used + 4 * 256 ? a.size - a.used - 4 * 256 : 0;
    size_t ne = left / 44;
    if (ne > 0x7fffffff) ne = 0x7fffffff;
    CARVE(c->d_key, ne * 8);
    CARVE(c->d_vec, ne * 16);
    CARVE(c->d_rev, ne * sizeof(int));
    CARVE(c->d_fe, ne * 16);
    c->cap_edges = ne;
  }
#undef CARVE
  HIPCHK(hipMemcpyAsync(c->d_basis, &m->basis, sizeof(mg_basis_dev), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_mlp, m->mlp, (size_t)m->wstride * m->nspc * sizeof(float), hipMemcpyHostToDevice, c->st));
  if (m->mfma_ok) {
    HIPCHK(hipMemcpyAsync(c->d_frag, m->frag, m->frag_n * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_frag_bf, m->frag_bf, m->frag_bf_n * sizeof(uint16_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_frag_lo, m->frag_lo, m->frag_bf_n * sizeof(uint16_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_fvec, m->fvec, m->fvec_n * sizeof(float), hipMemcpyHostToDevice, c->st));
  }
  int off[2] = {0, N}, r;
  if ((r = mgh_upload(c, 1, off, &nd, &xt, lat, spc, x)) != MG_OK) return r;
  if ((r = mgh_eval_chain(c, 1, MG_CHAIN_SYNC)) != MG_OK) return r;
  if (E) HIPCHK(hipMemcpy(E, c->d_E, sizeof(double), hipMemcpyDeviceToHost));
  if (U) HIPCHK(hipMemcpy(U, c->d_U, 6 * sizeof(double), hipMemcpyDeviceToHost));
  if (F) HIPCHK(hipMemcpy(F, c->d_F, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (EA) {
    float *tmp = (float *)g_static_hwork;  /* the upload's temporaries are no longer needed */
    HIPCHK(hipMemcpy(tmp, c->d_Ei, n * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < N; i++) EA[i] = tmp[i];
  }
  return MG_OK;
}
