/* mg_host.h -- host-side internals of libmaise_gpu (model, context, kernel launchers). */
#ifndef MG_HOST_H
#define MG_HOST_H

#include <hip/hip_runtime_api.h>  /* gcc builds the host code with -D__HIP_PLATFORM_AMD__ (maise_amd/build.py) */
#include <stddef.h>
#include <stdint.h>

#include "../../include/maise_gpu.h"
#include "mg_internal.h"

struct mg_model {
  int nspc, spcz[MG_MAXSPC];
  char id[64];          /* model unique ID of the file header */
  int D, H1, H2, nsym, ng2, ng4, nw;
  double rc_list;       /* neighbour-list cutoff (A) */
  mg_basis_dev basis;   /* FP32 basis tables */
  float *mlp;           /* per species: W0' [H1][D], b0, W1 [H2][H1], b1, W2 [H2], b2, xoff [D] */
  double *mlp64;        /* same layout in FP64 */
  int wstride;          /* floats per species block (multiple of 4) */
  int pat;              /* G4 kernel specialisation: 0 generic, 1 43-function, 2 22-function table */
  int uni;              /* 1 when every species pair has the same G4 etas */
  int eta_nonneg;       /* 1 when no G4 eta is negative (every radial factor exp(-eta r^2) <= 1) */
  /* MFMA MLP: weight fragments in the kernel's operand order (FP32 and BF16 sets) and per-unit vectors */
  int mfma_ok;          /* H1, H2 <= 64 */
  int fdesc[20];
  float *frag;
  uint16_t *frag_bf, *frag_lo;
  float *fvec;
  size_t frag_n, frag_bf_n, fvec_n;
};

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return mg_fail(MG_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

enum { EV_START, EV_NBR, EV_DESC, EV_MLP, EV_FORCE, EV_END, EV_N };
#define MG_TRING 32  /* evaluations whose stage events can be outstanding before the oldest is read */

struct mg_ctx;
/* (Re)load S structures into the context: per-structure tables, bins, species tiles, positions (xt 0 = fractional) */
int mgh_upload(struct mg_ctx *c, int S, const int *atom_off, const int *nd, const int *xt, const double *lat,
               const int *spc, const double *x);

struct mg_ctx {
  const mg_model *m;
  int device;
  hipStream_t st;
  double urep;
  int prec;
  int timing;
  int wpb_desc, wpb_force; /* waves (atoms) per block of the descriptor and force kernels; 0 = by LDS size */
  int wpb_nbr;              /* waves (atoms) per block of the neighbour fill; 0 = 4, fewer when LDS is short */
  int autotune;             /* relaxation / MD tune the launch shape before their first capture */
  /* skin neighbour lists (fixed-cell MD and FIRE in graph mode): candidates within rc + skin, rebuilt when an atom
   * has moved skin/2 since their build; every step filters them with LIST's own test */
  double skin;              /* A; 0 = off */
  int skin_on;              /* this run uses them (set by relaxation / MD for fixed cells) */
  int ccap_row;             /* candidate row capacity (0 = buffers not sized) */
  int skin_need;            /* longest candidate row seen by an overflowed step */
  void *d_ckey;             /* u64[N * ccap_row] candidate keys */
  int *d_crow;              /* int[N] candidate row lengths */
  double *d_xb;             /* double[3N] wrapped positions at the candidates' build */
  unsigned *d_mimg;         /* unsigned[N] packed image shifts since the build */
  size_t cap_skin_atoms, cap_ckey;
  int tuned;                /* the shape was tuned on this context's batch */
  int gcap_edges, gcap_row; /* graph-mode capacities: edges allocated, longest row the LDS tiles hold */
  int graph_enable;         /* relaxation / MD replay a captured iteration (default 1) */
  double graph_headroom;    /* fractional headroom of the graph-mode capacities (default 1/8) */
  int pair_store;           /* the descriptor kernel stores its G4 pair lists for the force kernel (default 1) */
  size_t pair_budget;       /* largest pair-list allocation in bytes (1/16 of the device memory) */
  size_t gpstride;          /* pair-list entries per atom at the graph-mode row capacity (0: not stored) */
  hipEvent_t tev[MG_TRING][EV_N]; /* stage events of the last MG_TRING timed evaluations (ring) */
  int tev_ok, tcount, tread;      /* events created; evaluations recorded; evaluations read */
  double tsum[6];                 /* stage milliseconds summed over the evaluations read */
  int tn;
  /* asynchronous evaluation: the flags of the last enqueued chain land in pinned h_flags; the first call that
   * reads results checks them (and evaluates again synchronised if the capacities overflowed) */
  int pending, pending_forces;
  int *h_flags;
  /* structures */
  int S, N;
  int *h_atom_off, *h_nd;
  double *h_lat;
  mg_struct_dev *h_str;
  int nbins;
  /* device */
  mg_basis_dev *d_basis;
  float *d_mlp;
  mg_struct_dev *d_str;
  int *d_atom_str, *d_spc;
  double *d_xin, *d_x;
  int *d_bin, *d_bin_count, *d_bin_start;
  int bins_dirty;           /* bin counts must be cleared before the next evaluation */
  void *d_bin_rec;          /* atoms in bin order: position, species, index (32 B each) */
  /* misc: [0] longest row of the evaluation, [1] error flag (reverse edge missing), [2] converged-structure count,
   * [3] capacity overflow flag, [4] largest edge total and [5] longest row seen since misc was last cleared,
   * [6] edge total of a synchronised evaluation */
  int *d_cnt, *d_rowcnt, *d_row_off, *d_misc;
  void *d_key, *d_vec;
  int *d_rev;
  uint16_t *d_rhash;        /* per-row hash of neighbour (atom, image) -> position + 1, mgk_rev_slots entries per row */
  size_t cap_rhash;
  float *d_G, *d_dEdG, *d_Ei, *d_erep, *d_vir;
  void *d_fself, *d_fe;
  unsigned *d_gpl;          /* stored G4 pair lists, pstride entries per atom (j | k << 12 | tag << 24) */
  int *d_npl;               /* pairs per atom and species class [N][MG_NCLS] */
  size_t cap_gpl;
  double *d_F, *d_E, *d_U, *d_part;
  float *d_frag, *d_fvec;
  void *d_frag_bf, *d_frag_lo;
  int *d_perm, *d_tiles, ntiles;
  size_t cap_tiles;
  int *d_chunk_str, nchunk;
  int *d_chunk_cnt;         /* per structure (at its first chunk): finished chunk blocks of k_assemble, zero between */
  size_t cap_chunk;
  void *d_tmp;
  size_t tmp_bytes;
  /* capacities (elements) */
  size_t cap_atoms, cap_str, cap_bins, cap_edges;
  int nedges, maxrow;
  int have_forces, have_g;
  /* fixed-scratch evaluation (mg_eval_static): every device buffer is a slice of the library's static device array
   * and the host temporaries of the upload come from h_work; any attempt to grow a buffer fails instead */
  int no_alloc;
  unsigned char *h_work;
  size_t h_work_bytes;
  /* dynamics state (mg_dyn.c) */
  double *d_v;          /* velocities [3N] (A/fs) */
  double *d_mass;       /* per-atom mass [N] (eV fs^2 / A^2) */
  double *d_dyn;        /* scalars: [0] Nose-Hoover xi, [1] kinetic energy, ... */
  double *d_fire;       /* per-structure FIRE state [S][4]: dt, alpha, npos, done */
  double *d_trace;      /* per-step energies */
  double *d_cellv;       /* cell-relaxation velocities [S][9] */
  double *d_snap_x, *d_snap_v, *d_snap_fire, *d_snap_dyn, *d_snap_cellv; /* state at the last graph checkpoint */
  mg_struct_dev *d_snap_str;                                             /* (lattices change in cell relaxation) */
  size_t cap_dyn_atoms, cap_dyn_str, cap_trace;
  /* L-BFGS relaxation (mg_relax_lbfgs): history ring [m][3N] of steps (d_lb_s) and force changes (d_lb_y), forces
   * at the previous positions [3N], 1/(y.s) [S][m], their snapshots at the last graph checkpoint, a gather buffer
   * and an index map for re-packing the batch; lb_stride = 3N of the batch the rings are laid out for */
  double *d_lb_s, *d_lb_y, *d_lb_f0, *d_lb_rho, *d_lb_tmp;
  double *d_snap_lb_s, *d_snap_lb_y, *d_snap_lb_f0, *d_snap_lb_rho;
  int *d_lb_map;
  size_t cap_lb_ring, cap_lb_atoms, cap_lb_str;
  int lb_on, lb_m;
  size_t lb_stride;
  /* last relaxation's wall time by segment between host checks: (iterations done, active structures, active
   * atoms, seconds) per record */
  double *rlog;
  int rlog_n, rlog_cap;
};


int mg_fail(int code, const char *fmt, ...);
/* per-device address and size of the static scratch array (mg_static.hip) */
int mgs_scratch(int device, void **ptr, size_t *bytes);
int mg_dev_alloc(void **p, size_t bytes);

/* kernel launchers (mg_kernels.hip) */
int mgk_scan(const int *in, int *out, int n, void *tmp, size_t *tmp_bytes, hipStream_t st);
int mgk_nbr_cand(int N, const double *x, const int *spc, const int *atom_str, const mg_struct_dev *str, const int *bin,
                 const int *bin_start, const void *brec, int *crow, void *ckey, int *misc, int ccap, double skin,
                 hipStream_t st);
int mgk_skin_commit(int N, const double *x, double *xb, unsigned *mimg, int *misc, hipStream_t st);
int mgk_nbr_skin(int N, const double *x, const int *atom_str, const mg_struct_dev *str, const void *ckey,
                 const int *crow, int ccap, const unsigned *mimg, int *row_off, int *rowcnt, int *cnt, void *key,
                 void *vec, int *rev, uint16_t *rhash, int *misc, int maxrow, int wpb_req, hipStream_t st);
int mgk_prepare(int N, const double *xin, const int *atom_str, const mg_struct_dev *str, double *x, int *bin,
                int *bin_count, int *misc, int nclear, const double *xb, unsigned *mimg, double lim2,
                hipStream_t st);
int mgk_bins(int N, int nbins, const int *bin, const int *bin_start, int *bin_count, const double *x, const int *spc,
             void *brec, hipStream_t st);
int mgk_nbr_count(int N, const double *x, const int *spc, const int *atom_str, const mg_struct_dev *str, const int *bin,
                  const int *bin_start, const void *brec, int *cnt, int *rowcnt, int *misc, hipStream_t st);
int mgh_autotune_once(struct mg_ctx *c);
int mgk_nbr_fill(int N, const double *x, const int *spc, const int *atom_str, const mg_struct_dev *str, const int *bin,
                 const int *bin_start, const void *brec, int *row_off, int *rowcnt, int *cnt, void *key, void *vec,
                 int *rev, uint16_t *rhash, int *misc, int maxrow, int padded, int wpb, hipStream_t st);
int mgk_rev_slots(int maxrow);
int mgk_desc(int N, int S, int variant, const int *spc, const int *row_off, const int *rowcnt, const int *cnt,
             const void *vec,
             const mg_basis_dev *bp, float *G, float *erep, float urep, int maxrow, int wpb, unsigned *gpl, int *npl,
             size_t pstride, const void *key, const uint16_t *rhash, int *rev, int *err, const int *ovf,
             hipStream_t st);
int mgk_mlp_mfma(int ntiles, const void *tiles, const int *perm, const float *G, const float *frag, const void *frag_bf,
                 const void *frag_lo, const float *vecs, const int *fdesc, int prec, float *Ei, float *dEdG,
                 hipStream_t st);
int mgk_mlp(int N, const int *spc, const float *G, const float *W, int wstride, int D, int H1, int H2, float *Ei,
            float *dEdG, hipStream_t st);
int mgk_force(int N, int S, int variant, const int *spc, const int *row_off, const int *rowcnt, const int *cnt,
              const void *vec,
              const mg_basis_dev *bp, const float *dEdG, void *fself, void *fe, float *vir, float urep, int maxrow,
              int wpb, const unsigned *gpl, const int *npl, size_t pstride, const int *rev, int *err, const int *ovf,
              hipStream_t st);
int mgk_clear(int *a, int na, int *b, int nb, int *c, int nc, hipStream_t st);
int mgk_check_caps(int N, const int *row_off, int *misc, int cap_edges, int cap_row, hipStream_t st);

/* the evaluation chain.  MG_CHAIN_SYNC reads the edge count back and sizes the edge buffers exactly (one host
 * synchronisation in the middle, one at the end); MG_CHAIN_GRAPH uses the capacities gcap_edges / gcap_row set
 * beforehand, never synchronises or allocates, and leaves the overflow flag misc[3] for the caller to check, so
 * it can be captured into a HIP graph. */
enum { MG_CHAIN_SYNC = 0, MG_CHAIN_GRAPH = 1, MG_CHAIN_CAPS = 2, MG_CHAIN_ASYNC = 3 };
/* CAPS: graph-mode capacities, not captured, flags checked at the end; ASYNC: CAPS whose flags are copied to the
 * host without waiting (checked by mgh_settle) */
int mgh_eval_chain(struct mg_ctx *c, int forces, int mode);
/* reciprocal rows R = (L1xL2, L2xL0, L0xL1) / (L0.(L1xL2)) in the reference's operation order */
void mgh_reciprocal(const double *L, double *R);
/* evaluate; async = 0 waits and checks the flags before returning (every internal caller) */
int mgh_eval(struct mg_ctx *c, int flags, int async);
/* finish the pending asynchronous evaluation, if any: wait, check its flags, evaluate again synchronised when the
 * capacities overflowed */
int mgh_settle(struct mg_ctx *c);
/* size the graph-mode capacities from the last evaluation with headroom, allocating edge buffers */
int mgh_graph_caps(struct mg_ctx *c, int need_edges, int need_row);
int mgk_assemble(int N, int forces, const int *row_off, const int *rowcnt, const void *fself, const void *fe, double *F,
                 const int *ovf, int nstr, int nchunk, const int *chunk_str, int *chunk_cnt, const mg_struct_dev *str,
                 const float *Ei, const float *erep, const float *vir, double *part, double *E, double *U,
                 hipStream_t st);

#endif
