/* mg_host.h -- host-side internals of libmaise_gpu (model, context, kernel launchers). */
#ifndef MG_HOST_H
#define MG_HOST_H

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__
#endif
#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "../../include/maise_gpu.h"
#include "mg_internal.h"

struct mg_model {
  int nspc, spcz[MG_MAXSPC];
  int D, H1, H2, nsym, ng2, ng4, nw;
  double rc_list;       /* neighbour-list cutoff (A) */
  mg_basis_dev basis;   /* FP32 basis tables */
  float *mlp;           /* per species: W0' [H1][D], b0, W1 [H2][H1], b1, W2 [H2], b2, xoff [D] */
  double *mlp64;        /* same layout in FP64 */
  int wstride;          /* floats per species block (multiple of 4) */
};

int mg_fail(int code, const char *fmt, ...);

/* kernel launchers (mg_kernels.hip) */
int mgk_scan(const int *in, int *out, int n, void *tmp, size_t *tmp_bytes, hipStream_t st);
int mgk_prepare(int N, const double *xin, const int *atom_str, const mg_struct_dev *str, double *x, int *bin,
                int *bin_count, hipStream_t st);
int mgk_bins(int N, int nbins, const int *bin, const int *bin_start, int *bin_cur, int *bin_atoms, hipStream_t st);
int mgk_nbr_count(int N, const double *x, const int *spc, const int *atom_str, const mg_struct_dev *str, const int *bin,
                  const int *bin_start, const int *bin_atoms, int *cnt, int *rowcnt, int *maxrow, hipStream_t st);
int mgk_nbr_fill(int N, const double *x, const int *spc, const int *atom_str, const mg_struct_dev *str, const int *bin,
                 const int *bin_start, const int *bin_atoms, const int *row_off, void *key_u, void *vec_u, void *key,
                 void *vec, int *rev, int *err, int maxrow, hipStream_t st);
int mgk_desc(int N, int S, int ne, const int *spc, const int *row_off, const int *cnt, const void *vec,
             const mg_basis_dev *bp, float *G, float *erep, float urep, int maxrow, hipStream_t st);
int mgk_mlp(int N, const int *spc, const float *G, const float *W, int wstride, int D, int H1, int H2, float *Ei,
            float *dEdG, hipStream_t st);
int mgk_force(int N, int S, int ne, const int *spc, const int *row_off, const int *cnt, const void *vec,
              const mg_basis_dev *bp, const float *dEdG, void *fself, void *fe, float *vir, float urep, int maxrow,
              const int *rev, double *F, hipStream_t st);
int mgk_reduce(int nstr, const mg_struct_dev *str, const float *Ei, const float *erep, const float *vir, int have_vir,
               double *E, double *U, hipStream_t st);

#endif
