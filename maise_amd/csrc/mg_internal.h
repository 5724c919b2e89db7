/*
 * mg_internal.h -- data structures shared by the host C code and the HIP kernels of libmaise_gpu.
 *
 * Device layout (one context = one structure or a batch of structures, atoms concatenated):
 *   x        double[3N]   wrapped Cartesian positions (FP64: the neighbour test must round like the reference)
 *   spc      int[N]       species index; atom_str int[N] structure id
 *   row_off  int[N+1]     CSR rows of the neighbour list; each row sorted by (species, j, image)
 *   nbr_j    int[M]       neighbour atom; nbr_img int[M] packed image (10 bits per axis, offset 512)
 *   nbr_vec  float4[M]    (dx, dy, dz, r) = X_j - X_i + image shift, in FP32
 *   rev      int[M]       index of the reverse edge (j -> i, -image)
 *   G, dEdG  float[N][D]  descriptors and dE/dG
 *   fself    float4[N], fe float4[M]   force on the centre atom and on each neighbour slot
 *   vir      float[N][6]  per-atom virial partials; E/U per structure reduced in FP64
 */
#ifndef MG_INTERNAL_H
#define MG_INTERNAL_H

#include <stdint.h>

#define MG_MAXSPC 4
#define MG_NPAIR (MG_MAXSPC * MG_MAXSPC)
#define MG_NCLS (MG_MAXSPC * (MG_MAXSPC + 1) / 2) /* unordered species classes (sa <= sb) of G4 pairs */
#define MG_MAXG2 16
#define MG_MAXE 8        /* distinct G4 eta values */
#define MG_MAXZ 4        /* distinct zeta values */
#define MG_MAXL 2        /* distinct lambda values */
#define MG_GRID (MG_MAXE * MG_MAXZ * MG_MAXL) /* 64: G4 functions live on an (eta, zeta, lambda) grid */
#define MG_MAXG4 MG_GRID
#define MG_IMG_OFF 512
#define MG_LOG2E 1.4426950408889634
#define MG_RCHUNK 512               /* atoms per FP64 reduction chunk */

#ifdef __cplusplus
extern "C" {
#endif

/* status of the launches just issued: 0, or -1 with the HIP error kept for mgk_last_error() (the error the
 * runtime reports once is consumed, so messages built afterwards would otherwise read "no error") */
int mgk_status(void);
const char *mgk_last_error(void);

/* Basis tables, in FP32, indexed by the ordered species pair p = s0*MG_MAXSPC + s1 (symmetric). */
typedef struct {
  int nspc, D, ng2, ng4;
  int ne, nz, nl;              /* counts of eta4, zeta, lambda values */
  int ind2[MG_MAXSPC];         /* G2 block offset by ds1 (src/nprs.c:484-498) */
  int ind4[MG_MAXSPC][MG_MAXSPC]; /* G4 block offset by (ds1, ds2) */
  float rc[MG_NPAIR];          /* cutoff per pair (G2 and G4 cutoffs are equal; checked on load) */
  float eta2[MG_NPAIR][MG_MAXG2]; /* G2 eta (A^-2) */
  float rs2[MG_NPAIR][MG_MAXG2];  /* G2 Rs (A) */
  float eta4[MG_NPAIR][MG_MAXE];  /* G4 eta (A^-2) */
  int zeta[MG_MAXZ];           /* integer zeta values */
  float lam[MG_MAXL];
  float pref[MG_MAXZ];         /* 2^(1 - zeta) */
  int g4map[MG_GRID];          /* grid cell (e*MAXZ + z)*MAXL + l -> G4 function index in the block, or -1 */
  int g4cell[MG_MAXG4];        /* G4 function index -> grid cell */
  float urep_rm[MG_MAXSPC];    /* UREP radius mindist[Z] per species (src/util.c:13) */
} mg_basis_dev;

typedef struct {
  double L[9];                 /* lattice rows (A) */
  double R[9];                 /* reciprocal rows, R = (L1xL2, L2xL0, L0xL1)/det (src/cell.c:432-444) */
  double rc2;                  /* neighbour-list cutoff squared */
  int atom0, natom, nd, xt;
  int nb[3];                   /* bins per lattice direction */
  int K[3];                    /* stencil half width per direction */
  int Ks[3];                   /* the same for the skin-candidate radius (rc + skin) */
  int bin0, nbins;
  int chunk0, nchunk;           /* reduction chunks (MG_RCHUNK atoms each) of this structure */
} mg_struct_dev;

#ifdef __cplusplus
}
#endif
#endif
