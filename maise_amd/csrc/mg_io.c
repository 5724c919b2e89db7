/*
 * mg_io.c -- the structure and output files of the reference's workflows, host code only:
 *   POSCAR / CONTCAR read ........ READ_CELL, src/util.c:820-953
 *   CONTCAR write ................ SAVE_CELL, src/util.c:955-1030 (text format kept: CONTCARs written here and by
 *                                  the reference are interchangeable and diff clean)
 *   OUTCAR / OSZICAR blocks ...... CELL_OUT, src/cmod.c:618-687, with the cell volume of CELL_VOL
 *                                  (src/cutl.c:178-188) and, for clusters, NP_VOL (src/nutl.c:846-903)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mg_host.h"

#define MG_NSYMB 101

static const char *const MG_SYMB[MG_NSYMB] = { /* element symbols by Z (atom_symb, src/nutl.c:741) */
  "--", "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl",
  "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
  "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",
  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu",
  "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac",
  "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm"};

/* atomic masses (amu) by Z, the same table as the dynamics (src/util.c:11) */
extern const double MG_AMU_TABLE[96];

/* atomic number of a symbol written as in the table or all in lower case ("Cu", "cu"); 0 if unknown
 * (symb_atom, src/nutl.c:635) */
static int symbol_z(const char *s) {
  for (int z = 1; z < MG_NSYMB; z++) {
    const char *e = MG_SYMB[z];
    if (strcmp(s, e) == 0) return z;
    if (s[0] == (char)(e[0] - 'A' + 'a') && strcmp(s + 1, e + 1) == 0) return z;
  }
  return 0;
}

const char *mg_element_symbol(int z) { return (z > 0 && z < MG_NSYMB) ? MG_SYMB[z] : MG_SYMB[0]; }
int mg_element_z(const char *symbol) { return symbol ? symbol_z(symbol) : 0; }

static double atom_mass(int z) { return (z > 0 && z < 96) ? MG_AMU_TABLE[z] : 0.0; }

/* shift fractional coordinates so that their mean is o along each lattice direction (CENTER, src/cutl.c:38-52) */
static void center_frac(int N, double *f, double o) {
  for (int q = 0; q < 3; q++) {
    double r = 0.0;
    for (int i = 0; i < N; i++) r += f[3 * i + q] / (double)N;
    for (int i = 0; i < N; i++) f[3 * i + q] += -r + o;
  }
}

/* fractional <- Cartesian with the reciprocal rows (Relative, src/cell.c:446-466) */
static void to_frac(const double *R, const double *x, double *f) {
  for (int q = 0; q < 3; q++) {
    double t = 0.0;
    for (int j = 0; j < 3; j++) t += x[j] * R[3 * q + j];
    f[q] = t;
  }
}

/* Cartesian <- fractional with the lattice rows (Real, src/cell.c:468-487) */
static void to_cart(const double *L, const double *f, double *x) {
  for (int q = 0; q < 3; q++) {
    double t = 0.0;
    for (int j = 0; j < 3; j++) t += f[j] * L[3 * j + q];
    x[q] = t;
  }
}

int mg_poscar_read(const char *path, int nd, int cap, char *title, double *lat, int *nspc, int *spcz, int *spcn,
                   int *natoms, double *x, int *ff, double *v, int *has_v) {
  if (!path || !lat || !nspc || !spcz || !spcn || !natoms) return mg_fail(MG_ERR_ARG, "mg_poscar_read: bad argument");
  FILE *in = fopen(path, "r");
  if (!in) return mg_fail(MG_ERR_IO, "cannot open %s", path);
  char buf[256], s[128], t[10][64];
  int rc = MG_OK;
  double R0 = 1.0, L[9];
  if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: empty file", path); }
  buf[strcspn(buf, "\n")] = '\0';
  if (title) { strncpy(title, buf, 199); title[199] = '\0'; }
  if (fscanf(in, "%lf", &R0) != 1) { fclose(in); return mg_fail(MG_ERR_IO, "%s: no scale factor", path); }
  for (int k = 0; k < 9; k++)
    if (fscanf(in, "%lf", &L[k]) != 1) { fclose(in); return mg_fail(MG_ERR_IO, "%s: incomplete lattice", path); }
  if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: truncated", path); }  /* rest of line 5 */
  /* species symbols (required, as in the reference), then the counts */
  if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: truncated", path); }
  int ns_sym = 0;
  if (sscanf(buf, "%127s", s) == 1 && !(s[0] >= '0' && s[0] <= '9')) {
    ns_sym = sscanf(buf, "%63s %63s %63s %63s %63s %63s %63s %63s %63s %63s", t[0], t[1], t[2], t[3], t[4], t[5], t[6],
                    t[7], t[8], t[9]);
    if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: truncated", path); }
  }
  int TN[10];
  const int ns = sscanf(buf, "%d %d %d %d %d %d %d %d %d %d", &TN[0], &TN[1], &TN[2], &TN[3], &TN[4], &TN[5], &TN[6],
                        &TN[7], &TN[8], &TN[9]);
  if (ns < 1) { fclose(in); return mg_fail(MG_ERR_IO, "%s: no atom counts", path); }
  if (ns_sym != ns) { fclose(in); return mg_fail(MG_ERR_ARG, "%s: the species symbols are missing", path); }
  int N = 0;
  for (int k = 0; k < ns; k++) {
    if (TN[k] < 0 || TN[k] > 100000000 - N) { fclose(in); return mg_fail(MG_ERR_IO, "%s: bad atom counts", path); }
    spcn[k] = TN[k];
    spcz[k] = symbol_z(t[k]);
    if (spcz[k] == 0) { fclose(in); return mg_fail(MG_ERR_ARG, "%s: unknown species %s", path, t[k]); }
    N += TN[k];
  }
  *nspc = ns;
  *natoms = N;
  for (int k = 0; k < 9; k++) lat[k] = L[k] * R0;
  if (cap < N || !x) { fclose(in); return MG_OK; }  /* header and atom count only */
  if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: truncated", path); }
  if (buf[0] == 'S' || buf[0] == 's')
    if (!fgets(buf, sizeof buf, in)) { fclose(in); return mg_fail(MG_ERR_IO, "%s: truncated", path); }
  const int direct = buf[strspn(buf, " \t")] == 'd' || buf[strspn(buf, " \t")] == 'D';
  for (int i = 0; i < N; i++) {
    char f3[3][8] = {"T", "T", "T"};
    double p[3];
    if (!fgets(buf, sizeof buf, in) ||
        sscanf(buf, "%lf %lf %lf %7s %7s %7s", &p[0], &p[1], &p[2], f3[0], f3[1], f3[2]) < 3) {
      rc = mg_fail(MG_ERR_IO, "%s: position %d missing", path, i);
      break;
    }
    if (direct) to_cart(lat, p, x + 3 * i);
    else for (int q = 0; q < 3; q++) x[3 * i + q] = p[q];
    if (ff)
      for (int q = 0; q < 3; q++) ff[3 * i + q] = strcmp(f3[q], "F") == 0 ? 0 : 1;
  }
  /* velocities, when three numbers follow for every atom */
  int got = 0, i = 0;
  double vv[3];
  while (rc == MG_OK && i < N && fgets(buf, sizeof buf, in)) {
    const int k = sscanf(buf, "%lf %lf %lf", &vv[0], &vv[1], &vv[2]);
    if (k <= 0 && got == 0) continue;  /* the blank separator line */
    if (k != 3) break;
    if (v) for (int q = 0; q < 3; q++) v[3 * i + q] = vv[q];
    got += 3;
    i++;
  }
  if (has_v) *has_v = (rc == MG_OK && got == 3 * N);
  fclose(in);
  if (rc == MG_OK && nd == 0 && N > 0) { /* clusters come back centred on the origin (CENTER(C, 0.0)) */
    double R[9], f[3];
    mgh_reciprocal(lat, R);
    double *fr = (double *)malloc(sizeof(double) * 3 * N);
    if (!fr) return mg_fail(MG_ERR_ALLOC, "mg_poscar_read: out of memory");
    for (int a = 0; a < N; a++) to_frac(R, x + 3 * a, fr + 3 * a);
    center_frac(N, fr, 0.0);
    for (int a = 0; a < N; a++) {
      to_cart(lat, fr + 3 * a, f);
      for (int q = 0; q < 3; q++) x[3 * a + q] = f[q];
    }
    free(fr);
  }
  return rc;
}

int mg_poscar_write(const char *path, const char *title, double scale, const double *lat, int nd, int nspc,
                    const int *spcz, const int *spcn, int natoms, const double *x, const int *ff, int sd,
                    const double *v) {
  if (!path || !lat || !spcz || !spcn || !x || nspc < 1 || nspc > 10 || scale == 0.0)
    return mg_fail(MG_ERR_ARG, "mg_poscar_write: bad argument");
  int N = 0;
  for (int k = 0; k < nspc; k++) N += spcn[k];
  if (N != natoms) return mg_fail(MG_ERR_ARG, "mg_poscar_write: species counts do not add up to %d", natoms);
  double R[9];
  mgh_reciprocal(lat, R);
  double *f = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
  if (!f) return mg_fail(MG_ERR_ALLOC, "mg_poscar_write: out of memory");
  for (int i = 0; i < N; i++) to_frac(R, x + 3 * i, f + 3 * i);
  if (nd == 0 && N > 0) center_frac(N, f, 0.5); /* clusters are written centred in the box (CENTER(C, 0.5)) */
  FILE *out = fopen(path, "w");
  if (!out) { free(f); return mg_fail(MG_ERR_IO, "cannot write %s", path); }
  fprintf(out, "%s\n", title ? title : "");
  fprintf(out, "% 17.14lf\n", scale);
  for (int i = 0; i < 3; i++) {
    for (int q = 0; q < 3; q++) fprintf(out, "% 19.14lf ", lat[3 * i + q] / scale);
    fprintf(out, "\n");
  }
  if (spcz[0] > 0) {
    for (int k = 0; k < nspc; k++)
      if (spcn[k] > 0) fprintf(out, "%s ", mg_element_symbol(spcz[k]));
    fprintf(out, "\n");
  }
  for (int k = 0; k < nspc; k++)
    if (spcn[k] > 0) fprintf(out, "%d ", spcn[k]);
  /* the header line follows the caller's sd only; fixed coordinates in ff add the T/F columns without it, as the
   * reference writes them (its reader takes the columns either way) */
  if (sd) fprintf(out, "\nSelective dynamics");
  fprintf(out, "\ndirect\n");
  if (!sd && ff)
    for (int i = 0; i < 3 * N; i++)
      if (ff[i] == 0) sd = 1;
  for (int i = 0; i < N; i++) {
    for (int q = 0; q < 3; q++) fprintf(out, "% 19.14lf ", f[3 * i + q]);
    if (sd)
      for (int q = 0; q < 3; q++) fprintf(out, (ff && ff[3 * i + q] == 0) ? "  F" : "  T");
    fprintf(out, "\n");
  }
  if (v) {
    fprintf(out, "\n");
    for (int i = 0; i < N; i++) fprintf(out, "% 19.14lf % 19.14lf % 19.14lf\n", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  }
  const int bad = ferror(out);
  fclose(out);
  free(f);
  return bad ? mg_fail(MG_ERR_IO, "write error on %s", path) : MG_OK;
}

/* cluster volume from the inertia tensor (NP_VOL): (4 pi / 3) (2.5 / M)^1.5 sqrt(w0 w1 w2), where the product of the
 * eigenvalues w is the tensor's determinant */
static double cluster_volume(int N, const int *z, const double *x) {
  double M = 0.0, a = 0.0, b = 0.0, c = 0.0;
  for (int i = 0; i < N; i++) M += atom_mass(z[i]);
  for (int i = 0; i < N; i++) {
    const double m = atom_mass(z[i]);
    a += m * x[3 * i]; b += m * x[3 * i + 1]; c += m * x[3 * i + 2];
  }
  double I[3][3];
  I[0][0] = -(b * b / M) - (c * c / M);
  I[1][1] = -(a * a / M) - (c * c / M);
  I[2][2] = -(b * b / M) - (a * a / M);
  I[0][1] = a * b / M; I[0][2] = a * c / M; I[1][2] = b * c / M;
  for (int i = 0; i < N; i++) {
    const double m = atom_mass(z[i]), x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    I[0][0] += m * x1 * x1 + m * x2 * x2;
    I[1][1] += m * x0 * x0 + m * x2 * x2;
    I[2][2] += m * x1 * x1 + m * x0 * x0;
    I[0][1] -= m * x0 * x1;
    I[0][2] -= m * x0 * x2;
    I[1][2] -= m * x1 * x2;
  }
  I[1][0] = I[0][1]; I[2][0] = I[0][2]; I[2][1] = I[1][2];
  const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) - I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                     I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
  return (4.0 * M_PI / 3.0) * pow(2.5 / M, 1.5) * pow(det, 0.5);
}

static double lattice_volume(const double *L) { /* CELL_VOL, src/cutl.c:183-187 */
  const double V = L[0] * L[4] * L[8] + L[2] * L[3] * L[7] + L[1] * L[5] * L[6] - L[0] * L[5] * L[7] -
                   L[1] * L[3] * L[8] - L[2] * L[4] * L[6];
  return fabs(V);
}

int mg_outcar_write(const char *dir, int it, const char *version, const char *model_id, int mint, int rlxt,
                    double p_gpa, int nd, int nspc, const int *spcz, const int *spcn, int natoms, const double *lat,
                    const double *x, const double *F, const double *EA, const double *U, double E, double H,
                    int detail) {
  const double ev2gpa = 160.2176565;
  if (!dir || !lat || !spcz || !spcn || nspc < 1 || nspc > 10 || natoms < 1 || (detail && (!x || !F || !EA || !U)))
    return mg_fail(MG_ERR_ARG, "mg_outcar_write: bad argument");
  char path[4096], t[256];
  const char *b = "----------------------------------------------------------------------------------------------------------\n";
  snprintf(path, sizeof path, "%s/OUTCAR", dir);
  FILE *out = fopen(path, "a");
  if (!out) return mg_fail(MG_ERR_IO, "cannot append to %s", path);
  if (it == 0) {
    time_t raw = time(NULL);
    snprintf(t, sizeof t, " %s", asctime(localtime(&raw)));
    t[strcspn(t, "\n")] = '\0';
    fprintf(out, "%s    Cell optimization    with %s   model %s   on %s                     \n%s", b,
            version ? version : "", model_id ? model_id : "", t, b);
    snprintf(t, sizeof t, " %d", mint);
    fprintf(out, "   optimizer type         %s\n", t);
    snprintf(t, sizeof t, " %d", rlxt);
    fprintf(out, "   relaxation type        %s\n", t);
    snprintf(t, sizeof t, " %1.2lf GPa", p_gpa);
    fprintf(out, "   target pressure        %s\n", t);
    snprintf(t, sizeof t, " %d", nd);
    fprintf(out, "   dimensionality         %s\n", t);
    snprintf(t, sizeof t, " ");
    for (int i = 0; i < nspc; i++) snprintf(t + strlen(t), sizeof t - strlen(t), "%s    ", mg_element_symbol(spcz[i]));
    fprintf(out, "   species                %s\n", t);
    /* the counts go in columns of 6 from column 27 of a blank-padded line */
    snprintf(t, sizeof t, "   atoms of each species                     ");
    for (int i = 0; i < nspc; i++) snprintf(t + 27 + 6 * i, sizeof t - 27 - 6 * i, "%d           ", spcn[i]);
    fprintf(out, "%s\n", t);
    snprintf(t, sizeof t, " %d", natoms);
    fprintf(out, "   total number of atoms  %s\n", t);
    fprintf(out, "%s\n\n\n", b);
  }
  int *z = (int *)malloc(sizeof(int) * natoms);
  if (!z) { fclose(out); return mg_fail(MG_ERR_ALLOC, "mg_outcar_write: out of memory"); }
  for (int k = 0, i = 0; k < nspc; k++)
    for (int q = 0; q < spcn[k]; q++) z[i++] = spcz[k];
  const double V = (nd == 0 && x) ? cluster_volume(natoms, z, x) : lattice_volume(lat);
  free(z);
  fprintf(out, "%s                   LATTICE VECTORS                                                 VOLUME (Angst^3/atom) \n%s", b, b);
  fprintf(out, "      % 12.6lf  % 12.6lf  % 12.6lf  % 45.6lf vol\n", lat[0], lat[1], lat[2], V / (double)natoms);
  fprintf(out, "      % 12.6lf  % 12.6lf  % 12.6lf\n", lat[3], lat[4], lat[5]);
  fprintf(out, "      % 12.6lf  % 12.6lf  % 12.6lf\n", lat[6], lat[7], lat[8]);
  if (detail) {
    fprintf(out, "%s                 POSITION (Angst)                     TOTAL-FORCE (eV/Angst)           ATOM ENERGY (eV)\n%s", b, b);
    for (int i = 0; i < natoms; i++)
      fprintf(out, "   % 12.6lf % 12.6lf % 12.6lf   % 12.6lf % 12.6lf % 12.6lf % 17.12lf\n", x[3 * i], x[3 * i + 1],
              x[3 * i + 2], F[3 * i], F[3 * i + 1], F[3 * i + 2], EA[i]);
    fprintf(out, "%s", b);
    fprintf(out, "  Total   ");
    for (int i = 0; i < 6; i++) fprintf(out, "% 15.8lf ", U[i] / V);
    fprintf(out, "\n");
    fprintf(out, "  in kB   ");
    for (int i = 0; i < 6; i++) fprintf(out, "% 15.8lf ", U[i] * ev2gpa * 10.0 / V);
    fprintf(out, "\n%s", b);
  }
  fprintf(out, "  iter %3d   total enthalpy= % 14.8lf   energy=  % 14.8lf   % 14.8lf  % 14.8lf\n%s\n\n\n", it, H, E,
          H / (double)natoms, E / (double)natoms, b);
  int bad = ferror(out);
  fclose(out);
  snprintf(path, sizeof path, "%s/OSZICAR", dir);
  out = fopen(path, "a");
  if (!out) return mg_fail(MG_ERR_IO, "cannot append to %s", path);
  fprintf(out, "                            % 1.8lf\n", H);
  bad |= ferror(out);
  fclose(out);
  return bad ? mg_fail(MG_ERR_IO, "write error on OUTCAR/OSZICAR in %s", dir) : MG_OK;
}
