/*
 * ref_driver.c -- a caller of the reference MAISE library's C entry point, used to produce golden
 * vectors and the reference CPU timing.  TEST INFRASTRUCTURE ONLY.
 *
 * It is linked against objects compiled from the reference's own sources (see oracle/Makefile, which
 * builds into oracle/_ref/); like interface/c/c2maise.c it declares the caller-owned structs and calls
 * CALL_MAISE (inc/cmod.h:30).  The two fields the library path leaves uninitialised (R.UREP, P.IO) are
 * set explicitly.
 *
 * usage: maise_ref <input> <output> [repeats] [threads]
 *   The model is read from ./model (the reference always reads the current directory).
 *   input : line 1 "N ND XT NM UREP P_GPa"; line 2 nine lattice numbers (rows); then N lines "species x y z".
 *   output: "H", "E", "EA", "F", "U", "NN" (neighbour counts), "NBR" (j ic0 ic1 ic2 per entry, in the
 *           reference's list order), "G" (descriptors), "X" (wrapped positions) blocks, %.17g.
 *   repeats > 1 times that many further calls after the first (model loading excluded) and prints
 *           "TIME <seconds per call>" on stdout.
 */
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>
#include "mlib.h"

static double now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s input output [repeats] [threads]\n", argv[0]);
    return 2;
  }
  int repeats = argc > 3 ? atoi(argv[3]) : 1;
  int threads = argc > 4 ? atoi(argv[4]) : 1;
  FILE *in = fopen(argv[1], "r");
  if (!in) { fprintf(stderr, "cannot open %s\n", argv[1]); return 2; }
  int N, ND, XT, NM;
  double urep, pgpa;
  if (fscanf(in, "%d %d %d %d %lf %lf", &N, &ND, &XT, &NM, &urep, &pgpa) != 6) return 2;
  double LAT[9];
  for (int i = 0; i < 9; i++) if (fscanf(in, "%lf", &LAT[i]) != 1) return 2;
  int *ATMN = malloc(sizeof(int) * N);
  double *X = malloc(sizeof(double) * 3 * N), *F = malloc(sizeof(double) * 3 * N);
  for (int i = 0; i < N; i++)
    if (fscanf(in, "%d %lf %lf %lf", &ATMN[i], &X[3 * i], &X[3 * i + 1], &X[3 * i + 2]) != 4) return 2;
  fclose(in);

  static ANN R;
  static PRS P;
  static PRS W[9];
  static LNK L;
  static Cell C;
  double STR[6];
  R.UREP = urep;
  P.IO = 0;
  L.B = 0;
  C.p = pgpa;
  double H = CALL_MAISE(&R, &P, W, &L, &C, 0, N, NM, ND, threads, XT, ATMN, LAT, X, F, STR);
  if (repeats > 1) {
    double t0 = now();
    for (int r = 0; r < repeats; r++) {
      C.p = pgpa;
      H = CALL_MAISE(&R, &P, W, &L, &C, 0, N, NM, ND, threads, XT, ATMN, LAT, X, F, STR);
    }
    printf("TIME %.9f\n", (now() - t0) / repeats);
  }
  FILE *out = fopen(argv[2], "w");
  fprintf(out, "H %.17g\nE %.17g\n", H, C.E);
  fprintf(out, "EA\n");
  for (int i = 0; i < N; i++) fprintf(out, "%.17g\n", C.EA[i]);
  fprintf(out, "F\n");
  for (int i = 0; i < N; i++) fprintf(out, "%.17g %.17g %.17g\n", F[3 * i], F[3 * i + 1], F[3 * i + 2]);
  fprintf(out, "U\n");
  for (int q = 0; q < 6; q++) fprintf(out, "%.17g\n", STR[q]);
  fprintf(out, "NN\n");
  for (int i = 0; i < N; i++) fprintf(out, "%d\n", C.Nn[i]);
  fprintf(out, "NBR\n");
  for (int i = 0; i < N; i++)
    for (int k = 0; k < C.Nn[i]; k++) {
      int ic[3];
      for (int q = 0; q < 3; q++) {
        /* recover integer image indices from the stored shift S = sum_q1 ic_q1 L_q1 */
        ic[q] = 0;
      }
      {
        /* solve S = ic . L using the reciprocal rows R (C.R holds them after LIST) */
        for (int q = 0; q < 3; q++) {
          double s = 0.0;
          for (int k2 = 0; k2 < 3; k2++) s += C.S[i][k][k2] * C.R[q][k2];
          ic[q] = (int)(s >= 0 ? s + 0.5 : s - 0.5);
        }
      }
      fprintf(out, "%d %d %d %d\n", C.Ni[i][k], ic[0], ic[1], ic[2]);
    }
  fprintf(out, "G\n");
  for (int i = 0; i < N; i++) {
    for (int n = 0; n < P.D; n++) fprintf(out, "%.17g ", L.Cn[i][n]);
    fprintf(out, "\n");
  }
  fprintf(out, "X\n");
  for (int i = 0; i < N; i++) fprintf(out, "%.17g %.17g %.17g\n", C.X[i][0], C.X[i][1], C.X[i][2]);
  fclose(out);
  return 0;
}
