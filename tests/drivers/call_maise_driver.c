/* A caller written against include/mlib.h exactly as the reference's interface/c/c2maise.c is written
 * against its mlib.h: caller-owned structs, C.p and L.B set, one CALL_MAISE per input structure.
 * usage: call_maise_driver <input> [repeats]; the model is ./model (or $MAISE_MODEL).
 * input: "N ND XT NM P_GPa" / lattice (9) / N lines "species x y z"; prints H, F, STR with %.17g. */
#include <stdio.h>
#include <stdlib.h>
#include "mlib.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  int repeats = argc > 2 ? atoi(argv[2]) : 1;
  FILE *in = fopen(argv[1], "r");
  if (!in) return 2;
  int N, ND, XT, NM;
  double p;
  if (fscanf(in, "%d %d %d %d %lf", &N, &ND, &XT, &NM, &p) != 5) return 2;
  double LAT[9];
  for (int k = 0; k < 9; k++) if (fscanf(in, "%lf", &LAT[k]) != 1) return 2;
  int *ATMN = malloc(sizeof(int) * N);
  double *X = malloc(sizeof(double) * 3 * N), *FRC = malloc(sizeof(double) * 3 * N);
  for (int i = 0; i < N; i++)
    if (fscanf(in, "%d %lf %lf %lf", &ATMN[i], &X[3 * i], &X[3 * i + 1], &X[3 * i + 2]) != 4) return 2;
  fclose(in);
  ANN R; PRS P; PRS W[9]; LNK L; Cell C;
  double STR[6], H = 0.0;
  L.B = 0;
  for (int r = 0; r < repeats; r++) {
    C.p = p;
    H = CALL_MAISE(&R, &P, W, &L, &C, 0, N, NM, ND, 4, XT, ATMN, LAT, X, FRC, STR);
  }
  printf("H %.17g\n", H);
  for (int i = 0; i < N; i++) printf("F %.17g %.17g %.17g\n", FRC[3 * i], FRC[3 * i + 1], FRC[3 * i + 2]);
  for (int q = 0; q < 6; q++) printf("S %.17g\n", STR[q]);
  maise_gpu_free(&L);
  return 0;
}
