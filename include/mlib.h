/*
 * mlib.h -- drop-in header for programs written against MAISE's library interface.
 *
 * The reference's interface drivers (interface/c/c2maise.c, interface/cpp/cpp2maise.cpp) include "mlib.h",
 * declare `ANN R; PRS P; PRS W[9]; LNK L; Cell C;` on the stack, set `C.p` (GPa) and `L.B = 0`, and call
 * CALL_MAISE.  This header gives those four types the reference's exact sizes and alignment (x86-64:
 * ANN 4704 B, PRS 160 B, LNK 376 B, Cell 10624 B) with the two caller-visible fields at the reference's
 * offsets (LNK.B at 152, Cell.p at 1736), so the drivers compile and link unchanged against
 * libmaise_gpu.  The rest of each struct is storage owned by the library (it keeps its GPU context in LNK).
 *
 * Entry point replaced: CALL_MAISE, inc/cmod.h:30 / src/cmod.c:1164-1224.  Semantics kept:
 *   - first call (L.B == 0) loads ./model (env MAISE_MODEL overrides the path) and sets L.B = 1;
 *   - C.p is divided by eV2GPa = 160.2176565 on EVERY call, as the reference does (src/cmod.c:1181):
 *     a caller that does not reset C.p between calls sees its pressure shrink;
 *   - returns H = E (+ p V for ND = 3) in eV; FRC[3N] in eV/A; STR[6] = U in eV (xx yy zz xy yz zx),
 *     components with |x| < 1e-8 zeroed (src/cmod.c:177-190);
 *   - UREP (left uninitialised by the reference's library path) is 100 eV, the executable default
 *     (src/util.c:1431); env MAISE_UREP overrides it.  No files are ever written (P->IO = 0).
 *   - errors (missing model, neighbour overflow, device failure) print a message and exit(1), as the
 *     reference does; CODE and NP are accepted and ignored.
 */
#ifndef MAISE_GPU_MLIB_H
#define MAISE_GPU_MLIB_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "maise_gpu.h"

#ifdef __cplusplus
extern "C" {
#define MG_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define MG_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

typedef struct { double storage_[588]; } ANN;                         /* 4704 B */
typedef struct { double storage_[20]; } PRS;                          /* 160 B */
typedef struct { void *storage_[19]; int B; int pad_; void *tail_[27]; } LNK;  /* 376 B, B at 152 */
typedef struct { double storage_[217]; double p; double tail_[1110]; } Cell;   /* 10624 B, p at 1736 */

MG_STATIC_ASSERT(sizeof(ANN) == 4704, "ANN size");
MG_STATIC_ASSERT(sizeof(PRS) == 160, "PRS size");
MG_STATIC_ASSERT(sizeof(LNK) == 376, "LNK size");
MG_STATIC_ASSERT(sizeof(Cell) == 10624, "Cell size");

double CALL_MAISE(ANN *R, PRS *P, PRS *W, LNK *L, Cell *C, int CODE, int N, int NM, int ND, int NP, int XT,
                  int *ATMN, double *LAT, double *X, double *FRC, double *STR);

/* Release the GPU context CALL_MAISE keeps inside L (the reference never frees; optional here). */
void maise_gpu_free(LNK *L);

#ifdef __cplusplus
}
#endif
#endif
