/*
 * mg_evo.c -- host C operators of the evolutionary search: cut-and-splice crossover (after BULK_MATE,
 * src/efnc.c:2021-2310, with the lattice choice of MATE_LT, src/eutl.c:139-172), Gaussian mutation and the
 * periodic minimum-distance check.  Plain C on the host: the search spends its GPU time in the batched
 * relaxations, and these operators must not stall it.
 *
 * Structures are fixed-size (N atoms, species grouped); lattices are rows (A); positions Cartesian (A).
 * Random numbers come from splitmix64 streams, so a seed fixes every offspring.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mg_host.h"

static unsigned long long sm64(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double urand(unsigned long long *s) { return (sm64(s) >> 11) * (1.0 / 9007199254740992.0); }
static double nrand(unsigned long long *s) {
  double u = urand(s), v = urand(s);
  if (u < 1e-300) u = 1e-300;
  return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

static double det3(const double *L) {
  return L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) + L[2] * (L[3] * L[7] - L[4] * L[6]);
}

/* inverse of the row-lattice: frac = x * inv */
static void inv3(const double *L, double *I) {
  const double d = det3(L);
  I[0] = (L[4] * L[8] - L[5] * L[7]) / d; I[1] = (L[2] * L[7] - L[1] * L[8]) / d; I[2] = (L[1] * L[5] - L[2] * L[4]) / d;
  I[3] = (L[5] * L[6] - L[3] * L[8]) / d; I[4] = (L[0] * L[8] - L[2] * L[6]) / d; I[5] = (L[2] * L[3] - L[0] * L[5]) / d;
  I[6] = (L[3] * L[7] - L[4] * L[6]) / d; I[7] = (L[1] * L[6] - L[0] * L[7]) / d; I[8] = (L[0] * L[4] - L[1] * L[3]) / d;
}

static void to_frac(const double *L, const double *x, int n, double *f) {
  double I[9];
  inv3(L, I);
  for (int i = 0; i < n; i++)
    for (int q = 0; q < 3; q++)
      f[3 * i + q] = x[3 * i] * I[q] + x[3 * i + 1] * I[3 + q] + x[3 * i + 2] * I[6 + q];
}

static void params(const double *L, double *len, double *ang) {
  for (int k = 0; k < 3; k++) len[k] = sqrt(L[3 * k] * L[3 * k] + L[3 * k + 1] * L[3 * k + 1] + L[3 * k + 2] * L[3 * k + 2]);
#define DOT(a, b) (L[3 * (a)] * L[3 * (b)] + L[3 * (a) + 1] * L[3 * (b) + 1] + L[3 * (a) + 2] * L[3 * (b) + 2])
  ang[0] = acos(DOT(1, 2) / (len[1] * len[2]));
  ang[1] = acos(DOT(0, 2) / (len[0] * len[2]));
  ang[2] = acos(DOT(0, 1) / (len[0] * len[1]));
#undef DOT
}

static int from_params(const double *len, const double *ang, double *L) {
  const double ca = cos(ang[0]), cb = cos(ang[1]), cg = cos(ang[2]), sg = sin(ang[2]);
  const double cx = len[2] * cb, cy = len[2] * (ca - cb * cg) / sg;
  const double cz2 = len[2] * len[2] - cx * cx - cy * cy;
  if (cz2 <= 1e-6) return 0;
  L[0] = len[0]; L[1] = 0.0; L[2] = 0.0;
  L[3] = len[1] * cg; L[4] = len[1] * sg; L[5] = 0.0;
  L[6] = cx; L[7] = cy; L[8] = sqrt(cz2);
  return 1;
}

/* smallest periodic distance between atoms of fractional positions f (27 images of the rounded difference) */
double mg_evo_min_distance(const double *L, const double *x, int n) {
  double *f = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  to_frac(L, x, n, f);
  double best = INFINITY;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double d[3];
      for (int q = 0; q < 3; q++) { d[q] = f[3 * j + q] - f[3 * i + q]; d[q] -= nearbyint(d[q]); }
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int c = -1; c <= 1; c++) {
            const double u = d[0] + a, v = d[1] + b, w = d[2] + c;
            const double X = u * L[0] + v * L[3] + w * L[6], Y = u * L[1] + v * L[4] + w * L[7], Z = u * L[2] + v * L[5] + w * L[8];
            const double r2 = X * X + Y * Y + Z * Z;
            if (r2 < best) best = r2;
          }
    }
  free(f);
  return sqrt(best);
}

/* Soft repulsion that pushes too-close atoms apart, after ADJT_CL (src/eutl.c:97-135): at most `iters` steps;
 * before each, give up (return 0) when a pair is closer than 0.7 dmin; a step moves every atom by 0.2 F_i with
 * F_i = sum_j (x_i - x_j)/R (r_ij/R)^-13 over the periodic images within 2R (R = dmin, the pair's hard-core
 * distance), |F_i| capped at 2.  Returns 1 when every distance ends >= dmin.  f and F hold 3N doubles. */
static int push_apart(const double *L, double *x, int N, double dmin, int iters, double *f, double *F) {
  const double R = dmin, rc2 = 4.0 * R * R;
  for (int it = 0; it < iters; it++) {
    if (mg_evo_min_distance(L, x, N) < 0.7 * dmin) return 0;
    to_frac(L, x, N, f);
    memset(F, 0, sizeof(double) * 3 * N);
    for (int i = 0; i < N; i++)
      for (int j = i + 1; j < N; j++) {
        double d[3];
        for (int q = 0; q < 3; q++) { d[q] = f[3 * j + q] - f[3 * i + q]; d[q] -= nearbyint(d[q]); }
        for (int a = -1; a <= 1; a++)
          for (int b = -1; b <= 1; b++)
            for (int c = -1; c <= 1; c++) {
              const double u = d[0] + a, v = d[1] + b, w = d[2] + c;
              const double X = u * L[0] + v * L[3] + w * L[6], Y = u * L[1] + v * L[4] + w * L[7],
                           Z = u * L[2] + v * L[5] + w * L[8];
              const double r2 = X * X + Y * Y + Z * Z;
              if (r2 >= rc2 || r2 == 0.0) continue;
              const double g = pow(sqrt(r2) / R, -13.0) / R;   /* x_j - x_i = (X, Y, Z) */
              F[3 * i] -= g * X; F[3 * i + 1] -= g * Y; F[3 * i + 2] -= g * Z;
              F[3 * j] += g * X; F[3 * j + 1] += g * Y; F[3 * j + 2] += g * Z;
            }
      }
    for (int i = 0; i < N; i++) {
      const double n = sqrt(F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] + F[3 * i + 2] * F[3 * i + 2]);
      const double sc = n > 2.0 ? 2.0 / n : 1.0;
      for (int q = 0; q < 3; q++) x[3 * i + q] += 0.2 * sc * F[3 * i + q];
    }
  }
  return mg_evo_min_distance(L, x, N) >= dmin;
}

/* random vector in the unit ball (RAND_VC, src/eutl.c:222-234) */
static void ball(unsigned long long *s, double *a) {
  do {
    for (int q = 0; q < 3; q++) a[q] = 2.0 * (0.5 - urand(s));
  } while (sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) > 1.0);
}

/* one crossover; returns 1 and fills (spc, L, x) of the child, or 0 when a test fails */
static int mate(int N, int nspc, const int *s1, const double *L1, const double *x1, const int *s2, const double *L2,
                const double *x2, unsigned long long *rs, int n_offsets, int *spc, double *L, double *x, double *f1,
                double *f2, int *side) {
  double sh1[3], sh2[3], a[3], b[3];
  for (int q = 0; q < 3; q++) { sh1[q] = urand(rs); sh2[q] = urand(rs); }
  to_frac(L1, x1, N, f1);
  to_frac(L2, x2, N, f2);
  for (int i = 0; i < 3 * N; i++) {
    f1[i] = fmod(f1[i] + sh1[i % 3], 1.0); if (f1[i] < 0) f1[i] += 1.0;
    f2[i] = fmod(f2[i] + sh2[i % 3], 1.0); if (f2[i] < 0) f2[i] += 1.0;
  }
  ball(rs, a);
  for (int q = 0; q < 3; q++) a[q] = 0.5 + 0.25 * a[q];
  double nb;
  do { ball(rs, b); nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]); } while (nb == 0.0);
  for (int q = 0; q < 3; q++) b[q] /= nb;
  int n1 = 0, have[MG_MAXSPC] = {0, 0, 0, 0}, want[MG_MAXSPC] = {0, 0, 0, 0};
  for (int i = 0; i < N; i++) {
    want[s1[i]]++;
    const double d = (f1[3 * i] - a[0]) * b[0] + (f1[3 * i + 1] - a[1]) * b[1] + (f1[3 * i + 2] - a[2]) * b[2];
    side[i] = d < 0.0;
    if (side[i]) { n1++; have[s1[i]]++; }
  }
  /* lattice (MATE_LT): the second parent's, cyclically permuted to best match the first, with probability n1/N */
  double l1[3], a1[3], l2[3], a2[3];
  params(L1, l1, a1);
  params(L2, l2, a2);
  int bk = 0;
  double bw = INFINITY;
  for (int k = 0; k < 3; k++) {
    double w = 0.0;
    for (int q = 0; q < 3; q++) {
      const double t1 = (l1[q] - l2[(q + k) % 3]) / (l1[q] + l2[(q + k) % 3]);
      const double t2 = (a1[q] - a2[(q + k) % 3]) / (a1[q] + a2[(q + k) % 3]);
      w += t1 * t1 + t2 * t2;
    }
    if (w < bw) { bw = w; bk = k; }
  }
  if (urand(rs) < (double)n1 / N) {
    double lp[3], ap[3];
    for (int q = 0; q < 3; q++) { lp[q] = l2[(q + bk) % 3]; ap[q] = a2[(q + bk) % 3]; }
    if (!from_params(lp, ap, L)) return 0;
  } else {
    memcpy(L, L1, 9 * sizeof(double));
  }
  const double V = fabs(det3(L)), Vm = 0.5 * (fabs(det3(L1)) + fabs(det3(L2)));
  if (fabs(V - Vm) / Vm >= 0.1) return 0;
  /* second slab, plane offset stepped until count and composition match */
  int found = 0;
  for (int o = -n_offsets / 2; o <= n_offsets / 2 && !found; o++) {
    int n2 = 0, h2[MG_MAXSPC] = {0, 0, 0, 0};
    const double off = 0.2 * (double)o / n_offsets;
    for (int i = 0; i < N; i++) {
      const double d = (f2[3 * i] - a[0] - b[0] * off) * b[0] + (f2[3 * i + 1] - a[1] - b[1] * off) * b[1] +
                       (f2[3 * i + 2] - a[2] - b[2] * off) * b[2];
      side[N + i] = d > 0.0;
      if (side[N + i]) { n2++; h2[s2[i]]++; }
    }
    if (n1 + n2 != N) continue;
    found = 1;
    for (int q = 0; q < nspc; q++)
      if (have[q] + h2[q] != want[q]) found = 0;
  }
  if (!found) return 0;
  /* child: species-grouped, fractional positions carried into the chosen lattice */
  int k = 0;
  for (int q = 0; q < nspc; q++) {
    for (int i = 0; i < N; i++)
      if (side[i] && s1[i] == q) {
        spc[k] = q;
        for (int c = 0; c < 3; c++) x[3 * k + c] = f1[3 * i] * L[c] + f1[3 * i + 1] * L[3 + c] + f1[3 * i + 2] * L[6 + c];
        k++;
      }
    for (int i = 0; i < N; i++)
      if (side[N + i] && s2[i] == q) {
        spc[k] = q;
        for (int c = 0; c < 3; c++) x[3 * k + c] = f2[3 * i] * L[c] + f2[3 * i + 1] * L[3 + c] + f2[3 * i + 2] * L[6 + c];
        k++;
      }
  }
  return k == N;
}

/* count offspring of N atoms each from npar parents (parent p: spc[p*N..], lat[9p..], x[3pN..]) drawn with the
 * given weights; each child is a crossover, then a Gaussian displacement (sigma, A) of every atom, and must keep
 * all periodic distances >= dmin, after at most N/4 steps of push_apart as BULK_MATE repairs its children with
 * ADJT_CL (src/efnc.c:2194).  A child whose operators fail max_tries times in a row is flagged in failed[k]
 * (its slot is left unset).  Returns the number of failed slots, or < 0 on a bad argument. */
int mg_evo_offspring(int npar, int N, int nspc, const int *spc, const double *lat, const double *x,
                     const double *weights, int count, double sigma, double dmin, unsigned long long seed,
                     int max_tries, int *out_spc, double *out_lat, double *out_x, int *failed) {
  if (npar < 2 || N < 1 || nspc < 1 || nspc > MG_MAXSPC || count < 0 || !spc || !lat || !x || !weights || !out_spc ||
      !out_lat || !out_x || !failed)
    return mg_fail(MG_ERR_ARG, "mg_evo_offspring: bad argument");
  for (int i = 0; i < npar * N; i++)
    if (spc[i] < 0 || spc[i] >= nspc) return mg_fail(MG_ERR_ARG, "mg_evo_offspring: species out of range");
  double *cum = (double *)malloc(sizeof(double) * npar);
  double tot = 0.0;
  int npos = 0;
  for (int p = 0; p < npar; p++) {
    npos += weights[p] > 0.0;
    tot += weights[p] > 0.0 ? weights[p] : 0.0;
    cum[p] = tot;
  }
  if (npos < 2) { free(cum); return mg_fail(MG_ERR_ARG, "mg_evo_offspring: at least two parents need positive weight"); }
  int nfail = 0;
  /* every child draws from its own stream (seed, k): the result does not depend on the thread count */
#pragma omp parallel reduction(+ : nfail)
  {
  double *f1 = (double *)malloc(sizeof(double) * 3 * N), *f2 = (double *)malloc(sizeof(double) * 3 * N);
  double *F = (double *)malloc(sizeof(double) * 3 * N);
  int *side = (int *)malloc(sizeof(int) * 2 * N);
#pragma omp for schedule(dynamic, 4)
  for (int k = 0; k < count; k++) {
    unsigned long long rs = seed ^ (0xD1B54A32D192ED03ull * (unsigned long long)(k + 1));
    int ok = 0;
    for (int t = 0; t < max_tries && !ok; t++) {
      int k1, k2;
      {
        double r = urand(&rs) * tot;
        for (k1 = 0; k1 < npar - 1 && (cum[k1] <= r || !(weights[k1] > 0.0)); k1++) ;
        do {
          r = urand(&rs) * tot;
          for (k2 = 0; k2 < npar - 1 && (cum[k2] <= r || !(weights[k2] > 0.0)); k2++) ;
        } while (k2 == k1);
      }
      int *cs = out_spc + (size_t)k * N;
      double *cl = out_lat + 9 * (size_t)k, *cx = out_x + 3 * (size_t)k * N;
      if (!mate(N, nspc, spc + (size_t)k1 * N, lat + 9 * k1, x + 3 * (size_t)k1 * N, spc + (size_t)k2 * N, lat + 9 * k2,
                x + 3 * (size_t)k2 * N, &rs, 20, cs, cl, cx, f1, f2, side))
        continue;
      for (int i = 0; i < 3 * N; i++) cx[i] += sigma * nrand(&rs);
      ok = mg_evo_min_distance(cl, cx, N) >= dmin || push_apart(cl, cx, N, dmin, N / 4, f1, F);
    }
    failed[k] = !ok;
    nfail += !ok;
  }
  free(f1);
  free(f2);
  free(F);
  free(side);
  }
  free(cum);
  return nfail;
}

/* Gaussian deviate by Marsaglia's polar method (the reference's RANG, src/util.c:169-185, draws the same
 * distribution from its own generator) */
static double grand(unsigned long long *s) {
  double x1, x2, w;
  do {
    x1 = 2.0 * urand(s) - 1.0;
    x2 = 2.0 * urand(s) - 1.0;
    w = x1 * x1 + x2 * x2;
  } while (w >= 1.0 || w == 0.0);
  return x1 * sqrt(-2.0 * log(w) / w);
}

/* lattice reduction in the manner of SHRT_LV (src/cell.c:514-545): while a row has a projection onto another row of
 * more than half that row's squared length, subtract the nearest integer multiple of the other row; atoms keep their
 * Cartesian positions (the lattice still spans the same periodic crystal).  Returns 0 if a row is shorter than lmin. */
static int reduce_rows(double *L, double lmin) {
  for (int it = 0; it < 100; it++) {
    int changed = 0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        if (a == b) continue;
        const double *u = L + 3 * b;
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double pu = L[3 * a] * u[0] + L[3 * a + 1] * u[1] + L[3 * a + 2] * u[2];
        if (fabs(pu) > 0.5 * uu * (1.0 + 1e-12)) {
          const double m = nearbyint(pu / uu);
          for (int q = 0; q < 3; q++) L[3 * a + q] -= m * u[q];
          changed = 1;
        }
      }
    if (!changed) break;
  }
  for (int a = 0; a < 3; a++)
    if (sqrt(L[3 * a] * L[3 * a] + L[3 * a + 1] * L[3 * a + 1] + L[3 * a + 2] * L[3 * a + 2]) < lmin) return 0;
  return 1;
}

/* count mutants of N atoms each, after BULK_MUTE (src/efnc.c:2316-2386) with SHKE_CL (src/eutl.c:176-218): a parent
 * drawn with the given weights (roulette), its lattice strained by e = I + cl g (diagonal) and 0.5 cl g (symmetric
 * off-diagonal), g standard Gaussian, with the fractional coordinates kept; every coordinate then moved by a
 * Gaussian of ca A; the rows reduced (SHRT_LV) and no row shorter than 1 A; the volume kept within a factor 1.5 of the
 * parent's; then, with more than one species, each atom's position swapped with probability cs with that of a random
 * atom of another species (species are grouped); finally every periodic distance must be >= dmin, after at most
 * N/4 steps of push_apart as BULK_MUTE repairs its mutants with ADJT_CL (src/efnc.c:2360).  A slot whose tries
 * all fail is flagged in failed[k].  Layout as mg_evo_offspring.  Returns the number of failed slots, or < 0. */
int mg_evo_mutate(int npar, int N, int nspc, const int *spc, const double *lat, const double *x, const double *weights,
                  int count, double cl, double ca, double cs, double dmin, unsigned long long seed, int max_tries,
                  int *out_spc, double *out_lat, double *out_x, int *failed) {
  if (npar < 1 || N < 1 || nspc < 1 || nspc > MG_MAXSPC || count < 0 || !spc || !lat || !x || !weights || !out_spc ||
      !out_lat || !out_x || !failed || cl < 0.0 || ca < 0.0 || cs < 0.0 || cs > 1.0)
    return mg_fail(MG_ERR_ARG, "mg_evo_mutate: bad argument");
  for (int i = 0; i < npar * N; i++)
    if (spc[i] < 0 || spc[i] >= nspc) return mg_fail(MG_ERR_ARG, "mg_evo_mutate: species out of range");
  double *cum = (double *)malloc(sizeof(double) * npar);
  double tot = 0.0;
  for (int p = 0; p < npar; p++) {
    tot += weights[p] > 0.0 ? weights[p] : 0.0;
    cum[p] = tot;
  }
  if (!(tot > 0.0)) { free(cum); return mg_fail(MG_ERR_ARG, "mg_evo_mutate: no parent has positive weight"); }
  int nfail = 0;
#pragma omp parallel reduction(+ : nfail)
  {
  double *f = (double *)malloc(sizeof(double) * 3 * N), *F = (double *)malloc(sizeof(double) * 3 * N);
#pragma omp for schedule(dynamic, 4)
  for (int k = 0; k < count; k++) {
    unsigned long long rs = seed ^ (0x9E6C63D0676A9A99ull * (unsigned long long)(k + 1));
    int ok = 0;
    int *cs_ = out_spc + (size_t)k * N;
    double *cL = out_lat + 9 * (size_t)k, *cx = out_x + 3 * (size_t)k * N;
    for (int t = 0; t < max_tries && !ok; t++) {
      int p;
      const double r = urand(&rs) * tot;
      for (p = 0; p < npar - 1 && (cum[p] <= r || !(weights[p] > 0.0)); p++) ;
      const int *ps = spc + (size_t)p * N;
      const double *pL = lat + 9 * (size_t)p, *px = x + 3 * (size_t)p * N;
      memcpy(cs_, ps, sizeof(int) * N);
      to_frac(pL, px, N, f);
      /* strain e (symmetric), rows L'_i = sum_k e_ik L_k, fractional coordinates kept */
      double e[9];
      e[0] = 1.0 + cl * grand(&rs); e[4] = 1.0 + cl * grand(&rs); e[8] = 1.0 + cl * grand(&rs);
      e[1] = e[3] = 0.5 * cl * grand(&rs);
      e[2] = e[6] = 0.5 * cl * grand(&rs);
      e[5] = e[7] = 0.5 * cl * grand(&rs);
      for (int i = 0; i < 3; i++)
        for (int q = 0; q < 3; q++) cL[3 * i + q] = e[3 * i] * pL[q] + e[3 * i + 1] * pL[3 + q] + e[3 * i + 2] * pL[6 + q];
      for (int i = 0; i < N; i++)
        for (int q = 0; q < 3; q++)
          cx[3 * i + q] = f[3 * i] * cL[q] + f[3 * i + 1] * cL[3 + q] + f[3 * i + 2] * cL[6 + q] + ca * grand(&rs);
      if (!reduce_rows(cL, 1.0)) continue;
      const double V0 = fabs(det3(pL)), V1 = fabs(det3(cL));
      if (!(V0 > 0.0) || fabs(V1 / V0 - 1.0) > 0.5) continue;
      if (nspc > 1)
        for (int i = 0; i < N; i++)
          if (urand(&rs) < cs) {
            int n_i = 0, j;
            for (int a = 0; a < N; a++) n_i += cs_[a] == cs_[i];
            if (n_i >= N) continue;
            for (j = i; cs_[j % N] == cs_[i]; j++) ;  /* first atom after i's species block */
            j = (j + (int)(urand(&rs) * (double)(N - n_i))) % N;
            for (int q = 0; q < 3; q++) {
              const double tmp = cx[3 * i + q];
              cx[3 * i + q] = cx[3 * j + q];
              cx[3 * j + q] = tmp;
            }
          }
      ok = mg_evo_min_distance(cL, cx, N) >= dmin || push_apart(cL, cx, N, dmin, N / 4, f, F);
    }
    failed[k] = !ok;
    nfail += !ok;
  }
  free(f);
  free(F);
  }
  free(cum);
  return nfail;
}

/* Structure fingerprints after RDF(C, 1) (src/cell.c:369-415): per structure s (atoms atom_off[s]..atom_off[s+1],
 * lattice rows lat[9s], Cartesian x, nd = 3), for every ordered species pair m <= n a radial distribution on nrdf bins
 * of width dr = rmax / nrdf: each pair (i, j image) with spc[j] >= spc[i] at distance r adds
 * f(x_k) exp(-(r - x_k)^2 / (2 sigma^2)) to the bins x_k = k dr within 3 sigma of r, f = 1 below rmin and
 * cos(pi/2 (x - rmin)/(rmax - rmin)) above; same-species bins are halved; the whole fingerprint is scaled to unit
 * length.  out holds nspc(nspc+1)/2 * nrdf floats per structure, pair-major (pairs in the order (0,0), (0,1), ...,
 * (1,1), ...).  The dot product of two fingerprints is CxC (src/cell.c:417-427). */
int mg_evo_fingerprint(int S, const int *atom_off, int nspc, const int *spc, const double *lat, const double *x,
                       int nrdf, double rmax, double rmin, double sigma, float *out) {
  if (S < 0 || !atom_off || nspc < 1 || nspc > MG_MAXSPC || !spc || !lat || !x || nrdf < 1 || !(rmax > 0.0) ||
      !(sigma > 0.0) || !(rmin < rmax) || !out)
    return mg_fail(MG_ERR_ARG, "mg_evo_fingerprint: bad argument");
  const int npair = nspc * (nspc + 1) / 2;
  int pidx[MG_MAXSPC][MG_MAXSPC];
  for (int m = 0, p = 0; m < nspc; m++)
    for (int n = m; n < nspc; n++) pidx[m][n] = pidx[n][m] = p++;
  const double dr = rmax / (double)nrdf;
  const int Do = (int)(3.0 * sigma / dr);
  const double DR = 0.5 / (sigma * sigma);
  const double rcut = rmax + 3.0 * sigma + dr;
  const double cstep = exp(-2.0 * DR * dr * dr);
  /* the taper f(x_k) of every bin, once */
  double *fk = (double *)malloc(sizeof(double) * (size_t)nrdf);
  for (int k = 0; k < nrdf; k++) {
    const double xk = (double)k * dr;
    fk[k] = xk < rmin ? 1.0 : cos(0.5 * 3.14159265358979323846 * (xk - rmin) / (rmax - rmin));
  }
  /* structures in parallel on the host cores, each with its own accumulator */
#pragma omp parallel
  {
  double *acc = (double *)malloc(sizeof(double) * (size_t)npair * nrdf);
  double *fr = NULL;
  int fr_cap = 0;
#pragma omp for schedule(dynamic, 4)
  for (int s = 0; s < S; s++) {
    const int a0 = atom_off[s], n = atom_off[s + 1] - a0;
    const double *L = lat + 9 * (size_t)s, *X = x + 3 * (size_t)a0;
    const int *sp = spc + a0;
    memset(acc, 0, sizeof(double) * (size_t)npair * nrdf);
    double I[9], rr[3];
    inv3(L, I);
    for (int q = 0; q < 3; q++)  /* a pair closer than rcut has fractional separation below rcut |column q of L^-1| */
      rr[q] = rcut * sqrt(I[q] * I[q] + I[3 + q] * I[3 + q] + I[6 + q] * I[6 + q]);
    if (n > fr_cap) {
      free(fr);
      fr_cap = n;
      fr = (double *)malloc(sizeof(double) * 3 * (size_t)fr_cap);
    }
    to_frac(L, X, n, fr);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        if (sp[j] < sp[i]) continue;
        const int p = pidx[sp[i]][sp[j]];
        /* only the images whose fractional separation can be below rr (the others are out of range) */
        int lo[3], hi[3];
        for (int q = 0; q < 3; q++) {
          const double df = fr[3 * j + q] - fr[3 * i + q];
          lo[q] = (int)ceil(-rr[q] - df) - 1;
          hi[q] = (int)floor(rr[q] - df) + 1;
        }
        for (int a = lo[0]; a <= hi[0]; a++)
          for (int b = lo[1]; b <= hi[1]; b++)
            for (int c = lo[2]; c <= hi[2]; c++) {
              if (i == j && a == 0 && b == 0 && c == 0) continue;
              double d[3], r2 = 0.0;
              for (int q = 0; q < 3; q++) {
                d[q] = X[3 * j + q] - X[3 * i + q] + a * L[q] + b * L[3 + q] + c * L[6 + q];
                r2 += d[q] * d[q];
              }
              if (r2 >= rcut * rcut) continue;
              const double r = sqrt(r2);
              const int o = (int)(r / dr);
              /* exp(-DR (r - x_k)^2) along the bins by recurrence: g_{k+1} = g_k q_k, q_{k+1} = q_k c */
              const int k0 = o - Do;
              double xk = (double)k0 * dr;
              double g = exp(-(r - xk) * (r - xk) * DR), qk = exp(2.0 * DR * dr * (r - xk) - DR * dr * dr);
              for (int k = k0; k <= o + Do; k++) {
                if (k >= 0 && k < nrdf) acc[(size_t)p * nrdf + k] += fk[k] * g;
                g *= qk;
                qk *= cstep;
              }
            }
      }
    for (int m = 0; m < nspc; m++)
      for (int k = 0; k < nrdf; k++) acc[(size_t)pidx[m][m] * nrdf + k] *= 0.5;
    double nrm = 0.0;
    for (size_t q = 0; q < (size_t)npair * nrdf; q++) nrm += acc[q] * acc[q];
    nrm = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    float *o = out + (size_t)s * npair * nrdf;
    for (size_t q = 0; q < (size_t)npair * nrdf; q++) o[q] = (float)(acc[q] * nrm);
  }
  free(acc);
  free(fr);
  }
  free(fk);
  return MG_OK;
}

/* Duplicate walk of the population update: structures in increasing energy (ties by index, non-finite energies
 * skipped); a structure is kept unless COMP_CL (src/eutl.c:394-413) against one already kept exceeds cut:
 *   w = 0.8 CxC + 0.1 exp(-dE^2 / (2 0.05^2)) + 0.1 exp(-dV^2 / (2 0.01^2)),  dV = (V_a - V_b) / (2 (V_a + V_b)),
 * scored only for pairs with E_k - E_kept <= de_win and |dV| <= dv_win (outside them w cannot exceed cut when
 * cut >= 0.95).  The first n_unique structures are free of duplicates among themselves, so two of them are never
 * compared.  G[(k - n_unique) * n + j] = CxC of structure k >= n_unique with structure j.  Stops at limit kept;
 * returns their number (indices in keep, in the order kept). */
typedef struct { double e; int k; } ek_pair;
static int ek_cmp(const void *a, const void *b) {
  const ek_pair *x = (const ek_pair *)a, *y = (const ek_pair *)b;
  if (x->e < y->e) return -1;
  if (x->e > y->e) return 1;
  return x->k < y->k ? -1 : (x->k > y->k);
}

int mg_evo_unique(int n, const double *E, const double *V, const float *G, int n_unique, double cut, double de_win,
                  double dv_win, int limit, int *keep) {
  if (n < 0 || (n > 0 && (!E || !V || !keep)) || n_unique < 0 || n_unique > n || limit < 0)
    return mg_fail(MG_ERR_ARG, "mg_evo_unique: bad argument");
  const int test = cut < 1.0 && n > n_unique;
  if (test && !G) return mg_fail(MG_ERR_ARG, "mg_evo_unique: fingerprint products missing");
  ek_pair *ord = (ek_pair *)malloc(sizeof(ek_pair) * (n > 0 ? n : 1));
  if (!ord) return mg_fail(MG_ERR_ALLOC, "mg_evo_unique: out of memory");
  int m = 0;
  for (int k = 0; k < n; k++)
    if (isfinite(E[k])) { ord[m].e = E[k]; ord[m].k = k; m++; }
  qsort(ord, m, sizeof(ek_pair), ek_cmp);
  int nk = 0;
  const double se = 1.0 / (0.05 * 0.05), sv = 1.0 / (0.01 * 0.01);
  for (int t = 0; t < m && nk < limit; t++) {
    const int k = ord[t].k;
    int dup = 0;
    if (test)
      for (int c = 0; c < nk && !dup; c++) {
        const int j = keep[c];
        if (k < n_unique && j < n_unique) continue;  /* two members of the duplicate-free block */
        const double de = E[j] - E[k];
        const double dv = 0.5 * (V[k] - V[j]) / (V[k] + V[j]);
        if (-de > de_win || fabs(dv) > dv_win) continue;
        const double cxc = k >= n_unique ? (double)G[(size_t)(k - n_unique) * n + j]
                                         : (double)G[(size_t)(j - n_unique) * n + k];
        const double w = 0.8 * cxc + 0.1 * exp(-0.5 * de * de * se) + 0.1 * exp(-0.5 * dv * dv * sv);
        if (w > cut) dup = 1;
      }
    if (!dup) keep[nk++] = k;
  }
  free(ord);
  return nk;
}
