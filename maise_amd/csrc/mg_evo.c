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
 * all periodic distances >= dmin.  A child whose operators fail max_tries times in a row is flagged in failed[k]
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
  double *f1 = (double *)malloc(sizeof(double) * 3 * N), *f2 = (double *)malloc(sizeof(double) * 3 * N);
  int *side = (int *)malloc(sizeof(int) * 2 * N);
  int nfail = 0;
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
      ok = mg_evo_min_distance(cl, cx, N) >= dmin;
    }
    failed[k] = !ok;
    nfail += !ok;
  }
  free(cum);
  free(f1);
  free(f2);
  free(side);
  return nfail;
}
