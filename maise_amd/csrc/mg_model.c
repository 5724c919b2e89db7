/*
 * mg_model.c -- reads a MAISE nn_* model file (network + appended Behler-Parrinello basis) and derives the
 * FP32 device tables.
 *
 * File semantics follow the reference readers:
 *   version gate ............ check_ver, src/util.c:1948-1996 (>= 2.4.00, line 2 must say "neural")
 *   header fields ........... READ_MODEL / READ_ANN, src/util.c:1233-1335, src/nutl.c:502-633 (25-char tags)
 *   weight order ............ V2W, src/nmlp.c:591-608 (per species: biases of each layer, then [out][in])
 *   input-scaling fold ...... W0 *= range, Rmin /= range, src/nutl.c:621-630 (done in FP64, then cast)
 *   basis block ............. Build_PRS, src/nprs.c:154-231 (NSYM rows, 7 parameter rows, unit scaling)
 *   species-pair mixing ..... src/nprs.c:233-252 (Rc, Rs averaged; eta = (mean of eta^-1/2)^-2)
 *   block offsets ........... src/nprs.c:484-498
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

#include "mg_host.h"

static const double MG_MINDIST[96] = {  /* UREP radii per element, src/util.c:13 */
  0.0,1.0,1.0,1.040,0.773,0.584,0.499,1.0,1.0,1.0,1.0,1.266,1.114,1.0,1.0,1.0,1.0,1.0,1.0,1.610,1.366,1.123,1.0,1.0,
  1.0,1.0,1.0,1.0,1.0,0.895,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,0.979,1.030,1.0,1.0,
  1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,
  0.984,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0,1.0};

typedef struct {
  char **line;
  int n;
} text_t;

static int text_read(const char *path, text_t *t) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  int cap = 1024;
  t->line = (char **)malloc(sizeof(char *) * cap);
  t->n = 0;
  char buf[4096];
  while (fgets(buf, sizeof buf, f)) {
    if (t->n == cap) {
      cap *= 2;
      t->line = (char **)realloc(t->line, sizeof(char *) * cap);
    }
    t->line[t->n++] = strdup(buf);
  }
  fclose(f);
  return 0;
}

static void text_free(text_t *t) {
  for (int i = 0; i < t->n; i++) free(t->line[i]);
  free(t->line);
}

static int find_tag(const text_t *t, int from, const char *tag, int len) {
  for (int i = from; i < t->n; i++)
    if (strncmp(t->line[i], tag, len) == 0) return i;
  return -1;
}

static int parse_ints(const char *s, int n, int *v) {
  int k = 0, used = 0;
  for (int i = 0; i < n; i++) {
    if (sscanf(s + k, "%d%n", &v[i], &used) != 1) return i;
    k += used;
  }
  return n;
}

/* version number n*1000 + k*100 + d from "maise.n.k.d" */
static int model_version(const text_t *t) {
  if (t->n < 2) return -1;
  char w[256] = {0};
  sscanf(t->line[1] + 2, "%255s", w);
  if (strncmp(w, "neural", 6) != 0) return -1;
  int li = find_tag(t, 0, "|  maise version", 16);
  if (li < 0) return 0;
  char v[256] = {0};
  sscanf(t->line[li], "|  maise version | %255s |", v);
  int part[3] = {0, 0, 0};
  size_t i = 6, L = strlen(v);
  for (int f = 0; f < 3 && i < L; f++) {
    char p[64];
    size_t len = strcspn(v + i, ".");
    if (len > 63) len = 63;
    memcpy(p, v + i, len);
    p[len] = 0;
    sscanf(p, "%d", &part[f]);
    i += len + 1;
  }
  return part[0] * 1000 + part[1] * 100 + part[2];
}

typedef struct {
  int gf[256][8];
  double gp[7][10];
  int gt[7];
} basis64_t;

static int parse_basis(const text_t *t, int at, int nsym, basis64_t *b, int *next) {
  /* at: line after "  B2A" header */
  double a = 0, r = 0;
  if (at >= t->n || sscanf(t->line[at], "%lf %lf", &a, &r) != 2) return -1;
  int li = at + 4;
  for (int n = 0; n < nsym; n++, li++) {
    int idx;
    if (li >= t->n) return -1;
    if (sscanf(t->line[li], "%d %d %d %d %d %d %d %d %d", &idx, &b->gf[n][0], &b->gf[n][1], &b->gf[n][2], &b->gf[n][3],
               &b->gf[n][4], &b->gf[n][5], &b->gf[n][6], &b->gf[n][7]) != 9)
      return -1;
  }
  li++; /* blank */
  for (int p = 0; p < 7; p++, li++) {
    char name[64];
    double *g = b->gp[p];
    for (int i = 0; i < 10; i++) g[i] = 0.0;
    if (li >= t->n) return -1;
    if (sscanf(t->line[li], "%63s %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", name, &b->gt[p], &g[0], &g[1], &g[2],
               &g[3], &g[4], &g[5], &g[6], &g[7], &g[8], &g[9]) < 3)
      return -1;
    if (b->gt[p] < 1 || b->gt[p] > 10) return -1;
  }
  for (int p = 0; p < 2; p++)
    for (int i = 0; i < 10; i++) b->gp[p][i] *= a * r;
  for (int p = 2; p < 4; p++)
    for (int i = 0; i < 10; i++) b->gp[p][i] /= a * a * r * r;
  *next = li;
  return 0;
}

void mg_model_free(mg_model *m) {
  if (!m) return;
  free(m->mlp);
  free(m->mlp64);
  free(m->frag);
  free(m->frag_bf);
  free(m->frag_lo);
  free(m->fvec);
  free(m);
}

static inline int acc_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

static uint16_t f32_to_bf16(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return (uint16_t)(b >> 16); /* inf/nan */
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

static float bf16_to_f32(uint16_t h) {
  uint32_t b = (uint32_t)h << 16;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

/* a = hi + lo with hi, lo BF16 (lo rounds the residual) */
static void put_split(uint16_t *hi, uint16_t *lo, size_t k, float a) {
  hi[k] = f32_to_bf16(a);
  lo[k] = f32_to_bf16(a - bf16_to_f32(hi[k]));
}

/* Weight fragments (A operands) for the MFMA MLP in the k order of the kernel (mg_kernels.hip, k_mlp_mfma):
 * accumulator register r of lane l holds unit 8(r/4) + 4(l/32) + r%4 of atom l%32 within a 32-unit block. */
static int pack_mfma(mg_model *m) {
  const int D = m->D, H1 = m->H1, H2 = m->H2;
  m->mfma_ok = H1 <= 64 && H2 <= 64;
  if (!m->mfma_ok) return MG_OK;
  const int hb1 = (H1 + 31) / 32, hb2 = (H2 + 31) / 32, db = (D + 31) / 32;
  const int k1f = (D + 1) / 2, k1b = (D + 15) / 16;
  int *fd = m->fdesc;
  fd[0] = D; fd[1] = H1; fd[2] = H2; fd[3] = hb1; fd[4] = hb2; fd[5] = db; fd[6] = k1b; fd[7] = k1f;
  /* bf16 set: offsets in slots */
  int ob[4], of[4];
  ob[0] = 0; ob[1] = ob[0] + hb1 * k1b; ob[2] = ob[1] + hb2 * hb1 * 2; ob[3] = ob[2] + hb1 * hb2 * 2;
  const int sb = ob[3] + db * hb1 * 2;
  of[0] = 0; of[1] = of[0] + hb1 * k1f; of[2] = of[1] + hb2 * hb1 * 16; of[3] = of[2] + hb1 * hb2 * 16;
  const int sf = of[3] + db * hb1 * 16;
  const int sv = hb1 * 32 + 2 * hb2 * 32 + 1 + D;
  for (int k = 0; k < 4; k++) { fd[8 + k] = ob[k]; fd[14 + k] = of[k]; }
  fd[12] = sb; fd[13] = sv; fd[18] = sf; fd[19] = sv;
  const int S = m->nspc;
  m->frag = (float *)calloc((size_t)S * sf * 64, sizeof(float));
  m->frag_bf = (uint16_t *)calloc((size_t)S * sb * 64 * 8, sizeof(uint16_t));
  m->frag_lo = (uint16_t *)calloc((size_t)S * sb * 64 * 8, sizeof(uint16_t));
  m->fvec = (float *)calloc((size_t)S * sv, sizeof(float));
  m->frag_n = (size_t)S * sf * 64;
  m->frag_bf_n = (size_t)S * sb * 64 * 8;
  m->fvec_n = (size_t)S * sv;
  for (int s = 0; s < S; s++) {
    const double *w = m->mlp64 + (size_t)s * m->wstride;
    const double *W0 = w, *b0 = W0 + H1 * D, *W1 = b0 + H1, *b1 = W1 + H2 * H1, *W2 = b1 + H2, *b2 = W2 + H2, *xo = b2 + 1;
#define W0F(u, k) ((u) < H1 && (k) < D ? W0[(u) * D + (k)] : 0.0)
#define W1F(o, u) ((o) < H2 && (u) < H1 ? W1[(o) * H1 + (u)] : 0.0)
    float *F = m->frag + (size_t)s * sf * 64;
    uint16_t *B = m->frag_bf + (size_t)s * sb * 64 * 8;
    uint16_t *Bl = m->frag_lo + (size_t)s * sb * 64 * 8;
    for (int l = 0; l < 64; l++) {
      const int c = l & 31, h = l >> 5;
      /* FP32 (K = 2 per step, one value per lane) */
      for (int rb = 0; rb < hb1; rb++)
        for (int st = 0; st < k1f; st++) F[(of[0] + rb * k1f + st) * 64 + l] = (float)W0F(rb * 32 + c, 2 * st + h);
      for (int rb2 = 0; rb2 < hb2; rb2++)
        for (int rb1 = 0; rb1 < hb1; rb1++)
          for (int r = 0; r < 16; r++)
            F[(of[1] + rb2 * hb1 * 16 + rb1 * 16 + r) * 64 + l] = (float)W1F(rb2 * 32 + c, rb1 * 32 + acc_row(r, h));
      for (int rb1 = 0; rb1 < hb1; rb1++)
        for (int rb2 = 0; rb2 < hb2; rb2++)
          for (int r = 0; r < 16; r++)
            F[(of[2] + rb1 * hb2 * 16 + rb2 * 16 + r) * 64 + l] = (float)W1F(rb2 * 32 + acc_row(r, h), rb1 * 32 + c);
      for (int rbd = 0; rbd < db; rbd++)
        for (int rb1 = 0; rb1 < hb1; rb1++)
          for (int r = 0; r < 16; r++)
            F[(of[3] + rbd * hb1 * 16 + rb1 * 16 + r) * 64 + l] = (float)W0F(rb1 * 32 + acc_row(r, h), rbd * 32 + c);
      /* BF16 (K = 16 per step, eight values per lane) */
      for (int v = 0; v < 8; v++) {
        for (int rb = 0; rb < hb1; rb++)
          for (int st = 0; st < k1b; st++)
            put_split(B, Bl, ((ob[0] + rb * k1b + st) * 64 + l) * 8 + v, (float)W0F(rb * 32 + c, 16 * st + 8 * h + v));
        for (int rb2 = 0; rb2 < hb2; rb2++)
          for (int rb1 = 0; rb1 < hb1; rb1++)
            for (int q = 0; q < 2; q++)
              put_split(B, Bl, ((ob[1] + rb2 * hb1 * 2 + rb1 * 2 + q) * 64 + l) * 8 + v,
                        (float)W1F(rb2 * 32 + c, rb1 * 32 + acc_row(8 * q + v, h)));
        for (int rb1 = 0; rb1 < hb1; rb1++)
          for (int rb2 = 0; rb2 < hb2; rb2++)
            for (int q = 0; q < 2; q++)
              put_split(B, Bl, ((ob[2] + rb1 * hb2 * 2 + rb2 * 2 + q) * 64 + l) * 8 + v,
                        (float)W1F(rb2 * 32 + acc_row(8 * q + v, h), rb1 * 32 + c));
        for (int rbd = 0; rbd < db; rbd++)
          for (int rb1 = 0; rb1 < hb1; rb1++)
            for (int q = 0; q < 2; q++)
              put_split(B, Bl, ((ob[3] + rbd * hb1 * 2 + rb1 * 2 + q) * 64 + l) * 8 + v,
                        (float)W0F(rb1 * 32 + acc_row(8 * q + v, h), rbd * 32 + c));
      }
    }
#undef W0F
#undef W1F
    float *V = m->fvec + (size_t)s * sv;
    for (int u = 0; u < hb1 * 32; u++) V[u] = u < H1 ? (float)b0[u] : 0.0f;
    for (int u = 0; u < hb2 * 32; u++) V[hb1 * 32 + u] = u < H2 ? (float)b1[u] : 0.0f;
    for (int u = 0; u < hb2 * 32; u++) V[hb1 * 32 + hb2 * 32 + u] = u < H2 ? (float)W2[u] : 0.0f;
    V[hb1 * 32 + 2 * hb2 * 32] = (float)b2[0];
    for (int k = 0; k < D; k++) V[hb1 * 32 + 2 * hb2 * 32 + 1 + k] = (float)xo[k];
  }
  return MG_OK;
}

static int pack_mfma(mg_model *m);

int mg_model_load(const char *path, mg_model **out) {
  if (!path || !out) return mg_fail(MG_ERR_ARG, "mg_model_load: null argument");
  *out = NULL;
  text_t t;
  if (text_read(path, &t) != 0) return mg_fail(MG_ERR_MODEL, "cannot open model file '%s'", path);
  int rc = MG_ERR_MODEL;
  mg_model *m = (mg_model *)calloc(1, sizeof(mg_model));
  double *V = NULL, *rmin = NULL, *rmax = NULL, *dr = NULL;
  int ver = model_version(&t);
  if (ver < 2400) { mg_fail(rc, "'%s' is not a neural-network model of version >= 2.4.00", path); goto done; }
  int li;
  if ((li = find_tag(&t, 0, "|  model unique ID      |", 25)) >= 0) sscanf(t.line[li] + 26, "%63s", m->id);
  if ((li = find_tag(&t, 0, "|  number of species    |", 25)) < 0 || sscanf(t.line[li] + 26, "%d", &m->nspc) != 1 ||
      m->nspc < 1 || m->nspc > MG_MAXSPC) {
    mg_fail(rc, "model: bad or unsupported number of species (1..%d)", MG_MAXSPC); goto done;
  }
  if ((li = find_tag(&t, 0, "|  species types        |", 25)) < 0 || parse_ints(t.line[li] + 26, m->nspc, m->spcz) != m->nspc) {
    mg_fail(rc, "model: bad species types"); goto done;
  }
  int nl = 0, nu[4] = {0, 0, 0, 0};
  if ((li = find_tag(&t, 0, "|  number of layers     |", 25)) < 0 || sscanf(t.line[li] + 26, "%d", &nl) != 1 || nl != 4) {
    mg_fail(MG_ERR_UNSUPPORTED, "model: only D-H1-H2-1 networks (4 layers) are supported"); rc = MG_ERR_UNSUPPORTED; goto done;
  }
  if ((li = find_tag(&t, 0, "|  architecture         |", 25)) < 0 || parse_ints(t.line[li] + 26, 4, nu) != 4 || nu[3] != 1) {
    mg_fail(rc, "model: bad architecture"); goto done;
  }
  m->D = nu[0]; m->H1 = nu[1]; m->H2 = nu[2];
  if (m->D < 1 || m->H1 < 1 || m->H2 < 1 || m->D > 4096 || m->H1 > 1024 || m->H2 > 1024) { mg_fail(rc, "model: bad layer sizes"); goto done; }
  if ((li = find_tag(&t, 0, "|  number of weights    |", 25)) < 0 || sscanf(t.line[li] + 26, "%d", &m->nw) != 1) {
    mg_fail(rc, "model: missing number of weights"); goto done;
  }
  {
    int nw = m->nspc * ((m->D + 1) * m->H1 + (m->H1 + 1) * m->H2 + (m->H2 + 1));
    if (nw != m->nw) { mg_fail(rc, "model: %d weights declared, architecture needs %d", m->nw, nw); goto done; }
  }
  /* input ranges and weights */
  if ((li = find_tag(&t, 0, "        minimum          ", 25)) < 0) { mg_fail(rc, "model: missing input range block"); goto done; }
  li++;
  {
    size_t nd = (size_t)m->nspc * m->D;
    rmin = (double *)malloc(sizeof(double) * nd);
    rmax = (double *)malloc(sizeof(double) * nd);
    dr = (double *)malloc(sizeof(double) * nd);
    for (size_t k = 0; k < nd; k++, li++)
      if (li >= t.n || sscanf(t.line[li], "%lf %lf %lf", &rmin[k], &rmax[k], &dr[k]) != 3) {
        mg_fail(rc, "model: short input range block"); goto done;
      }
    li += 3;
    V = (double *)malloc(sizeof(double) * m->nw);
    for (int k = 0; k < m->nw; k++, li++)
      if (li >= t.n || sscanf(t.line[li], "%lf", &V[k]) != 1) { mg_fail(rc, "model: short weight list"); goto done; }
  }
  /* basis */
  int b2a = find_tag(&t, 0, "  B2A", 5);
  if (b2a < 0) { mg_fail(rc, "model: no basis block (\"  B2A\")"); goto done; }
  {
    int rcl = find_tag(&t, b2a + 5, " Rc  ", 5);
    if (rcl < 0) { mg_fail(rc, "model: basis has no Rc row"); goto done; }
    m->nsym = rcl - (b2a + 5) - 1;
    if (m->nsym < 1 || m->nsym > 256) { mg_fail(rc, "model: bad number of symmetry functions"); goto done; }
  }
  basis64_t *bs = (basis64_t *)calloc(MG_MAXSPC * MG_MAXSPC + 4, sizeof(basis64_t)); /* by s0*s0+s1*s1 (max 18) */
  {
    int at = b2a + 1;
    for (int w0 = 0; w0 < m->nspc; w0++) {
      int next = 0;
      if (parse_basis(&t, at, m->nsym, &bs[2 * w0 * w0], &next) != 0) { free(bs); mg_fail(rc, "model: cannot parse basis block"); goto done; }
      /* a further block (another "  B2A" header right after) is species-specific; otherwise reuse block 0 */
      int nb = find_tag(&t, next, "  B2A", 5);
      at = nb >= 0 ? nb + 1 : b2a + 1;
    }
  }
  basis64_t *b0 = &bs[0];
  m->ng2 = m->ng4 = 0;
  for (int n = 0; n < m->nsym; n++) {
    if (b0->gf[n][0] == 2) m->ng2++;
    else if (b0->gf[n][0] == 4) m->ng4++;
    else { free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: only G2/G4 symmetry functions are supported"); goto done; }
  }
  for (int n = 0; n < m->ng2; n++)
    if (b0->gf[n][0] != 2) { free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: G2 functions must precede G4 functions"); goto done; }
  {
    int c = m->nspc * (m->nspc + 1) / 2;
    if (m->ng2 * m->nspc + m->ng4 * c != m->D) {
      free(bs); mg_fail(rc, "model: descriptor width %d does not match the basis (%d)", m->D, m->ng2 * m->nspc + m->ng4 * c); goto done;
    }
  }
  if (m->ng2 > MG_MAXG2 || m->ng4 > MG_MAXG4) { free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: too many functions"); goto done; }
  /* mixing of cross-species parameters */
  for (int w0 = 0; w0 < m->nspc; w0++)
    for (int w1 = 0; w1 < m->nspc; w1++)
      if (w0 != w1) {
        basis64_t *a = &bs[2 * w0 * w0], *b = &bs[2 * w1 * w1], *c = &bs[w0 * w0 + w1 * w1];
        *c = *b0;
        for (int p = 0; p < 2; p++)
          for (int i = 0; i < 10; i++) c->gp[p][i] = 0.5 * (a->gp[p][i] + b->gp[p][i]);
        for (int p = 2; p < 4; p++)
          for (int i = 0; i < 10; i++) c->gp[p][i] = pow(0.5 * (pow(a->gp[p][i], -0.5) + pow(b->gp[p][i], -0.5)), -2.0);
      }
  /* list cutoff: smallest function cutoff of the same-species blocks */
  m->rc_list = b0->gp[0][b0->gf[0][1]];
  for (int s0 = 0; s0 < m->nspc; s0++)
    for (int n = 0; n < m->nsym; n++) {
      double v = bs[2 * s0 * s0].gp[0][b0->gf[n][1]];
      if (m->rc_list > v) m->rc_list = v;
    }
  /* device tables */
  {
    mg_basis_dev *B = &m->basis;
    memset(B, 0, sizeof *B);
    B->nspc = m->nspc; B->D = m->D; B->ng2 = m->ng2; B->ng4 = m->ng4;
    B->ne = b0->gt[3]; B->nz = b0->gt[5]; B->nl = b0->gt[6];
    if (m->ng4 > 0 && (B->ne > MG_MAXE || B->nz > MG_MAXZ || B->nl > MG_MAXL)) {
      free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: G4 grid too large (eta %d, zeta %d, lambda %d)", B->ne, B->nz, B->nl); goto done;
    }
    int ind2[4] = {0, m->ng2 + m->ng4, 2 * m->ng2 + 3 * m->ng4, 3 * m->ng2 + 6 * m->ng4};
    int ind4[20];
    for (int k = 0; k < 20; k++) ind4[k] = 0;
    ind4[0] = m->ng2; ind4[1] = 2 * m->ng2 + m->ng4; ind4[2] = 2 * m->ng2 + 2 * m->ng4;
    ind4[4] = 3 * m->ng2 + 3 * m->ng4; ind4[8] = 3 * m->ng2 + 4 * m->ng4; ind4[5] = 3 * m->ng2 + 5 * m->ng4;
    ind4[9] = 4 * m->ng2 + 6 * m->ng4; ind4[18] = 4 * m->ng2 + 7 * m->ng4; ind4[10] = 4 * m->ng2 + 8 * m->ng4;
    ind4[13] = 4 * m->ng2 + 9 * m->ng4;
    for (int a = 0; a < MG_MAXSPC; a++) {
      B->ind2[a] = ind2[a];
      for (int b = 0; b < MG_MAXSPC; b++) B->ind4[a][b] = ind4[a * a + b * b];
    }
    int g2first = m->ng2;
    for (int s0 = 0; s0 < m->nspc; s0++)
      for (int s1 = 0; s1 < m->nspc; s1++) {
        basis64_t *w = &bs[s0 * s0 + s1 * s1];
        int p = s0 * MG_MAXSPC + s1;
        double rcg2 = w->gp[0][b0->gf[0][1]];
        double rcg4 = m->ng4 ? w->gp[0][b0->gf[g2first][1]] : rcg2;
        if (fabs(rcg2 - rcg4) > 1e-12 * rcg2) {
          free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: different G2 and G4 cutoffs are not supported"); goto done;
        }
        B->rc[p] = (float)rcg2;
        for (int n = 0, c = 0; n < m->nsym; n++)
          if (b0->gf[n][0] == 2) {
            B->eta2[p][c] = (float)w->gp[2][b0->gf[n][3]];
            B->rs2[p][c] = (float)w->gp[1][b0->gf[n][2]];
            c++;
          }
        for (int e = 0; e < B->ne && e < MG_MAXE; e++) B->eta4[p][e] = (float)w->gp[3][e];
      }
    for (int z = 0; z < B->nz && z < MG_MAXZ; z++) {
      double zv = b0->gp[5][z];
      if (m->ng4 > 0 && (zv < 1.0 || zv > 64.0 || fabs(zv - floor(zv)) > 0)) {
        free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: only integer zeta in [1, 64] is supported"); goto done;
      }
      B->zeta[z] = (int)zv;
      B->pref[z] = (float)pow(2.0, 1.0 - zv);
    }
    for (int z = B->nz; z < MG_MAXZ; z++) { B->zeta[z] = 1; B->pref[z] = 0.0f; }
    for (int l = 0; l < B->nl && l < MG_MAXL; l++) B->lam[l] = (float)b0->gp[6][l];
    for (int q = 0; q < MG_GRID; q++) B->g4map[q] = -1;
    for (int n = 0, c = 0; n < m->nsym; n++)
      if (b0->gf[n][0] == 4) {
        int e = b0->gf[n][4], z = b0->gf[n][6], l = b0->gf[n][7];
        if (e < 0 || e >= B->ne || z < 0 || z >= B->nz || l < 0 || l >= B->nl) {
          free(bs); mg_fail(rc, "model: G4 function %d indexes a missing parameter", n); goto done;
        }
        int cell = (e * MG_MAXZ + z) * MG_MAXL + l;
        if (B->g4map[cell] >= 0) { free(bs); rc = MG_ERR_UNSUPPORTED; mg_fail(rc, "model: duplicate G4 function"); goto done; }
        B->g4map[cell] = c;
        B->g4cell[c] = cell;
        c++;
      }
    /* kernel specialisation: the two G4 tables of the shipped models (43 and 22 functions) with
       zeta = {1,2,4,16}, lambda = {-1,+1}; anything else runs the generic dense-grid kernels */
    m->pat = 0;
    if (B->nz == 4 && B->nl == 2 && B->zeta[0] == 1 && B->zeta[1] == 2 && B->zeta[2] == 4 && B->zeta[3] == 16 &&
        B->lam[0] == -1.0f && B->lam[1] == 1.0f) {
      static const unsigned m43[8] = {0x0F, 0x0F, 0x0F, 0xFF, 0xFF, 0xFF, 0x7F, 0x00};
      static const unsigned m22[8] = {0x0F, 0x0F, 0x0A, 0xAA, 0xAA, 0xAA, 0x00, 0x00};
      for (int cand = 1; cand <= 2 && m->pat == 0; cand++) {
        const unsigned *mk = cand == 1 ? m43 : m22;
        /* eta rows past the table's (7 or 6) must be unused, which the mask comparison below checks */
        int ok = (cand == 1 && B->ne >= 7) || (cand == 2 && B->ne >= 6);
        for (int q = 0; q < MG_GRID && ok; q++) {
          int used = (mk[q / 8] >> (q % 8)) & 1u;
          if (used != (B->g4map[q] >= 0)) ok = 0;
        }
        if (ok) m->pat = cand;
      }
    }
    /* all species pairs share the G4 etas (single-basis models): one exp per triplet and eta */
    m->uni = 1;
    for (int s0 = 0; s0 < m->nspc; s0++)
      for (int s1 = 0; s1 < m->nspc; s1++)
        for (int e = 0; e < B->ne && e < MG_MAXE; e++)
          if (B->eta4[s0 * MG_MAXSPC + s1][e] != B->eta4[0][e]) m->uni = 0;
    m->eta_nonneg = 1;
    for (int p = 0; p < MG_MAXSPC * MG_MAXSPC; p++)
      for (int e = 0; e < B->ne && e < MG_MAXE; e++)
        if (B->eta4[p][e] < 0.0f) m->eta_nonneg = 0;
    for (int s = 0; s < m->nspc; s++) B->urep_rm[s] = (m->spcz[s] >= 0 && m->spcz[s] < 96) ? (float)MG_MINDIST[m->spcz[s]] : 1.0f;
  }
  free(bs);
  /* MLP: fold the input scaling in FP64, pack per species */
  {
    const int D = m->D, H1 = m->H1, H2 = m->H2;
    int stride = H1 * D + H1 + H2 * H1 + H2 + H2 + 1 + D;
    stride = (stride + 3) & ~3;
    m->wstride = stride;
    m->mlp = (float *)calloc((size_t)stride * m->nspc, sizeof(float));
    m->mlp64 = (double *)calloc((size_t)stride * m->nspc, sizeof(double));
    int v = 0;
    for (int s = 0; s < m->nspc; s++) {
      double *w = m->mlp64 + (size_t)s * stride;
      double *W0 = w, *b0v = W0 + H1 * D, *W1 = b0v + H1, *b1v = W1 + H2 * H1, *W2 = b1v + H2, *b2v = W2 + H2, *xo = b2v + 1;
      for (int o = 0; o < H1; o++) b0v[o] = V[v++];
      for (int o = 0; o < H2; o++) b1v[o] = V[v++];
      b2v[0] = V[v++];
      for (int o = 0; o < H1; o++)
        for (int n = 0; n < D; n++) W0[o * D + n] = V[v++] * dr[(size_t)s * D + n];
      for (int o = 0; o < H2; o++)
        for (int n = 0; n < H1; n++) W1[o * H1 + n] = V[v++];
      for (int n = 0; n < H2; n++) W2[n] = V[v++];
      for (int n = 0; n < D; n++) xo[n] = (rmin[(size_t)s * D + n] / dr[(size_t)s * D + n]) * dr[(size_t)s * D + n];
      for (int k = 0; k < stride; k++) m->mlp[(size_t)s * stride + k] = (float)w[k];
    }
  }
  if ((rc = pack_mfma(m)) != MG_OK) goto done;
  rc = MG_OK;
done:
  free(V); free(rmin); free(rmax); free(dr);
  text_free(&t);
  if (rc != MG_OK) { mg_model_free(m); return rc; }
  *out = m;
  return MG_OK;
}

int mg_model_info(const mg_model *m, int *info, double *rc) {
  if (!m || !info) return mg_fail(MG_ERR_ARG, "mg_model_info: null argument");
  info[0] = m->nspc; info[1] = m->D; info[2] = m->H1; info[3] = m->H2;
  info[4] = m->nsym; info[5] = m->ng2; info[6] = m->ng4; info[7] = m->nw;
  for (int s = 0; s < 4; s++) info[8 + s] = s < m->nspc ? m->spcz[s] : 0;
  info[12] = m->pat;   /* G4 kernel specialisation: 0 generic, 1 43-function table, 2 22-function table */
  info[13] = m->uni;   /* 1: every species pair shares the G4 etas (one exp per triplet) */
  info[14] = m->mfma_ok;
  info[15] = 0;
  if (rc) *rc = m->rc_list;
  return MG_OK;
}

int mg_model_id(const mg_model *m, char *buf, int len) {
  if (!m || !buf || len < 1) return mg_fail(MG_ERR_ARG, "mg_model_id: bad argument");
  snprintf(buf, (size_t)len, "%s", m->id);
  return MG_OK;
}
