/*
 * mg_sur.c -- host side of the screening surrogate (kernels in mg_sur.hip): parameter state, seeded
 * initialisation, normalisation, mini-batch Adam training, full-batch gradients and prediction, and the
 * per-structure descriptor features drawn from a context's last evaluation.
 *
 * Training data lives on the device for the duration of one training run; the parameters, the Adam moments and
 * the normalisation stay on the device between calls.  A run can be left in flight (mg_sur_train_begin) while the
 * host does other work; every other call on the surrogate waits for it first.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mg_host.h"

#define SUR_H 64
#define SUR_MAXIN 1024
#define SUR_W 4        /* samples per block of the gradient kernel */

int mgk_sur_features(int S, const mg_struct_dev *str, const int *spc, const float *G, int D, int nspc, float *feat,
                     hipStream_t st);
int mgk_sur_grad(int B, int nblk, int nin, const int *idx, const float *X, const float *Y, const float *p,
                 const float *norm, float *part, float *part_loss, hipStream_t st);
int mgk_sur_adam(int P, int nblk, const float *part, float *p, float *m, float *v, float *grad, float lr, float b1,
                 float b2, float eps, float c1, float c2, hipStream_t st);
int mgk_sur_predict(int n, int nin, const float *X, const float *p, const float *norm, float *out, hipStream_t st);

struct mg_sur {
  int device, nin, P;
  long step;                   /* Adam steps taken */
  hipStream_t st;
  float *d_p, *d_m, *d_v, *d_norm, *d_grad;
  float *d_part, *d_ploss;
  size_t cap_blk;
  float *d_X, *d_Y, *d_out;
  int *d_idx;
  size_t cap_n, cap_idx;
  /* a training run in flight (mg_sur_train_begin): its samples, sample indices and loss partials in pinned host
   * memory until mg_sur_train_end has waited for it */
  int pending, pend_n, pend_nblk;
  float *h_X, *h_Y, *h_ploss;
  int *h_perm;
  size_t cap_hx, cap_hperm, cap_hploss;
  float last_loss;
};

static int settle(mg_sur *s);

static unsigned long long splitmix64(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double unif(unsigned long long *s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

static double gauss(unsigned long long *s) {
  const double u = unif(s), v = unif(s);
  return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

int mg_sur_nparams(int nin) { return SUR_H * nin + SUR_H * SUR_H + 3 * SUR_H + 1; }

int mg_sur_create(int device, int nin, unsigned long long seed, mg_sur **out) {
  if (!out || nin < 1 || nin > SUR_MAXIN) return mg_fail(MG_ERR_ARG, "mg_sur_create: nin must be 1..%d", SUR_MAXIN);
  HIPCHK(hipSetDevice(device));
  mg_sur *s = (mg_sur *)calloc(1, sizeof *s);
  s->device = device;
  s->nin = nin;
  s->P = mg_sur_nparams(nin);
  int r;
  if ((r = mg_dev_alloc((void **)&s->d_p, sizeof(float) * s->P)) || (r = mg_dev_alloc((void **)&s->d_m, sizeof(float) * s->P)) ||
      (r = mg_dev_alloc((void **)&s->d_v, sizeof(float) * s->P)) || (r = mg_dev_alloc((void **)&s->d_grad, sizeof(float) * s->P)) ||
      (r = mg_dev_alloc((void **)&s->d_norm, sizeof(float) * (2 * nin + 2)))) {
    mg_sur_free(s);
    return r;
  }
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    mg_sur_free(s);
    return mg_fail(MG_ERR_HIP, "mg_sur_create: stream");
  }
  /* weights ~ N(0, 1/fan_in), biases 0 (as the synthetic potentials) */
  float *p = (float *)calloc(s->P, sizeof(float));
  unsigned long long st = seed;
  for (int k = 0; k < SUR_H * nin; k++) p[k] = (float)(gauss(&st) / sqrt((double)nin));
  float *W1 = p + SUR_H * nin + SUR_H;
  for (int k = 0; k < SUR_H * SUR_H; k++) W1[k] = (float)(gauss(&st) / sqrt((double)SUR_H));
  float *w2 = W1 + SUR_H * SUR_H + SUR_H;
  for (int k = 0; k < SUR_H; k++) w2[k] = (float)(gauss(&st) / sqrt((double)SUR_H));
  float *norm = (float *)malloc(sizeof(float) * (2 * nin + 2));
  for (int k = 0; k < nin; k++) { norm[k] = 0.0f; norm[nin + k] = 1.0f; }
  norm[2 * nin] = 0.0f;
  norm[2 * nin + 1] = 1.0f;
  hipError_t e1 = hipMemcpy(s->d_p, p, sizeof(float) * s->P, hipMemcpyHostToDevice);
  hipError_t e2 = hipMemcpy(s->d_norm, norm, sizeof(float) * (2 * nin + 2), hipMemcpyHostToDevice);
  hipError_t e3 = hipMemset(s->d_m, 0, sizeof(float) * s->P);
  hipError_t e4 = hipMemset(s->d_v, 0, sizeof(float) * s->P);
  free(p);
  free(norm);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
    mg_sur_free(s);
    return mg_fail(MG_ERR_HIP, "mg_sur_create: upload");
  }
  *out = s;
  return MG_OK;
}

void mg_sur_free(mg_sur *s) {
  if (!s) return;
  hipSetDevice(s->device);
  (void)settle(s);
  if (s->h_X) hipHostFree(s->h_X);
  if (s->h_perm) hipHostFree(s->h_perm);
  if (s->h_ploss) hipHostFree(s->h_ploss);
  float *bufs[] = {s->d_p, s->d_m, s->d_v, s->d_norm, s->d_grad, s->d_part, s->d_ploss, s->d_X, s->d_Y, s->d_out};
  for (size_t k = 0; k < sizeof bufs / sizeof bufs[0]; k++)
    if (bufs[k]) hipFree(bufs[k]);
  if (s->d_idx) hipFree(s->d_idx);
  if (s->st) hipStreamDestroy(s->st);
  free(s);
}

int mg_sur_get_params(mg_sur *s, float *p) {
  if (!s || !p) return mg_fail(MG_ERR_ARG, "mg_sur_get_params: bad argument");
  HIPCHK(hipSetDevice(s->device));
  { const int r_ = settle(s); if (r_) return r_; }
  HIPCHK(hipMemcpy(p, s->d_p, sizeof(float) * s->P, hipMemcpyDeviceToHost));
  return MG_OK;
}

int mg_sur_set_params(mg_sur *s, const float *p) {
  if (!s || !p) return mg_fail(MG_ERR_ARG, "mg_sur_set_params: bad argument");
  HIPCHK(hipSetDevice(s->device));
  { const int r_ = settle(s); if (r_) return r_; }
  HIPCHK(hipMemcpy(s->d_p, p, sizeof(float) * s->P, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(s->d_m, 0, sizeof(float) * s->P));
  HIPCHK(hipMemset(s->d_v, 0, sizeof(float) * s->P));
  s->step = 0;
  return MG_OK;
}

int mg_sur_set_norm(mg_sur *s, const float *mu, const float *sig, float ymu, float ysig) {
  if (!s || !mu || !sig || !(ysig > 0.0f)) return mg_fail(MG_ERR_ARG, "mg_sur_set_norm: bad argument");
  HIPCHK(hipSetDevice(s->device));
  { const int r_ = settle(s); if (r_) return r_; }
  for (int k = 0; k < s->nin; k++)
    if (!(sig[k] > 0.0f)) return mg_fail(MG_ERR_ARG, "mg_sur_set_norm: sig[%d] must be positive", k);
  float *norm = (float *)malloc(sizeof(float) * (2 * s->nin + 2));
  memcpy(norm, mu, sizeof(float) * s->nin);
  memcpy(norm + s->nin, sig, sizeof(float) * s->nin);
  norm[2 * s->nin] = ymu;
  norm[2 * s->nin + 1] = ysig;
  HIPCHK(hipSetDevice(s->device));
  hipError_t e = hipMemcpy(s->d_norm, norm, sizeof(float) * (2 * s->nin + 2), hipMemcpyHostToDevice);
  free(norm);
  if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "mg_sur_set_norm: %s", hipGetErrorString(e));
  return MG_OK;
}

#define SUR_MAXBLK 256  /* blocks of a full-data gradient pass (each takes more rounds as n grows) */
#define SUR_IDX_CAP (1 << 22)  /* sample indices uploaded at once: the permutations of that many epochs */

static int grad_blocks(int n) {
  const int b = (n + SUR_W - 1) / SUR_W;
  return b < SUR_MAXBLK ? b : SUR_MAXBLK;
}

/* device room for n samples, nidx sample indices and the block partials of batches of `batch` */
static int ensure_data(mg_sur *s, int n, size_t nidx, int batch) {
  int r;
  if ((size_t)n > s->cap_n) {
    size_t cap = (size_t)n + 256;
    s->cap_n = 0;
    if ((r = mg_dev_alloc((void **)&s->d_X, sizeof(float) * cap * s->nin)) ||
        (r = mg_dev_alloc((void **)&s->d_Y, sizeof(float) * cap)) ||
        (r = mg_dev_alloc((void **)&s->d_out, sizeof(float) * cap)))
      return r;
    s->cap_n = cap;
  }
  if (nidx > s->cap_idx) {
    s->cap_idx = 0;
    if ((r = mg_dev_alloc((void **)&s->d_idx, sizeof(int) * nidx))) return r;
    s->cap_idx = nidx;
  }
  size_t nblk = (size_t)(batch + SUR_W - 1) / SUR_W;
  if (nblk < SUR_MAXBLK) nblk = SUR_MAXBLK;
  if (nblk > s->cap_blk) {
    if ((r = mg_dev_alloc((void **)&s->d_part, sizeof(float) * nblk * s->P)) ||
        (r = mg_dev_alloc((void **)&s->d_ploss, sizeof(float) * nblk)))
      return r;
    s->cap_blk = nblk;
  }
  return MG_OK;
}

static int upload_xy(mg_sur *s, const float *X, const float *y, int n) {
  HIPCHK(hipMemcpyAsync(s->d_X, X, sizeof(float) * (size_t)n * s->nin, hipMemcpyHostToDevice, s->st));
  if (y) HIPCHK(hipMemcpyAsync(s->d_Y, y, sizeof(float) * n, hipMemcpyHostToDevice, s->st));
  return MG_OK;
}

/* pinned host buffer of at least `bytes` (grown on demand) */
static int ensure_pinned(void **p, size_t *cap, size_t bytes) {
  if (*cap >= bytes && *p) return MG_OK;
  if (*p) hipHostFree(*p);
  *p = NULL;
  *cap = 0;
  HIPCHK(hipHostMalloc(p, bytes, 0));
  *cap = bytes;
  return MG_OK;
}

/* mean squared error (normalised units) of the current parameters over the n samples now on the device: the
 * kernels and the copy of the per-block partials to pinned memory are enqueued; loss_finish sums them */
static int loss_enqueue(mg_sur *s, int n, float *grad) {
  const int nblk = grad_blocks(n);
  int r;
  if ((r = ensure_pinned((void **)&s->h_ploss, &s->cap_hploss, sizeof(float) * nblk))) return r;
  if (mgk_sur_grad(n, nblk, s->nin, NULL, s->d_X, s->d_Y, s->d_p, s->d_norm, s->d_part, s->d_ploss, s->st) ||
      mgk_sur_adam(s->P, nblk, s->d_part, s->d_p, s->d_m, s->d_v, s->d_grad, 0.0f, 0.f, 0.f, 0.f, 1.f, 1.f, s->st))
    return mg_fail(MG_ERR_HIP, "surrogate kernel launch failed");
  HIPCHK(hipMemcpyAsync(s->h_ploss, s->d_ploss, sizeof(float) * nblk, hipMemcpyDeviceToHost, s->st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, s->d_grad, sizeof(float) * s->P, hipMemcpyDeviceToHost, s->st));
  s->pend_n = n;
  s->pend_nblk = nblk;
  return MG_OK;
}

static int loss_finish(mg_sur *s, float *loss) {
  HIPCHK(hipStreamSynchronize(s->st));
  double sum = 0.0;
  for (int b = 0; b < s->pend_nblk; b++) sum += s->h_ploss[b];  /* fixed order */
  if (loss) *loss = (float)(sum / s->pend_n);
  return MG_OK;
}

static int full_loss(mg_sur *s, int n, float *loss, float *grad) {
  int r;
  if ((r = loss_enqueue(s, n, grad))) return r;
  return loss_finish(s, loss);
}

/* wait for a training run still in flight; its final loss is kept for mg_sur_train_end */
static int settle(mg_sur *s) {
  if (!s->pending) return MG_OK;
  s->pending = 0;
  return loss_finish(s, &s->last_loss);
}

int mg_sur_grad(mg_sur *s, const float *X, const float *y, int n, float *grad, float *loss) {
  if (!s || !X || !y || n < 1) return mg_fail(MG_ERR_ARG, "mg_sur_grad: bad argument");
  HIPCHK(hipSetDevice(s->device));
  int r;
  if ((r = settle(s))) return r;
  if ((r = ensure_data(s, n, 0, 1)) || (r = upload_xy(s, X, y, n))) return r;
  return full_loss(s, n, loss, grad);
}

int mg_sur_train_begin(mg_sur *s, const float *X, const float *y, int n, int epochs, int batch, float lr,
                       unsigned long long seed) {
  if (!s || !X || !y || n < 1 || epochs < 0 || batch < 1 || !(lr > 0.0f))
    return mg_fail(MG_ERR_ARG, "mg_sur_train: bad argument");
  HIPCHK(hipSetDevice(s->device));
  int r;
  if ((r = settle(s))) return r;
  if (batch > n) batch = n;
  /* the seeded permutations of a group of epochs go up in one copy; the steps then run without a host wait */
  int group = SUR_IDX_CAP / n;
  if (group < 1) group = 1;
  if (group > epochs) group = epochs > 0 ? epochs : 1;
  if ((r = ensure_data(s, n, (size_t)group * n, batch))) return r;
  /* the samples go through pinned memory owned by the surrogate, so the caller's arrays are free on return */
  if ((r = ensure_pinned((void **)&s->h_X, &s->cap_hx, sizeof(float) * (size_t)n * (s->nin + 1))) ||
      (r = ensure_pinned((void **)&s->h_perm, &s->cap_hperm, sizeof(int) * (size_t)group * n)))
    return r;
  s->h_Y = s->h_X + (size_t)n * s->nin;
  memcpy(s->h_X, X, sizeof(float) * (size_t)n * s->nin);
  memcpy(s->h_Y, y, sizeof(float) * n);
  if ((r = upload_xy(s, s->h_X, s->h_Y, n))) return r;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int *perm = s->h_perm;
  unsigned long long st = seed;
  for (int ep0 = 0; ep0 < epochs; ep0 += group) {
    const int ng = epochs - ep0 < group ? epochs - ep0 : group;
    /* the previous group's kernels read d_idx and its index copy reads perm: wait for both */
    if (ep0 > 0) HIPCHK(hipStreamSynchronize(s->st));
    for (int g = 0; g < ng; g++) {
      int *pg = perm + (size_t)g * n;
      for (int i = 0; i < n; i++) pg[i] = i;
      for (int i = n - 1; i > 0; i--) { /* Fisher-Yates, seeded */
        int j = (int)(splitmix64(&st) % (unsigned long long)(i + 1));
        int t = pg[i]; pg[i] = pg[j]; pg[j] = t;
      }
    }
    HIPCHK(hipMemcpyAsync(s->d_idx, perm, sizeof(int) * (size_t)ng * n, hipMemcpyHostToDevice, s->st));
    for (int g = 0; g < ng; g++)
      for (int b0 = 0; b0 < n; b0 += batch) {
        const int B = n - b0 < batch ? n - b0 : batch;
        const int nblk = (B + SUR_W - 1) / SUR_W;
        s->step++;
        const float c1 = 1.0f - powf(b1, (float)s->step), c2 = 1.0f - powf(b2, (float)s->step);
        if (mgk_sur_grad(B, nblk, s->nin, s->d_idx + (size_t)g * n + b0, s->d_X, s->d_Y, s->d_p, s->d_norm,
                         s->d_part, s->d_ploss, s->st) ||
            mgk_sur_adam(s->P, nblk, s->d_part, s->d_p, s->d_m, s->d_v, NULL, lr, b1, b2, eps, c1, c2, s->st))
          return mg_fail(MG_ERR_HIP, "surrogate kernel launch failed");
      }
  }
  if ((r = loss_enqueue(s, n, NULL))) return r;
  s->pending = 1;
  return MG_OK;
}

int mg_sur_train_end(mg_sur *s, float *loss) {
  if (!s) return mg_fail(MG_ERR_ARG, "mg_sur_train_end: bad argument");
  HIPCHK(hipSetDevice(s->device));
  int r;
  if ((r = settle(s))) return r;
  if (loss) *loss = s->last_loss;
  return MG_OK;
}

int mg_sur_train(mg_sur *s, const float *X, const float *y, int n, int epochs, int batch, float lr,
                 unsigned long long seed, float *loss) {
  int r;
  if ((r = mg_sur_train_begin(s, X, y, n, epochs, batch, lr, seed))) return r;
  return mg_sur_train_end(s, loss);
}

int mg_sur_predict(mg_sur *s, const float *X, int n, float *y) {
  if (!s || !X || !y || n < 0) return mg_fail(MG_ERR_ARG, "mg_sur_predict: bad argument");
  if (n == 0) return MG_OK;
  HIPCHK(hipSetDevice(s->device));
  int r;
  /* a training run in flight is waited for first: the prediction uses its parameters and its sample buffer */
  if ((r = settle(s))) return r;
  if ((r = ensure_data(s, n, 0, 1)) || (r = upload_xy(s, X, NULL, n))) return r;
  if (mgk_sur_predict(n, s->nin, s->d_X, s->d_p, s->d_norm, s->d_out, s->st))
    return mg_fail(MG_ERR_HIP, "surrogate kernel launch failed");
  HIPCHK(hipMemcpyAsync(y, s->d_out, sizeof(float) * n, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return MG_OK;
}

int mg_features(mg_ctx *c, float *feat) {
  if (!c || !feat) return mg_fail(MG_ERR_ARG, "mg_features: bad argument");
  if (!c->have_g) return mg_fail(MG_ERR_ARG, "mg_features: no evaluation yet");
  const mg_model *m = c->m;
  HIPCHK(hipSetDevice(c->device));
  { int rs_ = mgh_settle(c); if (rs_ != MG_OK) return rs_; }
  const int nf = m->nspc * m->D;
  float *d_feat = NULL;
  int r;
  if ((r = mg_dev_alloc((void **)&d_feat, sizeof(float) * (size_t)c->S * nf))) return r;
  if (mgk_sur_features(c->S, c->d_str, c->d_spc, c->d_G, m->D, m->nspc, d_feat, c->st)) {
    hipFree(d_feat);
    return mg_fail(MG_ERR_HIP, "feature kernel launch failed");
  }
  hipError_t e = hipMemcpyAsync(feat, d_feat, sizeof(float) * (size_t)c->S * nf, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  hipFree(d_feat);
  if (e != hipSuccess) return mg_fail(MG_ERR_HIP, "mg_features: %s", hipGetErrorString(e));
  return MG_OK;
}
