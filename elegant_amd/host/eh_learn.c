/*
 * eh_learn.c - public C ABI of the dynamic-aperture surrogate network
 * (kernels in elegant_amd/csrc/eh_mlp.hip).  The reference has no such
 * component; its aperture search tracks every point (aperture_search.c:1064-1120).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "eh_host.h"

typedef struct ehd_mlp ehd_mlp;
int ehd_mlp_create(ehd_dev *d, int32_t nin, int32_t hidden, const float *init, ehd_mlp **out);
int64_t ehd_mlp_nparams(const ehd_mlp *m);
void ehd_mlp_destroy(ehd_mlp *m);
int ehd_mlp_grad(ehd_mlp *m, const float *X, const float *y, int32_t B, float *grad, float *loss);
int ehd_mlp_train(ehd_mlp *m, const float *X, const float *y, int64_t n, const int64_t *perm, int32_t epochs, int32_t B,
                  float lr, float *losses);
int ehd_mlp_predict(ehd_mlp *m, const float *X, int64_t n, float *prob);
int ehd_mlp_params(ehd_mlp *m, float *buf, int set);
ehd_dev *eh_ctx_dev(eh_ctx *c);

struct eh_mlp {
  ehd_mlp *m;
  int32_t nin, hidden;
};

static uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double uni(uint64_t *s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

int eh_mlp_create(eh_ctx *ctx, int32_t n_in, int32_t hidden, uint64_t seed, eh_mlp **out) {
  eh_mlp *w;
  float *init;
  int64_t n = 0, off = 0;
  int rc;
  int dims[5];
  if (!ctx || !out || n_in < 1 || hidden < 1)
    return EH_ERR_INVALID;
  dims[0] = n_in;
  dims[1] = dims[2] = dims[3] = hidden;
  dims[4] = 1;
  for (int l = 0; l < 4; l++)
    n += (int64_t)dims[l] * dims[l + 1] + dims[l + 1];
  init = (float *)malloc(sizeof(float) * n);
  w = (eh_mlp *)calloc(1, sizeof(eh_mlp));
  if (!init || !w) {
    free(init);
    free(w);
    return EH_ERR_NOMEM;
  }
  /* He-uniform weights U(-sqrt(6/fan_in), +), zero biases */
  for (int l = 0; l < 4; l++) {
    double a = sqrt(6.0 / dims[l]);
    for (int64_t k = 0; k < (int64_t)dims[l] * dims[l + 1]; k++)
      init[off++] = (float)((2 * uni(&seed) - 1) * a);
    for (int k = 0; k < dims[l + 1]; k++)
      init[off++] = 0.f;
  }
  rc = ehd_mlp_create(eh_ctx_dev(ctx), n_in, hidden, init, &w->m);
  free(init);
  if (rc) {
    free(w);
    return rc;
  }
  w->nin = n_in;
  w->hidden = hidden;
  *out = w;
  return EH_OK;
}

int eh_mlp_destroy(eh_mlp *w) {
  if (!w)
    return EH_OK;
  ehd_mlp_destroy(w->m);
  free(w);
  return EH_OK;
}

int64_t eh_mlp_num_params(const eh_mlp *w) { return w ? ehd_mlp_nparams(w->m) : EH_ERR_INVALID; }

int eh_mlp_get_params(eh_mlp *w, float *p) { return (w && p) ? ehd_mlp_params(w->m, p, 0) : EH_ERR_INVALID; }

int eh_mlp_set_params(eh_mlp *w, const float *p) {
  return (w && p) ? ehd_mlp_params(w->m, (float *)p, 1) : EH_ERR_INVALID;
}

int eh_mlp_gradient(eh_mlp *w, const float *X, const float *y, int32_t n, float *grad, float *loss) {
  if (!w || !X || !y || n < 1 || !grad)
    return EH_ERR_INVALID;
  return ehd_mlp_grad(w->m, X, y, n, grad, loss);
}

int eh_mlp_train(eh_mlp *w, const float *X, const float *y, int64_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float *losses) {
  int64_t *perm;
  int rc;
  if (!w || !X || !y || n < batch || epochs < 1 || batch < 1)
    return EH_ERR_INVALID;
  perm = (int64_t *)malloc(sizeof(int64_t) * n * epochs);
  if (!perm)
    return EH_ERR_NOMEM;
  for (int e = 0; e < epochs; e++) {
    int64_t *p = perm + (int64_t)e * n;
    for (int64_t i = 0; i < n; i++)
      p[i] = i;
    for (int64_t i = n - 1; i > 0; i--) { /* Fisher-Yates */
      int64_t j = (int64_t)(uni(&seed) * (i + 1));
      int64_t t = p[i];
      p[i] = p[j];
      p[j] = t;
    }
  }
  rc = ehd_mlp_train(w->m, X, y, n, perm, epochs, batch, lr, losses);
  free(perm);
  return rc;
}

int eh_mlp_predict(eh_mlp *w, const float *X, int64_t n, float *prob) {
  if (!w || !X || !prob || n < 0)
    return EH_ERR_INVALID;
  return ehd_mlp_predict(w->m, X, n, prob);
}
