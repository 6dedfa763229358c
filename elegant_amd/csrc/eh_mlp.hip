/*
 * ehd_mlp.hip - a small multilayer perceptron trained and evaluated on the GPU.
 *
 * It proposes dynamic-aperture boundaries: inputs are a test particle's start
 * (x0, y0, delta) and the lattice variant's error summary, the output is the
 * probability that the particle survives.  FP32 throughout; FP64 tracking of
 * the proposed points decides the answer.
 *
 * Kernels: one LDS-tiled FP32 GEMM (64x64 output tile per 256-thread
 * workgroup, 4x4 per thread, K staged in 16-wide slabs) with transpose
 * variants for the forward pass (Y = X W^T + b), the data gradient
 * (dX = dY W) and the weight gradient (dW = dY^T X); fused bias/ReLU
 * epilogues; the loss gradient of binary cross-entropy on logits; Adam.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/elegant_hip.h"
#include "eh_device.h"
#include "ehd_internal.h"

#define TM 64
#define TN 64
#define TK 16

/* C[M,N] = op(A) op(B) (+ bias[N]) (ReLU), row-major.
 * TA = 0: A is [M,K] (lda);  TA = 1: A is [K,M] (lda), use A^T.
 * TB = 0: B is [K,N] (ldb);  TB = 1: B is [N,K] (ldb), use B^T. */
template <int TA, int TB>
__global__ void __launch_bounds__(256) gemm_f32(const float *__restrict__ A, const float *__restrict__ B,
                                                float *__restrict__ C, const float *__restrict__ bias, int M, int N,
                                                int K, int lda, int ldb, int ldc, int relu, int accumulate) {
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float acc[4][4] = {{0}};
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int e = threadIdx.x; e < TK * TM; e += 256) {
      int kk = e / TM, mm = e % TM;
      int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < K)
        v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[kk][mm] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      int kk = e / TN, nn = e % TN;
      int gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < K)
        v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++)
        b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int gm = m0 + ty * 4 + i;
    if (gm >= M)
      continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int gn = n0 + tx * 4 + j;
      if (gn >= N)
        continue;
      float v = acc[i][j];
      if (bias)
        v += bias[gn];
      if (relu)
        v = v > 0.f ? v : 0.f;
      float *dst = &C[(size_t)gm * ldc + gn];
      *dst = accumulate ? *dst + v : v;
    }
  }
}

/* dZ *= (H > 0): ReLU derivative with the layer's output H */
__global__ void relu_mask(float *dz, const float *h, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(h[i] > 0.f))
    dz[i] = 0.f;
}

/* column sums: db[j] = sum_i dz[i, j] (one thread per column) */
__global__ void col_sum(const float *dz, float *db, int rows, int cols) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols)
    return;
  float s = 0.f;
  for (int i = 0; i < rows; i++)
    s += dz[(size_t)i * cols + j];
  db[j] = s;
}

/* binary cross-entropy on logits, mean over the batch: loss partials and dz = (sigmoid(z) - y)/B */
__global__ void bce_grad(const float *z, const float *y, float *dz, float *loss_part, int B) {
  __shared__ float red[256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < B) {
    float zi = z[i], yi = y[i];
    float sp = zi > 0 ? zi + log1pf(expf(-zi)) : log1pf(expf(zi)); /* softplus */
    l = sp - yi * zi;
    float sg = 1.f / (1.f + expf(-zi));
    dz[i] = (sg - yi) / (float)B;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    loss_part[blockIdx.x] = red[0] / (float)B;
}

__global__ void sigmoid_k(const float *z, float *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 1.f / (1.f + expf(-z[i]));
}

__global__ void adam_k(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* ------------------------------------------------------------ host side */

#define MLP_LAYERS 4

struct ehd_mlp {
  ehd_dev *d;
  int32_t nin, hidden;
  int32_t in_dim[MLP_LAYERS], out_dim[MLP_LAYERS];
  int64_t woff[MLP_LAYERS], boff[MLP_LAYERS], nparams;
  float *params, *grads, *adam_m, *adam_v;
  float *act[MLP_LAYERS + 1]; /* act[0] = input, act[l+1] = output of layer l */
  float *dact[MLP_LAYERS + 1];
  float *labels, *loss_part;
  int32_t cap;
  int64_t step;
  char msg[256];
};

extern "C" {

static int mlp_fail(ehd_mlp *m, hipError_t e, const char *what) {
  snprintf(m->d->msg, sizeof(m->d->msg), "%s: %s", what, hipGetErrorString(e));
  return EH_ERR_HIP;
}

#define MCHK(call)                                                                                                     \
  do {                                                                                                                 \
    hipError_t e_ = (call);                                                                                            \
    if (e_ != hipSuccess)                                                                                              \
      return mlp_fail(m, e_, #call);                                                                                   \
  } while (0)

static void launch_gemm(ehd_mlp *m, int ta, int tb, const float *A, const float *B, float *C, const float *bias, int M,
                        int N, int K, int lda, int ldb, int ldc, int relu, int acc) {
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  if (!ta && tb)
    hipLaunchKernelGGL((gemm_f32<0, 1>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc);
  else if (!ta && !tb)
    hipLaunchKernelGGL((gemm_f32<0, 0>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc);
  else
    hipLaunchKernelGGL((gemm_f32<1, 0>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc);
}

static int ensure_batch(ehd_mlp *m, int32_t B) {
  if (B <= m->cap)
    return EH_OK;
  for (int l = 0; l <= MLP_LAYERS; l++) {
    if (m->act[l])
      MCHK(hipFree(m->act[l]));
    if (m->dact[l])
      MCHK(hipFree(m->dact[l]));
    m->act[l] = m->dact[l] = NULL;
  }
  if (m->labels)
    MCHK(hipFree(m->labels));
  if (m->loss_part)
    MCHK(hipFree(m->loss_part));
  for (int l = 0; l <= MLP_LAYERS; l++) {
    int w = l == 0 ? m->nin : m->out_dim[l - 1];
    MCHK(hipMalloc((void **)&m->act[l], sizeof(float) * (size_t)B * w));
    MCHK(hipMalloc((void **)&m->dact[l], sizeof(float) * (size_t)B * w));
  }
  MCHK(hipMalloc((void **)&m->labels, sizeof(float) * B));
  MCHK(hipMalloc((void **)&m->loss_part, sizeof(float) * ((B + 255) / 256)));
  m->cap = B;
  return EH_OK;
}

void ehd_mlp_destroy(ehd_mlp *m);

int ehd_mlp_create(ehd_dev *d, int32_t nin, int32_t hidden, const float *init, ehd_mlp **out) {
  ehd_mlp *m = (ehd_mlp *)calloc(1, sizeof(ehd_mlp));
  if (!m)
    return EH_ERR_NOMEM;
  m->d = d;
  m->nin = nin;
  m->hidden = hidden;
  int64_t off = 0;
  for (int l = 0; l < MLP_LAYERS; l++) {
    m->in_dim[l] = l == 0 ? nin : hidden;
    m->out_dim[l] = l == MLP_LAYERS - 1 ? 1 : hidden;
    m->woff[l] = off;
    off += (int64_t)m->in_dim[l] * m->out_dim[l];
    m->boff[l] = off;
    off += m->out_dim[l];
  }
  m->nparams = off;
  if (hipSetDevice(d->device) != hipSuccess || hipMalloc((void **)&m->params, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void **)&m->grads, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void **)&m->adam_m, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void **)&m->adam_v, sizeof(float) * off) != hipSuccess) {
    ehd_mlp_destroy(m);
    return EH_ERR_NOMEM;
  }
  if (hipMemcpy(m->params, init, sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->adam_m, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(m->adam_v, 0, sizeof(float) * off) != hipSuccess) {
    ehd_mlp_destroy(m);
    return EH_ERR_HIP;
  }
  *out = m;
  return EH_OK;
}

int64_t ehd_mlp_nparams(const ehd_mlp *m) { return m->nparams; }

/* teardown: errors are ignored, there is nothing left to undo */
void ehd_mlp_destroy(ehd_mlp *m) {
  if (!m)
    return;
  (void)hipSetDevice(m->d->device);
  (void)hipStreamSynchronize(m->d->stream);
  void *bufs[4] = {m->params, m->grads, m->adam_m, m->adam_v};
  for (int k = 0; k < 4; k++)
    if (bufs[k])
      (void)hipFree(bufs[k]);
  for (int l = 0; l <= MLP_LAYERS; l++) {
    if (m->act[l])
      (void)hipFree(m->act[l]);
    if (m->dact[l])
      (void)hipFree(m->dact[l]);
  }
  if (m->labels)
    (void)hipFree(m->labels);
  if (m->loss_part)
    (void)hipFree(m->loss_part);
  free(m);
}

static void forward(ehd_mlp *m, int B) {
  for (int l = 0; l < MLP_LAYERS; l++) {
    int last = l == MLP_LAYERS - 1;
    launch_gemm(m, 0, 1, m->act[l], m->params + m->woff[l], m->act[l + 1], m->params + m->boff[l], B,
                m->out_dim[l], m->in_dim[l], m->in_dim[l], m->in_dim[l], m->out_dim[l], last ? 0 : 1, 0);
  }
}

/* gradient of the mean BCE over the batch into m->grads; returns the loss */
static int backward(ehd_mlp *m, int B, float *loss) {
  int nb = (B + 255) / 256;
  hipLaunchKernelGGL(bce_grad, dim3(nb), dim3(256), 0, m->d->stream, m->act[MLP_LAYERS], m->labels,
                     m->dact[MLP_LAYERS], m->loss_part, B);
  for (int l = MLP_LAYERS - 1; l >= 0; l--) {
    /* dW = dZ^T X, db = colsum(dZ) */
    launch_gemm(m, 1, 0, m->dact[l + 1], m->act[l], m->grads + m->woff[l], NULL, m->out_dim[l], m->in_dim[l], B,
                m->out_dim[l], m->in_dim[l], m->in_dim[l], 0, 0);
    hipLaunchKernelGGL(col_sum, dim3((m->out_dim[l] + 255) / 256), dim3(256), 0, m->d->stream, m->dact[l + 1],
                       m->grads + m->boff[l], B, m->out_dim[l]);
    if (l > 0) {
      /* dX = dZ W, then the ReLU mask of the previous layer's output */
      launch_gemm(m, 0, 0, m->dact[l + 1], m->params + m->woff[l], m->dact[l], NULL, B, m->in_dim[l],
                  m->out_dim[l], m->out_dim[l], m->in_dim[l], m->in_dim[l], 0, 0);
      int64_t n = (int64_t)B * m->in_dim[l];
      hipLaunchKernelGGL(relu_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->d->stream, m->dact[l],
                         m->act[l], n);
    }
  }
  if (loss) {
    float *h = (float *)malloc(sizeof(float) * nb);
    if (!h)
      return EH_ERR_NOMEM;
    MCHK(hipMemcpyAsync(h, m->loss_part, sizeof(float) * nb, hipMemcpyDeviceToHost, m->d->stream));
    MCHK(hipStreamSynchronize(m->d->stream));
    double s = 0;
    for (int k = 0; k < nb; k++)
      s += h[k];
    *loss = (float)s;
    free(h);
  }
  return EH_OK;
}

int ehd_mlp_grad(ehd_mlp *m, const float *X, const float *y, int32_t B, float *grad, float *loss) {
  int rc;
  MCHK(hipSetDevice(m->d->device));
  if ((rc = ensure_batch(m, B)))
    return rc;
  MCHK(hipMemcpyAsync(m->act[0], X, sizeof(float) * (size_t)B * m->nin, hipMemcpyHostToDevice, m->d->stream));
  MCHK(hipMemcpyAsync(m->labels, y, sizeof(float) * B, hipMemcpyHostToDevice, m->d->stream));
  forward(m, B);
  if ((rc = backward(m, B, loss)))
    return rc;
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(grad, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->d->stream));
  MCHK(hipStreamSynchronize(m->d->stream));
  return EH_OK;
}

/* minibatch Adam over n samples (X row-major [n, nin]) for the given epochs;
 * the sample order per epoch is the given permutation table (epochs x n) */
int ehd_mlp_train(ehd_mlp *m, const float *X, const float *y, int64_t n, const int64_t *perm, int32_t epochs,
                  int32_t B, float lr, float *losses) {
  int rc;
  float *xb, *yb;
  MCHK(hipSetDevice(m->d->device));
  if ((rc = ensure_batch(m, B)))
    return rc;
  xb = (float *)malloc(sizeof(float) * (size_t)B * m->nin);
  yb = (float *)malloc(sizeof(float) * B);
  if (!xb || !yb) {
    free(xb);
    free(yb);
    return EH_ERR_NOMEM;
  }
  for (int ep = 0; ep < epochs; ep++) {
    double tot = 0;
    int nbatch = 0;
    for (int64_t s0 = 0; s0 + B <= n; s0 += B) {
      for (int i = 0; i < B; i++) {
        int64_t j = perm[(int64_t)ep * n + s0 + i];
        memcpy(xb + (size_t)i * m->nin, X + (size_t)j * m->nin, sizeof(float) * m->nin);
        yb[i] = y[j];
      }
      MCHK(hipMemcpyAsync(m->act[0], xb, sizeof(float) * (size_t)B * m->nin, hipMemcpyHostToDevice, m->d->stream));
      MCHK(hipMemcpyAsync(m->labels, yb, sizeof(float) * B, hipMemcpyHostToDevice, m->d->stream));
      forward(m, B);
      float l = 0;
      if ((rc = backward(m, B, &l))) {
        free(xb);
        free(yb);
        return rc;
      }
      m->step++;
      float c1 = 1.f - powf(0.9f, (float)m->step), c2 = 1.f - powf(0.999f, (float)m->step);
      hipLaunchKernelGGL(adam_k, dim3((unsigned)((m->nparams + 255) / 256)), dim3(256), 0, m->d->stream, m->params,
                         m->grads, m->adam_m, m->adam_v, m->nparams, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
      MCHK(hipGetLastError());
      tot += l;
      nbatch++;
    }
    if (losses)
      losses[ep] = nbatch ? (float)(tot / nbatch) : 0.f;
  }
  MCHK(hipStreamSynchronize(m->d->stream));
  free(xb);
  free(yb);
  return EH_OK;
}

int ehd_mlp_predict(ehd_mlp *m, const float *X, int64_t n, float *prob) {
  int rc;
  const int32_t B = 8192;
  MCHK(hipSetDevice(m->d->device));
  if ((rc = ensure_batch(m, B)))
    return rc;
  for (int64_t s0 = 0; s0 < n; s0 += B) {
    int32_t b = (int32_t)(n - s0 < B ? n - s0 : B);
    MCHK(hipMemcpyAsync(m->act[0], X + (size_t)s0 * m->nin, sizeof(float) * (size_t)b * m->nin,
                        hipMemcpyHostToDevice, m->d->stream));
    forward(m, b);
    hipLaunchKernelGGL(sigmoid_k, dim3((b + 255) / 256), dim3(256), 0, m->d->stream, m->act[MLP_LAYERS],
                       m->dact[MLP_LAYERS], (int64_t)b);
    MCHK(hipMemcpyAsync(prob + s0, m->dact[MLP_LAYERS], sizeof(float) * b, hipMemcpyDeviceToHost, m->d->stream));
  }
  MCHK(hipStreamSynchronize(m->d->stream));
  return EH_OK;
}

int ehd_mlp_params(ehd_mlp *m, float *buf, int set) {
  MCHK(hipSetDevice(m->d->device));
  if (set)
    MCHK(hipMemcpy(m->params, buf, sizeof(float) * m->nparams, hipMemcpyHostToDevice));
  else
    MCHK(hipMemcpy(buf, m->params, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
  return EH_OK;
}
}
