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
                                                int K, int lda, int ldb, int ldc, int relu, int accumulate,
                                                int kslice, int64_t c_slice) {
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  /* split K: slice z covers [z kslice, (z+1) kslice) and writes its partial
   * product to C + z c_slice */
  const int kb = blockIdx.z * kslice, ke = kb + kslice < K ? kb + kslice : K;
  C += blockIdx.z * c_slice;
  float acc[4][4] = {{0}};
  for (int k0 = kb; k0 < ke; k0 += TK) {
    for (int e = threadIdx.x; e < TK * TM; e += 256) {
      int kk = e / TM, mm = e % TM;
      int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < ke)
        v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[kk][mm] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      int kk = e / TN, nn = e % TN;
      int gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < ke)
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

/* out[i] = sum over s of part[s][i], s in order (deterministic split-K) */
__global__ void reduce_slices(const float *__restrict__ part, int S, int64_t count, float *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  float v = 0.f;
  for (int s = 0; s < S; s++)
    v += part[(int64_t)s * count + i];
  out[i] = v;
}

/* partial column sums of dz[rows, cols]: slice blockIdx.y of rowsl rows,
 * 64 columns per workgroup, 4 row lanes per column combined in LDS */
__global__ void __launch_bounds__(256) col_sum_part(const float *__restrict__ dz, float *part, int rows, int cols,
                                                    int rowsl) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), lane = threadIdx.x >> 6;
  const int r0 = blockIdx.y * rowsl, r1 = r0 + rowsl < rows ? r0 + rowsl : rows;
  float v = 0.f;
  if (c < cols)
    for (int r = r0 + lane; r < r1; r += 4)
      v += dz[(size_t)r * cols + c];
  red[lane][threadIdx.x & 63] = v;
  __syncthreads();
  if (lane == 0 && c < cols)
    part[(size_t)blockIdx.y * cols + c] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                                          red[3][threadIdx.x];
}

/* the minibatch from the resident training set: rows perm[0..B) of X, y */
__global__ void gather_batch(const float *__restrict__ X, const float *__restrict__ y,
                             const int32_t *__restrict__ perm, int B, int nin, float *xb, float *yb) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)B * nin)
    return;
  const int i = (int)(e / nin), k = (int)(e % nin);
  const int32_t j = perm[i];
  xb[e] = X[(size_t)j * nin + k];
  if (k == 0)
    yb[i] = y[j];
}

/* acc[0] += the batch loss (sum of the block partials in order) */
__global__ void loss_accumulate(const float *__restrict__ part, int nb, float *acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  float s = 0.f;
  for (int k = 0; k < nb; k++)
    s += part[k];
  acc[0] += s;
}

/* dZ *= (H > 0): ReLU derivative with the layer's output H */
__global__ void relu_mask(float *dz, const float *h, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(h[i] > 0.f))
    dz[i] = 0.f;
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
  float *part;           /* split-K and column-sum partials */
  int64_t part_cap;
  float *Xd, *yd;        /* the resident training set */
  int32_t *permd;        /* its sample order, epochs x n */
  float *eloss;          /* per-epoch loss sums */
  int64_t data_cap, y_cap, perm_cap, eloss_cap;
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

/* S > 1 splits K into S slices whose partial products go to C + z c_slice */
static void launch_gemm_s(ehd_mlp *m, int ta, int tb, const float *A, const float *B, float *C, const float *bias,
                          int M, int N, int K, int lda, int ldb, int ldc, int relu, int acc, int S, int64_t c_slice) {
  const int kslice = (K + S - 1) / S;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, S);
  if (!ta && tb)
    hipLaunchKernelGGL((gemm_f32<0, 1>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc, kslice, c_slice);
  else if (!ta && !tb)
    hipLaunchKernelGGL((gemm_f32<0, 0>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc, kslice, c_slice);
  else
    hipLaunchKernelGGL((gemm_f32<1, 0>), grid, dim3(256), 0, m->d->stream, A, B, C, bias, M, N, K, lda, ldb, ldc,
                       relu, acc, kslice, c_slice);
}

static void launch_gemm(ehd_mlp *m, int ta, int tb, const float *A, const float *B, float *C, const float *bias, int M,
                        int N, int K, int lda, int ldb, int ldc, int relu, int acc) {
  launch_gemm_s(m, ta, tb, A, B, C, bias, M, N, K, lda, ldb, ldc, relu, acc, 1, 0);
}

/* K slices for a weight-gradient GEMM of M x N over K = batch: enough
 * workgroups to fill the device (about 256), slices of at least 64 */
static int ksplit(int M, int N, int K) {
  int tiles = ((M + TM - 1) / TM) * ((N + TN - 1) / TN);
  int S = 256 / tiles;
  if (S > K / 64)
    S = K / 64;
  return S < 1 ? 1 : S;
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
  {
    /* partials of the split-K weight gradients and of the column sums */
    int64_t need = 0;
    for (int l = 0; l < MLP_LAYERS; l++) {
      int64_t w = (int64_t)ksplit(m->out_dim[l], m->in_dim[l], B) * m->out_dim[l] * m->in_dim[l];
      int64_t c = (int64_t)((B + 63) / 64) * m->out_dim[l];
      need = w > need ? w : need;
      need = c > need ? c : need;
    }
    if (need > m->part_cap) {
      if (m->part)
        MCHK(hipFree(m->part));
      m->part = NULL;
      m->part_cap = 0;
      MCHK(hipMalloc((void **)&m->part, sizeof(float) * need));
      m->part_cap = need;
    }
  }
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
  void *bufs[9] = {m->params, m->grads, m->adam_m, m->adam_v, m->part, m->Xd, m->yd, m->permd, m->eloss};
  for (int k = 0; k < 9; k++)
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

/* gradient of the mean BCE over the batch into m->grads.  The loss: into
 * *loss (synchronizing) when loss is given, else added to the device word
 * loss_acc when that is given.  Weight gradients dW = dZ^T X are split over
 * K (the batch) so that even the 19- and 1-wide layers fill the device; the
 * slices and the bias gradients' row slices are summed in a fixed order. */
static int backward(ehd_mlp *m, int B, float *loss, float *loss_acc) {
  int nb = (B + 255) / 256;
  hipLaunchKernelGGL(bce_grad, dim3(nb), dim3(256), 0, m->d->stream, m->act[MLP_LAYERS], m->labels,
                     m->dact[MLP_LAYERS], m->loss_part, B);
  if (loss_acc)
    hipLaunchKernelGGL(loss_accumulate, dim3(1), dim3(64), 0, m->d->stream, m->loss_part, nb, loss_acc);
  for (int l = MLP_LAYERS - 1; l >= 0; l--) {
    const int M = m->out_dim[l], N = m->in_dim[l];
    const int S = ksplit(M, N, B);
    const int64_t cnt = (int64_t)M * N;
    /* dW = dZ^T X */
    if (S > 1) {
      launch_gemm_s(m, 1, 0, m->dact[l + 1], m->act[l], m->part, NULL, M, N, B, M, N, N, 0, 0, S, cnt);
      hipLaunchKernelGGL(reduce_slices, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m->d->stream, m->part, S,
                         cnt, m->grads + m->woff[l]);
    } else
      launch_gemm(m, 1, 0, m->dact[l + 1], m->act[l], m->grads + m->woff[l], NULL, M, N, B, M, N, N, 0, 0);
    /* db = column sums of dZ, in row slices of 64 */
    {
      const int rs = (B + 63) / 64;
      hipLaunchKernelGGL(col_sum_part, dim3((unsigned)((M + 63) / 64), (unsigned)rs), dim3(256), 0, m->d->stream,
                         m->dact[l + 1], m->part, B, M, 64);
      hipLaunchKernelGGL(reduce_slices, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, m->d->stream, m->part, rs,
                         (int64_t)M, m->grads + m->boff[l]);
    }
    if (l > 0) {
      /* dX = dZ W, then the ReLU mask of the previous layer's output */
      launch_gemm(m, 0, 0, m->dact[l + 1], m->params + m->woff[l], m->dact[l], NULL, B, N, M, M, N, N, 0, 0);
      int64_t n = (int64_t)B * N;
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
  if ((rc = backward(m, B, loss, NULL)))
    return rc;
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(grad, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->d->stream));
  MCHK(hipStreamSynchronize(m->d->stream));
  return EH_OK;
}

/* grow a device buffer to hold count elements of size sz */
static int grow_dev(ehd_mlp *m, void **p, int64_t *cap, int64_t count, size_t sz) {
  if (count <= *cap)
    return EH_OK;
  if (*p)
    MCHK(hipFree(*p));
  *p = NULL;
  *cap = 0;
  MCHK(hipMalloc(p, sz * (size_t)(count > 0 ? count : 1)));
  *cap = count;
  return EH_OK;
}

/* minibatch Adam over n samples (X row-major [n, nin]) for the given epochs;
 * the sample order per epoch is the given permutation table (epochs x n).
 * The training set and the order are copied to the device once; each step
 * gathers its minibatch on the device, the losses are summed on the device,
 * and the host waits once, at the end. */
int ehd_mlp_train(ehd_mlp *m, const float *X, const float *y, int64_t n, const int64_t *perm, int32_t epochs,
                  int32_t B, float lr, float *losses) {
  int rc;
  const int64_t nsteps = n / B;
  MCHK(hipSetDevice(m->d->device));
  if (n > 0x7fffffff)
    return EH_ERR_INVALID;
  if ((rc = ensure_batch(m, B)))
    return rc;
  if ((rc = grow_dev(m, (void **)&m->Xd, &m->data_cap, n * m->nin, sizeof(float))) ||
      (rc = grow_dev(m, (void **)&m->yd, &m->y_cap, n, sizeof(float))) ||
      (rc = grow_dev(m, (void **)&m->permd, &m->perm_cap, (int64_t)epochs * n, sizeof(int32_t))) ||
      (rc = grow_dev(m, (void **)&m->eloss, &m->eloss_cap, epochs, sizeof(float))))
    return rc;
  {
    int32_t *p32 = (int32_t *)malloc(sizeof(int32_t) * (size_t)((int64_t)epochs * n > 0 ? (int64_t)epochs * n : 1));
    if (!p32)
      return EH_ERR_NOMEM;
    for (int64_t k = 0; k < (int64_t)epochs * n; k++)
      p32[k] = (int32_t)perm[k];
    MCHK(hipMemcpyAsync(m->Xd, X, sizeof(float) * (size_t)(n * m->nin), hipMemcpyHostToDevice, m->d->stream));
    MCHK(hipMemcpyAsync(m->yd, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->d->stream));
    MCHK(hipMemcpyAsync(m->permd, p32, sizeof(int32_t) * (size_t)((int64_t)epochs * n), hipMemcpyHostToDevice,
                        m->d->stream));
    MCHK(hipStreamSynchronize(m->d->stream));
    free(p32);
  }
  if (epochs > 0)
    hipLaunchKernelGGL(reduce_slices, dim3((unsigned)((epochs + 255) / 256)), dim3(256), 0, m->d->stream, m->eloss,
                       0, (int64_t)epochs, m->eloss); /* zero the per-epoch sums (a kernel, no memset) */
  for (int ep = 0; ep < epochs; ep++) {
    for (int64_t st = 0; st < nsteps; st++) {
      const int64_t e = (int64_t)B * m->nin;
      hipLaunchKernelGGL(gather_batch, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, m->d->stream, m->Xd, m->yd,
                         m->permd + (int64_t)ep * n + st * B, B, m->nin, m->act[0], m->labels);
      forward(m, B);
      if ((rc = backward(m, B, NULL, m->eloss + ep)))
        return rc;
      m->step++;
      float c1 = 1.f - powf(0.9f, (float)m->step), c2 = 1.f - powf(0.999f, (float)m->step);
      hipLaunchKernelGGL(adam_k, dim3((unsigned)((m->nparams + 255) / 256)), dim3(256), 0, m->d->stream, m->params,
                         m->grads, m->adam_m, m->adam_v, m->nparams, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    }
  }
  MCHK(hipGetLastError());
  if (losses && epochs > 0) {
    MCHK(hipMemcpyAsync(losses, m->eloss, sizeof(float) * epochs, hipMemcpyDeviceToHost, m->d->stream));
    MCHK(hipStreamSynchronize(m->d->stream));
    for (int ep = 0; ep < epochs; ep++)
      losses[ep] = nsteps ? losses[ep] / (float)nsteps : 0.f;
  }
  MCHK(hipStreamSynchronize(m->d->stream));
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
