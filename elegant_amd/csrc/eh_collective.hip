/*
 * eh_collective.hip - the WAKE and ZLONGIT steps (bin -> convolve/FFT -> kick).
 *
 * The reference computes each particle's arrival time, reduces min/max (and
 * for ZLONGIT the mean), histograms the times, turns the histogram into a
 * voltage per bin and kicks every particle with its bin's voltage
 * (wake.c:23-303, zlongit.c:56-557).  Here:
 *  1. time_minmax: t = s sqrt(P^2+1)/(cP) for alive particles, exact min/max
 *     with ordered-integer atomics, alive count;
 *  2. (ZLONGIT) time_fsum + fsum_final: the order-independent fixed-point
 *     mean of eh_fixedsum.h;
 *  3. histogram: integer counts in LDS per workgroup, merged with integer
 *     atomics (bit-exact whatever the order);
 *  4. WAKE: wake_convolve, one lane per output bin summing in the
 *     reference's index order; ZLONGIT: zlongit_fft, one workgroup doing the
 *     packed real FFT, V = Z I, inverse FFT in LDS with the same butterflies
 *     as the CPU checker;
 *  5. kick: add_to_particle_energy with the bin voltage, then the aperture
 *     and NaN checks do_tracking applies after the element.
 * Reading (s, delta) for 1-3 and (x', y', s, delta) in 5 is the HBM traffic.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include "../../include/elegant_hip.h"
#include "eh_aperture.h"
#include "eh_consts.h"
#include "eh_device.h"
#include "eh_fixedsum.h"
#include "eh_program.h"
#include "ehd_internal.h"

#pragma clang fp contract(off)

#define DEV static __device__ __forceinline__
#define CB 256 /* threads per workgroup for particle sweeps */
#define SWEEP_U 4 /* particles per lane per step of the grid-stride sweeps */

DEV double sqd(double v) { return v * v; }

/* computeTimeCoordinatesOnly (trwake.c:519-527) */
DEV double arrival_time(double s, double dp, double Po) {
  double P = Po * (dp + 1);
  return s * __builtin_sqrt(sqd(P) + 1) / (EH_C_MKS * P);
}

DEV unsigned long long order_key(double v) {
  unsigned long long b = (unsigned long long)fs_bits(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

static __host__ __device__ double key_value(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffULL) : ~k;
  double d;
  memcpy(&d, &b, 8);
  return d;
}

/* The step's decisions live in device memory, so the whole step is queued
 * without waiting for the host: the reductions land in red[], the planning
 * kernels turn them into the bin layout, and every later kernel reads it. */
typedef struct {
  unsigned long long red[4]; /* min key, max key, alive count, unused */
  double tmin, tmax, tmean;
  int32_t nb;                /* bins of this step; <= 0: nothing to bin or kick */
  int32_t e;                 /* fixed-point exponent of the mean */
  int32_t err;               /* EH_ERR_LONGBEAM when the bunch does not fit */
  int32_t pad[3];
} coll_plan;

/* after the min/max pass: the decisions of wake.c:120-182 and
 * zlongit.c:195-236 that do not need the mean */
__global__ void __launch_bounds__(CB) coll_plan_minmax(coll_plan *plan, const unsigned long long *part,
                                                       int64_t nparts, int32_t kind, int32_t n_fixed, int32_t zl_nb,
                                                       double dt, double wake_span, int32_t allow_long,
                                                       int32_t nb_cap, unsigned int *hist, int32_t nclear,
                                                       int32_t mean_given, double mean) {
  __shared__ unsigned long long smin[CB], smax[CB], scnt[CB];
  unsigned long long kmin = ~0ULL, kmax = 0, cnt = 0;
  /* the step starts here: clear the histogram (a kernel rather than a
   * memset, so that the step can be captured in a graph) */
  for (int k = threadIdx.x; k < nclear; k += CB)
    hist[k] = 0;
  if (threadIdx.x == 0) {
    plan->nb = 0;
    plan->err = 0;
  }
  for (int64_t k = threadIdx.x; k < nparts; k += CB) {
    unsigned long long a = part[3 * k], b = part[3 * k + 1];
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
    cnt += part[3 * k + 2];
  }
  smin[threadIdx.x] = kmin;
  smax[threadIdx.x] = kmax;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      unsigned long long a = smin[threadIdx.x + w], b = smax[threadIdx.x + w];
      smin[threadIdx.x] = a < smin[threadIdx.x] ? a : smin[threadIdx.x];
      smax[threadIdx.x] = b > smax[threadIdx.x] ? b : smax[threadIdx.x];
      scnt[threadIdx.x] += scnt[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0)
    return;
  coll_plan *p = plan;
  p->red[0] = smin[0];
  p->red[1] = smax[0];
  p->red[2] = scnt[0];
  if (kind < 0)
    return; /* the time span only */
  if (p->red[2] == 0) {
    p->nb = 0;
    return;
  }
  double tmin = key_value(p->red[0]), tmax = key_value(p->red[1]);
  p->tmin = tmin;
  p->tmax = tmax;
  if (kind == EH_OP_ZLONGIT || n_fixed) {
    double am = __builtin_fabs(tmin) > __builtin_fabs(tmax) ? __builtin_fabs(tmin) : __builtin_fabs(tmax);
    p->e = eh_fsum_exponent(am);
  }
  if (kind == EH_OP_WAKE) {
    if ((tmax - tmin) > wake_span && !allow_long) {
      p->err = EH_ERR_LONGBEAM;
      return;
    }
    if (n_fixed) {
      p->nb = n_fixed; /* tmin set from the mean: the fixed-point one, or the host's */
      if (mean_given) {
        p->tmean = mean;
        p->tmin = mean - dt * p->nb / 2.0;
      }
    } else {
      int32_t nb = (int32_t)(int64_t)((tmax - tmin) / dt + 3);
      p->tmin = tmin - dt;
      p->nb = nb;
    }
  } else {
    p->nb = zl_nb;
    if ((tmax - tmin) * 2 > zl_nb * dt && !allow_long) {
      p->err = EH_ERR_LONGBEAM;
      return;
    }
  }
  if (p->nb > nb_cap)
    p->err = EH_ERR_LONGBEAM;
}

/* per-workgroup (min key, max key, alive count) triples in red[3 k ..];
 * grid-stride, so there are at most a thousand of them for the planning
 * kernel to combine (no global atomics) */
/* a particle takes part in the step when it is alive (lp < 0, or no lp)
 * and, for one bunch of a bunched beam, carries that bunch index */
DEV bool in_step(const int32_t *__restrict__ lp, const int32_t *__restrict__ bunch, int32_t target, int64_t i) {
  return (!lp || lp[i] < 0) && (!bunch || bunch[i] == target);
}

__global__ void __launch_bounds__(CB) time_minmax(const double *__restrict__ s, const double *__restrict__ dp,
                                                  const int32_t *__restrict__ lp, int64_t n, double Po,
                                                  unsigned long long *red, const int32_t *__restrict__ bunch,
                                                  int32_t target) {
  __shared__ unsigned long long smin[CB / 64], smax[CB / 64], scnt[CB / 64];
  unsigned long long kmin = ~0ULL, kmax = 0, cnt = 0;
  const int64_t stride = (int64_t)gridDim.x * CB;
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  /* SWEEP_U particles per lane per step, their loads issued together */
  for (; i + (SWEEP_U - 1) * stride < n; i += SWEEP_U * stride) {
    double S[SWEEP_U], D[SWEEP_U];
    bool live[SWEEP_U];
#pragma unroll
    for (int u = 0; u < SWEEP_U; u++) {
      live[u] = in_step(lp, bunch, target, i + u * stride);
      S[u] = s[i + u * stride];
      D[u] = dp[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < SWEEP_U; u++)
      if (live[u]) {
        unsigned long long k = order_key(arrival_time(S[u], D[u], Po));
        kmin = k < kmin ? k : kmin;
        kmax = k > kmax ? k : kmax;
        cnt++;
      }
  }
  for (; i < n; i += stride) {
    if (in_step(lp, bunch, target, i)) {
      unsigned long long k = order_key(arrival_time(s[i], dp[i], Po));
      kmin = k < kmin ? k : kmin;
      kmax = k > kmax ? k : kmax;
      cnt++;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long a = __shfl_xor(kmin, off), b = __shfl_xor(kmax, off), c = __shfl_xor(cnt, off);
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
    cnt += c;
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smin[w] = kmin;
    smax[w] = kmax;
    scnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < CB / 64; k++) {
      kmin = smin[k] < kmin ? smin[k] : kmin;
      kmax = smax[k] > kmax ? smax[k] : kmax;
      cnt += scnt[k];
    }
    red[3 * blockIdx.x] = kmin;
    red[3 * blockIdx.x + 1] = kmax;
    red[3 * blockIdx.x + 2] = cnt;
  }
}

/* per-workgroup fixed-point partial sums of the arrival times */
__global__ void __launch_bounds__(CB) time_fsum(const double *__restrict__ s, const double *__restrict__ dp,
                                                const int32_t *__restrict__ lp, int64_t n, double Po,
                                                const coll_plan *plan, eh_fsum *partial,
                                                const int32_t *__restrict__ bunch, int32_t target) {
  __shared__ eh_fsum sh[CB];
  if (plan->nb <= 0 || plan->err)
    return;
  const int32_t e = plan->e;
  eh_fsum a;
  eh_fsum_init(&a, e);
  const int64_t stride = (int64_t)gridDim.x * CB;
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  for (; i + (SWEEP_U - 1) * stride < n; i += SWEEP_U * stride) {
    double S[SWEEP_U], D[SWEEP_U];
    bool live[SWEEP_U];
#pragma unroll
    for (int u = 0; u < SWEEP_U; u++) {
      live[u] = in_step(lp, bunch, target, i + u * stride);
      S[u] = s[i + u * stride];
      D[u] = dp[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < SWEEP_U; u++)
      if (live[u])
        eh_fsum_add(&a, arrival_time(S[u], D[u], Po));
  }
  for (; i < n; i += stride)
    if (in_step(lp, bunch, target, i))
      eh_fsum_add(&a, arrival_time(s[i], dp[i], Po));
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      eh_fsum_merge(&sh[threadIdx.x], &sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partial[blockIdx.x] = sh[0];
}

/* the mean from the partials, then tmin = tmean - dt nb / 2 (wake.c:165-168
 * for fixed N_BINS, zlongit.c:202-206) */
__global__ void __launch_bounds__(CB) fsum_final(eh_fsum *partial, int64_t nparts, coll_plan *plan, double dt) {
  __shared__ eh_fsum sh[CB];
  if (plan->nb <= 0 || plan->err)
    return;
  const int32_t e = plan->e;
  eh_fsum a;
  eh_fsum_init(&a, e);
  for (int64_t k = threadIdx.x; k < nparts; k += CB)
    eh_fsum_merge(&a, &partial[k]);
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      eh_fsum_merge(&sh[threadIdx.x], &sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double tmean = eh_fsum_to_double(&sh[0]) / (double)plan->red[2];
    plan->tmean = tmean;
    plan->tmin = tmean - dt * plan->nb / 2.0;
  }
}

/* bin index of wake.c:507 (round) or zlongit.c:267 (truncate) */
DEV int64_t bin_of(double t, double tmin, double dt, int round_half) {
  double v = round_half ? (t - tmin) / dt + 0.5 : (t - tmin) / dt;
  return (int64_t)v;
}

/* HB threads per histogram workgroup, one workgroup per CU: 16 waves keep
 * enough loads in flight per CU, and the merge still costs one global
 * atomic per bin per CU */
#define HB 1024
#define HIST_LDS_WORDS 16384 /* 64 KB of counts */
#define HIST_U 4 /* particles per lane per step of the histogram sweep (8 measured the same) */

__global__ void __launch_bounds__(HB) histogram(const double *__restrict__ s, const double *__restrict__ dp,
                                                const int32_t *__restrict__ lp, int64_t n, double Po,
                                                const coll_plan *plan, double dt, int32_t round_half,
                                                int32_t lds_bins, unsigned int *hist, const int32_t *__restrict__ bunch,
                                                int32_t target) {
  extern __shared__ unsigned int lh[];
  const int32_t nb = plan->nb;
  const double tmin = plan->tmin;
  if (nb <= 0 || plan->err)
    return;
  /* counts in LDS: as many copies (a power of 2, at most one per wave) as
   * fit, the waves sharing them in turn (a bunch puts most particles in a
   * few central bins, so fewer lanes share an address); else straight to
   * memory */
  int copies = HB / 64;
  while (copies > 1 && nb * copies > lds_bins)
    copies >>= 1;
  if (nb > lds_bins)
    copies = 0;
  unsigned int *h = copies ? lh + ((threadIdx.x >> 6) & (copies - 1)) * nb : hist;
  if (copies) {
    for (int k = threadIdx.x; k < nb * copies; k += HB)
      lh[k] = 0;
    __syncthreads();
  }
  const int64_t stride = (int64_t)gridDim.x * HB;
  int64_t i = (int64_t)blockIdx.x * HB + threadIdx.x;
  for (; i + (HIST_U - 1) * stride < n; i += HIST_U * stride) {
    double S[HIST_U], D[HIST_U];
    bool live[HIST_U];
#pragma unroll
    for (int u = 0; u < HIST_U; u++) {
      live[u] = in_step(lp, bunch, target, i + u * stride);
      S[u] = s[i + u * stride];
      D[u] = dp[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < HIST_U; u++)
      if (live[u]) {
        int64_t ib = bin_of(arrival_time(S[u], D[u], Po), tmin, dt, round_half);
        if (ib >= 0 && ib <= nb - 1)
          atomicAdd(&h[ib], 1u);
      }
  }
  for (; i < n; i += stride) {
    if (!in_step(lp, bunch, target, i))
      continue;
    int64_t ib = bin_of(arrival_time(s[i], dp[i], Po), tmin, dt, round_half);
    if (ib >= 0 && ib <= nb - 1)
      atomicAdd(&h[ib], 1u);
  }
  if (copies) {
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += HB) {
      unsigned int c = 0;
      for (int w = 0; w < copies; w++)
        c += lh[w * nb + k];
      if (c)
        atomicAdd(&hist[k], c);
    }
  }
}

/* convolveArrays with di2 = 0 (wake.c:479-495): V[ib] = sum_{j} I[j] W[ib-j]
 * over j = max(0, ib - wpoints + 1) .. ib, summed in increasing j, then
 * V *= factor (wake.c:258-260).  One lane per output bin.  The counts (as
 * doubles) and the window of W a workgroup needs are staged through LDS in
 * tiles of CONV_T counts, with zeros wherever W is outside [0, wpoints) and
 * past the last count.  Every lane then runs the tile's whole index range
 * j = t0 .. t1 in increasing j: a term outside the lane's own range is
 * count x 0 = +0, and adding +0 leaves the sum unchanged (it starts at +0
 * and is never -0), so each lane's sum is the reference's, and the index is
 * the same in every lane, which makes the count reads broadcasts. */
#define CONV_T 1024
#define CONV_B 64   /* output bins (lanes) per workgroup: one wave, so that a wave has its CU's LDS to itself */
#define CONV_PAD 48 /* zeros below the W window: the last steps of a tile and the loads two
                       steps ahead reach up to 39 below */
#define CONV_HPAD 48 /* zeros past the counts, read only by the loads ahead */
__global__ void __launch_bounds__(CONV_B) wake_convolve(const unsigned int *__restrict__ hist,
                                                        const double *__restrict__ W, int64_t wpoints,
                                                        const coll_plan *plan, double factor, double *volt) {
  __shared__ double hs[CONV_T + CONV_HPAD];
  __shared__ double ws[CONV_PAD + CONV_T + CONV_B];
  const int32_t nb = plan->nb;
  const int b0 = blockIdx.x * CONV_B;
  if (plan->err || nb <= 0 || b0 > nb)
    return;
  const int tid = (int)threadIdx.x;
  const int ib = b0 + tid;
  const int wp = (int)wpoints;
  const int last = (b0 + CONV_B - 1 < nb - 1) ? b0 + CONV_B - 1 : nb - 1;
  const int first = b0 - wp + 1 > 0 ? b0 - wp + 1 : 0;
  if (tid < CONV_PAD)
    ws[tid] = 0.0;
  if (tid < CONV_HPAD)
    hs[CONV_T + tid] = 0.0; /* read only by the loads ahead */
  /* staging of a tile [t0, t1]: every global load issued before the first
   * is used, with indices clamped into the arrays so that no load is
   * conditional (the values outside the ranges become zeros when stored);
   * the next tile's loads are in flight while the current tile is summed */
  unsigned int hv[CONV_T / CONV_B];
  double wv[(CONV_T + CONV_B) / CONV_B];
  auto stage_load = [&](int t0) {
    const int t1 = (t0 + CONV_T - 1 < last) ? t0 + CONV_T - 1 : last;
    const int len = t1 - t0 + 1, w_lo = b0 - t1;
#pragma unroll
    for (int r = 0; r < CONV_T / CONV_B; r++) {
      const int k = tid + r * CONV_B;
      hv[r] = hist[t0 + (k < len ? k : len - 1)];
    }
#pragma unroll
    for (int r = 0; r < (CONV_T + CONV_B) / CONV_B; r++) {
      const int w = w_lo + tid + r * CONV_B;
      wv[r] = W[w < 0 ? 0 : (w < wp ? w : wp - 1)];
    }
  };
  stage_load(first);
  double v = 0;
  for (int t0 = first; t0 <= last; t0 += CONV_T) {
    const int t1 = (t0 + CONV_T - 1 < last) ? t0 + CONV_T - 1 : last;
    const int len = t1 - t0 + 1;
    const int lenp = (len + 23) / 24 * 24;
    const int w_lo = b0 - t1; /* W index of ws[CONV_PAD] */
#pragma unroll
    for (int r = 0; r < CONV_T / CONV_B; r++) {
      const int k = tid + r * CONV_B;
      hs[k] = k < len ? (double)hv[r] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < (CONV_T + CONV_B) / CONV_B; r++) {
      const int w = w_lo + tid + r * CONV_B;
      ws[CONV_PAD + tid + r * CONV_B] = (w >= 0 && w < wp) ? wv[r] : 0.0;
    }
    __syncthreads();
    if (t0 + CONV_T <= last)
      stage_load(t0 + CONV_T);
    /* term t0 + q of lane tid uses W[ib - t0 - q] = ws[CONV_PAD + tid + len - 1 - q];
     * 24 terms per step in three groups of eight, each group's operands
     * loaded two groups before it is added (the loads ahead stay inside hs
     * and ws, in their zero padding) */
    const double *wq = ws + CONV_PAD + tid + len - 1;
    double ca[8], wa[8], cb[8], wb[8], cc[8], wc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      ca[u] = hs[u];
      wa[u] = wq[-u];
      cb[u] = hs[8 + u];
      wb[u] = wq[-(8 + u)];
    }
    for (int q = 0; q < lenp; q += 24) {
#pragma unroll
      for (int u = 0; u < 8; u++) {
        cc[u] = hs[q + 16 + u];
        wc[u] = wq[-(q + 16 + u)];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        v += ca[u] * wa[u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        ca[u] = hs[q + 24 + u];
        wa[u] = wq[-(q + 24 + u)];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        v += cb[u] * wb[u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        cb[u] = hs[q + 32 + u];
        wb[u] = wq[-(q + 32 + u)];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        v += cc[u] * wc[u];
    }
    __syncthreads();
  }
  if (ib < nb)
    volt[ib] = v * factor;
  else if (ib == nb)
    volt[nb] = 0; /* the slot past the last bin, read by interpolation */
}

/* packed real FFT, V = Z I factor, inverse FFT (zlongit.c:373-394), for
 * nb = 2m <= 8192 in one workgroup.  Same algorithm and operation order as
 * or_real_fft in the CPU checker: m-point radix-2 DIT complex FFT of
 * (x[2j], x[2j+1]) plus the real split, twiddles cos/sin(2 pi k/nb).
 * Everything stays in LDS: the points, the twiddle tables and, between the
 * two transforms, the packed spectrum (slot k holds the reference's
 * (V[2k-1], V[2k]), slot 0 holds (V[0], V[nb-1])).  Each LDS pass applies
 * two consecutive radix-2 stages to four points held in registers; the
 * butterflies and their operands are those of the stage-by-stage loop. */
#define FFT_T 1024
#define FFT_MAXM 4096
#define FFT_LDS_BINS (2 * FFT_MAXM) /* larger transforms take the global-memory path */
/* LDS index of point i: one padding word after every 16 and every 256
 * points, which spreads the strided accesses of the first passes over the
 * banks (the stage loop's patterns need 3328 bank cycles per 16-lane group
 * sum instead of 7680) */
#define FP(i) ((i) + ((i) >> 4) + ((i) >> 8))
#define FFT_PADM (FP(FFT_MAXM - 1) + 1)

/* one radix-2 butterfly (a, b) -> (a + t, a - t), t = b w or b conj(w) */
DEV void fft_bfly(double &ar, double &ai, double &br, double &bi, double c, double s, int inverse) {
  double tr, ti;
  if (!inverse) {
    tr = br * c + bi * s;
    ti = bi * c - br * s;
  } else {
    tr = br * c - bi * s;
    ti = bi * c + br * s;
  }
  const double xr = ar, xi = ai;
  ar = xr + tr;
  ai = xi + ti;
  br = xr - tr;
  bi = xi - ti;
}

/* the butterfly stages of one m-point transform on (zr, zi), bit-reversed
 * input; stage h has butterflies (g + j, g + j + h) with the twiddle of
 * table index j * m / h, which wr/wi hold at h + j (each stage's twiddles
 * contiguous, so that neighbouring lanes read neighbouring words); two
 * stages per pass */
DEV void fft_stages(double *zr, double *zi, const double *wr, const double *wi, int m, int bits, int inverse) {
  const int tid = (int)threadIdx.x;
  int lh = 0;
  for (; lh + 1 < bits; lh += 2) {
    const int h = 1 << lh;
    for (int grp = tid; grp < m / 4; grp += FFT_T) {
      const int j = grp & (h - 1);
      const int i0 = ((grp >> lh) << (lh + 2)) + j, i1 = i0 + h, i2 = i1 + h, i3 = i2 + h;
      double r0 = zr[FP(i0)], q0 = zi[FP(i0)], r1 = zr[FP(i1)], q1 = zi[FP(i1)];
      double r2 = zr[FP(i2)], q2 = zi[FP(i2)], r3 = zr[FP(i3)], q3 = zi[FP(i3)];
      const int ta = h + j, tb = 2 * h + j, tcix = 3 * h + j;
      /* stage h: (i0, i1) and (i2, i3), both at twiddle index j m / h */
      fft_bfly(r0, q0, r1, q1, wr[ta], wi[ta], inverse);
      fft_bfly(r2, q2, r3, q3, wr[ta], wi[ta], inverse);
      /* stage 2h: (i0, i2) at j m / 2h and (i1, i3) at (j + h) m / 2h */
      fft_bfly(r0, q0, r2, q2, wr[tb], wi[tb], inverse);
      fft_bfly(r1, q1, r3, q3, wr[tcix], wi[tcix], inverse);
      zr[FP(i0)] = r0;
      zi[FP(i0)] = q0;
      zr[FP(i1)] = r1;
      zi[FP(i1)] = q1;
      zr[FP(i2)] = r2;
      zi[FP(i2)] = q2;
      zr[FP(i3)] = r3;
      zi[FP(i3)] = q3;
    }
    __syncthreads();
  }
  if (lh < bits) {
    /* an odd stage count: the last stage alone */
    const int h = 1 << lh;
    for (int b = tid; b < m / 2; b += FFT_T) {
      const int j = b & (h - 1);
      const int i0 = ((b >> lh) << (lh + 1)) + j, i1 = i0 + h;
      const int t = h + j;
      double r0 = zr[FP(i0)], q0 = zi[FP(i0)], r1 = zr[FP(i1)], q1 = zi[FP(i1)];
      fft_bfly(r0, q0, r1, q1, wr[t], wi[t], inverse);
      zr[FP(i0)] = r0;
      zi[FP(i0)] = q0;
      zr[FP(i1)] = r1;
      zi[FP(i1)] = q1;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(FFT_T) zlongit_fft(const unsigned int *__restrict__ hist,
                                                     const double *__restrict__ Z, const double *__restrict__ tc_g,
                                                     const double *__restrict__ ts_g, int32_t nb, int32_t bits,
                                                     double factor, const coll_plan *plan, double *volt) {
  __shared__ double zr[FFT_PADM], zi[FFT_PADM], wr[FFT_MAXM], wi[FFT_MAXM];
  if (plan->nb <= 0 || plan->err)
    return;
  const int m = nb / 2;
  const int tid = (int)threadIdx.x;
  constexpr int KPT = FFT_MAXM / FFT_T; /* points per thread in the pointwise phases */
  /* Every global value the thread needs, loaded up front with indices
   * clamped into range so that no load waits behind a branch: for its points
   * k = tid + u FFT_T the stage-table entry (wr[h + j] = tc[j m / h] for
   * k = h + j, j < h), the pair (x[2k], x[2k+1]) loaded into bit-reversed
   * position, cos/sin(2 pi k / nb) of the real split, and Z[2k-1], Z[2k]
   * (Z[0], Z[nb-1] for k = 0). */
  double tck[KPT], tsk[KPT], zkr[KPT], zki[KPT];
  {
    double a[KPT], b[KPT];
    unsigned int x0[KPT], x1[KPT];
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      const int kc = k < m ? k : m - 1; /* clamped, for the loads only */
      const int k1 = kc < 1 ? 1 : kc;
      const int lh = 31 - __clz(k1), j = k1 - (1 << lh);
      a[u] = tc_g[j << (bits - lh)];
      b[u] = ts_g[j << (bits - lh)];
      x0[u] = hist[2 * kc];
      x1[u] = hist[2 * kc + 1];
      tck[u] = tc_g[kc];
      tsk[u] = ts_g[kc];
      zkr[u] = Z[kc == 0 ? 0 : 2 * kc - 1];
      zki[u] = Z[kc == 0 ? nb - 1 : 2 * kc];
    }
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      if (k >= 1 && k < m) {
        wr[k] = a[u];
        wi[k] = b[u];
      }
      if (k < m) {
        const int r = (int)(__brev((unsigned)k) >> (32 - bits));
        zr[FP(r)] = (double)x0[u];
        zi[FP(r)] = (double)x1[u];
      }
    }
  }
  __syncthreads();
  fft_stages(zr, zi, wr, wi, m, bits, 0);
  {
    /* real split -> packed spectrum (scaled by 1/nb), then V = Z I factor
     * (zlongit.c:379-391), kept in registers until every thread has read
     * the points it pairs with */
    const double scale = 1.0 / nb;
    double vr[KPT], vi[KPT];
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      if (k >= m)
        break;
      double pr, pi;
      if (k == 0) {
        pr = (zr[FP(0)] + zi[FP(0)]) * scale; /* V[0] */
        pi = (zr[FP(0)] - zi[FP(0)]) * scale; /* V[nb - 1] */
        vr[u] = pr * zkr[u] * factor;          /* Z[0] */
        vi[u] = pi * zki[u] * factor;          /* Z[nb - 1] */
      } else {
        double ar = zr[FP(k)], ai = zi[FP(k)], br = zr[FP(m - k)], bi = -zi[FP(m - k)];
        double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
        double dr = ar - br, di = ai - bi;
        double o_r = di * 0.5, o_i = -dr * 0.5;
        double c = tck[u], s = tsk[u];
        pr = (er + (c * o_r + s * o_i)) * scale; /* V[2k - 1] */
        pi = (ei + (c * o_i - s * o_r)) * scale; /* V[2k] */
        vr[u] = (pr * zkr[u] - pi * zki[u]) * factor; /* zkr = Z[2k-1], zki = Z[2k] */
        vi[u] = (pr * zki[u] + pi * zkr[u]) * factor;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      if (k >= m)
        break;
      zr[FP(k)] = vr[u];
      zi[FP(k)] = vi[u];
    }
    __syncthreads();
    /* the inverse transform's input, Z_k = (X_k + conj X_{m-k}) +
     * i w^-k (X_k - conj X_{m-k}), placed in bit-reversed order */
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      if (k >= m)
        break;
      if (k == 0) {
        double x0 = zr[FP(0)], xm = zi[FP(0)];
        vr[u] = x0 + xm;
        vi[u] = x0 - xm;
      } else {
        double ar = zr[FP(k)], ai = zi[FP(k)];
        double br = zr[FP(m - k)], bi = -zi[FP(m - k)];
        double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
        double c = tck[u], s = tsk[u];
        vr[u] = sr - (c * di + s * dr);
        vi[u] = si + (c * dr - s * di);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int k = tid + u * FFT_T;
      if (k >= m)
        break;
      int r = (int)(__brev((unsigned)k) >> (32 - bits));
      zr[FP(r)] = vr[u];
      zi[FP(r)] = vi[u];
    }
    __syncthreads();
  }
  fft_stages(zr, zi, wr, wi, m, bits, 1);
  for (int j = tid; j < m; j += FFT_T) {
    volt[2 * j] = zr[FP(j)];
    volt[2 * j + 1] = zi[FP(j)];
  }
  if (tid == 0)
    volt[nb] = 0;
}

/* The same transform for nb > 8192 (up to the 2^20 bins the bin optimiser
 * allows), where the points do not fit in LDS: the points live in global
 * memory (zr, zi, m doubles each, plus a second pair sr, si for the packed
 * spectrum) and each kernel below is one step of zlongit_fft, with the same
 * butterflies, twiddle indices and pointwise formulas; a step's kernel
 * boundary orders it after the previous one.  FG_B threads per workgroup,
 * one point, group or butterfly per thread. */
#define FG_B 256

__global__ void __launch_bounds__(FG_B) fftg_load(const unsigned int *__restrict__ hist, double *zr, double *zi,
                                                  int m, int bits, const coll_plan *plan) {
  const int k = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || k >= m)
    return;
  const int r = (int)(__brev((unsigned)k) >> (32 - bits));
  zr[r] = (double)hist[2 * k];
  zi[r] = (double)hist[2 * k + 1];
}

/* stages h = 2^lh and 2h on the groups of four points of fft_stages */
__global__ void __launch_bounds__(FG_B) fftg_stage2(double *zr, double *zi, const double *__restrict__ tc,
                                                    const double *__restrict__ ts, int m, int bits, int lh,
                                                    int inverse, const coll_plan *plan) {
  const int grp = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || grp >= m / 4)
    return;
  const int h = 1 << lh;
  const int j = grp & (h - 1);
  const int i0 = ((grp >> lh) << (lh + 2)) + j, i1 = i0 + h, i2 = i1 + h, i3 = i2 + h;
  double r0 = zr[i0], q0 = zi[i0], r1 = zr[i1], q1 = zi[i1];
  double r2 = zr[i2], q2 = zi[i2], r3 = zr[i3], q3 = zi[i3];
  const int ta = j << (bits - lh), tb = j << (bits - lh - 1), tcix = (j + h) << (bits - lh - 1);
  fft_bfly(r0, q0, r1, q1, tc[ta], ts[ta], inverse);
  fft_bfly(r2, q2, r3, q3, tc[ta], ts[ta], inverse);
  fft_bfly(r0, q0, r2, q2, tc[tb], ts[tb], inverse);
  fft_bfly(r1, q1, r3, q3, tc[tcix], ts[tcix], inverse);
  zr[i0] = r0;
  zi[i0] = q0;
  zr[i1] = r1;
  zi[i1] = q1;
  zr[i2] = r2;
  zi[i2] = q2;
  zr[i3] = r3;
  zi[i3] = q3;
}

/* the last stage alone, for an odd stage count */
__global__ void __launch_bounds__(FG_B) fftg_stage1(double *zr, double *zi, const double *__restrict__ tc,
                                                    const double *__restrict__ ts, int m, int bits, int lh,
                                                    int inverse, const coll_plan *plan) {
  const int b = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || b >= m / 2)
    return;
  const int h = 1 << lh;
  const int j = b & (h - 1);
  const int i0 = ((b >> lh) << (lh + 1)) + j, i1 = i0 + h;
  const int t = j << (bits - lh);
  double r0 = zr[i0], q0 = zi[i0], r1 = zr[i1], q1 = zi[i1];
  fft_bfly(r0, q0, r1, q1, tc[t], ts[t], inverse);
  zr[i0] = r0;
  zi[i0] = q0;
  zr[i1] = r1;
  zi[i1] = q1;
}

/* real split -> packed spectrum, V = Z I factor (slot k of sr, si as in
 * zlongit_fft) */
__global__ void __launch_bounds__(FG_B) fftg_split(const double *__restrict__ zr, const double *__restrict__ zi,
                                                   double *sr, double *si, const double *__restrict__ tc,
                                                   const double *__restrict__ ts, const double *__restrict__ Z,
                                                   int nb, double factor, const coll_plan *plan) {
  const int m = nb / 2;
  const int k = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || k >= m)
    return;
  const double scale = 1.0 / nb;
  if (k == 0) {
    const double pr = (zr[0] + zi[0]) * scale, pi = (zr[0] - zi[0]) * scale;
    sr[0] = pr * Z[0] * factor;
    si[0] = pi * Z[nb - 1] * factor;
  } else {
    double ar = zr[k], ai = zi[k], br = zr[m - k], bi = -zi[m - k];
    double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
    double dr = ar - br, di = ai - bi;
    double o_r = di * 0.5, o_i = -dr * 0.5;
    double c = tc[k], s = ts[k];
    const double pr = (er + (c * o_r + s * o_i)) * scale, pi = (ei + (c * o_i - s * o_r)) * scale;
    const int iR = 2 * k - 1, iI = iR + 1;
    sr[k] = (pr * Z[iR] - pi * Z[iI]) * factor;
    si[k] = (pr * Z[iI] + pi * Z[iR]) * factor;
  }
}

/* the inverse transform's input in bit-reversed order */
__global__ void __launch_bounds__(FG_B) fftg_prep(const double *__restrict__ sr, const double *__restrict__ si,
                                                  double *zr, double *zi, const double *__restrict__ tc,
                                                  const double *__restrict__ ts, int m, int bits,
                                                  const coll_plan *plan) {
  const int k = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || k >= m)
    return;
  double zrk, zik;
  if (k == 0) {
    const double x0 = sr[0], xm = si[0];
    zrk = x0 + xm;
    zik = x0 - xm;
  } else {
    double ar = sr[k], ai = si[k];
    double br = sr[m - k], bi = -si[m - k];
    double sr_ = ar + br, si_ = ai + bi, dr = ar - br, di = ai - bi;
    double c = tc[k], s = ts[k];
    zrk = sr_ - (c * di + s * dr);
    zik = si_ + (c * dr - s * di);
  }
  const int r = (int)(__brev((unsigned)k) >> (32 - bits));
  zr[r] = zrk;
  zi[r] = zik;
}

__global__ void __launch_bounds__(FG_B) fftg_out(const double *__restrict__ zr, const double *__restrict__ zi,
                                                 double *volt, int m, const coll_plan *plan) {
  const int j = blockIdx.x * FG_B + threadIdx.x;
  if (plan->nb <= 0 || plan->err || j >= m)
    return;
  volt[2 * j] = zr[j];
  volt[2 * j + 1] = zi[j];
  if (j == 0)
    volt[2 * m] = 0;
}

/* host side of the global-memory path: scratch holds zr, zi, sr, si */
static void fftg_launch(hipStream_t st, const unsigned int *hist, const double *Z, const double *tc, const double *ts,
                        int nb, int bits, double factor, const coll_plan *plan, double *volt, double *scratch) {
  const int m = nb / 2;
  double *zr = scratch, *zi = scratch + m, *sr = scratch + 2 * m, *si = scratch + 3 * m;
  const unsigned gm = (unsigned)((m + FG_B - 1) / FG_B);
  const unsigned g2 = (unsigned)((m / 2 + FG_B - 1) / FG_B);
  const unsigned g4 = (unsigned)((m / 4 + FG_B - 1) / FG_B);
  hipLaunchKernelGGL(fftg_load, dim3(gm), dim3(FG_B), 0, st, hist, zr, zi, m, bits, plan);
  for (int inverse = 0; inverse < 2; inverse++) {
    int lh = 0;
    for (; lh + 1 < bits; lh += 2)
      hipLaunchKernelGGL(fftg_stage2, dim3(g4), dim3(FG_B), 0, st, zr, zi, tc, ts, m, bits, lh, inverse, plan);
    if (lh < bits)
      hipLaunchKernelGGL(fftg_stage1, dim3(g2), dim3(FG_B), 0, st, zr, zi, tc, ts, m, bits, lh, inverse, plan);
    if (!inverse) {
      hipLaunchKernelGGL(fftg_split, dim3(gm), dim3(FG_B), 0, st, zr, zi, sr, si, tc, ts, Z, nb, factor, plan);
      hipLaunchKernelGGL(fftg_prep, dim3(gm), dim3(FG_B), 0, st, sr, si, zr, zi, tc, ts, m, bits, plan);
    }
  }
  hipLaunchKernelGGL(fftg_out, dim3(gm), dim3(FG_B), 0, st, zr, zi, volt, m, plan);
}

/* add_to_particle_energy (simple_rfca.c:987-1007) */
DEV void add_energy(double &xp, double &yp, double &s, double &dp, double tof, double Po, double dgamma) {
  double P = Po * (1 + dp);
  double gamma = __builtin_sqrt(P * P + 1);
  double gamma1 = gamma + dgamma;
  if (gamma1 <= 1)
    gamma1 = 1 + 1e-7;
  double P1 = __builtin_sqrt(gamma1 * gamma1 - 1);
  dp = (P1 - Po) / Po;
  s = tof * EH_C_MKS * P1 / gamma1;
  double Pz = P / __builtin_sqrt(1 + sqd(xp) + sqd(yp));
  double Pz1 = __builtin_sqrt(Pz * Pz + gamma1 * gamma1 - gamma * gamma);
  double ratio = Pz / Pz1;
  xp *= ratio;
  yp *= ratio;
}

/* applyLongitudinalWakeKicks (wake.c:305-329) / the ZLONGIT kick loop
 * (zlongit.c:454-486), then the post-element checks of do_tracking.  kind < 0:
 * the checks alone (a bunch the element leaves out), plan and volt unused. */
__global__ void __launch_bounds__(CB) kick(double *x, double *xp, double *y, double *yp, double *s, double *dp,
                                           int32_t *lp, int32_t *lc, int64_t n, double Po, const coll_plan *plan,
                                           double dt, const double *__restrict__ volt, int32_t kind,
                                           int32_t interpolate, double z_end, double ap_x, double ap_y,
                                           int32_t ap_on, int32_t ap_code, int64_t pass, int32_t code_base,
                                           const int32_t *__restrict__ bunch, int32_t target, double mc2) {
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (i >= n || !in_step(lp, bunch, target, i))
    return;
  int32_t nb = 0;
  double tmin = 0;
  if (kind >= 0) {
    nb = plan->nb;
    tmin = plan->tmin;
    if (nb <= 0 || plan->err)
      return;
  }
  double S = s[i], D = dp[i];
  double XP = xp[i], YP = yp[i];
  bool changed = false;
  double t = 0;
  int64_t ib = -1;
  if (kind >= 0) {
    t = arrival_time(S, D, Po);
    ib = bin_of(t, tmin, dt, kind == EH_OP_WAKE);
  }
  if (ib >= 0 && ib <= nb - 1) {
    double dgam;
    bool skip = false;
    if (kind == EH_OP_WAKE) {
      if (interpolate) {
        double dt1 = t - (tmin + dt * ib);
        if ((dt1 < 0 && ib) || ib == nb - 1) {
          ib--;
          dt1 += dt;
        }
        dgam = (volt[ib] + (volt[ib + 1] - volt[ib]) / dt * dt1) / mc2;
      } else
        dgam = volt[ib] / mc2;
    } else {
      if (interpolate && ib > 0) {
        double dt1 = t - (tmin + dt / 2.0 + dt * ib);
        if (dt1 < 0)
          dt1 += dt;
        else
          ib += 1;
        if (ib < nb)
          dgam = (volt[ib - 1] + (volt[ib] - volt[ib - 1]) / dt * dt1) / mc2;
        else {
          skip = true;
          dgam = 0;
        }
      } else
        dgam = volt[ib] / mc2;
    }
    if (!skip && dgam) {
      add_energy(XP, YP, S, D, t, Po, -dgam);
      changed = true;
    }
  }
  if (changed) {
    xp[i] = XP;
    yp[i] = YP;
    s[i] = S;
    dp[i] = D;
  }
  /* limit_amplitudes with the MAXAMP aperture, then the NaN/Inf cull */
  if (ap_on < 0)
    return;
  double X = x[i], Y = y[i];
  if (ap_on && ap_code) {
    /* a shaped MAXAMP (ellipse or open side); no drift extrapolation here */
    double S1 = S, D1 = D;
    if (ap_maxamp_shaped(X, XP, Y, YP, S1, D1, ap_x, ap_y, ap_code, false, z_end, Po)) {
      x[i] = X;
      y[i] = Y;
      s[i] = S1;
      dp[i] = D1;
      lp[i] = (int32_t)pass;
      lc[i] = code_base + 1;
      return;
    }
    ap_on = 0;
  }
  for (int stage = 1; stage <= 2; stage++) {
    double xmax = stage == 1 ? ap_x : 1.7976931348623157e308;
    double ymax = stage == 1 ? ap_y : 1.7976931348623157e308;
    if (stage == 1 && !ap_on)
      continue;
    int is_out = 0;
    if (xmax > 0 && __builtin_fabs(X) > xmax)
      is_out += 1;
    if (ymax > 0 && __builtin_fabs(Y) > ymax)
      is_out += 2;
    if (__builtin_isinf(X) || __builtin_isinf(Y) || __builtin_isnan(X) || __builtin_isnan(Y))
      is_out += 4;
    if (is_out) {
      /* no drift extrapolation for zero-length elements; dz = 0 */
      x[i] = X + 0 * XP;
      y[i] = Y + 0 * YP;
      s[i] = z_end + 0;
      dp[i] = Po * (1 + D);
      lp[i] = (int32_t)pass;
      lc[i] = code_base + stage;
      return;
    }
  }
}

/* ------------------------------------------------------------ host side */

extern "C" {

static int cfail(ehd_dev *d, hipError_t e, const char *what) {
  snprintf(d->msg, sizeof(d->msg), "%s: %s", what, hipGetErrorString(e));
  return EH_ERR_HIP;
}

#define CHK(call)                                                                                                      \
  do {                                                                                                                 \
    hipError_t e_ = (call);                                                                                            \
    if (e_ != hipSuccess)                                                                                              \
      return cfail(d, e_, #call);                                                                                      \
  } while (0)

/* teardown: errors are ignored, there is nothing left to undo */
void ehd_free_collective(ehd_dev *d) {
  for (int k = 0; k < EHD_MAX_COLL; k++) {
    ehd_coll_slot *sl = &d->slot[k];
    void *bufs[4] = {sl->W, sl->Z, sl->tc, sl->ts};
    for (int b = 0; b < 4; b++)
      if (bufs[b])
        (void)hipFree(bufs[b]);
    memset(sl, 0, sizeof(*sl));
  }
  if (d->cwork)
    (void)hipFree(d->cwork);
  d->cwork = NULL;
  d->cwork_bytes = 0;
  if (d->bunch)
    (void)hipFree(d->bunch);
  d->bunch = NULL;
  d->bunch_cap = 0;
}

int ehd_put_wake(ehd_dev *d, int32_t slot, const double *W, int64_t npoints) {
  if (slot < 0 || slot >= EHD_MAX_COLL || npoints < 2)
    return EH_ERR_INVALID;
  ehd_coll_slot *sl = &d->slot[slot];
  CHK(hipSetDevice(d->device));
  if (sl->W)
    CHK(hipFree(sl->W));
  sl->W = NULL;
  CHK(hipMalloc((void **)&sl->W, sizeof(double) * npoints));
  CHK(hipMemcpy(sl->W, W, sizeof(double) * npoints, hipMemcpyHostToDevice));
  sl->wpoints = npoints;
  return EH_OK;
}

int ehd_put_zlongit(ehd_dev *d, int32_t slot, const double *Z, const double *tc, const double *ts, int32_t nb) {
  if (slot < 0 || slot >= EHD_MAX_COLL || nb < 4 || nb > EHD_ZL_MAX_BINS || (nb & (nb - 1)))
    return EH_ERR_INVALID;
  ehd_coll_slot *sl = &d->slot[slot];
  CHK(hipSetDevice(d->device));
  if (sl->Z)
    CHK(hipFree(sl->Z));
  if (sl->tc)
    CHK(hipFree(sl->tc));
  if (sl->ts)
    CHK(hipFree(sl->ts));
  sl->Z = sl->tc = sl->ts = NULL;
  CHK(hipMalloc((void **)&sl->Z, sizeof(double) * nb));
  CHK(hipMalloc((void **)&sl->tc, sizeof(double) * (nb / 2 + 1)));
  CHK(hipMalloc((void **)&sl->ts, sizeof(double) * (nb / 2 + 1)));
  CHK(hipMemcpy(sl->Z, Z, sizeof(double) * nb, hipMemcpyHostToDevice));
  CHK(hipMemcpy(sl->tc, tc, sizeof(double) * (nb / 2 + 1), hipMemcpyHostToDevice));
  CHK(hipMemcpy(sl->ts, ts, sizeof(double) * (nb / 2 + 1), hipMemcpyHostToDevice));
  sl->nb = nb;
  return EH_OK;
}

/* work area: [0,128) the step plan, then the histogram (uint32 x (cap+8)),
 * the voltage (double x (cap+8)), the fixed-sum partials, and for ZLONGIT
 * beyond the LDS transform's 8192 bins the global transform's points
 * (4 x nb/2 doubles, 256-byte aligned) */
#define PLAN_BYTES 128
static int64_t hist_bytes(int64_t nb_cap) { return ((4 * (nb_cap + 8) + 7) / 8) * 8; }
static int64_t fft_offset(int64_t nb_cap, int64_t nparts) {
  int64_t o = PLAN_BYTES + hist_bytes(nb_cap) + 8 * (nb_cap + 8) + (int64_t)sizeof(eh_fsum) * (nparts + 1);
  return (o + 255) / 256 * 256;
}
static int64_t fft_bytes(int64_t nb) { return nb > FFT_LDS_BINS ? 16 * nb : 0; }

static int ensure_work(ehd_dev *d, int64_t nb_cap, int64_t nparts, int64_t nb_fft) {
  int64_t need = fft_offset(nb_cap, nparts) + fft_bytes(nb_fft) + 256;
  if (need <= d->cwork_bytes)
    return EH_OK;
  if (d->cwork)
    CHK(hipFree(d->cwork));
  d->cwork = NULL;
  d->cwork_bytes = 0;
  CHK(hipMalloc(&d->cwork, need));
  d->cwork_bytes = need;
  return EH_OK;
}

/* workgroups of the grid-stride reductions: enough to fill the device, few
 * enough that the per-workgroup atomics and partials stay cheap */
static int64_t reduce_grid(int64_t n) {
  int64_t g = (n + CB - 1) / CB;
  if (g > 1024)
    g = 1024;
  return g < 1 ? 1 : g;
}

/* One WAKE or ZLONGIT step on the particle arrays crd[0..5] (alive: lp[i] < 0,
 * or every particle when lp is NULL), queued as one sequence of kernels with
 * no host round trip: reset, min/max, plan, [mean], histogram, convolution
 * or FFT, kick.  work is the step's scratch (plan, histogram, voltage,
 * partials) laid out for nb_max bins; the host waits once at the end for the
 * plan (error, bin layout), which lands in *hp. */
static int coll_launch(ehd_dev *d, const ehd_collective *c, double *const crd[6], int32_t *lp, int32_t *lc,
                       int64_t n, const double *W, int64_t wpoints, const ehd_coll_slot *sl, int64_t nb_max,
                       void *work, coll_plan *hp) {
  const int64_t grid = (n + CB - 1) / CB;
  const int64_t rgrid = reduce_grid(n);
  const int64_t nb_cap = nb_max + 8;
  coll_plan *plan = (coll_plan *)work;
  unsigned int *hist = (unsigned int *)((char *)work + PLAN_BYTES);
  double *volt = (double *)((char *)hist + hist_bytes(nb_cap));
  eh_fsum *parts = (eh_fsum *)((char *)volt + 8 * (nb_cap + 8));
  const int32_t kind = c->kind;
  const int32_t fixed = kind == EH_OP_WAKE ? c->n_bins_fixed : 0;
  const int32_t nclear = (int32_t)(nb_cap + 8);
  /* without a loss-pass array nothing may be marked lost: no checks */
  const int32_t ap_on = (!lp || (c->flags & EHD_COLL_NOCHECKS)) ? -1 : ((c->flags & EH_F_AP) ? 1 : 0);
  /* one bunch of a bunched beam: only its particles take part */
  const int32_t *bunch = c->use_bunch ? d->bunch : NULL;
  const int32_t target = c->bunch_target;
  const int32_t mean_given = kind == EH_OP_WAKE && fixed && c->mean_given;
  CHK(hipEventRecord(d->cev0, d->stream));
  hipLaunchKernelGGL(time_minmax, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, crd[4], crd[5], lp, n, c->Po,
                     (unsigned long long *)parts, bunch, target);
  hipLaunchKernelGGL(coll_plan_minmax, dim3(1), dim3(CB), 0, d->stream, plan, (const unsigned long long *)parts, rgrid,
                     kind, fixed, c->nb, c->dt, c->wake_span, c->allow_long_beam, (int32_t)nb_max, hist, nclear,
                     mean_given, c->mean);
  if ((kind == EH_OP_ZLONGIT || fixed) && !mean_given) {
    hipLaunchKernelGGL(time_fsum, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, crd[4], crd[5], lp, n, c->Po, plan,
                       parts, bunch, target);
    hipLaunchKernelGGL(fsum_final, dim3(1), dim3(CB), 0, d->stream, parts, rgrid, plan, c->dt);
  }
  {
    /* one workgroup per CU: each sees enough particles that its LDS counts
     * aggregate many hits per bin before the one global atomic per bin (four
     * per CU measured slower: the merge's global atomics grow with the
     * number of workgroups) */
    if (!d->num_cu) {
      int cu = 0;
      CHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d->device));
      d->num_cu = cu > 0 ? cu : 256;
    }
    int hgrid = (int)((n + HB - 1) / HB);
    if (hgrid < 1)
      hgrid = 1;
    if (hgrid > d->num_cu)
      hgrid = d->num_cu;
    /* room for one copy of the counts per wave, at most 64 KB of LDS */
    const int32_t lds_bins =
        nb_max * (HB / 64) < HIST_LDS_WORDS ? (int32_t)(nb_max * (HB / 64)) : HIST_LDS_WORDS;
    hipLaunchKernelGGL(histogram, dim3(hgrid), dim3(HB), sizeof(unsigned int) * lds_bins, d->stream, crd[4], crd[5],
                       lp, n, c->Po, plan, c->dt, kind == EH_OP_WAKE ? 1 : 0, lds_bins, hist, bunch, target);
  }
  if (kind == EH_OP_WAKE) {
    hipLaunchKernelGGL(wake_convolve, dim3((unsigned)((nb_max + 1 + CONV_B - 1) / CONV_B)), dim3(CONV_B), 0, d->stream,
                       hist, W, wpoints, plan, c->factor, volt);
  } else {
    int bits = 0;
    while ((1 << bits) < c->nb / 2)
      bits++;
    if (c->nb <= FFT_LDS_BINS)
      hipLaunchKernelGGL(zlongit_fft, dim3(1), dim3(FFT_T), 0, d->stream, hist, sl->Z, sl->tc, sl->ts, c->nb, bits,
                         c->factor, plan, volt);
    else
      fftg_launch(d->stream, hist, sl->Z, sl->tc, sl->ts, c->nb, bits, c->factor, plan, volt,
                  (double *)((char *)work + fft_offset(nb_cap, rgrid)));
  }
  if (n > 0)
    hipLaunchKernelGGL(kick, dim3((unsigned)grid), dim3(CB), 0, d->stream, crd[0], crd[1], crd[2], crd[3], crd[4],
                       crd[5], lp, lc, n, c->Po, plan, c->dt, volt, kind, c->interpolate, c->z_end, c->ap_xmax,
                       c->ap_ymax, ap_on, c->ap_code, c->pass, 4 * c->op_index, bunch, target, c->mc2);
  CHK(hipGetLastError());
  CHK(hipEventRecord(d->cev1, d->stream));
  CHK(hipMemcpyAsync(hp, plan, sizeof(*hp), hipMemcpyDeviceToHost, d->stream));
  CHK(hipStreamSynchronize(d->stream));
  float ms = 0;
  CHK(hipEventElapsedTime(&ms, d->cev0, d->cev1));
  d->coll_ms += ms;
  d->coll_launches += 1;
  return EH_OK;
}

/* the post-element checks alone for the particles of one bunch that a
 * WAKE/ZLONGIT leaves out (START_BUNCH/END_BUNCH) */
int ehd_collective_checks(ehd_dev *d, const ehd_collective *c) {
  const int64_t n = d->n;
  CHK(hipSetDevice(d->device));
  if (n <= 0 || !d->loss_pass || (c->flags & EHD_COLL_NOCHECKS))
    return EH_OK;
  const int32_t ap_on = (c->flags & EH_F_AP) ? 1 : 0;
  const int32_t *bunch = c->use_bunch ? d->bunch : NULL;
  hipLaunchKernelGGL(kick, dim3((unsigned)((n + CB - 1) / CB)), dim3(CB), 0, d->stream, d->coord[0], d->coord[1],
                     d->coord[2], d->coord[3], d->coord[4], d->coord[5], d->loss_pass, d->loss_code, n, c->Po,
                     (const coll_plan *)NULL, 0.0, (const double *)NULL, -1, 0, c->z_end, c->ap_xmax, c->ap_ymax,
                     ap_on, c->ap_code, c->pass, 4 * c->op_index, bunch, c->bunch_target, c->mc2);
  CHK(hipGetLastError());
  CHK(hipStreamSynchronize(d->stream));
  return EH_OK;
}

int ehd_collective_step(ehd_dev *d, int32_t slot, const ehd_collective *c, double *hist_out, double *volt_out,
                        int64_t hist_cap, int64_t *nb_out, double *tmin_out) {
  (void)hist_cap;
  if (slot < 0 || slot >= EHD_MAX_COLL)
    return EH_ERR_INVALID;
  ehd_coll_slot *sl = &d->slot[slot];
  const int64_t n = d->n;
  CHK(hipSetDevice(d->device));
  const int64_t nb_max = c->kind == EH_OP_WAKE ? sl->wpoints : sl->nb; /* most bins this element can use */
  const int64_t nb_cap = nb_max + 8;
  int rc = ensure_work(d, nb_cap, reduce_grid(n), c->kind == EH_OP_ZLONGIT ? nb_max : 0);
  if (rc)
    return rc;
  coll_plan hp;
  rc = coll_launch(d, c, d->coord, d->loss_pass, d->loss_code, n, sl->W, sl->wpoints, sl, nb_max, d->cwork, &hp);
  if (rc)
    return rc;
  if (hp.err)
    return hp.err;
  unsigned int *hist = (unsigned int *)((char *)d->cwork + PLAN_BYTES);
  double *volt = (double *)((char *)hist + hist_bytes(nb_cap));
  const int32_t nb = hp.red[2] ? hp.nb : 0;
  if (nb > 0 && (hist_out || volt_out)) {
    unsigned int *h = (unsigned int *)malloc(sizeof(unsigned int) * nb);
    if (!h)
      return EH_ERR_NOMEM;
    CHK(hipMemcpyAsync(h, hist, sizeof(unsigned int) * nb, hipMemcpyDeviceToHost, d->stream));
    if (volt_out)
      CHK(hipMemcpyAsync(volt_out, volt, sizeof(double) * nb, hipMemcpyDeviceToHost, d->stream));
    CHK(hipStreamSynchronize(d->stream));
    if (hist_out)
      for (int k = 0; k < nb; k++)
        hist_out[k] = (double)h[k];
    free(h);
  }
  if (nb_out)
    *nb_out = nb;
  if (tmin_out)
    *tmin_out = hp.tmin;
  return EH_OK;
}

/* ---- the WAKE step without allocation ----
 * Scratch of the step (plan, histogram, voltage, partials) in a static
 * device array instead of hipMalloc'ed memory.  Each GPU holds its own copy
 * from module load until the process ends.  Host code never uses the
 * array's name as an address (on the host it is only a stand-in): ehd_create
 * resolves the device address once per context, with the device current,
 * through hipGetSymbolAddress and keeps it in d->noalloc. */
static __device__ __align__(256) unsigned char scratch[EHD_NOALLOC_BYTES];

int ehd_noalloc_init(ehd_dev *d) {
  void *ptr = NULL;
  CHK(hipSetDevice(d->device));
  CHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  d->noalloc = ptr;
  return EH_OK;
}

int64_t ehd_noalloc_bins_max(void) { return EHD_NOALLOC_BINS; }

int ehd_wake_step_noalloc(ehd_dev *d, const ehd_collective *c, double *const crd[6], int64_t n, const double *W,
                          int64_t wpoints) {
  if (c->kind != EH_OP_WAKE || !W || wpoints < 2 || n < 0)
    return EH_ERR_INVALID;
  if (wpoints > EHD_NOALLOC_BINS || n > ((int64_t)1 << 40)) {
    snprintf(d->msg, sizeof(d->msg), "no-allocation WAKE step: %lld wake points exceed the scratch limit of %d",
             (long long)wpoints, EHD_NOALLOC_BINS);
    return EH_ERR_INVALID;
  }
  if (!d->noalloc) {
    snprintf(d->msg, sizeof(d->msg), "no-allocation WAKE step: scratch address not resolved");
    return EH_ERR_STATE;
  }
  const int64_t nb_cap = wpoints + 8;
  const int64_t need = PLAN_BYTES + hist_bytes(nb_cap) + 8 * (nb_cap + 8) +
                       (int64_t)sizeof(eh_fsum) * (reduce_grid(n) + 1) + 256;
  if (need > EHD_NOALLOC_BYTES)
    return EH_ERR_INVALID; /* not reached for wpoints <= EHD_NOALLOC_BINS; kept as a guard */
  if (n == 0)
    return EH_OK;
  CHK(hipSetDevice(d->device));
  coll_plan hp;
  int rc = coll_launch(d, c, crd, NULL, NULL, n, W, wpoints, NULL, wpoints, d->noalloc, &hp);
  if (rc)
    return rc;
  return hp.err ? hp.err : EH_OK;
}

int ehd_time_span(ehd_dev *d, double Po, int32_t use_bunch, int32_t target, double *tmin, double *tmax,
                  int64_t *count) {
  const int64_t n = d->n;
  const int64_t rgrid = reduce_grid(n);
  CHK(hipSetDevice(d->device));
  int rc = ensure_work(d, 64, rgrid, 0);
  if (rc)
    return rc;
  coll_plan *plan = (coll_plan *)d->cwork;
  unsigned int *hist = (unsigned int *)((char *)d->cwork + PLAN_BYTES);
  unsigned long long *mm = (unsigned long long *)((char *)hist + hist_bytes(64) + 8 * (64 + 8));
  hipLaunchKernelGGL(time_minmax, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, d->coord[4], d->coord[5],
                     d->loss_pass, n, Po, mm, use_bunch ? d->bunch : NULL, target);
  hipLaunchKernelGGL(coll_plan_minmax, dim3(1), dim3(CB), 0, d->stream, plan, (const unsigned long long *)mm, rgrid,
                     -1, 0, 0, 1.0, 0.0, 0, 0, hist, 0, 0, 0.0);
  CHK(hipGetLastError());
  coll_plan hp;
  CHK(hipMemcpyAsync(&hp, plan, sizeof(hp), hipMemcpyDeviceToHost, d->stream));
  CHK(hipStreamSynchronize(d->stream));
  *count = (int64_t)hp.red[2];
  *tmin = key_value(hp.red[0]);
  *tmax = key_value(hp.red[1]);
  return EH_OK;
}

/* the bunch index of every particle (column 8 of the rows), for the
 * collective steps of a bunched beam */
int ehd_put_bunch(ehd_dev *d, const int32_t *bunch, int64_t n) {
  if (n != d->n || (n > 0 && !bunch))
    return EH_ERR_INVALID;
  CHK(hipSetDevice(d->device));
  if (n > d->bunch_cap) {
    if (d->bunch)
      CHK(hipFree(d->bunch));
    d->bunch = NULL;
    d->bunch_cap = 0;
    CHK(hipMalloc((void **)&d->bunch, sizeof(int32_t) * (size_t)n));
    d->bunch_cap = n;
  }
  if (n > 0)
    CHK(hipMemcpy(d->bunch, bunch, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  return EH_OK;
}

int ehd_collective_timing(ehd_dev *d, int64_t *launches, double *ms) {
  *launches = d->coll_launches;
  *ms = d->coll_ms;
  return EH_OK;
}
}
