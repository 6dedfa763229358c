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
                                                       int32_t nb_cap, unsigned int *hist, int32_t nclear) {
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
      p->nb = n_fixed; /* tmin set from the mean */
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
__global__ void __launch_bounds__(CB) time_minmax(const double *__restrict__ s, const double *__restrict__ dp,
                                                  const int32_t *__restrict__ lp, int64_t n, double Po,
                                                  unsigned long long *red) {
  __shared__ unsigned long long smin[CB / 64], smax[CB / 64], scnt[CB / 64];
  unsigned long long kmin = ~0ULL, kmax = 0, cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB) {
    if (!lp || lp[i] < 0) {
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
                                                const coll_plan *plan, eh_fsum *partial) {
  __shared__ eh_fsum sh[CB];
  if (plan->nb <= 0 || plan->err)
    return;
  const int32_t e = plan->e;
  eh_fsum a;
  eh_fsum_init(&a, e);
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB)
    if (!lp || lp[i] < 0)
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

__global__ void __launch_bounds__(CB) histogram(const double *__restrict__ s, const double *__restrict__ dp,
                                                const int32_t *__restrict__ lp, int64_t n, double Po,
                                                const coll_plan *plan, double dt, int32_t round_half,
                                                int32_t lds_bins, unsigned int *hist) {
  extern __shared__ unsigned int lh[];
  const int32_t nb = plan->nb;
  const double tmin = plan->tmin;
  if (nb <= 0 || plan->err)
    return;
  /* counts in LDS, one copy per wave when they fit (a bunch puts most
   * particles in a few central bins, so fewer lanes share an address), one
   * per workgroup when only that fits, else straight to memory */
  const int copies = nb * (CB / 64) <= lds_bins ? CB / 64 : (nb <= lds_bins ? 1 : 0);
  unsigned int *h = copies ? lh + (copies > 1 ? (threadIdx.x >> 6) * nb : 0) : hist;
  if (copies) {
    for (int k = threadIdx.x; k < nb * copies; k += CB)
      lh[k] = 0;
    __syncthreads();
  }
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB) {
    if (lp && lp[i] >= 0)
      continue;
    int64_t ib = bin_of(arrival_time(s[i], dp[i], Po), tmin, dt, round_half);
    if (ib >= 0 && ib <= nb - 1)
      atomicAdd(&h[ib], 1u);
  }
  if (copies) {
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += CB) {
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
 * V *= factor (wake.c:258-260).  One lane per output bin;
 * the counts and the window of W a workgroup needs are staged through LDS
 * in tiles of CONV_T counts, which keeps every lane's sum in the reference
 * order. */
#define CONV_T 1024
__global__ void __launch_bounds__(CB) wake_convolve(const unsigned int *__restrict__ hist,
                                                    const double *__restrict__ W, int64_t wpoints,
                                                    const coll_plan *plan, double factor, double *volt) {
  __shared__ unsigned int hs[CONV_T];
  __shared__ double ws[CONV_T + CB];
  const int32_t nb = plan->nb;
  const int b0 = blockIdx.x * CB;
  if (plan->err || nb <= 0 || b0 > nb)
    return;
  const int ib = b0 + (int)threadIdx.x;
  const int wp = (int)wpoints;
  const int last = (b0 + CB - 1 < nb - 1) ? b0 + CB - 1 : nb - 1;
  const int first = b0 - wp + 1 > 0 ? b0 - wp + 1 : 0;
  double v = 0;
  for (int t0 = first; t0 <= last; t0 += CONV_T) {
    const int t1 = (t0 + CONV_T - 1 < last) ? t0 + CONV_T - 1 : last;
    const int w_lo = b0 - t1; /* W index of ws[0] */
    for (int k = threadIdx.x; k <= t1 - t0; k += CB)
      hs[k] = hist[t0 + k];
    for (int k = threadIdx.x; k < t1 - t0 + CB; k += CB) {
      int w = w_lo + k;
      ws[k] = (w >= 0 && w < wp) ? W[w] : 0.0;
    }
    __syncthreads();
    if (ib < nb) {
      int lo = ib - wp + 1 > t0 ? ib - wp + 1 : t0;
      int hi = ib < t1 ? ib : t1;
      int j = lo;
      /* eight terms per step, software-pipelined: the next eight pairs are
       * loaded while the current eight are added, the additions stay in
       * order; a zero count adds +-0, which leaves the sum (never -0)
       * unchanged, exactly as the reference's unconditional loop */
      if (j + 7 <= hi) {
        double c[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          c[u] = (double)hs[j + u - t0];
          w[u] = ws[ib - (j + u) - w_lo];
        }
        for (; j + 15 <= hi; j += 8) {
          double c2[8], w2[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            c2[u] = (double)hs[j + 8 + u - t0];
            w2[u] = ws[ib - (j + 8 + u) - w_lo];
          }
#pragma unroll
          for (int u = 0; u < 8; u++)
            v += c[u] * w[u];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            c[u] = c2[u];
            w[u] = w2[u];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
          v += c[u] * w[u];
        j += 8;
      }
      for (; j <= hi; j++)
        v += (double)hs[j - t0] * ws[ib - j - w_lo];
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
 * (x[2j], x[2j+1]) plus the real split, twiddles cos/sin(2 pi k/nb). */
#define FFT_T 1024
__global__ void __launch_bounds__(FFT_T) zlongit_fft(const unsigned int *__restrict__ hist,
                                                     const double *__restrict__ Z, const double *__restrict__ tc,
                                                     const double *__restrict__ ts, int32_t nb, int32_t bits,
                                                     double factor, const coll_plan *plan, double *volt) {
  extern __shared__ double sm[];
  if (plan->nb <= 0 || plan->err)
    return;
  const int m = nb / 2;
  double *zr = sm, *zi = sm + m;
  const int tid = threadIdx.x;
  /* load with bit reversal: z[rev(j)] = (x[2j], x[2j+1]) */
  for (int j = tid; j < m; j += FFT_T) {
    int r = (int)(__brev((unsigned)j) >> (32 - bits));
    zr[r] = (double)hist[2 * j];
    zi[r] = (double)hist[2 * j + 1];
  }
  __syncthreads();
  for (int dir = 0; dir < 2; dir++) {
    for (int h = 1; h < m; h <<= 1) {
      const int stride = nb / (2 * h);
      for (int b = tid; b < m / 2; b += FFT_T) {
        int g = (b / h) * 2 * h, j = b % h;
        double c = tc[j * stride], s = ts[j * stride];
        double br = zr[g + j + h], bi = zi[g + j + h];
        double tr, ti;
        if (dir == 0) {
          tr = br * c + bi * s;
          ti = bi * c - br * s;
        } else {
          tr = br * c - bi * s;
          ti = bi * c + br * s;
        }
        double ar = zr[g + j], ai = zi[g + j];
        zr[g + j] = ar + tr;
        zi[g + j] = ai + ti;
        zr[g + j + h] = ar - tr;
        zi[g + j + h] = ai - ti;
      }
      __syncthreads();
    }
    if (dir == 1)
      break;
    /* real split -> packed spectrum in volt[], then V = Z I factor, then the
     * inverse pre-processing back into (zr, zi) with bit reversal */
    const double scale = 1.0 / nb;
    for (int k = tid; k <= m; k += FFT_T) {
      if (k == 0) {
        volt[0] = (zr[0] + zi[0]) * scale;
        volt[nb - 1] = (zr[0] - zi[0]) * scale;
      } else if (k < m) {
        double ar = zr[k], ai = zi[k], br = zr[m - k], bi = -zi[m - k];
        double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
        double dr = ar - br, di = ai - bi;
        double o_r = di * 0.5, o_i = -dr * 0.5;
        double c = tc[k], s = ts[k];
        volt[2 * k - 1] = (er + (c * o_r + s * o_i)) * scale;
        volt[2 * k] = (ei + (c * o_i - s * o_r)) * scale;
      }
    }
    __syncthreads();
    {
      /* V(f) = Z(f) I(f) factor, zlongit.c:379-391 (Ifreq read before any write) */
      double vr_[8], vi_[8];
      int cnt = 0;
      for (int k = tid; k < m; k += FFT_T, cnt++) {
        if (k == 0) {
          vr_[cnt] = volt[0] * Z[0] * factor;
          vi_[cnt] = volt[nb - 1] * Z[nb - 1] * factor;
        } else {
          int iR = 2 * k - 1, iI = iR + 1;
          vr_[cnt] = (volt[iR] * Z[iR] - volt[iI] * Z[iI]) * factor;
          vi_[cnt] = (volt[iR] * Z[iI] + volt[iI] * Z[iR]) * factor;
        }
      }
      __syncthreads();
      cnt = 0;
      for (int k = tid; k < m; k += FFT_T, cnt++) {
        if (k == 0) {
          volt[0] = vr_[cnt];
          volt[nb - 1] = vi_[cnt];
        } else {
          volt[2 * k - 1] = vr_[cnt];
          volt[2 * k] = vi_[cnt];
        }
      }
    }
    __syncthreads();
    for (int k = tid; k < m; k += FFT_T) {
      double zrk, zik;
      if (k == 0) {
        double x0 = volt[0], xm = volt[nb - 1];
        zrk = x0 + xm;
        zik = x0 - xm;
      } else {
        double ar = volt[2 * k - 1], ai = volt[2 * k];
        double br = volt[2 * (m - k) - 1], bi = -volt[2 * (m - k)];
        double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
        double c = tc[k], s = ts[k];
        zrk = sr - (c * di + s * dr);
        zik = si + (c * dr - s * di);
      }
      int r = (int)(__brev((unsigned)k) >> (32 - bits));
      zr[r] = zrk;
      zi[r] = zik;
    }
    __syncthreads();
  }
  for (int j = tid; j < m; j += FFT_T) {
    volt[2 * j] = zr[j];
    volt[2 * j + 1] = zi[j];
  }
  if (tid == 0)
    volt[nb] = 0;
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
 * (zlongit.c:454-486), then the post-element checks of do_tracking */
__global__ void __launch_bounds__(CB) kick(double *x, double *xp, double *y, double *yp, double *s, double *dp,
                                           int32_t *lp, int32_t *lc, int64_t n, double Po, const coll_plan *plan,
                                           double dt, const double *__restrict__ volt, int32_t kind,
                                           int32_t interpolate, double z_end, double ap_x, double ap_y,
                                           int32_t ap_on, int32_t ap_code, int64_t pass, int32_t code_base) {
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  const int32_t nb = plan->nb;
  const double tmin = plan->tmin;
  if (nb <= 0 || plan->err || i >= n || (lp && lp[i] >= 0))
    return;
  const double mc2 = 1e6 * EH_ME_MEV * EH_REL_SIGN;
  double S = s[i], D = dp[i];
  double t = arrival_time(S, D, Po);
  int64_t ib = bin_of(t, tmin, dt, kind == EH_OP_WAKE);
  double XP = xp[i], YP = yp[i];
  bool changed = false;
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
  if (slot < 0 || slot >= EHD_MAX_COLL || nb < 4 || nb > 8192 || (nb & (nb - 1)))
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
 * the voltage (double x (cap+8)), the fixed-sum partials */
#define PLAN_BYTES 128
static int64_t hist_bytes(int64_t nb_cap) { return ((4 * (nb_cap + 8) + 7) / 8) * 8; }

static int ensure_work(ehd_dev *d, int64_t nb_cap, int64_t nparts) {
  int64_t need = PLAN_BYTES + hist_bytes(nb_cap) + 8 * (nb_cap + 8) + (int64_t)sizeof(eh_fsum) * (nparts + 1) + 256;
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
  CHK(hipEventRecord(d->cev0, d->stream));
  hipLaunchKernelGGL(time_minmax, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, crd[4], crd[5], lp, n, c->Po,
                     (unsigned long long *)parts);
  hipLaunchKernelGGL(coll_plan_minmax, dim3(1), dim3(CB), 0, d->stream, plan, (const unsigned long long *)parts, rgrid,
                     kind, fixed, c->nb, c->dt, c->wake_span, c->allow_long_beam, (int32_t)nb_max, hist, nclear);
  if (kind == EH_OP_ZLONGIT || fixed) {
    hipLaunchKernelGGL(time_fsum, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, crd[4], crd[5], lp, n, c->Po, plan,
                       parts);
    hipLaunchKernelGGL(fsum_final, dim3(1), dim3(CB), 0, d->stream, parts, rgrid, plan, c->dt);
  }
  {
    /* one workgroup per CU: each sees enough particles that its LDS counts
     * aggregate many hits per bin before the one global atomic per bin */
    if (!d->num_cu) {
      int cu = 0;
      CHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d->device));
      d->num_cu = cu > 0 ? cu : 256;
    }
    int hgrid = (int)((n + CB - 1) / CB);
    if (hgrid < 1)
      hgrid = 1;
    if (hgrid > d->num_cu)
      hgrid = d->num_cu;
    /* room for one copy of the counts per wave, at most 48 KB of LDS */
    const int32_t lds_bins = nb_max * (CB / 64) < 12288 ? (int32_t)(nb_max * (CB / 64)) : 12288;
    hipLaunchKernelGGL(histogram, dim3(hgrid), dim3(CB), sizeof(unsigned int) * lds_bins, d->stream, crd[4], crd[5],
                       lp, n, c->Po, plan, c->dt, kind == EH_OP_WAKE ? 1 : 0, lds_bins, hist);
  }
  if (kind == EH_OP_WAKE) {
    hipLaunchKernelGGL(wake_convolve, dim3((unsigned)((nb_max + 1 + CB - 1) / CB)), dim3(CB), 0, d->stream, hist, W,
                       wpoints, plan, c->factor, volt);
  } else {
    int bits = 0;
    while ((1 << bits) < c->nb / 2)
      bits++;
    hipLaunchKernelGGL(zlongit_fft, dim3(1), dim3(FFT_T), sizeof(double) * c->nb, d->stream, hist, sl->Z, sl->tc,
                       sl->ts, c->nb, bits, c->factor, plan, volt);
  }
  if (n > 0)
    hipLaunchKernelGGL(kick, dim3((unsigned)grid), dim3(CB), 0, d->stream, crd[0], crd[1], crd[2], crd[3], crd[4],
                       crd[5], lp, lc, n, c->Po, plan, c->dt, volt, kind, c->interpolate, c->z_end, c->ap_xmax,
                       c->ap_ymax, ap_on, c->ap_code, c->pass, 4 * c->op_index);
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
  int rc = ensure_work(d, nb_cap, reduce_grid(n));
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

int ehd_time_span(ehd_dev *d, double Po, double *tmin, double *tmax, int64_t *count) {
  const int64_t n = d->n;
  const int64_t rgrid = reduce_grid(n);
  CHK(hipSetDevice(d->device));
  int rc = ensure_work(d, 64, rgrid);
  if (rc)
    return rc;
  coll_plan *plan = (coll_plan *)d->cwork;
  unsigned int *hist = (unsigned int *)((char *)d->cwork + PLAN_BYTES);
  unsigned long long *mm = (unsigned long long *)((char *)hist + hist_bytes(64) + 8 * (64 + 8));
  hipLaunchKernelGGL(time_minmax, dim3((unsigned)rgrid), dim3(CB), 0, d->stream, d->coord[4], d->coord[5],
                     d->loss_pass, n, Po, mm);
  hipLaunchKernelGGL(coll_plan_minmax, dim3(1), dim3(CB), 0, d->stream, plan, (const unsigned long long *)mm, rgrid,
                     -1, 0, 0, 1.0, 0.0, 0, 0, hist, 0);
  CHK(hipGetLastError());
  coll_plan hp;
  CHK(hipMemcpyAsync(&hp, plan, sizeof(hp), hipMemcpyDeviceToHost, d->stream));
  CHK(hipStreamSynchronize(d->stream));
  *count = (int64_t)hp.red[2];
  *tmin = key_value(hp.red[0]);
  *tmax = key_value(hp.red[1]);
  return EH_OK;
}

int ehd_collective_timing(ehd_dev *d, int64_t *launches, double *ms) {
  *launches = d->coll_launches;
  *ms = d->coll_ms;
  return EH_OK;
}
}
