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

/* red[0] = min key, red[1] = max key, red[2] = alive count */
__global__ void __launch_bounds__(CB) time_minmax(const double *__restrict__ s, const double *__restrict__ dp,
                                                  const int32_t *__restrict__ lp, int64_t n, double Po,
                                                  unsigned long long *red) {
  __shared__ unsigned long long smin[CB / 64], smax[CB / 64], scnt[CB / 64];
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  unsigned long long kmin = ~0ULL, kmax = 0, cnt = 0;
  if (i < n && lp[i] < 0) {
    unsigned long long k = order_key(arrival_time(s[i], dp[i], Po));
    kmin = kmax = k;
    cnt = 1;
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
    if (cnt) {
      atomicMin(&red[0], kmin);
      atomicMax(&red[1], kmax);
      atomicAdd(&red[2], cnt);
    }
  }
}

/* per-workgroup fixed-point partial sums of the arrival times */
__global__ void __launch_bounds__(CB) time_fsum(const double *__restrict__ s, const double *__restrict__ dp,
                                                const int32_t *__restrict__ lp, int64_t n, double Po, int32_t e,
                                                eh_fsum *partial) {
  __shared__ eh_fsum sh[CB];
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  eh_fsum a;
  eh_fsum_init(&a, e);
  if (i < n && lp[i] < 0)
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

__global__ void __launch_bounds__(CB) fsum_final(eh_fsum *partial, int64_t nparts, int32_t e, double count,
                                                 double *mean_out) {
  __shared__ eh_fsum sh[CB];
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
  if (threadIdx.x == 0)
    mean_out[0] = eh_fsum_to_double(&sh[0]) / count;
}

/* bin index of wake.c:507 (round) or zlongit.c:267 (truncate) */
DEV int64_t bin_of(double t, double tmin, double dt, int round_half) {
  double v = round_half ? (t - tmin) / dt + 0.5 : (t - tmin) / dt;
  return (int64_t)v;
}

__global__ void __launch_bounds__(CB) histogram(const double *__restrict__ s, const double *__restrict__ dp,
                                                const int32_t *__restrict__ lp, int64_t n, double Po, double tmin,
                                                double dt, int32_t nb, int32_t round_half, unsigned int *hist) {
  extern __shared__ unsigned int lh[];
  for (int k = threadIdx.x; k < nb; k += CB)
    lh[k] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB) {
    if (lp[i] >= 0)
      continue;
    int64_t ib = bin_of(arrival_time(s[i], dp[i], Po), tmin, dt, round_half);
    if (ib >= 0 && ib <= nb - 1)
      atomicAdd(&lh[ib], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nb; k += CB)
    if (lh[k])
      atomicAdd(&hist[k], lh[k]);
}

/* convolveArrays with di2 = 0 (wake.c:479-495): V[ib] = sum_{j=0..ib} I[j] W[ib-j],
 * summed in increasing j; zero counts are skipped (adding 0*W = +-0 to a
 * sum that starts at +0 leaves it unchanged), then V *= factor (wake.c:258-260) */
__global__ void __launch_bounds__(CB) wake_convolve(const unsigned int *__restrict__ hist,
                                                    const double *__restrict__ W, int64_t wpoints, int32_t nb,
                                                    double factor, double *volt) {
  int ib = blockIdx.x * CB + threadIdx.x;
  if (ib >= nb)
    return;
  int64_t ib1 = 0, ib2 = ib;
  if (ib2 >= wpoints) {
    int64_t di = ib2 - wpoints + 1;
    ib1 += di;
    ib2 -= di;
  }
  double v = 0;
  for (; ib1 < nb && ib2 >= 0; ib1++, ib2--) {
    unsigned int c = hist[ib1];
    if (c)
      v += (double)c * W[ib2];
  }
  volt[ib] = v * factor;
}

/* packed real FFT, V = Z I factor, inverse FFT (zlongit.c:373-394), for
 * nb = 2m <= 8192 in one workgroup.  Same algorithm and operation order as
 * or_real_fft in the CPU checker: m-point radix-2 DIT complex FFT of
 * (x[2j], x[2j+1]) plus the real split, twiddles cos/sin(2 pi k/nb). */
#define FFT_T 1024
__global__ void __launch_bounds__(FFT_T) zlongit_fft(const unsigned int *__restrict__ hist,
                                                     const double *__restrict__ Z, const double *__restrict__ tc,
                                                     const double *__restrict__ ts, int32_t nb, int32_t bits,
                                                     double factor, double *volt) {
  extern __shared__ double sm[];
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
                                           int32_t *lp, int32_t *lc, int64_t n, double Po, double tmin, double dt,
                                           int32_t nb, const double *__restrict__ volt, int32_t kind,
                                           int32_t interpolate, double z_end, double ap_x, double ap_y,
                                           int32_t ap_on, int64_t pass, int32_t code_base) {
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (i >= n || lp[i] >= 0)
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

void ehd_free_collective(ehd_dev *d) {
  for (int k = 0; k < EHD_MAX_COLL; k++) {
    ehd_coll_slot *sl = &d->slot[k];
    if (sl->W)
      hipFree(sl->W);
    if (sl->Z)
      hipFree(sl->Z);
    if (sl->tc)
      hipFree(sl->tc);
    if (sl->ts)
      hipFree(sl->ts);
    memset(sl, 0, sizeof(*sl));
  }
  if (d->cwork)
    hipFree(d->cwork);
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

/* work area: [0,64) reductions, then histogram (uint32 x cap), voltage
 * (double x (cap+1)), fixed-sum partials */
static int ensure_work(ehd_dev *d, int64_t nb_cap, int64_t nparts) {
  int64_t need = 64 + 8 * (nb_cap + 8) + 8 * (nb_cap + 8) + (int64_t)sizeof(eh_fsum) * (nparts + 1) + 256;
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

int ehd_collective_step(ehd_dev *d, int32_t slot, const ehd_collective *c, double *hist_out, double *volt_out,
                        int64_t hist_cap, int64_t *nb_out, double *tmin_out) {
  (void)hist_cap;
  if (slot < 0 || slot >= EHD_MAX_COLL)
    return EH_ERR_INVALID;
  ehd_coll_slot *sl = &d->slot[slot];
  const int64_t n = d->n;
  const int64_t grid = (n + CB - 1) / CB;
  CHK(hipSetDevice(d->device));
  int64_t nb_cap = c->kind == EH_OP_WAKE ? sl->wpoints + 8 : sl->nb;
  int rc = ensure_work(d, nb_cap, grid);
  if (rc)
    return rc;
  unsigned long long *red = (unsigned long long *)d->cwork;
  unsigned int *hist = (unsigned int *)((char *)d->cwork + 64);
  double *volt = (double *)((char *)d->cwork + 64 + 8 * (nb_cap + 8));
  eh_fsum *parts = (eh_fsum *)((char *)volt + 8 * (nb_cap + 8));
  double *mean_dev = (double *)((char *)d->cwork + 32);
  unsigned long long init[4] = {~0ULL, 0ULL, 0ULL, 0ULL};
  unsigned long long hred[4];
  CHK(hipEventRecord(d->cev0, d->stream));
  CHK(hipMemcpyAsync(red, init, sizeof(init), hipMemcpyHostToDevice, d->stream));
  if (n > 0)
    hipLaunchKernelGGL(time_minmax, dim3((unsigned)grid), dim3(CB), 0, d->stream, d->coord[4], d->coord[5],
                       d->loss_pass, n, c->Po, red);
  CHK(hipGetLastError());
  CHK(hipMemcpyAsync(hred, red, sizeof(hred), hipMemcpyDeviceToHost, d->stream));
  CHK(hipStreamSynchronize(d->stream));
  int64_t count = (int64_t)hred[2];
  if (count == 0) {
    if (nb_out)
      *nb_out = 0;
    return EH_OK;
  }
  double tmin = key_value(hred[0]), tmax = key_value(hred[1]);
  double dt = c->dt, tmean = 0;
  int32_t nb;
  if (c->kind == EH_OP_ZLONGIT || c->n_bins_fixed) {
    double am = __builtin_fabs(tmin) > __builtin_fabs(tmax) ? __builtin_fabs(tmin) : __builtin_fabs(tmax);
    int32_t e = eh_fsum_exponent(am);
    hipLaunchKernelGGL(time_fsum, dim3((unsigned)grid), dim3(CB), 0, d->stream, d->coord[4], d->coord[5],
                       d->loss_pass, n, c->Po, e, parts);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(fsum_final, dim3(1), dim3(CB), 0, d->stream, parts, grid, e, (double)count, mean_dev);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(&tmean, mean_dev, sizeof(double), hipMemcpyDeviceToHost, d->stream));
    CHK(hipStreamSynchronize(d->stream));
  }
  if (c->kind == EH_OP_WAKE) {
    if ((tmax - tmin) > c->wake_span && !c->allow_long_beam)
      return EH_ERR_LONGBEAM;
    if (c->n_bins_fixed) {
      nb = c->n_bins_fixed;
      tmin = tmean - dt * nb / 2.0;
    } else {
      nb = (int32_t)(int64_t)((tmax - tmin) / dt + 3);
      tmin -= dt;
    }
    if (nb <= 0)
      return EH_OK;
    if (nb > nb_cap - 8)
      return EH_ERR_LONGBEAM;
  } else {
    nb = c->nb;
    if ((tmax - tmin) * 2 > nb * dt && !c->allow_long_beam)
      return EH_ERR_LONGBEAM;
    tmin = tmean - dt * nb / 2.0;
  }
  CHK(hipMemsetAsync(hist, 0, sizeof(unsigned int) * (nb + 8), d->stream));
  {
    int hgrid = (int)((n + CB * 8 - 1) / (CB * 8));
    if (hgrid < 1)
      hgrid = 1;
    if (hgrid > 2048)
      hgrid = 2048;
    hipLaunchKernelGGL(histogram, dim3(hgrid), dim3(CB), sizeof(unsigned int) * nb, d->stream, d->coord[4],
                       d->coord[5], d->loss_pass, n, c->Po, tmin, dt, nb, c->kind == EH_OP_WAKE ? 1 : 0, hist);
    CHK(hipGetLastError());
  }
  if (c->kind == EH_OP_WAKE) {
    hipLaunchKernelGGL(wake_convolve, dim3((nb + CB - 1) / CB), dim3(CB), 0, d->stream, hist, sl->W, sl->wpoints,
                       nb, c->factor, volt);
    CHK(hipGetLastError());
    CHK(hipMemsetAsync(volt + nb, 0, sizeof(double), d->stream));
  } else {
    int bits = 0;
    while ((1 << bits) < nb / 2)
      bits++;
    hipLaunchKernelGGL(zlongit_fft, dim3(1), dim3(FFT_T), sizeof(double) * nb, d->stream, hist, sl->Z, sl->tc,
                       sl->ts, nb, bits, c->factor, volt);
    CHK(hipGetLastError());
  }
  hipLaunchKernelGGL(kick, dim3((unsigned)grid), dim3(CB), 0, d->stream, d->coord[0], d->coord[1], d->coord[2],
                     d->coord[3], d->coord[4], d->coord[5], d->loss_pass, d->loss_code, n, c->Po, tmin, dt, nb, volt,
                     c->kind, c->interpolate, c->z_end, c->ap_xmax, c->ap_ymax,
                     (c->flags & EHD_COLL_NOCHECKS) ? -1 : ((c->flags & EH_F_AP) ? 1 : 0), c->pass, 4 * c->op_index);
  CHK(hipGetLastError());
  CHK(hipEventRecord(d->cev1, d->stream));
  if (hist_out || volt_out) {
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
  CHK(hipEventSynchronize(d->cev1));
  float ms = 0;
  CHK(hipEventElapsedTime(&ms, d->cev0, d->cev1));
  d->coll_ms += ms;
  d->coll_launches += 1;
  if (nb_out)
    *nb_out = nb;
  if (tmin_out)
    *tmin_out = tmin;
  return EH_OK;
}

int ehd_time_span(ehd_dev *d, double Po, double *tmin, double *tmax, int64_t *count) {
  const int64_t n = d->n;
  const int64_t grid = (n + CB - 1) / CB;
  CHK(hipSetDevice(d->device));
  int rc = ensure_work(d, 64, grid);
  if (rc)
    return rc;
  unsigned long long *red = (unsigned long long *)d->cwork;
  unsigned long long init[4] = {~0ULL, 0ULL, 0ULL, 0ULL}, hred[4];
  CHK(hipMemcpyAsync(red, init, sizeof(init), hipMemcpyHostToDevice, d->stream));
  if (n > 0)
    hipLaunchKernelGGL(time_minmax, dim3((unsigned)grid), dim3(CB), 0, d->stream, d->coord[4], d->coord[5],
                       d->loss_pass, n, Po, red);
  CHK(hipGetLastError());
  CHK(hipMemcpyAsync(hred, red, sizeof(hred), hipMemcpyDeviceToHost, d->stream));
  CHK(hipStreamSynchronize(d->stream));
  *count = (int64_t)hred[2];
  *tmin = key_value(hred[0]);
  *tmax = key_value(hred[1]);
  return EH_OK;
}

int ehd_collective_timing(ehd_dev *d, int64_t *launches, double *ms) {
  *launches = d->coll_launches;
  *ms = d->coll_ms;
  return EH_OK;
}
}
