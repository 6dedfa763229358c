/*
 * emath.h - deterministic FP64 elementary functions shared by the gfx950
 * kernels and the CPU checker build.
 *
 * Why this exists: the per-particle integrators call sin, cos, asin, tan and
 * log (csbend.c:1567-1574, simple_rfca.c:846-879).  The device math library
 * and glibc differ by up to one ULP on some arguments, so a kernel that calls
 * the device library can never be compared bit-for-bit against a CPU run.
 * Every function here is written only with IEEE-754 operations that are
 * correctly rounded on both x86-64 and CDNA4 (+, -, *, /, sqrt, explicit fma)
 * and is compiled with -ffp-contract=off on both sides, so the same source
 * gives the same bits on the host and on the GPU.
 *
 * Methods (all public, none taken from a vendor library):
 *  - argument reduction modulo pi/2: Cody & Waite (1980) with pi/2 split into
 *    seven 24-bit pieces, accumulated in double-double (Knuth TwoSum), valid
 *    for |x| < 2^28; larger |x| returns NaN (documented domain limit);
 *  - sin/cos on [-pi/4, pi/4]: Taylor series to degree 17/16 (coefficients
 *    are 1/n! rounded to double), shortened to degree 9/8 for |x| < 2^-5,
 *    with the compensated 1 - z/2 step for cos;
 *  - asin: Taylor series of asin in x^2 on |x| <= 0.5 (5, 9 or 25 terms by
 *    band) and the reflection asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above;
 *  - every polynomial by Horner's rule in fused multiply-adds (correctly
 *    rounded on both sides: the C library's fma on the host, v_fma_f64 on
 *    the device);
 *  - log: x = 2^e m, log(m) = 2 atanh(f/(2+f)) with f = m - 1.
 * Constants were derived from exact rational arithmetic (pi from Machin's
 * formula, ln 2 from sum 1/(k 2^k)) and are written as hex literals.
 */
#ifndef ELEGANT_AMD_EMATH_H
#define ELEGANT_AMD_EMATH_H

#include <stdint.h>

#ifdef __HIPCC__
#define EM_FN static inline __host__ __device__
#else
#define EM_FN static inline
#endif

#define EM_PIO4 0x1.921fb54442d18p-1
#define EM_PIO2_HI 0x1.921fb54442d18p+0
#define EM_PIO2_LO 0x1.1a62633145c07p-54
#define EM_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define EM_SHIFT 0x1.8p+52
#define EM_LN2_HI 0x1.62e42fee00000p-1
#define EM_LN2_LO 0x1.a39ef35600000p-33

/* A polynomial coefficient at its point of use.  On the device the empty
 * volatile asm keeps each 64-bit constant a scalar-register operand
 * materialized where it is used; without it the compiler hoists every
 * coefficient of every inlined polynomial into vector registers for the
 * whole kernel, which costs occupancy.  The value is unchanged. */
#if defined(__HIP_DEVICE_COMPILE__)
EM_FN double em_k(double c) {
  asm volatile("" : "=s"(c) : "0"(c));
  return c;
}
#else
EM_FN double em_k(double c) { return c; }
#endif

EM_FN uint64_t em_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, sizeof u);
  return u;
}

EM_FN double em_from_bits(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, sizeof d);
  return d;
}

EM_FN double em_fabs(double x) { return em_from_bits(em_bits(x) & 0x7fffffffffffffffULL); }

EM_FN double em_nan(void) { return em_from_bits(0x7ff8000000000000ULL); }

/* a - b as an unevaluated sum s + e (Knuth's TwoSum, no ordering needed). */
EM_FN void em_two_diff(double a, double b, double *s, double *e) {
  double t = a - b;
  double bb = t - a;
  *e = (a - (t - bb)) - (b + bb);
  *s = t;
}

/* x = k*(pi/2) + (rh + rl), |rh| <= ~pi/4.  Returns k (mod 4 is what matters).
 * Returns -1 with *rh = NaN if x is NaN, infinite or |x| >= 2^28. */
EM_FN int em_reduce_pio2(double x, double *rh, double *rl) {
  const double c1 = 0x1.921fb40000000p+0, c2 = 0x1.4442d00000000p-24, c3 = 0x1.8469880000000p-48,
               c4 = 0x1.8cc5160000000p-72, c5 = 0x1.01b8380000000p-96, c6 = 0x1.a252040000000p-120,
               c7 = 0x1.3822280000000p-145;
  double ax = em_fabs(x);
  if (ax <= EM_PIO4) {
    *rh = x;
    *rl = 0.0;
    return 0;
  }
  if (!(ax < 0x1.0p+28)) {
    *rh = em_nan();
    *rl = 0.0;
    return -1;
  }
  /* round x*2/pi to the nearest integer (|k| < 2^28, so k*ci is exact) */
  double kd = (x * EM_TWO_OVER_PI + EM_SHIFT) - EM_SHIFT;
  double hi = x - kd * c1; /* exact: x and k*c1 agree to within a factor 2 */
  double lo = 0.0, e, s;
  em_two_diff(hi, kd * c2, &s, &e);
  hi = s;
  lo += e;
  em_two_diff(hi, kd * c3, &s, &e);
  hi = s;
  lo += e;
  em_two_diff(hi, kd * c4, &s, &e);
  hi = s;
  lo += e;
  lo -= kd * c5;
  lo -= kd * c6;
  lo -= kd * c7;
  s = hi + lo;
  *rl = lo - (s - hi);
  *rh = s;
  return (int)(long long)kd;
}

/* Polynomial cores, evaluated by Horner's rule in fused multiply-adds.  The
 * band |x| < 2^-5 keeps the series through x^9 (sin) and x^8 (cos): the
 * first dropped term is below 2^-64 relative, so both bands meet the same
 * accuracy while the narrow one costs half the operations. */
#define EM_S1 (-0x1.5555555555555p-3)
#define EM_SMALL_Z 0x1.0p-10

/* (sin(x) - x - S1 x^3) / x^5 as a polynomial in z = x*x; the narrow band
 * is its own function so that branch-free callers evaluate the same bits */
EM_FN double em_sin_r_small(double z) {
  const double s2 = 0x1.1111111111111p-7, s3 = -0x1.a01a01a01a01ap-13, s4 = 0x1.71de3a556c734p-19;
  return __builtin_fma(z, __builtin_fma(z, em_k(s4), em_k(s3)), em_k(s2));
}

EM_FN double em_sin_r(double z) {
  const double s2 = 0x1.1111111111111p-7, s3 = -0x1.a01a01a01a01ap-13, s4 = 0x1.71de3a556c734p-19,
               s5 = -0x1.ae64567f544e4p-26, s6 = 0x1.6124613a86d09p-33, s7 = -0x1.ae7f3e733b81fp-41,
               s8 = 0x1.952c77030ad4ap-49;
  if (z < EM_SMALL_Z)
    return em_sin_r_small(z);
  double r = __builtin_fma(z, em_k(s8), em_k(s7));
  r = __builtin_fma(z, r, em_k(s6));
  r = __builtin_fma(z, r, em_k(s5));
  r = __builtin_fma(z, r, em_k(s4));
  r = __builtin_fma(z, r, em_k(s3));
  return __builtin_fma(z, r, em_k(s2));
}

/* (cos(x) - 1 + z/2) / z^2 as a polynomial in z = x*x */
EM_FN double em_cos_r_small(double z) {
  const double c2 = 0x1.5555555555555p-5, c3 = -0x1.6c16c16c16c17p-10, c4 = 0x1.a01a01a01a01ap-16;
  return __builtin_fma(z, __builtin_fma(z, em_k(c4), em_k(c3)), em_k(c2));
}

EM_FN double em_cos_r(double z) {
  const double c2 = 0x1.5555555555555p-5, c3 = -0x1.6c16c16c16c17p-10, c4 = 0x1.a01a01a01a01ap-16,
               c5 = -0x1.27e4fb7789f5cp-22, c6 = 0x1.1eed8eff8d898p-29, c7 = -0x1.93974a8c07c9dp-37,
               c8 = 0x1.ae7f3e733b81fp-45;
  if (z < EM_SMALL_Z)
    return em_cos_r_small(z);
  double r = __builtin_fma(z, em_k(c8), em_k(c7));
  r = __builtin_fma(z, r, em_k(c6));
  r = __builtin_fma(z, r, em_k(c5));
  r = __builtin_fma(z, r, em_k(c4));
  r = __builtin_fma(z, r, em_k(c3));
  return __builtin_fma(z, r, em_k(c2));
}

/* sin(x + y) for |x| <= ~pi/4, |y| <= 2^-50 |x| */
EM_FN double em_sin_kernel(double x, double y) {
  double z = x * x;
  double v = z * x;
  return x + ((y - 0.5 * z * y) + v * __builtin_fma(z, em_sin_r(z), em_k(EM_S1)));
}

/* cos(x + y) for |x| <= ~pi/4, |y| <= 2^-50 |x|; 1 - z/2 is split so that
 * its rounding error is carried into the tail */
EM_FN double em_cos_kernel(double x, double y) {
  double z = x * x;
  double r = z * z * em_cos_r(z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (r - x * y));
}

/* the same kernels with y = 0, for arguments that need no reduction */
EM_FN double em_sin_kernel0(double x) {
  double z = x * x;
  return x + (z * x) * __builtin_fma(z, em_sin_r(z), em_k(EM_S1));
}

EM_FN double em_cos_kernel0(double x) {
  double z = x * x;
  double r = z * z * em_cos_r(z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + r);
}

/* the kernels on the narrow band, z = x*x < 2^-10: em_sin(x) for x != 0 and
 * em_cos(x) there, term for term */
EM_FN double em_sin_small(double x, double z) {
  return x + (z * x) * __builtin_fma(z, em_sin_r_small(z), em_k(EM_S1));
}

EM_FN double em_cos_small(double x, double z) {
  double r = z * z * em_cos_r_small(z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + r);
}

EM_FN void em_sincos(double x, double *sp, double *cp) {
  double rh, rl, s, c;
  int k;
  if (x == 0.0) {
    *sp = x;
    *cp = 1.0;
    return;
  }
  if (em_fabs(x) <= EM_PIO4) {
    *sp = em_sin_kernel0(x);
    *cp = em_cos_kernel0(x);
    return;
  }
  k = em_reduce_pio2(x, &rh, &rl);
  if (k < 0 && rh != rh) {
    *sp = *cp = em_nan();
    return;
  }
  s = em_sin_kernel(rh, rl);
  c = em_cos_kernel(rh, rl);
  switch (k & 3) {
  case 0:
    *sp = s;
    *cp = c;
    break;
  case 1:
    *sp = c;
    *cp = -s;
    break;
  case 2:
    *sp = -s;
    *cp = -c;
    break;
  default:
    *sp = -c;
    *cp = s;
    break;
  }
}

EM_FN double em_sin(double x) {
  double rh, rl, v;
  int k;
  if (x == 0.0)
    return x;
  if (em_fabs(x) <= EM_PIO4)
    return em_sin_kernel0(x);
  k = em_reduce_pio2(x, &rh, &rl);
  if (k < 0 && rh != rh)
    return em_nan();
  v = (k & 1) ? em_cos_kernel(rh, rl) : em_sin_kernel(rh, rl);
  return (k & 2) ? -v : v;
}

EM_FN double em_cos(double x) {
  double rh, rl, v;
  int k;
  if (em_fabs(x) <= EM_PIO4)
    return x == 0.0 ? 1.0 : em_cos_kernel0(x);
  k = em_reduce_pio2(x, &rh, &rl);
  if (k < 0 && rh != rh)
    return em_nan();
  v = (k & 1) ? em_sin_kernel(rh, rl) : em_cos_kernel(rh, rl);
  return ((k + 1) & 2) ? -v : v;
}

EM_FN double em_tan(double x) {
  double s, c;
  em_sincos(x, &s, &c);
  return s / c;
}

/* the five-term band of em_asin_poly_small */
EM_FN double em_asin_poly5(double z) {
  const double a1 = 0x1.5555555555555p-3, a2 = 0x1.3333333333333p-4, a3 = 0x1.6db6db6db6db7p-5,
               a4 = 0x1.f1c71c71c71c7p-6, a5 = 0x1.6e8ba2e8ba2e9p-6;
  double p = __builtin_fma(z, em_k(a5), em_k(a4));
  p = __builtin_fma(z, p, em_k(a3));
  p = __builtin_fma(z, p, em_k(a2));
  p = __builtin_fma(z, p, em_k(a1));
  return z * p;
}

/* (asin(x) - x)/x as a polynomial in z = x*x: five terms for z < 2^-10
 * (next term below 2^-65 relative), nine for z <= 1/64 */
EM_FN double em_asin_poly_small(double z) {
  const double a1 = 0x1.5555555555555p-3, a2 = 0x1.3333333333333p-4, a3 = 0x1.6db6db6db6db7p-5,
               a4 = 0x1.f1c71c71c71c7p-6, a5 = 0x1.6e8ba2e8ba2e9p-6, a6 = 0x1.1c4ec4ec4ec4fp-6,
               a7 = 0x1.c99999999999ap-7, a8 = 0x1.7a87878787878p-7, a9 = 0x1.3fde50d79435ep-7;
  double p;
  if (z < EM_SMALL_Z)
    return em_asin_poly5(z);
  p = __builtin_fma(z, em_k(a9), em_k(a8));
  p = __builtin_fma(z, p, em_k(a7));
  p = __builtin_fma(z, p, em_k(a6));
  p = __builtin_fma(z, p, em_k(a5));
  p = __builtin_fma(z, p, em_k(a4));
  p = __builtin_fma(z, p, em_k(a3));
  p = __builtin_fma(z, p, em_k(a2));
  p = __builtin_fma(z, p, em_k(a1));
  return z * p;
}

/* em_asin(x) for z = x*x < 2^-10 */
EM_FN double em_asin_small(double x, double z) { return __builtin_fma(x, em_asin_poly5(z), x); }

/* same series, long form for z <= 1/4 (25 terms) */
EM_FN double em_asin_poly_large(double z) {
  const double a[25] = {0x1.5555555555555p-3, 0x1.3333333333333p-4, 0x1.6db6db6db6db7p-5, 0x1.f1c71c71c71c7p-6,
                        0x1.6e8ba2e8ba2e9p-6, 0x1.1c4ec4ec4ec4fp-6, 0x1.c99999999999ap-7, 0x1.7a87878787878p-7,
                        0x1.3fde50d79435ep-7, 0x1.12ef3cf3cf3cfp-7, 0x1.df3bd37a6f4dfp-8, 0x1.a6863d70a3d71p-8,
                        0x1.782dda12f684cp-8, 0x1.51ba308d3dcb1p-8, 0x1.31683bdef7bdfp-8, 0x1.15ee9d45d1746p-8,
                        0x1.fcaf8fb6db6dbp-9, 0x1.d3d2a8e0dd67dp-9, 0x1.b026f57b13b14p-9, 0x1.90cb77f60c7cep-9,
                        0x1.750de64d7d05fp-9, 0x1.5c5f56efaaaabp-9, 0x1.464c0950f7d47p-9, 0x1.3275586c5f2f0p-9,
                        0x1.208d3570ae5a6p-9};
  double p = em_k(a[24]);
  for (int i = 23; i >= 0; i--)
    p = __builtin_fma(z, p, em_k(a[i]));
  return z * p;
}

EM_FN double em_asin(double x) {
  double ax = em_fabs(x);
  if (!(ax <= 1.0))
    return em_nan();
  if (ax <= 0.125)
    return __builtin_fma(x, em_asin_poly_small(x * x), x);
  if (ax <= 0.5)
    return __builtin_fma(x, em_asin_poly_large(x * x), x);
  /* asin(ax) = pi/2 - 2 asin(sqrt(t)), t = (1 - ax)/2 exact */
  double t = (1.0 - ax) * 0.5;
  double s = __builtin_sqrt(t);
  double corr = s > 0 ? __builtin_fma(-s, s, t) / (s + s) : 0.0; /* sqrt(t) = s + corr to ~2^-105 */
  double p = em_asin_poly_large(t);
  double r = EM_PIO2_HI - (2.0 * s + (2.0 * corr + 2.0 * s * p - EM_PIO2_LO));
  return x < 0 ? -r : r;
}

EM_FN double em_log(double x) {
  const double l1 = 0x1.5555555555555p-1, l2 = 0x1.999999999999ap-2, l3 = 0x1.2492492492492p-2,
               l4 = 0x1.c71c71c71c71cp-3, l5 = 0x1.745d1745d1746p-3, l6 = 0x1.3b13b13b13b14p-3,
               l7 = 0x1.1111111111111p-3, l8 = 0x1.e1e1e1e1e1e1ep-4, l9 = 0x1.af286bca1af28p-4,
               l10 = 0x1.8618618618618p-4, l11 = 0x1.642c8590b2164p-4, l12 = 0x1.47ae147ae147bp-4;
  uint64_t u;
  int e;
  double m, f, s, z, r, hfsq, ed;
  if (x != x)
    return x;
  if (x < 0.0)
    return em_nan();
  if (x == 0.0)
    return -em_from_bits(0x7ff0000000000000ULL);
  if (x == em_from_bits(0x7ff0000000000000ULL))
    return x;
  e = 0;
  if (x < 0x1.0p-1022) { /* subnormal: scale into the normal range */
    x *= 0x1.0p+54;
    e = -54;
  }
  u = em_bits(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  m = em_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1, 2) */
  if (m > 0x1.6a09e667f3bcdp+0) {
    m *= 0.5;
    e += 1;
  }
  f = m - 1.0; /* exact */
  s = f / (2.0 + f);
  z = s * s;
  r = z * (l1 + z * (l2 + z * (l3 + z * (l4 + z * (l5 + z * (l6 + z * (l7 + z * (l8 + z * (l9 + z * (l10 + z * (l11 + z * l12)))))))))));
  hfsq = 0.5 * f * f;
  ed = (double)e;
  return ed * EM_LN2_HI + (f - (hfsq - (s * (hfsq + r) + ed * EM_LN2_LO)));
}

/* ---- division and square root without range scaling ----
 *
 * The device's correctly rounded FP64 division a/b is
 *   y = rcp(b); e = fma(-b, y, 1); y = fma(y, e, y); e = fma(-b, y, 1);
 *   y = fma(y, e, y); q = a*y; r = fma(-b, q, a); q = fma(r, y, q)
 * (v_rcp_f64, Newton steps, one residual correction), run on operands that
 * v_div_scale_f64 multiplies by a power of two near the ends of the exponent
 * range, followed by v_div_fixup_f64 for zeros, infinities and NaN.  For
 * 2^-64 <= |b| <= 2^64 and 2^-600 <= |a| <= 2^600 no step overflows or
 * underflows, so the scaling cannot change a rounding and the unscaled
 * sequence gives the same bits.  Its reciprocal part depends on b alone:
 * em_recip(b) computes it once for several divisions by the same b,
 * em_recip_ok(b) says whether b is in range, and em_div(a, b, y, ok) runs
 * the quotient part when a is in range too and the full division otherwise
 * (which also keeps the sign of a zero quotient and every special value).
 * The square root likewise: the device's sequence is v_rsq_f64 and Newton-
 * Goldschmidt steps on an argument scaled by 2^256 below 2^-767, then a
 * select for +-0 and +inf; em_sqrt skips the scaling and the select for
 * 2^-700 <= x <= 2^1000.  On the host these are the C operators, which give
 * the same correctly rounded bits. */
EM_FN int em_recip_ok(double b) {
  double ab = em_fabs(b);
  return ab >= 0x1.0p-64 && ab <= 0x1.0p+64;
}

EM_FN double em_recip(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, e, y);
  /* computed once here: without the barrier the compiler may sink the steps
   * into the branch of each quotient and repeat them */
  asm volatile("" : "+v"(y));
  return y;
#else
  return 1.0 / b;
#endif
}

/* a / b; y = em_recip(b), ok = em_recip_ok(b) */
EM_FN double em_div(double a, double b, double y, int ok) {
#if defined(__HIP_DEVICE_COMPILE__)
  double aa = em_fabs(a);
  if (ok && aa >= 0x1.0p-600 && aa <= 0x1.0p+600) {
    double q = a * y;
    double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
  }
#else
  (void)y;
  (void)ok;
#endif
  return a / b;
}

/* The same sequences for callers that have established the ranges
 * themselves: em_div_pos for 2^-64 <= b <= 2^64 (positive) and a = 0 or
 * 2^-950 <= |a| <= 2^900 (then a*y is normal and the residual exact),
 * written with the negated residual so that a zero dividend keeps its sign
 * (RN(b q - a) = -RN(a - b q), so the quotient is em_div's otherwise);
 * em_sqrt_mid for 2^-700 <= x <= 2^1000. */
EM_FN double em_div_pos(double a, double b, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
  double q = a * y;
  double r = __builtin_fma(b, q, -a);
  return __builtin_fma(-r, y, q);
#else
  (void)y;
  return a / b;
#endif
}

EM_FN double em_sqrt_mid(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = r * 0.5;
  double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, e, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
#else
  return __builtin_sqrt(x);
#endif
}

EM_FN double em_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (x >= 0x1.0p-700 && x <= 0x1.0p+1000)
    return em_sqrt_mid(x);
#endif
  return __builtin_sqrt(x);
}

#endif
