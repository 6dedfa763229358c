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
 * Correct rounding.  sin and cos for |x| >= 2^-5, tan for |x| >= 2^-27 and
 * log everywhere return the double nearest to the exact value, for every
 * argument, except where the exact value lies within 2^-100 (relative) of
 * the midpoint between two doubles.  Method: Ziv's strategy (A. Ziv, "Fast
 * evaluation of elementary mathematical functions with correctly rounded
 * last bit", ACM TOMS 17, 1991): a fast evaluation into an unevaluated pair
 * h + l with a stated error bound u; when h + l - u and h + l + u round to
 * the same double that double is the result, else an accurate evaluation in
 * double-double arithmetic (relative error below 2^-100) decides.
 *  - argument reduction modulo pi/2: Cody & Waite (1980) with pi/2 split
 *    into seven 24-bit pieces, accumulated in double-double (Knuth TwoSum),
 *    for |x| < 2^28; beyond that, and in every accurate evaluation,
 *    Payne & Hanek (1983) with the bits of 2/pi in a table: the product of
 *    the 53-bit significand with the 256 bits of 2/pi that matter modulo 4,
 *    in exact integer arithmetic (K. C. Ng, "Argument reduction for huge
 *    arguments: good to the last bit", 1992);
 *  - sin/cos of the reduced argument r: r = j pi/64 + d, |d| <= pi/128,
 *    sin(r) = S_j cos d + C_j sin d and cos(r) = C_j cos d - S_j sin d with
 *    S_j, C_j tabulated as double-double pairs and Taylor polynomials in d
 *    (double precision in the fast path, double-double in the accurate one);
 *    tan = sin/cos by a double-double quotient;
 *  - log: x = 2^e m, m in [sqrt(2)/2, sqrt(2)], m r_i = 1 + u with
 *    r_i = RN(128/i) for the i nearest 128 m (the product m r_i split exactly
 *    by an FMA), log x = e ln2 - log r_i + log1p(u), -log r_i tabulated as a
 *    pair and log1p(u) by its series (|u| < 2^-7.4);
 *  - the small-argument bands, which the bend integrators use branch-free:
 *    sin/cos for |x| < 2^-5 by Taylor series to degree 9/8 (coefficients
 *    1/n! rounded to double) with the compensated 1 - z/2 step for cos, and
 *    asin everywhere (Taylor series in x^2 on |x| <= 0.5 by band, and the
 *    reflection asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above, the reduction
 *    fdlibm's e_asin.c also uses) are faithful
 *    evaluations that round correctly except on a small fraction of
 *    arguments (measured against 200-bit evaluations in tests/test_emath.py);
 *  - every polynomial by Horner's rule in fused multiply-adds (correctly
 *    rounded on both sides: the C library's fma on the host, v_fma_f64 on
 *    the device).
 * Constants and tables were derived in exact integer arithmetic (pi from
 * Machin's formula, logarithms from atanh series; gen_emath_tables.py writes
 * emath_tables.h) and are written as hex literals.
 */
#ifndef ELEGANT_AMD_EMATH_H
#define ELEGANT_AMD_EMATH_H

#include <stdint.h>

#ifdef __HIPCC__
#define EM_FN static inline __host__ __device__
#define EM_TABLE static const __device__
#else
#define EM_FN static inline
#define EM_TABLE static const
#endif

#include "emath_tables.h"

#define EM_PIO4 0x1.921fb54442d18p-1
#define EM_PIO2_HI 0x1.921fb54442d18p+0
#define EM_PIO2_LO 0x1.1a62633145c07p-54
#define EM_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define EM_SHIFT 0x1.8p+52

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

/* The small-argument band |x| < 2^-5 (z = x*x < 2^-10): sin through x^9
 * and cos through x^8; the first dropped term is below 2^-64 relative. */
#define EM_S1 (-0x1.5555555555555p-3)
#define EM_SMALL_Z 0x1.0p-10

/* (sin(x) - x - S1 x^3) / x^5 as a polynomial in z = x*x */
EM_FN double em_sin_r_small(double z) {
  const double s2 = 0x1.1111111111111p-7, s3 = -0x1.a01a01a01a01ap-13, s4 = 0x1.71de3a556c734p-19;
  return __builtin_fma(z, __builtin_fma(z, em_k(s4), em_k(s3)), em_k(s2));
}

/* (cos(x) - 1 + z/2) / z^2 as a polynomial in z = x*x */
EM_FN double em_cos_r_small(double z) {
  const double c2 = 0x1.5555555555555p-5, c3 = -0x1.6c16c16c16c17p-10, c4 = 0x1.a01a01a01a01ap-16;
  return __builtin_fma(z, __builtin_fma(z, em_k(c4), em_k(c3)), em_k(c2));
}

/* em_sin(x) and em_cos(x) on the band, z = x*x < 2^-10 (x != 0 for sin) */
EM_FN double em_sin_small(double x, double z) {
  return x + (z * x) * __builtin_fma(z, em_sin_r_small(z), em_k(EM_S1));
}

/* The compensated form w + ((1 - w) - z/2 + r), w = 1 - z/2, recovers the
 * rounding error of 1 - z/2; it is the one the cosine kernel of Sun's fdlibm
 * (k_cos.c) uses. */
EM_FN double em_cos_small(double x, double z) {
  double r = z * z * em_cos_r_small(z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + r);
}

/* ---- double-double arithmetic ---- */
/* a + b = s + e exactly, |a| >= |b| or a = 0 (Dekker's Fast2Sum) */
EM_FN void em_fast2sum(double a, double b, double *s, double *e) {
  double t = a + b;
  *e = b - (t - a);
  *s = t;
}

/* a + b = s + e exactly (Knuth's TwoSum) */
EM_FN void em_two_sum(double a, double b, double *s, double *e) {
  double t = a + b;
  double bb = t - a;
  *e = (a - (t - bb)) + (b - bb);
  *s = t;
}

/* a b = p + e exactly */
EM_FN void em_two_prod(double a, double b, double *p, double *e) {
  double t = a * b;
  *e = __builtin_fma(a, b, -t);
  *p = t;
}

EM_FN void em_dd_mul(double ah, double al, double bh, double bl, double *rh, double *rl) {
  double p, e;
  em_two_prod(ah, bh, &p, &e);
  e += ah * bl + al * bh;
  em_fast2sum(p, e, rh, rl);
}

EM_FN void em_dd_add(double ah, double al, double bh, double bl, double *rh, double *rl) {
  double s, e;
  em_two_sum(ah, bh, &s, &e);
  e += al + bl;
  em_fast2sum(s, e, rh, rl);
}

/* Ziv's rounding test: 1 when every value within u of h + l (h = RN(h + l))
 * rounds to h.  u is enlarged by 2^-10 of itself, which covers the rounding
 * of l +- u, so the two floating-point sums decide for the exact interval. */
EM_FN int em_rounds_to(double h, double l, double u) {
  u += u * 0x1.0p-10;
  return h == h + (l + u) && h == h + (l - u);
}

/* ---- Payne-Hanek reduction ---- */
/* word wi of the 320-bit little-endian integer p, zero outside [0, 10); a
 * select chain rather than an indexed load, so that p stays in registers */
EM_FN uint32_t em_word(const uint32_t *p, int wi) {
  uint32_t v = 0;
  for (int k = 0; k < 10; k++)
    v = k == wi ? p[k] : v;
  return v;
}

/* bits [pos, pos + 32) of p, pos >= 0 */
EM_FN uint32_t em_bits32(const uint32_t *p, int pos) {
  int wi = pos >> 5, sh = pos & 31;
  uint32_t v = em_word(p, wi) >> sh;
  if (sh)
    v |= em_word(p, wi + 1) << (32 - sh);
  return v;
}

/* bits [pos, pos + 48) of p */
EM_FN uint64_t em_bits48(const uint32_t *p, int pos) {
  return (((uint64_t)em_bits32(p, pos + 32) & 0xffffULL) << 32) | em_bits32(p, pos);
}

/* x = k pi/2 + (rh + rl) for finite |x| >= 2^-5: k mod 4 is returned, the
 * pair has a relative error below 2^-103 and |rh| <= pi/4 (1 + 2^-50).
 * |x| = M 2^E with M a 53-bit integer; bits b_i (weight 2^-i) of 2/pi with
 * i <= E - 2 contribute multiples of 4 to x 2/pi and are skipped, the next
 * 256 bits are multiplied by M exactly (32-bit limbs), the truncation error
 * is below 2^-200, and the integer part of the product is k, its fraction
 * (rounded to nearest k) times pi/2 is r. */
EM_FN int em_reduce_ph(double x, double *rh, double *rl) {
  uint64_t u = em_bits(x);
  int neg = (int)(u >> 63);
  int E = (int)((u >> 52) & 0x7ff) - 1075;
  uint64_t M = (u & 0x000fffffffffffffULL) | 0x0010000000000000ULL;
  int i0 = E - 1 > 1 ? E - 1 : 1;
  int q = (i0 - 1) >> 5, sh = (i0 - 1) & 31;
  uint32_t w[8], p[10], m[2];
  for (int k = 0; k < 8; k++) {
    uint32_t a = em_two_over_pi_bits[q + 7 - k], b = em_two_over_pi_bits[q + 8 - k];
    w[k] = sh ? (a << sh) | (b >> (32 - sh)) : a;
  }
  m[0] = (uint32_t)M;
  m[1] = (uint32_t)(M >> 32);
  for (int k = 0; k < 10; k++)
    p[k] = 0;
  for (int a = 0; a < 2; a++) {
    uint64_t carry = 0;
    for (int b = 0; b < 8; b++) {
      uint64_t t = (uint64_t)m[a] * w[b] + p[a + b] + carry;
      p[a + b] = (uint32_t)t;
      carry = t >> 32;
    }
    p[a + 8] = (uint32_t)carry;
  }
  int F = i0 + 255 - E; /* binary point of x 2/pi in p */
  int k = (int)(em_bits32(p, F) & 3);
  uint64_t g1 = em_bits48(p, F - 48);
  double d1 = (double)g1 * 0x1.0p-48, d2 = (double)em_bits48(p, F - 96) * 0x1.0p-96;
  double d3 = (double)em_bits48(p, F - 144) * 0x1.0p-144, d4 = (double)em_bits48(p, F - 192) * 0x1.0p-192;
  if (g1 >> 47) { /* fraction >= 1/2: round k up */
    d1 -= 1.0;
    k += 1;
  }
  double h, l, l2;
  em_two_sum(d1, d2, &h, &l);
  em_two_sum(h, d3 + d4, &h, &l2);
  l += l2;
  em_fast2sum(h, l, &h, &l);
  em_dd_mul(h, l, EM_PIO2_1, EM_PIO2_2, rh, rl);
  if (neg) {
    *rh = -*rh;
    *rl = -*rl;
    k = -k;
  }
  return k & 3;
}

/* reduced argument of a finite |x| >= 2^-5 for the fast evaluations */
EM_FN int em_reduce_fast(double x, double *rh, double *rl) {
  double ax = em_fabs(x);
  if (ax <= EM_PIO4) {
    *rh = x;
    *rl = 0.0;
    return 0;
  }
  if (ax < 0x1.0p+28)
    return em_reduce_pio2(x, rh, rl) & 3;
  return em_reduce_ph(x, rh, rl);
}

/* ---- sin and cos of a reduced argument ---- */
/* r = rh + rl = j pi/64 + d: j, and d as a pair (dh, dl) */
EM_FN int em_split_pio64(double rh, double rl, double *dh, double *dl) {
  double jd = (rh * EM_64_OVER_PI + EM_SHIFT) - EM_SHIFT;
  double d = rh - jd * EM_PIO64_A; /* exact: j A has 53 bits and lies within a factor 2 of rh */
  double t, te;
  em_two_prod(jd, EM_PIO64_B, &t, &te);
  em_two_sum(d, -t, dh, dl);
  *dl += (rl - te) - jd * EM_PIO64_C;
  return (int)jd;
}

EM_FN void em_table_sc(int j, double *Sh, double *Sl, double *Ch, double *Cl) {
  int aj = j < 0 ? -j : j;
  *Sh = em_sin_tab_hi[aj];
  *Sl = em_sin_tab_lo[aj];
  *Ch = em_cos_tab_hi[aj];
  *Cl = em_cos_tab_lo[aj];
  if (j < 0) {
    *Sh = -*Sh;
    *Sl = -*Sl;
  }
}

/* fast: sin r and cos r as pairs, |r| <= pi/4 (1 + 2^-50); relative error
 * below 2^-60 plus an absolute 2^-101 from the reduced argument.  The
 * dominant error is the rounding of S_j (cos d - 1), below 2^-62 of the
 * result; the table and the polynomials (through d^9 and d^8, |d| <= pi/128)
 * contribute below 2^-75. */
EM_FN void em_sincos_fast(double rh, double rl, double *sh, double *sl, double *ch, double *cl) {
  double dh, dl, Sh, Sl, Ch, Cl, p, pe, s, se, tail;
  int j = em_split_pio64(rh, rl, &dh, &dl);
  em_two_sum(dh, dl, &dh, &dl);
  double z = dh * dh;
  double ps = (dh * z) * __builtin_fma(z, __builtin_fma(z, __builtin_fma(z, em_k(0x1.71de3a556c734p-19),
                                                                            em_k(-0x1.a01a01a01a01ap-13)),
                                                        em_k(0x1.1111111111111p-7)),
                                       em_k(-0x1.5555555555555p-3));
  double pc = z * __builtin_fma(z, __builtin_fma(z, __builtin_fma(z, em_k(0x1.a01a01a01a01ap-16),
                                                                 em_k(-0x1.6c16c16c16c17p-10)),
                                                 em_k(0x1.5555555555555p-5)),
                                -0.5) -
              dh * dl;
  em_table_sc(j, &Sh, &Sl, &Ch, &Cl);
  /* sin(a + d) = S + C d + (S (cos d - 1) + C (sin d - d)) */
  em_two_prod(Ch, dh, &p, &pe);
  em_two_sum(Sh, p, &s, &se);
  tail = se + pe + Sl + Cl * dh + Sh * pc + Ch * (dl + ps);
  em_fast2sum(s, tail, sh, sl);
  /* cos(a + d) = C - S d + (C (cos d - 1) - S (sin d - d)) */
  em_two_prod(Sh, dh, &p, &pe);
  em_two_sum(Ch, -p, &s, &se);
  tail = se - pe + Cl - Sl * dh + Ch * pc - Sh * (dl + ps);
  em_fast2sum(s, tail, ch, cl);
}

/* accurate: the same in double-double, relative error below 2^-100
 * (series of sin d - d and cos d - 1 through d^13 and d^12 with 1/n! as
 * pairs; the first dropped terms are below 2^-112 relative) */
EM_FN void em_sincos_acc(double rh, double rl, double *sh, double *sl, double *ch, double *cl) {
  double dh, dl, zh, zl, qh, ql, ph, pl, th, tl, Sh, Sl, Ch, Cl, ah, al, bh, bl;
  int j = em_split_pio64(rh, rl, &dh, &dl);
  em_two_sum(dh, dl, &dh, &dl);
  em_dd_mul(dh, dl, dh, dl, &zh, &zl);
  /* Q(z) = sum_{k=1..6} (-1)^k z^(k-1) / (2k+1)!, P(z) = sum_{k=1..6} (-1)^k z^(k-1) / (2k)! */
  qh = em_inv_fact[13][0];
  ql = em_inv_fact[13][1];
  ph = em_inv_fact[12][0];
  pl = em_inv_fact[12][1];
  for (int k = 5; k >= 1; k--) {
    double sg = (k & 1) ? -1.0 : 1.0;
    em_dd_mul(qh, ql, zh, zl, &qh, &ql);
    em_dd_add(qh, ql, sg * em_inv_fact[2 * k + 1][0], sg * em_inv_fact[2 * k + 1][1], &qh, &ql);
    em_dd_mul(ph, pl, zh, zl, &ph, &pl);
    em_dd_add(ph, pl, sg * em_inv_fact[2 * k][0], sg * em_inv_fact[2 * k][1], &ph, &pl);
  }
  /* sin d = d + d z Q, cos d - 1 = z P */
  em_dd_mul(qh, ql, zh, zl, &th, &tl);
  em_dd_mul(th, tl, dh, dl, &th, &tl);
  em_dd_add(dh, dl, th, tl, &qh, &ql); /* q = sin d */
  em_dd_mul(ph, pl, zh, zl, &ph, &pl); /* p = cos d - 1 */
  em_table_sc(j, &Sh, &Sl, &Ch, &Cl);
  em_dd_mul(Sh, Sl, ph, pl, &ah, &al);
  em_dd_mul(Ch, Cl, qh, ql, &bh, &bl);
  em_dd_add(Sh, Sl, bh, bl, &th, &tl);
  em_dd_add(th, tl, ah, al, sh, sl);
  em_dd_mul(Ch, Cl, ph, pl, &ah, &al);
  em_dd_mul(Sh, Sl, qh, ql, &bh, &bl);
  em_dd_add(Ch, Cl, -bh, -bl, &th, &tl);
  em_dd_add(th, tl, ah, al, ch, cl);
}

/* the accurate evaluation of a finite |x| >= 2^-5: sin (fn 0), cos (fn 1)
 * or tan (fn 2), correctly rounded unless within 2^-100 of a midpoint */
EM_FN double em_trig_acc(double x, int fn) {
  double rh, rl, sh, sl, ch, cl;
  int k;
  if (em_fabs(x) <= EM_PIO4) {
    rh = x;
    rl = 0.0;
    k = 0;
  } else
    k = em_reduce_ph(x, &rh, &rl);
  em_sincos_acc(rh, rl, &sh, &sl, &ch, &cl);
  if (fn == 2) {
    /* tan = sin r / cos r (k even) or -cos r / sin r (k odd), pair quotient */
    double nh = (k & 1) ? -ch : sh, nl = (k & 1) ? -cl : sl;
    double dh = (k & 1) ? sh : ch, dl = (k & 1) ? sl : cl;
    double q = nh / dh;
    double r = __builtin_fma(-q, dh, nh); /* exact remainder */
    r = (r + nl) - q * dl;
    double h, l;
    em_fast2sum(q, r / dh, &h, &l);
    return h;
  }
  int qd = (k + fn) & 3; /* cos x = sin(x + pi/2) */
  double h = (qd & 1) ? ch : sh;
  return (qd & 2) ? -h : h;
}

/* sin (fn 0) or cos (fn 1) of a finite |x| >= 2^-5, correctly rounded */
EM_FN double em_sincos_cr1(double x, int fn) {
  double rh, rl, sh, sl, ch, cl;
  int k = em_reduce_fast(x, &rh, &rl);
  em_sincos_fast(rh, rl, &sh, &sl, &ch, &cl);
  int qd = (k + fn) & 3;
  double h = (qd & 1) ? ch : sh, l = (qd & 1) ? cl : sl;
  if (qd & 2) {
    h = -h;
    l = -l;
  }
  if (em_rounds_to(h, l, 0x1.0p-60 * em_fabs(h) + 0x1.0p-101))
    return h;
  return em_trig_acc(x, fn);
}

EM_FN double em_sin(double x) {
  double ax = em_fabs(x);
  if (ax < 0x1.0p-5)
    return x == 0.0 ? x : em_sin_small(x, x * x);
  if (!(ax <= 0x1.fffffffffffffp+1023))
    return em_nan();
  return em_sincos_cr1(x, 0);
}

EM_FN double em_cos(double x) {
  double ax = em_fabs(x);
  if (ax < 0x1.0p-5)
    return x == 0.0 ? 1.0 : em_cos_small(x, x * x);
  if (!(ax <= 0x1.fffffffffffffp+1023))
    return em_nan();
  return em_sincos_cr1(x, 1);
}

EM_FN void em_sincos(double x, double *sp, double *cp) {
  double ax = em_fabs(x);
  if (ax < 0x1.0p-5) {
    double z = x * x;
    *sp = x == 0.0 ? x : em_sin_small(x, z);
    *cp = x == 0.0 ? 1.0 : em_cos_small(x, z);
    return;
  }
  if (!(ax <= 0x1.fffffffffffffp+1023)) {
    *sp = *cp = em_nan();
    return;
  }
  double rh, rl, sh, sl, ch, cl;
  int k = em_reduce_fast(x, &rh, &rl);
  em_sincos_fast(rh, rl, &sh, &sl, &ch, &cl);
  const double u0 = 0x1.0p-101;
  /* sin x: quadrant k; cos x: quadrant k + 1 */
  for (int fn = 0; fn < 2; fn++) {
    int qd = (k + fn) & 3;
    double h = (qd & 1) ? ch : sh, l = (qd & 1) ? cl : sl;
    if (qd & 2) {
      h = -h;
      l = -l;
    }
    double v = em_rounds_to(h, l, 0x1.0p-60 * em_fabs(h) + u0) ? h : em_trig_acc(x, fn);
    if (fn == 0)
      *sp = v;
    else
      *cp = v;
  }
}

/* tan x: x itself below 2^-27 (x^3/3 is below half an ulp of x there);
 * else sin r / cos r or -cos r / sin r from the fast pairs by a pair
 * quotient, whose relative error is below 2^-59 plus 2^-101 / |sin r|. */
EM_FN double em_tan(double x) {
  double ax = em_fabs(x);
  if (ax < 0x1.0p-27)
    return x;
  if (!(ax <= 0x1.fffffffffffffp+1023))
    return em_nan();
  double rh, rl, sh, sl, ch, cl;
  int k = em_reduce_fast(x, &rh, &rl);
  em_sincos_fast(rh, rl, &sh, &sl, &ch, &cl);
  double nh = (k & 1) ? -ch : sh, nl = (k & 1) ? -cl : sl;
  double dh = (k & 1) ? sh : ch, dl = (k & 1) ? sl : cl;
  double q = nh / dh;
  double r = __builtin_fma(-q, dh, nh);
  r = (r + nl) - q * dl;
  double h, l;
  em_fast2sum(q, r / dh, &h, &l);
  if (em_rounds_to(h, l, em_fabs(h) * (0x1.0p-59 + 0x1.0p-101 / em_fabs(sh))))
    return h;
  return em_trig_acc(x, 2);
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

/* log x = e ln2 - log r_i + log1p(u) with u = m r_i - 1 as an exact pair.
 * Fast: the large terms summed exactly (TwoSum), log1p's u^2/2 as an exact
 * product, u^3 (1/3 - u/4 + ... + u^8/11) in double (next term below 2^-80
 * of u); relative error below 2^-66.  Accurate: log1p(u) by its series
 * through u^15 in double-double and ln 2 to 148 bits. */
EM_FN double em_log_acc(int e, double uh, double ul, int i) {
  double ph = em_inv_int[15][0], pl = em_inv_int[15][1], th, tl, ah, al;
  for (int k = 14; k >= 1; k--) {
    double sg = (k & 1) ? 1.0 : -1.0;
    em_dd_mul(ph, pl, uh, ul, &ph, &pl);
    em_dd_add(ph, pl, sg * em_inv_int[k][0], sg * em_inv_int[k][1], &ph, &pl);
  }
  em_dd_mul(ph, pl, uh, ul, &th, &tl); /* log1p(u) */
  double ed = (double)e, bh, bl;
  em_two_prod(ed, EM_LN2_2, &bh, &bl);
  bl += ed * EM_LN2_3;
  em_dd_add(ed * EM_LN2_1, 0.0, bh, bl, &ah, &al); /* e ln 2 */
  em_dd_add(ah, al, em_log_hi[i], em_log_lo[i], &ah, &al);
  em_dd_add(ah, al, th, tl, &ah, &al);
  return ah;
}

EM_FN double em_log(double x) {
  uint64_t u;
  int e, i;
  double m, r, p, pe, uh, ul, q, qe, ed, s, t1, t2, t3, poly, small, h, l;
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
  i = (int)(m * 128.0 + 0.5) - 91; /* 0..90 */
  r = em_log_r[i];
  em_two_prod(m, r, &p, &pe);
  em_two_sum(p - 1.0, pe, &uh, &ul); /* p - 1 is exact (p in [0.99, 1.01]) */
  em_two_prod(uh, uh, &q, &qe);
  ed = (double)e;
  em_two_sum(ed * EM_LN2_1, em_log_hi[i], &s, &t1); /* e EM_LN2_1 is exact */
  em_two_sum(s, uh, &s, &t2);
  em_two_sum(s, -0.5 * q, &s, &t3);
  /* u^3 (1/3 - u/4 + u^2/5 - ... + u^8/11) */
  poly = __builtin_fma(uh, em_k(0x1.745d1745d1746p-4), em_k(-0x1.999999999999ap-4));
  poly = __builtin_fma(uh, poly, em_k(0x1.c71c71c71c71cp-4));
  poly = __builtin_fma(uh, poly, em_k(-0x1.0p-3));
  poly = __builtin_fma(uh, poly, em_k(0x1.2492492492492p-3));
  poly = __builtin_fma(uh, poly, em_k(-0x1.5555555555555p-3));
  poly = __builtin_fma(uh, poly, em_k(0x1.999999999999ap-3));
  poly = __builtin_fma(uh, poly, em_k(-0x1.0p-2));
  poly = __builtin_fma(uh, poly, em_k(0x1.5555555555555p-2));
  poly *= uh * q;
  /* u^2/2 = q/2 + qe/2 + uh ul (+ ul^2/2, below 2^-120 of u) */
  small = t1 + t2 + t3 + (ed * EM_LN2_2 + em_log_lo[i]) + ul - 0.5 * qe - uh * ul + poly;
  em_fast2sum(s, small, &h, &l);
  if (em_rounds_to(h, l, 0x1.0p-63 * em_fabs(h)))
    return h;
  return em_log_acc(e, uh, ul, i);
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
