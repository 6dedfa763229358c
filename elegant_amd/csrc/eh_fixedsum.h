/*
 * eh_fixedsum.h - order-independent sum of doubles in fixed point.
 *
 * Used for the bunch-mean arrival time of ZLONGIT and WAKE with N_BINS > 0
 * (compute_average at zlongit.c:202 and wake.c:94).  The reference takes
 * compute_average from the SDDS toolkit (not vendored).  Here the mean is
 * defined as: every value v is converted to the integer trunc(|v| 2^(128-E))
 * where 2^E bounds the largest magnitude, positive and negative parts are
 * summed exactly in 192-bit integers, the difference is rounded once to
 * double (round to nearest even) and divided by n.  The integer sum does not
 * depend on the order of the values, so a GPU reduction in any order and the
 * CPU loop give the same bits.  Only bits below 2^(E-128) are dropped.
 */
#ifndef ELEGANT_AMD_EH_FIXEDSUM_H
#define ELEGANT_AMD_EH_FIXEDSUM_H

#include <stdint.h>

#ifdef __HIPCC__
#define FS_FN static inline __host__ __device__
#else
#define FS_FN static inline
#endif

#define EH_FSUM_FRAC 128

typedef struct {
  uint64_t pos[3];
  uint64_t neg[3];
  int32_t e; /* 2^e bounds every |v| */
  int32_t pad;
} eh_fsum;

FS_FN uint64_t fs_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, sizeof u);
  return u;
}

FS_FN double fs_from_bits(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, sizeof d);
  return d;
}

/* smallest e with |v| < 2^e (v finite, nonzero); 0 handled as -1074 */
FS_FN int32_t eh_fsum_exponent(double v) {
  uint64_t u = fs_bits(v) & 0x7fffffffffffffffULL;
  int32_t be = (int32_t)(u >> 52);
  if (u == 0)
    return -1074;
  if (be == 0) { /* subnormal */
    uint64_t m = u;
    int32_t k = 0;
    while (m) {
      m >>= 1;
      k++;
    }
    return -1074 + k;
  }
  return be - 1022;
}

FS_FN void eh_fsum_init(eh_fsum *a, int32_t e) {
  a->pos[0] = a->pos[1] = a->pos[2] = 0;
  a->neg[0] = a->neg[1] = a->neg[2] = 0;
  a->e = e;
  a->pad = 0;
}

FS_FN void fs_add3(uint64_t *acc, const uint64_t *v) {
  uint64_t s0 = acc[0] + v[0];
  uint64_t c0 = s0 < v[0];
  uint64_t s1 = acc[1] + v[1];
  uint64_t c1 = s1 < v[1];
  uint64_t s1c = s1 + c0;
  c1 += s1c < c0;
  acc[0] = s0;
  acc[1] = s1c;
  acc[2] = acc[2] + v[2] + c1;
}

/* the 192-bit integer trunc(|v| 2^(FRAC - e)) in limbs[3] */
FS_FN void eh_fsum_limbs(double v, int32_t e, uint64_t *limbs) {
  uint64_t u = fs_bits(v) & 0x7fffffffffffffffULL;
  int32_t be = (int32_t)(u >> 52);
  uint64_t m = u & 0x000fffffffffffffULL;
  int32_t ev, shift;
  limbs[0] = limbs[1] = limbs[2] = 0;
  if (u == 0)
    return;
  if (be == 0)
    ev = -1074;
  else {
    m |= 0x0010000000000000ULL;
    ev = be - 1075;
  }
  /* |v| = m 2^ev;  integer = m 2^(ev - e + FRAC) */
  shift = ev - e + EH_FSUM_FRAC;
  if (shift <= -64)
    return;
  if (shift < 0) {
    limbs[0] = m >> (-shift);
    return;
  }
  {
    int32_t w = shift >> 6, b = shift & 63;
    if (w < 3)
      limbs[w] = m << b;
    if (b && w + 1 < 3)
      limbs[w + 1] = m >> (64 - b);
  }
}

FS_FN void eh_fsum_add(eh_fsum *a, double v) {
  uint64_t l[3];
  eh_fsum_limbs(v, a->e, l);
  if (fs_bits(v) >> 63)
    fs_add3(a->neg, l);
  else
    fs_add3(a->pos, l);
}

FS_FN void eh_fsum_merge(eh_fsum *a, const eh_fsum *b) {
  fs_add3(a->pos, b->pos);
  fs_add3(a->neg, b->neg);
}

/* 2^k as a double for -1022 <= k <= 1023 */
FS_FN double fs_pow2(int32_t k) { return fs_from_bits((uint64_t)(k + 1023) << 52); }

FS_FN double eh_fsum_to_double(const eh_fsum *a) {
  uint64_t d[3], b0, b1;
  int neg = 0, top, i;
  /* d = |pos - neg| */
  const uint64_t *x = a->pos, *y = a->neg;
  for (i = 2; i >= 0; i--) {
    if (a->pos[i] != a->neg[i]) {
      if (a->pos[i] < a->neg[i]) {
        x = a->neg;
        y = a->pos;
        neg = 1;
      }
      break;
    }
  }
  b0 = x[0] < y[0];
  d[0] = x[0] - y[0];
  b1 = (x[1] < y[1]) || (x[1] - y[1] < b0);
  d[1] = x[1] - y[1] - b0;
  d[2] = x[2] - y[2] - b1;
  top = -1;
  for (i = 2; i >= 0 && top < 0; i--) {
    if (d[i]) {
      uint64_t t = d[i];
      int p = 0;
      while (t >>= 1)
        p++;
      top = i * 64 + p;
    }
  }
  if (top < 0)
    return 0.0;
  {
    /* take bits [top-52, top] as mantissa, round to nearest even */
    uint64_t mant = 0, rnd = 0, sticky = 0;
    int lo = top - 52, bit;
    for (bit = top; bit >= lo && bit >= 0; bit--)
      mant = (mant << 1) | ((d[bit >> 6] >> (bit & 63)) & 1);
    if (lo < 0)
      mant <<= -lo;
    if (lo - 1 >= 0) {
      rnd = (d[(lo - 1) >> 6] >> ((lo - 1) & 63)) & 1;
      for (bit = lo - 2; bit >= 0; bit--)
        if ((d[bit >> 6] >> (bit & 63)) & 1) {
          sticky = 1;
          break;
        }
    }
    if (rnd && (sticky || (mant & 1)))
      mant++;
    {
      /* value = mant 2^(lo + e - FRAC) (mant may have become 2^53: still exact) */
      int32_t k = lo + a->e - EH_FSUM_FRAC;
      double r = (double)mant;
      while (k > 1000) {
        r *= fs_pow2(1000);
        k -= 1000;
      }
      while (k < -1000) {
        r *= fs_pow2(-1000);
        k += 1000;
      }
      r *= fs_pow2(k);
      return neg ? -r : r;
    }
  }
}

#endif
