/*
 * eh_gen.c - seeded input generators: the uniform stream of the error and
 * beam generators, Gaussian deviates with a cutoff, element-error
 * perturbations and the Gaussian bunch in Twiss coordinates.
 *
 * Restated from the reference where it ships the code (DLARAN in
 * drand_oag.c:118-180 and its seeding at :94-116; perturbation() in
 * error.c:496-534; gaussian_distribution() and the normalized-coordinate
 * transform in generate_bunch.c:557-658, 1194-1212, with the plane set-up of
 * generate_bunch.c:138-160, 202-218, 315-360).  Two pieces come from the
 * SDDS toolkit, which the reference does not ship: the bit permutation of
 * the seed (permuteSeedBitOrder) and gauss_rn().  The seed is used as given
 * (elegant itself skips the permutation for the seed 987654321), and the
 * Gaussian deviate is the polar method of Marsaglia & Bray (1964) with the
 * second deviate of each pair saved for the next call, the published
 * algorithm gauss_rn implements.  Streams are therefore reproducible here
 * but are not bit-identical to elegant's except where noted.
 */
#include <math.h>
#include <string.h>

#include "eh_host.h"

/* random_1_elegant(-iseed) seeding (drand_oag.c:94-116), no permutation:
 * make the seed odd and split it into four 12-bit limbs. */
int eh_rng_seed(eh_rng *r, int64_t iseed) {
  if (!r)
    return EH_ERR_INVALID;
  memset(r, 0, sizeof(*r));
  if (iseed < 0)
    iseed = -iseed;
  iseed = (iseed / 2) * 2 + 1;
  r->seed[3] = iseed & 4095;
  r->seed[2] = (iseed >>= 12) & 4095;
  r->seed[1] = (iseed >>= 12) & 4095;
  r->seed[0] = (iseed >>= 12) & 4095;
  return EH_OK;
}

/* dlaran_OAG (drand_oag.c:118-180): seed *= 33952834046453 mod 2^48 in
 * 12-bit limbs (multiplier limbs 494, 322, 2508, 2549), value = seed/2^48 */
double eh_rng_uniform(eh_rng *r) {
  int64_t it1, it2, it3, it4;
  int64_t *s = r->seed;
  it4 = s[3] * 2549;
  it3 = it4 / 4096;
  it4 -= it3 << 12;
  it3 = it3 + s[2] * 2549 + s[3] * 2508;
  it2 = it3 / 4096;
  it3 -= it2 << 12;
  it2 = it2 + s[1] * 2549 + s[2] * 2508 + s[3] * 322;
  it1 = it2 / 4096;
  it2 -= it1 << 12;
  it1 = it1 + s[0] * 2549 + s[1] * 2508 + s[2] * 322 + s[3] * 494;
  it1 %= 4096;
  s[0] = it1;
  s[1] = it2;
  s[2] = it3;
  s[3] = it4;
  return ((double)it1 + ((double)it2 + ((double)it3 + (double)it4 * 2.44140625e-4) * 2.44140625e-4) * 2.44140625e-4) *
         2.44140625e-4;
}

/* unit Gaussian deviate: polar method, pairs with the second one saved */
double eh_gauss_rn(eh_rng *r) {
  double u1, u2, rsq, f;
  if (r->have_saved) {
    r->have_saved = 0;
    return r->saved;
  }
  do {
    u1 = 2 * eh_rng_uniform(r) - 1;
    u2 = 2 * eh_rng_uniform(r) - 1;
    rsq = u1 * u1 + u2 * u2;
  } while (rsq >= 1 || rsq == 0);
  f = sqrt(-2 * log(rsq) / rsq);
  r->saved = u2 * f;
  r->have_saved = 1;
  return u1 * f;
}

/* gauss_rn_lim(mean, sigma, limit): redraw while |z| > limit (limit <= 0: no cut) */
double eh_gauss_rn_lim(eh_rng *r, double mean, double sigma, double limit) {
  double z;
  do {
    z = eh_gauss_rn(r);
  } while (limit > 0 && fabs(z) > limit);
  return z * sigma + mean;
}

/* perturbation() (error.c:496-534) for the uniform, Gaussian and
 * plus-or-minus error types (0, 1, 2; correct.c:107-112) */
double eh_perturbation(eh_rng *r, double amplitude, double cutoff, int32_t error_type) {
  switch (error_type) {
  case 0:
    return 2 * amplitude * (eh_rng_uniform(r) - 0.5);
  case 1:
    return eh_gauss_rn_lim(r, 0.0, amplitude, cutoff);
  case 2:
    return amplitude * (eh_rng_uniform(r) > 0.5 ? 1.0 : -1.0);
  default:
    return NAN;
  }
}

int eh_perturbations(eh_rng *r, double amplitude, double cutoff, int32_t error_type, double *out, int64_t n) {
  if (!r || (!out && n > 0) || n < 0 || error_type < 0 || error_type > 2)
    return EH_ERR_INVALID;
  for (int64_t i = 0; i < n; i++)
    out[i] = eh_perturbation(r, amplitude, cutoff, error_type);
  return EH_OK;
}

/* gaussian_distribution() without Halton sequences (generate_bunch.c:629-657):
 * pairs (u, u') with sigmas s1, s2, kept when inside the cut (|u| <= s1 limit
 * and |u'| <= s2 limit, or u^2 + (beta u')^2 <= limit_invar when that is
 * nonzero), or outside it for a halo; symmetrize adds the three mirror
 * images of each pair. */
int eh_gaussian_distribution(eh_rng *r, double **particle, int64_t n, int32_t offset, double s1, double s2,
                             int32_t symmetrize, double limit, double limit_invar, double beta, int32_t halo) {
  double x1, x2, limit1 = s1 * limit, limit2 = s2 * limit;
  int flag;
  if (!r || !particle || n < 0 || offset < 0 || (!limit_invar && !(limit > 0)))
    return EH_ERR_INVALID; /* a zero cut would reject every pair */
  for (int64_t i = 0; i < n; i++) {
    do {
      particle[i][0 + offset] = x1 = eh_gauss_rn_lim(r, 0.0, s1, 0.0);
      particle[i][1 + offset] = x2 = eh_gauss_rn_lim(r, 0.0, s2, 0.0);
      if (limit_invar)
        flag = (x1 * x1 + (beta * x2) * (beta * x2)) <= limit_invar;
      else
        flag = fabs(x1) <= limit1 && fabs(x2) <= limit2;
      if ((!halo && flag) || (halo && !flag))
        break;
    } while (1);
    if (symmetrize) {
      if (++i >= n)
        break;
      particle[i][0 + offset] = x1;
      particle[i][1 + offset] = -x2;
      if (++i >= n)
        break;
      particle[i][0 + offset] = -x1;
      particle[i][1 + offset] = x2;
      if (++i >= n)
        break;
      particle[i][0 + offset] = -x1;
      particle[i][1 + offset] = -x2;
    }
  }
  return EH_OK;
}

/* transform_from_normalized_coordinates (generate_bunch.c:1194-1212) */
void eh_transform_from_normalized(double **particle, int64_t n, int32_t offset, double beta, double alpha) {
  double c1 = sqrt(beta), c2 = alpha / c1;
  for (int64_t i = 0; i < n; i++) {
    double u = particle[i][0 + offset], up = particle[i][1 + offset];
    particle[i][0 + offset] = c1 * u;
    particle[i][1 + offset] = c1 * up - c2 * u;
  }
}

/* A Gaussian bunch: transverse planes from (emit, beta, alpha, cutoff), the
 * longitudinal plane from (sigma_s, sigma_dp, alpha_z, cutoff) as in
 * generate_bunch.c:138-160, 202-218 and 315-360 (no symmetrization, no
 * Halton sequences, no rms enforcement, no dispersion or centroid offsets);
 * column 6 = particle ID from first_id, column 7 = -1 when nprops > 7. */
int eh_gaussian_bunch(eh_rng *r, double **particle, int64_t n, int32_t nprops, const double *x_plane,
                      const double *y_plane, const double *longit, int64_t first_id) {
  int rc;
  double s1, s2, beta, alpha, emit, s56;
  if (!r || !particle || n < 0 || nprops < 7 || !x_plane || !y_plane || !longit)
    return EH_ERR_INVALID;
  /* x_plane, y_plane: {emit, beta, alpha, cutoff} */
  for (int plane = 0; plane < 2; plane++) {
    const double *p = plane ? y_plane : x_plane;
    s1 = sqrt(p[0]);
    s2 = s1 / p[1];
    if ((rc = eh_gaussian_distribution(r, particle, n, 2 * plane, s1, s2, 0, p[3], 0.0, p[1], 0)))
      return rc;
    eh_transform_from_normalized(particle, n, 2 * plane, p[1], p[2]);
  }
  /* longit: {sigma_s, sigma_dp, alpha_z, cutoff} */
  if (!longit[2])
    s56 = 0;
  else
    s56 = -longit[1] * longit[0] * longit[2] / sqrt(1 + longit[2] * longit[2]);
  emit = sqrt((longit[1] * longit[0]) * (longit[1] * longit[0]) - s56 * s56);
  if (isnan(emit) || isinf(emit))
    emit = 0;
  if (emit) {
    beta = longit[0] * longit[0] / emit;
    alpha = longit[2] ? longit[2] : -s56 / emit;
    s1 = sqrt(emit);
    s2 = s1 / beta;
  } else {
    s1 = longit[0];
    s2 = longit[1];
    beta = 0;
    alpha = 0;
  }
  if ((rc = eh_gaussian_distribution(r, particle, n, 4, s1, s2, 0, longit[3], 0.0, beta, 0)))
    return rc;
  if (emit)
    eh_transform_from_normalized(particle, n, 4, beta, alpha);
  for (int64_t i = 0; i < n; i++) {
    particle[i][6] = (double)(first_id + i);
    if (nprops > 7)
      particle[i][7] = -1;
  }
  return EH_OK;
}
