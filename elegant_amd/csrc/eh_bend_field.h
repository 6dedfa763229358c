/*
 * eh_bend_field.h - the field of a CSBEND with normal multipoles (host side,
 * shared by the library's element compiler and the CPU checker).
 *
 * computeCSBENDFieldCoefficients (csbend.c:186-660) for b = (xKick/angle,
 * K1 rho0, ..., K8 rho0) and no skew terms: the coefficients Fx[i][j],
 * Fy[i][j] of x^i y^j in the normalised field, from the exact rational
 * expansion that eh_csbend_fields.h tabulates (Maxwell's equations in the
 * bend's curvilinear frame, derived in gen_csbend_fields.py), each evaluated
 * term by term, (n h^p b_m)/d, left to right in increasing m.  The kick
 * evaluates them as computeCSBENDFields (csbend.c:100-141) does:
 *   sum over i of (sum over j of (F[i][j] x^i) y^j), j odd for Fx and even
 *   for Fy, i + j <= expansion order.
 */
#ifndef ELEGANT_AMD_EH_BEND_FIELD_H
#define ELEGANT_AMD_EH_BEND_FIELD_H

#include <string.h>

#include "eh_csbend_fields.h"

#define EH_FIELD_MAX1 11 /* expansion order + 1, at most (csbend.c:201-203) */

/* b[0..8] as described above; expansion_order 0 picks elegant's default (the
 * highest nonzero b index + 2, at least 4; csbend.c:191-199).  Returns the
 * expansion order + 1 (the loop bound of the field sums), 0 when every b is
 * zero (the pure dipole: Fx = 0, Fy = 1 without a sum, csbend.c:105-109), or
 * -1 for an order above 10. */
static int eh_bend_field_setup(const double *b, double h1, int nonlinear, int expansion_order,
                               double Fx[EH_FIELD_MAX1][EH_FIELD_MAX1], double Fy[EH_FIELD_MAX1][EH_FIELD_MAX1]) {
  double h[20];
  int i, has_normal = 0;
  if (expansion_order == 0) {
    for (i = 8; i >= 0; i--)
      if (b[i])
        break;
    expansion_order = i + 2 < 4 ? 4 : i + 2;
  }
  if (expansion_order + 1 > EH_FIELD_MAX1 || expansion_order < 0)
    return -1;
  for (i = 0; i < 9; i++)
    if (b[i])
      has_normal = 1;
  memset(Fx, 0, sizeof(double) * EH_FIELD_MAX1 * EH_FIELD_MAX1);
  memset(Fy, 0, sizeof(double) * EH_FIELD_MAX1 * EH_FIELD_MAX1);
  h[0] = 1;
  for (i = 1; i < 20; i++)
    h[i] = h[i - 1] * h1;
  Fy[0][0] = 1 - b[0];
  Fx[0][1] = b[1];
  Fy[1][0] = b[1];
  if (nonlinear) {
    int prev = -1;
    for (int k = 0; k < EH_FIELD_TERMS; k++) {
      const eh_field_term *t = &eh_field_terms[k];
      double v = ((t->n * h[t->p]) * b[t->m]) / t->d;
      int key = (t->field * 16 + t->i) * 16 + t->j;
      double *F = t->field ? &Fx[(int)t->i][(int)t->j] : &Fy[(int)t->i][(int)t->j];
      if (key != prev)
        *F = v;
      else
        *F += v;
      prev = key;
    }
  }
  return has_normal ? expansion_order + 1 : 0;
}

/* the coefficients in the order the field sums read them: for each i, Fx
 * for odd j, then Fy for even j (j < order1 - i); returns how many */
static int eh_bend_field_flatten(int order1, double Fx[EH_FIELD_MAX1][EH_FIELD_MAX1],
                                 double Fy[EH_FIELD_MAX1][EH_FIELD_MAX1], double *out) {
  int k = 0;
  for (int i = 0; i < order1; i++) {
    for (int j = 1; j < order1 - i; j += 2)
      out[k++] = Fx[i][j];
    for (int j = 0; j < order1 - i; j += 2)
      out[k++] = Fy[i][j];
  }
  return k;
}

/* how many coefficients eh_bend_field_flatten writes */
static int eh_bend_field_count(int order1) {
  int k = 0;
  for (int i = 0; i < order1; i++)
    k += order1 - i;
  return k;
}

#endif
