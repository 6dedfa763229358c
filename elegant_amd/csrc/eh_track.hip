/*
 * eh_track.hip - gfx950 tracking kernel: one particle per lane, FP64.
 *
 * Each lane loads its particle's (x, x', y, y', s, delta) into registers,
 * runs a contiguous run of element-program records for one or many passes,
 * and writes the coordinates back once.  A particle that is lost stops, with
 * its row frozen exactly as the reference leaves a lost row (s = loss
 * position, delta = Po(1+delta)) and its pass and element/stage recorded;
 * the host reproduces the reference's row order from those records.
 *
 * Arithmetic follows the reference routines' operation order with no FP
 * contraction (built with -ffp-contract=off); transcendentals come from
 * emath.h so the CPU checker can reproduce every bit.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include "../../include/elegant_hip.h"
#include "eh_consts.h"
#include "eh_device.h"
#include "eh_aperture.h"
#include "eh_program.h"
#include "emath.h"
#include "ehd_internal.h"

#pragma clang fp contract(off)

#define DEV static __device__ __forceinline__

DEV double sq(double v) { return v * v; }

/* A wave-uniform read of the element program through the constant address
 * space: the records are read-only while a kernel runs, so a uniform address
 * becomes a scalar load instead of a vector load and wait. */
DEV double opc(const double *p) { return *(const __attribute__((address_space(4))) double *)p; }

struct particle {
  double x, xp, y, yp, s, dp;
};

/* true on every active lane: a scalar test, so the rarely taken side of a
 * branch costs one uniform jump instead of a divergent if/else */
#define ALL_LANES(c) (__builtin_amdgcn_ballot_w64(!(c)) == 0)

DEV int imin3(int a, int b, int c) {
  int m = a < b ? a : b;
  return m < c ? m : c;
}

/* ---- frozen rows of lost particles ---- */

/* limit_amplitudes loss (limit_amplitudes.c:232-268): extrapolate along the
 * preceding drift if asked, then s = z + dz, delta = Po(1 + delta) */
DEV void aperture_loss(particle &p, double xmax, double ymax, int is_out, bool extrap, double z, double Po) {
  double dz = 0;
  if (is_out && !(is_out & 4) && extrap) {
    double dzx = -1.7976931348623157e308, dzy = -1.7976931348623157e308;
    if ((is_out & 1) && p.xp != 0)
      dzx = (p.x - (p.xp > 0 ? 1 : -1) * xmax) / p.xp;
    if ((is_out & 2) && p.yp != 0)
      dzy = (p.y - (p.yp > 0 ? 1 : -1) * ymax) / p.yp;
    if (dzx > dzy)
      dz = -dzx;
    else
      dz = -dzy;
    if (dz == -1.7976931348623157e308)
      dz = 0;
  }
  p.x += dz * p.xp;
  p.y += dz * p.yp;
  p.s = z + dz;
  p.dp = Po * (1 + p.dp);
}

/* the two checks that follow every element (do_tracking.c:2459-2477, 2628-2637).
 * Returns 0 if the particle survives, else stage + 1 with the row frozen.
 * In the SH instance an element with a shaped MAXAMP (ap_code != 0) takes
 * elimit_amplitudes or the open-sided limit_amplitudes instead of the
 * rectangle. */
template <bool SH>
DEV int post_element_checks(particle &p, const eh_op &op, const double *z) {
  if (op.flags & EH_F_ALONE)
    return 0;
  if (SH && (op.flags & EH_F_AP) && op.ap_code) {
    if (ap_maxamp_shaped(p.x, p.xp, p.y, p.yp, p.s, p.dp, op.ap_xmax, op.ap_ymax, op.ap_code,
                         (op.flags & EH_F_EXTRAP) != 0, opc(&z[1]), op.Po))
      return EH_STAGE_MAXAMP + 1;
  } else if (op.flags & EH_F_AP) {
    double xmax = op.ap_xmax, ymax = op.ap_ymax;
    int is_out = 0;
    if (xmax > 0 && __builtin_fabs(p.x) > xmax)
      is_out += 1;
    if (ymax > 0 && __builtin_fabs(p.y) > ymax)
      is_out += 2;
    if (__builtin_isinf(p.x) || __builtin_isinf(p.y) || __builtin_isnan(p.x) || __builtin_isnan(p.y))
      is_out += 4;
    if (is_out) {
      aperture_loss(p, xmax, ymax, is_out, (op.flags & EH_F_EXTRAP) != 0, opc(&z[1]), op.Po);
      return EH_STAGE_MAXAMP + 1;
    }
  }
  {
    const double big = 1.7976931348623157e308;
    int is_out = 0;
    if (__builtin_fabs(p.x) > big)
      is_out += 1;
    if (__builtin_fabs(p.y) > big)
      is_out += 2;
    if (__builtin_isinf(p.x) || __builtin_isinf(p.y) || __builtin_isnan(p.x) || __builtin_isnan(p.y))
      is_out += 4;
    if (is_out) {
      aperture_loss(p, big, big, is_out, false, opc(&z[1]), op.Po);
      return EH_STAGE_NAN + 1;
    }
  }
  return 0;
}

/* ---- drifts ---- */

/* The drift's matrix applied as the sparse order-2 loop of matrix.c:389-429
 * does it: for each output i (5..0), sum = C[i], then for each input j
 * (5..0) with coord[j] != 0 (NaN counts as nonzero), sum1 = R[i][j] plus
 * T[i][j][k] coord[k] for k = j..0, and sum += sum1 coord[j].  For the drift
 * (R = 1 + L at (0,1) and (2,3), T[4][1][1] = T[4][3][3] = L/2, C[4] = L)
 * every zero entry still multiplies a coordinate, so an infinite or NaN
 * coordinate turns others into NaN, and zero sums take the sign the loop
 * gives them.  drift2_exact is that loop term for term; drift2 is the same
 * result in a few operations for rows where the zero terms cannot matter. */
DEV double drift_R(int i, int j, double L) {
  return i == j ? 1.0 : (((i == 0 && j == 1) || (i == 2 && j == 3)) ? L : 0.0);
}

DEV void drift2_exact(particle &p, double L, double hL) {
  const double v[6] = {p.x, p.xp, p.y, p.yp, p.s, p.dp};
  double out[6];
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    double sum = i == 4 ? L : 0.0;
#pragma unroll
    for (int j = 5; j >= 0; j--) {
      if (v[j] != 0 || v[j] != v[j]) {
        double sum1 = drift_R(i, j, L);
#pragma unroll
        for (int k = j; k >= 0; k--)
          sum1 += ((i == 4 && ((j == 1 && k == 1) || (j == 3 && k == 3))) ? hL : 0.0) * v[k];
        sum += sum1 * v[j];
      }
    }
    out[i] = sum;
  }
  p.x = out[0];
  p.xp = out[1];
  p.y = out[2];
  p.yp = out[3];
  p.s = out[4];
  p.dp = out[5];
}

/* drift2 equals drift2_exact when every coordinate is finite, |x'| and |y'|
 * are 0 or within [2^-300, 2^300] and the element's 2^-300 <= |L| <= 2^300
 * (EH_F_DFAST): then each zero term is a +0 that leaves a sum unchanged
 * (the sums start at +0 or L), no product of a nonzero slope rounds to zero,
 * and a zero x', y' or delta comes out as +0 (the loop skips it and keeps
 * sum = C[i] = +0). */
DEV bool drift2_ok(const particle &p) {
  const double fin = (__builtin_fabs(p.x) + __builtin_fabs(p.y)) + (__builtin_fabs(p.s) + __builtin_fabs(p.dp));
  return fin <= 0x1.0p+1000 && __builtin_fabs(p.xp) + __builtin_fabs(p.yp) <= 0x1.0p+300 &&
         imin3(__builtin_amdgcn_frexp_exp(p.xp), __builtin_amdgcn_frexp_exp(p.yp), 0) >= -299;
}

DEV void drift2(particle &p, double L, double hL) {
  double s = L;
  if (p.s != 0)
    s = s + p.s;
  if (p.yp != 0)
    s = s + (hL * p.yp) * p.yp;
  if (p.xp != 0)
    s = s + (hL * p.xp) * p.xp;
  double y = (p.yp != 0) ? L * p.yp : 0.0;
  if (p.y != 0)
    y = y + p.y;
  double x = (p.xp != 0) ? L * p.xp : 0.0;
  if (p.x != 0)
    x = x + p.x;
  p.x = x;
  p.y = y;
  p.s = s;
  p.xp = p.xp + 0.0;
  p.yp = p.yp + 0.0;
  p.dp = p.dp + 0.0;
}

/* order-1 matrix branch (matrix.c:430-455): sum = C[i] + R[i][j] ini[j] over
 * j = 5..0 for every output, zeros included */
DEV void drift1(particle &p, double L) {
  const double x0 = p.x, xp0 = p.xp, y0 = p.y, yp0 = p.yp, s0 = p.s, dp0 = p.dp;
  p.dp = 0.0 + 1.0 * dp0 + 0.0 * s0 + 0.0 * yp0 + 0.0 * y0 + 0.0 * xp0 + 0.0 * x0;
  p.s = L + 0.0 * dp0 + 1.0 * s0 + 0.0 * yp0 + 0.0 * y0 + 0.0 * xp0 + 0.0 * x0;
  p.yp = 0.0 + 0.0 * dp0 + 0.0 * s0 + 1.0 * yp0 + 0.0 * y0 + 0.0 * xp0 + 0.0 * x0;
  p.y = 0.0 + 0.0 * dp0 + 0.0 * s0 + L * yp0 + 1.0 * y0 + 0.0 * xp0 + 0.0 * x0;
  p.xp = 0.0 + 0.0 * dp0 + 0.0 * s0 + 0.0 * yp0 + 0.0 * y0 + 1.0 * xp0 + 0.0 * x0;
  p.x = 0.0 + 0.0 * dp0 + 0.0 * s0 + 0.0 * yp0 + 0.0 * y0 + L * xp0 + 1.0 * x0;
}

/* track_particles with a first-order matrix (matrix.c:430-458): each
 * coordinate is C[i] plus R[i][j] x[j] added from j = 5 down to 0 */
DEV void matrix1(particle &p, const double *m) {
  const double x[6] = {p.x, p.xp, p.y, p.yp, p.s, p.dp};
  double o[6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    double sum = m[EH_D_M1_C + i];
#pragma unroll
    for (int j = 5; j >= 0; j--)
      sum += m[EH_D_M1_R + 6 * i + j] * x[j];
    o[i] = sum;
  }
  p.x = o[0];
  p.xp = o[1];
  p.y = o[2];
  p.yp = o[3];
  p.s = o[4];
  p.dp = o[5];
}

DEV void exact_drift(particle &p, double L) {
  p.x += p.xp * L;
  p.y += p.yp * L;
  p.s += L * em_sqrt(1 + sq(p.xp) + sq(p.yp));
}

/* ---- misalignment (malign_mat.c:172-182, tilt_matrices.c:156-188) ---- */

DEV void offset(particle &p, double dx, double dy, double dz) {
  p.s += dz * em_sqrt(1 + sq(p.xp) + sq(p.yp));
  p.x = p.x - dx + dz * p.xp;
  p.y = p.y - dy + dz * p.yp;
}

DEV void rotate(particle &p, double c, double s) {
  double x = p.x, xp = p.xp, y = p.y, yp = p.yp;
  p.x = x * c + y * s;
  p.y = -x * s + y * c;
  p.xp = xp * c + yp * s;
  p.yp = -xp * s + yp * c;
}

/* ---- multipoles ---- */


/* checkMultAperture (multipole.c:1786-1833) for the closed rectangle: the
 * only shape the branch-free bodies see */
DEV bool in_rect(double x, double y, const eh_op &op) {
  if (!(op.flags & EH_F_INAP))
    return true;
  if (op.flags & EH_F_INTILT) {
    double x0 = x, y0 = y;
    x = x0 * op.in_cs0 + y0 * op.in_cs1;
    y = -x0 * op.in_cs1 + y0 * op.in_cs0;
  }
  if ((op.in_xmax > 0 && __builtin_fabs(x) > op.in_xmax) || (op.in_ymax > 0 && __builtin_fabs(y) > op.in_ymax))
    return false;
  return true;
}

/* checkMultAperture for any MAXAMP shape (in_code != 0: ellipse or open
 * side).  SH: the kernel instance built for programs with such shapes; the
 * other instance keeps the closed rectangle only, so the code of the common
 * case is the same as without them. */
template <bool SH>
DEV bool in_aperture(double x, double y, const eh_op &op) {
  if (!SH || !op.in_code)
    return in_rect(x, y, op);
  if (!(op.flags & EH_F_INAP))
    return true;
  if (op.flags & EH_F_INTILT) {
    double x0 = x, y0 = y;
    x = x0 * op.in_cs0 + y0 * op.in_cs1;
    y = -x0 * op.in_cs1 + y0 * op.in_cs0;
  }
  return ap_inside_shaped(x, y, op.in_xmax, op.in_ymax, op.in_code);
}

/* 1 + delta and its reciprocal for the IEEE quotients of emath.h, fixed for
 * the whole element */
struct opd_recip {
  double v, r;
  int ok;
};

DEV opd_recip make_opd(double dp) {
  opd_recip o;
  o.v = 1 + dp;
  o.r = em_recip(o.v);
  o.ok = em_recip_ok(o.v);
  return o;
}

/* convertMomentaToSlopes (multipole.h:266-281) */
DEV bool momenta_to_slopes(double &xp, double &yp, double qx, double qy, double dp, bool expand,
                           const opd_recip &o) {
  if (expand) {
    xp = em_div(qx, o.v, o.r, o.ok);
    yp = em_div(qy, o.v, o.r, o.ok);
    return true;
  }
  double denom = sq(1 + dp) - sq(qx) - sq(qy);
  if (denom <= 0)
    return false;
  denom = em_sqrt(denom);
  const double r = em_recip(denom);
  const int ok = em_recip_ok(denom);
  xp = em_div(qx, denom, r, ok);
  yp = em_div(qy, denom, r, ok);
  return true;
}

/* sum_Fx, sum_Fy of apply_canonical_multipole_kicks (multipole.h:208-241)
 * for order n, written out: coefficients (-1)^floor(i/2)/(i!(n-i)!) are
 * 1,1 | 1/2,1,-1/2 | 1/6,1/2,-1/2,-1/6, and products with 1 or sums with 0
 * in the reference loop are exact. */
DEV void multipole_fields(int order, double x, double y, double &Fx, double &Fy) {
  if (order == 1) {
    Fx = y;
    Fy = x;
  } else if (order == 2) {
    Fx = x * y;
    Fy = 0.5 * (x * x) + -0.5 * (y * y);
  } else if (order == 3) {
    double x2 = x * x, y2 = y * y;
    Fx = (0.5 * x2) * y + (-1.0 / 6.0) * (y2 * y);
    Fy = (1.0 / 6.0) * (x2 * x) + (-0.5 * x) * y2;
  } else { /* order 0 */
    Fx = 0;
    Fy = 1;
  }
}

DEV void apply_kick(double &qx, double &qy, int order, int skew, double KnL, double x, double y) {
  double Fx, Fy;
  multipole_fields(order, x, y, Fx, Fy);
  if (skew) {
    double t = Fx;
    Fx = -Fy;
    Fy = t;
  }
  qx -= KnL * Fy;
  qy += KnL * Fx;
}

/* The slices of integrate_kick_multipole_ordn (multipole.c:1122-1410), with
 * op.nsub drift substeps per slice (per-substep coefficients are scalar
 * loads of the element record).
 * Returns 0, or the loss kind: 1 aperture (x, y hold the loss point),
 * 2 slopes (the row is left as received). */
template <bool SH>
DEV int multipole_body(double &x, double &xp, double &y, double &yp, double &s, double &qx, double &qy,
                       double &dzLoss, const double dp, const bool expand, const opd_recip &o, const eh_op &op) {
  const int nsub = op.nsub;
  const int flags = op.flags;
  const double dxa = op.d[EH_D_DX], dya = op.d[EH_D_DY];
  const bool has_drift = op.d[EH_D_DRIFT] != 0;
  const int nko = op.nkick_orders;
  for (int ik = 0; ik < op.nslices; ik++) {
    if (!in_aperture<SH>(x + dxa, y + dya, op))
      return 1;
    for (int step = 0; step < nsub; step++) {
      if (has_drift) {
        const double dsh = opc(&op.d[EH_D_DSH + step]);
        x += xp * dsh;
        y += yp * dsh;
        if (expand)
          s += dsh * (1 + (sq(xp) + sq(yp)) / 2);
        else
          s += dsh * em_sqrt(1 + sq(xp) + sq(yp));
        dzLoss += dsh;
      }
      if (step == nsub - 1)
        break; /* kickFrac[last] == 0 */
      if (nko > 0)
        apply_kick(qx, qy, op.order0, op.skew0, opc(&op.d[EH_D_KICK + step]), x, y);
      if (nko > 1)
        apply_kick(qx, qy, op.order1, op.skew1, opc(&op.d[EH_D_KICK + 8 + step]), x, y);
      if (nko > 2)
        apply_kick(qx, qy, op.order2, op.skew2, opc(&op.d[EH_D_KICK + 16 + step]), x, y);
      if (flags & EH_F_XKICK)
        apply_kick(qx, qy, 0, 0, opc(&op.d[EH_D_XK + step]), x, y);
      if (flags & EH_F_YKICK)
        apply_kick(qx, qy, 0, 1, opc(&op.d[EH_D_YK + step]), x, y);
      if (!momenta_to_slopes(xp, yp, qx, qy, dp, expand, o))
        return 2;
      if (has_drift) {
        /* rad_coef = 0, isr_coef < 0: the q/(1+dp)*(1+dp) round trip (multipole.c:1321-1351) */
        qx = em_div(qx, o.v, o.r, o.ok);
        qy = em_div(qy, o.v, o.r, o.ok);
        qx *= (1 + dp);
        qy *= (1 + dp);
        if (!momenta_to_slopes(xp, yp, qx, qy, dp, expand, o))
          return 2;
      }
    }
  }
  return 0;
}

/* multipole_body on its common path (EXPAND_HAMILTONIAN=0), with no
 * per-lane branches: the unscaled square roots and quotients by positive
 * divisors (emath.h em_sqrt_mid, em_div_pos), no loss.  Returns false for a
 * lane where that may not hold; the caller then recomputes the lane from
 * the row it received with the general code.  Checked per substep, into
 * one lane mask:
 *   - 1 + x'^2 + y'^2 <= 2^1000 for the drift's square root (a NaN fails);
 *   - after the kicks, |qx| + |qy| <= 2^700 and qx, qy are 0 or >= 2^-700
 *     in magnitude (binary exponent >= -699), so every dividend below is 0
 *     or within [2^-900, 2^900];
 *   - (1 + delta)^2 - qx^2 - qy^2 >= 2^-128 at both conversions to slopes:
 *     no slope loss, the root lies in [2^-64, 2^64];
 *   - 1 + delta in [2^-64, 2^64]; the aperture holds at every slice.
 * dzLoss is not accumulated: the caller sets the full length's sum.
 * Arithmetic is otherwise multipole_body's, operation for operation. */
template <int ORD, int NSUB>
DEV bool multipole_body_fast(double &x, double &xp, double &y, double &yp, double &s, double &qx, double &qy,
                             const double dp, const opd_recip &o, const eh_op &op) {
  const int nsub = NSUB ? NSUB : op.nsub;
  const int flags = op.flags;
  const double dxa = op.d[EH_D_DX], dya = op.d[EH_D_DY];
  const bool has_drift = ORD > 0 || op.d[EH_D_DRIFT] != 0;
  const int nko = op.nkick_orders;
  bool good = o.v >= 0x1.0p-64 && o.v <= 0x1.0p+64;
  for (int ik = 0; ik < op.nslices; ik++) {
    good = good && in_rect(x + dxa, y + dya, op);
#pragma unroll
    for (int step = 0; step < nsub; step++) {
      if (has_drift) {
        const double dsh = opc(&op.d[EH_D_DSH + step]);
        x += xp * dsh;
        y += yp * dsh;
        const double t = 1 + sq(xp) + sq(yp);
        good = good && t <= 0x1.0p+1000;
        s += dsh * em_sqrt_mid(t);
      }
      if (step == nsub - 1)
        break; /* kickFrac[last] == 0 */
      if (ORD > 0) {
        apply_kick(qx, qy, ORD, 0, opc(&op.d[EH_D_KICK + step]), x, y);
      } else {
        if (nko > 0)
          apply_kick(qx, qy, op.order0, op.skew0, opc(&op.d[EH_D_KICK + step]), x, y);
        if (nko > 1)
          apply_kick(qx, qy, op.order1, op.skew1, opc(&op.d[EH_D_KICK + 8 + step]), x, y);
        if (nko > 2)
          apply_kick(qx, qy, op.order2, op.skew2, opc(&op.d[EH_D_KICK + 16 + step]), x, y);
        if (flags & EH_F_XKICK)
          apply_kick(qx, qy, 0, 0, opc(&op.d[EH_D_XK + step]), x, y);
        if (flags & EH_F_YKICK)
          apply_kick(qx, qy, 0, 1, opc(&op.d[EH_D_YK + step]), x, y);
      }
      good = good && __builtin_fabs(qx) + __builtin_fabs(qy) <= 0x1.0p+700 &&
             imin3(__builtin_amdgcn_frexp_exp(qx), __builtin_amdgcn_frexp_exp(qy), 0) >= -699;
      double den = sq(1 + dp) - sq(qx) - sq(qy);
      double dmin = den;
      den = em_sqrt_mid(den);
      double r = em_recip(den);
      xp = em_div_pos(qx, den, r);
      yp = em_div_pos(qy, den, r);
      if (has_drift) {
        /* rad_coef = 0, isr_coef < 0: the q/(1+dp)*(1+dp) round trip (multipole.c:1321-1351) */
        qx = em_div_pos(qx, o.v, o.r);
        qy = em_div_pos(qy, o.v, o.r);
        qx *= (1 + dp);
        qy *= (1 + dp);
        den = sq(1 + dp) - sq(qx) - sq(qy);
        dmin = __builtin_fmin(dmin, den);
        den = em_sqrt_mid(den);
        r = em_recip(den);
        xp = em_div_pos(qx, den, r);
        yp = em_div_pos(qy, den, r);
      }
      good = good && dmin >= 0x1.0p-128;
    }
  }
  return good;
}

/* the integrator's entry (multipole.c:1176-1205): limits, slopes to
 * momenta and back.  Returns 0, or 2 for a slope loss (dzLoss set). */
DEV int multipole_entry(const particle &p, double &xp, double &yp, double &qx, double &qy, double &dzLoss,
                        const bool expand, const opd_recip &o, const eh_op &op) {
  const double dp = p.dp;
  xp = p.xp;
  yp = p.yp;
  if (__builtin_isnan(p.x) || __builtin_isnan(xp) || __builtin_isnan(p.y) || __builtin_isnan(yp) ||
      __builtin_fabs(p.x) > EH_COORD_LIMIT || __builtin_fabs(p.y) > EH_COORD_LIMIT ||
      __builtin_fabs(xp) > EH_SLOPE_LIMIT || __builtin_fabs(yp) > EH_SLOPE_LIMIT) {
    dzLoss = op.d[EH_D_TOTLEN];
    return 2;
  }
  if (expand) {
    qx = (1 + dp) * xp;
    qy = (1 + dp) * yp;
  } else {
    double denom = em_sqrt(1 + sq(xp) + sq(yp));
    qx = (1 + dp) * xp / denom;
    qy = (1 + dp) * yp / denom;
  }
  if (!momenta_to_slopes(xp, yp, qx, qy, dp, expand, o)) {
    dzLoss = op.d[EH_D_TOTLEN];
    return 2;
  }
  return 0;
}

/* multipole_tracking2 per particle: entry misalignment, the integrator of
 * integrate_kick_multipole_ordn (multipole.c:1122-1410), exit misalignment.
 * Returns 0 if alive, EH_STAGE_ELEMENT + 1 if lost (row frozen). */
template <bool SH>
DEV int track_multipole(particle &p, const eh_op &op, const double *z) {
  const int flags = op.flags;
  const bool expand = (flags & EH_F_EXPAND) != 0;
  if (flags & EH_F_OFFSET)
    offset(p, op.d[EH_D_DX], op.d[EH_D_DY], op.d[EH_D_DZ]);
  if (flags & EH_F_ROTATE)
    rotate(p, op.d[EH_D_COS_IN], op.d[EH_D_SIN_IN]);
  /* KQUAD LEFFECTIVE: an exact drift before the body (multipole.c:1028-1030) */
  if (flags & EH_F_LEND)
    exact_drift(p, op.d[EH_D_LEND]);

  /* the row as the integrator receives it stays in p until the end */
  const double dp = p.dp;
  const opd_recip o = make_opd(dp);
  double x = p.x, xp = 0, y = p.y, yp = 0, s = 0, qx = 0, qy = 0, dzLoss = 0;
  const double dxa = op.d[EH_D_DX], dya = op.d[EH_D_DY];
  /* 1: aperture (x, y written), 2: slopes (nothing written), 3: final limits */
  int lost = multipole_entry(p, xp, yp, qx, qy, dzLoss, expand, o, op);
  if (!lost) {
    bool done = false;
    if (!expand && !(SH && op.in_code)) {
      /* the common thick element with one normal multipole and no steering
       * gets a body specialised to its order: the kick is straight-line code */
      const int ord = (op.nkick_orders == 1 && !op.skew0 && !(flags & (EH_F_XKICK | EH_F_YKICK)) &&
                       op.d[EH_D_DRIFT] != 0)
                          ? op.order0
                          : 0;
      const bool four = op.nsub == 4;
      if (ord == 1 && four)
        done = multipole_body_fast<1, 4>(x, xp, y, yp, s, qx, qy, dp, o, op);
      else if (ord == 2 && four)
        done = multipole_body_fast<2, 4>(x, xp, y, yp, s, qx, qy, dp, o, op);
      else if (ord == 3 && four)
        done = multipole_body_fast<3, 4>(x, xp, y, yp, s, qx, qy, dp, o, op);
      else
        done = multipole_body_fast<0, 0>(x, xp, y, yp, s, qx, qy, dp, o, op);
      /* the checks after the integrator may still lose the particle: at the
       * element's full length, the sum the general code would have made */
      dzLoss = done ? op.d[EH_D_TOTLEN] : 0.0;
      if (!ALL_LANES(done) && !done) {
        x = p.x;
        y = p.y;
        s = 0;
        multipole_entry(p, xp, yp, qx, qy, dzLoss, expand, o, op);
      }
    }
    if (!ALL_LANES(done) && !done)
      lost = multipole_body<SH>(x, xp, y, yp, s, qx, qy, dzLoss, dp, expand, o, op);
    if (!lost) {
      if (!in_aperture<SH>(x + dxa, y + dya, op))
        lost = 1;
      else if (!momenta_to_slopes(xp, yp, qx, qy, dp, expand, o))
        lost = 2;
      else {
        p.x = x;
        p.xp = xp;
        p.y = y;
        p.yp = yp;
        p.s += s;
        p.dp = dp;
        if (__builtin_isnan(x) || __builtin_isnan(xp) || __builtin_isnan(y) || __builtin_isnan(yp) ||
            __builtin_fabs(x) > EH_COORD_LIMIT || __builtin_fabs(y) > EH_COORD_LIMIT ||
            __builtin_fabs(xp) > EH_SLOPE_LIMIT || __builtin_fabs(yp) > EH_SLOPE_LIMIT)
          lost = 3;
      }
    }
  }
  if (lost) {
    if (lost == 1) {
      p.x = x;
      p.y = y;
    }
    /* lost == 2: row as received; lost == 3: row fully written */
    p.s = opc(&z[0]) + dzLoss;
    p.dp = op.Po * (1 + p.dp);
  }
  /* the exit drift of LEFFECTIVE and the exit misalignment, applied to
   * lost rows too (multipole.c:1098-1104) */
  if (flags & EH_F_LEND)
    exact_drift(p, op.d[EH_D_LEND]);
  if (flags & EH_F_ROTATE)
    rotate(p, op.d[EH_D_COS_OUT], op.d[EH_D_SIN_OUT]);
  if (flags & EH_F_OFFSET)
    offset(p, -op.d[EH_D_DX], -op.d[EH_D_DY], -op.d[EH_D_DZ]);
  return lost ? EH_STAGE_ELEMENT + 1 : 0;
}

/* ---- CSBEND ---- */

/* The slices of integrate_csbend_ordn (csbend.c:1434-1693) for a pure
 * dipole (Fx = 0, Fy = 1, csbend.c:106-110), op.nsub drift substeps per
 * slice; per-substep coefficients are scalar loads of the element record.  Returns true if the particle is lost; X, QX, Y, QY, dist
 * and dz_lost then hold their values at the loss, as the reference leaves
 * them. */
/* computeCSBENDFields (csbend.c:100-141), normal multipoles: the
 * coefficients in the order eh_bend_field_flatten lays them out, each term
 * (F x^i) y^j with the powers of y built by repeated multiplication as the
 * reference's table of powers is */
DEV void bend_fields(const eh_op &op, double x, double y, double &Fx, double &Fy) {
  const int n1 = op.nfield;
  double sumFx = 0, sumFy = 0, xt = 1;
  int k = 0;
  for (int i = 0; i < n1; i++) {
    double p = y; /* y^1 */
    for (int j = 1; j < n1 - i; j += 2) {
      sumFx += opc(&op.d[EH_D_FIELD + k]) * xt * p;
      k++;
      p *= y;
      p *= y;
    }
    p = 1.0; /* y^0 */
    for (int j = 0; j < n1 - i; j += 2) {
      sumFy += opc(&op.d[EH_D_FIELD + k]) * xt * p;
      k++;
      p *= y;
      p *= y;
    }
    xt *= x;
  }
  Fx = sumFx;
  Fy = sumFy;
}

template <bool SH>
DEV bool csbend_body(double &X, double &QX, double &Y, double &QY, const double DP, double &dist, double &dz_lost,
                     const eh_op &op) {
  const double rho0 = op.d[EH_D_RHO0], rhoa = op.d[EH_D_RHOA], rsgn = op.d[EH_D_RHOA_SGN];
  const double opd = 1 + DP;
  const int nsub = op.nsub;
  /* reciprocals shared by the divisions below (IEEE quotients, emath.h);
   * those of rho0 and rho_actual are rounded on the host */
  const double r_opd = em_recip(opd);
  const int ok_opd = em_recip_ok(opd);
  const int ok_rho = (op.flags & EH_F_RCP) != 0;
  for (int i = 0; i < op.nslices; i++) {
    if (!in_aperture<SH>(X, Y, op))
      return true;
    for (int j = 0; j < nsub; j++) {
      double f = sq(opd) - sq(QY);
      if (f <= 0)
        return true;
      f = em_sqrt(f);
      const double r_f = em_recip(f);
      const int ok_f = em_recip_ok(f);
      double sin_phi = em_div(QX, f, r_f, ok_f);
      if (__builtin_fabs(sin_phi) > 1)
        return true;
      double phi = em_asin(sin_phi);
      double sine, cosi;
      em_sincos(opc(&op.d[EH_D_BDSHR + j]) + phi, &sine, &cosi);
      if (cosi == 0)
        return true;
      const double r_c = em_recip(cosi);
      const int ok_c = em_recip_ok(cosi);
      double tang = em_div(sine, cosi, r_c, ok_c);
      double cos_phi = em_cos(phi);
      QX = f * sine;
      double factor = em_div((rho0 + X) * cos_phi, f, r_f, ok_f) * (tang - sin_phi / cos_phi);
      Y += QY * factor;
      dist += factor * opd;
      dz_lost += opc(&op.d[EH_D_BDSH + j]);
      f = em_div(cos_phi, cosi, r_c, ok_c);
      X = rho0 * (f - 1) + f * X;
      if (j == nsub - 1)
        break; /* kickFrac[last] == 0 */
      double ds = opc(&op.d[EH_D_BDS + j]);
      double g = 1 + em_div(X, rho0, opc(&op.d[EH_D_RRHO0]), ok_rho);
      if (op.flags & EH_F_FIELD) {
        double Fx, Fy;
        bend_fields(op, X, Y, Fx, Fy);
        QX += em_div(-ds * g * Fy, rhoa, opc(&op.d[EH_D_RRHOA]), ok_rho);
        QY += em_div(ds * g * Fx, rhoa, opc(&op.d[EH_D_RRHOA]), ok_rho);
      } else {
        QX += em_div(-ds * g * 1.0, rhoa, opc(&op.d[EH_D_RRHOA]), ok_rho);
        /* Fx = 0: ds*g*0.0 is +-0 or NaN, and dividing it by rho_actual only
         * sets its sign, which the multiplication by sign(rho_actual) does */
        QY += ds * g * 0.0 * rsgn;
      }
      QX = em_div(QX, opd, r_opd, ok_opd);
      QY = em_div(QY, opd, r_opd, ok_opd);
      QX *= opd;
      QY *= opd;
    }
  }
  return false;
}


/* csbend_body on its common path, with no per-lane branches: every substep
 * takes the narrow bands of asin/sin/cos (z < 2^-10, emath.h *_small), the
 * unscaled quotients by positive divisors (em_div_pos) and the unscaled
 * square root, and no loss occurs.  Returns false for a lane where any of
 * that may not hold; the caller then discards the lane's results and runs
 * csbend_body from the saved state.  What makes every step valid, checked
 * once at the end from running extrema (which add no branches):
 *   - |X| + |QX| + |QY| <= 2^700 at every substep (a NaN or infinity in the
 *     state fails it, and the state is the only source of one);
 *   - X, QX, QY are 0 or >= 2^-700 in magnitude at every substep (binary
 *     exponent >= -699); then every dividend below is 0 or within
 *     [2^-900, 2^900]: sums of such terms with the radii or with 1 are 0 or
 *     at least an ulp of 2^-764;
 *   - opd^2 - QY^2 >= 2^-128: the square root is in range, f lies in
 *     [2^-64, 2^64] and is positive, and the f <= 0 loss cannot occur;
 *   - sin(phi)^2, (angle)^2 and phi^2 < 2^-10 (the narrow bands; also rules
 *     out |sin(phi)| > 1 and cos(angle) = 0), |angle| >= 2^-900, which given
 *     the rest means angle != 0 (em_sincos's zero case), so cos(angle) and
 *     cos(phi) exceed 0.999 (this bound and the one above share a running
 *     minimum, of (opd^2 - QY^2) 2^-772 and |angle|);
 *   - dz_lost is not accumulated: it only matters for a lost particle, and
 *     the fast path never loses one;
 *   - 1 + delta in [2^-64, 2^64]; the radii, angles and kick lengths are
 *     checked on the host (EH_F_FAST);
 *   - the aperture holds at every slice.
 * Arithmetic is otherwise csbend_body's, operation for operation. */
template <int NSUB>
DEV bool csbend_body_fast(double &X, double &QX, double &Y, double &QY, const double DP, double &dist,
                          const eh_op &op) {
  const double rho0 = op.d[EH_D_RHO0], rhoa = op.d[EH_D_RHOA];
  const double opd = 1 + DP;
  const int nsub = NSUB ? NSUB : op.nsub;
  bool good = opd >= 0x1.0p-64 && opd <= 0x1.0p+64;
  const double r_opd = em_recip(opd);
  double acc = 0, zmax = 0, lo = 1;
  int emin = 0;
  for (int i = 0; i < op.nslices; i++) {
    good = good && in_rect(X, Y, op);
#pragma unroll
    for (int j = 0; j < nsub; j++) {
      acc += (__builtin_fabs(X) + __builtin_fabs(QX)) + __builtin_fabs(QY);
      emin = imin3(emin, __builtin_amdgcn_frexp_exp(X), imin3(emin, __builtin_amdgcn_frexp_exp(QX),
                                                              __builtin_amdgcn_frexp_exp(QY)));
      double f = sq(opd) - sq(QY);
      const double f_in = f;
      f = em_sqrt_mid(f);
      const double r_f = em_recip(f);
      double sin_phi = em_div_pos(QX, f, r_f);
      const double zs = sin_phi * sin_phi;
      double phi = em_asin_small(sin_phi, zs);
      const double ang = opc(&op.d[EH_D_BDSHR + j]) + phi;
      const double za = ang * ang;
      lo = __builtin_fmin(lo, __builtin_fmin(f_in * 0x1.0p-772, __builtin_fabs(ang)));
      double sine = em_sin_small(ang, za), cosi = em_cos_small(ang, za);
      const double r_c = em_recip(cosi);
      double tang = em_div_pos(sine, cosi, r_c);
      const double zp = phi * phi;
      zmax = __builtin_fmax(zmax, __builtin_fmax(zs, __builtin_fmax(za, zp)));
      double cos_phi = em_cos_small(phi, zp);
      QX = f * sine;
      double factor = em_div_pos((rho0 + X) * cos_phi, f, r_f) *
                      (tang - em_div_pos(sin_phi, cos_phi, em_recip(cos_phi)));
      Y += QY * factor;
      dist += factor * opd;
      f = em_div_pos(cos_phi, cosi, r_c);
      X = rho0 * (f - 1) + f * X;
      if (j == nsub - 1)
        break; /* kickFrac[last] == 0 */
      double ds = opc(&op.d[EH_D_BDS + j]);
      double g = 1 + em_div_pos(X, rho0, opc(&op.d[EH_D_RRHO0]));
      /* (-ds) g 1.0 = -(ds g) and rsgn = +1 here (rho_actual > 0): the same
       * bits as the general code's products, one product fewer */
      const double dsg = ds * g;
      QX += em_div_pos(-dsg, rhoa, opc(&op.d[EH_D_RRHOA]));
      QY += dsg * 0.0;
      QX = em_div_pos(QX, opd, r_opd);
      QY = em_div_pos(QY, opd, r_opd);
      QX *= opd;
      QY *= opd;
    }
  }
  /* the last substep's X and the last kick's QX, QY need no check: no
   * quotient takes them before the next element re-checks */
  return good && acc <= 0x1.0p+700 && emin >= -699 && lo >= 0x1.0p-900 && zmax < EM_SMALL_Z;
}

template <bool SH>
DEV int track_csbend(particle &p, const eh_op &op, const double *z) {
  const int flags = op.flags;
  const double ct = op.d[EH_D_CT], st = op.d[EH_D_ST];
  const double rho0 = op.d[EH_D_RHO0], rhoa = op.d[EH_D_RHOA];
  double x, xp, y, yp, s, dp, dp0;
  /* entry: misalignment then rotation by the total tilt (csbend.c:1120-1131) */
  p.s += op.d[EH_D_DZI] * em_sqrt(1 + sq(p.xp) + sq(p.yp));
  p.x = p.x + op.d[EH_D_DXI] + op.d[EH_D_DZI] * p.xp;
  p.y = p.y + op.d[EH_D_DYI] + op.d[EH_D_DZI] * p.yp;
  x = p.x * ct + p.y * st;
  y = -p.x * st + p.y * ct;
  xp = p.xp * ct + p.yp * st;
  yp = -p.xp * st + p.yp * ct;
  s = p.s;
  dp = dp0 = p.dp;
  if (flags & EH_F_EDGE1) {
    double rho = (1 + dp) * rhoa;
    double delta_xp = op.d[EH_D_TANE1] / rho * x;
    if ((flags & EH_F_KLIM1) && __builtin_fabs(delta_xp) > op.d[EH_D_KLIM1])
      delta_xp = (delta_xp > 0 ? 1 : -1) * op.d[EH_D_KLIM1];
    xp += delta_xp;
    double te = em_tan(op.d[EH_D_E1] - op.d[EH_D_PSI1] / (1 + dp));
    yp -= te / rho * y;
  }
  /* canonical (csbend.c:1411-1419) */
  double X = x, QX, Y = y, QY, S = 0, DP = dp;
  {
    double f = (1 + DP) / em_sqrt(1 + sq(xp) + sq(yp));
    QX = xp * f;
    QY = yp * f;
  }
  /* integrate_csbend_ordn (csbend.c:1434-1693), Fx = 0, Fy = 1 */
  double dist = 0, dz_lost = 0;
  bool lost = false, done = false;
  if ((flags & EH_F_FAST) && !(SH && op.in_code)) {
    const double X0 = X, QX0 = QX, Y0 = Y, QY0 = QY;
    /* the 4th-order integrator's four substeps of a slice unrolled (NSUB =
     * 4): the substep coefficients load once per slice and the scheduler
     * overlaps neighbouring substeps */
    done = op.nsub == 4 ? csbend_body_fast<4>(X, QX, Y, QY, DP, dist, op)
                        : csbend_body_fast<0>(X, QX, Y, QY, DP, dist, op);
    if (!ALL_LANES(done) && !done) {
      X = X0;
      QX = QX0;
      Y = Y0;
      QY = QY0;
      dist = 0;
    }
  }
  if (!ALL_LANES(done) && !done)
    lost = csbend_body<SH>(X, QX, Y, QY, DP, dist, dz_lost, op);
  if (!lost) {
    dz_lost = op.d[EH_D_NS];
    if (!in_aperture<SH>(X, Y, op))
      lost = true;
    else
      S += dist;
  }
  {
    /* convertFromDipoleCanonicalCoordinates (csbend.c:1421-1429) */
    double f = 1 / em_sqrt(sq(1 + DP) - sq(QX) - sq(QY));
    QX *= f;
    QY *= f;
  }
  if (lost) {
    /* memcpy(Qf) + convertFromCSBendCoords (csbend.c:1237-1239, 4502-4524) */
    p.x = X * ct - Y * st;
    p.y = X * st + Y * ct;
    p.xp = QX * ct - QY * st;
    p.yp = QX * st + QY * ct;
    p.s = opc(&z[0]) + dz_lost;
    p.dp = op.Po * (1 + DP);
    return EH_STAGE_ELEMENT + 1;
  }
  /* velocity rescaling of s (csbend.c:1280-1292) */
  {
    double p0 = op.Po * (1 + dp0);
    double beta0 = p0 / em_sqrt(sq(p0) + 1);
    double p1 = op.Po * (1 + DP);
    double beta1 = p1 / em_sqrt(sq(p1) + 1);
    s = beta1 * s / beta0 + S;
  }
  x = X;
  xp = QX;
  y = Y;
  yp = QY;
  dp = DP;
  if (flags & EH_F_EDGE2) {
    double rho = (1 + dp) * rhoa;
    double delta_xp = op.d[EH_D_TANE2] / rho * x;
    if ((flags & EH_F_KLIM2) && __builtin_fabs(delta_xp) > op.d[EH_D_KLIM2])
      delta_xp = (delta_xp > 0 ? 1 : -1) * op.d[EH_D_KLIM2];
    xp += delta_xp;
    double te = em_tan(op.d[EH_D_E2] - op.d[EH_D_PSI2] / (1 + dp));
    yp -= te / rho * y;
  }
  /* the error tilt's exit offsets, zero without one (csbend.c:1359-1364) */
  double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0, o4 = 0.0;
  if (flags & EH_F_ETILT) {
    o0 = opc(&op.d[EH_D_ET]);
    o1 = opc(&op.d[EH_D_ET + 1]);
    o2 = opc(&op.d[EH_D_ET + 2]);
    o3 = opc(&op.d[EH_D_ET + 3]);
    o4 = opc(&op.d[EH_D_ET + 4]);
  }
  p.x = x * ct - y * st + o0;
  p.y = x * st + y * ct + o2;
  p.xp = xp * ct - yp * st + o1;
  p.yp = xp * st + yp * ct + o3;
  p.s = s + o4;
  p.dp = dp;
  p.x += op.d[EH_D_DXF] + op.d[EH_D_DZF] * p.xp;
  p.y += op.d[EH_D_DYF] + op.d[EH_D_DZF] * p.yp;
  p.s += op.d[EH_D_DZF] * em_sqrt(1 + sq(p.xp) + sq(p.yp));
  return 0;
}

/* ---- RFCA ---- */

/* add_to_particle_energy (simple_rfca.c:987-1007) */
DEV void add_energy(particle &p, double tof, double Po, double dgamma) {
  double P = Po * (1 + p.dp);
  double gamma = em_sqrt(P * P + 1);
  double gamma1 = gamma + dgamma;
  if (gamma1 <= 1)
    gamma1 = 1 + 1e-7;
  double P1 = em_sqrt(gamma1 * gamma1 - 1);
  p.dp = (P1 - Po) / Po;
  p.s = tof * EH_C_MKS * P1 / gamma1;
  double Pz = P / em_sqrt(1 + sq(p.xp) + sq(p.yp));
  double Pz1 = em_sqrt(Pz * Pz + gamma1 * gamma1 - gamma * gamma);
  double ratio = Pz / Pz1;
  p.xp *= ratio;
  p.yp *= ratio;
}

/* trackRfCavityWithWakes for one particle, forward tracking, no wakes
 * (simple_rfca.c:557-984) followed by removeInvalidParticles (:975) */
DEV int track_rfca(particle &p, const eh_op &op, const double *z) {
  const double Po = op.Po;
  const double volt = op.d[EH_D_VOLT], omega = op.d[EH_D_OMEGA], phase = op.d[EH_D_PHASE];
  const double length = op.d[EH_D_LENGTH], dtLight = op.d[EH_D_DTLIGHT];
  const int nKicks = op.nkicks_rf;
  /* CHANGE_T (simple_rfca.c:488-497): whole RF periods in the arrival time
   * of the beam's first row, which the host supplies, or of the particle
   * itself when every particle is its own beam */
  double timeOffset = op.d[EH_D_TOFF];
  if (op.flags & EH_F_CHT_SELF) {
    const double To = op.d[EH_D_TO];
    double P = Po * (1 + p.dp);
    double beta_i = P / em_sqrt(sq(P) + 1);
    double t = p.s / (EH_C_MKS * beta_i);
    timeOffset = t > (0.9 * To) ? (double)(int64_t)(t / To + 0.5) * To : 0.0;
  }
  const bool change_t = (op.flags & EH_F_CHANGE_T) && timeOffset != 0;
  p.x -= op.d[EH_D_RFDX];
  p.y -= op.d[EH_D_RFDY];
  if (!(op.flags & EH_F_MATRIX)) {
    double inverseF = 0;
    for (int ik = 0; ik < nKicks; ik++) {
      if (p.dp != -1) {
        double dc4 = length ? length / 2 * em_sqrt(1 + sq(p.xp) + sq(p.yp)) : 0;
        double P = Po * (1 + p.dp);
        double gamma = em_sqrt(sq(P) + 1);
        double beta_i = P / gamma;
        double t = (p.s + dc4) / (EH_C_MKS * beta_i) - timeOffset;
        if (ik == 0 && change_t)
          p.s = t * EH_C_MKS * beta_i - dc4;
        double phi = omega * (t - 0 - ik * dtLight) + phase;
        double dgamma = volt * em_sin(phi) * 1;
        if (length) {
          if ((op.flags & EH_F_END1FOCUS) && ik == 0) {
            inverseF = dgamma / (2 * gamma * length);
            p.xp -= p.x * inverseF;
            p.yp -= p.y * inverseF;
          }
          p.x += p.xp * length / 2;
          p.y += p.yp * length / 2;
          p.s += length / 2 * em_sqrt(1 + sq(p.xp) + sq(p.yp));
        }
        add_energy(p, t, Po, dgamma);
        double gamma1 = gamma + dgamma;
        if (gamma1 <= 1)
          p.dp = -1;
        else
          inverseF = -dgamma / (2 * gamma1 * length);
      }
      if (length) {
        p.x += p.xp * length / 2;
        p.y += p.yp * length / 2;
        p.s += length / 2 * em_sqrt(1 + sq(p.xp) + sq(p.yp));
        if ((op.flags & EH_F_END2FOCUS) && ik == nKicks - 1) {
          p.xp -= p.x * inverseF;
          p.yp -= p.y * inverseF;
        }
      }
    }
  } else {
    for (int ik = 0; ik < nKicks; ik++) {
      double P = Po * (1 + p.dp);
      double gamma = em_sqrt(sq(P) + 1);
      double beta_i = P / gamma;
      double ds1 = length / 2 * em_sqrt(1 + sq(p.xp) + sq(p.yp));
      double t = (p.s + ds1) / (EH_C_MKS * beta_i) - timeOffset;
      if (change_t)
        p.s = t * EH_C_MKS * beta_i - ds1;
      double sin_phase = em_sin(omega * (t - 0 - ik * dtLight) + phase);
      double dgamma = volt * 1 * sin_phase;
      if ((op.flags & EH_F_END1FOCUS) && length && ik == 0) {
        double inverseF = dgamma / (2 * gamma * length);
        p.xp -= p.x * inverseF;
        p.yp -= p.y * inverseF;
      }
      double dP = em_sqrt(sq(gamma + dgamma) - 1) - P;
      double R22 = 1 / (1 + dP / P), R12, R11 = 1, R21 = 0;
      if (__builtin_fabs(dP / P) > 1e-14)
        R12 = length * (P / dP * em_log(1 + dP / P));
      else
        R12 = length;
      p.s += ds1;
      double xx = p.x, xxp = p.xp;
      p.x = xx * R11 + xxp * R12;
      p.xp = xx * R21 + xxp * R22;
      xx = p.y;
      xxp = p.yp;
      p.y = xx * R11 + xxp * R12;
      p.yp = xx * R21 + xxp * R22;
      p.s += length / 2 * em_sqrt(1 + sq(p.xp) + sq(p.yp));
      p.dp = (P + dP - Po) / Po;
      if ((gamma += dgamma) <= 1)
        p.dp = -1;
      if ((op.flags & EH_F_END2FOCUS) && length && ik == nKicks - 1) {
        double inverseF = -dgamma / (2 * gamma * length);
        p.xp -= p.x * inverseF;
        p.yp -= p.y * inverseF;
      }
      p.s = (P + dP) / gamma * p.s / beta_i;
    }
  }
  p.x += op.d[EH_D_RFDX];
  p.y += op.d[EH_D_RFDY];
  /* removeInvalidParticles (limit_amplitudes.c:274-320) */
  if (__builtin_isnan(p.x) || __builtin_isnan(p.xp) || __builtin_isnan(p.y) || __builtin_isnan(p.yp) ||
      __builtin_isnan(p.s) || __builtin_isnan(p.dp) || p.dp <= -1) {
    p.s = opc(&z[1]);
    p.dp = Po * (1 + p.dp);
    return EH_STAGE_ELEMENT + 1;
  }
  return 0;
}

/* offset_beam (do_tracking.c:3243-3306) for one particle, on pass ON_PASS
 * or, with -1, every pass (do_tracking.c:1818-1822) */
DEV void track_malign(particle &p, const eh_op &op, int64_t pass) {
  const int32_t on_pass = op.nkicks_rf;
  if (!(on_pass == -1 || on_pass == pass))
    return;
  const double dz = op.d[EH_D_MA_DZ];
  const double dt = op.d[EH_D_MA_DT], ddp = op.d[EH_D_MA_DP], de = op.d[EH_D_MA_DE];
  const double Po = op.Po;
  double ds = dz ? dz * em_sqrt(1 + sq(p.xp) + sq(p.yp)) : 0;
  p.x += op.d[EH_D_MA_DX] + dz * p.xp;
  p.xp += op.d[EH_D_MA_DXP];
  p.y += op.d[EH_D_MA_DY] + dz * p.yp;
  p.yp += op.d[EH_D_MA_DYP];
  p.s += ds;
  if (dt || ddp || de) {
    double pc = Po * (1 + p.dp);
    double gamma = em_sqrt(1 + pc * pc);
    double beta = pc / gamma;
    const double t = p.s / (beta * EH_C_MKS) + dt;
    if (ddp) {
      p.dp += ddp;
      pc = Po * (1 + p.dp);
      beta = pc / em_sqrt(1 + pc * pc);
    }
    if (de) {
      gamma += de * gamma;
      pc = em_sqrt(gamma * gamma - 1);
      beta = pc / gamma;
      p.dp = (pc - Po) / Po;
    }
    p.s = t * beta * EH_C_MKS;
  }
}

/* rectangular_collimator (limit_amplitudes.c:32-185), the entrance check:
 * lost outside the half-widths (an open side may spare it) or with a
 * non-finite x or y, INVERT swapping the verdict; the loss at the
 * element's start */
DEV int track_rcol_in(particle &p, const eh_op &op, const double *z) {
  const double xsize = op.d[EH_D_RC_XMAX], ysize = op.d[EH_D_RC_YMAX];
  const double dx = p.x - op.d[EH_D_RC_DX], dy = p.y - op.d[EH_D_RC_DY];
  const int open = op.order1;
  bool lost = false;
  if ((xsize > 0 && __builtin_fabs(dx) > xsize) || (ysize > 0 && __builtin_fabs(dy) > ysize))
    lost = open ? ap_open_lost(open, dx, dy, xsize, ysize) : true;
  else if (__builtin_isinf(p.x) || __builtin_isinf(p.y) || __builtin_isnan(p.x) || __builtin_isnan(p.y))
    lost = true;
  if (op.order0)
    lost = !lost;
  if (lost) {
    p.s = opc(&z[0]);
    p.dp = op.Po * (1 + p.dp);
    return EH_STAGE_ELEMENT + 1;
  }
  return 0;
}

/* the collimator's drift and exit check: a particle outside at the exit
 * is lost where its straight orbit met the jaw (no open side), a
 * non-finite exit position at the exit; survivors take the drift */
DEV int track_rcol_body(particle &p, const eh_op &op, const double *z) {
  const double length = op.d[EH_D_RC_LEN];
  const double xsize = op.d[EH_D_RC_XMAX], ysize = op.d[EH_D_RC_YMAX];
  const double xc = op.d[EH_D_RC_DX], yc = op.d[EH_D_RC_DY];
  const int open = op.order1;
  const double x1 = p.x + length * p.xp, y1 = p.y + length * p.yp;
  const double dx = x1 - xc, dy = y1 - yc;
  const double z0 = opc(&z[0]);
  int is_out = 0;
  if (xsize > 0 && __builtin_fabs(dx) > xsize)
    is_out += 1 * (open ? (ap_open_lost(open, dx, 0, xsize, ysize) ? 1 : 0) : 1);
  if (ysize > 0 && __builtin_fabs(dy) > ysize)
    is_out += 2 * (open ? (ap_open_lost(open, 0, dy, xsize, ysize) ? 1 : 0) : 1);
  if (__builtin_isinf(x1) || __builtin_isinf(y1) || __builtin_isnan(x1) || __builtin_isnan(y1))
    is_out += 4;
  if (is_out & 4) {
    p.s = z0 + length;
    p.x = x1;
    p.y = y1;
    p.dp = op.Po * (1 + p.dp);
    return EH_STAGE_ELEMENT + 1;
  }
  if ((is_out && !op.order0) || (!is_out && op.order0)) {
    if (!open) {
      double zx = 1.7976931348623157e308, zy = 1.7976931348623157e308;
      if ((is_out & 1) && p.xp != 0)
        zx = ((p.xp > 0 ? 1 : -1) * xsize - (p.x - xc)) / p.xp;
      if ((is_out & 2) && p.yp != 0)
        zy = ((p.yp > 0 ? 1 : -1) * ysize - (p.y - yc)) / p.yp;
      const double zz = zx < zy ? zx : zy;
      p.x += p.xp * zz;
      p.y += p.yp * zz;
      p.s = z0 + zz;
    }
    p.dp = op.Po * (1 + p.dp);
    return EH_STAGE_ELEMENT + 1;
  }
  p.s += length * em_sqrt(1 + sq(p.xp) + sq(p.yp));
  p.x = x1;
  p.y = y1;
  return 0;
}

/* One element on one particle.  z points at the element's (z_start, z_end),
 * which only the loss paths read: in the element record itself, or in the
 * position table of a program whose repeated cells share one record. */
template <bool SH>
DEV int run_op(particle &p, const eh_op &op, const double *z, int64_t pass) {
  int lost = 0;
  /* elements before the recirculation point act on pass 0 only */
  if ((op.flags & EH_F_FIRSTPASS) && pass != 0)
    return 0;
  switch (op.op) {
  case EH_OP_DRIFT2:
    if ((op.flags & EH_F_DFAST) && ALL_LANES(drift2_ok(p)))
      drift2(p, op.d[0], op.d[1]);
    else
      drift2_exact(p, op.d[0], op.d[1]);
    break;
  case EH_OP_DRIFT1:
    drift1(p, op.d[0]);
    break;
  case EH_OP_EXACT_DRIFT:
    exact_drift(p, op.d[0]);
    break;
  case EH_OP_MATRIX1:
    matrix1(p, op.d);
    break;
  case EH_OP_MULT:
    lost = track_multipole<SH>(p, op, z);
    break;
  case EH_OP_CSBEND:
    lost = track_csbend<SH>(p, op, z);
    break;
  case EH_OP_RFCA:
    lost = track_rfca(p, op, z);
    break;
  case EH_OP_MALIGN:
    track_malign(p, op, pass);
    break;
  case EH_OP_RCOL_IN:
    lost = track_rcol_in(p, op, z);
    break;
  case EH_OP_RCOL_BODY:
    lost = track_rcol_body(p, op, z);
    break;
  case EH_OP_RFCA_DRIFT:
    /* VOLT=0: a drift when the cavity has length, else nothing (simple_rfca.c:389-402) */
    if (op.d[EH_D_RFLEN_TOTAL] != 0) {
      p.x += p.xp * op.d[EH_D_RFLEN_TOTAL];
      p.y += p.yp * op.d[EH_D_RFLEN_TOTAL];
      p.s += op.d[EH_D_RFLEN_TOTAL] * em_sqrt(1 + sq(p.xp) + sq(p.yp));
    }
    break;
  default:
    break;
  }
  if (lost)
    return lost;
  return post_element_checks<SH>(p, op, z);
}

/* Four waves per SIMD (up to 128 VGPRs): the branch-free bend body needs
 * about 111 registers in this kernel; squeezing it into five or six waves
 * spills to scratch or parks state in LDS, and measured no faster on the
 * storage ring (29.0-29.3 M against 29.4 M particle-turns/s).
 * SH: built with the elliptical / open-sided MAXAMP checks; launched only for
 * programs that have such an aperture, so the rectangle-only instance keeps
 * the registers and schedule of the common case.
 * The beamline of nflat elements arrives as runs (eh_run): a body of
 * records executed reps times in a row, so a ring of identical cells keeps
 * one cell's records hot in the scalar cache instead of streaming every
 * element's record through it each pass.  Loss codes and op_first/op_last
 * count positions in the flat beamline; zpos holds (z_start, z_end) for
 * each of them. */
template <bool SH>
__global__ void __launch_bounds__(256, 4) eh_track_kernel(eh_state st, const eh_op *__restrict__ ops,
                                                       const eh_run *__restrict__ runs, int32_t nruns,
                                                       const double *__restrict__ zpos, int64_t pass0,
                                                       int32_t npasses, int32_t op_first, int32_t op_last,
                                                       int32_t nflat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n)
    return;
  if (st.loss_pass[i] >= 0)
    return;
  particle p;
  p.x = st.x[i];
  p.xp = st.xp[i];
  p.y = st.y[i];
  p.yp = st.yp[i];
  p.s = st.s[i];
  p.dp = st.dp[i];
  int lost = 0, lost_op = 0;
  int64_t lost_pass = 0;
  for (int32_t ip = 0; ip < npasses && !lost; ip++) {
    /* element positions [ob, oe) of the flat beamline in this pass */
    const int32_t ob = (ip == 0) ? op_first : 0;
    const int32_t oe = (ip == npasses - 1) ? op_last : nflat;
    for (int32_t r = 0; r < nruns && !lost; r++) {
      const int32_t fs = runs[r].flat, len = runs[r].len, first = runs[r].first;
      const int32_t fe = fs + len * runs[r].reps;
      if (fe <= ob)
        continue;
      if (fs >= oe)
        break;
      int32_t f = fs > ob ? fs : ob;
      const int32_t f1 = fe < oe ? fe : oe;
      int32_t k = (f - fs) % len;
      for (; f < f1; f++) {
        lost = run_op<SH>(p, ops[first + k], zpos + 2 * f, pass0 + ip);
        if (lost) {
          lost_op = f;
          lost_pass = pass0 + ip;
          break;
        }
        if (++k == len)
          k = 0;
      }
    }
  }
  st.x[i] = p.x;
  st.xp[i] = p.xp;
  st.y[i] = p.y;
  st.yp[i] = p.yp;
  st.s[i] = p.s;
  st.dp[i] = p.dp;
  if (lost) {
    st.loss_pass[i] = (int32_t)lost_pass;
    st.loss_code[i] = 4 * lost_op + (lost - 1);
  }
}

/* Persistent batch kernel: each workgroup pulls chunks (a variant and a run
 * of its particles) from a queue until it is empty; every lane tracks PPL
 * particles of the chunk through the variant's program.  The chunk queue is
 * ordered by variant so that workgroups resident at the same time share a few
 * programs in L2.  SH as for eh_track_kernel. */
template <int PPL, bool SH>
__global__ void __launch_bounds__(256) eh_batch_kernel(eh_state st, const eh_op *__restrict__ ops,
                                                       const eh_vdesc *__restrict__ vars,
                                                       const eh_chunk *__restrict__ chunks,
                                                       const int32_t *__restrict__ nchunks_p,
                                                       const int32_t *__restrict__ perm, int64_t pass0,
                                                       int32_t npasses, unsigned int *queue) {
  __shared__ int s_chunk;
  const int32_t nchunks = *nchunks_p; /* device-resident: a queue rebuilt on the device sets it */
  for (;;) {
    if (threadIdx.x == 0)
      s_chunk = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int c = s_chunk;
    __syncthreads();
    if (c >= nchunks)
      return;
    const eh_chunk ch = chunks[c];
    const eh_vdesc vd = vars[ch.variant];
    const eh_op *vops = ops + vd.op_offset;
    const int32_t nops = vd.nops;
    particle p[PPL];
    int lost[PPL], lost_op[PPL];
    int64_t lost_pass[PPL], idx[PPL];
    bool live[PPL];
    for (int k = 0; k < PPL; k++) {
      const int64_t local = threadIdx.x + (int64_t)k * blockDim.x;
      const bool in = local < ch.count;
      idx[k] = perm ? (in ? (int64_t)perm[ch.first + local] : 0) : ch.first + local;
      live[k] = in && st.loss_pass[idx[k]] < 0;
      lost[k] = 0;
      lost_op[k] = 0;
      lost_pass[k] = 0;
      if (live[k]) {
        p[k].x = st.x[idx[k]];
        p[k].xp = st.xp[idx[k]];
        p[k].y = st.y[idx[k]];
        p[k].yp = st.yp[idx[k]];
        p[k].s = st.s[idx[k]];
        p[k].dp = st.dp[idx[k]];
      }
    }
    for (int32_t ip = 0; ip < npasses; ip++) {
      bool any = false;
      for (int k = 0; k < PPL; k++)
        any |= live[k];
      if (!any)
        break;
      for (int32_t io = 0; io < nops; io++) {
        const eh_op &op = vops[io];
        for (int k = 0; k < PPL; k++) {
          if (!live[k])
            continue;
          int l = run_op<SH>(p[k], op, &op.z_start, pass0 + ip);
          if (l) {
            live[k] = false;
            lost[k] = l;
            lost_op[k] = io;
            lost_pass[k] = pass0 + ip;
          }
        }
      }
    }
    for (int k = 0; k < PPL; k++) {
      const int64_t local = threadIdx.x + (int64_t)k * blockDim.x;
      if (local >= ch.count)
        continue;
      if (live[k] || lost[k]) {
        const int64_t i = idx[k];
        st.x[i] = p[k].x;
        st.xp[i] = p[k].xp;
        st.y[i] = p[k].y;
        st.yp[i] = p[k].yp;
        st.s[i] = p[k].s;
        st.dp[i] = p[k].dp;
        if (lost[k]) {
          st.loss_pass[i] = (int32_t)lost_pass[k];
          st.loss_code[i] = 4 * lost_op[k] + (lost[k] - 1);
        }
      }
    }
  }
}

__global__ void eh_queue_reset(unsigned int *queue) { *queue = 0u; }

/* ------------------------------------------------------------ host side */

extern "C" {

static int hip_fail(ehd_dev *d, hipError_t e, const char *what) {
  if (d)
    snprintf(d->msg, sizeof(d->msg), "%s: %s", what, hipGetErrorString(e));
  return EH_ERR_HIP;
}

#define HIPCHK(d, call)                                                                                                \
  do {                                                                                                                 \
    hipError_t e_ = (call);                                                                                            \
    if (e_ != hipSuccess)                                                                                              \
      return hip_fail((d), e_, #call);                                                                                 \
  } while (0)

int ehd_create(int device, ehd_dev **out, char *msg, int msglen) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    if (msg && msglen > 0)
      snprintf(msg, msglen, "no HIP device available (%s)", hipGetErrorString(e));
    return EH_ERR_NODEVICE;
  }
  if (device < 0 || device >= count) {
    if (msg && msglen > 0)
      snprintf(msg, msglen, "device %d out of range (%d devices)", device, count);
    return EH_ERR_INVALID;
  }
  ehd_dev *d = (ehd_dev *)calloc(1, sizeof(ehd_dev));
  if (!d)
    return EH_ERR_NOMEM;
  d->device = device;
  d->block = 0; /* chosen per launch from the particle count */
  e = hipSetDevice(device);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreate(&d->ev0);
  if (e == hipSuccess)
    e = hipEventCreate(&d->ev1);
  if (e == hipSuccess)
    e = hipEventCreate(&d->cev0);
  if (e == hipSuccess)
    e = hipEventCreate(&d->cev1);
  if (e != hipSuccess) {
    if (msg && msglen > 0)
      snprintf(msg, msglen, "HIP set-up failed: %s", hipGetErrorString(e));
    free(d);
    return EH_ERR_HIP;
  }
  int rc = ehd_noalloc_init(d);
  if (rc) {
    if (msg && msglen > 0)
      snprintf(msg, msglen, "%s", d->msg);
    ehd_destroy(d);
    return rc;
  }
  *out = d;
  return EH_OK;
}

void ehd_free_collective(ehd_dev *d);

/* teardown: errors are ignored, there is nothing left to undo */
void ehd_destroy(ehd_dev *d) {
  if (!d)
    return;
  (void)hipSetDevice(d->device);
  (void)hipStreamSynchronize(d->stream);
  void *bufs[] = {d->coord[0], d->coord[1], d->coord[2], d->coord[3], d->coord[4], d->coord[5],
                  d->loss_pass, d->loss_code, d->ops, d->counter, d->vdesc, d->chunks, d->perm, d->queue,
                  d->snap[0], d->snap[1], d->snap[2], d->snap[3], d->snap[4], d->snap[5], d->snap_lp, d->snap_lc,
                  d->tile_sum, d->tops, d->runs, d->zpos};
  for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++)
    if (bufs[k])
      (void)hipFree(bufs[k]);
  ehd_free_collective(d);
  free(d->host_prog);
  (void)hipEventDestroy(d->ev0);
  (void)hipEventDestroy(d->ev1);
  (void)hipEventDestroy(d->cev0);
  (void)hipEventDestroy(d->cev1);
  (void)hipStreamDestroy(d->stream);
  free(d);
}

int ehd_alloc_state(ehd_dev *d, int64_t n) {
  HIPCHK(d, hipSetDevice(d->device));
  if (n > d->cap) {
    for (int k = 0; k < 6; k++) {
      if (d->coord[k])
        HIPCHK(d, hipFree(d->coord[k]));
      d->coord[k] = NULL;
    }
    if (d->loss_pass)
      HIPCHK(d, hipFree(d->loss_pass));
    if (d->loss_code)
      HIPCHK(d, hipFree(d->loss_code));
    d->loss_pass = d->loss_code = NULL;
    d->cap = 0;
    for (int k = 0; k < 6; k++)
      HIPCHK(d, hipMalloc((void **)&d->coord[k], sizeof(double) * (n > 0 ? n : 1)));
    HIPCHK(d, hipMalloc((void **)&d->loss_pass, sizeof(int32_t) * (n > 0 ? n : 1)));
    HIPCHK(d, hipMalloc((void **)&d->loss_code, sizeof(int32_t) * (n > 0 ? n : 1)));
    d->cap = n;
  }
  if (!d->counter)
    HIPCHK(d, hipMalloc((void **)&d->counter, sizeof(unsigned long long) * 8));
  d->n = n;
  return EH_OK;
}

int64_t ehd_state_n(const ehd_dev *d) { return d->n; }

int ehd_put_state(ehd_dev *d, const double *x, const double *xp, const double *y, const double *yp, const double *s,
                  const double *dp, const int32_t *loss_pass, const int32_t *loss_code, int64_t n) {
  const double *src[6] = {x, xp, y, yp, s, dp};
  if (n != d->n)
    return EH_ERR_STATE;
  HIPCHK(d, hipSetDevice(d->device));
  for (int k = 0; k < 6; k++)
    HIPCHK(d, hipMemcpyAsync(d->coord[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->loss_pass, loss_pass, sizeof(int32_t) * n, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->loss_code, loss_code, sizeof(int32_t) * n, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

int ehd_get_state(ehd_dev *d, double *x, double *xp, double *y, double *yp, double *s, double *dp,
                  int32_t *loss_pass, int32_t *loss_code, int64_t n) {
  double *dst[6] = {x, xp, y, yp, s, dp};
  if (n != d->n)
    return EH_ERR_STATE;
  HIPCHK(d, hipSetDevice(d->device));
  for (int k = 0; k < 6; k++)
    if (dst[k])
      HIPCHK(d, hipMemcpyAsync(dst[k], d->coord[k], sizeof(double) * n, hipMemcpyDeviceToHost, d->stream));
  if (loss_pass)
    HIPCHK(d, hipMemcpyAsync(loss_pass, d->loss_pass, sizeof(int32_t) * n, hipMemcpyDeviceToHost, d->stream));
  if (loss_code)
    HIPCHK(d, hipMemcpyAsync(loss_code, d->loss_code, sizeof(int32_t) * n, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

/* s, delta and loss pass of particle slot i alone (any may be NULL) */
int ehd_get_one(ehd_dev *d, int64_t i, double *s, double *dp, int32_t *loss_pass) {
  if (i < 0 || i >= d->n)
    return EH_ERR_INVALID;
  HIPCHK(d, hipSetDevice(d->device));
  if (s)
    HIPCHK(d, hipMemcpyAsync(s, d->coord[4] + i, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (dp)
    HIPCHK(d, hipMemcpyAsync(dp, d->coord[5] + i, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (loss_pass)
    HIPCHK(d, hipMemcpyAsync(loss_pass, d->loss_pass + i, sizeof(int32_t), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

int ehd_get_sdp(ehd_dev *d, double *s, double *dp, int32_t *loss_pass, int64_t n) {
  return ehd_get_state(d, NULL, NULL, NULL, NULL, s, dp, loss_pass, NULL, n);
}

static int tops_shaped(const eh_op *op) { return op->ap_code || op->in_code; }

int ehd_put_ops(ehd_dev *d, const eh_op *ops, int32_t nops) {
  HIPCHK(d, hipSetDevice(d->device));
  if (nops > d->ops_cap) {
    if (d->ops)
      HIPCHK(d, hipFree(d->ops));
    d->ops = NULL;
    HIPCHK(d, hipMalloc((void **)&d->ops, sizeof(eh_op) * (nops > 0 ? nops : 1)));
    d->ops_cap = nops;
  }
  if (nops > 0)
    HIPCHK(d, hipMemcpyAsync(d->ops, ops, sizeof(eh_op) * nops, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  d->nops = nops;
  /* programs with an elliptical or open-sided MAXAMP run on the kernel
   * instance that evaluates those shapes */
  d->shaped = 0;
  for (int32_t k = 0; k < nops; k++)
    if (tops_shaped(&ops[k]))
      d->shaped = 1;
  return EH_OK;
}

/* grows a device buffer to hold n elements of size sz (contents dropped) */
static int grow(ehd_dev *d, void **buf, int32_t *cap, int32_t n, size_t sz) {
  if (n <= *cap)
    return EH_OK;
  if (*buf)
    HIPCHK(d, hipFree(*buf));
  *buf = NULL;
  *cap = 0;
  HIPCHK(d, hipMalloc(buf, sz * (size_t)(n > 0 ? n : 1)));
  *cap = n;
  return EH_OK;
}

/* folds the host copy into runs and uploads bodies, runs and positions */
static int upload_program(ehd_dev *d) {
  const int32_t n = d->nflat;
  const size_t m = (size_t)(n > 0 ? n : 1);
  eh_run *runs = (eh_run *)malloc(sizeof(eh_run) * m);
  int32_t *body = (int32_t *)malloc(sizeof(int32_t) * m);
  eh_op *tops = (eh_op *)malloc(sizeof(eh_op) * m);
  double *zpos = (double *)malloc(sizeof(double) * 2 * m);
  int32_t nb = 0, nr = -1;
  int rc = EH_OK;
  if (runs && body && tops && zpos)
    nr = eh_fold_runs(d->host_prog, n, EH_RUN_MAX_LEN, runs, body, &nb);
  if (nr < 0) {
    rc = EH_ERR_NOMEM;
    snprintf(d->msg, sizeof d->msg, "out of host memory folding the program");
  }
  if (!rc) {
    for (int32_t j = 0; j < nb; j++)
      tops[j] = d->host_prog[body[j]];
    for (int32_t f = 0; f < n; f++) {
      zpos[2 * f] = d->host_prog[f].z_start;
      zpos[2 * f + 1] = d->host_prog[f].z_end;
    }
    rc = grow(d, (void **)&d->tops, &d->tops_cap, nb, sizeof(eh_op));
  }
  if (!rc)
    rc = grow(d, (void **)&d->runs, &d->runs_cap, nr, sizeof(eh_run));
  if (!rc)
    rc = grow(d, (void **)&d->zpos, &d->zpos_cap, 2 * n, sizeof(double));
  hipError_t e = hipSuccess;
  if (!rc && nb > 0)
    e = hipMemcpyAsync(d->tops, tops, sizeof(eh_op) * nb, hipMemcpyHostToDevice, d->stream);
  if (!rc && !e && nr > 0)
    e = hipMemcpyAsync(d->runs, runs, sizeof(eh_run) * nr, hipMemcpyHostToDevice, d->stream);
  if (!rc && !e && n > 0)
    e = hipMemcpyAsync(d->zpos, zpos, sizeof(double) * 2 * n, hipMemcpyHostToDevice, d->stream);
  /* the host buffers are the copies' sources: wait before freeing them */
  hipError_t es = hipStreamSynchronize(d->stream);
  if (!rc && (e || es))
    rc = hip_fail(d, e ? e : es, "uploading the tracking program");
  free(runs);
  free(body);
  free(tops);
  free(zpos);
  if (rc) {
    d->nruns = d->ntops = 0;
    d->nflat = 0;
    return rc;
  }
  d->nruns = nr;
  d->ntops = nb;
  d->tshaped = 0;
  for (int32_t f = 0; f < n; f++)
    if (tops_shaped(&d->host_prog[f]))
      d->tshaped = 1;
  return EH_OK;
}

int ehd_put_program(ehd_dev *d, const eh_op *ops, int32_t nops) {
  if (nops < 0 || (nops > 0 && !ops))
    return EH_ERR_INVALID;
  HIPCHK(d, hipSetDevice(d->device));
  if (nops > d->host_prog_cap) {
    eh_op *h = (eh_op *)realloc(d->host_prog, sizeof(eh_op) * (size_t)nops);
    if (!h)
      return EH_ERR_NOMEM;
    d->host_prog = h;
    d->host_prog_cap = nops;
  }
  if (nops > 0)
    memcpy(d->host_prog, ops, sizeof(eh_op) * (size_t)nops);
  d->nflat = nops;
  return upload_program(d);
}

int ehd_patch_op(ehd_dev *d, int32_t index, const eh_op *op) {
  if (index < 0 || index >= d->nflat)
    return EH_ERR_INVALID;
  HIPCHK(d, hipSetDevice(d->device));
  d->host_prog[index] = *op;
  return upload_program(d);
}

int32_t ehd_program_runs(const ehd_dev *d) { return d->nruns; }

/* 64, 128, 192 or 256 lanes per workgroup; 0 = chosen per launch */
void ehd_set_block(ehd_dev *d, int32_t block) {
  if (block == 0 || (block >= 64 && block <= 256 && block % 64 == 0))
    d->block = block;
}

int ehd_track(ehd_dev *d, int64_t pass0, int32_t npasses, int32_t op_first, int32_t op_last) {
  if (npasses <= 0 || d->n == 0)
    return EH_OK;
  if (op_first < 0 || op_last > d->nflat || op_first > d->nflat || (npasses == 1 && op_first > op_last))
    return EH_ERR_INVALID;
  HIPCHK(d, hipSetDevice(d->device));
  eh_state st;
  st.x = d->coord[0];
  st.xp = d->coord[1];
  st.y = d->coord[2];
  st.yp = d->coord[3];
  st.s = d->coord[4];
  st.dp = d->coord[5];
  st.loss_pass = d->loss_pass;
  st.loss_code = d->loss_code;
  st.n = d->n;
  int block = d->block;
  if (block == 0) {
    /* 256 lanes per workgroup, unless the particles would fill at most one
     * such workgroup per CU: then 64, so the waves spread over all CUs
     * instead of sharing a few (10^4 particles: 157 CUs busy instead of 40,
     * 6 % faster on the FODO ring) */
    if (!d->num_cu) {
      int cu = 0;
      HIPCHK(d, hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d->device));
      d->num_cu = cu > 0 ? cu : 256;
    }
    block = d->n <= (int64_t)d->num_cu * 256 ? 64 : 256;
  }
  const int64_t grid = (d->n + block - 1) / block;
  HIPCHK(d, hipEventRecord(d->ev0, d->stream));
  if (d->tshaped)
    hipLaunchKernelGGL(eh_track_kernel<true>, dim3((unsigned)grid), dim3(block), 0, d->stream, st, d->tops, d->runs,
                       d->nruns, d->zpos, pass0, npasses, op_first, op_last, d->nflat);
  else
    hipLaunchKernelGGL(eh_track_kernel<false>, dim3((unsigned)grid), dim3(block), 0, d->stream, st, d->tops, d->runs,
                       d->nruns, d->zpos, pass0, npasses, op_first, op_last, d->nflat);
  HIPCHK(d, hipGetLastError());
  HIPCHK(d, hipEventRecord(d->ev1, d->stream));
  HIPCHK(d, hipEventSynchronize(d->ev1));
  float ms = 0;
  HIPCHK(d, hipEventElapsedTime(&ms, d->ev0, d->ev1));
  d->track_ms += ms;
  d->track_launches += 1;
  return EH_OK;
}

/* do_match_energy's re-expression of delta against the new central
 * momentum, delta = dp + delta dr, on the particles still in the beam
 * (do_tracking.c:3408-3414) */
__global__ void eh_rescale_dp_kernel(double *dp_, const int32_t *loss_pass, int64_t n, double dp, double dr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && loss_pass[i] < 0)
    dp_[i] = dp + dp_[i] * dr;
}

int ehd_rescale_dp(ehd_dev *d, double dp, double dr) {
  if (d->n == 0)
    return EH_OK;
  HIPCHK(d, hipSetDevice(d->device));
  const int64_t grid = (d->n + 255) / 256;
  hipLaunchKernelGGL(eh_rescale_dp_kernel, dim3((unsigned)grid), dim3(256), 0, d->stream, d->coord[5], d->loss_pass,
                     d->n, dp, dr);
  HIPCHK(d, hipGetLastError());
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

__global__ void eh_count_alive_kernel(const int32_t *loss_pass, int64_t n, unsigned long long *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int alive = (i < n && loss_pass[i] < 0) ? 1 : 0;
  unsigned long long b = __ballot(alive);
  if ((threadIdx.x & 63) == 0 && b)
    atomicAdd(out, (unsigned long long)__popcll(b));
}

int ehd_count_alive(ehd_dev *d, int64_t *n_alive) {
  HIPCHK(d, hipSetDevice(d->device));
  unsigned long long h = 0;
  HIPCHK(d, hipMemcpyAsync(d->counter, &h, sizeof(h), hipMemcpyHostToDevice, d->stream));
  if (d->n > 0) {
    int64_t grid = (d->n + 255) / 256;
    hipLaunchKernelGGL(eh_count_alive_kernel, dim3((unsigned)grid), dim3(256), 0, d->stream, d->loss_pass, d->n,
                       d->counter);
    HIPCHK(d, hipGetLastError());
  }
  HIPCHK(d, hipMemcpyAsync(&h, d->counter, sizeof(h), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  *n_alive = (int64_t)h;
  return EH_OK;
}

int ehd_sync(ehd_dev *d) {
  HIPCHK(d, hipSetDevice(d->device));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

void ehd_timing_reset(ehd_dev *d) {
  d->track_ms = 0;
  d->track_launches = 0;
  d->coll_ms = 0;
  d->coll_launches = 0;
}

int ehd_timing(ehd_dev *d, int64_t *launches, double *ms) {
  *launches = d->track_launches;
  *ms = d->track_ms;
  return EH_OK;
}

const char *ehd_message(const ehd_dev *d) { return d ? d->msg : ""; }

int ehd_batch_setup(ehd_dev *d, const eh_vdesc *v, int32_t nvar, const eh_chunk *chunks, int32_t nchunks) {
  HIPCHK(d, hipSetDevice(d->device));
  if (nvar > d->var_cap) {
    if (d->vdesc)
      HIPCHK(d, hipFree(d->vdesc));
    d->vdesc = NULL;
    HIPCHK(d, hipMalloc((void **)&d->vdesc, sizeof(eh_vdesc) * nvar));
    d->var_cap = nvar;
  }
  if (nchunks > d->chunk_cap) {
    if (d->chunks)
      HIPCHK(d, hipFree(d->chunks));
    d->chunks = NULL;
    HIPCHK(d, hipMalloc((void **)&d->chunks, sizeof(eh_chunk) * nchunks));
    d->chunk_cap = nchunks;
  }
  if (!d->queue)
    HIPCHK(d, hipMalloc((void **)&d->queue, 256));
  if (!d->num_cu) {
    int cu = 0;
    HIPCHK(d, hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d->device));
    d->num_cu = cu > 0 ? cu : 256;
  }
  d->dev_nchunks = (int32_t *)d->queue + 32;
  HIPCHK(d, hipMemcpyAsync(d->vdesc, v, sizeof(eh_vdesc) * nvar, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->chunks, chunks, sizeof(eh_chunk) * nchunks, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->dev_nchunks, &nchunks, sizeof(int32_t), hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  d->nvar = nvar;
  d->nchunks = nchunks;
  d->use_perm = 0;
  return EH_OK;
}

int ehd_batch_compact(ehd_dev *d, const int32_t *perm, int64_t nperm, const eh_chunk *chunks, int32_t nchunks) {
  HIPCHK(d, hipSetDevice(d->device));
  if (nperm > d->perm_cap) {
    if (d->perm)
      HIPCHK(d, hipFree(d->perm));
    d->perm = NULL;
    d->perm_cap = 0;
    HIPCHK(d, hipMalloc((void **)&d->perm, sizeof(int32_t) * (nperm > 0 ? nperm : 1)));
    d->perm_cap = nperm > 0 ? nperm : 1;
  }
  if (nchunks > d->chunk_cap)
    return EH_ERR_INVALID; /* compaction never adds chunks */
  if (nperm > 0)
    HIPCHK(d, hipMemcpyAsync(d->perm, perm, sizeof(int32_t) * nperm, hipMemcpyHostToDevice, d->stream));
  if (nchunks > 0)
    HIPCHK(d, hipMemcpyAsync(d->chunks, chunks, sizeof(eh_chunk) * nchunks, hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->dev_nchunks, &nchunks, sizeof(int32_t), hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  d->nchunks = nchunks;
  d->use_perm = 1;
  return EH_OK;
}

/* ---- the work queue rebuilt on the device ----
 * The surviving particles, in index order (so grouped by variant, whose
 * particle ranges are contiguous), are listed in perm, and each variant's
 * list is cut into chunks of csize; the count lands in dev_nchunks.  Six
 * small kernels, no host round trip and no memset, so the sequence can sit
 * between the tracking launches of one captured graph. */
#define CT 1024 /* particles per tile of the compaction */

/* alive particles per tile */
__global__ void __launch_bounds__(256) compact_count(const int32_t *__restrict__ lp, int64_t n, int32_t *tile_sum) {
  __shared__ int32_t s[4];
  const int64_t base = (int64_t)blockIdx.x * CT;
  int32_t c = 0;
  for (int k = threadIdx.x; k < CT; k += 256) {
    const int64_t i = base + k;
    c += (i < n && lp[i] < 0) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1)
    c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0)
    s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    tile_sum[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

/* exclusive scan of the tile sums in place (one workgroup); tile_sum[nt] = total */
__global__ void __launch_bounds__(256) compact_scan(int32_t *tile_sum, int32_t nt) {
  __shared__ int32_t s[256];
  __shared__ int32_t carry;
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  for (int32_t base = 0; base < nt; base += 256) {
    const int32_t i = base + threadIdx.x;
    const int32_t v = i < nt ? tile_sum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int32_t t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    const int32_t incl = s[threadIdx.x];
    const int32_t c0 = carry;
    if (i < nt)
      tile_sum[i] = c0 + incl - v;
    __syncthreads();
    if (threadIdx.x == 255)
      carry = c0 + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0)
    tile_sum[nt] = carry;
}

/* perm[position] = index of each alive particle, in index order */
__global__ void __launch_bounds__(256) compact_scatter(const int32_t *__restrict__ lp, int64_t n,
                                                       const int32_t *__restrict__ tile_sum, int32_t *perm) {
  __shared__ int32_t wsum[4];
  const int64_t base = (int64_t)blockIdx.x * CT;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t off = tile_sum[blockIdx.x];
  for (int k0 = 0; k0 < CT; k0 += 256) {
    const int64_t i = base + k0 + threadIdx.x;
    const bool a = i < n && lp[i] < 0;
    const unsigned long long m = __ballot(a);
    const int before = __popcll(m & ((1ULL << lane) - 1ULL));
    if (lane == 0)
      wsum[w] = __popcll(m);
    __syncthreads();
    int32_t wbase = 0;
    for (int q = 0; q < w; q++)
      wbase += wsum[q];
    const int32_t tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (a)
      perm[off + wbase + before] = (int32_t)i;
    __syncthreads();
    off += tot;
  }
}

/* alive particles with index below k, by the workgroup: the tile's prefix
 * plus a count over the start of the tile */
DEV int32_t alive_below(const int32_t *__restrict__ lp, int64_t n, const int32_t *__restrict__ tile_sum, int64_t k,
                        int32_t *red) {
  const int64_t tile = k / CT, base = tile * CT;
  int32_t c = 0;
  for (int64_t i = base + threadIdx.x; i < k && i < n; i += 256)
    c += lp[i] < 0 ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1)
    c += __shfl_xor(c, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = c;
  __syncthreads();
  return tile_sum[tile] + red[0] + red[1] + red[2] + red[3];
}

/* per variant (one workgroup each): vinfo[4v..] = start in perm, alive, chunks */
__global__ void __launch_bounds__(256) compact_variants(const int32_t *__restrict__ lp, int64_t n,
                                                        const int32_t *__restrict__ tile_sum,
                                                        const eh_vdesc *__restrict__ vars, int32_t csize,
                                                        int32_t *vinfo) {
  __shared__ int32_t red[4];
  const eh_vdesc vd = vars[blockIdx.x];
  const int32_t p0 = alive_below(lp, n, tile_sum, vd.first, red);
  const int32_t p1 = alive_below(lp, n, tile_sum, vd.first + vd.n, red);
  if (threadIdx.x == 0) {
    vinfo[4 * blockIdx.x] = p0;
    vinfo[4 * blockIdx.x + 1] = p1 - p0;
    vinfo[4 * blockIdx.x + 2] = (p1 - p0 + csize - 1) / csize;
  }
}

/* first chunk of each variant (exclusive scan, one workgroup) and the count */
__global__ void __launch_bounds__(256) compact_chunk_offsets(int32_t *vinfo, int32_t nvar, int32_t *nchunks_out) {
  __shared__ int32_t s[256];
  __shared__ int32_t carry;
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  for (int32_t base = 0; base < nvar; base += 256) {
    const int32_t v = base + threadIdx.x;
    const int32_t c = v < nvar ? vinfo[4 * v + 2] : 0;
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int32_t t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    const int32_t incl = s[threadIdx.x];
    const int32_t c0 = carry;
    if (v < nvar)
      vinfo[4 * v + 3] = c0 + incl - c;
    __syncthreads();
    if (threadIdx.x == 255)
      carry = c0 + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *nchunks_out = carry;
}

/* the chunks of each variant (one workgroup each) */
__global__ void __launch_bounds__(256) compact_write_chunks(const int32_t *__restrict__ vinfo, int32_t csize,
                                                            eh_chunk *chunks) {
  const int32_t v = blockIdx.x;
  const int32_t start = vinfo[4 * v], alive = vinfo[4 * v + 1], nch = vinfo[4 * v + 2], c0 = vinfo[4 * v + 3];
  for (int32_t c = threadIdx.x; c < nch; c += 256) {
    eh_chunk ch;
    ch.variant = v;
    ch.first = (int64_t)start + (int64_t)c * csize;
    ch.count = alive - c * csize < csize ? alive - c * csize : csize;
    chunks[c0 + c] = ch;
  }
}

static int compact_ensure(ehd_dev *d) {
  const int64_t nt = (d->n + CT - 1) / CT;
  if (d->perm_cap < d->n) {
    if (d->perm)
      HIPCHK(d, hipFree(d->perm));
    d->perm = NULL;
    d->perm_cap = 0;
    HIPCHK(d, hipMalloc((void **)&d->perm, sizeof(int32_t) * (d->n > 0 ? d->n : 1)));
    d->perm_cap = d->n > 0 ? d->n : 1;
  }
  if (d->tile_cap < nt + 1 || d->vinfo_cap < d->nvar) {
    if (d->tile_sum)
      HIPCHK(d, hipFree(d->tile_sum));
    d->tile_sum = NULL;
    d->tile_cap = d->vinfo_cap = 0;
    HIPCHK(d, hipMalloc((void **)&d->tile_sum, sizeof(int32_t) * (nt + 1 + 4 * (int64_t)d->nvar)));
    d->tile_cap = nt + 1;
    d->vinfo_cap = d->nvar;
  }
  return EH_OK;
}

static void compact_launch(ehd_dev *d, int32_t csize) {
  const int64_t nt = (d->n + CT - 1) / CT;
  int32_t *vinfo = d->tile_sum + d->tile_cap;
  hipLaunchKernelGGL(compact_count, dim3((unsigned)nt), dim3(256), 0, d->stream, d->loss_pass, d->n, d->tile_sum);
  hipLaunchKernelGGL(compact_scan, dim3(1), dim3(256), 0, d->stream, d->tile_sum, (int32_t)nt);
  hipLaunchKernelGGL(compact_scatter, dim3((unsigned)nt), dim3(256), 0, d->stream, d->loss_pass, d->n, d->tile_sum,
                     d->perm);
  hipLaunchKernelGGL(compact_variants, dim3((unsigned)d->nvar), dim3(256), 0, d->stream, d->loss_pass, d->n,
                     d->tile_sum, d->vdesc, csize, vinfo);
  hipLaunchKernelGGL(compact_chunk_offsets, dim3(1), dim3(256), 0, d->stream, vinfo, d->nvar, d->dev_nchunks);
  hipLaunchKernelGGL(compact_write_chunks, dim3((unsigned)d->nvar), dim3(256), 0, d->stream, vinfo, csize, d->chunks);
}

static int batch_launch(ehd_dev *d, eh_state st, int64_t pass0, int32_t npasses, int32_t ppl, int32_t block,
                        int grid) {
  hipLaunchKernelGGL(eh_queue_reset, dim3(1), dim3(1), 0, d->stream, d->queue);
  const int32_t *perm = d->use_perm ? d->perm : (const int32_t *)NULL;
  if (ppl == 2 && d->shaped)
    hipLaunchKernelGGL((eh_batch_kernel<2, true>), dim3(grid), dim3(block), 0, d->stream, st, d->ops, d->vdesc,
                       d->chunks, d->dev_nchunks, perm, pass0, npasses, d->queue);
  else if (ppl == 2)
    hipLaunchKernelGGL((eh_batch_kernel<2, false>), dim3(grid), dim3(block), 0, d->stream, st, d->ops, d->vdesc,
                       d->chunks, d->dev_nchunks, perm, pass0, npasses, d->queue);
  else if (d->shaped)
    hipLaunchKernelGGL((eh_batch_kernel<1, true>), dim3(grid), dim3(block), 0, d->stream, st, d->ops, d->vdesc,
                       d->chunks, d->dev_nchunks, perm, pass0, npasses, d->queue);
  else
    hipLaunchKernelGGL((eh_batch_kernel<1, false>), dim3(grid), dim3(block), 0, d->stream, st, d->ops, d->vdesc,
                       d->chunks, d->dev_nchunks, perm, pass0, npasses, d->queue);
  return EH_OK;
}

int ehd_batch_track(ehd_dev *d, int64_t pass0, int32_t npasses, int32_t passes_per_launch, int32_t ppl,
                    int32_t block, int32_t use_graph, int32_t device_compact, double *ms_out, int64_t *launches_out) {
  if (npasses <= 0 || d->nchunks <= 0) {
    *ms_out = 0;
    *launches_out = 0;
    return EH_OK;
  }
  if ((ppl != 1 && ppl != 2) || block < 64 || block > 256 || block % 64 || passes_per_launch < 1)
    return EH_ERR_INVALID;
  HIPCHK(d, hipSetDevice(d->device));
  if (device_compact) {
    int rc = compact_ensure(d);
    if (rc)
      return rc;
  }
  eh_state st;
  st.x = d->coord[0];
  st.xp = d->coord[1];
  st.y = d->coord[2];
  st.yp = d->coord[3];
  st.s = d->coord[4];
  st.dp = d->coord[5];
  st.loss_pass = d->loss_pass;
  st.loss_code = d->loss_code;
  st.n = d->n;
  int per_cu = 0;
  if (ppl == 2 && d->shaped)
    HIPCHK(d, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, eh_batch_kernel<2, true>, block, 0));
  else if (ppl == 2)
    HIPCHK(d, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, eh_batch_kernel<2, false>, block, 0));
  else if (d->shaped)
    HIPCHK(d, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, eh_batch_kernel<1, true>, block, 0));
  else
    HIPCHK(d, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, eh_batch_kernel<1, false>, block, 0));
  if (per_cu < 1)
    per_cu = 1;
  int grid = d->num_cu * per_cu;
  if (grid > d->nchunks)
    grid = d->nchunks;
  int64_t nl = 0;
  hipGraph_t graph = NULL;
  hipGraphExec_t exec = NULL;
  if (use_graph)
    HIPCHK(d, hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
  else
    HIPCHK(d, hipEventRecord(d->ev0, d->stream));
  if (!device_compact) {
    for (int32_t done = 0; done < npasses; done += passes_per_launch) {
      int32_t k = npasses - done < passes_per_launch ? npasses - done : passes_per_launch;
      batch_launch(d, st, pass0 + done, k, ppl, block, grid);
      nl++;
    }
  } else {
    /* launches of 1, 2, 4 ... passes up to passes_per_launch, the queue
     * rebuilt from the survivors on the device after each */
    int32_t done = 0, step = 1;
    while (done < npasses) {
      int32_t k = step < passes_per_launch ? step : passes_per_launch;
      if (k > npasses - done)
        k = npasses - done;
      batch_launch(d, st, pass0 + done, k, ppl, block, grid);
      nl++;
      done += k;
      if (step < passes_per_launch)
        step *= 2;
      if (done < npasses) {
        compact_launch(d, block * ppl);
        d->use_perm = 1;
      }
    }
  }
  if (use_graph) {
    HIPCHK(d, hipStreamEndCapture(d->stream, &graph));
    HIPCHK(d, hipGraphInstantiate(&exec, graph, NULL, NULL, 0));
    HIPCHK(d, hipEventRecord(d->ev0, d->stream));
    HIPCHK(d, hipGraphLaunch(exec, d->stream));
  }
  HIPCHK(d, hipGetLastError());
  HIPCHK(d, hipEventRecord(d->ev1, d->stream));
  HIPCHK(d, hipEventSynchronize(d->ev1));
  float ms = 0;
  HIPCHK(d, hipEventElapsedTime(&ms, d->ev0, d->ev1));
  if (exec)
    HIPCHK(d, hipGraphExecDestroy(exec));
  if (graph)
    HIPCHK(d, hipGraphDestroy(graph));
  d->track_ms += ms;
  d->track_launches += nl;
  *ms_out = ms;
  *launches_out = nl;
  return EH_OK;
}

/* copy of the particle state on the device (autotuning trials restore it) */
int ehd_state_snapshot(ehd_dev *d, int restore) {
  HIPCHK(d, hipSetDevice(d->device));
  if (!restore && d->n > d->snap_cap) {
    for (int k = 0; k < 6; k++) {
      if (d->snap[k])
        HIPCHK(d, hipFree(d->snap[k]));
      d->snap[k] = NULL;
    }
    if (d->snap_lp)
      HIPCHK(d, hipFree(d->snap_lp));
    if (d->snap_lc)
      HIPCHK(d, hipFree(d->snap_lc));
    d->snap_lp = d->snap_lc = NULL;
    for (int k = 0; k < 6; k++)
      HIPCHK(d, hipMalloc((void **)&d->snap[k], sizeof(double) * d->n));
    HIPCHK(d, hipMalloc((void **)&d->snap_lp, sizeof(int32_t) * d->n));
    HIPCHK(d, hipMalloc((void **)&d->snap_lc, sizeof(int32_t) * d->n));
    d->snap_cap = d->n;
  }
  if (restore && d->snap_cap < d->n)
    return EH_ERR_STATE;
  for (int k = 0; k < 6; k++)
    HIPCHK(d, hipMemcpyAsync(restore ? d->coord[k] : d->snap[k], restore ? d->snap[k] : d->coord[k],
                             sizeof(double) * d->n, hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(restore ? d->loss_pass : d->snap_lp, restore ? d->snap_lp : d->loss_pass,
                           sizeof(int32_t) * d->n, hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(restore ? d->loss_code : d->snap_lc, restore ? d->snap_lc : d->loss_code,
                           sizeof(int32_t) * d->n, hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return EH_OK;
}

/* ---- device buffers for tests of the device-pointer entry points ---- */
int ehd_test_buffer(ehd_dev *d, int32_t op, void **ptr, void *host, int64_t bytes) {
  HIPCHK(d, hipSetDevice(d->device));
  switch (op) {
  case 0: /* allocate */
    HIPCHK(d, hipMalloc(ptr, bytes > 0 ? bytes : 8));
    return EH_OK;
  case 1: /* free */
    HIPCHK(d, hipFree(*ptr));
    *ptr = NULL;
    return EH_OK;
  case 2: /* host -> device */
    HIPCHK(d, hipMemcpy(*ptr, host, bytes, hipMemcpyHostToDevice));
    return EH_OK;
  case 3: /* device -> host */
    HIPCHK(d, hipMemcpy(host, *ptr, bytes, hipMemcpyDeviceToHost));
    return EH_OK;
  }
  return EH_ERR_INVALID;
}

/* ---- deterministic-math probe (used by tests to compare with the CPU) ---- */
__global__ void eh_math_probe_kernel(int32_t fn, const double *in, const double *in2, double *out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double v = in[i], r;
  switch (fn) {
  case 0:
    r = em_sin(v);
    break;
  case 1:
    r = em_cos(v);
    break;
  case 2:
    r = em_asin(v);
    break;
  case 3:
    r = em_log(v);
    break;
  case 4:
    r = em_tan(v);
    break;
  case 5:
    r = __builtin_sqrt(v);
    break;
  case 7: /* the unscaled square root the tracking kernels use */
    r = em_sqrt(v);
    break;
  case 8: /* the shared-reciprocal quotient the tracking kernels use */
    r = em_div(v, in2[i], em_recip(in2[i]), em_recip_ok(in2[i]));
    break;
  case 9: /* the same quotient from the rounded reciprocal (bend radii) */
    r = em_div(v, in2[i], 1.0 / in2[i], em_recip_ok(in2[i]));
    break;
  case 10: /* the aperture's integer power, exponent in in2 (0..EH_MAX_AP_EXPONENT) */
    r = ap_ipow(v, (int)in2[i]);
    break;
  default:
    r = v / in2[i];
    break;
  }
  out[i] = r;
}

int ehd_math_probe(ehd_dev *d, int32_t fn, const double *in, const double *in2, double *out, int64_t n) {
  double *din = NULL, *din2 = NULL, *dout = NULL;
  if (n <= 0)
    return EH_OK;
  HIPCHK(d, hipSetDevice(d->device));
  HIPCHK(d, hipMalloc((void **)&din, sizeof(double) * n));
  HIPCHK(d, hipMalloc((void **)&din2, sizeof(double) * n));
  HIPCHK(d, hipMalloc((void **)&dout, sizeof(double) * n));
  HIPCHK(d, hipMemcpyAsync(din, in, sizeof(double) * n, hipMemcpyHostToDevice, d->stream));
  if (in2)
    HIPCHK(d, hipMemcpyAsync(din2, in2, sizeof(double) * n, hipMemcpyHostToDevice, d->stream));
  else
    HIPCHK(d, hipMemcpyAsync(din2, in, sizeof(double) * n, hipMemcpyHostToDevice, d->stream));
  hipLaunchKernelGGL(eh_math_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, fn, din, din2,
                     dout, n);
  HIPCHK(d, hipGetLastError());
  HIPCHK(d, hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipStreamSynchronize(d->stream));
  HIPCHK(d, hipFree(din));
  HIPCHK(d, hipFree(din2));
  HIPCHK(d, hipFree(dout));
  return EH_OK;
}

} /* extern "C" */
