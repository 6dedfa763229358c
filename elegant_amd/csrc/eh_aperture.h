/*
 * eh_aperture.h - MAXAMP shapes other than the closed rectangle: the
 * (super-)ellipse and the open side, for the checks after an element and
 * inside an integrator.  Device code shared by the tracking and collective
 * kernels.  Every lane of a wave sees the same shape code (it comes from the
 * element record), so the branches on it are uniform.
 */
#ifndef EH_APERTURE_H
#define EH_APERTURE_H

#include <hip/hip_runtime.h>

#include "../../include/elegant_hip.h"
#include "eh_program.h"
#include "emath.h"

#define EHA_DEV static __device__ __forceinline__

/* x^p (0 <= p <= EH_MAX_AP_EXPONENT) with the multiplications of the SDDS
 * toolkit's ipow, as the reference's device copy orders them
 * (gpuElegant/gpu_bin_transverse_distribution.cu:8-56): explicit products
 * for p <= 8; above, ipow(x, p/2) squared, times x for odd p.  The
 * recursion is unrolled: p >> L is the first exponent <= 8 of the chain
 * p, p/2, p/4, ..., and each level back up squares and, when that level's
 * exponent is odd, multiplies by x.  p = 0 gives 1; otherwise x = 0 gives +0. */
EHA_DEV double ap_ipow(double x, int p) {
  if (p <= 0)
    return 1.0;
  if (x == 0)
    return 0.0;
  int L = 0;
  while ((p >> L) > 8)
    L++;
  double h, t;
  switch (p >> L) {
  case 1:
    h = x;
    break;
  case 2:
    h = x * x;
    break;
  case 3:
    t = x * x;
    h = t * x;
    break;
  case 4:
    t = x * x;
    h = t * t;
    break;
  case 5:
    t = x * x;
    h = t * t * x;
    break;
  case 6:
    t = x * x;
    h = t * t * t;
    break;
  case 7:
    t = x * x * x;
    h = t * t * x;
    break;
  default: /* 8 */
    t = x * x;
    t = t * t;
    h = t * t;
    break;
  }
  for (int l = L - 1; l >= 0; l--) {
    h = h * h;
    if ((p >> l) & 1)
      h = h * x;
  }
  return h;
}

/* evaluateLostWithOpenSides (limit_amplitudes.c:1105-1128): false when the
 * particle left through the open side within the other plane's half-width
 * (or with that plane unlimited) */
EHA_DEV bool ap_open_lost(int open, double dx, double dy, double xsize, double ysize) {
  switch (open) {
  case EH_OPEN_PLUS_X:
    return !(dx > 0 && (__builtin_fabs(dy) < ysize || ysize <= 0));
  case EH_OPEN_MINUS_X:
    return !(dx < 0 && (__builtin_fabs(dy) < ysize || ysize <= 0));
  case EH_OPEN_PLUS_Y:
    return !(dy > 0 && (__builtin_fabs(dx) < xsize || xsize <= 0));
  case EH_OPEN_MINUS_Y:
    return !(dy < 0 && (__builtin_fabs(dx) < xsize || xsize <= 0));
  default:
    return true;
  }
}

/* checkMultAperture's test on the already centred and rotated (x, y)
 * (multipole.c:1807-1827): the rectangle (strict >) or, with EH_AP_ELL,
 * (x/a)^p + (y/b)^q >= 1; an open side may spare the particle.  True if
 * the particle is inside. */
EHA_DEV bool ap_inside_shaped(double x, double y, double xMax, double yMax, int code) {
  bool out;
  if (!(code & EH_AP_ELL))
    out = (xMax > 0 && __builtin_fabs(x) > xMax) || (yMax > 0 && __builtin_fabs(y) > yMax);
  else {
    const double xa = x / xMax, yb = y / yMax;
    out = (ap_ipow(xa, EH_AP_XE(code)) + ap_ipow(yb, EH_AP_YE(code))) >= 1;
  }
  const int open = EH_AP_OPEN(code);
  return !(out && (open == 0 || ap_open_lost(open, x, y, xMax, yMax)));
}

/* The MAXAMP check after an element for a shaped aperture (code != 0).
 * Without EH_AP_ELL: limit_amplitudes with an open side
 * (limit_amplitudes.c:217-268; no loss-point extrapolation); with it:
 * elimit_amplitudes (limit_amplitudes.c:531-596), whose lost rows hold
 * sqrt((Po(1+delta))^2 + 1) in place of delta, and whose NaN/infinity test
 * comes first and leaves x and y as they are.  True if the particle is lost;
 * the row then holds the loss values. */
EHA_DEV bool ap_maxamp_shaped(double &x, double xp, double &y, double yp, double &s, double &dp, double xmax,
                              double ymax, int code, bool extrap, double z, double Po) {
  const int open = EH_AP_OPEN(code);
  const bool bad = __builtin_isinf(x) || __builtin_isinf(y) || __builtin_isnan(x) || __builtin_isnan(y);
  if (!(code & EH_AP_ELL)) {
    int is_out = 0;
    if (xmax > 0 && __builtin_fabs(x) > xmax)
      is_out += 1;
    if (ymax > 0 && __builtin_fabs(y) > ymax)
      is_out += 2;
    if (open && !ap_open_lost(open, x, y, xmax, ymax))
      is_out = 0;
    if (bad)
      is_out += 4;
    if (!is_out)
      return false;
    const double dz = 0;
    x += dz * xp;
    y += dz * yp;
    s = z + dz;
    dp = Po * (1 + dp);
    return true;
  }
  if (bad) {
    const double P = Po * (1 + dp);
    s = z;
    dp = em_sqrt(P * P + 1);
    return true;
  }
  const int xe = EH_AP_XE(code), ye = EH_AP_YE(code);
  const double a2 = ap_ipow(xmax, xe), b2 = ap_ipow(ymax, ye);
  if (!((ap_ipow(x, xe) / a2 + ap_ipow(y, ye) / b2) > 1))
    return false;
  if (open && !ap_open_lost(open, x, y, xmax, ymax))
    return false;
  double dz = 0;
  if (extrap && !open && xe == 2 && ye == 2) {
    const double c0 = x * x / a2 + y * y / b2 - 1;
    const double c1 = 2 * (x * xp / a2 + y * yp / b2);
    const double c2 = xp * xp / a2 + yp * yp / b2;
    const double det = c1 * c1 - 4 * c0 * c2;
    if (z > 0 && c2 != 0 && det >= 0) {
      dz = (-c1 + em_sqrt(det)) / (2 * c2);
      if (dz > 0)
        dz = (-c1 - em_sqrt(det)) / (2 * c2);
      if ((z + dz) < 0)
        dz = -z;
      x += dz * xp;
      y += dz * yp;
    }
  }
  const double P = Po * (1 + dp);
  s = z + dz;
  dp = em_sqrt(P * P + 1);
  return true;
}

#endif
