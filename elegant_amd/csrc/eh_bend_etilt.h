/*
 * eh_bend_etilt.h - the exit offsets of a CSBEND with an error tilt (host
 * side, shared by the library's element compiler and the CPU checker).
 *
 * computeEtiltCentroidOffset (csbend.c:4723-4776): the bend rotated by
 * ETILT about its entrance axis ends its arc away from the nominal exit; the
 * offsets (x, x', y, y', s) of that orbit where it crosses the nominal exit
 * plane, then rotated into the lab frame by the bend's TILT
 * (rotateCoordinatesForMisalignment with -tilt, tilt_matrices.c:123-149).
 * The tracking kernel adds them to every particle leaving the bend
 * (csbend.c:1359-1364).  libm on the host, in the reference's order.
 */
#ifndef ELEGANT_AMD_EH_BEND_ETILT_H
#define ELEGANT_AMD_EH_BEND_ETILT_H

#include <math.h>

#include "eh_consts.h"

static void eh_etilt_offsets(double rho0, double angle, double etilt, double tilt, double out[5]) {
  double q1a, q2a, q3a, q1b, q2b, q3b, qp1, qp2, qp3, dz, tan_alpha, k, d[6];
  if (!etilt) {
    for (int i = 0; i < 5; i++)
      out[i] = 0;
    return;
  }
  etilt *= -1;
  q1a = (1 - cos(angle)) * rho0 * (cos(etilt) - 1);
  q2a = 0;
  q3a = (1 - cos(angle)) * rho0 * sin(etilt);
  qp1 = sin(angle) * cos(etilt);
  qp2 = cos(angle);
  k = sqrt(qp1 * qp1 + qp2 * qp2);
  qp1 /= k;
  qp2 /= k;
  qp3 = sin(angle) * sin(etilt) / k;
  tan_alpha = 1. / tan(angle) / cos(etilt);
  q1b = q1a * tan_alpha / (tan(angle) + tan_alpha);
  q2b = -q1b * tan(angle);
  dz = sqrt((q1b - q1a) * (q1b - q1a) + (q2b - q2a) * (q2b - q2a));
  q3b = q3a + qp3 * dz;
  d[0] = sqrt(q1b * q1b + q2b * q2b);
  d[1] = tan(atan(tan_alpha) - (EH_PIo2 - angle));
  d[2] = q3b;
  d[3] = qp3;
  d[4] = dz * sqrt(1 + qp3 * qp3);
  d[5] = 0;
  (void)qp1;
  (void)qp2;
  /* rotation by -tilt, with the exact cosines and sines at multiples of pi/2 */
  {
    const double a = -tilt;
    double c, s;
    if (a && !(fabs(fabs(a) - EH_PIx2) < 1e-12)) {
      if (fabs(fabs(a) - EH_PI) < 1e-12) {
        c = -1;
        s = 0;
      } else if (fabs(a - EH_PIo2) < 1e-12) {
        c = 0;
        s = 1;
      } else if (fabs(a + EH_PIo2) < 1e-12) {
        c = 0;
        s = -1;
      } else {
        c = cos(a);
        s = sin(a);
      }
      const double x = d[0], xp = d[1], y = d[2], yp = d[3];
      d[0] = x * c + y * s;
      d[2] = -x * s + y * c;
      d[1] = xp * c + yp * s;
      d[3] = -xp * s + yp * c;
    }
  }
  for (int i = 0; i < 5; i++)
    out[i] = d[i];
}

#endif
