/*
 * eh_da.c - dynamic-aperture area from per-ray limits.
 *
 * Restates trimApertureSearchResult (trimApertureSearchResult.c:16-117):
 * clip "islands" so that the boundary is monotone in each quadrant, repeat
 * until the trapezoid-rule area stops changing, return the area.  The limits
 * arrays are modified in place, as in the reference.
 */
#include "eh_host.h"

static double trim_once(int64_t lines, double *xLimit, double *yLimit, const double *dxFactor, const double *dyFactor,
                        int32_t fullPlane) {
  int64_t line;
  double area;
  if (!fullPlane) {
    /* x increases monotonically with the line index from both ends toward the middle */
    for (line = 0; line < lines / 2; line++)
      if (xLimit[line + 1] < xLimit[line])
        xLimit[line + 1] = xLimit[line];
    for (line = lines - 1; line > lines / 2; line--)
      if (xLimit[line - 1] > xLimit[line])
        xLimit[line - 1] = xLimit[line];
    /* x < 0: y increases with x; x > 0: y decreases with x */
    for (line = lines / 2; line > 0; line--)
      if (yLimit[line - 1] > yLimit[line])
        yLimit[line - 1] = yLimit[line];
    for (line = lines / 2; line < (lines - 1); line++)
      if (yLimit[line + 1] > yLimit[line])
        yLimit[line + 1] = yLimit[line];
  } else {
    for (line = 0; line < lines - 1; line++) {
      if (dxFactor[line] > 0 && dxFactor[line + 1] > 0) {
        if (dyFactor[line] > 0 && dyFactor[line + 1] > 0) {
          if (xLimit[line + 1] < xLimit[line])
            xLimit[line] = xLimit[line + 1];
        } else if (dyFactor[line] < 0 && dyFactor[line + 1] < 0) {
          if (xLimit[line + 1] > xLimit[line])
            xLimit[line + 1] = xLimit[line];
        }
      } else if (dxFactor[line] < 0 && dxFactor[line + 1] < 0) {
        if (dyFactor[line] > 0 && dyFactor[line + 1] > 0) {
          if (xLimit[line + 1] < xLimit[line])
            xLimit[line + 1] = xLimit[line];
        } else if (dyFactor[line] < 0 && dyFactor[line + 1] < 0) {
          if (xLimit[line + 1] > xLimit[line])
            xLimit[line] = xLimit[line + 1];
        }
      }
    }
    for (line = 0; line < lines - 1; line++) {
      if (dxFactor[line] > 0 && dxFactor[line + 1] > 0) {
        if (dyFactor[line] > 0 && dyFactor[line + 1] > 0) {
          if (yLimit[line + 1] > yLimit[line])
            yLimit[line + 1] = yLimit[line];
        } else if (dyFactor[line] < 0 && dyFactor[line + 1] < 0) {
          if (yLimit[line] < yLimit[line + 1])
            yLimit[line] = yLimit[line + 1];
        }
      } else if (dxFactor[line] < 0 && dxFactor[line + 1] < 0) {
        if (dyFactor[line] > 0 && dyFactor[line + 1] > 0) {
          if (yLimit[line + 1] < yLimit[line])
            yLimit[line] = yLimit[line + 1];
        } else if (dyFactor[line] < 0 && dyFactor[line + 1] < 0) {
          if (yLimit[line + 1] < yLimit[line])
            yLimit[line + 1] = yLimit[line];
        }
      }
    }
  }
  area = 0;
  for (line = 0; line < lines - 1; line++)
    area += (xLimit[line + 1] - xLimit[line]) * (yLimit[line + 1] + yLimit[line]) / 2;
  return area;
}

double eh_da_trim_area(int64_t lines, double *xLimit, double *yLimit, const double *dxFactor, const double *dyFactor,
                       int32_t fullPlane) {
  double area1 = -1, area2 = -1;
  int64_t count = lines;
  if (lines < 2 || !xLimit || !yLimit || !dxFactor || !dyFactor)
    return -1;
  while (count-- && (area2 = trim_once(lines, xLimit, yLimit, dxFactor, dyFactor, fullPlane)) != area1)
    area1 = area2;
  return area1;
}
