/*
 * oracle.c - CPU restatement of elegant's 6D tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the timed CPU
 * baseline.  Every routine follows the cited reference routine's operation
 * order (ISO C, no FP contraction), including its quirks:
 *  - the "radiation" round trip q /= (1+dp) ... q *= (1+dp) taken whenever
 *    isr_coef != 0, which is always (ISR=0 makes it negative,
 *    multipole.c:790-793; csbend.c:972-975);
 *  - swap-to-tail compaction of lost rows with s -> loss position and
 *    delta -> Po(1+delta);
 *  - exit misalignment applied to all n_part rows, lost ones included
 *    (multipole.c:1103-1104).
 * Dead computations that only feed sigmaDelta2 (NULL on the do_tracking
 * path, do_tracking.c:2011) are left out; they change no coordinate.
 *
 * What pins each routine (tests/test_oracle.py and friends): closed-form
 * maps (drift, thick quadrupole, thin multipoles, sector bend and edges,
 * RF energy gain, drift-aperture loss point), invariants (symplecticity,
 * roll and offset symmetries, conservation in add_to_particle_energy) and
 * brute force (Kahan and the fixed-point mean against exact rationals, the
 * FFT against numpy and a direct DFT, histogram and convolution against
 * numpy, WAKE against ZLONGIT for one resonator).
 * Parity unpinned: the SDDS-toolkit routines the reference calls but does
 * not ship (realFFT's packing and normalisation, compute_average's
 * summation, find_min_max) are restated from their published behaviour and
 * checked against mathematics, not against the toolkit's bits; the physical
 * constants (CODATA 2018, eh_consts.h) are this build's choice; and the
 * bit-level operation order inside each restated routine rests on review of
 * the cited lines, since the reference cannot be compiled here.  The
 * toolkit's ipow (aperture powers) is the exception among the toolkit
 * routines: the reference ships a device copy of it whose multiplication
 * order is followed here.
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "eh_bend_etilt.h"
#include "eh_bend_field.h"
#include "eh_consts.h"
#include "eh_fixedsum.h"

#ifdef ORACLE_EMATH
#include "emath.h"
#define P_SIN em_sin
#define P_COS em_cos
#define P_ASIN em_asin
#define P_TAN em_tan
#define P_LOG em_log
#define ORACLE_MODE 1
#else
#define P_SIN sin
#define P_COS cos
#define P_ASIN asin
#define P_TAN tan
#define P_LOG log
#define ORACLE_MODE 0
#endif

int or_math_mode(void) { return ORACLE_MODE; }

int64_t or_ctx_effort(const or_ctx *c) { return c->effort; }

/* evaluate the per-particle math of this build on an array (fn: 0 sin, 1 cos,
 * 2 asin, 3 log, 4 tan, 5 sqrt, 6 in/in2) */
static double or_ipow(double x, long p);

void or_math_eval(int fn, const double *in, const double *in2, double *out, long n) {
  for (long i = 0; i < n; i++) {
    double v = in[i];
    switch (fn) {
    case 0:
      out[i] = P_SIN(v);
      break;
    case 1:
      out[i] = P_COS(v);
      break;
    case 2:
      out[i] = P_ASIN(v);
      break;
    case 3:
      out[i] = P_LOG(v);
      break;
    case 4:
      out[i] = P_TAN(v);
      break;
    case 5:
      out[i] = sqrt(v);
      break;
    case 7:
      out[i] = or_ipow(v, (long)in2[i]);
      break;
    default:
      out[i] = v / in2[i];
      break;
    }
  }
}

static double sqr(double x) { return x * x; }

or_ctx *or_ctx_new(void) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  return c;
}

void or_ctx_free(or_ctx *c) {
  if (!c)
    return;
  free(c->last_hist);
  free(c->last_volt);
  free(c);
}

static void swap_rows(double *a, double *b, int nprops) {
  double t;
  for (int i = 0; i < nprops; i++) {
    t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}

/* rows carry at least 8 columns: x xp y yp s dp id lossPass */
/* columns swapped with a row (elegant's totalPropertiesPerParticle): 8, or
 * more when the rows carry the bunch index (column 8) and further properties */
/* the particle species: elegant's globals particleMass, particleCharge,
 * particleMassMV, particleRadius and particleRelSign (abrat.c:182-186 set the
 * electron's; change_particle, elegant.c:2933-2991, changes them) */
static double particleMass = EH_ME_MKS, particleCharge = EH_E_MKS, particleMassMV = EH_ME_MEV,
              particleRadius = EH_RE_MKS, particleRelSign = 1;
void or_set_species(double mass, double charge, double massMV, double radius, double relSign) {
  particleMass = mass;
  particleCharge = charge;
  particleMassMV = massMV;
  particleRadius = radius;
  particleRelSign = relSign;
}

static int or_nprops = 8;
#define OR_NPROPS or_nprops
void or_set_row_width(long nprops) { or_nprops = nprops < 8 ? 8 : (int)nprops; }
long or_row_width(void) { return or_nprops; }

/* ---------------------------------------------------------------- drift */

/* track_particles, order-2 branch (matrix.c:389-429), with the drift matrix
 * of compute_matrices.c:1101-1104: C[4]=L, R=identity with R[0][1]=R[2][3]=L,
 * T[4][1][1]=T[4][3][3]=L/2.  The sparse loop is executed literally. */
/* track_particles for a zero-length corrector with a kick: corrector_matrix
 * (bend_matrix.c:300-330) turned by tilt_matrices (tilt_matrices.c:18-73,
 * IRot (M Rot) by concat_matrices, concat_mat.c:62-97), applied as the
 * first-order case of track_particles (matrix.c:430-458) */
static void or_mat1_mul(double C2[6], double R2[6][6], const double C1[6], const double R1[6][6],
                        const double C0[6], const double R0[6][6]) {
  long i, j, m;
  double sum;
  for (i = 0; i < 6; i++) {
    for (j = 0, sum = 0; j < 6; j++)
      sum += R1[i][j] * C0[j];
    C2[i] = sum + C1[i];
  }
  for (i = 0; i < 6; i++)
    for (m = 0; m < 6; m++) {
      for (j = 0, sum = 0; j < 6; j++)
        sum += R1[i][j] * R0[j][m];
      R2[i][m] = sum;
    }
}

void or_thin_corrector(double **coord, long np, double kick, double tilt) {
  double C[6] = {0, kick, 0, 0, 0, 0}, R[6][6] = {{0}}, sum, tmp[6];
  long i, j, ip;
  for (i = 0; i < 6; i++)
    R[i][i] = 1;
  R[1][5] = -sin(kick);
  if (tilt) {
    double st = sin(tilt), ct = cos(tilt), Z[6] = {0}, Rot[6][6] = {{0}}, IRot[6][6] = {{0}}, Cr[6], Rr[6][6];
    if (fabs(tilt - EH_PIo2) < 1e-12)
      st = 1, ct = 0;
    else if (fabs(tilt + EH_PIo2) < 1e-12)
      st = -1, ct = 0;
    else if (fabs(tilt - EH_PI) < 1e-12 || fabs(tilt + EH_PI) < 1e-12)
      st = 0, ct = -1;
    Rot[0][0] = Rot[2][2] = Rot[1][1] = Rot[3][3] = ct;
    IRot[0][0] = IRot[2][2] = IRot[1][1] = IRot[3][3] = ct;
    Rot[0][2] = Rot[1][3] = st;
    IRot[0][2] = IRot[1][3] = -st;
    Rot[2][0] = Rot[3][1] = -st;
    IRot[2][0] = IRot[3][1] = -(-st);
    Rot[4][4] = Rot[5][5] = IRot[4][4] = IRot[5][5] = 1;
    or_mat1_mul(Cr, Rr, C, R, Z, Rot);
    or_mat1_mul(C, R, Z, IRot, Cr, Rr);
  }
  for (ip = np - 1; ip >= 0; ip--) {
    double *x = coord[ip];
    for (i = 5; i >= 0; i--) {
      sum = C[i];
      for (j = 5; j >= 0; j--)
        sum += R[i][j] * x[j];
      tmp[i] = sum;
    }
    for (i = 5; i >= 0; i--)
      x[i] = tmp[i];
  }
}

void or_track_particles_drift(double **coord, long np, double length, int order) {
  double C[6] = {0, 0, 0, 0, 0, 0};
  double R[6][6];
  double T[6][6][6];
  double temp[6];
  memset(R, 0, sizeof(R));
  memset(T, 0, sizeof(T));
  C[4] = length;
  for (int i = 0; i < 6; i++)
    R[i][i] = 1;
  R[0][1] = R[2][3] = length;
  if (order > 1)
    T[4][1][1] = T[4][3][3] = length / 2;
  for (long ip = np - 1; ip >= 0; ip--) {
    double *ini = coord[ip];
    for (int i = 5; i >= 0; i--) {
      double sum = C[i];
      if (order > 1) {
        for (int j = 5; j >= 0; j--) {
          double coord_j = ini[j];
          if (coord_j) {
            double sum1 = R[i][j];
            for (int k = j; k >= 0; k--)
              sum1 += T[i][j][k] * ini[k];
            sum += sum1 * coord_j;
          }
        }
      } else {
        for (int j = 5; j >= 0; j--)
          sum += R[i][j] * ini[j];
      }
      temp[i] = sum;
    }
    for (int i = 5; i >= 0; i--)
      ini[i] = temp[i];
  }
}

/* offset_beam (do_tracking.c:3243-3306) over every particle (no PID range),
 * the rows from the last to the first as the reference loops */
void or_offset_beam(double **coord, long nToTrack, const EH_MALIGN *offset, double P_central) {
  long i_part;
  double *part, pc, beta, gamma, t, ds;
  for (i_part = nToTrack - 1; i_part >= 0; i_part--) {
    part = coord[i_part];
    if (offset->dz)
      ds = offset->dz * sqrt(1 + sqr(part[1]) + sqr(part[3]));
    else
      ds = 0;
    part[0] += offset->dx + offset->dz * part[1];
    part[1] += offset->dxp;
    part[2] += offset->dy + offset->dz * part[3];
    part[3] += offset->dyp;
    part[4] += ds;
    if (offset->dt || offset->dp || offset->de) {
      pc = P_central * (1 + part[5]);
      beta = pc / (gamma = sqrt(1 + pc * pc));
      t = part[4] / (beta * EH_C_MKS) + offset->dt;
      if (offset->dp) {
        part[5] += offset->dp;
        pc = P_central * (1 + part[5]);
        beta = pc / sqrt(1 + pc * pc);
      }
      if (offset->de) {
        gamma += offset->de * gamma;
        pc = sqrt(gamma * gamma - 1);
        beta = pc / gamma;
        part[5] = (pc - P_central) / P_central;
      }
      part[4] = t * beta * EH_C_MKS;
    }
  }
}

/* rectangular_collimator (limit_amplitudes.c:32-185): the entrance loop
 * and, for a positive length, the drift loop with the exit check, each
 * compacting its losses by swaps */
static long or_open_side_lost(long code, double dx, double dy, double xsize, double ysize);
void or_exact_drift(double **part, long np, double length);
long or_rectangular_collimator(double **initial, const EH_RCOL *rcol, long np, double z, double Po) {
  double length, *ini, xsize = rcol->x_max, ysize = rcol->y_max, x_center = rcol->dx, y_center = rcol->dy;
  double x1, y1, zx, zy, dx, dy;
  long ip, itop, is_out, lost, openCode = rcol->openSide;
  if (xsize <= 0 && ysize <= 0) {
    or_exact_drift(initial, np, rcol->length);
    return np;
  }
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    ini = initial[ip];
    dx = ini[0] - x_center;
    dy = ini[2] - y_center;
    lost = 0;
    if ((xsize > 0 && fabs(dx) > xsize) || (ysize > 0 && fabs(dy) > ysize))
      lost = openCode ? or_open_side_lost(openCode, dx, dy, xsize, ysize) : 1;
    else if (isinf(ini[0]) || isinf(ini[2]) || isnan(ini[0]) || isnan(ini[2]))
      lost = 1;
    if (rcol->invert)
      lost = !lost;
    if (lost) {
      swap_rows(initial[ip], initial[itop], OR_NPROPS);
      initial[itop][4] = z;
      initial[itop][5] = Po * (1 + initial[itop][5]);
      --itop;
      --ip;
      --np;
    }
  }
  if (np == 0 || (length = rcol->length) <= 0)
    return np;
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    ini = initial[ip];
    x1 = ini[0] + length * ini[1];
    y1 = ini[2] + length * ini[3];
    dx = x1 - x_center;
    dy = y1 - y_center;
    is_out = 0;
    if (xsize > 0 && fabs(dx) > xsize)
      is_out += 1 * (openCode ? or_open_side_lost(openCode, dx, 0, xsize, ysize) : 1);
    if (ysize > 0 && fabs(dy) > ysize)
      is_out += 2 * (openCode ? or_open_side_lost(openCode, 0, dy, xsize, ysize) : 1);
    if (isinf(x1) || isinf(y1) || isnan(x1) || isnan(y1))
      is_out += 4;
    if (is_out & 4) {
      ini[4] = z + length;
      ini[0] = x1;
      ini[2] = y1;
      ini[5] = Po * (1 + ini[5]);
      swap_rows(initial[ip], initial[itop], OR_NPROPS);
      --itop;
      --ip;
      --np;
    } else if ((is_out && !rcol->invert) || (!is_out && rcol->invert)) {
      if (!openCode) {
        zx = zy = DBL_MAX;
        if (is_out & 1 && ini[1] != 0)
          zx = ((ini[1] > 0 ? 1 : -1) * xsize - (ini[0] - x_center)) / ini[1];
        if (is_out & 2 && ini[3] != 0)
          zy = ((ini[3] > 0 ? 1 : -1) * ysize - (ini[2] - y_center)) / ini[3];
        if (zx < zy) {
          ini[0] += ini[1] * zx;
          ini[2] += ini[3] * zx;
          ini[4] = z + zx;
        } else {
          ini[0] += ini[1] * zy;
          ini[2] += ini[3] * zy;
          ini[4] = z + zy;
        }
      }
      ini[5] = Po * (1 + ini[5]);
      swap_rows(initial[ip], initial[itop], OR_NPROPS);
      --itop;
      --ip;
      --np;
    } else {
      ini[4] += length * sqrt(1 + sqr(ini[1]) + sqr(ini[3]));
      ini[0] = x1;
      ini[2] = y1;
    }
  }
  return np;
}

/* exactDrift (csbend.c:3767-3790) */
void or_exact_drift(double **part, long np, double length) {
  for (long i = 0; i < np; i++) {
    double *coord = part[i];
    coord[0] += coord[1] * length;
    coord[2] += coord[3] * length;
    coord[4] += length * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  }
}

/* --------------------------------------------------------- misalignment */

/* offsetBeamCoordinatesForMisalignment (malign_mat.c:172-182) */
static void offset_beam(double **coord, long np, double dx, double dy, double dz) {
  for (long ip = np - 1; ip >= 0; ip--) {
    double *part = coord[ip];
    part[4] += dz * sqrt(1 + sqr(part[1]) + sqr(part[3]));
    part[0] = part[0] - dx + dz * part[1];
    part[2] = part[2] - dy + dz * part[3];
  }
}

/* rotateBeamCoordinatesForMisalignment (tilt_matrices.c:156-188) */
static void rotate_beam(double **part, long np, double angle) {
  double sin_a, cos_a;
  if (!angle || fabs(fabs(angle) - EH_PIx2) < 1e-12)
    return;
  if (fabs(fabs(angle) - EH_PI) < 1e-12) {
    cos_a = -1;
    sin_a = 0;
  } else if (fabs(angle - EH_PIo2) < 1e-12) {
    cos_a = 0;
    sin_a = 1;
  } else if (fabs(angle + EH_PIo2) < 1e-12) {
    cos_a = 0;
    sin_a = -1;
  } else {
    cos_a = cos(angle);
    sin_a = sin(angle);
  }
  for (long i = 0; i < np; i++) {
    double *coord = part[i];
    double x = coord[0], xp = coord[1], y = coord[2], yp = coord[3];
    coord[0] = x * cos_a + y * sin_a;
    coord[2] = -x * sin_a + y * cos_a;
    coord[1] = xp * cos_a + yp * sin_a;
    coord[3] = -xp * sin_a + yp * cos_a;
  }
}

/* offsetParticlesForMisalignment mode 0 (malign_mat.c:480-535) */
static void offset_for_misalignment(double **coord, long np, double dx, double dy, double dz, double tilt, int face) {
  if (face == 1) {
    if (dx || dy || dz)
      offset_beam(coord, np, dx, dy, dz);
    if (tilt)
      rotate_beam(coord, np, tilt);
  } else {
    if (tilt)
      rotate_beam(coord, np, -tilt);
    if (dx || dy || dz)
      offset_beam(coord, np, -dx, -dy, -dz);
  }
}

/* ------------------------------------------------------------ apertures */

/* ipow: x^p by the fixed multiplication order of the SDDS toolkit's ipow,
 * as the reference's own device copy spells it out
 * (gpuElegant/gpu_bin_transverse_distribution.cu:8-56): explicit products
 * up to p = 8, then the square of ipow(x, p/2), times x when p is odd.
 * x = 0 gives 0 (p > 0).  Only p >= 1 occurs here. */
static double or_ipow(double x, long p) {
  double h;
  if (!x)
    return p == 0 ? 1.0 : 0.0;
  switch (p) {
  case 0:
    return 1.0;
  case 1:
    return x;
  case 2:
    return x * x;
  case 3:
    h = x * x;
    return h * x;
  case 4:
    h = x * x;
    return h * h;
  case 5:
    h = x * x;
    return h * h * x;
  case 6:
    h = x * x;
    return h * h * h;
  case 7:
    h = x * x * x;
    return h * h * x;
  case 8:
    h = x * x;
    h = h * h;
    return h * h;
  default:
    h = or_ipow(x, p / 2);
    return (p & 1) ? h * h * x : h * h;
  }
}

/* evaluateLostWithOpenSides (limit_amplitudes.c:1105-1128): a particle
 * outside the aperture on its open side, within the other plane's
 * half-width (or with that plane open), is not lost.  Codes as track.h:1515-1518. */
static long or_open_side_lost(long code, double dx, double dy, double xsize, double ysize) {
  long lost = 1;
  switch (code) {
  case EH_OPEN_PLUS_X:
    if (dx > 0 && (fabs(dy) < ysize || ysize <= 0))
      lost = 0;
    break;
  case EH_OPEN_MINUS_X:
    if (dx < 0 && (fabs(dy) < ysize || ysize <= 0))
      lost = 0;
    break;
  case EH_OPEN_PLUS_Y:
    if (dy > 0 && (fabs(dx) < xsize || xsize <= 0))
      lost = 0;
    break;
  case EH_OPEN_MINUS_Y:
    if (dy < 0 && (fabs(dx) < xsize || xsize <= 0))
      lost = 0;
    break;
  default:
    break;
  }
  return lost;
}

typedef struct {
  double xMax, yMax, xCen, yCen, reverseTilt, cs0, cs1;
  int present, elliptical, xExponent, yExponent, openSide;
} mult_ap;

/* setupMultApertureData from a MAXAMP (multipole.c:1725-1763; no
 * aperture file, no APCONTOUR) */
static void setup_mult_ap(mult_ap *ap, double reverseTilt, const EH_MAXAMP *maxamp) {
  ap->reverseTilt = reverseTilt;
  ap->cs0 = cos(reverseTilt);
  ap->cs1 = sin(reverseTilt);
  ap->xCen = ap->yCen = 0;
  ap->xMax = ap->yMax = 0;
  ap->present = ap->elliptical = ap->openSide = 0;
  ap->xExponent = ap->yExponent = 0;
  if (maxamp) {
    ap->xMax = maxamp->x_max;
    ap->yMax = maxamp->y_max;
    ap->elliptical = maxamp->elliptical;
    ap->present = ap->xMax > 0 || ap->yMax > 0;
    ap->xExponent = ap->yExponent = maxamp->exponent;
    if (maxamp->yExponent)
      ap->yExponent = maxamp->yExponent;
    ap->openSide = maxamp->openSide;
  }
}

/* checkMultAperture (multipole.c:1786-1833): rectangle (strict >), or
 * (super-)ellipse (x/a)^p + (y/b)^q >= 1 when elliptical with both
 * half-axes positive; an open side can spare the particle either way */
static int check_mult_ap(double xInput, double yInput, const mult_ap *ap) {
  double x = xInput - ap->xCen;
  double y = yInput - ap->yCen;
  if (ap->reverseTilt) {
    double x0 = x, y0 = y;
    x = x0 * ap->cs0 + y0 * ap->cs1;
    y = -x0 * ap->cs1 + y0 * ap->cs0;
  }
  if (ap->elliptical == 0 || ap->xMax <= 0 || ap->yMax <= 0) {
    if ((ap->xMax > 0 && fabs(x) > ap->xMax) || (ap->yMax > 0 && fabs(y) > ap->yMax)) {
      if (ap->openSide == 0 || or_open_side_lost(ap->openSide, x, y, ap->xMax, ap->yMax))
        return 0;
    }
  } else {
    double xa = x / ap->xMax;
    double yb = y / ap->yMax;
    if ((or_ipow(xa, ap->xExponent) + or_ipow(yb, ap->yExponent)) >= 1) {
      if (ap->openSide == 0 || or_open_side_lost(ap->openSide, x, y, ap->xMax, ap->yMax))
        return 0;
    }
  }
  return 1;
}

/* limit_amplitudes (limit_amplitudes.c:193-272) */
long or_limit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                         long openCode) {
  long ip, itop, is_out;
  double *part, dz, dzx, dzy;
  if (xmax < 0 && ymax < 0)
    return np;
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    part = coord[ip];
    is_out = 0;
    if (xmax > 0 && fabs(part[0]) > xmax)
      is_out += 1;
    if (ymax > 0 && fabs(part[2]) > ymax)
      is_out += 2;
    if (openCode)
      is_out *= or_open_side_lost(openCode, part[0], part[2], xmax, ymax);
    if (isinf(part[0]) || isinf(part[2]) || isnan(part[0]) || isnan(part[2]))
      is_out += 4;
    dz = 0;
    if (is_out && !(is_out & 4) && !openCode && extrapolate_z) {
      dzx = dzy = -DBL_MAX;
      if (is_out & 1 && part[1] != 0)
        dzx = (part[0] - (part[1] > 0 ? 1 : -1) * xmax) / part[1];
      if (is_out & 2 && part[3] != 0)
        dzy = (part[2] - (part[3] > 0 ? 1 : -1) * ymax) / part[3];
      if (dzx > dzy)
        dz = -dzx;
      else
        dz = -dzy;
      if (dz == -DBL_MAX)
        dz = 0;
    }
    if (is_out) {
      if (ip != itop)
        swap_rows(coord[ip], coord[itop], OR_NPROPS);
      coord[itop][0] += dz * coord[itop][1];
      coord[itop][2] += dz * coord[itop][3];
      coord[itop][4] = z + dz;
      coord[itop][5] = Po * (1 + coord[itop][5]);
      --itop;
      --ip;
      np--;
    }
  }
  return np;
}

/* elimit_amplitudes (limit_amplitudes.c:472-598): the (super-)elliptical
 * MAXAMP after an element.  Its loss rows differ from limit_amplitudes':
 * delta becomes sqrt((Po(1+delta))^2 + 1), and the loss point is
 * extrapolated back along a drift only for a closed true ellipse.  Invalid
 * exponents are an error (the reference exits). */
long or_elimit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                          long openCode, long exponent, long yexponent) {
  long ip, itop, lost, xe, ye;
  double *part, a2, b2, c0, c1, c2, dz, det;
  if ((xe = exponent) < 2 || xe % 2)
    return EH_ERR_INVALID;
  ye = xe;
  if (yexponent) {
    if ((ye = yexponent) < 2 || ye % 2)
      return EH_ERR_INVALID;
  }
  if (xmax <= 0 || ymax <= 0) {
    if (xmax > 0 || ymax > 0)
      return or_limit_amplitudes(coord, xmax, ymax, np, z, Po, extrapolate_z, openCode);
    return np;
  }
  a2 = or_ipow(xmax, xe);
  b2 = or_ipow(ymax, ye);
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    part = coord[ip];
    if (isinf(part[0]) || isinf(part[2]) || isnan(part[0]) || isnan(part[2])) {
      swap_rows(coord[ip], coord[itop], OR_NPROPS);
      coord[itop][4] = z;
      coord[itop][5] = sqrt((Po * (1 + coord[itop][5])) * (Po * (1 + coord[itop][5])) + 1);
      --itop;
      --ip;
      np--;
      continue;
    }
    lost = 0;
    if ((or_ipow(part[0], xe) / a2 + or_ipow(part[2], ye) / b2) > 1)
      lost = 1;
    else
      continue;
    if (openCode)
      lost *= or_open_side_lost(openCode, part[0], part[2], xmax, ymax);
    if (lost) {
      dz = 0;
      if (extrapolate_z && !openCode && xe == 2 && ye == 2) {
        c0 = part[0] * part[0] / a2 + part[2] * part[2] / b2 - 1;
        c1 = 2 * (part[0] * part[1] / a2 + part[2] * part[3] / b2);
        c2 = part[1] * part[1] / a2 + part[3] * part[3] / b2;
        det = c1 * c1 - 4 * c0 * c2;
        if (z > 0 && c2 && det >= 0) {
          if ((dz = (-c1 + sqrt(det)) / (2 * c2)) > 0)
            dz = (-c1 - sqrt(det)) / (2 * c2);
          if ((z + dz) < 0)
            dz = -z;
          part[0] += dz * part[1];
          part[2] += dz * part[3];
        }
      }
      swap_rows(coord[ip], coord[itop], OR_NPROPS);
      coord[itop][4] = z + dz;
      coord[itop][5] = sqrt((Po * (1 + coord[itop][5])) * (Po * (1 + coord[itop][5])) + 1);
      --itop;
      --ip;
      np--;
    }
  }
  return np;
}

/* the MAXAMP check do_tracking makes after an element (do_tracking.c:2461-2470) */
static long or_maxamp_check(double **coord, const EH_MAXAMP *maxamp, double x_max, double y_max, long np, double z,
                            double Po, long extrapolate_z) {
  long openCode = maxamp ? maxamp->openSide : 0;
  if (!maxamp || !maxamp->elliptical)
    return or_limit_amplitudes(coord, x_max, y_max, np, z, Po, extrapolate_z, openCode);
  return or_elimit_amplitudes(coord, x_max, y_max, np, z, Po, extrapolate_z, openCode, maxamp->exponent,
                              maxamp->yExponent);
}

/* removeInvalidParticles (limit_amplitudes.c:274-320) */
long or_remove_invalid_particles(double **coord, long np, double z, double Po) {
  long ip, itop, is_out, ic;
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    double *part = coord[ip];
    is_out = 0;
    for (ic = 0; ic < 6; ic++)
      if (isnan(part[ic])) {
        is_out = 1;
        break;
      }
    if (part[5] <= -1)
      is_out = 1;
    if (is_out) {
      swap_rows(coord[ip], coord[itop], OR_NPROPS);
      coord[itop][4] = z;
      coord[itop][5] = Po * (1 + coord[itop][5]);
      --itop;
      --ip;
      np--;
    }
  }
  return np;
}

/* ---------------------------------------------------------- multipoles */

/* BETA is 2^(1/3); Forest-Ruth/Yoshida fractions (multipole.c:1119-1149) */
#define OR_BETA 1.25992104989487316477
static const double driftFrac2[2] = {0.5, 0.5};
static const double kickFrac2[2] = {1.0, 0.0};
static const double driftFrac4[4] = {0.5 / (2 - OR_BETA), (1 - OR_BETA) / (2 - OR_BETA) / 2,
                                     (1 - OR_BETA) / (2 - OR_BETA) / 2, 0.5 / (2 - OR_BETA)};
static const double kickFrac4[4] = {1. / (2 - OR_BETA), -OR_BETA / (2 - OR_BETA), 1 / (2 - OR_BETA), 0};
static const double driftFrac6[8] = {0.39225680523878,     0.5100434119184585, -0.47105338540975655,
                                     0.0687531682525181,   0.0687531682525181, -0.47105338540975655,
                                     0.5100434119184585,   0.39225680523878};
static const double kickFrac6[8] = {0.784513610477560,  0.235573213359357, -1.17767998417887,
                                    1.3151863206839063, -1.17767998417887, 0.235573213359357,
                                    0.784513610477560,  0};

static void integrator_fracs(int order, const double **drift, const double **kick, int *nsub) {
  if (order == 2) {
    *drift = driftFrac2;
    *kick = kickFrac2;
    *nsub = 2;
  } else if (order == 6) {
    *drift = driftFrac6;
    *kick = kickFrac6;
    *nsub = 8;
  } else {
    *drift = driftFrac4;
    *kick = kickFrac4;
    *nsub = 4;
  }
}

static double dfactorial(long n) {
  double f = 1;
  for (long i = 2; i <= n; i++)
    f *= i;
  return f;
}

/* expansion_coefficients (multipole.h:137-165) */
static void expansion_coefficients(long n, double *coef) {
  for (long i = 0; i <= n; i++)
    coef[i] = (((i / 2) % 2) ? -1.0 : 1.0) / (dfactorial(i) * dfactorial(n - i));
}

/* apply_canonical_multipole_kicks, TURBO_MODE>=2 form (multipole.h:208-241) */
static void apply_kicks(double *qx, double *qy, const double *xpow, const double *ypow, long order, double KnL,
                        int skew) {
  double coef[16], sum_Fx = 0, sum_Fy = 0;
  long i;
  expansion_coefficients(order, coef);
  for (i = 1; i <= order; i += 2)
    sum_Fx += coef[i] * xpow[order - i] * ypow[i];
  for (i = 0; i <= order; i += 2)
    sum_Fy += coef[i] * xpow[order - i] * ypow[i];
  if (skew) {
    double t = sum_Fx;
    sum_Fx = sum_Fy;
    sum_Fy = t;
    sum_Fx = -sum_Fx;
  }
  *qx -= KnL * sum_Fy;
  *qy += KnL * sum_Fx;
}

static void fill_power(double x, double *xpow, long order) {
  xpow[0] = 1;
  for (long i = 1; i <= order; i++)
    xpow[i] = xpow[i - 1] * x;
}

/* convertSlopesToMomenta / convertMomentaToSlopes (multipole.h:250-281) */
static int slopes_to_momenta(double *qx, double *qy, double xp, double yp, double delta, int expand) {
  if (expand) {
    *qx = (1 + delta) * xp;
    *qy = (1 + delta) * yp;
  } else {
    double denom = sqrt(1 + sqr(xp) + sqr(yp));
    *qx = (1 + delta) * xp / denom;
    *qy = (1 + delta) * yp / denom;
  }
  return 1;
}

static int momenta_to_slopes(double *xp, double *yp, double qx, double qy, double delta, int expand) {
  if (expand) {
    *xp = qx / (1 + delta);
    *yp = qy / (1 + delta);
  } else {
    double denom;
    if ((denom = sqr(1 + delta) - sqr(qx) - sqr(qy)) <= 0)
      return 0;
    denom = sqrt(denom);
    *xp = qx / denom;
    *yp = qy / denom;
  }
  return 1;
}

typedef struct {
  double KnL[3];
  long order[3];
  int skew[3];
  double dx, dy, dz, tilt, drift, xkick, ykick;
  long nSlices, integ_order;
  int expand;
  int doEndDrift; /* KQUAD LEFFECTIVE: exact drifts of lEnd around the body (multipole.c:766-770) */
  double lEnd;
} mult_setup;

/* integrate_kick_multipole_ordn (multipole.c:1122-1410) with no error,
 * edge or steering multipoles, radial=0, no obstructions, iSlice<0. */
static int integrate_kick_multipole(double *coord, const mult_setup *m, double Po, const mult_ap *ap,
                                    double *dzLoss) {
  double p, qx, qy, beta0, dp, s;
  double x, y, xp, yp;
  long i_kick, step, iOrder, maxOrder;
  double dsh, xpow[8], ypow[8];
  const double *driftFrac, *kickFrac;
  int nSubsteps;
  long n_parts = m->nSlices;
  double drift, xkick, ykick;
  (void)p;
  (void)beta0;

  integrator_fracs((int)m->integ_order, &driftFrac, &kickFrac, &nSubsteps);
  drift = m->drift / n_parts;
  xkick = m->xkick / n_parts;
  ykick = m->ykick / n_parts;

  x = coord[0];
  xp = coord[1];
  y = coord[2];
  yp = coord[3];
  s = 0;
  dp = coord[5];
  p = Po * (1 + dp);
  beta0 = p / sqrt(sqr(p) + 1);

  if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp))
    return 0;
  if (fabs(x) > EH_COORD_LIMIT || fabs(y) > EH_COORD_LIMIT || fabs(xp) > EH_SLOPE_LIMIT || fabs(yp) > EH_SLOPE_LIMIT)
    return 0;

  slopes_to_momenta(&qx, &qy, xp, yp, dp, m->expand);
  maxOrder = m->order[0];
  if (m->order[1] > maxOrder)
    maxOrder = m->order[1];
  if (m->order[2] > maxOrder)
    maxOrder = m->order[2];

  if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, m->expand))
    return 0;

  *dzLoss = 0;
  for (i_kick = 0; i_kick < n_parts; i_kick++) {
    if (!check_mult_ap(x + m->dx, y + m->dy, ap)) {
      coord[0] = x;
      coord[2] = y;
      return 0;
    }
    for (step = 0; step < nSubsteps; step++) {
      if (drift) {
        dsh = drift * driftFrac[step];
        x += xp * dsh;
        y += yp * dsh;
        if (m->expand)
          s += dsh * (1 + (sqr(xp) + sqr(yp)) / 2);
        else
          s += dsh * sqrt(1 + sqr(xp) + sqr(yp));
        *dzLoss += dsh;
      }
      if (!kickFrac[step])
        break;
      fill_power(x, xpow, maxOrder);
      fill_power(y, ypow, maxOrder);
      for (iOrder = 0; iOrder < 3; iOrder++)
        if (m->KnL[iOrder])
          apply_kicks(&qx, &qy, xpow, ypow, m->order[iOrder], m->KnL[iOrder] / n_parts * kickFrac[step],
                      m->skew[iOrder]);
      if (xkick)
        apply_kicks(&qx, &qy, xpow, ypow, 0, -xkick * kickFrac[step], 0);
      if (ykick)
        apply_kicks(&qx, &qy, xpow, ypow, 0, -ykick * kickFrac[step], 1);
      if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, m->expand))
        return 0;
      if (drift) {
        /* radiation branch with rad_coef = 0 and isr_coef < 0 (multipole.c:1321-1351) */
        qx /= (1 + dp);
        qy /= (1 + dp);
        qx *= (1 + dp);
        qy *= (1 + dp);
        if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, m->expand))
          return 0;
      }
    }
  }

  if (!check_mult_ap(x + m->dx, y + m->dy, ap)) {
    coord[0] = x;
    coord[2] = y;
    return 0;
  }
  if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, m->expand))
    return 0;

  coord[0] = x;
  coord[1] = xp;
  coord[2] = y;
  coord[3] = yp;
  coord[4] += s;
  coord[5] = dp;

  if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp))
    return 0;
  if (fabs(x) > EH_COORD_LIMIT || fabs(y) > EH_COORD_LIMIT || fabs(xp) > EH_SLOPE_LIMIT || fabs(yp) > EH_SLOPE_LIMIT)
    return 0;
  return 1;
}

/* the per-type set-up of multipole_tracking2 (multipole.c:765-1017) */
static int multipole_setup(const eh_element *elem, double Po, mult_setup *m) {
  long n_kicks = 0, nSlices = 0, integ_order = 4;
  memset(m, 0, sizeof(*m));
  switch (elem->type) {
  case EH_T_KQUAD: {
    const EH_KQUAD *q = &elem->p.kquad;
    double lEffective;
    if (q->synch_rad || q->isr || q->edge1_effects || q->edge2_effects || q->radial || q->malignMethod)
      return EH_ERR_UNSUPPORTED;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    m->expand = q->expandHamiltonian;
    m->order[0] = 1;
    if ((lEffective = q->lEffective) <= 0)
      lEffective = q->length;
    else {
      m->lEnd = (q->length - lEffective) / 2;
      m->doEndDrift = 1;
    }
    if (q->bore)
      m->KnL[0] = q->B / q->bore * (particleCharge / (particleMass * EH_C_MKS * Po)) * lEffective * (1 + q->fse);
    else
      m->KnL[0] = q->k1 * lEffective * (1 + q->fse);
    m->drift = lEffective;
    m->tilt = q->tilt;
    m->dx = q->dx;
    m->dy = q->dy;
    m->dz = q->dz;
    m->xkick = q->xkick * q->xKickCalibration;
    m->ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    break;
  }
  case EH_T_KSEXT: {
    const EH_KSEXT *q = &elem->p.ksext;
    if (q->synch_rad || q->isr || q->malignMethod)
      return EH_ERR_UNSUPPORTED;
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    m->expand = q->expandHamiltonian;
    m->order[0] = 2;
    if (q->bore)
      m->KnL[0] = 2 * q->B / sqr(q->bore) * (particleCharge / (particleMass * EH_C_MKS * Po)) * q->length * (1 + q->fse);
    else
      m->KnL[0] = q->k2 * q->length * (1 + q->fse);
    m->drift = q->length;
    m->tilt = q->tilt;
    m->dx = q->dx;
    m->dy = q->dy;
    m->dz = q->dz;
    m->xkick = q->xkick * q->xKickCalibration;
    m->ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    if (q->k1) {
      m->KnL[1] = q->k1 * q->length;
      m->order[1] = 1;
    }
    if (q->j1) {
      m->KnL[2] = q->j1 * q->length;
      m->order[2] = 1;
      m->skew[2] = 1;
    }
    break;
  }
  case EH_T_KOCT: {
    const EH_KOCT *q = &elem->p.koct;
    if (q->synch_rad || q->isr || q->malignMethod)
      return EH_ERR_UNSUPPORTED;
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    m->expand = q->expandHamiltonian;
    m->order[0] = 3;
    if (q->bore)
      m->KnL[0] = 6 * q->B / (q->bore * q->bore * q->bore) * (particleCharge / (particleMass * EH_C_MKS * Po)) * q->length *
                  (1 + q->fse);
    else
      m->KnL[0] = q->k3 * q->length * (1 + q->fse);
    m->drift = q->length;
    m->tilt = q->tilt;
    m->dx = q->dx;
    m->dy = q->dy;
    m->dz = q->dz;
    integ_order = q->integration_order;
    break;
  }
  default:
    return EH_ERR_INVALID;
  }
  if (integ_order != 2 && integ_order != 4 && integ_order != 6)
    return EH_ERR_INVALID;
  if (n_kicks <= 0) {
    if (nSlices <= 0)
      return EH_ERR_INVALID;
  } else {
    if (integ_order > 2) {
      if ((nSlices = (long)ceil(n_kicks / (1.0 * integ_order))) < 1)
        nSlices = 1;
    } else
      nSlices = n_kicks;
  }
  m->nSlices = nSlices;
  m->integ_order = integ_order;
  return EH_OK;
}

/* multipole_tracking2 (multipole.c:693-1117) */
long or_multipole_tracking2(double **particle, long n_part, const eh_element *elem, double Po, double z_start,
                            const EH_MAXAMP *maxamp) {
  mult_setup m;
  mult_ap ap;
  long i_part, i_top;
  double dzLoss = 0;
  int rc = multipole_setup(elem, Po, &m);
  if (rc)
    return rc;
  i_top = n_part - 1;
  setup_mult_ap(&ap, -m.tilt, maxamp);
  offset_for_misalignment(particle, n_part, m.dx, m.dy, m.dz, m.tilt, 1);
  if (m.doEndDrift)
    or_exact_drift(particle, n_part, m.lEnd);
  for (i_part = 0; i_part <= i_top; i_part++) {
    double *coord = particle[i_part];
    if (!integrate_kick_multipole(coord, &m, Po, &ap, &dzLoss)) {
      swap_rows(particle[i_part], particle[i_top], OR_NPROPS);
      particle[i_top][4] = z_start + dzLoss;
      particle[i_top][5] = Po * (1 + particle[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
  }
  /* the exit drift, like the exit misalignment, covers the lost rows too
   * (multipole.c:1098-1104) */
  if (m.doEndDrift)
    or_exact_drift(particle, n_part, m.lEnd);
  if (m.dx || m.dy || m.dz || m.tilt)
    offset_for_misalignment(particle, n_part, m.dx, m.dy, m.dz, m.tilt, 2);
  return i_top + 1;
}

/* --------------------------------------------------------------- CSBEND */

typedef struct {
  double rho0, rho_actual, e1, e2, psi1, psi2, tan_e1, tan_e2, cos_ttilt, sin_ttilt;
  double dxi, dyi, dzi, dxf, dyf, dzf, e1_kick_limit, e2_kick_limit;
  double det[5]; /* the error tilt's exit offsets (computeEtiltCentroidOffset) */
  int edge1, edge2;
  int order1; /* expansion order + 1 of the field sums; 0: pure dipole */
  double Fx[EH_FIELD_MAX1][EH_FIELD_MAX1], Fy[EH_FIELD_MAX1][EH_FIELD_MAX1];
} bend_setup;

/* computeCSBENDFields (csbend.c:100-141, the TURBO_MODE >= 5 form), normal
 * multipoles: for each power i of x, the odd powers j of y for Fx and the
 * even ones for Fy, i + j below the expansion order + 1 */
static void bend_fields(const bend_setup *b, double x, double y, double *Fx, double *Fy) {
  double yp[EH_FIELD_MAX1], sumFx = 0, sumFy = 0, xt = 1;
  long i, j;
  if (!b->order1) {
    *Fx = 0;
    *Fy = 1;
    return;
  }
  yp[0] = 1;
  for (i = 1; i < b->order1; i++)
    yp[i] = yp[i - 1] * y;
  for (i = 0; i < b->order1; i++) {
    for (j = 1; j < b->order1 - i; j += 2)
      sumFx += b->Fx[i][j] * xt * yp[j];
    for (j = 0; j < b->order1 - i; j += 2)
      sumFy += b->Fy[i][j] * xt * yp[j];
    xt *= x;
  }
  *Fx = sumFx;
  *Fy = sumFy;
}

/* set-up part of track_through_csbend with edge order 1 (csbend.c:785-1090):
 * the dipole with normal multipoles K1..K4 and XKICK */
static int csbend_setup(const EH_CSBEND *cb, bend_setup *b) {
  double angle, e1, e2, tilt, etilt, ttilt, fse, h, Kg1, Kg2, fint1, fint2;
  /* H1/H2 act only in the second-order edge map (edge_order > 1, refused below) */
  if (cb->yKick || cb->malignMethod || cb->referenceCorrection || cb->expandHamiltonian || cb->synch_rad ||
      cb->isr || cb->fseCorrection)
    return EH_ERR_UNSUPPORTED;
  if ((cb->edge1_effects != 0 && cb->edge1_effects != 1) || (cb->edge2_effects != 0 && cb->edge2_effects != 1) ||
      cb->edge_order > 1)
    return EH_ERR_UNSUPPORTED;
  if (cb->integration_order != 2 && cb->integration_order != 4 && cb->integration_order != 6)
    return EH_ERR_INVALID;
  if (cb->nSlices < 1)
    return EH_ERR_INVALID;
  if (cb->angle < 0) {
    angle = -cb->angle;
    e1 = -cb->e1;
    e2 = -cb->e2;
    etilt = cb->etilt * cb->etiltSign;
    tilt = cb->tilt + EH_PI;
  } else {
    angle = cb->angle;
    e1 = cb->e1;
    e2 = cb->e2;
    etilt = cb->etilt * cb->etiltSign;
    tilt = cb->tilt;
  }
  b->rho0 = cb->length / angle;
  if (fabs(b->rho0) > 1e6)
    return EH_ERR_UNSUPPORTED;
  {
    /* csbend.c:812-883: b_n = K_n rho0 (signed radius), FSE adjustments,
     * b_0 = XKICK/ANGLE, odd b_n negated for a negative angle */
    double bn[9];
    const double rho_s = cb->length / cb->angle;
    memset(bn, 0, sizeof(bn));
    bn[1] = cb->k1 * rho_s;
    bn[2] = cb->k2 * rho_s;
    bn[3] = cb->k3 * rho_s;
    bn[4] = cb->k4 * rho_s;
    bn[1] *= (1 + cb->fse + cb->fseQuadrupole) / (1 + cb->fse + cb->fseDipole);
    bn[2] *= (1 + cb->fse) / (1 + cb->fse + cb->fseDipole);
    bn[0] = cb->xKick / cb->angle;
    if (cb->angle < 0)
      for (int k = 1; k < 9; k += 2)
        bn[k] *= -1;
    b->order1 = eh_bend_field_setup(bn, 1 / b->rho0, cb->nonlinear, 0, b->Fx, b->Fy);
    if (b->order1 < 0)
      return EH_ERR_UNSUPPORTED;
  }
  fse = cb->fse + cb->fseDipole;
  h = 1 / b->rho0;
  if (fabs(fse + 1) < 1e-12)
    fse = -1 + 1e-12;
  b->rho_actual = 1 / ((1 + fse) * h);
  b->e1_kick_limit = cb->edge1_kick_limit;
  b->e2_kick_limit = cb->edge2_kick_limit;
  if (cb->kick_limit_scaling) {
    b->e1_kick_limit *= b->rho0 / b->rho_actual;
    b->e2_kick_limit *= b->rho0 / b->rho_actual;
  }
  fint1 = cb->fint1 >= 0 ? cb->fint1 : cb->fint;
  fint2 = cb->fint2 >= 0 ? cb->fint2 : cb->fint;
  Kg1 = 2 * cb->hgap * fint1 * (b->rho0 > 0 ? 1 : -1);
  b->psi1 = Kg1 / fabs(b->rho_actual) / cos(e1) * (1 + sqr(sin(e1)));
  Kg2 = 2 * cb->hgap * fint2 * (b->rho0 > 0 ? 1 : -1);
  b->psi2 = Kg2 / fabs(b->rho_actual) / cos(e2) * (1 + sqr(sin(e2)));
  if (cb->length < 0) {
    b->psi1 *= -1;
    b->psi2 *= -1;
  }
  b->e1 = e1;
  b->e2 = e2;
  b->tan_e1 = tan(e1);
  b->tan_e2 = tan(e2);
  ttilt = tilt + etilt;
  if (ttilt == 0) {
    b->cos_ttilt = 1;
    b->sin_ttilt = 0;
  } else if (fabs(fabs(ttilt) - EH_PI) < 1e-12) {
    b->cos_ttilt = -1;
    b->sin_ttilt = 0;
  } else if (fabs(ttilt - EH_PIo2) < 1e-12) {
    b->cos_ttilt = 0;
    b->sin_ttilt = 1;
  } else if (fabs(ttilt + EH_PIo2) < 1e-12) {
    b->cos_ttilt = 0;
    b->sin_ttilt = -1;
  } else {
    b->cos_ttilt = cos(ttilt);
    b->sin_ttilt = sin(ttilt);
  }
  eh_etilt_offsets(b->rho0, angle, etilt, tilt, b->det);
  b->dxi = -cb->dx;
  b->dzi = cb->dz;
  b->dyi = -cb->dy;
  b->dxf = cb->dx * cos(cb->angle) + cb->dz * sin(cb->angle);
  b->dzf = cb->dx * sin(cb->angle) - cb->dz * cos(cb->angle);
  b->dyf = cb->dy;
  b->edge1 = cb->edge1_effects;
  b->edge2 = cb->edge2_effects;
  return EH_OK;
}

/* integrate_csbend_ordn (csbend.c:1434-1693) with the fields of bend_fields,
 * and the rad_coef=0, isrConstant<0 round trip of the kick (csbend.c:1603-1624) */
static int integrate_csbend(double *Qf, const double *Qi, double length, long n, double rho0, double rho_actual,
                            double *dz_lost, const mult_ap *ap, int integration_order, const bend_setup *fb) {
  long i;
  double f, phi, ds, dsh, dist, s;
  double sine, cosi, tang, sin_phi, cos_phi, factor, Fx, Fy;
  const double *driftFrac, *kickFrac;
  int nSubsteps;
  integrator_fracs(integration_order, &driftFrac, &kickFrac, &nSubsteps);
  memcpy(Qf, Qi, sizeof(*Qi) * 6);
  dist = 0;
  s = length / n;
  *dz_lost = 0;
  for (i = 0; i < n; i++) {
    int j;
    if (!check_mult_ap(Qf[0], Qf[2], ap))
      return 0;
    for (j = 0; j < nSubsteps; j++) {
      dsh = s * driftFrac[j];
      f = sqr(1 + Qf[5]) - sqr(Qf[3]);
      if (f <= 0)
        return 0;
      f = sqrt(f);
      if (fabs(Qf[1] / f) > 1)
        return 0;
      sin_phi = Qf[1] / f;
      phi = P_ASIN(sin_phi);
      sine = P_SIN(dsh / rho0 + phi);
      cosi = P_COS(dsh / rho0 + phi);
      if (cosi == 0)
        return 0;
      tang = sine / cosi;
      cos_phi = P_COS(phi);
      Qf[1] = f * sine;
      factor = (rho0 + Qf[0]) * cos_phi / f * (tang - sin_phi / cos_phi);
      Qf[2] += Qf[3] * factor;
      dist += factor * (1 + Qf[5]);
      *dz_lost += dsh;
      f = cos_phi / cosi;
      Qf[0] = rho0 * (f - 1) + f * Qf[0];
      if (kickFrac[j] == 0)
        break;
      ds = s * kickFrac[j];
      bend_fields(fb, Qf[0], Qf[2], &Fx, &Fy);
      Qf[1] += -ds * (1 + Qf[0] / rho0) * Fy / rho_actual;
      Qf[3] += ds * (1 + Qf[0] / rho0) * Fx / rho_actual;
      Qf[1] /= (1 + Qf[5]);
      Qf[3] /= (1 + Qf[5]);
      Qf[1] *= (1 + Qf[5]);
      Qf[3] *= (1 + Qf[5]);
    }
  }
  *dz_lost = n * s;
  if (!check_mult_ap(Qf[0], Qf[2], ap))
    return 0;
  Qf[4] += dist;
  return 1;
}

/* track_through_csbend (csbend.c:665-1409) */
long or_track_through_csbend(double **part, long n_part, const EH_CSBEND *cb, double Po, double z_start,
                             const EH_MAXAMP *maxamp) {
  bend_setup b;
  mult_ap ap;
  long i_part, i_top;
  double Qi[6], Qf[6], dz_lost, p0, p1, beta0, beta1;
  int rc;
  if (cb->angle == 0) {
    or_exact_drift(part, n_part, cb->length);
    return n_part;
  }
  rc = csbend_setup(cb, &b);
  if (rc)
    return rc;
  setup_mult_ap(&ap, -(cb->angle < 0 ? cb->tilt + EH_PI : cb->tilt), maxamp);
  i_top = n_part - 1;
  for (i_part = 0; i_part <= i_top; i_part++) {
    double *coord = part[i_part];
    double x, xp, y, yp, s, dp, dp0, rho, delta_xp;
    int particle_lost;
    coord[4] += b.dzi * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
    coord[0] = coord[0] + b.dxi + b.dzi * coord[1];
    coord[2] = coord[2] + b.dyi + b.dzi * coord[3];
    x = coord[0] * b.cos_ttilt + coord[2] * b.sin_ttilt;
    y = -coord[0] * b.sin_ttilt + coord[2] * b.cos_ttilt;
    xp = coord[1] * b.cos_ttilt + coord[3] * b.sin_ttilt;
    yp = -coord[1] * b.sin_ttilt + coord[3] * b.cos_ttilt;
    s = coord[4];
    dp = dp0 = coord[5];
    if (b.edge1 == 1) {
      rho = (1 + dp) * b.rho_actual;
      delta_xp = b.tan_e1 / rho * x;
      if (b.e1_kick_limit > 0 && fabs(delta_xp) > b.e1_kick_limit)
        delta_xp = (delta_xp > 0 ? 1 : -1) * b.e1_kick_limit;
      xp += delta_xp;
      yp -= P_TAN(b.e1 - b.psi1 / (1 + dp)) / rho * y;
    }
    Qi[0] = x;
    Qi[1] = xp;
    Qi[2] = y;
    Qi[3] = yp;
    Qi[4] = 0;
    Qi[5] = dp;
    {
      /* convertToDipoleCanonicalCoordinates (csbend.c:1411-1419) */
      double f = (1 + Qi[5]) / sqrt(1 + sqr(Qi[1]) + sqr(Qi[3]));
      Qi[1] *= f;
      Qi[3] *= f;
    }
    particle_lost = !integrate_csbend(Qf, Qi, cb->length, cb->nSlices, b.rho0, b.rho_actual, &dz_lost, &ap,
                                      cb->integration_order, &b);
    {
      /* convertFromDipoleCanonicalCoordinates (csbend.c:1421-1429) */
      double f = 1 / sqrt(sqr(1 + Qf[5]) - sqr(Qf[1]) - sqr(Qf[3]));
      Qf[1] *= f;
      Qf[3] *= f;
    }
    if (particle_lost) {
      double X, Y, XP, YP;
      memcpy(coord, Qf, sizeof(double) * 6);
      /* convertFromCSBendCoords (csbend.c:4502-4524) */
      X = coord[0] * b.cos_ttilt - coord[2] * b.sin_ttilt;
      Y = coord[0] * b.sin_ttilt + coord[2] * b.cos_ttilt;
      XP = coord[1] * b.cos_ttilt - coord[3] * b.sin_ttilt;
      YP = coord[1] * b.sin_ttilt + coord[3] * b.cos_ttilt;
      coord[0] = X;
      coord[2] = Y;
      coord[1] = XP;
      coord[3] = YP;
      swap_rows(part[i_part], part[i_top], OR_NPROPS);
      part[i_top][4] = z_start + dz_lost;
      part[i_top][5] = Po * (1 + part[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
    /* isrConstant < 0 is nonzero: velocity rescaling of s (csbend.c:1280-1292) */
    p0 = Po * (1 + dp0);
    beta0 = p0 / sqrt(sqr(p0) + 1);
    p1 = Po * (1 + Qf[5]);
    beta1 = p1 / sqrt(sqr(p1) + 1);
    s = beta1 * s / beta0 + Qf[4];
    x = Qf[0];
    xp = Qf[1];
    y = Qf[2];
    yp = Qf[3];
    dp = Qf[5];
    if (b.edge2 == 1) {
      rho = (1 + dp) * b.rho_actual;
      delta_xp = b.tan_e2 / rho * x;
      if (b.e2_kick_limit > 0 && fabs(delta_xp) > b.e2_kick_limit)
        delta_xp = (delta_xp > 0 ? 1 : -1) * b.e2_kick_limit;
      xp += delta_xp;
      yp -= P_TAN(b.e2 - b.psi2 / (1 + dp)) / rho * y;
    }
    coord[0] = x * b.cos_ttilt - y * b.sin_ttilt + b.det[0];
    coord[2] = x * b.sin_ttilt + y * b.cos_ttilt + b.det[2];
    coord[1] = xp * b.cos_ttilt - yp * b.sin_ttilt + b.det[1];
    coord[3] = xp * b.sin_ttilt + yp * b.cos_ttilt + b.det[3];
    coord[4] = s + b.det[4];
    coord[5] = dp;
    coord[0] += b.dxf + b.dzf * coord[1];
    coord[2] += b.dyf + b.dzf * coord[3];
    coord[4] += b.dzf * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  }
  return i_top + 1;
}

/* ----------------------------------------------------------------- RFCA */

double or_kahan_plus(double oldSum, double b, double *error) {
  /* KahanPlus (kahan.c:27-36); volatile keeps the compensation from being
   * optimized away (the reference builds kahan.c at -O0, Makefile.OAG:632-636) */
  volatile double sum, Y;
  Y = b + *error;
  sum = oldSum + Y;
  *error = Y - (sum - oldSum);
  return sum;
}

static double beta_from_delta(double p, double delta) {
  p *= 1 + delta;
  return (p / sqrt(p * p + 1));
}

/* add_to_particle_energy (simple_rfca.c:987-1007) */
void or_add_to_particle_energy(double *coord, double timeOfFlight, double Po, double dgamma) {
  double gamma, gamma1, PRatio, P, P1, Pz1, Pz;
  P = Po * (1 + coord[5]);
  gamma1 = (gamma = sqrt(P * P + 1)) + dgamma;
  if (gamma1 <= 1)
    gamma1 = 1 + 1e-7;
  P1 = sqrt(gamma1 * gamma1 - 1);
  coord[5] = (P1 - Po) / Po;
  coord[4] = timeOfFlight * EH_C_MKS * P1 / gamma1;
  Pz = P / sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  Pz1 = sqrt(Pz * Pz + gamma1 * gamma1 - gamma * gamma);
  PRatio = Pz / Pz1;
  coord[1] *= PRatio;
  coord[3] *= PRatio;
}

/* findFiducialTime (simple_rfca.c:62-262), TMEAN and LIGHT modes */
static int find_fiducial_time(double **part, long np, double s0, double sOffset, double p0, int mode, double *tFid) {
  if (mode == EH_FID_LIGHT) {
    *tFid = (s0 + sOffset) / EH_C_MKS;
    return EH_OK;
  }
  if (mode == EH_FID_TMEAN) {
    double tsum = 0, error = 0.0;
    long ip, nsum = 0;
    for (ip = 0; ip < np; ip++) {
      tsum = or_kahan_plus(tsum, (part[ip][4] + sOffset) / (EH_C_MKS * beta_from_delta(p0, part[ip][5])), &error);
      nsum++;
    }
    if (nsum <= 0)
      return EH_ERR_FIDUCIAL;
    *tFid = tsum / nsum;
    return EH_OK;
  }
  return EH_ERR_UNSUPPORTED;
}

/* do_match_energy with change_beam = 0 (do_tracking.c:3308-3420, serial
 * build with USE_KAHAN): the central momentum becomes the KahanPlus mean of
 * P_central (1 + delta) over the np particles in row order, and every delta
 * is re-expressed against it, unless the relative change is at most 1e-14 */
void or_do_match_energy(double **coord, long np, double *P_central) {
  double P_average = 0, error = 0, dp, dr;
  if (!np)
    return;
  for (long ip = 0; ip < np; ip++)
    P_average = or_kahan_plus(P_average, (*P_central) * (1 + coord[ip][5]), &error);
  P_average /= np;
  if (fabs(P_average - (*P_central)) / (*P_central) > 1e-14) {
    dp = (*P_central - P_average) / P_average;
    dr = (*P_central) / P_average;
    for (long ip = 0; ip < np; ip++)
      coord[ip][5] = dp + coord[ip][5] * dr;
    *P_central = P_average;
  }
}

/* simple_rf_cavity -> trackRfCavityWithWakes without wakes, forward tracking
 * (simple_rfca.c:264-984): kick method and matrix method */
long or_simple_rf_cavity(or_ctx *c, long ei, double **part, long np, const EH_RFCA *rfca, double *P_central,
                         double zEnd) {
  long ip, nKicks, ik;
  double P, gamma, dgamma = 0.0, phase, length, dtLight, volt, beta_i, t, t0, omega, tau;
  int matrixMethod = 0;
  double timeOffset = 0, To;
  if (rfca->linearize || rfca->lockPhase || rfca->bodyFocusModel || rfca->standingWave || rfca->Q)
    return EH_ERR_UNSUPPORTED;
  if (np <= 0)
    return np;
  length = rfca->length;
  if (rfca->volt == 0) {
    if (rfca->length) {
      for (ip = 0; ip < np; ip++) {
        double *coord = part[ip];
        coord[0] += coord[1] * length;
        coord[2] += coord[3] * length;
        coord[4] += length * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
      }
    }
    return np;
  }
  omega = EH_PIx2 * rfca->freq;
  volt = rfca->volt / (1e6 * particleMassMV * particleRelSign);
  tau = 0;
  nKicks = length ? rfca->nKicks : 1;
  if (nKicks <= 0) {
    matrixMethod = 1;
    nKicks = 1;
  }
  length /= nKicks;
  volt /= nKicks;
  dtLight = length / EH_C_MKS;
  if (!c->fiducial_seen[ei]) {
    if (rfca->tReference != -1)
      t0 = rfca->tReference;
    else {
      int rc = find_fiducial_time(part, np, zEnd - rfca->length, length / 2, *P_central, rfca->fiducialMode, &t0);
      if (rc)
        return rc;
    }
    c->phase_fiducial[ei] = -omega * t0;
    c->fiducial_seen[ei] = 1;
  }
  phase = c->phase_fiducial[ei];
  if (omega) {
    t0 = -c->phase_fiducial[ei] / omega;
    To = EH_PIx2 / omega;
  } else
    t0 = To = 0;
  (void)t0;
  (void)tau;
  phase += rfca->phase * EH_PI / 180.0;

  /* CHANGE_T: whole RF periods in the first row's arrival time
   * (simple_rfca.c:488-497) */
  if (omega && rfca->change_t && np != 0) {
    double *coord = part[0];
    P = *P_central * (1 + coord[5]);
    beta_i = P / (gamma = sqrt(sqr(P) + 1));
    t = coord[4] / (EH_C_MKS * beta_i);
    if (omega != 0 && t > (0.9 * To) && rfca->change_t)
      timeOffset = ((long)(t / To + 0.5)) * To;
  }

  for (ip = 0; ip < np; ip++) {
    part[ip][0] -= rfca->dx;
    part[ip][2] -= rfca->dy;
  }

  if (!matrixMethod) {
    double *inverseF = (double *)calloc(np, sizeof(double));
    for (ik = 0; ik < nKicks; ik++) {
      for (ip = 0; ip < np; ip++) {
        double *coord = part[ip];
        double dc4, gamma1;
        if (coord[5] == -1)
          continue;
        if (length)
          dc4 = length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        else
          dc4 = 0;
        P = *P_central * (1 + coord[5]);
        beta_i = P / (gamma = sqrt(sqr(P) + 1));
        t = (coord[4] + dc4) / (EH_C_MKS * beta_i) - timeOffset;
        if (ik == 0 && timeOffset && rfca->change_t)
          coord[4] = t * EH_C_MKS * beta_i - dc4;
        {
          double phi = omega * (t - 0 - ik * dtLight) + phase;
          dgamma = volt * P_SIN(phi) * 1;
        }
        if (length) {
          if (rfca->end1Focus && ik == 0) {
            inverseF[ip] = dgamma / (2 * gamma * length);
            coord[1] -= coord[0] * inverseF[ip];
            coord[3] -= coord[2] * inverseF[ip];
          }
          coord[0] += coord[1] * length / 2;
          coord[2] += coord[3] * length / 2;
          coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        }
        or_add_to_particle_energy(coord, t, *P_central, dgamma);
        if ((gamma1 = gamma + dgamma) <= 1)
          coord[5] = -1;
        else
          inverseF[ip] = -dgamma / (2 * gamma1 * length);
      }
      if (length) {
        for (ip = 0; ip < np; ip++) {
          double *coord = part[ip];
          coord[0] += coord[1] * length / 2;
          coord[2] += coord[3] * length / 2;
          coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
          if (rfca->end2Focus && (ik == nKicks - 1)) {
            coord[1] -= coord[0] * inverseF[ip];
            coord[3] -= coord[2] * inverseF[ip];
          }
        }
      }
    }
    free(inverseF);
  } else {
    double R11 = 1, R21 = 0, R22, R12, dP, ds1, inverseF;
    for (ik = 0; ik < nKicks; ik++) {
      for (ip = 0; ip < np; ip++) {
        double *coord = part[ip];
        double x, xp;
        P = *P_central * (1 + coord[5]);
        beta_i = P / (gamma = sqrt(sqr(P) + 1));
        ds1 = length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        t = (coord[4] + ds1) / (EH_C_MKS * beta_i) - timeOffset;
        if (timeOffset && rfca->change_t)
          coord[4] = t * EH_C_MKS * beta_i - ds1;
        {
          double sin_phase = P_SIN(omega * (t - 0 - ik * dtLight) + phase);
          dgamma = volt * 1 * sin_phase;
        }
        if (rfca->end1Focus && length && ik == 0) {
          inverseF = dgamma / (2 * gamma * length);
          coord[1] -= coord[0] * inverseF;
          coord[3] -= coord[2] * inverseF;
        }
        dP = sqrt(sqr(gamma + dgamma) - 1) - P;
        R22 = 1 / (1 + dP / P);
        if (fabs(dP / P) > 1e-14)
          R12 = length * (P / dP * P_LOG(1 + dP / P));
        else
          R12 = length;
        coord[4] += ds1;
        x = coord[0];
        xp = coord[1];
        coord[0] = x * R11 + xp * R12;
        coord[1] = x * R21 + xp * R22;
        x = coord[2];
        xp = coord[3];
        coord[2] = x * R11 + xp * R12;
        coord[3] = x * R21 + xp * R22;
        coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        coord[5] = (P + dP - (*P_central)) / (*P_central);
        if ((gamma += dgamma) <= 1)
          coord[5] = -1;
        if (rfca->end2Focus && length && ik == (nKicks - 1)) {
          inverseF = -dgamma / (2 * gamma * length);
          coord[1] -= coord[0] * inverseF;
          coord[3] -= coord[2] * inverseF;
        }
        coord[4] = (P + dP) / gamma * coord[4] / beta_i;
      }
    }
  }
  for (ip = 0; ip < np; ip++) {
    part[ip][0] += rfca->dx;
    part[ip][2] += rfca->dy;
  }
  np = or_remove_invalid_particles(part, np, zEnd, *P_central);
  /* CHANGE_P0 (simple_rfca.c:977-982) */
  if (rfca->change_p0)
    or_do_match_energy(part, np, P_central);
  return np;
}

/* ---------------------------------------------------------- WAKE/ZLONGIT */

/* computeTimeCoordinatesOnly (trwake.c:519-527) */
static void time_coordinates(double *time, double Po, double **part, long np) {
  for (long ip = 0; ip < np; ip++) {
    double P = Po * (part[ip][5] + 1);
    time[ip] = part[ip][4] * sqrt(sqr(P) + 1) / (EH_C_MKS * P);
  }
}

/* find_min_max (mdbmth, not vendored): exact minimum and maximum */
static void find_min_max(double *min, double *max, const double *v, long n) {
  for (long i = 0; i < n; i++) {
    if (v[i] < *min)
      *min = v[i];
    if (v[i] > *max)
      *max = v[i];
  }
}

/* compute_average (mdbmth, not vendored).  Restated as the arithmetic mean
 * with the sum taken exactly in fixed point relative to the largest
 * magnitude (eh_fixedsum.h), rounded once, then divided by n: an
 * order-independent mean (parity unpinned against the toolkit). */
static double compute_average(const double *v, long n) {
  eh_fsum acc;
  int32_t e = -1074;
  for (long i = 0; i < n; i++) {
    int32_t ei = eh_fsum_exponent(v[i]);
    if (ei > e)
      e = ei;
  }
  eh_fsum_init(&acc, e);
  for (long i = 0; i < n; i++)
    eh_fsum_add(&acc, v[i]);
  return eh_fsum_to_double(&acc) / n;
}

static void keep_hist(or_ctx *c, const double *hist, const double *volt, long nb, double tmin, double dt) {
  if (nb > c->hist_cap) {
    free(c->last_hist);
    free(c->last_volt);
    c->last_hist = (double *)malloc(sizeof(double) * nb);
    c->last_volt = (double *)malloc(sizeof(double) * nb);
    c->hist_cap = nb;
  }
  memcpy(c->last_hist, hist, sizeof(double) * nb);
  memcpy(c->last_volt, volt, sizeof(double) * nb);
  c->last_nb = nb;
  c->last_tmin = tmin;
  c->last_dt = dt;
}

/* track_through_wake (wake.c:23-303) for one bunch, no smoothing, causal
 * Green function starting at t = 0 */
/* index_bunch_assignments (lrwake.c:17-233): with idSlotsPerBunch > 0 the
 * particles fall into buckets by the bunch index in column 8, bucket
 * ib - ibMin; bucket[] lists each bucket's rows in row order after the
 * first-row offsets first[b] (nbuckets + 1 entries).  Returns the number of
 * buckets (1 without bunch slots, 0 without particles). */
static long bunch_buckets(long id_slots, double **part, long np, long **first_out, long **rows_out) {
  long ibMin = LONG_MAX, ibMax = LONG_MIN, nb;
  *first_out = NULL;
  *rows_out = NULL;
  if (id_slots <= 0)
    return 1;
  for (long ip = 0; ip < np; ip++) {
    long ib = (long)part[ip][8];
    if (ib < ibMin)
      ibMin = ib;
    if (ib > ibMax)
      ibMax = ib;
  }
  if (ibMin == LONG_MAX || ibMax == LONG_MIN)
    return 0;
  nb = ibMax - ibMin + 1;
  long *first = (long *)calloc(nb + 1, sizeof(long));
  long *rows = (long *)malloc(sizeof(long) * (np > 0 ? np : 1));
  for (long ip = 0; ip < np; ip++)
    first[(long)part[ip][8] - ibMin + 1]++;
  for (long b = 0; b < nb; b++)
    first[b + 1] += first[b];
  {
    long *fill = (long *)malloc(sizeof(long) * nb);
    memcpy(fill, first, sizeof(long) * nb);
    for (long ip = 0; ip < np; ip++)
      rows[fill[(long)part[ip][8] - ibMin]++] = ip;
    free(fill);
  }
  *first_out = first;
  *rows_out = rows;
  return nb;
}

static long wake_bucket(or_ctx *c, double **part, long np, const EH_WAKE *w, double Po, long i_pass, int plain_mean);

/* START_BUNCH/END_BUNCH (wake.c:84-88, zlongit.c:142-146): bucket b takes
 * no wake or impedance */
static int bucket_left_out(long bunched, long start, long end, long b) {
  return bunched && ((start >= 0 && b < start) || (end >= 0 && b > end));
}

/* track_through_wake (wake.c:23-303): one bucket, or with a bunched beam one
 * pass of the binning, convolution and kick per bucket on its own rows
 * (wake.c:71-117, 285-295), the mean time of a bucket a plain sum in row
 * order */
static long track_through_wake_buckets(or_ctx *c, double **part, long np, const EH_WAKE *w, double Po, long i_pass);

/* track_through_wake with the central momentum it may move: CHANGE_P0 runs
 * do_match_energy over the element's rows after every bucket
 * (wake.c:288-289) */
long or_track_through_wake_p(or_ctx *c, double **part, long np, const EH_WAKE *w, double *Po, long i_pass) {
  long r = track_through_wake_buckets(c, part, np, w, *Po, i_pass);
  if (r >= 0 && w->change_p0)
    or_do_match_energy(part, np, Po);
  return r;
}

long or_track_through_wake(or_ctx *c, double **part, long np, const EH_WAKE *w, double Po, long i_pass) {
  return or_track_through_wake_p(c, part, np, w, &Po, i_pass);
}

static long track_through_wake_buckets(or_ctx *c, double **part, long np, const EH_WAKE *w, double Po, long i_pass) {
  long *first, *rows, nbk;
  if (np <= 0 || !(w->bunchedBeamMode && c->id_slots_per_bunch > 0))
    return bucket_left_out(w->bunchedBeamMode, w->startBunch, w->endBunch, 0) ? np
                                                                               : wake_bucket(c, part, np, w, Po, i_pass, 0);
  nbk = bunch_buckets(c->id_slots_per_bunch, part, np, &first, &rows);
  if (nbk == 1) {
    free(first);
    free(rows);
    return bucket_left_out(w->bunchedBeamMode, w->startBunch, w->endBunch, 0) ? np
                                                                               : wake_bucket(c, part, np, w, Po, i_pass, 0);
  }
  for (long b = 0; b < nbk; b++) {
    long n = first[b + 1] - first[b];
    if (bucket_left_out(w->bunchedBeamMode, w->startBunch, w->endBunch, b) || !n)
      continue;
    double **sub = (double **)malloc(sizeof(double *) * n);
    for (long k = 0; k < n; k++)
      sub[k] = part[rows[first[b] + k]];
    long rc = wake_bucket(c, sub, n, w, Po, i_pass, 1);
    free(sub);
    if (rc < 0) {
      free(first);
      free(rows);
      return rc;
    }
  }
  free(first);
  free(rows);
  return np;
}

static long wake_bucket(or_ctx *c, double **part, long np, const EH_WAKE *w, double Po, long i_pass, int plain_mean) {
  double tmin, tmax, dt, factor, rampFactor, wtmin = DBL_MAX, wtmax = -DBL_MAX, tmean = 0;
  long nb, ib, ip, n_binned;
  double *time, *Itime, *Vtime;
  long *pbin;
  if (w->smoothing)
    return EH_ERR_UNSUPPORTED;
  if (w->wakePoints < 2)
    return EH_ERR_INVALID;
  find_min_max(&wtmin, &wtmax, w->t, w->wakePoints);
  if (wtmin >= wtmax)
    return EH_ERR_INVALID;
  if (wtmin != 0)
    return EH_ERR_UNSUPPORTED;
  dt = (wtmax - wtmin) / (w->wakePoints - 1);
  if (np <= 0)
    return np;
  rampFactor = 0;
  if (i_pass >= (w->rampPasses - 1))
    rampFactor = 1;
  else
    rampFactor = (i_pass + 1.0) / w->rampPasses;
  time = (double *)malloc(sizeof(double) * np);
  pbin = (long *)malloc(sizeof(long) * np);
  time_coordinates(time, Po, part, np);
  if (w->n_bins) {
    if (plain_mean) {
      for (ip = 0; ip < np; ip++)
        tmean += time[ip];
      tmean /= np;
    } else
      tmean = compute_average(time, np);
  }
  tmax = -(tmin = DBL_MAX);
  find_min_max(&tmin, &tmax, time, np);
  if ((tmax - tmin) > (w->t[w->wakePoints - 1] - w->t[0]) && !w->allowLongBeam) {
    free(time);
    free(pbin);
    return EH_ERR_LONGBEAM;
  }
  if (w->n_bins) {
    nb = w->n_bins;
    tmin = tmean - dt * nb / 2.0;
    tmax = tmin + dt * (nb - 1);
  } else {
    nb = (long)((tmax - tmin) / dt + 3);
    tmin -= dt;
    tmax += dt;
  }
  if (nb <= 0) {
    free(time);
    free(pbin);
    return np;
  }
  Itime = (double *)calloc(2 * nb, sizeof(double));
  Vtime = (double *)calloc(2 * (nb + 1), sizeof(double));
  /* binTimeDistribution (wake.c:497-517) */
  for (ip = n_binned = 0; ip < np; ip++) {
    pbin[ip] = -1;
    ib = (long)((time[ip] - tmin) / dt + 0.5);
    if (ib < 0)
      continue;
    if (ib > nb - 1)
      continue;
    Itime[ib] += 1;
    pbin[ip] = ib;
    n_binned++;
  }
  /* convolveArrays (wake.c:479-495) with di2 = 0 */
  Vtime[nb] = 0;
  for (ib = 0; ib < nb; ib++) {
    long ib1 = 0, ib2 = ib, di;
    Vtime[ib] = 0;
    if (ib2 >= w->wakePoints) {
      di = ib2 - w->wakePoints + 1;
      ib1 += di;
      ib2 -= di;
    }
    for (; ib1 < nb && ib2 >= 0; ib1++, ib2--)
      Vtime[ib] += Itime[ib1] * w->W[ib2];
  }
  factor = c->macroParticleCharge * particleRelSign * w->factor * rampFactor;
  for (ib = 0; ib < nb; ib++)
    Vtime[ib] *= factor;
  keep_hist(c, Itime, Vtime, nb, tmin, dt);
  /* applyLongitudinalWakeKicks (wake.c:305-329) */
  for (ip = 0; ip < np; ip++) {
    double dgam, dt1;
    if ((ib = pbin[ip]) >= 0 && ib <= nb - 1) {
      if (w->interpolate) {
        dt1 = time[ip] - (tmin + dt * ib);
        if ((dt1 < 0 && ib) || ib == nb - 1) {
          ib--;
          dt1 += dt;
        }
        dgam = (Vtime[ib] + (Vtime[ib + 1] - Vtime[ib]) / dt * dt1) / (1e6 * particleMassMV * particleRelSign);
      } else
        dgam = Vtime[ib] / (1e6 * particleMassMV * particleRelSign);
      if (dgam)
        or_add_to_particle_energy(part[ip], time[ip], Po, -dgam);
    }
  }
  free(Itime);
  free(Vtime);
  free(time);
  free(pbin);
  return np;
}

/* realFFT (SDDS fftpack, not vendored).  Restated: forward transform
 * X_k = (1/n) sum_j x_j exp(-2 pi i jk/n), inverse x_j = sum_k X_k exp(+2 pi i jk/n),
 * packed as [X_0, Re X_1, Im X_1, ..., Re X_{n/2-1}, Im X_{n/2-1}, X_{n/2}]
 * (zlongit.c:379-391).  Algorithm: n/2-point complex radix-2 DIT FFT of the
 * even/odd pairs plus the standard real split; twiddles cos/sin(2 pi k/n)
 * from glibc.  The GPU kernel runs the same butterflies in the same order. */
void or_real_fft(double *data, long n, int inverse) {
  long m = n / 2, k, j, h, bits = 0;
  double *tc = (double *)malloc(sizeof(double) * (m + 1));
  double *ts = (double *)malloc(sizeof(double) * (m + 1));
  double *zr = (double *)malloc(sizeof(double) * m);
  double *zi = (double *)malloc(sizeof(double) * m);
  for (k = 0; k <= m; k++) {
    double a = EH_PIx2 * k / n;
    tc[k] = cos(a);
    ts[k] = sin(a);
  }
  while ((1L << bits) < m)
    bits++;
  if (!inverse) {
    for (j = 0; j < m; j++) {
      zr[j] = data[2 * j];
      zi[j] = data[2 * j + 1];
    }
  } else {
    /* Z_k = (X_k + conj X_{m-k}) + i w^{-k} (X_k - conj X_{m-k}) */
    double x0 = data[0], xm = data[n - 1];
    zr[0] = x0 + xm;
    zi[0] = x0 - xm;
    for (k = 1; k < m; k++) {
      double ar = data[2 * k - 1], ai = data[2 * k];
      double br = data[2 * (m - k) - 1], bi = -data[2 * (m - k)];
      double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
      double c = tc[k], s = ts[k];
      zr[k] = sr - (c * di + s * dr);
      zi[k] = si + (c * dr - s * di);
    }
  }
  /* bit-reversal permutation */
  for (j = 0; j < m; j++) {
    long r = 0;
    for (long b = 0; b < bits; b++)
      if (j & (1L << b))
        r |= 1L << (bits - 1 - b);
    if (r > j) {
      double t = zr[j];
      zr[j] = zr[r];
      zr[r] = t;
      t = zi[j];
      zi[j] = zi[r];
      zi[r] = t;
    }
  }
  /* butterflies; twiddle for stage half-size h, index j: angle 2 pi j/(2h) = table[j*(n/(2h))] */
  for (h = 1; h < m; h <<= 1) {
    long stride = n / (2 * h);
    for (long g = 0; g < m; g += 2 * h) {
      for (j = 0; j < h; j++) {
        double c = tc[j * stride], s = ts[j * stride];
        double br = zr[g + j + h], bi = zi[g + j + h];
        double tr, ti;
        if (!inverse) {
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
    }
  }
  if (!inverse) {
    double scale = 1.0 / n;
    data[0] = (zr[0] + zi[0]) * scale;
    data[n - 1] = (zr[0] - zi[0]) * scale;
    for (k = 1; k < m; k++) {
      double ar = zr[k], ai = zi[k], br = zr[m - k], bi = -zi[m - k];
      double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
      double dr = ar - br, di = ai - bi;
      double or_ = di * 0.5, oi = -dr * 0.5;
      double c = tc[k], s = ts[k];
      data[2 * k - 1] = (er + (c * or_ + s * oi)) * scale;
      data[2 * k] = (ei + (c * oi - s * or_)) * scale;
    }
  } else {
    for (j = 0; j < m; j++) {
      data[2 * j] = zr[j];
      data[2 * j + 1] = zi[j];
    }
  }
  free(tc);
  free(ts);
  free(zr);
  free(zi);
}

/* optimizeBinSettingsForImpedance (ztransverse.c:672-759) */
static int optimize_bins(double timeSpan, double freq, double Q, double *binSize, long *nBins, long maxBins) {
  long n_bins = *nBins, maxBins2;
  double bin_size = *binSize, factor;
  if (maxBins <= 0)
    maxBins2 = (long)pow(2, 20);
  else
    maxBins2 = (long)pow(2, (long)(log(maxBins) / log(2) + 1));
  if (1 / (2 * freq * bin_size) < 10)
    bin_size = 1. / freq / 20;
  if (2 * timeSpan > bin_size * n_bins) {
    n_bins = (long)pow(2, (long)(log(2 * timeSpan * 1.05 / bin_size) / log(2) + 1));
    if (maxBins2 < n_bins)
      return EH_ERR_LONGBEAM;
  }
  if (Q < 1)
    factor = 200 / (n_bins * bin_size * freq / Q);
  else
    factor = 200 / (n_bins * bin_size * freq);
  if (factor > 1) {
    if (n_bins * factor > maxBins2) {
      if ((n_bins * factor) / maxBins2 > 50)
        return EH_ERR_INVALID;
      n_bins = maxBins2;
    } else
      n_bins = (long)pow(2, (long)(log(n_bins * factor) / log(2) + 1));
  }
  *nBins = n_bins;
  *binSize = bin_size;
  return EH_OK;
}

/* track_through_zlongit (zlongit.c:56-557), broad-band resonator, one bunch */
static long zlongit_bucket(or_ctx *c, long ei, double **part, long np, const EH_ZLONGIT *zl, double Po,
                           long i_pass);

/* track_through_zlongit (zlongit.c:56-557), per bucket with a bunched beam
 * (zlongit.c:125-195, 491-497); the bin settings are optimized on the first
 * bucket binned */
long or_track_through_zlongit(or_ctx *c, long ei, double **part, long np, const EH_ZLONGIT *zl, double Po, long i_pass) {
  long *first, *rows, nbk;
  if (np <= 0 || !(zl->bunchedBeamMode && c->id_slots_per_bunch > 0))
    return bucket_left_out(zl->bunchedBeamMode, zl->startBunch, zl->endBunch, 0)
               ? np
               : zlongit_bucket(c, ei, part, np, zl, Po, i_pass);
  nbk = bunch_buckets(c->id_slots_per_bunch, part, np, &first, &rows);
  if (nbk == 1) {
    free(first);
    free(rows);
    return bucket_left_out(zl->bunchedBeamMode, zl->startBunch, zl->endBunch, 0)
               ? np
               : zlongit_bucket(c, ei, part, np, zl, Po, i_pass);
  }
  for (long b = 0; b < nbk; b++) {
    long n = first[b + 1] - first[b];
    if (bucket_left_out(zl->bunchedBeamMode, zl->startBunch, zl->endBunch, b) || !n)
      continue;
    double **sub = (double **)malloc(sizeof(double *) * n);
    for (long k = 0; k < n; k++)
      sub[k] = part[rows[first[b] + k]];
    long rc = zlongit_bucket(c, ei, sub, n, zl, Po, i_pass);
    free(sub);
    if (rc < 0) {
      free(first);
      free(rows);
      return rc;
    }
  }
  free(first);
  free(rows);
  return np;
}

/* checkPointSpacing (zlongit.c:788-803) */
static long check_point_spacing(const double *x, long n, double tolerance) {
  double dx, dx0, range;
  long i;
  if (n < 3)
    return 1;
  if ((range = x[n - 1] - x[0]) <= 0)
    return 0;
  dx0 = (x[1] - x[0]) / range;
  for (i = 1; i < n - 1; i++) {
    dx = (x[i + 1] - x[i]) / range;
    if (fabs(dx - dx0) > tolerance)
      return 0;
  }
  return 1;
}

/* set_up_zlongit from the ZREAL/ZIMAG tables (zlongit.c:634-700); the two
 * tables share the frequency array here.  The toolkit's power_of_2 is not
 * part of the reference: n_spect - 1 is taken as 2^n with n > 1, as the
 * reference's message states. */
static int zlongit_table(const EH_ZLONGIT *zl, double **Z_out, long *n_bins, double *bin_size) {
  long i, n_spect = (long)zl->n_spect;
  const double *Zr = zl->Zreal, *Zi = zl->Zimag;
  double df_spect, *Z;
  if (!zl->f || (!Zr && !Zi))
    return EH_ERR_INVALID; /* "you must either give broad_band=1, or Zreal and/or Zimag" */
  if (n_spect < 2 || !check_point_spacing(zl->f, n_spect, 1e-6))
    return EH_ERR_INVALID;
  df_spect = zl->f[1] - zl->f[0];
  if (Zi && Zi[0])
    return EH_ERR_INVALID; /* non-zero imaginary DC term */
  if (n_spect - 1 < 4 || ((n_spect - 1) & (n_spect - 2)))
    return EH_ERR_INVALID;
  *n_bins = 2 * (n_spect - 1);
  *bin_size = 1.0 / (*n_bins * df_spect);
  Z = (double *)malloc(sizeof(double) * 2 * *n_bins);
  for (i = 0; i < n_spect; i++) {
    const double zr = Zr ? Zr[i] : 0, zi = Zi ? Zi[i] : 0;
    if (i == 0)
      /* DC term (multiply by 2 for consistency with sddsfft) */
      Z[0] = 2 * zr;
    else if (i == n_spect - 1 && *n_bins % 2 == 0)
      /* Nyquist term */
      Z[2 * i - 1] = 2 * zr;
    else {
      Z[2 * i - 1] = zr;
      Z[2 * i] = zi;
    }
  }
  *Z_out = Z;
  return EH_OK;
}

static long zlongit_bucket(or_ctx *c, long ei, double **part, long np, const EH_ZLONGIT *zl, double Po,
                           long i_pass) {
  double tmin, tmax, tmean, dt, factor, rampFactor, Ra, bin_size, df;
  long nb, ib, ip, nfreq, n_bins;
  double *time, *Itime, *Vtime, *Z;
  long *pbin;
  int rc;
  if (zl->area_weight || zl->smoothing || zl->reverseTimeOrder)
    return EH_ERR_UNSUPPORTED;
  if ((i_pass -= zl->startOnPass) < 0 || zl->factor == 0)
    return np;
  rampFactor = 0;
  if (i_pass >= (zl->rampPasses - 1))
    rampFactor = 1;
  else
    rampFactor = (i_pass + 1.0) / zl->rampPasses;
  if (np <= 0)
    return np;
  time = (double *)malloc(sizeof(double) * np);
  pbin = (long *)malloc(sizeof(long) * np);
  time_coordinates(time, Po, part, np);
  tmax = -(tmin = DBL_MAX);
  find_min_max(&tmin, &tmax, time, np);
  tmean = compute_average(time, np);
  /* set_up_zlongit (zlongit.c:559-700), recomputed per call (pure function of the element) */
  if (!zl->broad_band) {
    rc = zlongit_table(zl, &Z, &n_bins, &bin_size);
    if (rc) {
      free(time);
      free(pbin);
      return rc;
    }
    nb = n_bins;
    dt = bin_size;
    nfreq = n_bins / 2 + 1;
    goto have_z;
  }
  if (zl->bin_size <= 0 || (zl->Ra && zl->Rs) || zl->n_bins % 2 != 0 || zl->Zreal || zl->Zimag) {
    free(time);
    free(pbin);
    return EH_ERR_INVALID;
  }
  Ra = zl->Ra ? zl->Ra : 2 * zl->Rs;
  if (!c->zl_initialized[ei]) {
    /* the bin settings are optimized once, on the first call (zlongit.c:583, 603) */
    bin_size = zl->bin_size;
    n_bins = zl->n_bins;
    rc = optimize_bins(tmax - tmin, zl->freq, zl->Q, &bin_size, &n_bins, zl->max_n_bins);
    if (rc) {
      free(time);
      free(pbin);
      return rc;
    }
    c->zl_bin_size[ei] = bin_size;
    c->zl_n_bins[ei] = n_bins;
    c->zl_initialized[ei] = 1;
  }
  bin_size = c->zl_bin_size[ei];
  n_bins = c->zl_n_bins[ei];
  nb = n_bins;
  dt = bin_size;
  df = 1 / (n_bins * bin_size) / (zl->freq);
  nfreq = n_bins / 2 + 1;
  Z = (double *)malloc(sizeof(double) * nb);
  Z[0] = 0;
  Z[nb - 1] = 0;
  for (long i = 1; i < nfreq - 1; i++) {
    double term = zl->Q * (1.0 / (i * df) - i * df);
    Z[2 * i - 1] = Ra / 2 / (1 + sqr(term));
    Z[2 * i] = Z[2 * i - 1] * term;
  }
have_z:
  if ((tmax - tmin) * 2 > nb * dt && !zl->allowLongBeam) {
    free(time);
    free(pbin);
    free(Z);
    return EH_ERR_LONGBEAM;
  }
  tmin = tmean - dt * nb / 2.0;
  Itime = (double *)calloc(2 * nb, sizeof(double));
  Vtime = (double *)calloc(2 * (nb + 1), sizeof(double));
  for (ip = 0; ip < np; ip++) {
    pbin[ip] = -1;
    ib = (long)((time[ip] - tmin) / dt);
    if (ib < 0)
      continue;
    if (ib > nb - 1)
      continue;
    Itime[ib] += 1;
    pbin[ip] = ib;
  }
  memcpy(Vtime, Itime, sizeof(double) * nb);
  or_real_fft(Vtime, nb, 0);
  factor = c->macroParticleCharge * particleRelSign / dt * zl->factor * rampFactor;
  {
    double *Ifreq = (double *)malloc(sizeof(double) * nb);
    memcpy(Ifreq, Vtime, sizeof(double) * nb);
    Vtime[0] = Ifreq[0] * Z[0] * factor;
    if (nb % 2 == 0)
      Vtime[nb - 1] = Ifreq[nb - 1] * Z[nb - 1] * factor;
    for (ib = 1; ib < nfreq - 1; ib++) {
      long iReal = 2 * ib - 1, iImag = iReal + 1;
      Vtime[iReal] = (Ifreq[iReal] * Z[iReal] - Ifreq[iImag] * Z[iImag]) * factor;
      Vtime[iImag] = (Ifreq[iReal] * Z[iImag] + Ifreq[iImag] * Z[iReal]) * factor;
    }
    free(Ifreq);
  }
  or_real_fft(Vtime, nb, 1);
  Vtime[nb] = 0;
  keep_hist(c, Itime, Vtime, nb, tmin, dt);
  for (ip = 0; ip < np; ip++) {
    double dgam, dt1;
    if ((ib = pbin[ip]) >= 0 && ib <= (nb - 1)) {
      if (zl->interpolate && ib > 0) {
        dt1 = time[ip] - (tmin + dt / 2.0 + dt * ib);
        if (dt1 < 0)
          dt1 += dt;
        else
          ib += 1;
        if (ib < nb)
          dgam = (Vtime[ib - 1] + (Vtime[ib] - Vtime[ib - 1]) / dt * dt1) / (1e6 * particleMassMV * particleRelSign);
        else
          continue;
      } else
        dgam = Vtime[ib] / (1e6 * particleMassMV * particleRelSign);
      if (dgam)
        or_add_to_particle_energy(part[ip], time[ip], Po, -dgam);
    }
  }
  free(Itime);
  free(Vtime);
  free(Z);
  free(time);
  free(pbin);
  return np;
}

/* ---------------------------------------------------------- do_tracking */

static int has_length(int type) {
  return type == EH_T_DRIF || type == EH_T_EDRIFT || type == EH_T_KQUAD || type == EH_T_KSEXT || type == EH_T_KOCT ||
         type == EH_T_CSBEND || type == EH_T_RFCA || type == EH_T_MONI || type == EH_T_HMON || type == EH_T_VMON ||
         type == EH_T_HKICK || type == EH_T_VKICK || type == EH_T_RCOL;
}

static double elem_length(const eh_element *e) {
  switch (e->type) {
  case EH_T_DRIF:
    return e->p.drift.length;
  case EH_T_EDRIFT:
    return e->p.edrift.length;
  case EH_T_KQUAD:
    return e->p.kquad.length;
  case EH_T_KSEXT:
    return e->p.ksext.length;
  case EH_T_KOCT:
    return e->p.koct.length;
  case EH_T_CSBEND:
    return e->p.csbend.length;
  case EH_T_RFCA:
    return e->p.rfca.length;
  case EH_T_MONI:
  case EH_T_HMON:
  case EH_T_VMON:
    return e->p.moni.length;
  case EH_T_HKICK:
  case EH_T_VKICK:
    return e->p.corrector.length;
  case EH_T_RCOL:
    return e->p.rcol.length;
  default:
    return 0;
  }
}

/* do_tracking (do_tracking.c:144-3241): pass loop (:488), element loop
 * (:816), aperture after each element unless NO_APERTURE (:2459-2477),
 * NaN/Inf cull and loss record (:2628-2637), effort (:2758-2765). */
long or_do_tracking(or_ctx *c, double **coord, long nOriginal, const eh_element *bl, long n_elems,
                    double *P_central, long n_passes, long passOffset) {
  long nToTrack = nOriginal, nLeft, i_pass;
  double x_max = 0, y_max = 0;
  const EH_MAXAMP *maxamp = NULL;
  if (n_elems > OR_MAX_ELEMS)
    return EH_ERR_INVALID;
  for (long ip = 0; ip < nOriginal; ip++)
    coord[ip][7] = -1;
  c->effort = 0;
  /* check_nan: limit_amplitudes(DBL_MAX) once before the pass loop, at z = 0,
   * losses recorded on pass 0 (do_tracking.c:398, 468-473) */
  nLeft = or_limit_amplitudes(coord, DBL_MAX, DBL_MAX, nToTrack, 0.0, *P_central, 0, 0);
  for (long ip = nLeft; ip < nToTrack; ip++)
    coord[ip][7] = 0;
  nToTrack = nLeft;
  /* the recirculation point (do_tracking.c:459-465, 1382-1387): passes
   * other than pass 0 start at the first RECIRC, at its z */
  long i_recirc = 0;
  double z_recirc = 0;
  for (long ie = 0; ie < n_elems; ie++) {
    if (bl[ie].type == EH_T_RECIRC) {
      i_recirc = ie;
      break;
    }
    if (has_length(bl[ie].type))
      z_recirc += elem_length(&bl[ie]);
    if (ie == n_elems - 1)
      z_recirc = 0; /* no RECIRC */
  }
  for (i_pass = passOffset; i_pass < n_passes + passOffset; i_pass++) {
    double z = i_pass ? z_recirc : 0, last_z = z;
    /* stop_tracking_particle_limit (do_tracking.c:500-520) */
    if (c->particle_limit > 0 && nToTrack < c->particle_limit)
      break;
    if (!nToTrack)
      break;
    for (long ie = i_pass ? i_recirc : 0; ie < n_elems && nToTrack; ie++) {
      const eh_element *e = &bl[ie];
      long rc = 0;
      /* NO_APERTURE elements (track_data.c:3928-3933) */
      int no_aperture = (e->type == EH_T_MARK || e->type == EH_T_CHARGE || e->type == EH_T_MONI ||
                         e->type == EH_T_HMON || e->type == EH_T_VMON || e->type == EH_T_RECIRC);
      last_z = z;
      if (has_length(e->type))
        z += elem_length(e);
      nLeft = nToTrack;
      switch (e->type) {
      case EH_T_MARK:
      case EH_T_RECIRC:
        break;
      case EH_T_CHARGE:
        if (i_pass == passOffset) {
          c->macroParticleCharge = 0;
          if (nOriginal)
            c->macroParticleCharge = e->p.charge.charge / nOriginal;
          c->charge_seen = 1;
        }
        if (e->p.charge.chargePerParticle)
          c->macroParticleCharge = e->p.charge.chargePerParticle;
        break;
      case EH_T_MAXAMP:
        if (e->p.maxamp.openSide < 0 || e->p.maxamp.openSide > 4)
          return EH_ERR_INVALID;
        maxamp = &e->p.maxamp;
        x_max = maxamp->x_max;
        y_max = maxamp->y_max;
        break;
      case EH_T_DRIF:
        or_track_particles_drift(coord, nToTrack, e->p.drift.length, e->p.drift.order ? e->p.drift.order : 2);
        break;
      case EH_T_EDRIFT:
        or_exact_drift(coord, nToTrack, e->p.edrift.length);
        break;
      case EH_T_MONI:
      case EH_T_HMON:
      case EH_T_VMON:
        /* MATRIX_TRACKING with the drift matrix (compute_matrices.c:1317-1328) */
        or_track_particles_drift(coord, nToTrack, e->p.moni.length, e->p.moni.order ? e->p.moni.order : 2);
        break;
      case EH_T_RCOL:
        rc = or_rectangular_collimator(coord, &e->p.rcol, nToTrack, last_z, *P_central);
        break;
      case EH_T_HKICK:
      case EH_T_VKICK:
        /* track_particles with corrector_matrix, which with no strength is
         * the drift matrix (bend_matrix.c:314-317, do_tracking.c:2404-2415) */
        if (e->p.corrector.synchRad || e->p.corrector.isr)
          return EH_ERR_UNSUPPORTED;
        if (e->p.corrector.kick * e->p.corrector.calibration != 0) {
          if (e->p.corrector.length != 0)
            return EH_ERR_UNSUPPORTED;
          or_thin_corrector(coord, nToTrack, e->p.corrector.kick * e->p.corrector.calibration,
                            e->type == EH_T_VKICK ? e->p.corrector.tilt + EH_PIo2 : e->p.corrector.tilt);
          break;
        }
        or_track_particles_drift(coord, nToTrack, e->p.corrector.length,
                                 e->p.corrector.order ? e->p.corrector.order : 2);
        break;
      case EH_T_MALIGN:
        /* do_tracking.c:1818-1822 */
        if (e->p.malign.startPID != -1 || e->p.malign.endPID != -1)
          return EH_ERR_UNSUPPORTED;
        if (e->p.malign.on_pass == -1 || e->p.malign.on_pass == i_pass)
          or_offset_beam(coord, nToTrack, &e->p.malign, *P_central);
        break;
      case EH_T_KQUAD:
      case EH_T_KSEXT:
      case EH_T_KOCT:
        rc = or_multipole_tracking2(coord, nToTrack, e, *P_central, last_z, maxamp);
        break;
      case EH_T_CSBEND:
        rc = or_track_through_csbend(coord, nToTrack, &e->p.csbend, *P_central, last_z, maxamp);
        break;
      case EH_T_RFCA:
        rc = or_simple_rf_cavity(c, ie, coord, nToTrack, &e->p.rfca, P_central, z);
        break;
      case EH_T_WAKE:
        if (!c->charge_seen)
          return EH_ERR_INVALID;
        rc = or_track_through_wake_p(c, coord, nToTrack, &e->p.wake, P_central, i_pass);
        break;
      case EH_T_ZLONGIT:
        if (!c->charge_seen)
          return EH_ERR_INVALID;
        rc = or_track_through_zlongit(c, ie, coord, nToTrack, &e->p.zlongit, *P_central, i_pass);
        break;
      default:
        return EH_ERR_INVALID;
      }
      if (rc < 0)
        return rc;
      if (e->type == EH_T_KQUAD || e->type == EH_T_KSEXT || e->type == EH_T_KOCT || e->type == EH_T_CSBEND ||
          e->type == EH_T_RFCA || e->type == EH_T_RCOL)
        nLeft = rc;
      if (!no_aperture && (x_max || y_max)) {
        nLeft = or_maxamp_check(coord, maxamp, x_max, y_max, nLeft, z, *P_central, e->type == EH_T_DRIF);
        if (nLeft < 0)
          return nLeft;
      }
      nLeft = or_limit_amplitudes(coord, DBL_MAX, DBL_MAX, nLeft, z, *P_central, 0, 0);
      for (long ip = nLeft; ip < nToTrack; ip++)
        coord[ip][7] = i_pass;
      nToTrack = nLeft;
    }
    c->effort += nToTrack;
  }
  return nToTrack;
}

/* computeEtiltCentroidOffset with the rotation by -tilt (test access) */
void or_etilt_offsets(double rho0, double angle, double etilt, double tilt, double *out) {
  eh_etilt_offsets(rho0, angle, etilt, tilt, out);
}

/* last WAKE/ZLONGIT histogram and voltage (test access) */
int64_t or_last_hist(const or_ctx *c, double *hist, double *volt, int64_t cap, double *tmin, double *dt) {
  int64_t nb = c->last_nb;
  if (hist && volt && nb <= cap) {
    memcpy(hist, c->last_hist, sizeof(double) * nb);
    memcpy(volt, c->last_volt, sizeof(double) * nb);
  }
  if (tmin)
    *tmin = c->last_tmin;
  if (dt)
    *dt = c->last_dt;
  return nb;
}

/* compute_average as restated (exact fixed-point sum / n) */
double or_compute_average(const double *v, long n) { return compute_average(v, n); }

/* the fiducial phase of the RFCA at beamline index elem_index, as set by an
 * earlier first pass (rfca->phase_fiducial, fiducial_seen; simple_rfca.c:466-471) */
void or_set_fiducial(or_ctx *c, long elem_index, double phase_fiducial) {
  if (elem_index < 0 || elem_index >= OR_MAX_ELEMS)
    return;
  c->phase_fiducial[elem_index] = phase_fiducial;
  c->fiducial_seen[elem_index] = 1;
}

/* the fiducial phase fixed so far for the RFCA at elem_index (NaN if none) */
double or_get_fiducial(const or_ctx *c, long elem_index) {
  if (elem_index < 0 || elem_index >= OR_MAX_ELEMS || !c->fiducial_seen[elem_index])
    return NAN;
  return c->phase_fiducial[elem_index];
}

void or_set_particle_limit(or_ctx *c, long limit) { c->particle_limit = limit; }

/* charge->idSlotsPerBunch = beam->id_slots_per_bunch (do_tracking.c:1341) */
void or_set_id_slots_per_bunch(or_ctx *c, long n) { c->id_slots_per_bunch = n; }

void or_set_macro_charge(or_ctx *c, double q) {
  c->macroParticleCharge = q;
  c->charge_seen = 1;
}
