/*
 * eh_compile.c - turns a flattened beamline into the kernel's element program.
 *
 * Every per-element constant the reference evaluates in its set-up code is
 * computed here, on the host, with glibc and in the reference's operation
 * order: the kernel then reproduces only the per-particle arithmetic.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "eh_consts.h"
#include "eh_host.h"
#include "eh_bend_etilt.h"
#include "eh_bend_field.h"
#include "emath.h"

static double sqr(double x) { return x * x; }

/* BETA = 2^(1/3); integrator fractions (multipole.c:1119-1149, csbend.c:1456-1480) */
#define EH_BETA 1.25992104989487316477
static const double driftFrac2[2] = {0.5, 0.5};
static const double kickFrac2[2] = {1.0, 0.0};
static const double driftFrac4[4] = {0.5 / (2 - EH_BETA), (1 - EH_BETA) / (2 - EH_BETA) / 2,
                                     (1 - EH_BETA) / (2 - EH_BETA) / 2, 0.5 / (2 - EH_BETA)};
static const double kickFrac4[4] = {1. / (2 - EH_BETA), -EH_BETA / (2 - EH_BETA), 1 / (2 - EH_BETA), 0};
static const double driftFrac6[8] = {0.39225680523878,   0.5100434119184585, -0.47105338540975655,
                                     0.0687531682525181, 0.0687531682525181, -0.47105338540975655,
                                     0.5100434119184585, 0.39225680523878};
static const double kickFrac6[8] = {0.784513610477560,  0.235573213359357, -1.17767998417887,
                                    1.3151863206839063, -1.17767998417887, 0.235573213359357,
                                    0.784513610477560,  0};

static int fracs(int order, const double **drift, const double **kick) {
  switch (order) {
  case 2:
    *drift = driftFrac2;
    *kick = kickFrac2;
    return 2;
  case 4:
    *drift = driftFrac4;
    *kick = kickFrac4;
    return 4;
  case 6:
    *drift = driftFrac6;
    *kick = kickFrac6;
    return 8;
  default:
    return 0;
  }
}

int eh_element_defaults(eh_element *e, int32_t type) {
  memset(e, 0, sizeof(*e));
  e->type = type;
  switch (type) {
  case EH_T_MARK:
  case EH_T_EDRIFT:
  case EH_T_CHARGE:
  case EH_T_RECIRC: /* track_data.c:1018-1020 */
  case EH_T_RCOL:   /* track_data.c:651-659: every parameter defaults to zero */
    break;
  case EH_T_MONI: /* track_data.c:600-647 */
  case EH_T_HMON:
  case EH_T_VMON:
    e->p.moni.weight = e->p.moni.xcalibration = e->p.moni.ycalibration = 1;
    break;
  case EH_T_HKICK: /* track_data.c:544-575 */
  case EH_T_VKICK:
    e->p.corrector.calibration = 1;
    e->p.corrector.steering = 1;
    break;
  case EH_T_MALIGN: /* track_data.c:926-940 */
    e->p.malign.on_pass = e->p.malign.startPID = e->p.malign.endPID = -1;
    break;
  case EH_T_DRIF:
    e->p.drift.order = 0;
    break;
  case EH_T_KQUAD: /* track_data.c:1222-1283 */
    e->p.kquad.xKickCalibration = e->p.kquad.yKickCalibration = 1;
    e->p.kquad.nSlices = 1;
    e->p.kquad.integration_order = 4;
    break;
  case EH_T_KSEXT: /* track_data.c:1120-1160 */
    e->p.ksext.xKickCalibration = e->p.ksext.yKickCalibration = 1;
    e->p.ksext.nSlices = 1;
    e->p.ksext.integration_order = 4;
    break;
  case EH_T_KOCT: /* track_data.c:1164-1187: N_KICKS = N_SLICES = 4 */
    e->p.koct.n_kicks = 4;
    e->p.koct.nSlices = 4;
    e->p.koct.integration_order = 4;
    break;
  case EH_T_CSBEND: /* track_data.c:1893-1980 */
    e->p.csbend.fint = 0.5;
    e->p.csbend.fint1 = e->p.csbend.fint2 = -1;
    e->p.csbend.nSlices = 4;
    e->p.csbend.etiltSign = 1;
    e->p.csbend.nonlinear = 1;
    e->p.csbend.edge1_effects = e->p.csbend.edge2_effects = 1;
    e->p.csbend.edge_order = 1;
    e->p.csbend.integration_order = 4;
    break;
  case EH_T_RFCA: /* track_data.c:576-597 */
    e->p.rfca.freq = 500e6;
    e->p.rfca.tReference = -1;
    e->p.rfca.fiducialMode = EH_FID_TMEAN;
    break;
  case EH_T_WAKE: /* track_data.c:2396-2414 */
    e->p.wake.factor = 1;
    e->p.wake.bunchedBeamMode = 1;
    e->p.wake.startBunch = e->p.wake.endBunch = -1;
    break;
  case EH_T_ZLONGIT: /* track_data.c:2251-2283 (IS_LONG defaults from the integer field) */
    e->p.zlongit.n_bins = 128;
    e->p.zlongit.factor = 1;
    e->p.zlongit.bunchedBeamMode = 1;
    e->p.zlongit.startBunch = e->p.zlongit.endBunch = -1;
    break;
  case EH_T_MAXAMP: /* track_data.c:990-997 */
    e->p.maxamp.exponent = 2;
    break;
  default:
    return EH_ERR_INVALID;
  }
  return EH_OK;
}

int64_t eh_struct_size(int32_t type) {
  switch (type) {
  case 0:
    return (int64_t)sizeof(eh_element);
  case EH_T_DRIF:
    return (int64_t)sizeof(EH_DRIFT);
  case EH_T_EDRIFT:
    return (int64_t)sizeof(EH_EDRIFT);
  case EH_T_KQUAD:
    return (int64_t)sizeof(EH_KQUAD);
  case EH_T_KSEXT:
    return (int64_t)sizeof(EH_KSEXT);
  case EH_T_KOCT:
    return (int64_t)sizeof(EH_KOCT);
  case EH_T_CSBEND:
    return (int64_t)sizeof(EH_CSBEND);
  case EH_T_RFCA:
    return (int64_t)sizeof(EH_RFCA);
  case EH_T_WAKE:
    return (int64_t)sizeof(EH_WAKE);
  case EH_T_ZLONGIT:
    return (int64_t)sizeof(EH_ZLONGIT);
  case EH_T_MAXAMP:
    return (int64_t)sizeof(EH_MAXAMP);
  case EH_T_CHARGE:
    return (int64_t)sizeof(EH_CHARGE);
  case EH_T_MALIGN:
    return (int64_t)sizeof(EH_MALIGN);
  case EH_T_MONI:
  case EH_T_HMON:
  case EH_T_VMON:
    return (int64_t)sizeof(EH_MONI);
  case EH_T_RCOL:
    return (int64_t)sizeof(EH_RCOL);
  case EH_T_HKICK:
  case EH_T_VKICK:
    return (int64_t)sizeof(EH_CORRECTOR);
  default:
    return -1;
  }
}

double eh_element_length(const eh_element *e) {
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

int eh_has_length(int32_t type) {
  return type == EH_T_DRIF || type == EH_T_EDRIFT || type == EH_T_KQUAD || type == EH_T_KSEXT ||
         type == EH_T_KOCT || type == EH_T_CSBEND || type == EH_T_RFCA || type == EH_T_MONI || type == EH_T_HMON ||
         type == EH_T_VMON || type == EH_T_HKICK || type == EH_T_VKICK || type == EH_T_RCOL;
}

/* rotateBeamCoordinatesForMisalignment's choice of cos/sin (tilt_matrices.c:161-176) */
static void rotation_cs(double angle, double *c, double *s) {
  if (fabs(fabs(angle) - EH_PI) < 1e-12) {
    *c = -1;
    *s = 0;
  } else if (fabs(angle - EH_PIo2) < 1e-12) {
    *c = 0;
    *s = 1;
  } else if (fabs(angle + EH_PIo2) < 1e-12) {
    *c = 0;
    *s = -1;
  } else {
    *c = cos(angle);
    *s = sin(angle);
  }
}

/* concat_matrices for two first-order matrices (concat_mat.c:62-97):
 * M2 = M1 M0, sums started at zero and taken over j = 0..5 */
static void concat1(double C2[6], double R2[6][6], const double C1[6], const double R1[6][6],
                    const double C0[6], const double R0[6][6]) {
  for (int i = 0; i < 6; i++) {
    double sum = 0;
    for (int j = 0; j < 6; j++)
      sum += R1[i][j] * C0[j];
    C2[i] = sum + C1[i];
  }
  for (int i = 0; i < 6; i++)
    for (int m = 0; m < 6; m++) {
      double sum = 0;
      for (int j = 0; j < 6; j++)
        sum += R1[i][j] * R0[j][m];
      R2[i][m] = sum;
    }
}

/* corrector_matrix of a zero-length corrector with a kick (bend_matrix.c:
 * 300-330): C[1] = kick and R the identity with R[1][5] = -sin(kick); then
 * tilt_matrices (tilt_matrices.c:18-73), M <- IRot (M Rot), with its exact
 * sine and cosine at +-pi/2 and +-pi.  Out: C[6] then R row by row */
static void thin_corrector_matrix(double kick, double tilt, double *out) {
  double C[6] = {0}, R[6][6] = {{0}};
  for (int i = 0; i < 6; i++)
    R[i][i] = 1;
  C[1] = kick;
  R[1][5] = -sin(kick);
  if (tilt) {
    double s, c, Cr[6] = {0}, Rot[6][6] = {{0}}, IRot[6][6] = {{0}}, Cm[6], Rm[6][6];
    if (fabs(tilt - EH_PIo2) < 1e-12) {
      s = 1;
      c = 0;
    } else if (fabs(tilt + EH_PIo2) < 1e-12) {
      s = -1;
      c = 0;
    } else if (fabs(tilt - EH_PI) < 1e-12 || fabs(tilt + EH_PI) < 1e-12) {
      s = 0;
      c = -1;
    } else {
      s = sin(tilt);
      c = cos(tilt);
    }
    for (int k = 0; k < 2; k++) { /* (x, y) at k = 0, (x', y') at k = 1 */
      IRot[k][k] = Rot[k][k] = c;
      IRot[k][k + 2] = -(Rot[k][k + 2] = s);
      IRot[k + 2][k] = -(Rot[k + 2][k] = -s);
      IRot[k + 2][k + 2] = Rot[k + 2][k + 2] = c;
    }
    IRot[4][4] = IRot[5][5] = Rot[4][4] = Rot[5][5] = 1;
    concat1(Cm, Rm, C, R, Cr, Rot);   /* Mr = M Rot */
    concat1(C, R, Cr, IRot, Cm, Rm);  /* M = IRot Mr */
  }
  for (int i = 0; i < 6; i++) {
    out[EH_D_M1_C + i] = C[i];
    for (int j = 0; j < 6; j++)
      out[EH_D_M1_R + 6 * i + j] = R[i][j];
  }
}

static int rotation_active(double angle) { return angle && !(fabs(fabs(angle) - EH_PIx2) < 1e-12); }

/* x and y exponents of an elliptical MAXAMP (YEXPONENT 0 = EXPONENT) */
static int32_t maxamp_xe(const EH_MAXAMP *m) { return m->exponent; }
static int32_t maxamp_ye(const EH_MAXAMP *m) { return m->yExponent ? m->yExponent : m->exponent; }

/* setupMultApertureData from a MAXAMP (multipole.c:1725-1763): the
 * integrators' checkMultAperture takes the elliptical branch only with both
 * half-axes positive (multipole.c:1807) */
static void set_in_aperture(eh_op *op, double reverseTilt, const EH_MAXAMP *maxamp) {
  op->in_cs0 = cos(reverseTilt);
  op->in_cs1 = sin(reverseTilt);
  if (reverseTilt)
    op->flags |= EH_F_INTILT;
  if (maxamp) {
    int ell = maxamp->elliptical && maxamp->x_max > 0 && maxamp->y_max > 0;
    op->in_xmax = maxamp->x_max;
    op->in_ymax = maxamp->y_max;
    if (op->in_xmax > 0 || op->in_ymax > 0)
      op->flags |= EH_F_INAP;
    op->in_code = ell ? EH_AP_CODE(1, maxamp->openSide, maxamp_xe(maxamp), maxamp_ye(maxamp))
                      : EH_AP_CODE(0, maxamp->openSide, 0, 0);
  }
}

/* MAXAMP parameters the checks can evaluate: an open-side code of
 * determineOpenSideCode (limit_amplitudes.c:1130-1162) and, for an ellipse,
 * even exponents >= 2 (elimit_amplitudes rejects others,
 * limit_amplitudes.c:507-521) */
static int check_maxamp(const eh_element *e, char *msg, int msglen) {
  const EH_MAXAMP *m = &e->p.maxamp;
  if (m->openSide < 0 || m->openSide > 4) {
    snprintf(msg, msglen, "%s: OPEN_SIDE code %d is not one of 0 (none), +x, +y, -x, -y", e->name, m->openSide);
    return EH_ERR_INVALID;
  }
  if (m->elliptical && (m->x_max || m->y_max)) {
    int32_t xe = maxamp_xe(m), ye = maxamp_ye(m);
    if (xe < 2 || xe % 2 || ye < 2 || ye % 2) {
      snprintf(msg, msglen, "%s: exponent=%d is not valid.  Give even integer >=2", e->name, xe < 2 || xe % 2 ? xe : ye);
      return EH_ERR_INVALID;
    }
    if (xe > EH_MAX_AP_EXPONENT || ye > EH_MAX_AP_EXPONENT) {
      snprintf(msg, msglen, "%s: exponents above %d are not supported", e->name, EH_MAX_AP_EXPONENT);
      return EH_ERR_UNSUPPORTED;
    }
  }
  return EH_OK;
}

/* The MAXAMP check do_tracking makes after an element (do_tracking.c:2461-2470),
 * with the early returns of limit_amplitudes (xmax < 0 and ymax < 0: no
 * check at all, limit_amplitudes.c:215-217) and elimit_amplitudes (a
 * non-positive half-axis: the rectangle in the other plane, or nothing,
 * limit_amplitudes.c:523-528) resolved here.  Returns 0 when no check is
 * made, else 1 with *code the EH_AP_* shape (0 = closed rectangle). */
static int post_aperture(const EH_MAXAMP *m, double x_max, double y_max, int32_t *code) {
  *code = 0;
  if (!(x_max || y_max))
    return 0;
  if (m && m->elliptical) {
    if (x_max <= 0 || y_max <= 0) {
      if (!(x_max > 0 || y_max > 0))
        return 0;
      *code = EH_AP_CODE(0, m->openSide, 0, 0);
      return 1;
    }
    *code = EH_AP_CODE(1, m->openSide, maxamp_xe(m), maxamp_ye(m));
    return 1;
  }
  if (x_max < 0 && y_max < 0)
    return 0;
  *code = m ? EH_AP_CODE(0, m->openSide, 0, 0) : 0;
  return 1;
}

/* multipole_tracking2 set-up (multipole.c:765-1017) */
static int compile_multipole(const eh_element *e, double Po, const eh_species *sp, const EH_MAXAMP *maxamp,
                             eh_op *op, char *msg, int msglen) {
  double KnL[3] = {0, 0, 0}, drift = 0, tilt = 0, dx = 0, dy = 0, dz = 0, xkick = 0, ykick = 0;
  long order[3] = {0, 0, 0}, n_kicks = 0, nSlices = 0, integ_order = 4;
  int skew[3] = {0, 0, 0}, expand = 0;
  const double *df, *kf;
  int nsub, k;
  switch (e->type) {
  case EH_T_KQUAD: {
    const EH_KQUAD *q = &e->p.kquad;
    if (q->synch_rad || q->isr || q->edge1_effects || q->edge2_effects || q->radial || q->malignMethod) {
      snprintf(msg, msglen, "%s: KQUAD with SYNCH_RAD/ISR/EDGE*_EFFECTS/RADIAL/MALIGN_METHOD", e->name);
      return EH_ERR_UNSUPPORTED;
    }
    double lEffective;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 1;
    /* LEFFECTIVE: the body takes the effective length, exact drifts the
     * rest (multipole.c:766-777) */
    if ((lEffective = q->lEffective) <= 0)
      lEffective = q->length;
    else {
      op->flags |= EH_F_LEND;
      op->d[EH_D_LEND] = (q->length - lEffective) / 2;
    }
    if (q->bore)
      KnL[0] = q->B / q->bore * (sp->charge / (sp->mass * EH_C_MKS * Po)) * lEffective * (1 + q->fse);
    else
      KnL[0] = q->k1 * lEffective * (1 + q->fse);
    drift = lEffective;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    xkick = q->xkick * q->xKickCalibration;
    ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    break;
  }
  case EH_T_KSEXT: {
    const EH_KSEXT *q = &e->p.ksext;
    if (q->synch_rad || q->isr || q->malignMethod) {
      snprintf(msg, msglen, "%s: KSEXT with SYNCH_RAD/ISR/MALIGN_METHOD", e->name);
      return EH_ERR_UNSUPPORTED;
    }
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    expand = q->expandHamiltonian;
    order[0] = 2;
    if (q->bore)
      KnL[0] = 2 * q->B / sqr(q->bore) * (sp->charge / (sp->mass * EH_C_MKS * Po)) * q->length * (1 + q->fse);
    else
      KnL[0] = q->k2 * q->length * (1 + q->fse);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    xkick = q->xkick * q->xKickCalibration;
    ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    if (q->k1) {
      KnL[1] = q->k1 * q->length;
      order[1] = 1;
    }
    if (q->j1) {
      KnL[2] = q->j1 * q->length;
      order[2] = 1;
      skew[2] = 1;
    }
    break;
  }
  case EH_T_KOCT: {
    const EH_KOCT *q = &e->p.koct;
    if (q->synch_rad || q->isr || q->malignMethod) {
      snprintf(msg, msglen, "%s: KOCT with SYNCH_RAD/ISR/MALIGN_METHOD", e->name);
      return EH_ERR_UNSUPPORTED;
    }
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    expand = q->expandHamiltonian;
    order[0] = 3;
    if (q->bore)
      KnL[0] = 6 * q->B / (q->bore * q->bore * q->bore) * (sp->charge / (sp->mass * EH_C_MKS * Po)) * q->length *
               (1 + q->fse);
    else
      KnL[0] = q->k3 * q->length * (1 + q->fse);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    integ_order = q->integration_order;
    break;
  }
  default:
    return EH_ERR_INVALID;
  }
  if ((nsub = fracs((int)integ_order, &df, &kf)) == 0) {
    snprintf(msg, msglen, "%s: INTEGRATION_ORDER must be 2, 4 or 6", e->name);
    return EH_ERR_INVALID;
  }
  if (n_kicks <= 0) {
    if (nSlices <= 0) {
      snprintf(msg, msglen, "%s: N_KICKS<=0 and N_SLICES<=0", e->name);
      return EH_ERR_INVALID;
    }
  } else {
    if (integ_order > 2) {
      if ((nSlices = (long)ceil(n_kicks / (1.0 * integ_order))) < 1)
        nSlices = 1;
    } else
      nSlices = n_kicks;
  }
  op->op = EH_OP_MULT;
  op->nslices = (int32_t)nSlices;
  op->nsub = nsub;
  if (expand)
    op->flags |= EH_F_EXPAND;
  /* the kernel applies the nonzero KnL terms in iOrder order */
  {
    int slot = 0;
    for (k = 0; k < 3; k++) {
      if (!KnL[k])
        continue;
      if (slot == 0) {
        op->order0 = (int32_t)order[k];
        op->skew0 = skew[k];
      } else if (slot == 1) {
        op->order1 = (int32_t)order[k];
        op->skew1 = skew[k];
      } else {
        op->order2 = (int32_t)order[k];
        op->skew2 = skew[k];
      }
      for (int step = 0; step < nsub; step++)
        op->d[EH_D_KICK + 8 * slot + step] = KnL[k] / nSlices * kf[step];
      slot++;
    }
    op->nkick_orders = slot;
  }
  {
    double driftn = drift / nSlices, xk = xkick / nSlices, yk = ykick / nSlices, tot = 0;
    op->d[EH_D_DRIFT] = driftn;
    for (int step = 0; step < nsub; step++) {
      op->d[EH_D_DSH + step] = driftn * df[step];
      op->d[EH_D_XK + step] = -xk * kf[step];
      op->d[EH_D_YK + step] = -yk * kf[step];
    }
    /* dzLoss of a particle that completed every substep: the running sum in step order */
    for (long ik = 0; ik < nSlices; ik++)
      for (int step = 0; step < nsub; step++) {
        if (driftn)
          tot += op->d[EH_D_DSH + step];
        if (!kf[step])
          break;
      }
    op->d[EH_D_TOTLEN] = tot;
    if (xk)
      op->flags |= EH_F_XKICK;
    if (yk)
      op->flags |= EH_F_YKICK;
  }
  op->d[EH_D_DX] = dx;
  op->d[EH_D_DY] = dy;
  op->d[EH_D_DZ] = dz;
  if (dx || dy || dz)
    op->flags |= EH_F_OFFSET;
  if (tilt && rotation_active(tilt)) {
    op->flags |= EH_F_ROTATE;
    rotation_cs(tilt, &op->d[EH_D_COS_IN], &op->d[EH_D_SIN_IN]);
    rotation_cs(-tilt, &op->d[EH_D_COS_OUT], &op->d[EH_D_SIN_OUT]);
  }
  set_in_aperture(op, -tilt, maxamp);
  return EH_OK;
}

/* track_through_csbend set-up for a pure dipole, edge order 1 (csbend.c:785-1090) */
static int compile_csbend(const eh_element *e, const EH_MAXAMP *maxamp, eh_op *op, char *msg, int msglen) {
  const EH_CSBEND *cb = &e->p.csbend;
  double angle, e1, e2, tilt, etilt, ttilt, fse, h, rho0, rho_actual, Kg1, Kg2, fint1, fint2, psi1, psi2;
  const double *df, *kf;
  int nsub;
  if (cb->angle == 0) {
    op->op = EH_OP_EXACT_DRIFT;
    op->d[0] = cb->length;
    return EH_OK;
  }
  /* H1/H2 (pole-face curvature) enter only the second-order edge map and
   * the K. Hwang fringe model (csbend.c:1150-1166, 1311-1325), which the
   * EDGE_ORDER check below excludes: at EDGE_ORDER <= 1 they change nothing */
  if (cb->yKick || cb->malignMethod || cb->referenceCorrection || cb->expandHamiltonian || cb->synch_rad ||
      cb->isr || cb->fseCorrection) {
    snprintf(msg, msglen, "%s: CSBEND without YKICK, MALIGN_METHOD, radiation or reference correction", e->name);
    return EH_ERR_UNSUPPORTED;
  }
  if ((cb->edge1_effects != 0 && cb->edge1_effects != 1) || (cb->edge2_effects != 0 && cb->edge2_effects != 1) ||
      cb->edge_order > 1) {
    snprintf(msg, msglen, "%s: CSBEND edge effects other than EDGE_ORDER=1, EDGE*_EFFECTS in {0,1}", e->name);
    return EH_ERR_UNSUPPORTED;
  }
  if ((nsub = fracs(cb->integration_order, &df, &kf)) == 0 || cb->nSlices < 1) {
    snprintf(msg, msglen, "%s: CSBEND INTEGRATION_ORDER or N_SLICES invalid", e->name);
    return EH_ERR_INVALID;
  }
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
  rho0 = cb->length / angle;
  if (fabs(rho0) > 1e6) {
    snprintf(msg, msglen, "%s: CSBEND with radius > 1e6", e->name);
    return EH_ERR_UNSUPPORTED;
  }
  {
    /* the multipole content (csbend.c:812-883): b_n = K_n rho0 with the
     * signed radius, the FSE adjustments of b_1 and b_2, b_0 = XKICK/ANGLE,
     * and for a negative angle the odd b_n change sign */
    double b[9], F[EH_D_FIELD_MAX], Fx[EH_FIELD_MAX1][EH_FIELD_MAX1], Fy[EH_FIELD_MAX1][EH_FIELD_MAX1];
    const double rho_s = cb->length / cb->angle;
    int order1;
    memset(b, 0, sizeof(b));
    b[1] = cb->k1 * rho_s;
    b[2] = cb->k2 * rho_s;
    b[3] = cb->k3 * rho_s;
    b[4] = cb->k4 * rho_s;
    b[1] *= (1 + cb->fse + cb->fseQuadrupole) / (1 + cb->fse + cb->fseDipole);
    b[2] *= (1 + cb->fse) / (1 + cb->fse + cb->fseDipole);
    b[0] = cb->xKick / cb->angle;
    if (cb->angle < 0)
      for (int k = 1; k < 9; k += 2)
        b[k] *= -1;
    order1 = eh_bend_field_setup(b, 1 / rho0, cb->nonlinear, 0, Fx, Fy);
    if (order1 < 0 || eh_bend_field_count(order1) > EH_D_FIELD_MAX) {
      snprintf(msg, msglen, "%s: CSBEND field expansion beyond order 6", e->name);
      return EH_ERR_UNSUPPORTED;
    }
    if (order1 > 0) {
      int k = eh_bend_field_flatten(order1, Fx, Fy, F);
      memcpy(&op->d[EH_D_FIELD], F, sizeof(double) * k);
      op->nfield = order1;
      op->flags |= EH_F_FIELD;
    }
  }
  fse = cb->fse + cb->fseDipole;
  h = 1 / rho0;
  if (fabs(fse + 1) < 1e-12)
    fse = -1 + 1e-12;
  rho_actual = 1 / ((1 + fse) * h);
  {
    double l1 = cb->edge1_kick_limit, l2 = cb->edge2_kick_limit;
    if (cb->kick_limit_scaling) {
      l1 *= rho0 / rho_actual;
      l2 *= rho0 / rho_actual;
    }
    op->d[EH_D_KLIM1] = l1;
    op->d[EH_D_KLIM2] = l2;
    if (l1 > 0)
      op->flags |= EH_F_KLIM1;
    if (l2 > 0)
      op->flags |= EH_F_KLIM2;
  }
  fint1 = cb->fint1 >= 0 ? cb->fint1 : cb->fint;
  fint2 = cb->fint2 >= 0 ? cb->fint2 : cb->fint;
  Kg1 = 2 * cb->hgap * fint1 * (rho0 > 0 ? 1 : -1);
  psi1 = Kg1 / fabs(rho_actual) / cos(e1) * (1 + sqr(sin(e1)));
  Kg2 = 2 * cb->hgap * fint2 * (rho0 > 0 ? 1 : -1);
  psi2 = Kg2 / fabs(rho_actual) / cos(e2) * (1 + sqr(sin(e2)));
  if (cb->length < 0) {
    psi1 *= -1;
    psi2 *= -1;
  }
  ttilt = tilt + etilt;
  if (ttilt == 0) {
    op->d[EH_D_CT] = 1;
    op->d[EH_D_ST] = 0;
  } else
    rotation_cs(ttilt, &op->d[EH_D_CT], &op->d[EH_D_ST]);
  if (etilt) {
    /* the error tilt's exit offsets (computeEtiltCentroidOffset, csbend.c:1075) */
    eh_etilt_offsets(rho0, angle, etilt, tilt, &op->d[EH_D_ET]);
    op->flags |= EH_F_ETILT;
  }
  op->op = EH_OP_CSBEND;
  op->nslices = cb->nSlices;
  op->nsub = nsub;
  op->d[EH_D_RHO0] = rho0;
  op->d[EH_D_RHOA] = rho_actual;
  op->d[EH_D_RHOA_SGN] = rho_actual < 0 ? -1.0 : 1.0;
  op->d[EH_D_RRHO0] = 1.0 / rho0;
  op->d[EH_D_RRHOA] = 1.0 / rho_actual;
  if (em_recip_ok(rho0) && em_recip_ok(rho_actual))
    op->flags |= EH_F_RCP;
  op->d[EH_D_TANE1] = tan(e1);
  op->d[EH_D_TANE2] = tan(e2);
  op->d[EH_D_E1] = e1;
  op->d[EH_D_E2] = e2;
  op->d[EH_D_PSI1] = psi1;
  op->d[EH_D_PSI2] = psi2;
  if (cb->edge1_effects == 1)
    op->flags |= EH_F_EDGE1;
  if (cb->edge2_effects == 1)
    op->flags |= EH_F_EDGE2;
  op->d[EH_D_DXI] = -cb->dx;
  op->d[EH_D_DZI] = cb->dz;
  op->d[EH_D_DYI] = -cb->dy;
  op->d[EH_D_DXF] = cb->dx * cos(cb->angle) + cb->dz * sin(cb->angle);
  op->d[EH_D_DZF] = cb->dx * sin(cb->angle) - cb->dz * cos(cb->angle);
  op->d[EH_D_DYF] = cb->dy;
  {
    double s = cb->length / cb->nSlices;
    op->d[EH_D_SSLICE] = s;
    op->d[EH_D_NS] = cb->nSlices * s;
    for (int j = 0; j < nsub; j++) {
      double dsh = s * df[j];
      op->d[EH_D_BDSH + j] = dsh;
      op->d[EH_D_BDSHR + j] = dsh / rho0;
      op->d[EH_D_BDS + j] = s * kf[j];
      op->d[EH_D_BKF + j] = kf[j];
    }
  }
  /* the branch-free common path of the kernel needs positive radii within
   * the shared-reciprocal range, substep angles that are 0 or moderate and
   * below 2^-5, and kick lengths that are 0 or moderate */
  if (rho0 > 0 && rho_actual > 0 && em_recip_ok(rho0) && em_recip_ok(rho_actual) && !(op->flags & EH_F_FIELD)) {
    int fast = 1;
    for (int j = 0; j < nsub; j++) {
      double a = fabs(op->d[EH_D_BDSHR + j]), k = fabs(op->d[EH_D_BDS + j]);
      if (!(a == 0 || (a >= 0x1.0p-700 && a < 0x1.0p-5)) || !(k == 0 || em_recip_ok(k)))
        fast = 0;
    }
    if (fast)
      op->flags |= EH_F_FAST;
  }
  set_in_aperture(op, -tilt, maxamp);
  return EH_OK;
}

/* trackRfCavityWithWakes set-up without wakes (simple_rfca.c:378-478); the
 * phase is completed by eh_rfca_set_phase once the fiducial is known */
static int compile_rfca(const eh_element *e, const eh_species *sp, eh_op *op, char *msg, int msglen) {
  const EH_RFCA *rf = &e->p.rfca;
  double length = rf->length, volt, omega;
  long nKicks;
  int matrixMethod = 0;
  if (rf->linearize || rf->lockPhase || rf->bodyFocusModel || rf->standingWave || rf->Q) {
    snprintf(msg, msglen, "%s: RFCA with LINEARIZE/LOCK_PHASE/BODY_FOCUS_MODEL/STANDING_WAVE/Q", e->name);
    return EH_ERR_UNSUPPORTED;
  }
  if (rf->fiducialMode != EH_FID_TMEAN && rf->fiducialMode != EH_FID_LIGHT) {
    snprintf(msg, msglen, "%s: RFCA fiducial mode must be tmean or light", e->name);
    return EH_ERR_UNSUPPORTED;
  }
  if (rf->volt == 0) {
    op->op = EH_OP_RFCA_DRIFT;
    op->d[EH_D_RFLEN_TOTAL] = rf->length;
    return EH_OK;
  }
  omega = EH_PIx2 * rf->freq;
  volt = rf->volt / (1e6 * sp->massMV * sp->relSign);
  nKicks = length ? rf->nKicks : 1;
  if (nKicks <= 0) {
    matrixMethod = 1;
    nKicks = 1;
  }
  length /= nKicks;
  volt /= nKicks;
  op->op = EH_OP_RFCA;
  op->nkicks_rf = (int32_t)nKicks;
  if (matrixMethod)
    op->flags |= EH_F_MATRIX;
  if (rf->end1Focus)
    op->flags |= EH_F_END1FOCUS;
  if (rf->end2Focus)
    op->flags |= EH_F_END2FOCUS;
  op->d[EH_D_VOLT] = volt;
  op->d[EH_D_OMEGA] = omega;
  op->d[EH_D_LENGTH] = length;
  op->d[EH_D_DTLIGHT] = length / EH_C_MKS;
  op->d[EH_D_RFDX] = rf->dx;
  op->d[EH_D_RFDY] = rf->dy;
  op->d[EH_D_PHASE] = NAN; /* set by eh_rfca_set_phase */
  /* To = PIx2 / omega (simple_rfca.c:473-478); the offset itself follows the
   * beam's first row at every pass (eh_rfca_time_offset) */
  op->d[EH_D_TO] = omega ? EH_PIx2 / omega : 0;
  op->d[EH_D_TOFF] = 0;
  if (rf->change_t && omega)
    op->flags |= EH_F_CHANGE_T;
  return EH_OK;
}

/* timeOffset of CHANGE_T (simple_rfca.c:488-497) for a beam whose first row
 * arrives with (s, delta): the whole RF periods in its arrival time, once
 * it exceeds 0.9 of a period */
double eh_rfca_time_offset(const eh_op *op, double Po, double s, double dp) {
  const double To = op->d[EH_D_TO];
  double P, beta_i, t;
  if (!(op->flags & EH_F_CHANGE_T))
    return 0;
  P = Po * (1 + dp);
  beta_i = P / sqrt(P * P + 1);
  t = s / (EH_C_MKS * beta_i);
  if (t > (0.9 * To))
    return ((long)(t / To + 0.5)) * To;
  return 0;
}

/* phase = phase_fiducial + PHASE*pi/180 (simple_rfca.c:466-478) */
void eh_rfca_set_phase(eh_op *op, const EH_RFCA *rf, double phase_fiducial) {
  double phase = phase_fiducial;
  phase += rf->phase * EH_PI / 180.0;
  op->d[EH_D_PHASE] = phase;
}

/* t0 from a TMEAN/LIGHT fiducialization (simple_rfca.c:412-471) -> phase_fiducial */
double eh_rfca_phase_fiducial(const EH_RFCA *rf, double t0) { return -(EH_PIx2 * rf->freq) * t0; }

int eh_compile(const eh_element *bl, int64_t n_elems, double Po, const eh_species *sp, eh_program *prog, char *msg,
               int msglen) {
  double z = 0, last_z;
  double x_max = 0, y_max = 0;
  const EH_MAXAMP *maxamp = NULL;
  int64_t i;
  prog->nops = 0;
  prog->n_rfca = 0;
  prog->n_coll = 0;
  prog->n_p0 = 0;
  prog->Po = Po;
  /* the recirculation point: passes after pass 0 start at the first RECIRC
   * (do_tracking.c:459-465), so the elements before it act on pass 0 only */
  int64_t i_recirc = -1;
  for (i = 0; i < n_elems && i_recirc < 0; i++)
    if (bl[i].type == EH_T_RECIRC)
      i_recirc = i;
  for (i = 0; i < n_elems; i++) {
    const eh_element *e = &bl[i];
    eh_op op;
    int rc = EH_OK, keep = 1, no_aperture = 0, rcol_body = 0;
    memset(&op, 0, sizeof(op));
    last_z = z;
    if (eh_has_length(e->type))
      z += eh_element_length(e);
    op.elem = (int32_t)i;
    op.z_start = last_z;
    op.z_end = z;
    op.Po = Po;
    switch (e->type) {
    case EH_T_MARK:
    case EH_T_CHARGE:
    case EH_T_RECIRC:
      keep = 0; /* NO_APERTURE; the NaN cull after them cannot fire (every element ends with it) */
      break;
    case EH_T_HKICK:
    case EH_T_VKICK: {
      /* corrector_matrix with no strength is the drift matrix
       * (bend_matrix.c:314-317); the checks of every element follow, the
       * aperture's loss point not extrapolated (do_tracking.c:2463) */
      const EH_CORRECTOR *k = &e->p.corrector;
      int order = k->order ? k->order : 2;
      double kick = k->kick * k->calibration;
      if (k->synchRad || k->isr) {
        snprintf(msg, msglen, "%s: steering corrector radiation", e->name);
        return EH_ERR_UNSUPPORTED;
      }
      if (kick != 0) {
        /* a zero-length corrector's kick matrix (bend_matrix.c:323-330);
         * VKICK turns it by pi/2 (compute_matrices.c:1341-1346) */
        if (k->length != 0) {
          snprintf(msg, msglen, "%s: steering corrector with a length and a kick", e->name);
          return EH_ERR_UNSUPPORTED;
        }
        op.op = EH_OP_MATRIX1;
        thin_corrector_matrix(kick, e->type == EH_T_VKICK ? k->tilt + EH_PIo2 : k->tilt, op.d);
        break;
      }
      op.op = order >= 2 ? EH_OP_DRIFT2 : EH_OP_DRIFT1;
      op.d[0] = k->length;
      op.d[1] = k->length / 2;
      if (fabs(op.d[0]) >= 0x1.0p-300 && fabs(op.d[0]) <= 0x1.0p+300)
        op.flags |= EH_F_DFAST;
      break;
    }
    case EH_T_RCOL: {
      /* rectangular_collimator (limit_amplitudes.c:32-185): the entrance
       * check, and with a length the drift and exit check as a second
       * record, each compacting its own losses as the two loops do */
      const EH_RCOL *r = &e->p.rcol;
      if (r->openSide < 0 || r->openSide > 4) {
        snprintf(msg, msglen, "%s: RCOL open side code out of range", e->name);
        return EH_ERR_INVALID;
      }
      if (r->x_max <= 0 && r->y_max <= 0) {
        op.op = EH_OP_EXACT_DRIFT; /* no jaws: exactDrift (limit_amplitudes.c:66-69) */
        op.d[0] = r->length;
        break;
      }
      op.op = EH_OP_RCOL_IN;
      op.d[EH_D_RC_LEN] = r->length;
      op.d[EH_D_RC_XMAX] = r->x_max;
      op.d[EH_D_RC_YMAX] = r->y_max;
      op.d[EH_D_RC_DX] = r->dx;
      op.d[EH_D_RC_DY] = r->dy;
      op.order0 = r->invert != 0;
      op.order1 = r->openSide;
      rcol_body = r->length > 0;
      break;
    }
    case EH_T_MALIGN: {
      /* offset_beam (do_tracking.c:3243-3306), zero length, followed by the
       * checks of every element (MALIGN is not NO_APERTURE) */
      const EH_MALIGN *m = &e->p.malign;
      if (m->startPID != -1 || m->endPID != -1) {
        snprintf(msg, msglen, "%s: MALIGN with START_PID/END_PID", e->name);
        return EH_ERR_UNSUPPORTED;
      }
      op.op = EH_OP_MALIGN;
      op.d[EH_D_MA_DXP] = m->dxp;
      op.d[EH_D_MA_DYP] = m->dyp;
      op.d[EH_D_MA_DX] = m->dx;
      op.d[EH_D_MA_DY] = m->dy;
      op.d[EH_D_MA_DZ] = m->dz;
      op.d[EH_D_MA_DT] = m->dt;
      op.d[EH_D_MA_DP] = m->dp;
      op.d[EH_D_MA_DE] = m->de;
      op.nkicks_rf = m->on_pass;
      break;
    }
    case EH_T_MAXAMP:
      rc = check_maxamp(e, msg, msglen);
      if (rc)
        return rc;
      maxamp = &e->p.maxamp;
      x_max = maxamp->x_max;
      y_max = maxamp->y_max;
      op.op = EH_OP_APERTURE;
      break;
    case EH_T_DRIF: {
      int order = e->p.drift.order ? e->p.drift.order : 2;
      op.op = order >= 2 ? EH_OP_DRIFT2 : EH_OP_DRIFT1;
      op.d[0] = e->p.drift.length;
      op.d[1] = e->p.drift.length / 2;
      op.flags |= EH_F_EXTRAP;
      if (fabs(op.d[0]) >= 0x1.0p-300 && fabs(op.d[0]) <= 0x1.0p+300)
        op.flags |= EH_F_DFAST;
      break;
    }
    case EH_T_EDRIFT:
      op.op = EH_OP_EXACT_DRIFT;
      op.d[0] = e->p.edrift.length;
      break;
    case EH_T_MONI:
    case EH_T_HMON:
    case EH_T_VMON: {
      /* the monitor's drift matrix (compute_matrices.c:1317-1328); NO_APERTURE:
       * only the NaN cull follows it */
      int order = e->p.moni.order ? e->p.moni.order : 2;
      op.op = order >= 2 ? EH_OP_DRIFT2 : EH_OP_DRIFT1;
      op.d[0] = e->p.moni.length;
      op.d[1] = e->p.moni.length / 2;
      if (fabs(op.d[0]) >= 0x1.0p-300 && fabs(op.d[0]) <= 0x1.0p+300)
        op.flags |= EH_F_DFAST;
      no_aperture = 1;
      break;
    }
    case EH_T_KQUAD:
    case EH_T_KSEXT:
    case EH_T_KOCT:
      rc = compile_multipole(e, Po, sp, maxamp, &op, msg, msglen);
      break;
    case EH_T_CSBEND:
      rc = compile_csbend(e, maxamp, &op, msg, msglen);
      break;
    case EH_T_RFCA:
      rc = compile_rfca(e, sp, &op, msg, msglen);
      if (rc == EH_OK && op.op == EH_OP_RFCA) {
        if (prog->n_rfca >= EH_MAX_RFCA) {
          snprintf(msg, msglen, "too many RFCA elements");
          return EH_ERR_UNSUPPORTED;
        }
        prog->rfca_op[prog->n_rfca] = (int32_t)prog->nops;
        prog->rfca_elem[prog->n_rfca] = (int32_t)i;
        prog->n_rfca++;
      }
      break;
    case EH_T_WAKE:
    case EH_T_ZLONGIT:
      if (prog->n_coll >= EH_MAX_COLL) {
        snprintf(msg, msglen, "too many WAKE/ZLONGIT elements");
        return EH_ERR_UNSUPPORTED;
      }
      op.op = e->type == EH_T_WAKE ? EH_OP_WAKE : EH_OP_ZLONGIT;
      prog->coll_op[prog->n_coll] = (int32_t)prog->nops;
      prog->coll_elem[prog->n_coll] = (int32_t)i;
      op.nkicks_rf = prog->n_coll; /* collective slot */
      prog->n_coll++;
      break;
    default:
      snprintf(msg, msglen, "element %lld (%s): unsupported type %d", (long long)i, e->name, e->type);
      return EH_ERR_UNSUPPORTED;
    }
    if (rc)
      return rc;
    if (!keep)
      continue;
    if (!no_aperture && post_aperture(maxamp, x_max, y_max, &op.ap_code)) {
      op.flags |= EH_F_AP;
      op.ap_xmax = x_max;
      op.ap_ymax = y_max;
    }
    if (prog->nops + 3 > EH_MAX_OPS) {
      snprintf(msg, msglen, "beamline longer than %d elements", EH_MAX_OPS);
      return EH_ERR_UNSUPPORTED;
    }
    if (i_recirc >= 0 && i < i_recirc) {
      if (op.op == EH_OP_WAKE || op.op == EH_OP_ZLONGIT ||
          (op.op == EH_OP_RFCA && (e->p.rfca.change_p0 || e->p.rfca.change_t))) {
        snprintf(msg, msglen, "%s: WAKE, ZLONGIT or RFCA with CHANGE_P0/CHANGE_T before the RECIRC point", e->name);
        return EH_ERR_UNSUPPORTED;
      }
      op.flags |= EH_F_FIRSTPASS;
    }
    if ((op.op == EH_OP_RFCA && e->p.rfca.change_p0) || (op.op == EH_OP_WAKE && e->p.wake.change_p0)) {
      /* CHANGE_P0: do_match_energy runs inside the cavity routine, after its
       * own removeInvalidParticles, and at the end of track_through_wake,
       * before do_tracking's checks that follow the element
       * (simple_rfca.c:975-982, wake.c:288-289).  Three records: the element
       * without those checks, the host's match step, and the checks alone at
       * the element's end, all carrying its element index. */
      eh_op step, post;
      memset(&step, 0, sizeof(step));
      step.op = EH_OP_MATCH_P0;
      step.elem = op.elem;
      step.z_start = step.z_end = op.z_end;
      step.Po = Po;
      step.flags = EH_F_ALONE;
      post = step;
      post.op = EH_OP_APERTURE;
      post.flags = op.flags & EH_F_AP; /* the MAXAMP check if any, then the NaN cull */
      post.ap_code = op.ap_code;
      post.ap_xmax = op.ap_xmax;
      post.ap_ymax = op.ap_ymax;
      op.flags = (op.flags & ~EH_F_AP) | EH_F_ALONE;
      if (prog->n_p0 >= EH_MAX_RFCA) {
        snprintf(msg, msglen, "too many RFCA/WAKE elements with CHANGE_P0");
        return EH_ERR_UNSUPPORTED;
      }
      prog->ops[prog->nops++] = op;
      prog->p0_op[prog->n_p0++] = (int32_t)prog->nops;
      prog->ops[prog->nops++] = step;
      prog->ops[prog->nops++] = post;
      continue;
    }
    if (rcol_body) {
      /* the entrance check alone, then the body with the element's checks */
      eh_op in = op;
      in.flags = (op.flags & ~EH_F_AP) | EH_F_ALONE;
      prog->ops[prog->nops++] = in;
      op.op = EH_OP_RCOL_BODY;
    }
    prog->ops[prog->nops++] = op;
  }
  return EH_OK;
}
