/*
 * eh_collective.c - host side of the WAKE and ZLONGIT steps: element
 * set-up that the reference does once per element (wake function spacing,
 * set_up_wake at wake.c:340-477; impedance table and bin settings,
 * set_up_zlongit at zlongit.c:559-632 and optimizeBinSettingsForImpedance at
 * ztransverse.c:672-759), the per-pass scale factors, and the
 * element-by-element entry points.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eh_consts.h"
#include "eh_host.h"

ehd_dev *eh_ctx_dev(eh_ctx *c);
const eh_program *eh_ctx_prog(eh_ctx *c);
const eh_element *eh_ctx_elems(eh_ctx *c);
double eh_ctx_macro_charge(eh_ctx *c, int *seen);
char *eh_ctx_msg(eh_ctx *c);
int *eh_ctx_coll_loaded(eh_ctx *c);
int32_t *eh_ctx_zl_nbins(eh_ctx *c);
double *eh_ctx_zl_binsize(eh_ctx *c);
int32_t eh_ctx_row_props(eh_ctx *c);
int eh_ctx_buckets(eh_ctx *c, long *ib_min);
int64_t eh_ctx_bucket_count(const eh_ctx *c, long b);
int eh_ctx_bunch_to_device(eh_ctx *c);
int eh_ctx_bucket_mean(eh_ctx *c, int64_t pass, int32_t code_lim, long b, double Po, double *mean);
const char *ehd_message(const ehd_dev *d);
int eh_upload(eh_ctx *c, double **coord, int64_t n, int32_t nprops);
int eh_upload_rows(eh_ctx *c, double **coord, int64_t n, int32_t nprops, int cull);
int eh_download(eh_ctx *c, double **coord, int64_t n, int32_t nprops, int64_t *nLeft);

static double sqr(double x) { return x * x; }

static int msgf(eh_ctx *c, int code, const char *m) {
  snprintf(eh_ctx_msg(c), 512, "%s", m);
  return code;
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

static double ramp_factor(long i_pass, long rampPasses) {
  if (i_pass >= (rampPasses - 1))
    return 1;
  return (i_pass + 1.0) / rampPasses;
}

static int wake_setup(eh_ctx *c, int slot, const EH_WAKE *w, double *dt, double *span) {
  double tmin = DBL_MAX, tmax = -DBL_MAX;
  if (w->smoothing)
    return msgf(c, EH_ERR_UNSUPPORTED, "WAKE with SMOOTHING");
  if (w->wakePoints < 2 || !w->t || !w->W)
    return msgf(c, EH_ERR_INVALID, "WAKE: too little wake data");
  for (int64_t i = 0; i < w->wakePoints; i++) {
    if (w->t[i] < tmin)
      tmin = w->t[i];
    if (w->t[i] > tmax)
      tmax = w->t[i];
  }
  if (tmin >= tmax)
    return msgf(c, EH_ERR_INVALID, "WAKE: zero or negative time span");
  if (tmin != 0)
    return msgf(c, EH_ERR_UNSUPPORTED, "WAKE: acausal Green functions (t[0] != 0)");
  *dt = (tmax - tmin) / (w->wakePoints - 1);
  *span = w->t[w->wakePoints - 1] - w->t[0];
  if (!eh_ctx_coll_loaded(c)[slot]) {
    int rc = ehd_put_wake(eh_ctx_dev(c), slot, w->W, w->wakePoints);
    if (rc)
      return msgf(c, rc, ehd_message(eh_ctx_dev(c)));
    eh_ctx_coll_loaded(c)[slot] = 1;
  }
  return EH_OK;
}

/* the impedance Z[n_bins] and the FFT twiddles cos/sin(2 pi k / n_bins) (the
 * table the CPU checker builds) to the device slot */
static int zlongit_upload(eh_ctx *c, int slot, const double *Z, long n_bins, double bin_size) {
  ehd_dev *d = eh_ctx_dev(c);
  double *tc = (double *)malloc(sizeof(double) * (n_bins / 2 + 1));
  double *ts = (double *)malloc(sizeof(double) * (n_bins / 2 + 1));
  int rc;
  if (!tc || !ts) {
    free(tc);
    free(ts);
    return EH_ERR_NOMEM;
  }
  for (long k = 0; k <= n_bins / 2; k++) {
    double a = EH_PIx2 * k / n_bins;
    tc[k] = cos(a);
    ts[k] = sin(a);
  }
  rc = ehd_put_zlongit(d, slot, Z, tc, ts, (int32_t)n_bins);
  free(tc);
  free(ts);
  if (rc)
    return msgf(c, rc, ehd_message(d));
  eh_ctx_zl_nbins(c)[slot] = (int32_t)n_bins;
  eh_ctx_zl_binsize(c)[slot] = bin_size;
  eh_ctx_coll_loaded(c)[slot] = 1;
  return EH_OK;
}

/* checkPointSpacing (zlongit.c:788-803): spacings equal to within the
 * tolerance, relative to the whole range */
static int spacing_ok(const double *x, int64_t n, double tolerance) {
  double dx0, range;
  if (n < 3)
    return 1;
  if ((range = x[n - 1] - x[0]) <= 0)
    return 0;
  dx0 = (x[1] - x[0]) / range;
  for (int64_t i = 1; i < n - 1; i++)
    if (fabs((x[i + 1] - x[i]) / range - dx0) > tolerance)
      return 0;
  return 1;
}

/* set_up_zlongit with Z(f) tables (zlongit.c:634-700): the packed spectrum
 * of the real FFT, the DC and Nyquist terms doubled (as sddsfft scales
 * them); Z holds 2 (n_spect - 1) values */
static int zlongit_table(eh_ctx *c, const EH_ZLONGIT *zl, double **Z_out, long *n_bins, double *bin_size) {
  const int64_t n = zl->n_spect;
  double *Z;
  if (!zl->f || (!zl->Zreal && !zl->Zimag))
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: give BROAD_BAND=1, or ZREAL and/or ZIMAG");
  if (n < 2)
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: too little data in the impedance tables");
  if (!spacing_ok(zl->f, n, 1e-6))
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: impedance frequencies not equally spaced");
  if (zl->Zimag && zl->Zimag[0])
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: impedance spectrum has a non-zero imaginary DC term");
  if (n - 1 < 4 || ((n - 1) & (n - 2)))
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: the number of spectrum points must be 2^n+1, n>1");
  if (2 * (n - 1) > (1L << 20))
    return msgf(c, EH_ERR_UNSUPPORTED, "ZLONGIT: more than 2^20 bins");
  *n_bins = (long)(2 * (n - 1));
  *bin_size = 1.0 / (*n_bins * (zl->f[1] - zl->f[0]));
  Z = (double *)malloc(sizeof(double) * *n_bins);
  if (!Z)
    return EH_ERR_NOMEM;
  for (int64_t i = 0; i < n; i++) {
    const double zr = zl->Zreal ? zl->Zreal[i] : 0, zi = zl->Zimag ? zl->Zimag[i] : 0;
    if (i == 0)
      Z[0] = 2 * zr;
    else if (i == n - 1)
      Z[2 * i - 1] = 2 * zr; /* Nyquist: N_BINS is even */
    else {
      Z[2 * i - 1] = zr;
      Z[2 * i] = zi;
    }
  }
  *Z_out = Z;
  return EH_OK;
}

/* set_up_zlongit (zlongit.c:559-700), done once per element: the broad-band
 * resonator on bins the optimiser fits to the bunch, or the tables */
static int zlongit_setup(eh_ctx *c, int slot, const EH_ZLONGIT *zl, double Po, int32_t use_bunch, int32_t target) {
  ehd_dev *d = eh_ctx_dev(c);
  double tmin, tmax, bin_size, df, Ra;
  long n_bins, nfreq;
  int64_t count;
  int rc;
  if (eh_ctx_coll_loaded(c)[slot])
    return EH_OK;
  if (zl->area_weight || zl->smoothing || zl->reverseTimeOrder)
    return msgf(c, EH_ERR_UNSUPPORTED, "ZLONGIT: AREA_WEIGHT, SMOOTHING and REVERSE_TIME_ORDER are not supported");
  if (!zl->broad_band) {
    double *Z = NULL;
    rc = zlongit_table(c, zl, &Z, &n_bins, &bin_size);
    if (rc)
      return rc;
    rc = zlongit_upload(c, slot, Z, n_bins, bin_size);
    free(Z);
    return rc;
  }
  if (zl->Zreal || zl->Zimag)
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: both a broad-band impedance and Z(f) tables");
  if (zl->bin_size <= 0 || (zl->Ra && zl->Rs) || zl->n_bins % 2 != 0)
    return msgf(c, EH_ERR_INVALID, "ZLONGIT: BIN_SIZE, RA/RS or N_BINS invalid");
  rc = ehd_time_span(d, Po, use_bunch, target, &tmin, &tmax, &count);
  if (rc)
    return msgf(c, rc, ehd_message(d));
  Ra = zl->Ra ? zl->Ra : 2 * zl->Rs;
  bin_size = zl->bin_size;
  n_bins = zl->n_bins;
  rc = optimize_bins(tmax - tmin, zl->freq, zl->Q, &bin_size, &n_bins, zl->max_n_bins);
  if (rc)
    return msgf(c, rc, "ZLONGIT: bin settings cannot cover the bunch");
  /* powers of two up to 2^20 bins, the optimiser's default cap: at most 8192
   * take the one-workgroup LDS transform, more the global-memory one */
  if (n_bins > (1L << 20))
    return msgf(c, EH_ERR_UNSUPPORTED, "ZLONGIT: more than 2^20 bins");
  if (n_bins & (n_bins - 1))
    return msgf(c, EH_ERR_UNSUPPORTED, "ZLONGIT: a bin count that is not a power of 2");
  {
    double *Z = (double *)malloc(sizeof(double) * n_bins);
    if (!Z)
      return EH_ERR_NOMEM;
    df = 1 / (n_bins * bin_size) / (zl->freq);
    nfreq = n_bins / 2 + 1;
    Z[0] = 0;
    Z[n_bins - 1] = 0;
    for (long i = 1; i < nfreq - 1; i++) {
      double term = zl->Q * (1.0 / (i * df) - i * df);
      Z[2 * i - 1] = Ra / 2 / (1 + sqr(term));
      Z[2 * i] = Z[2 * i - 1] * term;
    }
    rc = zlongit_upload(c, slot, Z, n_bins, bin_size);
    free(Z);
    return rc;
  }
}

/* START_BUNCH/END_BUNCH (wake.c:84-88, zlongit.c:142-146): bucket b takes no
 * wake or impedance */
static int bucket_left_out(int32_t bunched, int32_t start, int32_t end, long b) {
  return bunched && ((start >= 0 && b < start) || (end >= 0 && b > end));
}

/* One WAKE/ZLONGIT element: the whole beam as one bucket, or with a bunched
 * beam (BUNCHED_BEAM_MODE and id slots per bunch) one step per bucket of the
 * bunch index in column 8, in bucket order (index_bunch_assignments,
 * lrwake.c:17-233; wake.c:84-117; zlongit.c:142-195).  A bucket's step sees
 * only its own particles; its post-element checks are those of the
 * element, particle by particle, so checking each bucket after its kick
 * marks the same losses with the same codes as checking all at the end. */
static int collective_run(eh_ctx *c, int slot, const eh_element *e, const eh_op *op, int op_index, int64_t pass,
                          double macro, int flags, double *hist, double *volt, int64_t *nb, double *tmin) {
  ehd_collective cc;
  ehd_dev *d = eh_ctx_dev(c);
  const int is_wake = e->type == EH_T_WAKE;
  const int32_t bunched = is_wake ? e->p.wake.bunchedBeamMode : e->p.zlongit.bunchedBeamMode;
  const int32_t b_start = is_wake ? e->p.wake.startBunch : e->p.zlongit.startBunch;
  const int32_t b_end = is_wake ? e->p.wake.endBunch : e->p.zlongit.endBunch;
  long ib_min = 0;
  int nbk = 1, rc;
  long i_pass = (long)pass;
  memset(&cc, 0, sizeof(cc));
  cc.kind = is_wake ? EH_OP_WAKE : EH_OP_ZLONGIT;
  cc.mc2 = 1e6 * eh_ctx_species(c)->massMV * eh_ctx_species(c)->relSign;
  cc.op_index = op_index;
  cc.pass = pass;
  cc.Po = op->Po;
  cc.z_end = op->z_end;
  cc.ap_xmax = op->ap_xmax;
  cc.ap_ymax = op->ap_ymax;
  cc.flags = (op->flags & EH_F_AP) | flags;
  /* CHANGE_P0: the checks after the element follow the match step */
  if (op->flags & EH_F_ALONE)
    cc.flags |= EHD_COLL_NOCHECKS;
  cc.ap_code = op->ap_code;
  if (is_wake) {
    const EH_WAKE *w = &e->p.wake;
    double dt, span;
    rc = wake_setup(c, slot, w, &dt, &span);
    if (rc)
      return rc;
    cc.dt = dt;
    cc.wake_span = span;
    cc.interpolate = w->interpolate;
    cc.n_bins_fixed = w->n_bins;
    cc.allow_long_beam = w->allowLongBeam;
    /* factor = macroParticleCharge * particleRelSign * FACTOR * rampFactor (wake.c:258) */
    cc.factor = macro * eh_ctx_species(c)->relSign * w->factor * ramp_factor((long)pass, w->rampPasses);
  } else {
    const EH_ZLONGIT *zl = &e->p.zlongit;
    i_pass = (long)pass - zl->startOnPass;
    if (i_pass < 0 || zl->factor == 0)
      return EH_OK;
    cc.interpolate = zl->interpolate;
    cc.allow_long_beam = zl->allowLongBeam;
  }
  if (bunched) {
    nbk = eh_ctx_buckets(c, &ib_min);
    if (nbk < 0)
      return nbk;
    if (nbk > 1) {
      rc = eh_ctx_bunch_to_device(c);
      if (rc)
        return rc;
    }
  }
  for (long b = 0; b < nbk; b++) {
    if (nbk > 1) {
      if (!eh_ctx_bucket_count(c, ib_min + b))
        continue;
      cc.use_bunch = 1;
      cc.bunch_target = (int32_t)(ib_min + b);
    }
    if (bucket_left_out(bunched, b_start, b_end, b)) {
      /* no step, but the element's checks still apply to the bucket */
      rc = ehd_collective_checks(d, &cc);
      if (rc)
        return msgf(c, rc, ehd_message(d));
      continue;
    }
    if (is_wake) {
      cc.mean_given = 0;
      if (nbk > 1 && cc.n_bins_fixed) {
        /* a bucket's mean time is a plain sum over its rows (wake.c:110-116) */
        rc = eh_ctx_bucket_mean(c, pass, 4 * op_index, ib_min + b, cc.Po, &cc.mean);
        if (rc)
          return rc;
        cc.mean_given = 1;
      }
    } else {
      const EH_ZLONGIT *zl = &e->p.zlongit;
      /* the bin settings follow the first bucket that takes the impedance */
      rc = zlongit_setup(c, slot, zl, op->Po, cc.use_bunch, cc.bunch_target);
      if (rc)
        return rc;
      cc.nb = eh_ctx_zl_nbins(c)[slot];
      cc.dt = eh_ctx_zl_binsize(c)[slot];
      /* factor = macroParticleCharge * particleRelSign / dt * FACTOR * rampFactor (zlongit.c:377) */
      cc.factor = macro * eh_ctx_species(c)->relSign / cc.dt * zl->factor * ramp_factor(i_pass, zl->rampPasses);
    }
    rc = ehd_collective_step(d, slot, &cc, hist, volt, 0, nb, tmin);
    if (rc == EH_ERR_LONGBEAM)
      return msgf(c, rc, "bunch longer than the wake/impedance window (wake.c:134-140, zlongit.c:219-236)");
    if (rc)
      return msgf(c, rc, ehd_message(d));
  }
  return EH_OK;
}

int eh_collective(eh_ctx *c, int slot, int64_t pass, int64_t n_alive) {
  const eh_program *p = eh_ctx_prog(c);
  const eh_element *e = &eh_ctx_elems(c)[p->coll_elem[slot]];
  const eh_op *op = &p->ops[p->coll_op[slot]];
  int seen;
  double macro = eh_ctx_macro_charge(c, &seen);
  (void)n_alive;
  if (!seen)
    return msgf(c, EH_ERR_INVALID, "WAKE/ZLONGIT need a CHARGE element before them");
  return collective_run(c, slot, e, op, p->coll_op[slot], pass, macro, 0, NULL, NULL, NULL, NULL);
}

/* last histogram/voltage of a collective element, for kernel-level tests */
int eh_collective_probe(eh_ctx *c, double **part, int64_t np, const eh_element *e, double Po, int64_t i_pass,
                        double macro, double *hist, double *volt, int64_t *nb, double *tmin) {
  eh_element bl[1];
  eh_program *prog = (eh_program *)eh_ctx_prog(c);
  int rc;
  int64_t nLeft;
  if (e->type == EH_T_WAKE && e->p.wake.change_p0)
    return msgf(c, EH_ERR_UNSUPPORTED, "one WAKE step with CHANGE_P0: use eh_do_tracking");
  bl[0] = *e;
  rc = eh_compile(bl, 1, Po, eh_ctx_species(c), prog, eh_ctx_msg(c), 512);
  if (rc)
    return rc;
  memset(eh_ctx_coll_loaded(c), 0, sizeof(int) * EH_MAX_COLL);
  /* the element routines leave NaN rows to their caller: no initial cull;
   * whole rows, so that a bunched beam's bunch index (column 8) is seen */
  rc = eh_upload_rows(c, part, np, eh_ctx_row_props(c), 0);
  if (rc)
    return rc;
  rc = collective_run(c, 0, e, &prog->ops[0], 0, i_pass, macro, EHD_COLL_NOCHECKS, hist, volt, nb, tmin);
  if (rc)
    return rc;
  {
    double **tmp = (double **)malloc(sizeof(double *) * (np > 0 ? np : 1));
    double *buf = (double *)malloc(sizeof(double) * 8 * (np > 0 ? np : 1));
    if (!tmp || !buf) {
      free(tmp);
      free(buf);
      return EH_ERR_NOMEM;
    }
    for (int64_t i = 0; i < np; i++)
      tmp[i] = buf + 8 * i;
    rc = eh_download(c, tmp, np, 8, &nLeft);
    if (!rc)
      for (int64_t i = 0; i < np; i++)
        memcpy(part[i], tmp[i], sizeof(double) * 7);
    free(tmp);
    free(buf);
  }
  return rc;
}

int64_t eh_track_through_wake(eh_ctx *c, double **part, int64_t np, const EH_WAKE *wake, double Po, int64_t i_pass,
                              double macroParticleCharge) {
  eh_element e;
  int rc;
  if (!c || !wake)
    return EH_ERR_INVALID;
  if (np <= 0)
    return np;
  if (wake->change_p0)
    /* the central momentum it moves has no way back through this signature;
     * eh_do_tracking runs WAKE with CHANGE_P0 */
    return msgf(c, EH_ERR_UNSUPPORTED, "eh_track_through_wake: CHANGE_P0 (use eh_do_tracking)");
  eh_element_defaults(&e, EH_T_WAKE);
  e.p.wake = *wake;
  rc = eh_collective_probe(c, part, np, &e, Po, i_pass, macroParticleCharge, NULL, NULL, NULL, NULL);
  return rc ? rc : np;
}

int64_t eh_track_through_zlongit(eh_ctx *c, double **part, int64_t np, const EH_ZLONGIT *zlongit, double Po,
                                 int64_t i_pass, double macroParticleCharge) {
  eh_element e;
  int rc;
  if (!c || !zlongit)
    return EH_ERR_INVALID;
  if (np <= 0)
    return np;
  eh_element_defaults(&e, EH_T_ZLONGIT);
  e.p.zlongit = *zlongit;
  rc = eh_collective_probe(c, part, np, &e, Po, i_pass, macroParticleCharge, NULL, NULL, NULL, NULL);
  return rc ? rc : np;
}

/* The WAKE step for callers that cannot allocate: track_through_wake's
 * arithmetic (wake.c:23-303) on device-resident particles, with the step's
 * scratch in the library's static per-GPU array.  Nothing here or below
 * allocates, host or device. */
int64_t eh_track_through_wake_noalloc(eh_ctx *c, double *const coord[6], int64_t np, const EH_WAKE *wake,
                                      const double *W_device, double Po, int64_t i_pass,
                                      double macroParticleCharge) {
  ehd_collective cc;
  double tmin = DBL_MAX, tmax = -DBL_MAX;
  int rc;
  if (!c || !wake || !coord || !W_device || np < 0)
    return EH_ERR_INVALID;
  for (int k = 0; k < 6; k++)
    if (!coord[k])
      return EH_ERR_INVALID;
  if (wake->smoothing || wake->change_p0)
    return msgf(c, EH_ERR_UNSUPPORTED, "WAKE with SMOOTHING or CHANGE_P0");
  if (wake->wakePoints < 2 || !wake->t)
    return msgf(c, EH_ERR_INVALID, "WAKE: too little wake data");
  if (wake->wakePoints > ehd_noalloc_bins_max())
    return msgf(c, EH_ERR_INVALID, "no-allocation WAKE step: more wake points than its scratch holds");
  for (int64_t i = 0; i < wake->wakePoints; i++) {
    if (wake->t[i] < tmin)
      tmin = wake->t[i];
    if (wake->t[i] > tmax)
      tmax = wake->t[i];
  }
  if (tmin >= tmax)
    return msgf(c, EH_ERR_INVALID, "WAKE: zero or negative time span");
  if (tmin != 0)
    return msgf(c, EH_ERR_UNSUPPORTED, "WAKE: acausal Green functions (t[0] != 0)");
  if (np == 0)
    return 0;
  /* the particles carry no bunch index here: one bucket, bucket 0 */
  if (bucket_left_out(wake->bunchedBeamMode, wake->startBunch, wake->endBunch, 0))
    return np;
  memset(&cc, 0, sizeof(cc));
  cc.kind = EH_OP_WAKE;
  cc.mc2 = 1e6 * eh_ctx_species(c)->massMV * eh_ctx_species(c)->relSign;
  cc.pass = i_pass;
  cc.Po = Po;
  cc.flags = EHD_COLL_NOCHECKS;
  cc.dt = (tmax - tmin) / (wake->wakePoints - 1);
  cc.wake_span = wake->t[wake->wakePoints - 1] - wake->t[0];
  cc.interpolate = wake->interpolate;
  cc.n_bins_fixed = wake->n_bins;
  cc.allow_long_beam = wake->allowLongBeam;
  cc.factor = macroParticleCharge * eh_ctx_species(c)->relSign * wake->factor * ramp_factor((long)i_pass, wake->rampPasses);
  rc = ehd_wake_step_noalloc(eh_ctx_dev(c), &cc, coord, np, W_device, wake->wakePoints);
  if (rc == EH_ERR_LONGBEAM)
    return msgf(c, rc, "bunch longer than the wake window (wake.c:134-140)");
  if (rc)
    return msgf(c, rc, ehd_message(eh_ctx_dev(c)));
  return np;
}
