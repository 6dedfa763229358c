/*
 * eh_batch.c - many lattice variants (error seeds) with their own particle
 * sets tracked in one persistent, graph-captured launch sequence.
 *
 * The reference runs each variant (and, for dynamic-aperture scans, each
 * particle) through its own do_tracking call (aperture_search.c:1080-1082).
 * Here every variant's element program and particles are uploaded together;
 * the persistent kernel pulls (variant, particle run) chunks from a queue, so
 * variants of very different sizes keep all compute units busy.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "eh_host.h"

ehd_dev *eh_ctx_dev(eh_ctx *c);
char *eh_ctx_msg(eh_ctx *c);
const char *ehd_message(const ehd_dev *d);

/* rows tried for a TMEAN fiducial before giving up */
#define EH_FID_TRIES 1024

struct eh_batch {
  eh_ctx *ctx;
  eh_program *prog; /* scratch for compiling one variant */
  eh_op *ops;
  int64_t nops, ops_cap;
  eh_vdesc *var;
  int32_t nvar, var_cap;
  double *soa[6];
  int32_t *lp, *lc;
  double *ids;
  int64_t n, cap;
  int uploaded;
  int64_t scheduled; /* particles in the current work queue */
  eh_batch_tuning tune;
};

static int bmsg(eh_batch *b, int code, const char *m) {
  snprintf(eh_ctx_msg(b->ctx), 512, "%s", m);
  return code;
}

int eh_batch_create(eh_ctx *ctx, eh_batch **out) {
  eh_batch *b;
  if (!ctx || !out)
    return EH_ERR_INVALID;
  b = (eh_batch *)calloc(1, sizeof(eh_batch));
  if (!b)
    return EH_ERR_NOMEM;
  b->prog = (eh_program *)calloc(1, sizeof(eh_program));
  if (!b->prog) {
    free(b);
    return EH_ERR_NOMEM;
  }
  b->ctx = ctx;
  b->tune.particles_per_lane = 1;
  b->tune.block = 256;
  b->tune.passes_per_launch = 64;
  b->tune.use_graph = 1;
  *out = b;
  return EH_OK;
}

int eh_batch_destroy(eh_batch *b) {
  if (!b)
    return EH_OK;
  for (int k = 0; k < 6; k++)
    free(b->soa[k]);
  free(b->lp);
  free(b->lc);
  free(b->ids);
  free(b->ops);
  free(b->var);
  free(b->prog);
  free(b);
  return EH_OK;
}

static int grow(void **p, int64_t *cap, int64_t need, size_t elem) {
  if (need <= *cap)
    return EH_OK;
  int64_t nc = *cap ? *cap : 1024;
  while (nc < need)
    nc *= 2;
  void *q = realloc(*p, elem * nc);
  if (!q)
    return EH_ERR_NOMEM;
  *p = q;
  *cap = nc;
  return EH_OK;
}

/* TMEAN fiducials of a variant's RF cavities.  elegant fixes a cavity's phase
 * on the first beam that reaches it (simple_rfca.c:179-214, fiducial_seen)
 * and keeps it for the rest of the step; an aperture search tracks its grid
 * points one do_tracking call at a time (aperture_search.c:1080-1082), so
 * the phase is that of the first point tracked, alone.  The batch tracks
 * every row as such a call, so each row is tried alone through the
 * variant's first pass on the context, in row order, until every cavity has
 * a phase; phases[k] gets cavity k's (NaN where no tried row reached it).
 * Uses the context's program and particle state, as the element entry
 * points do. */
static int batch_fiducials(eh_batch *b, const eh_element *elems, int64_t n_elems, double Po, double **coord,
                           int64_t n, int32_t nprops, double *phases) {
  eh_ctx *c = b->ctx;
  const eh_program *p = b->prog;
  int need = 0, rc;
  for (int k = 0; k < p->n_rfca; k++) {
    const EH_RFCA *rf = &elems[p->rfca_elem[k]].p.rfca;
    const int tmean = rf->tReference == -1 && rf->fiducialMode == EH_FID_TMEAN;
    phases[k] = NAN;
    need += tmean;
  }
  for (int64_t r = 0; need && r < n && r < EH_FID_TRIES; r++) {
    eh_stats st;
    double pc = Po;
    rc = eh_load_program(c, elems, n_elems, Po);
    if (!rc)
      rc = eh_upload(c, coord + r, 1, nprops);
    if (!rc)
      rc = eh_track(c, 0, 1, &pc, &st);
    if (rc)
      return rc;
    for (int k = 0; k < p->n_rfca; k++) {
      double ph;
      if (!isnan(phases[k]))
        continue;
      rc = eh_rfca_fiducial(c, p->rfca_elem[k], &ph);
      if (rc)
        return rc;
      if (!isnan(ph)) {
        const EH_RFCA *rf = &elems[p->rfca_elem[k]].p.rfca;
        phases[k] = ph;
        need -= rf->tReference == -1 && rf->fiducialMode == EH_FID_TMEAN;
      }
    }
  }
  return need && n ? bmsg(b, EH_ERR_FIDUCIAL, "batched RFCA (TMEAN): no row tried reaches the cavity on the first pass")
                   : EH_OK;
}

int eh_batch_add_variant(eh_batch *b, const eh_element *elems, int64_t n_elems, double Po, double **coord, int64_t n,
                         int32_t nprops) {
  double phases[EH_MAX_RFCA];
  int rc;
  if (!b || !elems || n_elems <= 0 || (!coord && n > 0) || n < 0 || nprops < 7 || !(Po > 0))
    return EH_ERR_INVALID;
  rc = eh_compile(elems, n_elems, Po, eh_ctx_species(b->ctx), b->prog, eh_ctx_msg(b->ctx), 512);
  if (rc)
    return rc;
  if (b->prog->n_coll)
    return bmsg(b, EH_ERR_UNSUPPORTED, "batched variants cannot contain WAKE/ZLONGIT");
  if (b->prog->n_p0)
    return bmsg(b, EH_ERR_UNSUPPORTED, "batched variants cannot contain RFCA with CHANGE_P0");
  /* the fiducial runs compile into the context's program, not b->prog */
  rc = batch_fiducials(b, elems, n_elems, Po, coord, n, nprops, phases);
  if (rc)
    return rc;
  for (int k = 0; k < b->prog->n_rfca; k++) {
    const EH_RFCA *rf = &elems[b->prog->rfca_elem[k]].p.rfca;
    eh_op *op = &b->prog->ops[b->prog->rfca_op[k]];
    double t0;
    /* CHANGE_T: every row is its own beam, so its own first row */
    if (op->flags & EH_F_CHANGE_T)
      op->flags |= EH_F_CHT_SELF;
    if (!isnan(phases[k])) {
      eh_rfca_set_phase(op, rf, phases[k]);
      continue;
    }
    if (rf->tReference == -1 && rf->fiducialMode == EH_FID_TMEAN)
      continue; /* a variant without particles: nothing reaches the cavity */
    if (rf->tReference != -1)
      t0 = rf->tReference;
    else {
      /* FIDUCIAL=light (simple_rfca.c:429-431) */
      double length = rf->length;
      long nKicks = length ? rf->nKicks : 1;
      if (nKicks <= 0)
        nKicks = 1;
      length /= nKicks;
      t0 = ((op->z_end - rf->length) + length / 2) / 2.99792458e8;
    }
    eh_rfca_set_phase(op, rf, eh_rfca_phase_fiducial(rf, t0));
  }
  {
    int64_t ocap = b->ops_cap, vcap = b->var_cap;
    if (grow((void **)&b->ops, &ocap, b->nops + b->prog->nops, sizeof(eh_op)) ||
        grow((void **)&b->var, &vcap, b->nvar + 1, sizeof(eh_vdesc)))
      return EH_ERR_NOMEM;
    b->ops_cap = ocap;
    b->var_cap = (int32_t)vcap;
  }
  if (b->n + n > b->cap) {
    int64_t nc = b->cap ? b->cap : 4096;
    while (nc < b->n + n)
      nc *= 2;
    for (int k = 0; k < 6; k++) {
      double *q = (double *)realloc(b->soa[k], sizeof(double) * nc);
      if (!q)
        return EH_ERR_NOMEM;
      b->soa[k] = q;
    }
    int32_t *a = (int32_t *)realloc(b->lp, sizeof(int32_t) * nc);
    if (!a)
      return EH_ERR_NOMEM;
    b->lp = a;
    a = (int32_t *)realloc(b->lc, sizeof(int32_t) * nc);
    if (!a)
      return EH_ERR_NOMEM;
    b->lc = a;
    double *q = (double *)realloc(b->ids, sizeof(double) * nc);
    if (!q)
      return EH_ERR_NOMEM;
    b->ids = q;
    b->cap = nc;
  }
  memcpy(b->ops + b->nops, b->prog->ops, sizeof(eh_op) * b->prog->nops);
  b->var[b->nvar].op_offset = (int32_t)b->nops;
  b->var[b->nvar].nops = b->prog->nops;
  b->var[b->nvar].first = b->n;
  b->var[b->nvar].n = n;
  for (int64_t i = 0; i < n; i++) {
    const double *r = coord[i];
    for (int k = 0; k < 6; k++)
      b->soa[k][b->n + i] = r[k];
    b->ids[b->n + i] = r[6];
    b->lp[b->n + i] = -1;
    b->lc[b->n + i] = 0;
    if (isinf(r[0]) || isinf(r[2]) || isnan(r[0]) || isnan(r[2])) {
      b->soa[4][b->n + i] = 0.0;
      b->soa[5][b->n + i] = Po * (1 + r[5]);
      b->lp[b->n + i] = 0;
      b->lc[b->n + i] = -1;
    }
  }
  b->nops += b->prog->nops;
  b->n += n;
  b->uploaded = 0;
  return b->nvar++;
}

/* The work queue: chunks of block x particles-per-lane particles, never
 * straddling variants.  Full: every particle of every variant in order.
 * Alive only (b->lp current): the surviving particles of each variant, listed
 * in an index array the chunks refer to, so the lanes of a wave stay busy
 * after most of a grid has been lost. */
static int build_queue(eh_batch *b, int alive_only) {
  ehd_dev *d = eh_ctx_dev(b->ctx);
  const int64_t csize = (int64_t)b->tune.block * b->tune.particles_per_lane;
  int64_t nchunks = 0, nperm = 0;
  int32_t *perm = NULL;
  eh_chunk *ch;
  int rc;
  if (alive_only) {
    perm = (int32_t *)malloc(sizeof(int32_t) * (b->n > 0 ? b->n : 1));
    if (!perm)
      return EH_ERR_NOMEM;
  }
  ch = (eh_chunk *)malloc(sizeof(eh_chunk) * (b->n / csize + b->nvar + 1));
  if (!ch) {
    free(perm);
    return EH_ERR_NOMEM;
  }
  for (int v = 0; v < b->nvar; v++) {
    const int64_t first = b->var[v].first, n = b->var[v].n;
    int64_t start = alive_only ? nperm : first, count = n;
    if (alive_only) {
      for (int64_t i = first; i < first + n; i++)
        if (b->lp[i] < 0)
          perm[nperm++] = (int32_t)i;
      count = nperm - start;
    }
    for (int64_t f = 0; f < count; f += csize) {
      ch[nchunks].variant = v;
      ch[nchunks].first = start + f;
      ch[nchunks].count = (int32_t)(count - f < csize ? count - f : csize);
      nchunks++;
    }
  }
  rc = alive_only ? ehd_batch_compact(d, perm, nperm, ch, (int32_t)nchunks)
                  : ehd_batch_setup(d, b->var, b->nvar, ch, (int32_t)nchunks);
  free(ch);
  free(perm);
  if (rc)
    return bmsg(b, rc, ehd_message(d));
  b->scheduled = alive_only ? nperm : b->n;
  return EH_OK;
}

static int upload(eh_batch *b) {
  ehd_dev *d = eh_ctx_dev(b->ctx);
  int rc;
  if (b->nops > 0x7fffffff || b->n > 0x7fffffff)
    return bmsg(b, EH_ERR_UNSUPPORTED, "batch programs or particle count too large");
  rc = ehd_alloc_state(d, b->n);
  if (!rc)
    rc = ehd_put_state(d, b->soa[0], b->soa[1], b->soa[2], b->soa[3], b->soa[4], b->soa[5], b->lp, b->lc, b->n);
  if (!rc)
    rc = ehd_put_ops(d, b->ops, (int32_t)b->nops);
  if (rc)
    return bmsg(b, rc, ehd_message(d));
  rc = build_queue(b, 0);
  if (rc)
    return rc;
  b->uploaded = 1;
  return EH_OK;
}

int eh_batch_set_tuning(eh_batch *b, const eh_batch_tuning *t) {
  if (!b || !t || (t->particles_per_lane != 1 && t->particles_per_lane != 2) || t->block < 64 || t->block > 256 ||
      t->block % 64 || t->passes_per_launch < 1 || t->compact < 0 || t->compact > 2)
    return EH_ERR_INVALID;
  if (t->block * t->particles_per_lane != b->tune.block * b->tune.particles_per_lane && b->uploaded) {
    /* the chunk size changed: rebuild the queue from the current device state */
    int rc;
    ehd_dev *d = eh_ctx_dev(b->ctx);
    rc = ehd_get_state(d, b->soa[0], b->soa[1], b->soa[2], b->soa[3], b->soa[4], b->soa[5], b->lp, b->lc, b->n);
    if (rc)
      return bmsg(b, rc, ehd_message(d));
    b->tune = *t;
    return upload(b);
  }
  b->tune = *t;
  return EH_OK;
}

int eh_batch_get_tuning(const eh_batch *b, eh_batch_tuning *t) {
  if (!b || !t)
    return EH_ERR_INVALID;
  *t = b->tune;
  return EH_OK;
}

/* the passes of one track call in the tuned work-queue mode:
 * compact 0: one fixed queue, launches of passes_per_launch (captured in a
 *   graph with use_graph);
 * compact 1: launches of 1, 2, 4 ... passes up to passes_per_launch, so the
 *   queue can shed the many particles a grid loses in its first passes; after
 *   each launch the host rebuilds the queue from the survivors once a quarter
 *   of the scheduled lanes are idle;
 * compact 2: the same schedule with the queue rebuilt on the device after
 *   every launch, the whole sequence captured in one graph with use_graph. */
static int run_passes(eh_batch *b, int64_t pass0, int64_t npasses, double *ms_out, int64_t *nl_out) {
  ehd_dev *d = eh_ctx_dev(b->ctx);
  double ms = 0;
  int64_t nl = 0;
  int rc;
  if (b->tune.compact != 1) {
    rc = ehd_batch_track(d, pass0, (int32_t)npasses, b->tune.passes_per_launch, b->tune.particles_per_lane,
                         b->tune.block, b->tune.use_graph, b->tune.compact == 2, &ms, &nl);
    if (rc)
      return bmsg(b, rc, ehd_message(d));
  } else {
    int64_t done = 0, step = 1;
    while (done < npasses) {
      double ms1 = 0;
      int64_t nl1 = 0, alive = 0;
      int64_t k = step < b->tune.passes_per_launch ? step : b->tune.passes_per_launch;
      if (k > npasses - done)
        k = npasses - done;
      rc = ehd_batch_track(d, pass0 + done, (int32_t)k, (int32_t)k, b->tune.particles_per_lane, b->tune.block, 0, 0,
                           &ms1, &nl1);
      if (rc)
        return bmsg(b, rc, ehd_message(d));
      ms += ms1;
      nl += nl1;
      done += k;
      if (step < b->tune.passes_per_launch)
        step *= 2;
      if (done >= npasses)
        break;
      rc = ehd_get_state(d, NULL, NULL, NULL, NULL, NULL, NULL, b->lp, b->lc, b->n);
      if (rc)
        return bmsg(b, rc, ehd_message(d));
      for (int64_t i = 0; i < b->n; i++)
        alive += b->lp[i] < 0;
      if (alive == 0)
        break;
      if (alive * 4 < b->scheduled * 3) {
        rc = build_queue(b, 1);
        if (rc)
          return rc;
      }
    }
  }
  *ms_out = ms;
  *nl_out = nl;
  return EH_OK;
}

int eh_batch_track(eh_batch *b, int64_t pass0, int64_t npasses, eh_stats *stats) {
  ehd_dev *d;
  double ms = 0;
  int64_t nl = 0;
  int rc;
  if (!b || npasses < 0 || npasses > 0x7fffffff)
    return EH_ERR_INVALID;
  d = eh_ctx_dev(b->ctx);
  if (!b->uploaded) {
    rc = upload(b);
    if (rc)
      return rc;
  }
  ehd_timing_reset(d);
  rc = run_passes(b, pass0, npasses, &ms, &nl);
  if (rc)
    return rc;
  if (stats) {
    int64_t effort = 0, left = 0;
    rc = ehd_get_state(d, NULL, NULL, NULL, NULL, NULL, NULL, b->lp, b->lc, b->n);
    if (rc)
      return bmsg(b, rc, ehd_message(d));
    for (int64_t i = 0; i < b->n; i++) {
      int64_t lp = b->lp[i];
      if (lp >= 0 && lp < pass0)
        continue;
      if (lp < 0) {
        effort += npasses;
        left++;
      } else
        effort += lp - pass0;
    }
    memset(stats, 0, sizeof(*stats));
    stats->effort = effort;
    stats->n_left = left;
    stats->passes_done = npasses;
    stats->kernel_launches = nl;
    stats->track_ms = ms;
  }
  return EH_OK;
}

static double wall_ms(void) {
  struct timespec ts;
  timespec_get(&ts, TIME_UTC);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* The autotuner, on samples of the loaded batch from its current state
 * (restored after every trial):
 *  1. particles per lane x workgroup size, on the fixed queue for
 *     sample_passes passes (kernel time);
 *  2. the work-queue mode and the passes fused per launch: the fixed queue
 *     in a graph, the host-rebuilt queue, and the device-rebuilt queue in a
 *     graph with launches of at most 8, 32 or P passes (P: about a second of
 *     work per launch), over 4 x sample_passes passes, on the host's wall
 *     clock (it includes queue rebuilds, copies and graph instantiation).
 * The fastest of each stage is kept. */
int eh_batch_autotune(eh_batch *b, int64_t sample_passes, eh_batch_tuning *best_out) {
  static const int ppl[2] = {1, 2};
  static const int blocks[3] = {64, 128, 256};
  ehd_dev *d;
  eh_batch_tuning best, t;
  double best_ms = 1e300;
  int rc;
  if (!b || sample_passes < 1)
    return EH_ERR_INVALID;
  d = eh_ctx_dev(b->ctx);
  if (!b->uploaded) {
    rc = upload(b);
    if (rc)
      return rc;
  } else if ((rc = build_queue(b, 0)))
    return rc;
  rc = ehd_state_snapshot(d, 0);
  if (rc)
    return bmsg(b, rc, ehd_message(d));
  best = b->tune;
  for (int a = 0; a < 2; a++)
    for (int c = 0; c < 3; c++) {
      double ms = 0;
      int64_t nl = 0;
      t = b->tune;
      t.particles_per_lane = ppl[a];
      t.block = blocks[c];
      t.passes_per_launch = (int32_t)sample_passes;
      t.compact = 0;
      t.use_graph = 0;
      rc = eh_batch_set_tuning(b, &t);
      if (!rc)
        rc = ehd_state_snapshot(d, 1);
      if (!rc)
        rc = build_queue(b, 0);
      if (!rc)
        rc = ehd_batch_track(d, 0, (int32_t)sample_passes, t.passes_per_launch, t.particles_per_lane, t.block, 0, 0,
                             &ms, &nl);
      if (rc)
        return bmsg(b, rc, ehd_message(d));
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  best.sample_ms = best_ms;
  {
    /* launches of about a second at most */
    double per_pass = best_ms / sample_passes;
    int32_t P = per_pass > 0 ? (int32_t)(1000.0 / per_pass) : 1024;
    const int64_t long_passes = 4 * sample_passes;
    struct {
      int32_t compact, graph, cap;
    } cand[5];
    double best_wall = 1e300;
    eh_batch_tuning pick = best;
    if (P < 1)
      P = 1;
    if (P > 1024)
      P = 1024;
    cand[0].compact = 0, cand[0].graph = 1, cand[0].cap = P;
    cand[1].compact = 1, cand[1].graph = 0, cand[1].cap = P;
    cand[2].compact = 2, cand[2].graph = 1, cand[2].cap = 8 < P ? 8 : P;
    cand[3].compact = 2, cand[3].graph = 1, cand[3].cap = 32 < P ? 32 : P;
    cand[4].compact = 2, cand[4].graph = 1, cand[4].cap = P;
    for (int k = 0; k < 5; k++) {
      double ms = 0, w0;
      int64_t nl = 0;
      t = best;
      t.compact = cand[k].compact;
      t.use_graph = cand[k].graph;
      t.passes_per_launch = cand[k].cap;
      rc = eh_batch_set_tuning(b, &t);
      if (!rc)
        rc = ehd_state_snapshot(d, 1);
      if (!rc)
        rc = build_queue(b, 0);
      if (rc)
        return bmsg(b, rc, ehd_message(d));
      w0 = wall_ms();
      rc = run_passes(b, 0, long_passes, &ms, &nl);
      if (rc)
        return rc;
      w0 = wall_ms() - w0;
      if (w0 < best_wall) {
        best_wall = w0;
        pick = t;
      }
    }
    best = pick;
    best.sample_ms = best_ms;
  }
  rc = eh_batch_set_tuning(b, &best);
  if (!rc)
    rc = ehd_state_snapshot(d, 1);
  if (rc)
    return bmsg(b, rc, ehd_message(d));
  if ((rc = build_queue(b, 0)))
    return rc;
  ehd_timing_reset(d);
  if (best_out)
    *best_out = best;
  return EH_OK;
}

int eh_batch_loss_passes(eh_batch *b, int32_t variant, int32_t *loss_pass, int64_t n) {
  ehd_dev *d;
  int rc;
  if (!b || variant < 0 || variant >= b->nvar || !loss_pass || n != b->var[variant].n)
    return EH_ERR_INVALID;
  d = eh_ctx_dev(b->ctx);
  if (!b->uploaded)
    memcpy(loss_pass, b->lp + b->var[variant].first, sizeof(int32_t) * n);
  else {
    rc = ehd_get_state(d, NULL, NULL, NULL, NULL, NULL, NULL, b->lp, b->lc, b->n);
    if (rc)
      return bmsg(b, rc, ehd_message(d));
    memcpy(loss_pass, b->lp + b->var[variant].first, sizeof(int32_t) * n);
  }
  return EH_OK;
}

typedef struct {
  int64_t pass;
  int32_t code;
  int64_t slot;
} bev;

static int cmp_bev(const void *a, const void *b) {
  const bev *x = (const bev *)a, *y = (const bev *)b;
  if (x->pass != y->pass)
    return x->pass < y->pass ? -1 : 1;
  if (x->code != y->code)
    return x->code < y->code ? -1 : 1;
  return 0;
}

/* one variant's rows in elegant's order (survivors, then lost rows) */
int eh_batch_download(eh_batch *b, int32_t variant, double **coord, int64_t n, int32_t nprops, int64_t *nLeft) {
  ehd_dev *d;
  int rc;
  int64_t first, m = 0, alive;
  int64_t *order, *pos, *lp;
  bev *ev;
  if (!b || variant < 0 || variant >= b->nvar || nprops < 8 || n != b->var[variant].n)
    return EH_ERR_INVALID;
  d = eh_ctx_dev(b->ctx);
  if (b->uploaded) {
    rc = ehd_get_state(d, b->soa[0], b->soa[1], b->soa[2], b->soa[3], b->soa[4], b->soa[5], b->lp, b->lc, b->n);
    if (rc)
      return bmsg(b, rc, ehd_message(d));
  }
  first = b->var[variant].first;
  order = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  pos = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  lp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  ev = (bev *)malloc(sizeof(bev) * (n + 1));
  if (!order || !pos || !lp || !ev) {
    free(order);
    free(pos);
    free(lp);
    free(ev);
    return EH_ERR_NOMEM;
  }
  for (int64_t i = 0; i < n; i++) {
    order[i] = pos[i] = i;
    if (b->lp[first + i] >= 0) {
      ev[m].pass = b->lp[first + i];
      ev[m].code = b->lc[first + i];
      ev[m].slot = i;
      m++;
    }
  }
  qsort(ev, (size_t)m, sizeof(bev), cmp_bev);
  alive = n;
  for (int64_t a = 0; a < m;) {
    int64_t e = a;
    while (e < m && ev[e].pass == ev[a].pass && ev[e].code == ev[a].code)
      e++;
    {
      int64_t *slots = (int64_t *)malloc(sizeof(int64_t) * (e - a));
      if (!slots) {
        free(order);
        free(pos);
        free(lp);
        free(ev);
        return EH_ERR_NOMEM;
      }
      for (int64_t j = a; j < e; j++)
        slots[j - a] = ev[j].slot;
      eh_compact_event(order, pos, alive, slots, e - a, lp);
      free(slots);
    }
    alive -= e - a;
    a = e;
  }
  for (int64_t j = 0; j < n; j++) {
    int64_t i = first + order[j];
    for (int k = 0; k < 6; k++)
      coord[j][k] = b->soa[k][i];
    coord[j][6] = b->ids[i];
    coord[j][7] = b->lp[i] >= 0 ? (double)b->lp[i] : -1.0;
  }
  if (nLeft)
    *nLeft = alive;
  free(order);
  free(pos);
  free(lp);
  free(ev);
  return EH_OK;
}

int32_t eh_batch_variants(const eh_batch *b) { return b ? b->nvar : 0; }
