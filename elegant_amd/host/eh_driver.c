/*
 * eh_driver.c - the public C ABI: context, upload/download, the
 * do_tracking-shaped pass loop, reference row order, and the
 * element-by-element entry points.
 *
 * The device keeps particles in fixed slots with a per-slot loss record
 * (pass, element, stage).  elegant instead moves each element's lost rows to
 * the tail of the array (swap-to-tail, e.g. multipole.c:1065-1080).  The
 * order that produces is a deterministic function of which rows were lost
 * at each element, so it is rebuilt here, on the host, whenever a caller or
 * an order-dependent reduction (the RFCA fiducial mean, simple_rfca.c:179-210)
 * needs it.
 */
#define _POSIX_C_SOURCE 200809L
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "eh_consts.h"
#include "eh_host.h"

const char *ehd_message(const ehd_dev *d);

struct eh_ctx {
  ehd_dev *dev;
  char msg[512];
  eh_program *prog;
  eh_element *elems;
  int64_t n_elems;
  double Po;
  int loaded;
  /* particles */
  int64_t n;
  double *ids;
  int64_t *order, *pos, *scratch;
  int64_t n_alive_ref;
  uint8_t *applied;
  int32_t *h_pass, *h_code;
  double *h_s, *h_dp;
  /* per-program state (reference element-struct state) */
  int fid_seen[EH_MAX_RFCA];
  double phase_fid[EH_MAX_RFCA];
  double macro_charge;
  int charge_seen;
  int coll_loaded[EH_MAX_COLL];
  int32_t zl_nbins[EH_MAX_COLL];
  double zl_binsize[EH_MAX_COLL];
  int64_t pass_offset0; /* first pass of the run (CHARGE acts on it) */
  int64_t n_original;
  int32_t chunk_passes;
  /* columns of a particle row (totalPropertiesPerParticle, elegant.c:906-912):
   * rows move whole, as swapParticles moves them (elegant.c:2663-2670) */
  int32_t row_props;
  /* columns 8.. of the uploaded rows (bunch, weight, ...), carried with
   * their particle to eh_download */
  double *extra;
  int32_t extra_w;
  /* bunched beams: charge->idSlotsPerBunch (do_tracking.c:1341), and
   * whether the device holds the bunch index of the particles uploaded */
  int64_t id_slots_per_bunch;
  int bunch_on_device;
  eh_species sp; /* the particle species (change_particle) */
  /* stopping conditions of the pass loop (do_tracking.c:494-520) */
  int64_t particle_limit;
  char *interrupt_path;
  double interrupt_mtime;
  volatile const int32_t *interrupt_flag;
};

static int fail(eh_ctx *c, int code, const char *m) {
  if (c)
    snprintf(c->msg, sizeof(c->msg), "%s", m);
  return code;
}

static int devfail(eh_ctx *c, int code) {
  if (c && code)
    snprintf(c->msg, sizeof(c->msg), "%s", ehd_message(c->dev));
  return code;
}

const char *eh_strerror(int code) {
  switch (code) {
  case EH_OK:
    return "ok";
  case EH_ERR_INVALID:
    return "invalid argument or element parameter";
  case EH_ERR_UNSUPPORTED:
    return "element feature outside the implemented subset";
  case EH_ERR_HIP:
    return "HIP runtime failure";
  case EH_ERR_NOMEM:
    return "out of memory";
  case EH_ERR_STATE:
    return "call out of order";
  case EH_ERR_NODEVICE:
    return "no GPU device";
  case EH_ERR_LONGBEAM:
    return "bunch longer than the wake or impedance window";
  case EH_ERR_FIDUCIAL:
    return "no particle available for RF fiducialization";
  default:
    return "unknown error";
  }
}

const char *eh_last_error(const eh_ctx *c) { return c ? c->msg : ""; }

int eh_ctx_create(int device, eh_ctx **out) {
  eh_ctx *c;
  int rc;
  if (!out)
    return EH_ERR_INVALID;
  *out = NULL;
  c = (eh_ctx *)calloc(1, sizeof(eh_ctx));
  if (!c)
    return EH_ERR_NOMEM;
  c->prog = (eh_program *)calloc(1, sizeof(eh_program));
  if (!c->prog) {
    free(c);
    return EH_ERR_NOMEM;
  }
  rc = ehd_create(device, &c->dev, c->msg, sizeof(c->msg));
  if (rc) {
    free(c->prog);
    free(c);
    return rc;
  }
  c->chunk_passes = 64;
  c->row_props = 8;
  eh_species_electron(&c->sp);
  *out = c;
  return EH_OK;
}

/* ---- particle species (process_particle_command, elegant.c:2933-2991) ---- */

void eh_species_electron(eh_species *sp) {
  /* "ensure exact backward compatibility" (elegant.c:2946-2953) */
  sp->mass = EH_ME_MKS;
  sp->charge = EH_E_MKS;
  sp->massMV = EH_ME_MEV;
  sp->radius = EH_RE_MKS;
  sp->relSign = 1;
}

int eh_species_named(const char *name, double mass_ratio, double charge_ratio, eh_species *sp) {
  double charge; /* the table's charge of the type, electron = -e (elegant.c:2928-2929) */
  if (!name || !sp)
    return EH_ERR_INVALID;
  if (!strcmp(name, "electron")) {
    eh_species_electron(sp);
    return EH_OK;
  }
  if (!strcmp(name, "positron")) {
    /* "ensure exact reversal from electron" (elegant.c:2954-2961) */
    eh_species_electron(sp);
    sp->charge = -EH_E_MKS;
    sp->relSign = -1;
    return EH_OK;
  }
  if (!strcmp(name, "proton")) {
    sp->mass = EH_MP_MKS;
    charge = EH_E_MKS;
  } else if (!strcmp(name, "muon")) {
    sp->mass = EH_MMU_MKS;
    charge = -EH_E_MKS;
  } else if (!strcmp(name, "custom")) {
    if (mass_ratio <= 0 || charge_ratio == 0)
      return EH_ERR_INVALID; /* "Must have mass_ratio>0 and charge_ratio nonzero" */
    sp->mass = EH_ME_MKS * mass_ratio;
    charge = -EH_E_MKS * charge_ratio;
  } else
    return EH_ERR_INVALID;
  /* the minus sign: a legacy of the electron-only code (elegant.c:2968-2977) */
  sp->charge = -charge;
  /* particleRelSign = -SIGN(particleCharge)/SIGN(charge[electron]), and the
   * electron's table charge is negative: the sign of particleCharge */
  sp->relSign = sp->charge > 0 ? 1.0 : -1.0;
  sp->massMV = sp->mass * (EH_C_MKS * EH_C_MKS) / fabs(sp->charge) / 1e6;
  sp->radius = (sp->charge * sp->charge) / (4 * EH_PI * EH_EPSILON_O * sp->mass * (EH_C_MKS * EH_C_MKS));
  return EH_OK;
}

int eh_set_species(eh_ctx *c, const eh_species *sp) {
  if (!c || !sp)
    return EH_ERR_INVALID;
  if (!(sp->mass > 0) || !(sp->massMV > 0) || sp->charge == 0 || (sp->relSign != 1 && sp->relSign != -1))
    return fail(c, EH_ERR_INVALID, "eh_set_species: mass and massMV > 0, charge != 0, relSign +-1");
  c->sp = *sp;
  return EH_OK;
}

const eh_species *eh_ctx_species(const eh_ctx *c) { return &c->sp; }

int eh_set_particle_properties(eh_ctx *c, int32_t nprops) {
  if (!c || nprops < 8 || nprops > 64)
    return fail(c, EH_ERR_INVALID, "eh_set_particle_properties: 8 <= nprops <= 64");
  c->row_props = nprops;
  return EH_OK;
}

static void free_particles(eh_ctx *c) {
  free(c->ids);
  free(c->order);
  free(c->pos);
  free(c->scratch);
  free(c->applied);
  free(c->h_pass);
  free(c->h_code);
  free(c->h_s);
  free(c->h_dp);
  free(c->extra);
  c->extra = NULL;
  c->extra_w = 0;
  c->ids = NULL;
  c->order = c->pos = c->scratch = NULL;
  c->applied = NULL;
  c->h_pass = c->h_code = NULL;
  c->h_s = c->h_dp = NULL;
  c->n = 0;
}

int eh_ctx_destroy(eh_ctx *c) {
  if (!c)
    return EH_OK;
  free_particles(c);
  free(c->elems);
  free(c->prog);
  free(c->interrupt_path);
  ehd_destroy(c->dev);
  free(c);
  return EH_OK;
}

void eh_set_chunk_passes(eh_ctx *c, int32_t passes) {
  if (c && passes > 0)
    c->chunk_passes = passes;
}

void eh_set_block(eh_ctx *c, int32_t block) {
  if (c)
    ehd_set_block(c->dev, block);
}

/* ------------------------------------------------------------ program */

int eh_load_program(eh_ctx *c, const eh_element *elems, int64_t n_elems, double Po) {
  int rc;
  if (!c || (!elems && n_elems > 0) || n_elems < 0 || !(Po > 0))
    return fail(c, EH_ERR_INVALID, "eh_load_program: bad arguments");
  c->msg[0] = 0;
  rc = eh_compile(elems, n_elems, Po, &c->sp, c->prog, c->msg, sizeof(c->msg));
  if (rc)
    return rc;
  free(c->elems);
  c->elems = (eh_element *)malloc(sizeof(eh_element) * (n_elems > 0 ? n_elems : 1));
  if (!c->elems)
    return EH_ERR_NOMEM;
  memcpy(c->elems, elems, sizeof(eh_element) * n_elems);
  c->n_elems = n_elems;
  c->Po = Po;
  memset(c->fid_seen, 0, sizeof(c->fid_seen));
  memset(c->coll_loaded, 0, sizeof(c->coll_loaded));
  c->charge_seen = 0;
  c->macro_charge = 0;
  /* RFCA with a given T_REFERENCE or LIGHT fiducial needs no particles (simple_rfca.c:429-431) */
  for (int k = 0; k < c->prog->n_rfca; k++) {
    const EH_RFCA *rf = &elems[c->prog->rfca_elem[k]].p.rfca;
    eh_op *op = &c->prog->ops[c->prog->rfca_op[k]];
    double t0;
    if (rf->tReference != -1 || rf->fiducialMode == EH_FID_LIGHT) {
      if (rf->tReference != -1)
        t0 = rf->tReference;
      else {
        double length = rf->length, zEnd = op->z_end;
        long nKicks = length ? rf->nKicks : 1;
        if (nKicks <= 0)
          nKicks = 1;
        length /= nKicks;
        t0 = ((zEnd - rf->length) + length / 2) / EH_C_MKS;
      }
      c->phase_fid[k] = eh_rfca_phase_fiducial(rf, t0);
      c->fid_seen[k] = 1;
      eh_rfca_set_phase(op, rf, c->phase_fid[k]);
    }
  }
  rc = ehd_put_program(c->dev, c->prog->ops, c->prog->nops);
  if (rc)
    return devfail(c, rc);
  c->loaded = 1;
  return EH_OK;
}

/* ------------------------------------------------------- particle state */

static int alloc_particles(eh_ctx *c, int64_t n) {
  free_particles(c);
  c->ids = (double *)malloc(sizeof(double) * (n + 1));
  c->order = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  c->pos = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  c->scratch = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  c->applied = (uint8_t *)calloc(n + 1, 1);
  c->h_pass = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  c->h_code = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  c->h_s = (double *)malloc(sizeof(double) * (n + 1));
  c->h_dp = (double *)malloc(sizeof(double) * (n + 1));
  if (!c->ids || !c->order || !c->pos || !c->scratch || !c->applied || !c->h_pass || !c->h_code || !c->h_s ||
      !c->h_dp) {
    free_particles(c);
    return EH_ERR_NOMEM;
  }
  c->n = n;
  for (int64_t i = 0; i < n; i++) {
    c->order[i] = i;
    c->pos[i] = i;
  }
  c->n_alive_ref = n;
  return EH_OK;
}

/* AoS rows -> SoA, with do_tracking's initial NaN/Inf cull (do_tracking.c:468-473)
 * when cull is set: such rows are recorded as lost on pass 0 before the first
 * element.  The single-element entry points upload without it, as the
 * element routines leave such rows to their caller. */
int eh_upload_rows(eh_ctx *c, double **coord, int64_t n, int32_t nprops, int cull) {
  int rc;
  double *soa;
  if (!c || (!coord && n > 0) || n < 0 || nprops < 7)
    return fail(c, EH_ERR_INVALID, "eh_upload: bad arguments");
  if (n > 0x7fffffff)
    return fail(c, EH_ERR_INVALID, "eh_upload: more than 2^31-1 particles");
  rc = alloc_particles(c, n);
  if (rc)
    return rc;
  soa = (double *)malloc(sizeof(double) * 6 * (n + 1));
  if (!soa)
    return EH_ERR_NOMEM;
  c->bunch_on_device = 0;
  if (nprops > 8) {
    c->extra_w = nprops - 8;
    c->extra = (double *)malloc(sizeof(double) * c->extra_w * (n + 1));
    if (!c->extra) {
      free(soa);
      return EH_ERR_NOMEM;
    }
    for (int64_t i = 0; i < n; i++)
      memcpy(c->extra + c->extra_w * i, coord[i] + 8, sizeof(double) * c->extra_w);
  }
  {
    int64_t nlost = 0;
    for (int64_t i = 0; i < n; i++) {
      double *r = coord[i];
      for (int k = 0; k < 6; k++)
        soa[k * n + i] = r[k];
      c->ids[i] = r[6];
      c->h_pass[i] = -1;
      c->h_code[i] = 0;
      if (cull && (isinf(r[0]) || isinf(r[2]) || isnan(r[0]) || isnan(r[2]))) {
        /* limit_amplitudes(DBL_MAX) at z = 0: s = 0, delta = Po(1+delta) */
        soa[4 * n + i] = 0.0;
        soa[5 * n + i] = c->Po * (1 + r[5]);
        c->h_pass[i] = 0;
        c->h_code[i] = -1;
        nlost++;
      }
    }
    (void)nlost;
  }
  rc = ehd_alloc_state(c->dev, n);
  if (!rc)
    rc = ehd_put_state(c->dev, soa, soa + n, soa + 2 * n, soa + 3 * n, soa + 4 * n, soa + 5 * n, c->h_pass, c->h_code,
                       n);
  free(soa);
  if (rc)
    return devfail(c, rc);
  return EH_OK;
}

int eh_upload(eh_ctx *c, double **coord, int64_t n, int32_t nprops) { return eh_upload_rows(c, coord, n, nprops, 1); }

/* ----------------------------------------------------- reference order */

/* Swap-to-tail compaction of one element call's losses
 * (multipole.c:1046-1081, limit_amplitudes.c:217-268):
 *   for (ip = 0; ip <= itop; ip++) if lost(ip) { swap(ip, itop); itop--; ip--; }
 * Survivors above the new nLeft fill the holes from the top down; the cost
 * is O(k log k) for k lost rows, not O(n). */
static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

void eh_compact_event(int64_t *order, int64_t *pos, int64_t n_alive, const int64_t *lost_slots, int64_t k,
                      int64_t *lp) {
  int64_t itop = n_alive - 1, li = 0, hi = k - 1;
  for (int64_t j = 0; j < k; j++)
    lp[j] = pos[lost_slots[j]];
  qsort(lp, (size_t)k, sizeof(int64_t), cmp_i64);
  while (li <= hi) {
    int64_t ip = lp[li];
    for (;;) {
      int64_t a = order[ip], b = order[itop];
      order[ip] = b;
      order[itop] = a;
      pos[b] = ip;
      pos[a] = itop;
      if (lp[hi] == itop) {
        /* the row brought down from the top is lost as well: examine it again */
        itop--;
        hi--;
        if (li > hi || ip > itop)
          break;
      } else {
        itop--;
        li++;
        break;
      }
    }
  }
}

typedef struct {
  int64_t pass;
  int32_t code;
  int64_t slot;
} loss_ev;

static int cmp_ev(const void *a, const void *b) {
  const loss_ev *x = (const loss_ev *)a, *y = (const loss_ev *)b;
  if (x->pass != y->pass)
    return x->pass < y->pass ? -1 : 1;
  if (x->code != y->code)
    return x->code < y->code ? -1 : 1;
  return 0;
}

/* Apply, in order, every loss event not yet applied whose (pass, code) is
 * strictly before (pass_lim, code_lim).  h_pass/h_code must be current. */
static int replay_until(eh_ctx *c, int64_t pass_lim, int32_t code_lim) {
  int64_t m = 0;
  loss_ev *ev;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] >= 0 && !c->applied[i])
      m++;
  if (!m)
    return EH_OK;
  ev = (loss_ev *)malloc(sizeof(loss_ev) * m);
  if (!ev)
    return EH_ERR_NOMEM;
  m = 0;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] >= 0 && !c->applied[i]) {
      if (c->h_pass[i] < pass_lim || (c->h_pass[i] == pass_lim && c->h_code[i] < code_lim)) {
        ev[m].pass = c->h_pass[i];
        ev[m].code = c->h_code[i];
        ev[m].slot = i;
        m++;
      }
    }
  qsort(ev, (size_t)m, sizeof(loss_ev), cmp_ev);
  {
    int64_t a = 0;
    while (a < m) {
      int64_t b = a;
      while (b < m && ev[b].pass == ev[a].pass && ev[b].code == ev[a].code)
        b++;
      for (int64_t j = a; j < b; j++)
        c->scratch[j - a] = ev[j].slot;
      {
        /* compaction needs the slot list and a position buffer */
        int64_t k = b - a;
        int64_t *lp = (int64_t *)malloc(sizeof(int64_t) * (k > 0 ? k : 1));
        if (!lp) {
          free(ev);
          return EH_ERR_NOMEM;
        }
        eh_compact_event(c->order, c->pos, c->n_alive_ref, c->scratch, k, lp);
        free(lp);
        c->n_alive_ref -= k;
      }
      for (int64_t j = a; j < b; j++)
        c->applied[ev[j].slot] = 1;
      a = b;
    }
  }
  free(ev);
  return EH_OK;
}

static int refresh_losses(eh_ctx *c) {
  int rc = ehd_get_state(c->dev, NULL, NULL, NULL, NULL, NULL, NULL, c->h_pass, c->h_code, c->n);
  return rc ? devfail(c, rc) : EH_OK;
}

/* ------------------------------------------------------------ tracking */

/* RFCA TMEAN fiducial on the first pass: KahanPlus mean of arrival times over
 * the particles in reference row order (simple_rfca.c:179-210, kahan.c:27-36) */
static double kahan_plus(double oldSum, double b, double *error) {
  volatile double sum, Y;
  Y = b + *error;
  sum = oldSum + Y;
  *error = Y - (sum - oldSum);
  return sum;
}

static int rfca_fiducial(eh_ctx *c, int k, int64_t pass) {
  const int32_t opi = c->prog->rfca_op[k];
  eh_op *op = &c->prog->ops[opi];
  const EH_RFCA *rf = &c->elems[c->prog->rfca_elem[k]].p.rfca;
  double length = rf->length, sOffset, tsum = 0, error = 0, t0;
  long nKicks = length ? rf->nKicks : 1;
  int64_t nsum = 0;
  int rc;
  if (nKicks <= 0)
    nKicks = 1;
  length /= nKicks;
  sOffset = length / 2;
  rc = ehd_get_state(c->dev, NULL, NULL, NULL, NULL, c->h_s, c->h_dp, c->h_pass, c->h_code, c->n);
  if (rc)
    return devfail(c, rc);
  rc = replay_until(c, pass, 4 * opi);
  if (rc)
    return rc;
  for (int64_t j = 0; j < c->n_alive_ref; j++) {
    int64_t i = c->order[j];
    double p = c->Po * (1 + c->h_dp[i]);
    double beta = p / sqrt(p * p + 1);
    tsum = kahan_plus(tsum, (c->h_s[i] + sOffset) / (EH_C_MKS * beta), &error);
    nsum++;
  }
  if (nsum <= 0)
    return EH_OK; /* no particle left: do_tracking's element loop stops before the
                     cavity (do_tracking.c:816), so the phase stays unset */
  t0 = tsum / nsum;
  c->phase_fid[k] = eh_rfca_phase_fiducial(rf, t0);
  c->fid_seen[k] = 1;
  eh_rfca_set_phase(op, rf, c->phase_fid[k]);
  rc = ehd_patch_op(c->dev, opi, op);
  return rc ? devfail(c, rc) : EH_OK;
}

int eh_collective(eh_ctx *c, int slot, int64_t pass, int64_t n_alive);

int eh_rfca_fiducial(eh_ctx *c, int64_t elem_index, double *phase_fiducial) {
  if (!c || !phase_fiducial)
    return EH_ERR_INVALID;
  if (!c->loaded)
    return EH_ERR_STATE;
  for (int k = 0; k < c->prog->n_rfca; k++)
    if (c->prog->rfca_elem[k] == elem_index) {
      *phase_fiducial = c->fid_seen[k] ? c->phase_fid[k] : NAN;
      return EH_OK;
    }
  return EH_ERR_INVALID;
}

/* CHANGE_T (simple_rfca.c:488-497): the cavity's time offset follows the
 * beam's first row (part[0]) at every pass.  The first row in reference
 * order moves only when the particle there is lost (swap compaction never
 * touches row 0 otherwise), so one loss flag is read per pass and the loss
 * events are replayed only when that particle has gone. */
static int rfca_time_offset(eh_ctx *c, int k, int64_t pass) {
  const int32_t opi = c->prog->rfca_op[k];
  eh_op *op = &c->prog->ops[opi];
  double s = 0, dp = 0, toff = 0;
  int32_t lp = 0;
  int rc;
  if (c->n_alive_ref > 0) {
    rc = ehd_get_one(c->dev, c->order[0], NULL, NULL, &lp);
    if (rc)
      return devfail(c, rc);
  }
  if (c->n_alive_ref <= 0 || lp >= 0) {
    rc = refresh_losses(c);
    if (!rc)
      rc = replay_until(c, pass, 4 * opi);
    if (rc)
      return rc;
  }
  if (c->n_alive_ref > 0) {
    rc = ehd_get_one(c->dev, c->order[0], &s, &dp, NULL);
    if (rc)
      return devfail(c, rc);
    toff = eh_rfca_time_offset(op, c->Po, s, dp);
  }
  if (op->d[EH_D_TOFF] != toff) {
    op->d[EH_D_TOFF] = toff;
    rc = ehd_patch_op(c->dev, opi, op);
    if (rc)
      return devfail(c, rc);
  }
  return EH_OK;
}

/* index of the first op >= from that needs the host on this pass, or nops */
static int32_t next_barrier(const eh_ctx *c, int32_t from) {
  const eh_program *p = c->prog;
  int32_t best = p->nops;
  for (int k = 0; k < p->n_rfca; k++)
    if ((!c->fid_seen[k] || (p->ops[p->rfca_op[k]].flags & EH_F_CHANGE_T)) && p->rfca_op[k] >= from &&
        p->rfca_op[k] < best)
      best = p->rfca_op[k];
  for (int k = 0; k < p->n_coll; k++)
    if (p->coll_op[k] >= from && p->coll_op[k] < best)
      best = p->coll_op[k];
  for (int k = 0; k < p->n_p0; k++)
    if (p->p0_op[k] >= from && p->p0_op[k] < best)
      best = p->p0_op[k];
  return best;
}

/* the program for the current central momentum c->Po, with the RFCA phases
 * fixed so far, compiled and uploaded (CHANGE_P0 moves the momentum every
 * element after the cavity sees, field-specified multipole strengths
 * included) */
static int recompile_program(eh_ctx *c) {
  int rc = eh_compile(c->elems, c->n_elems, c->Po, &c->sp, c->prog, c->msg, sizeof(c->msg));
  if (rc)
    return rc;
  for (int k = 0; k < c->prog->n_rfca; k++)
    if (c->fid_seen[k])
      eh_rfca_set_phase(&c->prog->ops[c->prog->rfca_op[k]], &c->elems[c->prog->rfca_elem[k]].p.rfca,
                        c->phase_fid[k]);
  rc = ehd_put_program(c->dev, c->prog->ops, c->prog->nops);
  return rc ? devfail(c, rc) : EH_OK;
}

/* do_match_energy after an RFCA with CHANGE_P0 (simple_rfca.c:977-982,
 * do_tracking.c:3371-3417): the KahanPlus mean of P_central (1 + delta)
 * over the particles left after the cavity's own losses, in reference row
 * order; when it moved by more than 1e-14 relative, delta is re-expressed
 * against it on the device and the central momentum becomes it (and the
 * program is recompiled for it if recompile is set).  opi is the
 * EH_OP_MATCH_P0 step. */
static int match_energy(eh_ctx *c, int32_t opi, int64_t pass, int recompile) {
  double P_average = 0, error = 0;
  int rc = ehd_get_state(c->dev, NULL, NULL, NULL, NULL, NULL, c->h_dp, c->h_pass, c->h_code, c->n);
  if (rc)
    return devfail(c, rc);
  rc = replay_until(c, pass, 4 * opi);
  if (rc)
    return rc;
  if (c->n_alive_ref <= 0)
    return EH_OK;
  for (int64_t j = 0; j < c->n_alive_ref; j++)
    P_average = kahan_plus(P_average, c->Po * (1 + c->h_dp[c->order[j]]), &error);
  P_average /= c->n_alive_ref;
  if (fabs(P_average - c->Po) / c->Po > 1e-14) {
    const double dp = (c->Po - P_average) / P_average, dr = c->Po / P_average;
    rc = ehd_rescale_dp(c->dev, dp, dr);
    if (rc)
      return devfail(c, rc);
    c->Po = P_average;
    if (recompile)
      return recompile_program(c);
  }
  return EH_OK;
}

static void charge_elements(eh_ctx *c, int64_t pass) {
  for (int64_t i = 0; i < c->n_elems; i++) {
    const eh_element *e = &c->elems[i];
    if (e->type != EH_T_CHARGE)
      continue;
    /* CHARGE (do_tracking.c:1324-1353) */
    if (pass == c->pass_offset0) {
      c->macro_charge = 0;
      if (c->n_original)
        c->macro_charge = e->p.charge.charge / c->n_original;
      c->charge_seen = 1;
    }
    if (e->p.charge.chargePerParticle)
      c->macro_charge = e->p.charge.chargePerParticle;
  }
}

/* ---- stopping conditions of do_tracking's pass loop (do_tracking.c:494-520) ---- */

int eh_set_particle_limit(eh_ctx *c, int64_t limit) {
  if (!c)
    return EH_ERR_INVALID;
  c->particle_limit = limit > 0 ? limit : 0;
  return EH_OK;
}

static double file_mtime(const char *path) {
  struct stat st;
  if (stat(path, &st) != 0)
    return -1;
  return (double)st.st_mtim.tv_sec + 1e-9 * (double)st.st_mtim.tv_nsec;
}

int eh_set_interrupt_file(eh_ctx *c, const char *path) {
  if (!c)
    return EH_ERR_INVALID;
  free(c->interrupt_path);
  c->interrupt_path = NULL;
  c->interrupt_mtime = 0;
  if (path && path[0]) {
    size_t len = strlen(path);
    c->interrupt_path = (char *)malloc(len + 1);
    if (!c->interrupt_path)
      return EH_ERR_NOMEM;
    memcpy(c->interrupt_path, path, len + 1);
    /* a file already there counts only once it is modified again (elegant.c:1133-1137) */
    double m = file_mtime(path);
    c->interrupt_mtime = m > 0 ? m : 0;
  }
  return EH_OK;
}

int eh_set_interrupt_flag(eh_ctx *c, volatile const int32_t *flag) {
  if (!c)
    return EH_ERR_INVALID;
  c->interrupt_flag = flag;
  return EH_OK;
}

static int interrupt_requested(const eh_ctx *c) {
  if (c->interrupt_flag && *c->interrupt_flag)
    return 1;
  if (c->interrupt_path) {
    double m = file_mtime(c->interrupt_path);
    if (m >= 0 && m > c->interrupt_mtime)
      return 1;
  }
  return 0;
}

/* particles not lost before pass q (h_pass current); rows removed by the
 * initial NaN cull (code -1) were lost before the first pass */
static int64_t alive_at(const eh_ctx *c, int64_t q) {
  int64_t k = 0;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] < 0 || (c->h_pass[i] >= q && c->h_code[i] != -1))
      k++;
  return k;
}

int eh_track(eh_ctx *c, int64_t pass0, int64_t npasses, double *P_central, eh_stats *stats) {
  int rc;
  int64_t n_alive0 = 0, done_end;
  if (!c || !P_central || npasses < 0)
    return fail(c, EH_ERR_INVALID, "eh_track: bad arguments");
  if (!c->loaded)
    return fail(c, EH_ERR_STATE, "eh_track: no program loaded");
  if (*P_central != c->Po)
    return fail(c, EH_ERR_INVALID, "eh_track: P_central differs from the program's Po");
  if (c->n == 0) {
    if (stats)
      memset(stats, 0, sizeof(*stats));
    return EH_OK;
  }
  ehd_timing_reset(c->dev);
  rc = refresh_losses(c);
  if (rc)
    return rc;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] < 0 || c->h_pass[i] >= pass0)
      n_alive0++;
  {
    int64_t pass = pass0, end = pass0 + npasses;
    while (pass < end) {
      int32_t b = next_barrier(c, 0);
      if (interrupt_requested(c))
        break;
      if (c->particle_limit > 0) {
        rc = refresh_losses(c);
        if (rc)
          return rc;
        if (alive_at(c, pass) < c->particle_limit)
          break;
      }
      charge_elements(c, pass);
      if (b >= c->prog->nops) {
        /* no host step in this pass: run many passes in one launch */
        int64_t chunk = end - pass;
        if (chunk > c->chunk_passes)
          chunk = c->chunk_passes;
        if (c->particle_limit > 0) {
          /* the limit may be crossed inside the launch: keep the state at its
           * start, and if a later pass of it would not have begun, go back
           * and run only the passes before that one */
          rc = ehd_state_snapshot(c->dev, 0);
          if (rc)
            return devfail(c, rc);
          rc = ehd_track(c->dev, pass, (int32_t)chunk, 0, c->prog->nops);
          if (rc)
            return devfail(c, rc);
          rc = refresh_losses(c);
          if (rc)
            return rc;
          int64_t stop = -1;
          for (int64_t q = pass + 1; q < pass + chunk && stop < 0; q++)
            if (alive_at(c, q) < c->particle_limit)
              stop = q;
          if (stop >= 0) {
            rc = ehd_state_snapshot(c->dev, 1);
            if (!rc)
              rc = ehd_track(c->dev, pass, (int32_t)(stop - pass), 0, c->prog->nops);
            if (rc)
              return devfail(c, rc);
            pass = stop;
            break;
          }
          pass += chunk;
          continue;
        }
        rc = ehd_track(c->dev, pass, (int32_t)chunk, 0, c->prog->nops);
        if (rc)
          return devfail(c, rc);
        pass += chunk;
        continue;
      }
      /* segmented pass */
      {
        int32_t cur = 0;
        while (cur < c->prog->nops) {
          b = next_barrier(c, cur);
          if (b > cur) {
            rc = ehd_track(c->dev, pass, 1, cur, b);
            if (rc)
              return devfail(c, rc);
          }
          if (b >= c->prog->nops)
            break;
          {
            const eh_op *op = &c->prog->ops[b];
            if (op->op == EH_OP_RFCA) {
              int k;
              for (k = 0; k < c->prog->n_rfca; k++)
                if (c->prog->rfca_op[k] == b)
                  break;
              rc = EH_OK;
              if (!c->fid_seen[k])
                rc = rfca_fiducial(c, k, pass);
              if (!rc && (op->flags & EH_F_CHANGE_T))
                rc = rfca_time_offset(c, k, pass);
              if (rc)
                return rc;
              cur = b; /* the RFCA itself runs in the next launch */
              if (next_barrier(c, cur) == cur) {
                rc = ehd_track(c->dev, pass, 1, cur, cur + 1);
                if (rc)
                  return devfail(c, rc);
                cur = cur + 1;
              }
            } else if (op->op == EH_OP_MATCH_P0) {
              rc = match_energy(c, b, pass, 1);
              if (rc)
                return rc;
              cur = b + 1;
            } else {
              rc = eh_collective(c, op->nkicks_rf, pass, 0);
              if (rc)
                return rc;
              cur = b + 1;
            }
          }
        }
        pass++;
      }
    }
    done_end = pass;
  }
  rc = refresh_losses(c);
  if (rc)
    return rc;
  if (stats) {
    int64_t effort = 0, left = 0;
    for (int64_t i = 0; i < c->n; i++) {
      int64_t lp = c->h_pass[i];
      if (lp >= 0 && lp < pass0)
        continue;
      if (lp < 0) {
        effort += done_end - pass0;
        left++;
      } else
        effort += lp - pass0;
    }
    memset(stats, 0, sizeof(*stats));
    stats->effort = effort;
    stats->n_left = left;
    stats->passes_done = done_end - pass0;
    ehd_timing(c->dev, &stats->kernel_launches, &stats->track_ms);
    ehd_collective_timing(c->dev, &left, &stats->collective_ms);
  }
  (void)n_alive0;
  *P_central = c->Po; /* moved by CHANGE_P0 cavities, as do_tracking's *P_central */
  return EH_OK;
}

int eh_last_kernel_timing(const eh_ctx *c, int64_t *launches, double *total_ms) {
  if (!c || !launches || !total_ms)
    return EH_ERR_INVALID;
  return ehd_timing(c->dev, launches, total_ms);
}

/* rows in reference order: survivors, then lost rows (tail order of the
 * compactions), column 7 = pass of loss (-1 for survivors) */
int eh_download(eh_ctx *c, double **coord, int64_t n, int32_t nprops, int64_t *nLeft) {
  double *soa;
  int rc;
  if (!c || (!coord && n > 0) || nprops < 8 || n != c->n)
    return fail(c, EH_ERR_INVALID, "eh_download: bad arguments");
  soa = (double *)malloc(sizeof(double) * 6 * (n + 1));
  if (!soa)
    return EH_ERR_NOMEM;
  rc = ehd_get_state(c->dev, soa, soa + n, soa + 2 * n, soa + 3 * n, soa + 4 * n, soa + 5 * n, c->h_pass, c->h_code,
                     n);
  if (rc) {
    free(soa);
    return devfail(c, rc);
  }
  rc = replay_until(c, INT64_MAX, INT32_MAX);
  if (rc) {
    free(soa);
    return rc;
  }
  for (int64_t j = 0; j < n; j++) {
    int64_t i = c->order[j];
    double *r = coord[j];
    for (int k = 0; k < 6; k++)
      r[k] = soa[k * n + i];
    r[6] = c->ids[i];
    r[7] = c->h_pass[i] >= 0 ? (double)c->h_pass[i] : -1.0;
    for (int k = 8; k < nprops && k - 8 < c->extra_w; k++)
      r[k] = c->extra[c->extra_w * i + k - 8];
  }
  if (nLeft)
    *nLeft = c->n_alive_ref;
  free(soa);
  return EH_OK;
}

int64_t eh_do_tracking(eh_ctx *c, double **coord, int64_t nOriginal, int64_t *effort, const eh_element *beamline,
                       int64_t n_elems, double *P_central, double **accepted, int64_t n_passes, int64_t passOffset) {
  eh_stats st;
  int64_t nLeft = 0;
  int rc;
  const int64_t W = c ? c->row_props : 0;
  double *initial = NULL;
  if (!c || !P_central || nOriginal < 0 || (nOriginal > 0 && (!coord || !beamline)))
    return fail(c, EH_ERR_INVALID, "eh_do_tracking: bad arguments");
  rc = eh_load_program(c, beamline, n_elems, *P_central);
  if (rc)
    return rc;
  if (accepted && nOriginal > 0) {
    /* accepted starts as a copy of the rows (do_tracking.c:360-361) and is
     * then permuted with them, so it ends holding each particle's initial
     * row in that particle's final row */
    initial = (double *)malloc(sizeof(double) * W * nOriginal);
    if (!initial)
      return EH_ERR_NOMEM;
    for (int64_t i = 0; i < nOriginal; i++)
      memcpy(initial + W * i, coord[i], sizeof(double) * W);
  }
  rc = eh_upload(c, coord, nOriginal, (int32_t)W);
  if (!rc) {
    c->pass_offset0 = passOffset;
    c->n_original = nOriginal;
    rc = eh_track(c, passOffset, n_passes, P_central, &st);
  }
  if (!rc)
    rc = eh_download(c, coord, nOriginal, (int32_t)W, &nLeft);
  if (!rc && initial)
    for (int64_t j = 0; j < nOriginal; j++)
      memcpy(accepted[j], initial + W * c->order[j], sizeof(double) * W);
  free(initial);
  if (rc)
    return rc;
  if (effort)
    *effort += st.effort;
  return nLeft;
}

/* ------------------------------------------------ element entry points */

/* The rows of a single-element call back to the caller in the reference's
 * order.  Whole rows move, as swapParticles moves row_props doubles
 * (elegant.c:2663-2670): the element's columns 0..6 come from the device,
 * the other columns keep their values and follow their particle, and the
 * accepted rows, if given, take the same permutation (multipole.c:1075-1076,
 * limit_amplitudes.c:263-264 and 308-309).  Rows were uploaded in their
 * order, so slot i is row i. */
static int writeback_rows(eh_ctx *c, double **rows, int64_t n, double **accepted, int64_t *nLeft) {
  const int64_t W = c->row_props;
  double **tmp = (double **)malloc(sizeof(double *) * n);
  double *buf = (double *)malloc(sizeof(double) * 8 * n);
  double *keep = (double *)malloc(sizeof(double) * W * n);
  double *acc = accepted ? (double *)malloc(sizeof(double) * W * n) : NULL;
  int rc;
  if (!tmp || !buf || !keep || (accepted && !acc)) {
    free(tmp);
    free(buf);
    free(keep);
    free(acc);
    return EH_ERR_NOMEM;
  }
  for (int64_t i = 0; i < n; i++)
    tmp[i] = buf + 8 * i;
  rc = eh_download(c, tmp, n, 8, nLeft);
  if (!rc) {
    for (int64_t i = 0; i < n; i++) {
      memcpy(keep + W * i, rows[i], sizeof(double) * W);
      if (acc)
        memcpy(acc + W * i, accepted[i], sizeof(double) * W);
    }
    for (int64_t j = 0; j < n; j++) {
      const int64_t src = c->order[j];
      memcpy(rows[j], tmp[j], sizeof(double) * 7);
      memcpy(rows[j] + 7, keep + W * src + 7, sizeof(double) * (W - 7));
      if (acc)
        memcpy(accepted[j], acc + W * src, sizeof(double) * W);
    }
  }
  free(tmp);
  free(buf);
  free(keep);
  free(acc);
  return rc;
}

/* run a one-element program on host rows and compact them like the
 * reference element routine (lost rows to the tail) */
static int64_t run_single(eh_ctx *c, double **rows, int64_t n, double **accepted, const eh_element *e, double Po,
                          double z_start, const EH_MAXAMP *maxamp, int64_t pass) {
  eh_element bl[2];
  int nb = 0, rc;
  int64_t nLeft = 0;
  eh_stats st;
  if (n <= 0)
    return 0;
  if (maxamp) {
    eh_element_defaults(&bl[nb], EH_T_MAXAMP);
    bl[nb].p.maxamp = *maxamp;
    nb++;
  }
  bl[nb++] = *e;
  rc = eh_compile(bl, nb, Po, &c->sp, c->prog, c->msg, sizeof(c->msg));
  if (rc)
    return rc;
  /* keep only the element itself, without the checks do_tracking adds after
   * it (the MAXAMP aperture and the NaN/infinity cull of limit_amplitudes):
   * the element routines leave such rows to their caller */
  c->prog->ops[0] = c->prog->ops[c->prog->nops - 1];
  c->prog->nops = 1;
  c->prog->ops[0].flags &= ~EH_F_AP;
  c->prog->ops[0].flags |= EH_F_ALONE;
  c->prog->ops[0].z_start = z_start;
  c->prog->ops[0].z_end = z_start + eh_element_length(e);
  c->Po = Po;
  c->loaded = 1;
  rc = ehd_put_program(c->dev, c->prog->ops, 1);
  if (rc)
    return devfail(c, rc);
  rc = eh_upload_rows(c, rows, n, 8, 0);
  if (rc)
    return rc;
  rc = ehd_track(c->dev, pass, 1, 0, 1);
  if (rc)
    return devfail(c, rc);
  rc = writeback_rows(c, rows, n, accepted, &nLeft);
  if (rc)
    return rc;
  (void)st;
  return nLeft;
}

int64_t eh_multipole_tracking2(eh_ctx *c, double **particle, int64_t n_part, const eh_element *elem, double Po,
                               double **accepted, double z_start, const EH_MAXAMP *maxamp) {
  if (!c || !elem || (elem->type != EH_T_KQUAD && elem->type != EH_T_KSEXT && elem->type != EH_T_KOCT))
    return fail(c, EH_ERR_INVALID, "eh_multipole_tracking2: element must be KQUAD, KSEXT or KOCT");
  return run_single(c, particle, n_part, accepted, elem, Po, z_start, maxamp, 0);
}

int64_t eh_track_through_csbend(eh_ctx *c, double **part, int64_t n_part, const EH_CSBEND *csbend, double Po,
                                double **accepted, double z_start, const EH_MAXAMP *maxamp) {
  eh_element e;
  if (!c || !csbend)
    return EH_ERR_INVALID;
  eh_element_defaults(&e, EH_T_CSBEND);
  e.p.csbend = *csbend;
  return run_single(c, part, n_part, accepted, &e, Po, z_start, maxamp, 0);
}

int64_t eh_track_particles_drift(eh_ctx *c, double **part, int64_t n_part, const EH_DRIFT *drift) {
  eh_element e;
  if (!c || !drift)
    return EH_ERR_INVALID;
  eh_element_defaults(&e, EH_T_DRIF);
  e.p.drift = *drift;
  return run_single(c, part, n_part, NULL, &e, 1.0, 0.0, NULL, 0);
}

int64_t eh_simple_rf_cavity(eh_ctx *c, double **part, int64_t np, const EH_RFCA *rfca, double **accepted,
                            double *P_central, double zEnd, double *phase_fiducial) {
  eh_element e;
  int64_t r;
  if (!c || !rfca || !P_central || !phase_fiducial)
    return EH_ERR_INVALID;
  if (np <= 0)
    return np;
  eh_element_defaults(&e, EH_T_RFCA);
  e.p.rfca = *rfca;
  if (isnan(*phase_fiducial)) {
    if (rfca->tReference != -1 || rfca->fiducialMode == EH_FID_LIGHT) {
      double length = rfca->length, t0;
      long nKicks = length ? rfca->nKicks : 1;
      if (nKicks <= 0)
        nKicks = 1;
      length /= nKicks;
      t0 = rfca->tReference != -1 ? rfca->tReference : ((zEnd - rfca->length) + length / 2) / EH_C_MKS;
      *phase_fiducial = eh_rfca_phase_fiducial(rfca, t0);
    } else {
      double length = rfca->length, tsum = 0, error = 0;
      long nKicks = length ? rfca->nKicks : 1;
      if (nKicks <= 0)
        nKicks = 1;
      length /= nKicks;
      for (int64_t i = 0; i < np; i++) {
        double p = *P_central * (1 + part[i][5]);
        tsum = kahan_plus(tsum, (part[i][4] + length / 2) / (EH_C_MKS * (p / sqrt(p * p + 1))), &error);
      }
      *phase_fiducial = eh_rfca_phase_fiducial(rfca, tsum / np);
    }
  }
  {
    /* compile the cavity with the known phase */
    eh_element bl[1];
    int rc;
    int64_t nLeft = 0;
    bl[0] = e;
    rc = eh_compile(bl, 1, *P_central, &c->sp, c->prog, c->msg, sizeof(c->msg));
    if (rc)
      return rc;
    if (c->prog->ops[0].op == EH_OP_RFCA) {
      eh_rfca_set_phase(&c->prog->ops[0], rfca, *phase_fiducial);
      /* CHANGE_T: the offset of the caller's first row (part[0]) */
      c->prog->ops[0].d[EH_D_TOFF] = eh_rfca_time_offset(&c->prog->ops[0], *P_central, part[0][4], part[0][5]);
    }
    /* the cavity's own removeInvalidParticles is the only cull, as in the
     * routine it replaces */
    c->prog->ops[0].flags &= ~EH_F_AP;
    c->prog->ops[0].flags |= EH_F_ALONE;
    c->prog->ops[0].z_end = zEnd;
    c->Po = *P_central;
    c->loaded = 1;
    rc = ehd_put_program(c->dev, c->prog->ops, 1);
    if (rc)
      return devfail(c, rc);
    rc = eh_upload_rows(c, part, np, 8, 0);
    if (rc)
      return rc;
    rc = ehd_track(c->dev, 0, 1, 0, 1);
    if (rc)
      return devfail(c, rc);
    if (c->prog->n_p0) {
      /* CHANGE_P0: the match step that follows the cavity's own culls */
      rc = match_energy(c, c->prog->p0_op[0], 0, 0);
      if (rc)
        return rc;
      *P_central = c->Po;
    }
    rc = writeback_rows(c, part, np, accepted, &nLeft);
    if (rc)
      return rc;
    r = nLeft;
  }
  return r;
}

/* one MAXAMP check on its own, at z: a zero-length step whose only work is
 * the check do_tracking makes after an element (do_tracking.c:2461-2470)
 * with the given aperture */
static int64_t run_aperture(eh_ctx *c, double **coord, int64_t np, double **accepted, const EH_MAXAMP *m, double z,
                            double Po, int32_t extrapolate_z) {
  eh_element bl[2];
  int rc;
  int64_t nLeft = 0;
  eh_element_defaults(&bl[0], EH_T_MAXAMP);
  bl[0].p.maxamp = *m;
  eh_element_defaults(&bl[1], EH_T_DRIF);
  bl[1].p.drift.length = 0;
  rc = eh_compile(bl, 2, Po, &c->sp, c->prog, c->msg, sizeof(c->msg));
  if (rc)
    return rc;
  c->prog->ops[0] = c->prog->ops[1];
  c->prog->ops[0].op = EH_OP_NOP;
  c->prog->ops[0].z_end = z;
  if (!extrapolate_z)
    c->prog->ops[0].flags &= ~EH_F_EXTRAP;
  c->prog->nops = 1;
  c->Po = Po;
  c->loaded = 1;
  rc = ehd_put_program(c->dev, c->prog->ops, 1);
  if (rc)
    return devfail(c, rc);
  /* NaN and infinite rows are the check's own to remove, at z */
  rc = eh_upload_rows(c, coord, np, 8, 0);
  if (rc)
    return rc;
  rc = ehd_track(c->dev, 0, 1, 0, 1);
  if (rc)
    return devfail(c, rc);
  rc = writeback_rows(c, coord, np, accepted, &nLeft);
  if (rc)
    return rc;
  return nLeft;
}

int64_t eh_limit_amplitudes(eh_ctx *c, double **coord, double xmax, double ymax, int64_t np, double **accepted,
                            double z, double Po, int32_t extrapolate_z, int32_t openCode) {
  EH_MAXAMP m;
  if (!c || openCode < 0 || openCode > 4)
    return EH_ERR_INVALID;
  /* limit_amplitudes.c:215-217 */
  if (np <= 0 || (xmax < 0 && ymax < 0))
    return np;
  memset(&m, 0, sizeof(m));
  m.x_max = xmax;
  m.y_max = ymax;
  m.exponent = 2;
  m.openSide = openCode;
  return run_aperture(c, coord, np, accepted, &m, z, Po, extrapolate_z);
}

int64_t eh_elimit_amplitudes(eh_ctx *c, double **coord, double xmax, double ymax, int64_t np, double **accepted,
                             double z, double Po, int32_t extrapolate_z, int32_t openCode, int32_t exponent,
                             int32_t yexponent) {
  EH_MAXAMP m;
  if (!c || openCode < 0 || openCode > 4)
    return EH_ERR_INVALID;
  /* limit_amplitudes.c:507-521 */
  if (exponent < 2 || exponent % 2 || (yexponent && (yexponent < 2 || yexponent % 2)))
    return EH_ERR_INVALID;
  /* a non-positive half-axis: the rectangle in the other plane, or no
   * check at all (limit_amplitudes.c:523-528) */
  if (xmax <= 0 || ymax <= 0) {
    if (xmax > 0 || ymax > 0)
      return eh_limit_amplitudes(c, coord, xmax, ymax, np, accepted, z, Po, extrapolate_z, openCode);
    return np;
  }
  if (np <= 0)
    return np;
  memset(&m, 0, sizeof(m));
  m.x_max = xmax;
  m.y_max = ymax;
  m.elliptical = 1;
  m.exponent = exponent;
  m.yExponent = yexponent;
  m.openSide = openCode;
  return run_aperture(c, coord, np, accepted, &m, z, Po, extrapolate_z);
}

/* debugging/measurement helpers (not in the public header) */
int eh_get_loss_records(eh_ctx *c, int32_t *pass, int32_t *code) {
  int rc = refresh_losses(c);
  if (rc)
    return rc;
  memcpy(pass, c->h_pass, sizeof(int32_t) * c->n);
  memcpy(code, c->h_code, sizeof(int32_t) * c->n);
  return EH_OK;
}

int eh_math_probe(eh_ctx *c, int32_t fn, const double *in, const double *in2, double *out, int64_t n) {
  int rc = ehd_math_probe(c->dev, fn, in, in2, out, n);
  return rc ? devfail(c, rc) : EH_OK;
}

/* device buffers for tests of the device-pointer entry points (not in the
 * public header) */
int eh_test_buffer(eh_ctx *c, int32_t op, void **ptr, void *host, int64_t bytes) {
  int rc = ehd_test_buffer(c->dev, op, ptr, host, bytes);
  return rc ? devfail(c, rc) : EH_OK;
}

/* the runs the tracking kernel walks for a beamline: compiles it on the host
 * and folds it (eh_runs.c) into out[4 k .. 4 k + 3] = first, len, reps,
 * flat of run k, for at most cap runs; returns the number of runs or a
 * negative EH_ERR_* code (not in the public header) */
int32_t eh_test_fold(const eh_element *bl, int64_t n_elems, double Po, int32_t *out, int32_t cap) {
  eh_program *prog = (eh_program *)malloc(sizeof(eh_program));
  char msg[256];
  int32_t nr = EH_ERR_NOMEM, nb = 0;
  eh_species sp;
  if (!prog)
    return EH_ERR_NOMEM;
  eh_species_electron(&sp);
  int rc = eh_compile(bl, n_elems, Po, &sp, prog, msg, sizeof msg);
  if (rc) {
    free(prog);
    return rc;
  }
  const size_t m = (size_t)(prog->nops > 0 ? prog->nops : 1);
  eh_run *runs = (eh_run *)malloc(sizeof(eh_run) * m);
  int32_t *body = (int32_t *)malloc(sizeof(int32_t) * m);
  if (runs && body)
    nr = eh_fold_runs(prog->ops, prog->nops, EH_RUN_MAX_LEN, runs, body, &nb);
  if (nr < 0)
    nr = EH_ERR_NOMEM;
  for (int32_t k = 0; k < nr && k < cap; k++) {
    out[4 * k] = runs[k].first;
    out[4 * k + 1] = runs[k].len;
    out[4 * k + 2] = runs[k].reps;
    out[4 * k + 3] = runs[k].flat;
  }
  free(runs);
  free(body);
  free(prog);
  return nr;
}

/* how many runs the loaded tracking program has on the device (not in the
 * public header) */
int32_t eh_test_program_runs(eh_ctx *c) { return ehd_program_runs(c->dev); }

/* ---- bunched beams (index_bunch_assignments, lrwake.c:17-233) ---- */

int eh_set_id_slots_per_bunch(eh_ctx *c, int64_t id_slots_per_bunch) {
  if (!c || id_slots_per_bunch < 0)
    return EH_ERR_INVALID;
  c->id_slots_per_bunch = id_slots_per_bunch;
  return EH_OK;
}

/* the bunch index of particle slot i: column 8 of its row, as a long */
static long bunch_of(const eh_ctx *c, int64_t i) { return (long)c->extra[(int64_t)c->extra_w * i]; }

/* Buckets of the particles still in the beam: 0 without particles, 1 when
 * the beam is not bunched (no id slots, or rows without a bunch column),
 * else ibMax - ibMin + 1 with *ib_min set.  The loss state is refreshed
 * from the device first. */
int eh_ctx_buckets(eh_ctx *c, long *ib_min) {
  long lo = LONG_MAX, hi = LONG_MIN;
  int rc;
  *ib_min = 0;
  if (c->id_slots_per_bunch <= 0 || c->extra_w < 1)
    return 1;
  rc = refresh_losses(c);
  if (rc)
    return rc;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] < 0) {
      long b = bunch_of(c, i);
      if (b < lo)
        lo = b;
      if (b > hi)
        hi = b;
    }
  if (lo == LONG_MAX)
    return 0;
  if (lo < INT32_MIN || hi > INT32_MAX || hi - lo >= INT32_MAX)
    return fail(c, EH_ERR_UNSUPPORTED, "bunch indices beyond the 32-bit range");
  *ib_min = lo;
  return (int)(hi - lo + 1);
}

/* particles still in the beam with bunch index b */
int64_t eh_ctx_bucket_count(const eh_ctx *c, long b) {
  int64_t k = 0;
  for (int64_t i = 0; i < c->n; i++)
    if (c->h_pass[i] < 0 && bunch_of(c, i) == b)
      k++;
  return k;
}

/* the device's copy of the bunch indices, uploaded once per particle upload */
int eh_ctx_bunch_to_device(eh_ctx *c) {
  if (c->bunch_on_device)
    return EH_OK;
  int32_t *b = (int32_t *)malloc(sizeof(int32_t) * (c->n > 0 ? c->n : 1));
  if (!b)
    return EH_ERR_NOMEM;
  for (int64_t i = 0; i < c->n; i++)
    b[i] = (int32_t)bunch_of(c, i);
  int rc = ehd_put_bunch(c->dev, b, c->n);
  free(b);
  if (rc)
    return devfail(c, rc);
  c->bunch_on_device = 1;
  return EH_OK;
}

/* WAKE with N_BINS over one bucket of a bunched beam: its mean arrival
 * time is a plain sum in row order (wake.c:99-104), taken here over the
 * bucket's particles in the reference's row order at the element (loss
 * events replayed up to code_lim of this pass), with the kernels' time
 * formula (computeTimeCoordinatesOnly, trwake.c:519-527) */
int eh_ctx_bucket_mean(eh_ctx *c, int64_t pass, int32_t code_lim, long b, double Po, double *mean) {
  double sum = 0;
  int64_t k = 0;
  int rc = ehd_get_state(c->dev, NULL, NULL, NULL, NULL, c->h_s, c->h_dp, c->h_pass, c->h_code, c->n);
  if (rc)
    return devfail(c, rc);
  rc = replay_until(c, pass, code_lim);
  if (rc)
    return rc;
  for (int64_t j = 0; j < c->n_alive_ref; j++) {
    int64_t i = c->order[j];
    if (bunch_of(c, i) != b)
      continue;
    double P = Po * (c->h_dp[i] + 1);
    sum += c->h_s[i] * sqrt(P * P + 1) / (EH_C_MKS * P);
    k++;
  }
  *mean = k ? sum / k : 0;
  return EH_OK;
}

/* accessors used by the collective step (eh_collective.c) */
ehd_dev *eh_ctx_dev(eh_ctx *c) { return c->dev; }
int32_t eh_ctx_row_props(eh_ctx *c) { return c->row_props; }
const eh_program *eh_ctx_prog(eh_ctx *c) { return c->prog; }
const eh_element *eh_ctx_elems(eh_ctx *c) { return c->elems; }
double eh_ctx_macro_charge(eh_ctx *c, int *seen) {
  *seen = c->charge_seen;
  return c->macro_charge;
}
char *eh_ctx_msg(eh_ctx *c) { return c->msg; }
int *eh_ctx_coll_loaded(eh_ctx *c) { return c->coll_loaded; }
int32_t *eh_ctx_zl_nbins(eh_ctx *c) { return c->zl_nbins; }
double *eh_ctx_zl_binsize(eh_ctx *c) { return c->zl_binsize; }
