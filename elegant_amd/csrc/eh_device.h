/*
 * eh_device.h - internal C ABI between the C host (elegant_amd/host) and the
 * HIP launch layer (the .hip sources in elegant_amd/csrc).  Not part of the
 * public API.
 */
#ifndef ELEGANT_AMD_EH_DEVICE_H
#define ELEGANT_AMD_EH_DEVICE_H

#include <stdint.h>

#include "eh_program.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ehd_dev ehd_dev;

/* returns 0 or a negative EH_ERR_* code; msg (if not NULL) gets a description */
int ehd_create(int device, ehd_dev **out, char *msg, int msglen);
void ehd_destroy(ehd_dev *d);

/* particle state: SoA doubles x, xp, y, yp, s, dp and int32 loss_pass/loss_code */
int ehd_alloc_state(ehd_dev *d, int64_t n);
int ehd_put_state(ehd_dev *d, const double *x, const double *xp, const double *y, const double *yp,
                  const double *s, const double *dp, const int32_t *loss_pass, const int32_t *loss_code, int64_t n);
int ehd_get_one(ehd_dev *d, int64_t i, double *s, double *dp, int32_t *loss_pass);
int ehd_get_state(ehd_dev *d, double *x, double *xp, double *y, double *yp, double *s, double *dp,
                  int32_t *loss_pass, int32_t *loss_code, int64_t n);
int ehd_get_sdp(ehd_dev *d, double *s, double *dp, int32_t *loss_pass, int64_t n);
int64_t ehd_state_n(const ehd_dev *d);

/* program: ehd_put_ops loads the flat record array the batch kernel reads
 * (eh_vdesc offsets into it); ehd_put_program loads the tracking kernel's
 * program, folded into runs of repeated cells (eh_fold_runs), and
 * ehd_patch_op replaces one of its records by flat position */
int ehd_put_ops(ehd_dev *d, const eh_op *ops, int32_t nops);
int ehd_put_program(ehd_dev *d, const eh_op *ops, int32_t nops);
int ehd_patch_op(ehd_dev *d, int32_t index, const eh_op *op);
int32_t ehd_program_runs(const ehd_dev *d);

/* host side (eh_runs.c): folds ops[0, n) into runs of repeated cells with
 * bodies of at most max_len records; see there */
int32_t eh_fold_runs(const eh_op *ops, int32_t n, int32_t max_len, eh_run *runs, int32_t *body, int32_t *nbody);
#define EH_RUN_MAX_LEN 256

/* run passes [pass0, pass0 + npasses): ops [op_first, nops) in the first pass,
 * [0, op_last) in the last, all ops in between.  Asynchronous on the
 * library stream; timing is accumulated with HIP events. */
int ehd_track(ehd_dev *d, int64_t pass0, int32_t npasses, int32_t op_first, int32_t op_last);

/* survivors after the last completed work (synchronizes) */
int ehd_count_alive(ehd_dev *d, int64_t *n_alive);
int ehd_sync(ehd_dev *d);

/* timing of the tracking kernel since the last reset */
void ehd_timing_reset(ehd_dev *d);
int ehd_timing(ehd_dev *d, int64_t *launches, double *ms);

/* launch tuning: particles per workgroup (multiple of 64) */
void ehd_set_block(ehd_dev *d, int32_t block);

/* ---- batched variants ---- */
/* vdesc/chunks describe variants and work items; ops are all programs concatenated */
int ehd_batch_setup(ehd_dev *d, const eh_vdesc *v, int32_t nvar, const eh_chunk *chunks, int32_t nchunks);
/* replace the work queue by chunks over a list of particle indices (the
 * surviving particles, grouped by variant); chunk.first indexes perm */
int ehd_batch_compact(ehd_dev *d, const int32_t *perm, int64_t nperm, const eh_chunk *chunks, int32_t nchunks);
/* passes [pass0, pass0+npasses) for every variant, npasses split into launches of
 * passes_per_launch; ppl particles per lane (1 or 2); graph = capture and replay;
 * device_compact: launches of 1, 2, 4 ... passes up to passes_per_launch with
 * the work queue rebuilt from the survivors on the device between them */
int ehd_batch_track(ehd_dev *d, int64_t pass0, int32_t npasses, int32_t passes_per_launch, int32_t ppl,
                    int32_t block, int32_t use_graph, int32_t device_compact, double *ms, int64_t *launches);
int ehd_state_snapshot(ehd_dev *d, int restore);
/* delta = dp + delta dr on the particles not lost (CHANGE_P0) */
int ehd_rescale_dp(ehd_dev *d, double dp, double dr);

/* ---- collective elements (WAKE / ZLONGIT) ---- */
typedef struct {
  int32_t kind;         /* EH_OP_WAKE or EH_OP_ZLONGIT */
  int32_t op_index;     /* position in the program (loss codes) */
  int32_t interpolate;
  int32_t n_bins_fixed; /* WAKE N_BINS (0 = autoscale) */
  int64_t pass;
  double Po, dt, factor, macro_charge, ramp;
  double z_end, ap_xmax, ap_ymax;
  int32_t flags;        /* EH_F_AP */
  int32_t nb;           /* ZLONGIT n_bins */
  double wake_span;     /* WAKE: t[last] - t[0] (long-beam check) */
  int32_t allow_long_beam;
  int32_t ap_code;      /* EH_AP_* shape of the MAXAMP check, 0 = rectangle */
  int32_t use_bunch;    /* one bunch of a bunched beam: the particles with bunch index bunch_target */
  int32_t bunch_target;
  int32_t mean_given;   /* WAKE with N_BINS: the bucket's mean time from the host (a plain sum) */
  double mean;
  double mc2;           /* 1e6 particleMassMV particleRelSign: voltage to gamma (wake.c:320-322) */
} ehd_collective;

/* WAKE Green function samples; ZLONGIT impedance Z[nb] and twiddles */
int ehd_put_wake(ehd_dev *d, int32_t slot, const double *W, int64_t npoints);
int ehd_put_zlongit(ehd_dev *d, int32_t slot, const double *Z, const double *tc, const double *ts, int32_t nb);
/* one WAKE or ZLONGIT step on all alive particles.  Returns EH_ERR_LONGBEAM if the
 * bunch is longer than allowed.  hist/volt (optional, may be NULL) receive the
 * histogram and the scaled voltage; *nb_out and *tmin_out describe them. */
int ehd_collective_step(ehd_dev *d, int32_t slot, const ehd_collective *c, double *hist, double *volt,
                        int64_t hist_cap, int64_t *nb_out, double *tmin_out);
int ehd_collective_timing(ehd_dev *d, int64_t *launches, double *ms);
/* the post-element checks alone, for the bunch a step leaves out */
int ehd_collective_checks(ehd_dev *d, const ehd_collective *c);
/* exact min/max of the alive particles' arrival times and their count */
int ehd_time_span(ehd_dev *d, double Po, int32_t use_bunch, int32_t target, double *tmin, double *tmax,
                  int64_t *count);
int ehd_put_bunch(ehd_dev *d, const int32_t *bunch, int64_t n);
/* The WAKE step without any allocation: device SoA particle arrays crd[0..5]
 * of n particles (all alive, no checks after the step) and the Green
 * function W[wpoints] on the device, owned by the caller; the step's scratch
 * is the library's static per-GPU array (at most ehd_noalloc_bins_max()
 * wake points).  Calls that overlap on one GPU share that array and race. */
int ehd_noalloc_init(ehd_dev *d);
int64_t ehd_noalloc_bins_max(void);
int ehd_wake_step_noalloc(ehd_dev *d, const ehd_collective *c, double *const crd[6], int64_t n, const double *W,
                          int64_t wpoints);
#define EHD_COLL_NOCHECKS (1 << 30) /* element entry points: no do_tracking checks after the step */

/* tests: op 0 allocate *ptr, 1 free it, 2 copy host -> *ptr, 3 copy *ptr -> host */
int ehd_test_buffer(ehd_dev *d, int32_t op, void **ptr, void *host, int64_t bytes);

/* deterministic math check: out[i] = f(in[i]) on the device, f = 0 sin, 1 cos, 2 asin, 3 log, 4 tan,
 * 5 sqrt, 6 division in[i]/in2[i] */
int ehd_math_probe(ehd_dev *d, int32_t fn, const double *in, const double *in2, double *out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
