/*
 * oracle.h - CPU restatement of elegant's 6D tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the GPU path is
 * compared against and the timed CPU baseline; nothing in the product links
 * it.  Each routine restates one reference routine in its operation order,
 * on AoS rows (double **coord, columns as in track.h:81-101), with
 * swap-to-tail loss compaction.
 *
 * Build modes (see oracle/Makefile):
 *  - liboracle_glibc.so: per-particle transcendentals from glibc, as the
 *    reference is built (Makefile.OAG:1-20: -O3, ISO C, no FP contraction);
 *  - liboracle_emath.so: per-particle transcendentals from the project's
 *    deterministic math header (elegant_amd/csrc/emath.h), the same source
 *    the GPU kernels use.  Per-element constants use glibc in both modes, as
 *    does the GPU path's host side.
 */
#ifndef ELEGANT_ORACLE_H
#define ELEGANT_ORACLE_H

#include <stdint.h>
#include "../include/elegant_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_ELEMS 4096

/* state that the reference keeps in element structs or file statics */
typedef struct {
  double macroParticleCharge;      /* CHARGE (do_tracking.c:1324-1353) */
  int charge_seen;
  double phase_fiducial[OR_MAX_ELEMS]; /* RFCA phase_fiducial (simple_rfca.c:466) */
  int fiducial_seen[OR_MAX_ELEMS];
  /* last WAKE/ZLONGIT histogram, for kernel-level parity tests */
  int64_t last_nb;
  double last_tmin, last_dt;
  double *last_hist;   /* [last_nb] */
  double *last_volt;   /* [last_nb] */
  int64_t hist_cap;
  int64_t effort;
  int zl_initialized[OR_MAX_ELEMS];  /* ZLONGIT bin settings (zlongit.c:583-632) */
  long zl_n_bins[OR_MAX_ELEMS];
  double zl_bin_size[OR_MAX_ELEMS];
  long particle_limit; /* stop_tracking_particle_limit (do_tracking.c:500-520), 0 = off */
  long id_slots_per_bunch; /* bunched beam: buckets by the bunch index in column 8 */
} or_ctx;

or_ctx *or_ctx_new(void);
void or_ctx_free(or_ctx *c);
int or_math_mode(void); /* 0 = glibc, 1 = emath */
int64_t or_ctx_effort(const or_ctx *c);
void or_math_eval(int fn, const double *in, const double *in2, double *out, long n);

/* element routines (reference names, or_ prefix) */
void or_track_particles_drift(double **coord, long np, double length, int order);
void or_thin_corrector(double **coord, long np, double kick, double tilt);
void or_exact_drift(double **part, long np, double length);
long or_multipole_tracking2(double **particle, long n_part, const eh_element *elem, double Po, double z_start,
                            const EH_MAXAMP *maxamp);
long or_track_through_csbend(double **part, long n_part, const EH_CSBEND *csbend, double Po, double z_start,
                             const EH_MAXAMP *maxamp);
long or_simple_rf_cavity(or_ctx *c, long elem_index, double **part, long np, const EH_RFCA *rfca, double *P_central,
                         double zEnd);
long or_track_through_wake(or_ctx *c, double **part, long np, const EH_WAKE *wake, double Po, long i_pass);
long or_track_through_wake_p(or_ctx *c, double **part, long np, const EH_WAKE *wake, double *Po, long i_pass);
long or_track_through_zlongit(or_ctx *c, long elem_index, double **part, long np, const EH_ZLONGIT *zl, double Po,
                              long i_pass);
long or_limit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                         long openCode);
long or_elimit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                          long openCode, long exponent, long yexponent);
long or_remove_invalid_particles(double **coord, long np, double z, double Po);
void or_add_to_particle_energy(double *coord, double timeOfFlight, double Po, double dgamma);
double or_kahan_plus(double oldSum, double b, double *error);
void or_do_match_energy(double **coord, long np, double *P_central);
void or_etilt_offsets(double rho0, double angle, double etilt, double tilt, double *out);
void or_set_row_width(long nprops);
long or_row_width(void);
/* elegant's species globals (change_particle, elegant.c:2933-2991) */
long or_rectangular_collimator(double **initial, const EH_RCOL *rcol, long np, double z, double Po);
void or_offset_beam(double **coord, long nToTrack, const EH_MALIGN *offset, double P_central);
void or_set_species(double mass, double charge, double massMV, double radius, double relSign);
void or_set_id_slots_per_bunch(or_ctx *c, long n);

/* do_tracking pass/element loop over a flattened beamline.  Returns nLeft,
 * or a negative EH_ERR_* code.  Lost rows carry the pass of loss in column 7. */
long or_do_tracking(or_ctx *c, double **coord, long nOriginal, const eh_element *bl, long n_elems,
                    double *P_central, long n_passes, long passOffset);

int64_t or_last_hist(const or_ctx *c, double *hist, double *volt, int64_t cap, double *tmin, double *dt);
double or_compute_average(const double *v, long n);
void or_set_macro_charge(or_ctx *c, double q);
void or_set_particle_limit(or_ctx *c, long limit);
void or_set_fiducial(or_ctx *c, long elem_index, double phase_fiducial);
double or_get_fiducial(const or_ctx *c, long elem_index);

/* FFT used by ZLONGIT (packed real layout, see oracle.c) */
void or_real_fft(double *data, long n, int inverse);

#ifdef __cplusplus
}
#endif
#endif
