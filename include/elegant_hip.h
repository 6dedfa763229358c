/*
 * elegant_hip.h - C ABI of the gfx950 6D tracking library (libelegant_hip.so).
 *
 * The library replaces the per-particle inner loops of elegant's tracking
 * core.  The reference seam is the top-of-routine hook each element routine
 * carries for an accelerator ("if (getElementOnGpu()) { ...; return nLeft; }",
 * multipole.c:731-742, csbend.c:700-712, simple_rfca.c:267-289, wake.c:47-59,
 * limit_amplitudes.c:200-213, matrix.c:306-321) and the lifecycle calls made
 * by do_tracking (gpuBaseInit at do_tracking.c:401-404, gpuBaseDealloc at
 * do_tracking.c:3104-3111).
 *
 * Conventions kept from the reference:
 *  - particles are rows of a caller-owned double **coord; columns 0..5 are
 *    (x, x', y, y', s, delta), column 6 the particle ID, column 7 the pass of
 *    loss (track.h:81-101); rows are mutated in place;
 *  - element routines return the survivor count nLeft; lost rows are moved to
 *    [nLeft, n) by elegant's swap-to-tail compaction, with s replaced by the
 *    loss position and delta by the absolute momentum Po(1+delta)
 *    (multipole.c:1065-1080, limit_amplitudes.c:232-268);
 *  - element parameters use elegant's names and defaults (track_data.c).
 * Differences: errors are negative return codes (never exit()), there is no
 * global mutable state, and one context per device/stream may be used from
 * one host thread.
 *
 * All functions return EH_OK (0) or a negative EH_ERR_* code unless stated.
 */
#ifndef ELEGANT_HIP_H
#define ELEGANT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EH_OK 0
#define EH_ERR_INVALID (-1)     /* bad argument or element parameter */
#define EH_ERR_UNSUPPORTED (-2) /* element feature outside the implemented subset */
#define EH_ERR_HIP (-3)         /* HIP runtime failure */
#define EH_ERR_NOMEM (-4)
#define EH_ERR_STATE (-5)       /* call out of order (e.g. track before upload) */
#define EH_ERR_NODEVICE (-6)
#define EH_ERR_LONGBEAM (-7)    /* bunch longer than wake/impedance window (wake.c:134-140, zlongit.c:219-236) */
#define EH_ERR_FIDUCIAL (-8)    /* no particle available for RF fiducialization (simple_rfca.c:213-214) */

/* element type codes, named after elegant's entity names (track_data.c:72-173) */
#define EH_T_MARK 1
#define EH_T_DRIF 2
#define EH_T_EDRIFT 3
#define EH_T_KQUAD 4
#define EH_T_KSEXT 5
#define EH_T_KOCT 6
#define EH_T_CSBEND 7
#define EH_T_RFCA 8
#define EH_T_WAKE 9
#define EH_T_ZLONGIT 10
#define EH_T_MAXAMP 11
#define EH_T_CHARGE 12
#define EH_T_MALIGN 13
#define EH_T_MONI 14
#define EH_T_HMON 15
#define EH_T_VMON 16
#define EH_T_HKICK 17
#define EH_T_VKICK 18
#define EH_T_RECIRC 19 /* recirculation point: passes after pass 0 start here (do_tracking.c:459-465) */
#define EH_T_RCOL 20

/* DRIF (track.h:1310-1316); tracked with the order-2 drift matrix
 * (compute_matrices.c:1101-1104, matrix.c:389-429) */
typedef struct {
  double length;
  int32_t order; /* 0 = run default (2); 1 or 2 */
  int32_t pad;
} EH_DRIFT;

/* EDRIFT: exact drift (csbend.c:3767-3790) */
typedef struct {
  double length;
} EH_EDRIFT;

/* KQUAD (track.h:2151-2175; defaults track_data.c:1222-1283) */
typedef struct {
  double length, k1, tilt, dx, dy, dz, fse;
  double xkick, ykick, xKickCalibration, yKickCalibration;
  double bore, B, lEffective;
  int32_t n_kicks, nSlices, integration_order, expandHamiltonian;
  int32_t synch_rad, isr, edge1_effects, edge2_effects;
  int32_t radial, malignMethod;
} EH_KQUAD;

/* KSEXT (track.h:2096-2113; defaults track_data.c:1120-1160) */
typedef struct {
  double length, k2, k1, j1, tilt, dx, dy, dz, fse;
  double xkick, ykick, xKickCalibration, yKickCalibration;
  double bore, B;
  int32_t n_kicks, nSlices, integration_order, expandHamiltonian;
  int32_t synch_rad, isr, malignMethod, pad;
} EH_KSEXT;

/* KOCT (track.h:2118-2129; defaults track_data.c:1164-1187) */
typedef struct {
  double length, k3, tilt, dx, dy, dz, fse, bore, B;
  int32_t n_kicks, nSlices, integration_order, expandHamiltonian;
  int32_t synch_rad, isr, malignMethod, pad;
} EH_KOCT;

/* CSBEND (track.h:2409-2448; defaults track_data.c:1893-1980) */
typedef struct {
  double length, angle, k1, k2, k3, k4;
  double e1, e2, tilt, h1, h2, hgap, fint, fint1, fint2;
  double dx, dy, dz, xKick, yKick;
  double fse, fseDipole, fseQuadrupole, etilt;
  double edge1_kick_limit, edge2_kick_limit;
  int32_t nSlices, etiltSign, nonlinear, synch_rad;
  int32_t edge1_effects, edge2_effects, edge_order, integration_order;
  int32_t expandHamiltonian, isr, malignMethod, referenceCorrection;
  int32_t kick_limit_scaling, fseCorrection;
} EH_CSBEND;

/* RFCA (track.h:1443-1457; defaults track_data.c:576-597) */
#define EH_FID_TMEAN 0
#define EH_FID_LIGHT 1
#define EH_FID_FIRST 2
#define EH_FID_PMAX 3
typedef struct {
  double length, volt, phase, freq, Q, dx, dy, tReference;
  int32_t nKicks, change_p0, change_t, end1Focus;
  int32_t end2Focus, standingWave, linearize, lockPhase;
  int32_t fiducialMode, bodyFocusModel; /* bodyFocusModel: 0 = none */
} EH_RFCA;

/* WAKE (track.h:2908-2924; defaults track_data.c:2396-2414).  The Green
 * function W(t) (V/C) is given as wakePoints samples t[], W[] starting at
 * t = 0 with uniform spacing (set_up_wake, wake.c:340-477). */
typedef struct {
  const double *t;
  const double *W;
  int64_t wakePoints;
  double factor;
  int32_t n_bins, interpolate, rampPasses, allowLongBeam;
  int32_t smoothing, change_p0, bunchedBeamMode;
  /* START_BUNCH/END_BUNCH: with a bunched beam, the buckets (counted from
   * the lowest bunch index present) outside [startBunch, endBunch] take no
   * wake (wake.c:84-88); -1, the reference's default, is no limit */
  int32_t startBunch, endBunch, pad;
} EH_WAKE;

/* ZLONGIT (track.h:2836-2865; defaults track_data.c:2251-2283): the
 * broad-band resonator (BROAD_BAND=1, zlongit.c:587-632) or an impedance
 * given as samples (zlongit.c:634-700) */
typedef struct {
  double Ra, Rs, Q, freq, bin_size, factor;
  int32_t broad_band, n_bins, max_n_bins, interpolate;
  int32_t area_weight, startOnPass, rampPasses, reverseTimeOrder;
  int32_t smoothing, allowLongBeam, bunchedBeamMode;
  int32_t startBunch, endBunch, pad; /* as in EH_WAKE (zlongit.c:142-146) */
  /* BROAD_BAND=0: the ZREAL and/or ZIMAG tables, n_spect samples at the
   * equally spaced frequencies f[] (Hz, from 0; the two files of the
   * reference must share count and spacing, so one array serves both),
   * n_spect - 1 a power of 2 of at least 4; Zreal or Zimag NULL: zero.
   * N_BINS becomes 2 (n_spect - 1) and BIN_SIZE 1 / (N_BINS (f[1] - f[0])). */
  const double *f, *Zreal, *Zimag;
  int64_t n_spect;
} EH_ZLONGIT;

/* MAXAMP (track.h:1955-1961; defaults track_data.c:990-997).  ELLIPTICAL
 * selects (x/x_max)^exponent + (y/y_max)^yExponent (yExponent 0: exponent;
 * even, >= 2, at most EH_MAX_AP_EXPONENT here).  OPEN_SIDE is held as the
 * code determineOpenSideCode returns (limit_amplitudes.c:1130-1162) for
 * "+x"/"+h", "+y"/"+v", "-x"/"-h", "-y"/"-v" (track.h:1515-1518); 0 = none. */
#define EH_OPEN_PLUS_X 1
#define EH_OPEN_PLUS_Y 2
#define EH_OPEN_MINUS_X 3
#define EH_OPEN_MINUS_Y 4
#define EH_MAX_AP_EXPONENT 254
typedef struct {
  double x_max, y_max;
  int32_t elliptical, exponent, yExponent, openSide;
} EH_MAXAMP;

/* CHARGE (track.h:3220-3226) */
typedef struct {
  double charge, chargePerParticle;
} EH_CHARGE;

/* MALIGN (track.h:1902-1905; defaults track_data.c:926-940): offsets of
 * the beam, applied on pass ON_PASS (-1: every pass) by offset_beam
 * (do_tracking.c:3243-3306, 1818-1822).  START_PID/END_PID other than -1
 * (a subset of particle IDs) are not supported here; FORCE_MODIFY_MATRIX and
 * FLOOR do not touch tracking. */
typedef struct {
  double dxp, dyp, dx, dy, dz, dt, dp, de;
  int32_t on_pass, startPID, endPID, pad;
} EH_MALIGN;

/* MONI, HMON, VMON (track.h monitor structs; defaults track_data.c:600-647):
 * beam-position monitors, tracked as the drift matrix of length L and
 * ORDER (0: the default order 2; compute_matrices.c:1317-1328) with no
 * aperture check after them (NO_APERTURE, track_data.c:3928-3930) but the
 * NaN cull that follows every element.  The other fields only shape
 * readouts and are kept for the caller. */
typedef struct {
  double length, dx, dy, weight, tilt, xcalibration, ycalibration, xsetpoint, ysetpoint;
  int32_t order, pad;
} EH_MONI;

/* HKICK, VKICK (HCOR/VCOR, track.h; defaults track_data.c:544-575): steering
 * correctors, tracked with corrector_matrix (bend_matrix.c:300-316) through
 * track_particles.  Supported here as they are with no strength: KICK x
 * CALIBRATION == 0 gives the drift matrix of length L and ORDER (0: 2),
 * followed by the checks of every element, the aperture's loss point not
 * extrapolated back (only DRIF's is).  A nonzero kick and SYNCH_RAD/ISR are
 * refused. */
typedef struct {
  double length, kick, tilt, b2, calibration, lEffRad;
  int32_t edge_effects, order, steering, synchRad, isr, pad;
} EH_CORRECTOR;

/* RCOL (track.h; defaults track_data.c:651-659): the rectangular
 * collimator of rectangular_collimator (limit_amplitudes.c:32-187): an
 * entrance check at the element's start, then for a positive length the
 * drift with an exit check whose losses are placed where the orbit met the
 * jaw.  OPEN_SIDE is held as determineOpenSideCode returns it (as in
 * EH_MAXAMP); INVERT swaps lost and transmitted. */
typedef struct {
  double length, x_max, y_max, dx, dy;
  int32_t openSide, invert;
} EH_RCOL;

/* One beamline element; mirrors ELEMENT_LIST's (type, p_elem) pair
 * (track.h:308-344) with the parameters held by value. */
typedef struct {
  int32_t type;
  int32_t pad;
  char name[32];
  union {
    EH_DRIFT drift;
    EH_EDRIFT edrift;
    EH_KQUAD kquad;
    EH_KSEXT ksext;
    EH_KOCT koct;
    EH_CSBEND csbend;
    EH_RFCA rfca;
    EH_WAKE wake;
    EH_ZLONGIT zlongit;
    EH_MAXAMP maxamp;
    EH_CHARGE charge;
    EH_MALIGN malign;
    EH_MONI moni; /* MONI, HMON and VMON */
    EH_CORRECTOR corrector; /* HKICK and VKICK */
    EH_RCOL rcol;
  } p;
} eh_element;

/* Fill an element with elegant's defaults for its type (track_data.c tables). */
int eh_element_defaults(eh_element *elem, int32_t type);

/* sizeof of eh_element (type 0) or of the parameter struct of an EH_T_*
 * type, so foreign-function bindings can check their layouts. */
int64_t eh_struct_size(int32_t type);

/* Run statistics returned by eh_track / eh_do_tracking. */
typedef struct {
  int64_t effort;          /* sum over passes of survivors (do_tracking.c:2758-2765) */
  int64_t n_left;          /* survivors after the last pass */
  int64_t passes_done;
  int64_t kernel_launches;
  double track_ms;         /* device time of the tracking kernels (HIP events) */
  double collective_ms;    /* device time of WAKE/ZLONGIT steps */
} eh_stats;

typedef struct eh_ctx eh_ctx;

/* ---- lifecycle (replaces gpuBaseInit / gpuBaseDealloc) ---- */
int eh_ctx_create(int device, eh_ctx **out);
int eh_ctx_destroy(eh_ctx *ctx);
const char *eh_strerror(int code);
/* Message describing the last error on this context (empty if none). */
const char *eh_last_error(const eh_ctx *ctx);

/* ---- particle species (elegant's particleMass, particleCharge,
 * particleMassMV, particleRadius, particleRelSign globals, set by the
 * change_particle command, elegant.c:2933-2991) ----
 * charge keeps elegant's sign convention: +e for electrons ("a legacy of
 * time when this was an electron-only code"), relSign = +1 for electrons and
 * -1 for positively charged particles.  A context starts with electrons.
 * On the tracked path the species enters field-given multipole strengths
 * (charge/mass), RF cavity voltages and wake/impedance kicks (massMV,
 * relSign); radius only enters radiation, which is not tracked here. */
typedef struct {
  double mass;    /* kg */
  double charge;  /* C */
  double massMV;  /* rest energy over the charge, MV */
  double radius;  /* classical radius, m */
  double relSign;
} eh_species;

/* The species of change_particle's NAME ("electron", "positron", "proton",
 * "muon" or "custom" with MASS_RATIO > 0 and CHARGE_RATIO != 0 in electron
 * units), with process_particle_command's arithmetic (elegant.c:2933-2991). */
int eh_species_named(const char *name, double mass_ratio, double charge_ratio, eh_species *out);
/* The species every later program load, element entry point and batch
 * variant of this context uses. */
int eh_set_species(eh_ctx *ctx, const eh_species *species);

/* ---- lattice: flattened beamline, all defaults resolved by the caller ---- */
int eh_load_program(eh_ctx *ctx, const eh_element *elems, int64_t n_elems, double Po);

/* ---- particle rows ----
 * Columns per particle row, elegant's totalPropertiesPerParticle
 * (elegant.c:906-912): 6 coordinates, ID, pass of loss, bunch, weight and,
 * with global loss coordinates, three more.  Rows (and accepted rows) are
 * permuted whole, as swapParticles moves them (elegant.c:2663-2670): the
 * library writes columns 0..7 and carries every other column with its
 * particle.  Default 8; 8 <= nprops <= 64. */
int eh_set_particle_properties(eh_ctx *ctx, int32_t nprops);

/* Bunched beams: charge->idSlotsPerBunch, the beam's id_slots_per_bunch
 * (do_tracking.c:1341).  When it is positive and the rows carry the bunch
 * index in column 8 (nprops >= 9), a WAKE or ZLONGIT with
 * BUNCHED_BEAM_MODE acts on each bucket of particles with one bunch index
 * on its own, in bucket order (index_bunch_assignments, lrwake.c:17-233).
 * Default 0: the beam is one bucket. */
int eh_set_id_slots_per_bunch(eh_ctx *ctx, int64_t id_slots_per_bunch);

/* ---- particle state ---- */
/* Copy AoS rows (>= 7 columns: x, x', y, y', s, delta, ID) to device SoA;
 * columns 8..nprops-1 are kept on the host and follow their particle. */
int eh_upload(eh_ctx *ctx, double **coord, int64_t n, int32_t nprops);
/* Track passes [pass0, pass0+npasses) through the loaded program (the pass
 * and element loops of do_tracking.c:488,816).  *P_central must equal the
 * program's momentum on entry and receives it on return: an RFCA with
 * CHANGE_P0 moves it every pass to the beam's mean momentum (do_match_energy,
 * do_tracking.c:3371-3417), and every element after it uses the new value.  The
 * pass loop stops early, before a pass, when a stopping condition below
 * holds (do_tracking.c:494-520); stats->passes_done then counts the passes
 * completed. */
int eh_track(eh_ctx *ctx, int64_t pass0, int64_t npasses, double *P_central, eh_stats *stats);
/* Write the particles back in elegant's row order: survivors first in the
 * order swap-to-tail compaction leaves them, then lost rows with s = loss
 * position, delta = Po(1+delta) and column 7 = pass of loss; columns 8 and
 * up (as far as both this nprops and the upload's reach) are the uploaded
 * ones of the same particle.  *nLeft gets the survivor count.  nprops >= 8. */
int eh_download(eh_ctx *ctx, double **coord, int64_t n, int32_t nprops, int64_t *nLeft);

/* ---- stopping conditions of the pass loop (do_tracking.c:494-520) ----
 * run_setup's stop_tracking_particle_limit (elegant.c:1131): stop before a
 * pass that would start with fewer than `limit` particles (<= 0: off).
 * Exact to the pass: a multi-pass launch that crosses the limit is rolled
 * back to its start and rerun up to the stopping pass. */
int eh_set_particle_limit(eh_ctx *ctx, int64_t limit);
/* track's interrupt_file (elegant.c:1133-1137): stop before a pass once the
 * file exists with a modification time later than it had when this was
 * called (or at all, if it did not exist).  NULL or "" turns it off.
 * Checked at pass boundaries between kernel launches (every pass when a
 * pass needs the host, else every chunk of fused passes). */
int eh_set_interrupt_file(eh_ctx *ctx, const char *path);
/* The same stop through a word the caller sets to nonzero (from another
 * thread or a signal handler); NULL turns it off.  No counterpart in the
 * reference, which polls the file only. */
int eh_set_interrupt_flag(eh_ctx *ctx, volatile const int32_t *flag);

/* do_tracking (do_tracking.c:144-3241) for a flattened beamline: upload,
 * track n_passes starting at passOffset, download.  Returns nLeft >= 0 or
 * a negative error code.  *effort accumulates survivors per pass.  accepted
 * (NULL or nOriginal rows of the same width) starts as a copy of the rows
 * (do_tracking.c:360-361) and is permuted with them, so it ends holding
 * each particle's initial row at that particle's final row. */
int64_t eh_do_tracking(eh_ctx *ctx, double **coord, int64_t nOriginal, int64_t *effort,
                       const eh_element *beamline, int64_t n_elems, double *P_central, double **accepted,
                       int64_t n_passes, int64_t passOffset);

/* ---- element-by-element entry points (host-row API of the reference) ----
 * Each uploads the rows, runs one element on the GPU and writes the rows
 * back with the reference's compaction: whole rows of the width set by
 * eh_set_particle_properties move, and `accepted` (NULL, or n rows of the
 * same width) is permuted with them, as the reference swaps accepted[i]
 * with particle[i] (multipole.c:1075-1076, limit_amplitudes.c:263-264,
 * 308-309).  The aperture is the MAXAMP in force (NULL for none).  They
 * return nLeft >= 0 or a negative error code.  Like the routines they
 * replace, they leave NaN and infinite coordinates to the caller:
 * do_tracking's culls (before the first element and
 * limit_amplitudes(DBL_MAX) after each one) belong to eh_do_tracking and
 * eh_track; eh_limit_amplitudes removes such rows at its own z. */
/* multipole_tracking2 (multipole.c:693-1117) for KQUAD/KSEXT/KOCT */
int64_t eh_multipole_tracking2(eh_ctx *ctx, double **particle, int64_t n_part, const eh_element *elem,
                               double Po, double **accepted, double z_start, const EH_MAXAMP *maxamp);
/* track_through_csbend (csbend.c:665-1409) */
int64_t eh_track_through_csbend(eh_ctx *ctx, double **part, int64_t n_part, const EH_CSBEND *csbend,
                                double Po, double **accepted, double z_start, const EH_MAXAMP *maxamp);
/* The fiducial phase that eh_do_tracking / eh_track fixed for the RFCA at
 * beamline index elem_index on its first pass (rfca->phase_fiducial once
 * fiducial_seen is set, simple_rfca.c:466-471); NaN while unset.  An
 * element-by-element caller continuing a fused run passes it on to
 * eh_simple_rf_cavity.  EH_ERR_INVALID if that element is not an RFCA of
 * the loaded beamline. */
int eh_rfca_fiducial(eh_ctx *ctx, int64_t elem_index, double *phase_fiducial);
/* simple_rf_cavity (simple_rfca.c:264-293); *phase_fiducial carries the
 * element's fiducial phase between calls (NaN = not yet set, as
 * rfca->fiducial_seen == 0).  With CHANGE_P0, *P_central becomes the mean
 * momentum of the rows left and their deltas are re-expressed against it
 * (simple_rfca.c:977-982). */
int64_t eh_simple_rf_cavity(eh_ctx *ctx, double **part, int64_t np, const EH_RFCA *rfca, double **accepted,
                            double *P_central, double zEnd, double *phase_fiducial);
/* track_particles with the order-2 drift matrix (matrix.c:294-468) */
int64_t eh_track_particles_drift(eh_ctx *ctx, double **part, int64_t n_part, const EH_DRIFT *drift);
/* limit_amplitudes (limit_amplitudes.c:193-272); openCode as EH_MAXAMP.openSide */
int64_t eh_limit_amplitudes(eh_ctx *ctx, double **coord, double xmax, double ymax, int64_t np,
                            double **accepted, double z, double Po, int32_t extrapolate_z, int32_t openCode);
/* elimit_amplitudes (limit_amplitudes.c:472-598): the (super-)elliptical
 * aperture; lost rows get delta = sqrt((Po(1+delta))^2 + 1) as there.
 * Odd or < 2 exponents give EH_ERR_INVALID (the reference exits). */
int64_t eh_elimit_amplitudes(eh_ctx *ctx, double **coord, double xmax, double ymax, int64_t np,
                             double **accepted, double z, double Po, int32_t extrapolate_z, int32_t openCode,
                             int32_t exponent, int32_t yexponent);
/* track_through_wake (wake.c:23-303); macroParticleCharge as set by CHARGE */
int64_t eh_track_through_wake(eh_ctx *ctx, double **part, int64_t np, const EH_WAKE *wake, double Po,
                              int64_t i_pass, double macroParticleCharge);
/* track_through_zlongit (zlongit.c:56-557) */
int64_t eh_track_through_zlongit(eh_ctx *ctx, double **part, int64_t np, const EH_ZLONGIT *zlongit, double Po,
                                 int64_t i_pass, double macroParticleCharge);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * eh_track_through_wake's step on device-resident particles, allocating
 * nothing on the host or the device.  coord[0..5] are the caller's device
 * arrays x, x', y, y', s, delta of np particles (all taken as alive; updated
 * in place, nothing is lost, as track_through_wake loses nothing); W_device
 * is the Green function W[wakePoints] on the device; wake->t stays a host
 * array (its spacing and span are read here) and wake->W is not used.  The
 * step's scratch is one static array per GPU (256 KiB, held until the
 * process ends): at most EH_NOALLOC_WAKE_POINTS wake points, larger tables
 * give EH_ERR_INVALID.  Calls that overlap on one GPU (from several
 * contexts or threads) share that scratch and race; each call leaves it
 * holding its own step and clears what it needs at its start.  Returns np
 * or a negative error code. */
#define EH_NOALLOC_WAKE_POINTS 16384
int64_t eh_track_through_wake_noalloc(eh_ctx *ctx, double *const coord[6], int64_t np, const EH_WAKE *wake,
                                      const double *W_device, double Po, int64_t i_pass,
                                      double macroParticleCharge);

/* ---- batched lattice variants (error-seed studies, dynamic-aperture grids) ----
 * No counterpart call in the reference, which runs each variant and each
 * aperture-scan particle through its own do_tracking call
 * (aperture_search.c:1080-1082, error.c seeds via vary_beamline).  Every
 * variant has its own beamline and particles; all are tracked together by a
 * persistent kernel over a queue of (variant, particle run) work items,
 * launched as a captured HIP graph.  Variants may not contain WAKE/ZLONGIT,
 * and RFCA elements need T_REFERENCE or FIDUCIAL=light. */
typedef struct eh_batch eh_batch;
typedef struct {
  int32_t particles_per_lane; /* 1 or 2 */
  int32_t block;              /* workgroup size: 64, 128 or 256 */
  int32_t passes_per_launch;  /* passes fused in one kernel launch */
  int32_t use_graph;          /* capture the launch sequence in a HIP graph (compact 0 and 2) */
  int32_t compact;            /* the work queue between launches: 0 fixed; 1 rebuilt by the host from the
                                 surviving particles when fewer than 3/4 of the scheduled lanes are alive
                                 (launches not captured); 2 rebuilt on the device after every launch, inside
                                 the captured graph.  With 1 and 2 launches run 1, 2, 4 ... passes up to
                                 passes_per_launch */
  int32_t pad;
  double sample_ms;           /* autotune: time of the best sample */
} eh_batch_tuning;

int eh_batch_create(eh_ctx *ctx, eh_batch **out);
int eh_batch_destroy(eh_batch *b);
/* Returns the new variant's index (>= 0) or a negative error code.  Every
 * row is tracked as its own do_tracking call would track it (the aperture
 * search's way, aperture_search.c:1080-1082): an RF cavity with the TMEAN
 * fiducial takes its phase from the first row, in row order, that reaches it
 * on the first pass when tracked alone (elegant's fiducial_seen,
 * simple_rfca.c:179-214); those rows are tried here, on the context, whose
 * program and particles they replace.  No CHANGE_P0 cavities, WAKE or
 * ZLONGIT. */
int eh_batch_add_variant(eh_batch *b, const eh_element *elems, int64_t n_elems, double Po, double **coord, int64_t n,
                         int32_t nprops);
/* time the candidate tunings on samples of the loaded batch (state restored after each) and keep the
 * fastest: particles per lane x workgroup size over sample_passes, then the work-queue mode and the
 * passes fused per launch over 4 x sample_passes on the host's wall clock */
int eh_batch_autotune(eh_batch *b, int64_t sample_passes, eh_batch_tuning *best);
int eh_batch_set_tuning(eh_batch *b, const eh_batch_tuning *t);
int eh_batch_get_tuning(const eh_batch *b, eh_batch_tuning *t);
int eh_batch_track(eh_batch *b, int64_t pass0, int64_t npasses, eh_stats *stats);
/* one variant's rows in elegant's order, column 7 = pass of loss */
int eh_batch_download(eh_batch *b, int32_t variant, double **coord, int64_t n, int32_t nprops, int64_t *nLeft);
/* pass of loss of each particle of a variant, in the order it was added (-1 = survived) */
int eh_batch_loss_passes(eh_batch *b, int32_t variant, int32_t *loss_pass, int64_t n);
int32_t eh_batch_variants(const eh_batch *b);

/* ---- dynamic-aperture surrogate network ----
 * No counterpart in the reference.  A 4-layer perceptron n_in -> hidden ->
 * hidden -> hidden -> 1 (ReLU, FP32) giving the probability that a test
 * particle survives; trained and evaluated on the GPU by hand-written
 * kernels.  Parameters are the flat float array [W1 (hidden x n_in), b1,
 * W2, b2, W3, b3, W4 (1 x hidden), b4], row-major. */
typedef struct eh_mlp eh_mlp;
int eh_mlp_create(eh_ctx *ctx, int32_t n_in, int32_t hidden, uint64_t seed, eh_mlp **out);
int eh_mlp_destroy(eh_mlp *m);
int64_t eh_mlp_num_params(const eh_mlp *m);
int eh_mlp_get_params(eh_mlp *m, float *params);
int eh_mlp_set_params(eh_mlp *m, const float *params);
/* gradient of the mean binary cross-entropy over n samples (X row-major n x n_in, y in {0,1}) */
int eh_mlp_gradient(eh_mlp *m, const float *X, const float *y, int32_t n, float *grad, float *loss);
/* Adam, minibatches of `batch`, `epochs` passes over shuffled data; losses[epochs] optional */
int eh_mlp_train(eh_mlp *m, const float *X, const float *y, int64_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float *losses);
int eh_mlp_predict(eh_mlp *m, const float *X, int64_t n, float *prob);

/* trimApertureSearchResult (trimApertureSearchResult.c:105-117): clip islands
 * from per-ray limits (modified in place) and return the trapezoid-rule area;
 * -1 for fewer than two lines. */
double eh_da_trim_area(int64_t lines, double *xLimit, double *yLimit, const double *dxFactor, const double *dyFactor,
                       int32_t fullPlane);

/* ---- seeded input generators (host only) ----
 * The uniform stream is DLARAN as elegant ships it (dlaran_OAG,
 * drand_oag.c:118-180; seeding as random_1_elegant, drand_oag.c:94-116,
 * without the toolkit's seed-bit permutation, which the reference does not
 * ship).  Gaussian deviates use the polar method (the toolkit's gauss_rn is
 * not shipped either), so only the uniform stream matches elegant's. */
typedef struct {
  int64_t seed[4];    /* 12-bit limbs of the 48-bit state, most significant first */
  int32_t have_saved; /* a second Gaussian deviate is pending */
  int32_t pad;
  double saved;
} eh_rng;
int eh_rng_seed(eh_rng *r, int64_t iseed);
double eh_rng_uniform(eh_rng *r);
double eh_gauss_rn(eh_rng *r);
/* gauss_rn_lim: mean + sigma z, z redrawn while |z| > limit (limit <= 0: no cut) */
double eh_gauss_rn_lim(eh_rng *r, double mean, double sigma, double limit);
/* perturbation() (error.c:496-534): error_type 0 uniform, 1 Gaussian with cutoff, 2 plus-or-minus */
double eh_perturbation(eh_rng *r, double amplitude, double cutoff, int32_t error_type);
/* n successive perturbation() values into out[0..n) (the calls an error
 * definition makes over n element occurrences); no counterpart of its own in
 * the reference, which calls perturbation() once per element */
int eh_perturbations(eh_rng *r, double amplitude, double cutoff, int32_t error_type, double *out, int64_t n);
/* gaussian_distribution() (generate_bunch.c:557-658, no Halton sequences) into columns offset, offset+1 */
int eh_gaussian_distribution(eh_rng *r, double **particle, int64_t n, int32_t offset, double s1, double s2,
                             int32_t symmetrize, double limit, double limit_invar, double beta, int32_t halo);
/* transform_from_normalized_coordinates (generate_bunch.c:1194-1212) */
void eh_transform_from_normalized(double **particle, int64_t n, int32_t offset, double beta, double alpha);
/* Gaussian bunch: x_plane/y_plane = {emit, beta, alpha, cutoff}, longit =
 * {sigma_s, sigma_dp, alpha_z, cutoff} (generate_bunch.c:138-160, 202-218,
 * 315-360); IDs from first_id in column 6, column 7 = -1 if nprops > 7 */
int eh_gaussian_bunch(eh_rng *r, double **particle, int64_t n, int32_t nprops, const double *x_plane,
                      const double *y_plane, const double *longit, int64_t first_id);

/* ---- device information for measurement ---- */
/* Device time of the last eh_track call's dominant kernel: number of
 * launches and summed milliseconds (HIP events on the library's stream). */
int eh_last_kernel_timing(const eh_ctx *ctx, int64_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif
