/*
 * eh_program.h - the element program executed by the tracking kernel.
 *
 * The C host flattens a beamline into an array of eh_op records.  Every
 * per-element quantity the reference computes inside its per-element set-up
 * (strengths KnL/n*kickFrac, drift fractions, bend radii, tan(e1), cos/sin of
 * tilts, aperture parameters, RF phase) is evaluated once on the host with
 * glibc in the reference's operation order, so the kernel only performs the
 * per-particle arithmetic.  The kernel reads records with wave-uniform
 * indices, which the compiler turns into scalar-cache loads.
 *
 * Layout: 16 int32 + 97 doubles = 840 bytes per record.
 */
#ifndef ELEGANT_AMD_EH_PROGRAM_H
#define ELEGANT_AMD_EH_PROGRAM_H

#include <stdint.h>

/* op codes */
#define EH_OP_NOP 0
#define EH_OP_DRIFT2 1     /* order-2 drift matrix (matrix.c:389-429) */
#define EH_OP_DRIFT1 2     /* order-1 drift matrix (matrix.c:430-455) */
#define EH_OP_EXACT_DRIFT 3 /* exactDrift (csbend.c:3767-3790) */
#define EH_OP_MULT 4       /* multipole_tracking2 (multipole.c:693-1117) */
#define EH_OP_CSBEND 5     /* track_through_csbend (csbend.c:665-1409) */
#define EH_OP_RFCA 6       /* trackRfCavityWithWakes, no wakes (simple_rfca.c:295-984) */
#define EH_OP_RFCA_DRIFT 7 /* RFCA with VOLT=0 (simple_rfca.c:389-402) */
#define EH_OP_APERTURE 8   /* MAXAMP: only the checks that follow every element */
#define EH_OP_WAKE 9       /* collective: handled outside the tracking kernel */
#define EH_OP_ZLONGIT 10   /* collective: handled outside the tracking kernel */
#define EH_OP_MALIGN 12    /* offset_beam (do_tracking.c:3243-3306) */
#define EH_OP_RCOL_IN 13   /* rectangular_collimator's entrance check (limit_amplitudes.c:70-110) */
#define EH_OP_RCOL_BODY 14 /* its drift and exit check (limit_amplitudes.c:111-186) */
#define EH_OP_MATRIX1 15   /* track_particles with a first-order matrix (matrix.c:430-458) */
#define EH_OP_MATCH_P0 11  /* CHANGE_P0 after an RFCA: do_match_energy on the host (do_tracking.c:3308-3420);
                              the kernel does nothing here */

/* flags */
#define EH_F_AP (1 << 0)       /* MAXAMP aperture after the element (do_tracking.c:2459-2467) */
#define EH_F_EXTRAP (1 << 1)   /* extrapolate loss position back along a drift (limit_amplitudes.c:232-248) */
#define EH_F_INAP (1 << 2)     /* aperture checked inside the integrator (multipole.c:1242-1249) */
#define EH_F_INTILT (1 << 3)   /* aperture frame rotated by -tilt (multipole.c:1789-1796) */
#define EH_F_OFFSET (1 << 4)   /* dx|dy|dz misalignment (malign_mat.c:480-500) */
#define EH_F_ROTATE (1 << 5)   /* tilt roll (tilt_matrices.c:156-188) */
#define EH_F_XKICK (1 << 6)
#define EH_F_YKICK (1 << 7)
#define EH_F_EDGE1 (1 << 8)    /* CSBEND entrance edge, edge_order 1 (csbend.c:1140-1150) */
#define EH_F_EDGE2 (1 << 9)
#define EH_F_PSI1 (1 << 10)    /* fringe angle psi1 != 0: tan evaluated per particle */
#define EH_F_PSI2 (1 << 11)
#define EH_F_KLIM1 (1 << 12)   /* edge kick limits */
#define EH_F_KLIM2 (1 << 13)
#define EH_F_EXPAND (1 << 14)  /* EXPAND_HAMILTONIAN=1 */
#define EH_F_END1FOCUS (1 << 15)
#define EH_F_END2FOCUS (1 << 16)
#define EH_F_MATRIX (1 << 17)  /* RFCA matrix method */
#define EH_F_HASLEN (1 << 18)  /* nonzero length (drift branches) */
#define EH_F_RCP (1 << 19)     /* CSBEND: rho0 and rho_actual within em_recip_ok's range (emath.h) */
#define EH_F_FAST (1 << 20)    /* CSBEND: radii, substep angles and kick lengths admit the branch-free common path */
#define EH_F_ALONE (1 << 21)   /* single-element call: no limit_amplitudes cull after the element */
#define EH_F_DFAST (1 << 22)   /* DRIF: 2^-300 <= |L| <= 2^300, the short form of the order-2 drift applies */
#define EH_F_FIELD (1 << 23)   /* CSBEND: multipole field expansion (nfield coefficients at EH_D_FIELD) */
#define EH_F_ETILT (1 << 24)   /* CSBEND: error tilt, exit offsets at EH_D_ET */
#define EH_F_CHANGE_T (1 << 25) /* RFCA CHANGE_T: EH_D_TOFF whole RF periods taken out of t and s */
#define EH_F_CHT_SELF (1 << 26) /* CHANGE_T with each particle as its own beam: the offset from its own time */
#define EH_F_LEND (1 << 27)    /* KQUAD LEFFECTIVE: exact drifts of EH_D_LEND around the body */
#define EH_F_FIRSTPASS (1 << 28) /* before the RECIRC point: tracked on pass 0 only (do_tracking.c:459-465) */

/* Aperture shapes other than the closed rectangle, packed in one int32:
 * bit 0 (super-)ellipse, bits 1-3 the open side (EH_OPEN_*), bits 8-15 the
 * x exponent, bits 16-23 the y exponent.  The checks after an element
 * follow limit_amplitudes / elimit_amplitudes (limit_amplitudes.c:193-272,
 * 472-598), the ones inside an integrator checkMultAperture
 * (multipole.c:1786-1833); 0 means the closed rectangle of the fast path. */
#define EH_AP_ELL 1
#define EH_AP_CODE(ell, open, xe, ye) ((ell) | ((open) << 1) | ((xe) << 8) | ((ye) << 16))
#define EH_AP_OPEN(code) (((code) >> 1) & 7)
#define EH_AP_XE(code) (((code) >> 8) & 0xff)
#define EH_AP_YE(code) (((code) >> 16) & 0xff)

/* loss stages recorded with the element index (loss_code = 4*op_index + stage) */
#define EH_STAGE_ELEMENT 0  /* inside the element routine's own compaction */
#define EH_STAGE_MAXAMP 1   /* limit_amplitudes with the MAXAMP aperture */
#define EH_STAGE_NAN 2      /* limit_amplitudes(DBL_MAX): NaN/Inf cull */

typedef struct {
  int32_t op, elem, flags, nslices;
  int32_t nsub, nkick_orders, order0, order1;
  int32_t order2, skew0, skew1, skew2;
  int32_t nkicks_rf, ap_code, in_code; /* ap_code, in_code: EH_AP_* shape of the two checks, 0 = rectangle */
  int32_t nfield;                       /* CSBEND: expansion order + 1 of the field sums (EH_F_FIELD) */
  /* common */
  double z_start, z_end, Po;
  double ap_xmax, ap_ymax;      /* post-element rectangle (0 = open in that plane) */
  double in_xmax, in_ymax;      /* in-element rectangle */
  double in_cs0, in_cs1;        /* cos/sin of the aperture's reverse tilt */
  /* element specific, indexed by the EH_D_* names below */
  double d[88];
} eh_op;

/* ---- MULT ---- */
#define EH_D_DX 0
#define EH_D_DY 1
#define EH_D_DZ 2
#define EH_D_COS_IN 3   /* rotation by +tilt at entry */
#define EH_D_SIN_IN 4
#define EH_D_COS_OUT 5  /* rotation by -tilt at exit */
#define EH_D_SIN_OUT 6
#define EH_D_DSH 8      /* [8]: drift * driftFrac[step] / n */
#define EH_D_KICK 16    /* [3][8]: KnL[o] / n * kickFrac[step] */
#define EH_D_XK 40      /* [8]: -(xkick/n) * kickFrac[step] */
#define EH_D_YK 48      /* [8]: -(ykick/n) * kickFrac[step] */
#define EH_D_DRIFT 56   /* drift / n (nonzero test) */
#define EH_D_TOTLEN 57  /* sum of all dsh in step order: dzLoss of a completed particle */
#define EH_D_LEND 58    /* (L - LEFFECTIVE) / 2 */

/* ---- RCOL: d[0..4] = length, x_max, y_max, dx, dy; order0 = INVERT, order1 = open-side code ---- */
#define EH_D_RC_LEN 0
#define EH_D_RC_XMAX 1
#define EH_D_RC_YMAX 2
#define EH_D_RC_DX 3
#define EH_D_RC_DY 4

/* ---- MALIGN: d[0..7] = dxp, dyp, dx, dy, dz, dt, dp, de; nkicks_rf = ON_PASS ---- */
/* EH_OP_MATRIX1: C[6] then R[6][6] row by row */
#define EH_D_M1_C 0
#define EH_D_M1_R 6
#define EH_D_MA_DXP 0
#define EH_D_MA_DYP 1
#define EH_D_MA_DX 2
#define EH_D_MA_DY 3
#define EH_D_MA_DZ 4
#define EH_D_MA_DT 5
#define EH_D_MA_DP 6
#define EH_D_MA_DE 7

/* ---- CSBEND ---- */
#define EH_D_RHO0 0
#define EH_D_RHOA 1       /* rho_actual */
#define EH_D_TANE1 2
#define EH_D_TANE2 3
#define EH_D_E1 4
#define EH_D_E2 5
#define EH_D_PSI1 6
#define EH_D_PSI2 7
#define EH_D_KLIM1 8
#define EH_D_KLIM2 9
#define EH_D_CT 10        /* cos_ttilt */
#define EH_D_ST 11        /* sin_ttilt */
#define EH_D_DXI 12
#define EH_D_DYI 13
#define EH_D_DZI 14
#define EH_D_DXF 15
#define EH_D_DYF 16
#define EH_D_DZF 17
#define EH_D_SSLICE 18    /* s = length / n */
#define EH_D_NS 19        /* n * s */
#define EH_D_BDSH 20      /* [8]: s * driftFrac[j] */
#define EH_D_BDSHR 28     /* [8]: (s * driftFrac[j]) / rho0 */
#define EH_D_BDS 36       /* [8]: s * kickFrac[j] */
#define EH_D_BKF 44       /* [8]: kickFrac[j] (0 ends the slice) */
#define EH_D_RHOA_SGN 52  /* +1 or -1: the sign of rho_actual (t / rho_actual for t = +-0 or NaN) */
#define EH_D_RRHO0 53     /* 1 / rho0 rounded to nearest: the quotients by rho0 use em_div */
#define EH_D_RRHOA 54     /* 1 / rho_actual rounded to nearest */
#define EH_D_FIELD 55     /* [<= 28]: field coefficients in the order the sums read them (eh_bend_field.h) */
#define EH_D_FIELD_MAX 28 /* expansion order 6: K1..K4 and XKICK */
#define EH_D_ET 83        /* [5]: ETILT exit offsets of x, x', y, y', s (eh_bend_etilt.h) */

/* ---- RFCA ---- */
#define EH_D_VOLT 0       /* volt / (1e6 me_mev relSign) / nKicks */
#define EH_D_OMEGA 1
#define EH_D_PHASE 2      /* phase_fiducial + PHASE*pi/180 */
#define EH_D_LENGTH 3     /* length / nKicks */
#define EH_D_DTLIGHT 4
#define EH_D_RFDX 5
#define EH_D_RFDY 6
#define EH_D_HALFLEN 7    /* length / 2 */
#define EH_D_RFLEN_TOTAL 8 /* element length (VOLT=0 drift) */
#define EH_D_TOFF 9       /* CHANGE_T timeOffset (set by the host from the beam's first row) */
#define EH_D_TO 10        /* RF period 2 pi / omega */

/* A run of the tracking kernel's program: records [first, first + len) of
 * the run table's record array, executed reps times in a row, standing for
 * the flat beamline's positions [flat, flat + len * reps).  Repeated cells
 * whose records differ only in their positions (z_start, z_end, elem) share
 * one body. */
typedef struct {
  int32_t first, len, reps, flat;
} eh_run;

/* batched variants: one element program and one particle range per variant,
 * processed as chunks pulled from a queue by a persistent kernel */
typedef struct {
  int32_t op_offset, nops;
  int64_t first, n;
} eh_vdesc;

typedef struct {
  int32_t variant, count;
  int64_t first;
} eh_chunk;

/* SoA particle state in HBM.  A particle is alive while loss_pass < 0. */
typedef struct {
  double *x, *xp, *y, *yp, *s, *dp;
  int32_t *loss_pass;
  int32_t *loss_code;
  int64_t n;
} eh_state;

#endif
