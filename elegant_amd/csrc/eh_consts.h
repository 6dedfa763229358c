/*
 * eh_consts.h - physical constants and tracking limits used by the GPU path
 * and by the CPU checker.
 *
 * The reference takes c_mks, e_mks, me_mks, me_mev and re_mks from the SDDS
 * toolkit (not vendored; used at elegant.c:2949-2961) and PI from mdb.h.
 * No version of those headers is recorded, so this build fixes its own
 * values: CODATA 2018 for the particle constants and the double nearest pi.
 * Electrons: particleRelSign = 1, particleMassMV = me_mev (elegant.c:2946-2953);
 * other species come in at run time (eh_species).
 */
#ifndef ELEGANT_AMD_EH_CONSTS_H
#define ELEGANT_AMD_EH_CONSTS_H

#define EH_PI 3.141592653589793238462643
#define EH_PIx2 (2 * EH_PI)
#define EH_PIo2 (EH_PI / 2)
#define EH_C_MKS 2.99792458e8
#define EH_E_MKS 1.602176634e-19
#define EH_ME_MKS 9.1093837015e-31
#define EH_ME_MEV 0.51099895000
#define EH_RE_MKS 2.8179403262e-15
#define EH_REL_SIGN 1.0
/* vacuum permittivity, CODATA 2018 (the toolkit's epsilon_o) */
#define EH_EPSILON_O 8.8541878128e-12
/* change_particle's mass table (elegant.c:2930-2931) */
#define EH_MP_MKS 1.6726485e-27
#define EH_MMU_MKS 1.88353109e-28
/* COORD_LIMIT and SLOPE_LIMIT (track.h:1204-1205, track_data.c:51-52) */
#define EH_COORD_LIMIT 10.0
#define EH_SLOPE_LIMIT 1.0

#endif
