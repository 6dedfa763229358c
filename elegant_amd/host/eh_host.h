/*
 * eh_host.h - internal declarations of the C host layer.
 */
#ifndef ELEGANT_AMD_EH_HOST_H
#define ELEGANT_AMD_EH_HOST_H

#include <stdint.h>

#include "../../include/elegant_hip.h"
#include "eh_device.h"
#include "eh_program.h"

#define EH_MAX_OPS 8192
#define EH_MAX_RFCA 16
#define EH_MAX_COLL 16

typedef struct {
  eh_op ops[EH_MAX_OPS];
  int32_t nops;
  double Po;
  int32_t n_rfca, rfca_op[EH_MAX_RFCA], rfca_elem[EH_MAX_RFCA];
  int32_t n_coll, coll_op[EH_MAX_COLL], coll_elem[EH_MAX_COLL];
  int32_t n_p0, p0_op[EH_MAX_RFCA]; /* EH_OP_MATCH_P0 steps (RFCA with CHANGE_P0) */
} eh_program;

int eh_compile(const eh_element *bl, int64_t n_elems, double Po, const eh_species *sp, eh_program *prog, char *msg,
               int msglen);
const eh_species *eh_ctx_species(const eh_ctx *c);
/* change_particle's electron (elegant.c:2946-2953) */
void eh_species_electron(eh_species *sp);
double eh_element_length(const eh_element *e);
int eh_has_length(int32_t type);
void eh_rfca_set_phase(eh_op *op, const EH_RFCA *rf, double phase_fiducial);
double eh_rfca_time_offset(const eh_op *op, double Po, double s, double dp);
double eh_rfca_phase_fiducial(const EH_RFCA *rf, double t0);

/* swap-to-tail compaction of one loss event on the reference row order */
void eh_compact_event(int64_t *order, int64_t *pos, int64_t n_alive, const int64_t *lost_slots, int64_t k,
                      int64_t *scratch);

#endif
