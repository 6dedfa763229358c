/*
 * ehd_internal.h - state of one device context (HIP side only).
 */
#ifndef ELEGANT_AMD_EHD_INTERNAL_H
#define ELEGANT_AMD_EHD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eh_program.h"

#define EHD_MAX_COLL 16
/* the no-allocation WAKE step: most wake points (bins) and the size of its
 * static scratch (plan 128 B + histogram and voltage for EHD_NOALLOC_BINS + 16
 * bins + 1,025 fixed-point partials, rounded up) */
#define EHD_ZL_MAX_BINS (1 << 20) /* ZLONGIT bins: the bin optimiser's default cap (ztransverse.c:672-759) */
#define EHD_NOALLOC_BINS 16384
#define EHD_NOALLOC_BYTES (256 * 1024)

typedef struct {
  double *W;          /* WAKE Green function */
  int64_t wpoints;
  double *Z, *tc, *ts; /* ZLONGIT impedance and FFT twiddles */
  int32_t nb;
} ehd_coll_slot;

struct ehd_dev {
  int device;
  hipStream_t stream;
  hipEvent_t ev0, ev1, cev0, cev1;
  double *coord[6];
  int32_t *loss_pass, *loss_code;
  int64_t n, cap;
  eh_op *ops;          /* flat records of the batch kernel's programs */
  int32_t nops, ops_cap;
  /* the tracking kernel's program (ehd_put_program) */
  eh_op *host_prog;    /* flat copy on the host: patches refold it */
  int32_t nflat, host_prog_cap;
  eh_op *tops;         /* run bodies */
  eh_run *runs;
  double *zpos;        /* (z_start, z_end) of each flat position */
  int32_t ntops, nruns, tops_cap, runs_cap, zpos_cap;
  int tshaped;         /* the tracking program has an elliptical or open-sided MAXAMP */
  int32_t block;
  unsigned long long *counter;
  double track_ms, coll_ms;
  int64_t track_launches, coll_launches;
  /* collective work buffers */
  ehd_coll_slot slot[EHD_MAX_COLL];
  void *cwork;       /* reductions, histogram, voltages */
  int32_t *bunch;    /* bunch index per particle (bunched beams) */
  int64_t bunch_cap;
  void *noalloc;     /* device address of this GPU's static no-allocation scratch */
  int64_t cwork_bytes;
  /* batched variants */
  eh_vdesc *vdesc;
  eh_chunk *chunks;
  int32_t nvar, nchunks, chunk_cap, var_cap;
  unsigned int *queue;
  int32_t *perm;     /* compacted particle indices the chunks refer to (NULL: identity) */
  int64_t perm_cap;
  int use_perm;
  int32_t *dev_nchunks; /* the chunk count the batch kernel reads (inside the queue buffer) */
  int32_t *tile_sum;    /* device compaction: tile prefix sums, then per-variant info */
  int64_t tile_cap;
  int32_t vinfo_cap;
  int shaped; /* the loaded program has an elliptical or open-sided MAXAMP */
  double *snap[6];
  int32_t *snap_lp, *snap_lc;
  int64_t snap_cap;
  int num_cu;
  char msg[512];
};

#endif
