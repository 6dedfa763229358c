/*
 * eh_runs.c - folds a flat element program into runs of repeated cells.
 *
 * A ring built from identical cells compiles to the same records over and
 * over, apart from each element's position (z_start, z_end) and its index
 * in the beamline.  The tracking kernel walks runs (eh_program.h): a body
 * of records executed reps times in a row, so one cell's records stay in
 * the scalar cache for the whole pass; positions come from a separate table
 * indexed by the flat position, so loss codes and loss coordinates are those
 * of the flat program.  This replaces nothing in the reference: its element
 * loop (do_tracking.c:816) walks the beamline's element list one element at
 * a time.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "eh_device.h"

/* the record with the fields that differ between cells cleared */
static void body_of(const eh_op *op, eh_op *out) {
  *out = *op;
  out->z_start = 0;
  out->z_end = 0;
  out->elem = 0;
}

static uint64_t fnv1a(const void *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  uint64_t h = 1469598103934665603ull;
  for (size_t k = 0; k < n; k++) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

/* ids[f]: equal for records whose bodies are byte-identical, numbered in
 * order of first appearance */
static int body_ids(const eh_op *ops, int32_t n, int32_t *ids) {
  int32_t size = 16;
  while (size < 2 * n)
    size *= 2;
  int32_t *slot = (int32_t *)malloc(sizeof(int32_t) * size); /* flat index of a record with that body */
  uint64_t *hs = (uint64_t *)malloc(sizeof(uint64_t) * size);
  eh_op a, b;
  int32_t next = 0;
  if (!slot || !hs) {
    free(slot);
    free(hs);
    return -1;
  }
  for (int32_t k = 0; k < size; k++)
    slot[k] = -1;
  for (int32_t f = 0; f < n; f++) {
    body_of(&ops[f], &a);
    const uint64_t h = fnv1a(&a, sizeof a);
    int32_t k = (int32_t)(h & (uint64_t)(size - 1));
    for (;;) {
      if (slot[k] < 0) {
        slot[k] = f;
        hs[k] = h;
        ids[f] = next++;
        break;
      }
      if (hs[k] == h) {
        body_of(&ops[slot[k]], &b);
        if (!memcmp(&a, &b, sizeof a)) {
          ids[f] = ids[slot[k]];
          break;
        }
      }
      k = (k + 1) & (size - 1);
    }
  }
  free(slot);
  free(hs);
  return 0;
}

/* Greedy left to right: at each position the repetition (a block of len
 * records followed by reps - 1 copies, len <= max_len) that covers the most
 * positions, the shortest block among equals; positions that start no
 * repetition gather into single-pass runs.  runs and body need room for n
 * entries; body[j] is the flat position whose record fills body slot j.
 * Returns the number of runs, or -1 when out of memory. */
int32_t eh_fold_runs(const eh_op *ops, int32_t n, int32_t max_len, eh_run *runs, int32_t *body, int32_t *nbody) {
  int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t nr = 0, nb = 0, lit = -1;
  if (!ids || body_ids(ops, n, ids)) {
    free(ids);
    return -1;
  }
  for (int32_t i = 0; i < n;) {
    int32_t best_len = 0, best_reps = 0;
    int64_t best_cover = 0;
    for (int32_t L = 1; L <= max_len && (int64_t)i + 2 * L <= n; L++) {
      int32_t R = 1;
      while ((int64_t)i + (int64_t)(R + 1) * L <= n && !memcmp(ids + i, ids + i + (int64_t)R * L, sizeof(int32_t) * L))
        R++;
      if (R >= 2 && (int64_t)L * R > best_cover) {
        best_cover = (int64_t)L * R;
        best_len = L;
        best_reps = R;
      }
    }
    if (best_len) {
      lit = -1;
      runs[nr].first = nb;
      runs[nr].len = best_len;
      runs[nr].reps = best_reps;
      runs[nr].flat = i;
      nr++;
      for (int32_t k = 0; k < best_len; k++)
        body[nb++] = i + k;
      i += best_len * best_reps;
    } else {
      if (lit < 0) {
        lit = nr++;
        runs[lit].first = nb;
        runs[lit].len = 0;
        runs[lit].reps = 1;
        runs[lit].flat = i;
      }
      runs[lit].len++;
      body[nb++] = i;
      i++;
    }
  }
  free(ids);
  *nbody = nb;
  return nr;
}
