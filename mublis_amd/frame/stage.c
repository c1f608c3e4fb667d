/*
 * Staging of caller operands for the device: a pointer that is not device
 * memory is copied into a device buffer covering its whole strided
 * footprint, and copied back the same way, so column gaps and unreferenced
 * triangles round-trip unchanged.  Device pointers are used in place.
 */
#include <stdint.h>
#include <string.h>

#include "frame_internal.h"
#include "mublis_gfx950.h"

/* Footprints from this size travel through the pinned pool
 * (mublis_rt_h2d_2d / _d2h_2d); smaller ones as one pageable copy. */
#define PINNED_MIN ((size_t)4 << 20)

/* Footprint, in elements, of an m x n view with strides (rs, cs). */
static int footprint(int m, int n, int64_t rs, int64_t cs, int64_t *elems) {
  if (m <= 0 || n <= 0) {
    *elems = 0;
    return 0;
  }
  if ((m > 1 && rs < 0) || (n > 1 && cs < 0)) return MUBLIS_EINVAL;
  *elems = (int64_t)(m - 1) * (m > 1 ? rs : 0) +
           (int64_t)(n - 1) * (n > 1 ? cs : 0) + 1;
  return 0;
}

int mublis_stage_in(const void *p, int m, int n, int64_t rs, int64_t cs,
                    size_t esz, int copy, stage_t *st, void *stream) {
  int64_t elems;
  memset(st, 0, sizeof *st);
  st->host = p;
  if (footprint(m, n, rs, cs, &elems)) return MUBLIS_EINVAL;
  st->bytes = (size_t)elems * esz;
  if (!p || elems == 0) return MUBLIS_OK;
  if (mublis_rt_is_device_ptr(p)) {
    st->dev = (void *)p;
    return MUBLIS_OK;
  }
  st->dev = mublis_rt_alloc(st->bytes);
  if (!st->dev) return MUBLIS_ENOMEM;
  st->owned = 1;
  st->esz = esz;
  if (!copy) return MUBLIS_OK;
  if (st->bytes >= PINNED_MIN)
    return mublis_rt_h2d_2d(st->dev, p, elems, 1, elems, esz, stream);
  return mublis_rt_h2d(st->dev, p, st->bytes, stream);
}

int mublis_stage_out(stage_t *st, void *stream) {
  if (!st->owned) return MUBLIS_OK;
  if (st->bytes >= PINNED_MIN) {
    const int64_t elems = (int64_t)(st->bytes / st->esz);
    return mublis_rt_d2h_2d((void *)st->host, st->dev, elems, 1, elems, st->esz, stream);
  }
  return mublis_rt_d2h((void *)st->host, st->dev, st->bytes, stream);
}

void mublis_stage_release(stage_t *st) {
  if (st->owned) mublis_rt_free(st->dev);
  st->owned = 0;
}

