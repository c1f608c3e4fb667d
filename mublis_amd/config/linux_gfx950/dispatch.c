/*
 * Config linux_gfx950: select the gfx950 target when the current HIP device
 * is an MI355X-class gfx950 GPU (the reference's configs return a fixed
 * target, config/linux_avx2/dispatch.c:4-8; this one checks the hardware
 * first and reports failure otherwise).
 */
#include "mublis_gfx950.h"
#include "mublis_instantiate.h"
#include "target_registry.h"

int mublis_get_context(mublis_context_t *context) {
  if (mublis_rt_init() != MUBLIS_OK) return 1;
  *context = gfx950_context;
  return 0;
}
