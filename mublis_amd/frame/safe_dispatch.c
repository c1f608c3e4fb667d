/*
 * Fetch the config's context once per process and validate it, caching the
 * result (success or failure) -- the contract of the reference's
 * mublis_get_safe_context (frame/base/safe_dispatch.c:10-71): every
 * micro-kernel pointer non-NULL, every block size positive, mc % mr,
 * nc % nr, kc % mr and kc % nr all zero.
 */
#include <pthread.h>

#include "frame_internal.h"
#include "mublis_instantiate.h"

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static mublis_context_t g_context;
static int g_status = 1;

#define SIZES_OK(x)                                                            \
  ((x).mr > 0 && (x).nr > 0 && (x).mc > 0 && (x).kc > 0 && (x).nc > 0 &&       \
   (x).mc % (x).mr == 0 && (x).nc % (x).nr == 0 && (x).kc % (x).mr == 0 &&     \
   (x).kc % (x).nr == 0)
#define KERNELS_OK(x) ((x).gemm_ukr && (x).gemmtrsml_ukr && (x).gemmtrsmu_ukr)

static void fetch(void) {
  if (mublis_get_context(&g_context) != 0) {
    g_status = 1;
    return;
  }
  g_status = (KERNELS_OK(g_context.s) && KERNELS_OK(g_context.d) &&
              SIZES_OK(g_context.s) && SIZES_OK(g_context.d))
                 ? 0
                 : 1;
}

int mublis_frame_context_ok(void) {
  pthread_once(&g_once, fetch);
  return g_status;
}

const mublis_context_t *mublis_frame_context(void) {
  return g_status == 0 ? &g_context : 0;
}
