/* Internal declarations shared by the host frame's translation units. */
#ifndef MUBLIS_FRAME_INTERNAL_H
#define MUBLIS_FRAME_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "mublis.h"
#include "mublis_instantiate.h"

#pragma GCC visibility push(hidden)

/* 0 once the config's context has been fetched and validated (cached for
 * the process, failure included); non-zero otherwise. */
int mublis_frame_context_ok(void);

/* The validated context (NULL before a successful mublis_frame_context_ok). */
const mublis_context_t *mublis_frame_context(void);

#include <stddef.h>
#include <stdint.h>

typedef struct {
  const void *host; /* caller's pointer */
  void *dev;        /* device pointer used by kernels */
  size_t bytes;     /* footprint in bytes */
  int owned;        /* dev is a staging buffer owned by the call */
  size_t esz;       /* element size of an owned buffer */
} stage_t;

/* Stage an m x n operand with strides (rs, cs); copy == 0 skips the H2D. */
int mublis_stage_in(const void *p, int m, int n, int64_t rs, int64_t cs,
                    size_t esz, int copy, stage_t *st, void *stream);
int mublis_stage_out(stage_t *st, void *stream);
void mublis_stage_release(stage_t *st);

/* The native FP32 MFMA SGEMM (mublis_sgemm dispatches to it by default). */
int mublis_sgemm_native(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n,
                        int k, float alpha, const float *a, int rs_a, int cs_a,
                        const float *b, int rs_b, int cs_b, float beta, float *c,
                        int rs_c, int cs_c);

#pragma GCC visibility pop

#endif
