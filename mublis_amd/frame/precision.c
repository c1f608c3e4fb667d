/*
 * Host entry points of the precision extensions (include/mublis_gfx950.h):
 * BF16x3 split SGEMM and FP64-refined STRSM, with the staging and argument
 * mapping of the regular mublis_* routines (frame/l3/gemm.c:14-24 for the
 * transposes, frame/l3/trsm.c:27-54 for TRSM's alpha == 0, Right and
 * transpose handling).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "frame_internal.h"
#include "mublis.h"
#include "mublis_gfx950.h"

static pthread_once_t g_mode_once = PTHREAD_ONCE_INIT;
static volatile int g_sgemm_mode = 0;

static void read_mode(void) {
  const char *e = getenv("MUBLIS_SGEMM_MODE");
  if (e && strcmp(e, "bf16x3") == 0) g_sgemm_mode = 1;
}

int mublis_get_sgemm_mode(void) {
  pthread_once(&g_mode_once, read_mode);
  return g_sgemm_mode;
}

int mublis_set_sgemm_mode(int mode) {
  pthread_once(&g_mode_once, read_mode);
  if (mode != 0 && mode != 1) return -1;
  int old = g_sgemm_mode;
  g_sgemm_mode = mode;
  return old;
}

static int begin(void **stream) {
  int st = mublis_frame_context_ok();
  if (st) return st;
  if ((st = mublis_rt_init())) return st;
  *stream = mublis_rt_stream();
  return *stream ? MUBLIS_OK : MUBLIS_EHIP;
}

int mublis_sgemm_bf16x3(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n,
                        int k, float alpha, const float *a, int rs_a, int cs_a,
                        const float *b, int rs_b, int cs_b, float beta, float *c,
                        int rs_c, int cs_c) {
  if (m == 0 || n == 0) return 0;
  if (m < 0 || n < 0 || k < 0) return MUBLIS_EINVAL;
  if (trans_a == MUBLIS_TRANSPOSE) { int t = rs_a; rs_a = cs_a; cs_a = t; }
  if (trans_b == MUBLIS_TRANSPOSE) { int t = rs_b; rs_b = cs_b; cs_b = t; }
  void *s;
  int st = begin(&s);
  if (st) return st;
  const int use = alpha != 0.0f && k > 0;
  stage_t sa, sb, sc;
  memset(&sa, 0, sizeof sa);
  memset(&sb, 0, sizeof sb);
  st = mublis_stage_in(c, m, n, rs_c, cs_c, sizeof(float), 1, &sc, s);
  if (!st && use) st = mublis_stage_in(a, m, k, rs_a, cs_a, sizeof(float), 1, &sa, s);
  if (!st && use) st = mublis_stage_in(b, k, n, rs_b, cs_b, sizeof(float), 1, &sb, s);
  if (!st)
    st = mublis_gfx950_sgemm_bf16x3_dev(m, n, use ? k : 0, use ? alpha : 0.0f,
                                        (const float *)sa.dev, rs_a, cs_a,
                                        (const float *)sb.dev, rs_b, cs_b, beta,
                                        (float *)sc.dev, rs_c, cs_c, s);
  if (!st) st = mublis_stage_out(&sc, s);
  int sy = mublis_rt_sync(s);
  if (!st) st = sy;
  mublis_stage_release(&sa);
  mublis_stage_release(&sb);
  mublis_stage_release(&sc);
  return st;
}

int mublis_strsm_refined(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                         mublis_diag_t diag, int m, int n, float alpha, const float *a,
                         int rs_a, int cs_a, float *b, int rs_b, int cs_b, int sweeps) {
  if (m == 0 || n == 0) return 0;
  if (m < 0 || n < 0 || sweeps < 0) return MUBLIS_EINVAL;
  void *s;
  int st = begin(&s);
  if (st) return st;
  stage_t sa, sb;
  memset(&sa, 0, sizeof sa);
  st = mublis_stage_in(b, m, n, rs_b, cs_b, sizeof(float), 1, &sb, s);
  const int sz = side == MUBLIS_LEFT ? m : n;
  if (!st && alpha != 0.0f)
    st = mublis_stage_in(a, sz, sz, rs_a, cs_a, sizeof(float), 1, &sa, s);
  int64_t rb = rs_b, cb = cs_b, ra = rs_a, ca = cs_a;
  int mm = m, nn = n, trans = trans_a == MUBLIS_TRANSPOSE;
  if (side == MUBLIS_RIGHT) {
    mm = n; nn = m; rb = cs_b; cb = rs_b; trans = !trans;
  }
  mublis_uplo_t u = uplo;
  if (trans) {
    int64_t t = ra; ra = ca; ca = t;
    u = u == MUBLIS_LOWER ? MUBLIS_UPPER : (u == MUBLIS_UPPER ? MUBLIS_LOWER : MUBLIS_DENSE);
  }
  if (!st)
    st = mublis_gfx950_strsm_refined_dev(u == MUBLIS_LOWER, diag == MUBLIS_UNIT_DIAG, mm, nn,
                                         alpha, (const float *)sa.dev, ra, ca,
                                         (float *)sb.dev, rb, cb, sweeps, s);
  if (!st) st = mublis_stage_out(&sb, s);
  int sy = mublis_rt_sync(s);
  if (!st) st = sy;
  mublis_stage_release(&sa);
  mublis_stage_release(&sb);
  return st;
}
