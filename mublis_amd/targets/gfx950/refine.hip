// STRSM with FP64-residual iterative refinement.
//
//   X0 := FP32 TRSM of alpha*B (the regular device schedule);
//   repeat `sweeps` times:
//     R  := alpha*B - op(A)*X     in FP64 (A, X and B are exact in FP64; the
//                                 product skips the zero triangle's K slices)
//     D  := FP32 TRSM of R
//     X  := X + D                 in FP64, and the FP32 result is rounded from it.
// FP64 is used only for the diagonal-block solves inside the TRSM schedule,
// the residual and the update; the O(m^2 n) solves run in FP32.  The
// refinement removes the growth of the FP32 solve's error with the matrix's
// conditioning, leaving the final FP32 rounding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gfx950_common.h"
#include "mublis_gfx950.h"

extern "C" int mublis_gfx950_strsm_refined_dev(int lower, int unit, int m, int n,
                                               float alpha, const float *a, int64_t rs_a,
                                               int64_t cs_a, float *b, int64_t rs_b,
                                               int64_t cs_b, int sweeps, void *stream) {
  if (m < 0 || n < 0 || sweeps < 0) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (alpha == 0.0f || sweeps == 0)
    return mublis_gfx950_strsm_dev(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                                   s);
  const size_t mn = (size_t)m * n;
  double *t64 = (double *)mublis_rt_alloc(sizeof(double) * (size_t)m * m);
  double *b64 = (double *)mublis_rt_alloc(sizeof(double) * mn);
  double *x64 = (double *)mublis_rt_alloc(sizeof(double) * mn);
  double *r64 = (double *)mublis_rt_alloc(sizeof(double) * mn);
  float *r32 = (float *)mublis_rt_alloc(sizeof(float) * mn);
  // one TRSM workspace for the three FP32 solves: the triangle's inverted
  // diagonal blocks are computed by the first and reused by the others
  const size_t tws_bytes = gfx950_strsm_keep_ws_bytes(m, n);
  void *tws = mublis_rt_alloc(tws_bytes);
  bool inv_ready = false;
  int st = (t64 && b64 && x64 && r64 && r32 && tws) ? MUBLIS_OK : MUBLIS_ENOMEM;
  // the right-hand side and the triangle, exactly, in FP64
  if (st == MUBLIS_OK) st = gfx950_convert_f2d(m, n, b, rs_b, cs_b, b64, 1, m, s);
  if (st == MUBLIS_OK)
    st = gfx950_pack_f2d(m, m, a, rs_a, cs_a, 2 /*TRIANGULAR*/, lower ? 1 : 2,
                         unit ? 1 : 0, t64, m, s);
  // X0 in FP32, in place
  if (st == MUBLIS_OK)
    st = gfx950_strsm_keep(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, tws,
                           tws_bytes, &inv_ready, s);
  if (st == MUBLIS_OK) st = gfx950_convert_f2d(m, n, b, rs_b, cs_b, x64, 1, m, s);
  for (int sw = 0; sw < sweeps && st == MUBLIS_OK; ++sw) {
    st = gfx950_copy<double>(m, n, b64, 1, m, r64, 1, m, s);
    if (st == MUBLIS_OK)
      st = mublis_gfx950_dgemm_tri_dev(m, n, m, -1.0, t64, 1, m, x64, 1, m, (double)alpha,
                                       r64, 1, m, lower ? 1 : 2, 0, s);
    if (st == MUBLIS_OK) st = gfx950_convert_d2f(m, n, r64, 1, m, r32, 1, m, s);
    if (st == MUBLIS_OK)
      st = gfx950_strsm_keep(lower, unit, m, n, 1.0f, a, rs_a, cs_a, r32, 1, m, tws,
                             tws_bytes, &inv_ready, s);
    if (st == MUBLIS_OK) st = gfx950_refine_update(m, n, r32, x64, b, rs_b, cs_b, s);
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  mublis_rt_free(t64);
  mublis_rt_free(b64);
  mublis_rt_free(x64);
  mublis_rt_free(r64);
  mublis_rt_free(r32);
  mublis_rt_free(tws);
  return st;
}
