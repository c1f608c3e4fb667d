/*
 * The gfx950 target's plugin context (include/mublis_instantiate.h:125-247).
 *
 * Block sizes describe the device tiling:
 *   mr x nr = 64 x 64   one wavefront's MFMA tile (4 x 4 tiles of 16 x 16);
 *   mc x nc = 128 x 128 one workgroup's output tile (2 x 2 wavefronts);
 *   kc      = 64        the TRSM diagonal block (inverted in FP64 as a whole).
 *
 * The micro-kernels honour the reference's packed-panel contract
 * (include/mublis_instantiate.h:34-111) by running the tile on the GPU: the
 * packed panels are staged to device memory if needed, the MR x NR tile is
 * computed by the same MFMA GEMM kernel the driver launches, and the
 * GEMMTRSM solve by the TRSM schedule's FP64 diagonal-block inversion and
 * inverse multiply.
 * The frame itself never calls them per tile; they exist so that code
 * written against the plugin seam works with this target.
 */
#include <stdint.h>

#include "frame_internal.h"
#include "mublis_gfx950.h"
#include "mublis_instantiate.h"

#define MR 64
#define NR 64


#define GEMM_UKR(T, name, gemm_dev)                                             \
  static void name(int k, T alpha, const T *MUBLIS_RESTRICT a,                 \
                   const T *MUBLIS_RESTRICT b, T beta, T *MUBLIS_RESTRICT c,   \
                   int rs_c, int cs_c, const mublis_auxinfo_t *aux) {          \
    (void)aux;                                                                 \
    if (mublis_rt_init()) return;                                              \
    void *s = mublis_rt_stream();                                              \
    stage_t sa, sb, sc;                                                        \
    int kk = k > 0 ? k : 0;                                                    \
    int st = mublis_stage_in(a, MR, kk, 1, MR, sizeof(T), 1, &sa, s);          \
    if (!st) st = mublis_stage_in(b, kk, NR, NR, 1, sizeof(T), 1, &sb, s);     \
    if (!st) st = mublis_stage_in(c, MR, NR, rs_c, cs_c, sizeof(T), 1, &sc, s);\
    if (!st)                                                                   \
      st = gemm_dev(MR, NR, kk, alpha, (const T *)sa.dev, 1, MR,               \
                    (const T *)sb.dev, NR, 1, (const T *)0, 0, 0,              \
                    (const T *)0, 0, 0, beta, (T *)sc.dev, rs_c, cs_c, 0, s);  \
    if (!st) st = mublis_stage_out(&sc, s);                                    \
    mublis_rt_sync(s);                                                         \
    mublis_stage_release(&sa);                                                 \
    mublis_stage_release(&sb);                                                 \
    mublis_stage_release(&sc);                                                 \
  }
GEMM_UKR(float, sgemm_ukr_gfx950, mublis_gfx950_sgemm_dev)
GEMM_UKR(double, dgemm_ukr_gfx950, mublis_gfx950_dgemm_dev)

/* R := alpha*B11 - A1x*Bx1 on the MFMA GEMM kernel, then the FP64 solve
 * with the pre-inverted diagonal, then C11 := X. */
#define GEMMTRSM_UKR(T, name, lower, gemm_dev, solve_dev, copy_dev)             \
  static void name(int k, T alpha, const T *MUBLIS_RESTRICT a1x,              \
                   const T *MUBLIS_RESTRICT a11, const T *MUBLIS_RESTRICT bx1, \
                   T *MUBLIS_RESTRICT b11, T *MUBLIS_RESTRICT c11, int rs_c,   \
                   int cs_c, const mublis_auxinfo_t *aux) {                    \
    (void)aux;                                                                 \
    if (mublis_rt_init()) return;                                              \
    void *s = mublis_rt_stream();                                              \
    stage_t sx, s11, sbx, sb, sc;                                              \
    int kk = k > 0 ? k : 0;                                                    \
    int st = mublis_stage_in(a1x, MR, kk, 1, MR, sizeof(T), 1, &sx, s);        \
    if (!st) st = mublis_stage_in(a11, MR, MR, 1, MR, sizeof(T), 1, &s11, s);  \
    if (!st) st = mublis_stage_in(bx1, kk, NR, NR, 1, sizeof(T), 1, &sbx, s);  \
    if (!st) st = mublis_stage_in(b11, MR, NR, NR, 1, sizeof(T), 1, &sb, s);   \
    if (!st)                                                                   \
      st = mublis_stage_in(c11, MR, NR, rs_c, cs_c, sizeof(T), 1, &sc, s);     \
    if (!st)                                                                   \
      st = gemm_dev(MR, NR, kk, (T)-1, (const T *)sx.dev, 1, MR,               \
                    (const T *)sbx.dev, NR, 1, (const T *)0, 0, 0,             \
                    (const T *)0, 0, 0, alpha, (T *)sb.dev, NR, 1, 0, s);      \
    if (!st)                                                                   \
      st = solve_dev(lower, MR, NR, (const T *)s11.dev, 1, MR, (T *)sb.dev,    \
                     NR, 1, s);                                                \
    if (!st)                                                                   \
      st = copy_dev(MR, NR, (const T *)sb.dev, NR, 1, (T *)sc.dev, rs_c, cs_c, \
                    s);                                                        \
    if (!st) st = mublis_stage_out(&sb, s);                                    \
    if (!st) st = mublis_stage_out(&sc, s);                                    \
    mublis_rt_sync(s);                                                         \
    mublis_stage_release(&sx);                                                 \
    mublis_stage_release(&s11);                                                \
    mublis_stage_release(&sbx);                                                \
    mublis_stage_release(&sb);                                                 \
    mublis_stage_release(&sc);                                                 \
  }
GEMMTRSM_UKR(float, sgemmtrsml_ukr_gfx950, 1, mublis_gfx950_sgemm_dev,
             mublis_gfx950_strsm_inverted_dev, mublis_gfx950_scopy_dev)
GEMMTRSM_UKR(float, sgemmtrsmu_ukr_gfx950, 0, mublis_gfx950_sgemm_dev,
             mublis_gfx950_strsm_inverted_dev, mublis_gfx950_scopy_dev)
GEMMTRSM_UKR(double, dgemmtrsml_ukr_gfx950, 1, mublis_gfx950_dgemm_dev,
             mublis_gfx950_dtrsm_inverted_dev, mublis_gfx950_dcopy_dev)
GEMMTRSM_UKR(double, dgemmtrsmu_ukr_gfx950, 0, mublis_gfx950_dgemm_dev,
             mublis_gfx950_dtrsm_inverted_dev, mublis_gfx950_dcopy_dev)

MUBLIS_CONTEXT_STAMP(gfx950_context,
                     sgemm_ukr_gfx950, sgemmtrsml_ukr_gfx950,
                     sgemmtrsmu_ukr_gfx950, MR, NR, 128, 64, 128,
                     dgemm_ukr_gfx950, dgemmtrsml_ukr_gfx950,
                     dgemmtrsmu_ukr_gfx950, MR, NR, 128, 64, 128)
