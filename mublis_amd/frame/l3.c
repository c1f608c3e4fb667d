/*
 * Host C frame of the gfx950 engine: the mublis_* stride API and the generic
 * L3 driver, routed to the device through include/mublis_gfx950.h.
 *
 * Operation mapping follows the reference shims:
 *   GEMM  - transpose is a stride swap (frame/l3/gemm.c:14-24);
 *   SYRK  - C = A * A' with B the stride-swapped A, ji from uplo
 *           (frame/l3/syrk.c:19-51);
 *   SYR2K - C = beta*C + alpha*A*B' + alpha*B*A'.  The reference makes two
 *           driver passes (frame/l3/syr2k.c:47-68); here both products are
 *           the two K segments of one launch;
 *   SYMM  - the symmetric operand takes the A role (Left) or the B role
 *           (Right) of the driver (frame/l3/symm.c:36-56);
 *   TRMM  - transpose swaps strides and flips uplo (frame/l3/trmm.c:44-49);
 *           alpha == 0 zeroes B (trmm.c:138-144);
 *   TRSM  - alpha == 0 zeroes B; Right is solved as the transposed Left
 *           problem; transpose swaps strides and flips uplo
 *           (frame/l3/trsm.c:27-54).
 *
 * Staging: a pointer that is not device memory is copied to a device
 * buffer covering its whole strided footprint (so gaps between columns and
 * the unreferenced triangle travel unchanged), and outputs are copied back
 * the same way; large footprints travel through the pinned pool.  Large
 * GEMMs with C in host memory stream C by column panels instead
 * (host_gemm_ok).  Every call synchronises its stream before returning.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mublis.h"
#include "mublis_gfx950.h"
#include "frame_internal.h"

/* Library init + context validation, once per process (the reference's
 * lazy first-call contract, frame/l3/l3_driver.c:64-71). */
static int begin(void **stream) {
  int st = mublis_frame_context_ok();
  if (st) return st;
  st = mublis_rt_init();
  if (st) return st;
  *stream = mublis_rt_stream();
  return *stream ? MUBLIS_OK : MUBLIS_EHIP;
}

static int finish(int st, void *stream, stage_t *out, stage_t **tmp, int ntmp) {
  if (st == MUBLIS_OK && out) st = mublis_stage_out(out, stream);
  int sy = mublis_rt_sync(stream);
  if (st == MUBLIS_OK) st = sy;
  if (out) mublis_stage_release(out);
  for (int i = 0; i < ntmp; i++)
    if (tmp[i]) mublis_stage_release(tmp[i]);
  return st;
}

/* ------------------------------------------------------ generic driver -- */
/*
 * Driver contract (frame/l3/l3_driver.c:49-249): C := alpha * Aview * Bview +
 * beta * C on the elements with ji(j, i).  The views put the diagonal at
 * i == p for A and p == j for B whatever the shape, as packm's diagoff = ic -
 * pc / jc - pc does (l3_driver.c:133-161, frame/l1m/packm.c:41): an m x k
 * TRIANGULAR A is a trapezoid when m != k, a SYMMETRIC one reads (p, i) for
 * an unstored (i, p).  jp / pi are applied per term (term p of C(i,j) only
 * where pi(p, i) and jp(j, p) hold) by zeroing the excluded part of the
 * operand while it is materialised.  The reference skips whole cache blocks
 * instead (l3_driver.c:27-47,130,148); the two agree whenever the relation
 * excludes only structural zeros, which is every descriptor the shims build.
 * Structured or masked operands are materialised densely first (pack_dev);
 * GENERAL unmasked ones are read in place.
 */

/* Extent of the storage an operand view reads (rows x cols of its strided
 * footprint).  A SYMMETRIC view also reads reflected elements, which lie
 * past the m x k box on the stored side when m != k. */
static void view_extent(const mublis_l3_operand_t *o, int m, int k, int *rows,
                        int *cols) {
  *rows = m;
  *cols = k;
  if (o->struc == MUBLIS_STRUC_SYMMETRIC && m != k && m > 0 && k > 0) {
    const int lo = m < k ? m : k, hi = m < k ? k : m;
    if (o->uplo == MUBLIS_LOWER) { *rows = hi; *cols = lo; }
    if (o->uplo == MUBLIS_UPPER) { *rows = lo; *cols = hi; }
  }
}

static int operand_ok(const mublis_l3_operand_t *o) {
  return (unsigned)o->struc <= 2u && (unsigned)o->uplo <= 2u && (unsigned)o->diag <= 2u;
}

/* Large GEMMs with C in host memory take the panel pipeline
 * (mublis_gfx950_?gemm_host), whose transfers overlap the kernels: C
 * column-major, host operands with a unit stride, columns short enough for
 * 128-column panels of the pinned pool, at least 16 MiB of operands. */
static int host_gemm_ok(int m, int n, int k, const void *a, int rs_a, int cs_a,
                        const void *b, int rs_b, int cs_b, const void *c, int rs_c,
                        int cs_c, size_t esz) {
  const double bytes = ((double)m * k + (double)k * n + (double)m * n) * (double)esz;
  const int tall = m > k ? m : k;
  if (rs_c != 1 || cs_c < m || bytes < 16.0 * (1 << 20) ||
      (double)tall * (double)esz * 128.0 > 64.0 * (1 << 20))
    return 0;
  if (mublis_rt_is_device_ptr(c)) return 0;
  if (rs_a != 1 && cs_a != 1 && m > 1 && k > 1 && !mublis_rt_is_device_ptr(a)) return 0;
  if (rs_b != 1 && cs_b != 1 && k > 1 && n > 1 && !mublis_rt_is_device_ptr(b)) return 0;
  return 1;
}

#define DRIVER_IMPL(T, name, gemm_dev, pack_dev, gemm_host)                     \
  int name(int m, int n, int k, T alpha, const T *a, int rs_a, int cs_a,       \
           const T *b, int rs_b, int cs_b, T beta, T *c, int rs_c, int cs_c,   \
           const mublis_l3_product_t *product) {                               \
    if (m == 0 || n == 0) return 0;                                            \
    if (m < 0 || n < 0 || k < 0 || !product) return MUBLIS_EINVAL;             \
    const mublis_l3_operand_t *oa = &product->a, *ob = &product->b;            \
    const mublis_l3_domain_t *dom = &product->domain;                          \
    if (!operand_ok(oa) || !operand_ok(ob) || (unsigned)dom->jp > 2u ||        \
        (unsigned)dom->ji > 2u || (unsigned)dom->pi > 2u)                      \
      return MUBLIS_EINVAL;                                                    \
    void *s;                                                                   \
    int st = begin(&s);                                                        \
    if (st) return st;                                                         \
    const int use = alpha != (T)0 && k > 0;                                    \
    /* pi(p, i) LOWER keeps p <= i of A(i, p); jp(j, p) LOWER keeps j <= p   \
     * of B(p, j): both are "column <= row", the pack's lower mask. */        \
    const int mask_a = (int)dom->pi, mask_b = (int)dom->jp;                    \
    const int pack_a = oa->struc != MUBLIS_STRUC_GENERAL || mask_a;            \
    const int pack_b = ob->struc != MUBLIS_STRUC_GENERAL || mask_b;            \
    if (use && !pack_a && !pack_b && dom->ji == 0 && c &&                      \
        host_gemm_ok(m, n, k, a, rs_a, cs_a, b, rs_b, cs_b, c, rs_c, cs_c,     \
                     sizeof(T))) {                                             \
      st = gemm_host(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c,    \
                     rs_c, cs_c, s);                                           \
      return finish(st, s, NULL, NULL, 0);                                     \
    }                                                                          \
    int ra_rows, ra_cols, rb_rows, rb_cols;                                    \
    view_extent(oa, m, k, &ra_rows, &ra_cols);                                 \
    view_extent(ob, k, n, &rb_rows, &rb_cols);                                 \
    stage_t sa, sb, sc;                                                        \
    memset(&sa, 0, sizeof sa);                                                 \
    memset(&sb, 0, sizeof sb);                                                 \
    T *pa = NULL, *pb = NULL;                                                  \
    stage_t *tmp[2] = {&sa, &sb};                                              \
    st = mublis_stage_in(c, m, n, rs_c, cs_c, sizeof(T), 1, &sc, s);           \
    if (st == MUBLIS_OK && use)                                                \
      st = mublis_stage_in(a, ra_rows, ra_cols, rs_a, cs_a, sizeof(T), 1, &sa, s); \
    if (st == MUBLIS_OK && use) {                                              \
      if ((const void *)b == (const void *)a && sa.dev && rb_rows == ra_cols &&\
          rb_cols == ra_rows && rs_b == cs_a && cs_b == rs_a) {                \
        sb = sa;                                                               \
        sb.owned = 0; /* the same storage (SYRK passes A twice) */             \
      } else                                                                   \
        st = mublis_stage_in(b, rb_rows, rb_cols, rs_b, cs_b, sizeof(T), 1, &sb, s); \
    }                                                                          \
    const T *da = (const T *)sa.dev, *db = (const T *)sb.dev;                  \
    int64_t ra = rs_a, ca = cs_a, rb = rs_b, cb = cs_b;                        \
    if (st == MUBLIS_OK && use && pack_a) {                                    \
      pa = (T *)mublis_rt_alloc(sizeof(T) * (size_t)m * (size_t)k);            \
      if (!pa) st = MUBLIS_ENOMEM;                                             \
      else st = pack_dev(m, k, da, rs_a, cs_a, oa->struc, oa->uplo, oa->diag,  \
                         0, mask_a, pa, m, s);                                 \
      da = pa; ra = 1; ca = m;                                                 \
    }                                                                          \
    if (st == MUBLIS_OK && use && pack_b) {                                    \
      pb = (T *)mublis_rt_alloc(sizeof(T) * (size_t)k * (size_t)n);            \
      if (!pb) st = MUBLIS_ENOMEM;                                             \
      else st = pack_dev(k, n, db, rs_b, cs_b, ob->struc, ob->uplo, ob->diag,  \
                         0, mask_b, pb, k, s);                                 \
      db = pb; rb = 1; cb = k;                                                 \
    }                                                                          \
    if (st == MUBLIS_OK)                                                       \
      st = gemm_dev(m, n, use ? k : 0, use ? alpha : (T)0, da, ra, ca, db, rb, \
                    cb, NULL, 0, 0, NULL, 0, 0, beta, (T *)sc.dev, rs_c, cs_c, \
                    (int)dom->ji, s);                                          \
    st = finish(st, s, &sc, tmp, 2);                                           \
    mublis_rt_free(pa);                                                        \
    mublis_rt_free(pb);                                                        \
    return st;                                                                 \
  }

DRIVER_IMPL(float, mublis_l3_sdriver, mublis_gfx950_sgemm_dev,
            mublis_gfx950_spack_dev, mublis_gfx950_sgemm_host)
DRIVER_IMPL(double, mublis_l3_ddriver, mublis_gfx950_dgemm_dev,
            mublis_gfx950_dpack_dev, mublis_gfx950_dgemm_host)

static const mublis_l3_product_t k_general = {
    {MUBLIS_L3_ALL, MUBLIS_L3_ALL, MUBLIS_L3_ALL},
    {MUBLIS_STRUC_GENERAL, MUBLIS_DENSE, MUBLIS_PACKM_DIAG_NONUNIT},
    {MUBLIS_STRUC_GENERAL, MUBLIS_DENSE, MUBLIS_PACKM_DIAG_NONUNIT}};

#define SWAP(x, y)                                                             \
  do {                                                                         \
    int t_ = (x);                                                              \
    (x) = (y);                                                                 \
    (y) = t_;                                                                  \
  } while (0)

/* ---------------------------------------------------------------- GEMM -- */
#define GEMM_IMPL(T, name, driver)                                              \
  int name(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n, int k, \
           T alpha, const T *a, int rs_a, int cs_a, const T *b, int rs_b,      \
           int cs_b, T beta, T *c, int rs_c, int cs_c) {                       \
    if (trans_a == MUBLIS_TRANSPOSE) SWAP(rs_a, cs_a);                         \
    if (trans_b == MUBLIS_TRANSPOSE) SWAP(rs_b, cs_b);                         \
    return driver(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, \
                  cs_c, &k_general);                                           \
  }
GEMM_IMPL(float, mublis_sgemm_native, mublis_l3_sdriver)
GEMM_IMPL(double, mublis_dgemm, mublis_l3_ddriver)

/* mublis_sgemm: native FP32 MFMA, or the BF16x3 split product when selected
 * with mublis_set_sgemm_mode(1) / MUBLIS_SGEMM_MODE=bf16x3. */
int mublis_sgemm(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n, int k,
                 float alpha, const float *a, int rs_a, int cs_a, const float *b,
                 int rs_b, int cs_b, float beta, float *c, int rs_c, int cs_c) {
  if (mublis_get_sgemm_mode() == 1)
    return mublis_sgemm_bf16x3(trans_a, trans_b, m, n, k, alpha, a, rs_a, cs_a, b, rs_b,
                               cs_b, beta, c, rs_c, cs_c);
  return mublis_sgemm_native(trans_a, trans_b, m, n, k, alpha, a, rs_a, cs_a, b, rs_b,
                             cs_b, beta, c, rs_c, cs_c);
}

/* ---------------------------------------------------------------- SYMM -- */
#define SYMM_IMPL(T, name, driver)                                              \
  int name(mublis_side_t side, mublis_uplo_t uplo, int m, int n, T alpha,      \
           const T *a, int rs_a, int cs_a, const T *b, int rs_b, int cs_b,     \
           T beta, T *c, int rs_c, int cs_c) {                                 \
    mublis_l3_product_t p = k_general;                                         \
    if (side == MUBLIS_LEFT) {                                                 \
      p.a.struc = MUBLIS_STRUC_SYMMETRIC;                                      \
      p.a.uplo = uplo;                                                         \
      return driver(m, n, m, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c,     \
                    rs_c, cs_c, &p);                                           \
    }                                                                          \
    p.b.struc = MUBLIS_STRUC_SYMMETRIC;                                        \
    p.b.uplo = uplo;                                                           \
    return driver(m, n, n, alpha, b, rs_b, cs_b, a, rs_a, cs_a, beta, c, rs_c, \
                  cs_c, &p);                                                   \
  }
SYMM_IMPL(float, mublis_ssymm, mublis_l3_sdriver)
SYMM_IMPL(double, mublis_dsymm, mublis_l3_ddriver)

/* ---------------------------------------------------------------- SYRK -- */
#define SYRK_IMPL(T, name, driver)                                              \
  int name(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k, T alpha,    \
           const T *a, int rs_a, int cs_a, T beta, T *c, int rs_c, int cs_c) { \
    if (trans == MUBLIS_TRANSPOSE) SWAP(rs_a, cs_a);                           \
    mublis_l3_product_t p = k_general;                                         \
    p.domain.ji = uplo == MUBLIS_LOWER                                         \
                      ? MUBLIS_L3_LOWER                                        \
                      : (uplo == MUBLIS_UPPER ? MUBLIS_L3_UPPER : MUBLIS_L3_ALL); \
    return driver(n, n, k, alpha, a, rs_a, cs_a, a, cs_a, rs_a, beta, c, rs_c, \
                  cs_c, &p);                                                   \
  }
SYRK_IMPL(float, mublis_ssyrk, mublis_l3_sdriver)
SYRK_IMPL(double, mublis_dsyrk, mublis_l3_ddriver)

/* --------------------------------------------------------------- SYR2K -- */
#define SYR2K_IMPL(T, name, gemm_dev)                                           \
  int name(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k, T alpha,    \
           const T *a, int rs_a, int cs_a, const T *b, int rs_b, int cs_b,     \
           T beta, T *c, int rs_c, int cs_c) {                                 \
    if (n == 0) return 0;                                                      \
    if (n < 0 || k < 0) return MUBLIS_EINVAL;                                  \
    if (trans == MUBLIS_TRANSPOSE) {                                           \
      SWAP(rs_a, cs_a);                                                        \
      SWAP(rs_b, cs_b);                                                        \
    }                                                                          \
    int ji = uplo == MUBLIS_LOWER ? 1 : (uplo == MUBLIS_UPPER ? 2 : 0);        \
    void *s;                                                                   \
    int st = begin(&s);                                                        \
    if (st) return st;                                                         \
    const int use = alpha != (T)0 && k > 0;                                    \
    stage_t sa, sb, sc;                                                        \
    memset(&sa, 0, sizeof sa);                                                 \
    memset(&sb, 0, sizeof sb);                                                 \
    stage_t *tmp[2] = {&sa, &sb};                                              \
    st = mublis_stage_in(c, n, n, rs_c, cs_c, sizeof(T), 1, &sc, s);                  \
    if (st == MUBLIS_OK && use)                                                \
      st = mublis_stage_in(a, n, k, rs_a, cs_a, sizeof(T), 1, &sa, s);                \
    if (st == MUBLIS_OK && use)                                                \
      st = mublis_stage_in(b, n, k, rs_b, cs_b, sizeof(T), 1, &sb, s);                \
    const T *da = (const T *)sa.dev, *db = (const T *)sb.dev;                  \
    if (st == MUBLIS_OK)                                                       \
      st = gemm_dev(n, n, use ? k : 0, use ? alpha : (T)0, da, rs_a, cs_a, db, \
                    cs_b, rs_b, use ? db : NULL, rs_b, cs_b,                   \
                    use ? da : NULL, cs_a, rs_a, beta, (T *)sc.dev, rs_c,      \
                    cs_c, ji, s);                                              \
    return finish(st, s, &sc, tmp, 2);                                         \
  }
SYR2K_IMPL(float, mublis_ssyr2k, mublis_gfx950_sgemm_dev)
SYR2K_IMPL(double, mublis_dsyr2k, mublis_gfx950_dgemm_dev)

/* ---------------------------------------------------------------- TRMM -- */
/* B := alpha * op(A) * B (Left) or alpha * B * op(A) (Right).  The
 * triangle is materialised with zeros outside it (and 1 on a unit diagonal,
 * never read), B is copied to a dense workspace, and one GEMM writes the
 * product over B, so the in-place hazard that dictates the reference's
 * traversal order (frame/l3/trmm.c:146-160) does not arise. */
#define TRMM_IMPL(T, name, gemm_tri_dev, pack_dev, copy_dev, fill_dev)          \
  int name(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,     \
           mublis_diag_t diag, int m, int n, T alpha, const T *a, int rs_a,    \
           int cs_a, T *b, int rs_b, int cs_b) {                               \
    if (m == 0 || n == 0) return 0;                                            \
    if (m < 0 || n < 0) return MUBLIS_EINVAL;                                  \
    if (trans_a == MUBLIS_TRANSPOSE) {                                         \
      SWAP(rs_a, cs_a);                                                        \
      uplo = uplo == MUBLIS_LOWER                                              \
                 ? MUBLIS_UPPER                                                \
                 : (uplo == MUBLIS_UPPER ? MUBLIS_LOWER : MUBLIS_DENSE);       \
    }                                                                          \
    void *s;                                                                   \
    int st = begin(&s);                                                        \
    if (st) return st;                                                         \
    const int sz = side == MUBLIS_LEFT ? m : n;                                \
    stage_t sa, sb;                                                            \
    memset(&sa, 0, sizeof sa);                                                 \
    stage_t *tmp[1] = {&sa};                                                   \
    T *tri = NULL, *bw = NULL;                                                 \
    st = mublis_stage_in(b, m, n, rs_b, cs_b, sizeof(T), 1, &sb, s);                  \
    if (st == MUBLIS_OK && alpha == (T)0) {                                    \
      st = fill_dev(m, n, (T)0, (T *)sb.dev, rs_b, cs_b, s);                   \
      return finish(st, s, &sb, tmp, 0);                                       \
    }                                                                          \
    if (st == MUBLIS_OK)                                                       \
      st = mublis_stage_in(a, sz, sz, rs_a, cs_a, sizeof(T), 1, &sa, s);              \
    if (st == MUBLIS_OK) {                                                     \
      tri = (T *)mublis_rt_alloc(sizeof(T) * (size_t)sz * (size_t)sz);         \
      bw = (T *)mublis_rt_alloc(sizeof(T) * (size_t)m * (size_t)n);            \
      if (!tri || !bw) st = MUBLIS_ENOMEM;                                     \
    }                                                                          \
    if (st == MUBLIS_OK)                                                       \
      st = pack_dev(sz, sz, (const T *)sa.dev, rs_a, cs_a,                     \
                    MUBLIS_STRUC_TRIANGULAR, uplo,                             \
                    diag == MUBLIS_UNIT_DIAG ? MUBLIS_PACKM_DIAG_UNIT          \
                                             : MUBLIS_PACKM_DIAG_NONUNIT,      \
                    0, 0, tri, sz, s);                                         \
    if (st == MUBLIS_OK)                                                       \
      st = copy_dev(m, n, (const T *)sb.dev, rs_b, cs_b, bw, 1, m, s);         \
    /* K slices outside the triangle are skipped (tri codes of            \
     * include/mublis_gfx950.h); uplo DENSE multiplies the full operand. */    \
    const int tcode = uplo == MUBLIS_LOWER ? 1 : (uplo == MUBLIS_UPPER ? 2 : 0); \
    if (st == MUBLIS_OK) {                                                     \
      if (side == MUBLIS_LEFT)                                                 \
        st = gemm_tri_dev(m, n, m, alpha, tri, 1, sz, bw, 1, m, (T)0,          \
                          (T *)sb.dev, rs_b, cs_b, tcode, 0, s);               \
      else                                                                     \
        st = gemm_tri_dev(m, n, n, alpha, bw, 1, m, tri, 1, sz, (T)0,          \
                          (T *)sb.dev, rs_b, cs_b, 0, tcode, s);               \
    }                                                                          \
    st = finish(st, s, &sb, tmp, 1);                                           \
    mublis_rt_free(tri);                                                       \
    mublis_rt_free(bw);                                                        \
    return st;                                                                 \
  }
TRMM_IMPL(float, mublis_strmm, mublis_gfx950_sgemm_tri_dev, mublis_gfx950_spack_dev,
          mublis_gfx950_scopy_dev, mublis_gfx950_sfill_dev)
TRMM_IMPL(double, mublis_dtrmm, mublis_gfx950_dgemm_tri_dev, mublis_gfx950_dpack_dev,
          mublis_gfx950_dcopy_dev, mublis_gfx950_dfill_dev)

/* ---------------------------------------------------------------- TRSM -- */
#define TRSM_IMPL(T, name, trsm_dev, fill_dev)                                  \
  int name(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,     \
           mublis_diag_t diag, int m, int n, T alpha, const T *a, int rs_a,    \
           int cs_a, T *b, int rs_b, int cs_b) {                               \
    if (m == 0 || n == 0) return 0;                                            \
    if (m < 0 || n < 0) return MUBLIS_EINVAL;                                  \
    void *s;                                                                   \
    int st = begin(&s);                                                        \
    if (st) return st;                                                         \
    stage_t sa, sb;                                                            \
    memset(&sa, 0, sizeof sa);                                                 \
    stage_t *tmp[1] = {&sa};                                                   \
    st = mublis_stage_in(b, m, n, rs_b, cs_b, sizeof(T), 1, &sb, s);                  \
    if (st == MUBLIS_OK && alpha == (T)0) {                                    \
      st = fill_dev(m, n, (T)0, (T *)sb.dev, rs_b, cs_b, s);                   \
      return finish(st, s, &sb, tmp, 0);                                       \
    }                                                                          \
    const int sz = side == MUBLIS_LEFT ? m : n;                                \
    if (st == MUBLIS_OK)                                                       \
      st = mublis_stage_in(a, sz, sz, rs_a, cs_a, sizeof(T), 1, &sa, s);              \
    int64_t rb = rs_b, cb = cs_b, ra = rs_a, ca = cs_a;                        \
    int mm = m, nn = n;                                                        \
    int trans = trans_a == MUBLIS_TRANSPOSE;                                   \
    if (side == MUBLIS_RIGHT) {                                                \
      mm = n;                                                                  \
      nn = m;                                                                  \
      rb = cs_b;                                                               \
      cb = rs_b;                                                               \
      trans = !trans;                                                          \
    }                                                                          \
    mublis_uplo_t u = uplo;                                                    \
    if (trans) {                                                               \
      int64_t t_ = ra;                                                         \
      ra = ca;                                                                 \
      ca = t_;                                                                 \
      u = u == MUBLIS_LOWER ? MUBLIS_UPPER                                     \
                            : (u == MUBLIS_UPPER ? MUBLIS_LOWER : MUBLIS_DENSE); \
    }                                                                          \
    const int lower = u == MUBLIS_LOWER;                                       \
    if (st == MUBLIS_OK)                                                       \
      st = trsm_dev(lower, diag == MUBLIS_UNIT_DIAG, mm, nn, alpha,            \
                    (const T *)sa.dev, ra, ca, (T *)sb.dev, rb, cb, s);        \
    return finish(st, s, &sb, tmp, 1);                                         \
  }
TRSM_IMPL(float, mublis_strsm, mublis_gfx950_strsm_dev, mublis_gfx950_sfill_dev)
TRSM_IMPL(double, mublis_dtrsm, mublis_gfx950_dtrsm_dev, mublis_gfx950_dfill_dev)
