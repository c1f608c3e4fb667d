/*
 * oracle.c -- CPU restatement of MuBLIS's L3 path, used ONLY as the test
 * checker (tests/, the smoke check and the CPU-baseline leg of bench.py).
 * Nothing in the product library links, loads or calls this file.
 *
 * What is restated, with the reference lines each part follows:
 *   - CBLAS argument checks and their xerbla parameter numbers, and the
 *     row-major -> column-major mapping (frame/compat/cblas/cblas.c:114-516);
 *   - the mublis_* shims: transpose = stride swap, SYRK's B = A' view, the
 *     SYR2K pair of products, SYMM left/right operand roles
 *     (frame/l3/gemm.c:14-24, syrk.c:19-51, syr2k.c:24-68, symm.c:4-56);
 *   - the generic driver's arithmetic contract: beta pre-pass (beta == 0
 *     stores zero, never multiplies), alpha == 0 exits after the pre-pass,
 *     only elements allowed by the ji relation are written
 *     (frame/l3/l3_driver.c:61-124, l3_utils.c:18-27);
 *   - operand views from the packing rules: GENERAL copies, SYMMETRIC reads
 *     the reflected element when not stored, TRIANGULAR is zero outside the
 *     stored triangle, UNIT puts 1 on the diagonal without reading it, INVERT
 *     puts 1/a there (frame/l1m/packm.c:5-84);
 *   - TRMM (frame/l3/trmm.c:41-144): transpose flips uplo, alpha == 0 zeroes B;
 *   - TRSM (frame/l3/trsm.c:24-85): alpha == 0 zeroes B, Right is solved as
 *     the transposed Left problem, transpose flips uplo, and the diagonal is
 *     used as a reciprocal that is multiplied, not divided
 *     (frame/l1m/packm.c:143-144, targets/reference/gemmtrsm_ukr.inc:29-40).
 *
 * The oracle is the mathematical definition computed with wider accumulation
 * (long double for double data, double for float data) and rounded once, so
 * that errors of the GPU path can be stated against it.  It does not copy the
 * reference's blocking: the reference's block-size dependent summation order
 * is not part of its contract (tests use normwise tolerances).
 *
 * Every function returns 0 on success, or the 1-based CBLAS parameter number
 * that the reference's cblas_xerbla would have reported.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { ROW = 101, COL = 102, NOTR = 111, TR = 112, UP = 121, LO = 122,
       NONUNIT = 131, UNIT = 132, LEFT = 141, RIGHT = 142 };

/* mublis-level enums (include/mublis.h values) */
enum { M_DENSE = 0, M_LOWER = 1, M_UPPER = 2 };
enum { S_GENERAL = 0, S_SYMMETRIC = 1, S_TRIANGULAR = 2 };
enum { D_NONUNIT = 0, D_UNIT = 1, D_INVERT = 2 };
enum { R_ALL = 0, R_LOWER = 1, R_UPPER = 2 };

static int max1(int v) { return v > 1 ? v : 1; }

/* relation_holds: frame/l3/l3_utils.c:18-27 (LOWER: lhs <= rhs). */
static int rel_holds(int rel, int lhs, int rhs) {
  if (rel == R_LOWER) return lhs <= rhs;
  if (rel == R_UPPER) return lhs >= rhs;
  return 1;
}

/* mublis_packm_is_stored: frame/l1m/packm.c:5-16 with diagoff = row - col. */
static int is_stored(int uplo, long delta) {
  if (uplo == M_LOWER) return delta >= 0;
  if (uplo == M_UPPER) return delta <= 0;
  return 1;
}

/* Element (i, p) of an operand view (frame/l1m/packm.c:40-71).  `*skip` is
 * set when the value is a structural zero, so a NaN elsewhere in storage is
 * never touched. */
#define VIEW_FN(T, W, name)                                                     \
  static W name(const T *x, long rs, long cs, long i, long p, int struc,       \
                int uplo, int diag) {                                          \
    long d = i - p;                                                            \
    if (struc == S_SYMMETRIC) {                                                \
      if (is_stored(uplo, d)) return (W)x[i * rs + p * cs];                    \
      return (W)x[p * rs + i * cs];                                            \
    }                                                                          \
    if (struc == S_TRIANGULAR) {                                               \
      if (!is_stored(uplo, d)) return (W)0;                                    \
      if (d == 0 && diag == D_UNIT) return (W)1;                               \
      if (d == 0 && diag == D_INVERT) return (W)1 / (W)x[i * rs + p * cs];     \
    }                                                                          \
    return (W)x[i * rs + p * cs];                                              \
  }
VIEW_FN(float, double, view_s)
VIEW_FN(double, long double, view_d)

/*
 * Generic driver contract (frame/l3/l3_driver.c:49-249, include/mublis.h):
 * C(i,j) := alpha * sum_p Aview(i,p) * Bview(p,j) + beta * C(i,j) for every
 * (i, j) with ji(j, i).  Operand views put the diagonal at i == p (A) and
 * p == j (B) whatever the shape: packm is called with diagoff = ic - pc and
 * jc - pc (l3_driver.c:133-161, packm.c:41), so an m x k TRIANGULAR A with
 * m != k is a trapezoid and a SYMMETRIC one reads the reflected element
 * (p, i) (packm.c:50-58).  B is packed through its transpose with flipped
 * uplo, which reads Bview(p, j) with B's own `uplo`.
 * jp and pi (include/mublis.h: "which (j, p) and (p, i) coordinates to
 * read") are applied per term: term p of C(i,j) is included only when
 * pi(p, i) and jp(j, p) hold.  The reference applies them per cache block
 * (l3_driver.c:27-47,130,148,163,171), which equals this whenever the
 * relation excludes only structural zeros (pi = uplo of a TRIANGULAR A, jp =
 * uplo of a TRIANGULAR B) -- the only descriptors whose reference result
 * does not depend on its block sizes.
 */
#define DRIVER_FN(T, W, name, view)                                             \
  int name(int m, int n, int k, T alpha, const T *a, int rs_a, int cs_a,       \
           const T *b, int rs_b, int cs_b, T beta, T *c, int rs_c, int cs_c,   \
           int jp, int ji, int pi, int a_struc, int a_uplo, int a_diag,        \
           int b_struc, int b_uplo, int b_diag) {                              \
    if (m == 0 || n == 0) return 0;                                            \
    for (long j = 0; j < n; j++)                                               \
      for (long i = 0; i < m; i++) {                                           \
        if (!rel_holds(ji, (int)j, (int)i)) continue;                          \
        T *cij = &c[i * rs_c + j * cs_c];                                      \
        W acc = 0;                                                             \
        if (alpha != 0)                                                        \
          for (long p = 0; p < k; p++) {                                       \
            if (!rel_holds(pi, (int)p, (int)i) || !rel_holds(jp, (int)j, (int)p)) \
              continue;                                                        \
            acc += view(a, rs_a, cs_a, i, p, a_struc, a_uplo, a_diag) *        \
                   view(b, rs_b, cs_b, p, j, b_struc, b_uplo, b_diag);         \
          }                                                                    \
        W prev = (beta == 0) ? (W)0 : (W)beta * (W)*cij;                        \
        *cij = (T)((W)alpha * acc + prev);                                     \
      }                                                                        \
    return 0;                                                                  \
  }
DRIVER_FN(float, double, orc_sdriver, view_s)
DRIVER_FN(double, long double, orc_ddriver, view_d)

/* ---------------------------------------------------------------- GEMM ---- */
/* frame/compat/cblas/cblas.c:114-201 and frame/l3/gemm.c:4-53 */
#define GEMM_FN(T, name, drv)                                                   \
  int name(int order, int ta, int tb, int M, int N, int K, T alpha,            \
           const T *A, int lda, const T *B, int ldb, T beta, T *C, int ldc) {  \
    if (order != ROW && order != COL) return 1;                                \
    if (ta != NOTR && ta != TR) return 2;                                      \
    if (tb != NOTR && tb != TR) return 3;                                      \
    if (M < 0) return 4;                                                       \
    if (N < 0) return 5;                                                       \
    if (K < 0) return 6;                                                       \
    int min_lda, min_ldb;                                                      \
    if (order == COL) {                                                        \
      min_lda = max1(ta == NOTR ? M : K);                                      \
      min_ldb = max1(tb == NOTR ? K : N);                                      \
    } else {                                                                   \
      min_lda = max1(ta == NOTR ? K : M);                                      \
      min_ldb = max1(tb == NOTR ? N : K);                                      \
    }                                                                          \
    if (lda < min_lda) return 9;                                               \
    if (ldb < min_ldb) return 11;                                              \
    if (ldc < max1(order == COL ? M : N)) return 14;                           \
    /* Row-major is the column-major problem C' = op(B)' op(A)'. */            \
    int m = M, n = N, t1 = ta, t2 = tb, l1 = lda, l2 = ldb;                    \
    const T *x1 = A, *x2 = B;                                                  \
    if (order == ROW) {                                                        \
      m = N; n = M; t1 = tb; t2 = ta; x1 = B; x2 = A; l1 = ldb; l2 = lda;      \
    }                                                                          \
    int rs1 = 1, cs1 = l1, rs2 = 1, cs2 = l2;                                  \
    if (t1 == TR) { rs1 = l1; cs1 = 1; }                                       \
    if (t2 == TR) { rs2 = l2; cs2 = 1; }                                       \
    return drv(m, n, K, alpha, x1, rs1, cs1, x2, rs2, cs2, beta, C, 1, ldc,   \
               R_ALL, R_ALL, R_ALL, S_GENERAL, M_DENSE, D_NONUNIT, S_GENERAL,  \
               M_DENSE, D_NONUNIT);                                            \
  }
GEMM_FN(float, orc_sgemm, orc_sdriver)
GEMM_FN(double, orc_dgemm, orc_ddriver)

/* ---------------------------------------------------------------- SYMM ---- */
/* frame/compat/cblas/cblas.c:203-278 and frame/l3/symm.c:4-56 */
#define SYMM_FN(T, name, drv)                                                   \
  int name(int order, int side, int uplo, int M, int N, T alpha, const T *A,   \
           int lda, const T *B, int ldb, T beta, T *C, int ldc) {              \
    if (order != ROW && order != COL) return 1;                                \
    if (side != LEFT && side != RIGHT) return 2;                               \
    if (uplo != UP && uplo != LO) return 3;                                    \
    if (M < 0) return 4;                                                       \
    if (N < 0) return 5;                                                       \
    if (lda < max1(side == LEFT ? M : N)) return 8;                            \
    int min_ldbc = max1(order == COL ? M : N);                                 \
    if (ldb < min_ldbc) return 10;                                             \
    if (ldc < min_ldbc) return 13;                                             \
    int m = M, n = N, sd = side, ul = uplo;                                    \
    if (order == ROW) {                                                        \
      m = N; n = M;                                                            \
      sd = side == LEFT ? RIGHT : LEFT;                                        \
      ul = uplo == UP ? LO : UP;                                               \
    }                                                                          \
    int mu = ul == UP ? M_UPPER : M_LOWER;                                     \
    if (sd == LEFT)                                                            \
      return drv(m, n, m, alpha, A, 1, lda, B, 1, ldb, beta, C, 1, ldc, R_ALL, \
                 R_ALL, R_ALL, S_SYMMETRIC, mu, D_NONUNIT, S_GENERAL, M_DENSE, \
                 D_NONUNIT);                                                   \
    return drv(m, n, n, alpha, B, 1, ldb, A, 1, lda, beta, C, 1, ldc, R_ALL,   \
               R_ALL, R_ALL, S_GENERAL, M_DENSE, D_NONUNIT, S_SYMMETRIC, mu,   \
               D_NONUNIT);                                                     \
  }
SYMM_FN(float, orc_ssymm, orc_sdriver)
SYMM_FN(double, orc_dsymm, orc_ddriver)

/* ------------------------------------------------------- SYRK / SYR2K ---- */
/* frame/compat/cblas/cblas.c:280-433, frame/l3/syrk.c:4-52, syr2k.c:3-69.
 * SYR2K's two driver passes are one exact sum here:
 * C := beta*C + alpha*(A*B' + B*A') on the ji triangle. */
#define SYRK_FN(T, W, name, two)                                                \
  int name(int order, int uplo, int trans, int N, int K, T alpha, const T *A,  \
           int lda, const T *B, int ldb, T beta, T *C, int ldc) {              \
    if (order != ROW && order != COL) return 1;                                \
    if (uplo != UP && uplo != LO) return 2;                                    \
    if (trans != NOTR && trans != TR) return 3;                                \
    if (N < 0) return 4;                                                       \
    if (K < 0) return 5;                                                       \
    int min_ld = (order == COL) ? max1(trans == NOTR ? N : K)                  \
                                : max1(trans == NOTR ? K : N);                 \
    if (lda < min_ld) return 8;                                                \
    if (two && ldb < min_ld) return 10;                                        \
    if (ldc < max1(N)) return two ? 13 : 11;                                   \
    int ul = uplo, tr = trans;                                                 \
    if (order == ROW) {                                                        \
      ul = uplo == UP ? LO : UP;                                               \
      tr = trans == NOTR ? TR : NOTR;                                          \
    }                                                                          \
    if (N == 0) return 0;                                                      \
    long rs = 1, cs = lda, rsb = 1, csb = ldb;                                 \
    if (tr == TR) { rs = lda; cs = 1; rsb = ldb; csb = 1; }                    \
    int ji = ul == LO ? R_LOWER : R_UPPER;                                     \
    for (long j = 0; j < N; j++)                                               \
      for (long i = 0; i < N; i++) {                                           \
        if (!rel_holds(ji, (int)j, (int)i)) continue;                          \
        T *cij = &C[i + j * (long)ldc];                                        \
        W acc = 0;                                                             \
        if (alpha != 0)                                                        \
          for (long p = 0; p < K; p++) {                                       \
            if (two)                                                           \
              acc += (W)A[i * rs + p * cs] * (W)B[j * rsb + p * csb] +         \
                     (W)B[i * rsb + p * csb] * (W)A[j * rs + p * cs];          \
            else                                                               \
              acc += (W)A[i * rs + p * cs] * (W)A[j * rs + p * cs];            \
          }                                                                    \
        W prev = (beta == 0) ? (W)0 : (W)beta * (W)*cij;                        \
        *cij = (T)((W)alpha * acc + prev);                                     \
      }                                                                        \
    return 0;                                                                  \
  }
SYRK_FN(float, double, orc_ssyr2k_impl, 1)
SYRK_FN(double, long double, orc_dsyr2k_impl, 1)
SYRK_FN(float, double, orc_ssyrk_impl, 0)
SYRK_FN(double, long double, orc_dsyrk_impl, 0)

int orc_ssyrk(int o, int u, int t, int N, int K, float al, const float *A,
              int lda, float be, float *C, int ldc) {
  return orc_ssyrk_impl(o, u, t, N, K, al, A, lda, A, lda, be, C, ldc);
}
int orc_dsyrk(int o, int u, int t, int N, int K, double al, const double *A,
              int lda, double be, double *C, int ldc) {
  return orc_dsyrk_impl(o, u, t, N, K, al, A, lda, A, lda, be, C, ldc);
}
int orc_ssyr2k(int o, int u, int t, int N, int K, float al, const float *A,
               int lda, const float *B, int ldb, float be, float *C, int ldc) {
  return orc_ssyr2k_impl(o, u, t, N, K, al, A, lda, B, ldb, be, C, ldc);
}
int orc_dsyr2k(int o, int u, int t, int N, int K, double al, const double *A,
               int lda, const double *B, int ldb, double be, double *C,
               int ldc) {
  return orc_dsyr2k_impl(o, u, t, N, K, al, A, lda, B, ldb, be, C, ldc);
}

/* -------------------------------------------------------- TRMM / TRSM ---- */
/* Shared argument checks and row-major mapping:
 * frame/compat/cblas/cblas.c:435-516. */
static int tr_check(int order, int side, int uplo, int ta, int diag, int M,
                    int N, int lda, int ldb) {
  if (order != ROW && order != COL) return 1;
  if (side != LEFT && side != RIGHT) return 2;
  if (uplo != UP && uplo != LO) return 3;
  if (ta != NOTR && ta != TR) return 4;
  if (diag != NONUNIT && diag != UNIT) return 5;
  if (M < 0) return 6;
  if (N < 0) return 7;
  if (lda < max1(side == LEFT ? M : N)) return 10;
  if (ldb < max1(order == COL ? M : N)) return 12;
  return 0;
}

/* Element (i, p) of op(A) for a triangular A with stored triangle `lower`. */
#define TRI_FN(T, W, name)                                                      \
  static W name(const T *A, long lda, int lower, int trans, int unit, long i,  \
                long p) {                                                      \
    long r = trans ? p : i, c = trans ? i : p;                                 \
    if (lower ? (r < c) : (r > c)) return 0;                                   \
    if (r == c && unit) return 1;                                              \
    return (W)A[r + c * lda];                                                  \
  }
TRI_FN(float, double, tri_s)
TRI_FN(double, long double, tri_d)

/* TRMM: B := alpha*op(A)*B (Left) or alpha*B*op(A) (Right), frame/l3/trmm.c. */
#define TRMM_FN(T, W, name, tri)                                                \
  int name(int order, int side, int uplo, int ta, int diag, int M, int N,      \
           T alpha, const T *A, int lda, T *B, int ldb) {                      \
    int e = tr_check(order, side, uplo, ta, diag, M, N, lda, ldb);             \
    if (e) return e;                                                           \
    int m = M, n = N, sd = side, ul = uplo;                                    \
    if (order == ROW) {                                                        \
      m = N; n = M;                                                            \
      sd = side == LEFT ? RIGHT : LEFT;                                        \
      ul = uplo == UP ? LO : UP;                                               \
    }                                                                          \
    if (m == 0 || n == 0) return 0;                                            \
    int lower = ul == LO, trans = ta == TR, unit = diag == UNIT;               \
    long L = ldb;                                                              \
    if (alpha == 0) {                                                          \
      for (long j = 0; j < n; j++)                                             \
        for (long i = 0; i < m; i++) B[i + j * L] = 0;                         \
      return 0;                                                                \
    }                                                                          \
    W *out = (W *)malloc(sizeof(W) * (size_t)m * (size_t)n);                   \
    if (!out) return -1;                                                       \
    for (long j = 0; j < n; j++)                                               \
      for (long i = 0; i < m; i++) {                                           \
        W acc = 0;                                                             \
        if (sd == LEFT) {                                                      \
          for (long p = 0; p < m; p++) {                                       \
            W v = tri(A, lda, lower, trans, unit, i, p);                       \
            if (v != 0) acc += v * (W)B[p + j * L];                            \
          }                                                                    \
        } else {                                                               \
          for (long p = 0; p < n; p++) {                                       \
            W v = tri(A, lda, lower, trans, unit, p, j);                       \
            if (v != 0) acc += (W)B[i + p * L] * v;                            \
          }                                                                    \
        }                                                                      \
        out[i + j * m] = (W)alpha * acc;                                       \
      }                                                                        \
    for (long j = 0; j < n; j++)                                               \
      for (long i = 0; i < m; i++) B[i + j * L] = (T)out[i + j * m];           \
    free(out);                                                                 \
    return 0;                                                                  \
  }
TRMM_FN(float, double, orc_strmm, tri_s)
TRMM_FN(double, long double, orc_dtrmm, tri_d)

/*
 * TRSM: X := alpha * op(A)^{-1} * B (Left) or alpha * B * op(A)^{-1} (Right).
 * Right is the transposed Left problem (frame/l3/trsm.c:35-47); each row (or
 * column) of the substitution multiplies by the reciprocal diagonal
 * (frame/l1m/packm.c:143-144, targets/reference/gemmtrsm_ukr.inc:29-40).
 */
#define TRSM_FN(T, W, name, tri)                                                \
  int name(int order, int side, int uplo, int ta, int diag, int M, int N,      \
           T alpha, const T *A, int lda, T *B, int ldb) {                      \
    int e = tr_check(order, side, uplo, ta, diag, M, N, lda, ldb);             \
    if (e) return e;                                                           \
    int m = M, n = N, sd = side, ul = uplo;                                    \
    if (order == ROW) {                                                        \
      m = N; n = M;                                                            \
      sd = side == LEFT ? RIGHT : LEFT;                                        \
      ul = uplo == UP ? LO : UP;                                               \
    }                                                                          \
    if (m == 0 || n == 0) return 0;                                            \
    long L = ldb;                                                              \
    if (alpha == 0) {                                                          \
      for (long j = 0; j < n; j++)                                             \
        for (long i = 0; i < m; i++) B[i + j * L] = 0;                         \
      return 0;                                                                \
    }                                                                          \
    int lower = ul == LO, trans = ta == TR, unit = diag == UNIT;               \
    /* Left: solve op(A) x = alpha*b per column.  Right: x op(A) = alpha*b    \
     * per row, i.e. op(A)' x' = alpha*b'.  Express both as a system          \
     * T x = r of size s with T(i,p) from tri() and the right transposition.*/ \
    int s = sd == LEFT ? m : n, cnt = sd == LEFT ? n : m;                      \
    W *x = (W *)malloc(sizeof(W) * (size_t)s);                                 \
    if (!x) return -1;                                                         \
    /* effective lower-ness of the system matrix */                            \
    int sys_trans = sd == LEFT ? trans : !trans;                               \
    int eff_lower = sys_trans ? !lower : lower;                                \
    for (long v = 0; v < cnt; v++) {                                           \
      for (long i = 0; i < s; i++)                                             \
        x[i] = (W)alpha *                                                      \
               (W)(sd == LEFT ? B[i + v * L] : B[v + i * L]);                  \
      for (long t = 0; t < s; t++) {                                           \
        long i = eff_lower ? t : s - 1 - t;                                    \
        W acc = x[i];                                                          \
        long p0 = eff_lower ? 0 : i + 1, p1 = eff_lower ? i : s;               \
        for (long p = p0; p < p1; p++) {                                       \
          W aip = tri(A, lda, lower, sys_trans, unit, i, p);                   \
          acc -= aip * x[p];                                                   \
        }                                                                      \
        W dii = tri(A, lda, lower, sys_trans, unit, i, i);                     \
        x[i] = acc * ((W)1 / dii);                                             \
      }                                                                        \
      for (long i = 0; i < s; i++) {                                           \
        if (sd == LEFT) B[i + v * L] = (T)x[i];                                \
        else B[v + i * L] = (T)x[i];                                           \
      }                                                                        \
    }                                                                          \
    free(x);                                                                   \
    return 0;                                                                  \
  }
TRSM_FN(float, double, orc_strsm, tri_s)
TRSM_FN(double, long double, orc_dtrsm, tri_d)

/* ------------------------------------------- probes for large problems ---- */
/*
 * Sampled GEMM entries: out[t] = alpha * sum_p op(A)(I[t],p) * op(B)(p,J[t])
 * + beta * C0(I[t],J[t]), column-major, accumulated in long double.  Used to
 * check n = 8192 results without an O(n^3) CPU product.
 */
void orc_dgemm_entries(int ta, int tb, int K, double alpha, const double *A,
                       int lda, const double *B, int ldb, double beta,
                       const double *C0, int ldc, int count, const int *I,
                       const int *J, double *out) {
  for (int t = 0; t < count; t++) {
    long i = I[t], j = J[t];
    long double acc = 0;
    for (long p = 0; p < K; p++) {
      long double a = ta == TR ? A[p + i * (long)lda] : A[i + p * (long)lda];
      long double b = tb == TR ? B[j + p * (long)ldb] : B[p + j * (long)ldb];
      acc += a * b;
    }
    long double prev = beta == 0 ? 0 : (long double)beta * C0[i + j * (long)ldc];
    out[t] = (double)((long double)alpha * acc + prev);
  }
}

void orc_sgemm_entries(int ta, int tb, int K, float alpha, const float *A,
                       int lda, const float *B, int ldb, float beta,
                       const float *C0, int ldc, int count, const int *I,
                       const int *J, double *out) {
  for (int t = 0; t < count; t++) {
    long i = I[t], j = J[t];
    long double acc = 0;
    for (long p = 0; p < K; p++) {
      long double a = ta == TR ? A[p + i * (long)lda] : A[i + p * (long)lda];
      long double b = tb == TR ? B[j + p * (long)ldb] : B[p + j * (long)ldb];
      acc += a * b;
    }
    long double prev = beta == 0 ? 0 : (long double)beta * C0[i + j * (long)ldc];
    out[t] = (double)((long double)alpha * acc + prev);
  }
}

/*
 * y := op(A) * x for a column-major triangular (or full, uplo == 0) m x m
 * matrix in long double; used to form TRSM residuals ||op(A) X v - alpha B v||
 * at sizes where the O(m^2 n) oracle solve is too slow.
 */
void orc_dtrmv_ld(int uplo, int ta, int diag, int m, const double *A, int lda,
                  const double *x, double *y) {
  for (long i = 0; i < m; i++) {
    long double acc = 0;
    for (long p = 0; p < m; p++) {
      long r = ta == TR ? p : i, c = ta == TR ? i : p;
      if (uplo == LO && r < c) continue;
      if (uplo == UP && r > c) continue;
      long double a = (r == c && diag == UNIT) ? 1.0L : A[r + c * (long)lda];
      acc += a * x[p];
    }
    y[i] = (double)acc;
  }
}

void orc_strmv_ld(int uplo, int ta, int diag, int m, const float *A, int lda,
                  const double *x, double *y) {
  for (long i = 0; i < m; i++) {
    long double acc = 0;
    for (long p = 0; p < m; p++) {
      long r = ta == TR ? p : i, c = ta == TR ? i : p;
      if (uplo == LO && r < c) continue;
      if (uplo == UP && r > c) continue;
      long double a = (r == c && diag == UNIT) ? 1.0L : A[r + c * (long)lda];
      acc += a * x[p];
    }
    y[i] = (double)acc;
  }
}

/* y := M * x for a general column-major m x n matrix, long double sums. */
void orc_dgemv_ld(int m, int n, const double *M, int ld, const double *x,
                  double *y) {
  for (long i = 0; i < m; i++) {
    long double acc = 0;
    for (long j = 0; j < n; j++) acc += (long double)M[i + j * (long)ld] * x[j];
    y[i] = (double)acc;
  }
}
void orc_sgemv_ld(int m, int n, const float *M, int ld, const double *x,
                  double *y) {
  for (long i = 0; i < m; i++) {
    long double acc = 0;
    for (long j = 0; j < n; j++) acc += (long double)M[i + j * (long)ld] * x[j];
    y[i] = (double)acc;
  }
}
