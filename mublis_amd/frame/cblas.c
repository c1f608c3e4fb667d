/*
 * CBLAS Level-3 front end for the gfx950 engine.
 *
 * Argument checks report the same 1-based parameter numbers as the
 * reference (frame/compat/cblas/cblas.c:130-171, 218-249, 294-325, 369-404,
 * 450-485).  Column-major calls pass straight through to the mublis_* stride
 * API with (rs, cs) = (1, ld).  A row-major matrix is the transpose of a
 * column-major one with the same ld, so row-major calls become the
 * transposed column-major problem: GEMM computes C' = op(B)' op(A)'; SYMM,
 * TRMM and TRSM flip side and uplo and swap M and N; SYRK and SYR2K flip
 * uplo and trans (cblas.c:188-200, 266-277, 341-351, 421-432, 503-515).
 * A non-zero backend status is reported as cblas_xerbla(0, ...).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include "cblas.h"
#include "mublis.h"

/* Weak default handler: print and exit, as the reference does
 * (frame/compat/cblas/cblas.c:14-32).  Applications may link their own. */
__attribute__((weak)) void cblas_xerbla(int p, const char *rout,
                                        const char *form, ...) {
  va_list ap;
  if (p > 0)
    fprintf(stderr, "Parameter %d to routine %s was incorrect\n", p, rout);
  else
    fprintf(stderr, "Routine %s failed\n", rout);
  va_start(ap, form);
  vfprintf(stderr, form, ap);
  va_end(ap);
  exit(EXIT_FAILURE);
}

static int at_least1(int v) { return v < 1 ? 1 : v; }
static int ok_order(int v) { return v == CblasRowMajor || v == CblasColMajor; }
static int ok_trans(int v) { return v == CblasNoTrans || v == CblasTrans; }
static int ok_uplo(int v) { return v == CblasUpper || v == CblasLower; }
static int ok_diag(int v) { return v == CblasNonUnit || v == CblasUnit; }
static int ok_side(int v) { return v == CblasLeft || v == CblasRight; }

static mublis_trans_t to_trans(int t, int flip) {
  return (t == CblasTrans) != (flip != 0) ? MUBLIS_TRANSPOSE
                                          : MUBLIS_NO_TRANSPOSE;
}
static mublis_uplo_t to_uplo(int u, int flip) {
  return (u == CblasUpper) != (flip != 0) ? MUBLIS_UPPER : MUBLIS_LOWER;
}
static mublis_side_t to_side(int s, int flip) {
  return (s == CblasLeft) != (flip != 0) ? MUBLIS_LEFT : MUBLIS_RIGHT;
}

/* Report argument p and leave the routine. */
#define REQUIRE(cond, p, ...)                                                  \
  do {                                                                         \
    if (!(cond)) {                                                             \
      cblas_xerbla((p), routine, __VA_ARGS__);                                 \
      return;                                                                  \
    }                                                                          \
  } while (0)

#define BACKEND(call)                                                          \
  do {                                                                         \
    int status_ = (call);                                                      \
    if (status_ != 0)                                                          \
      cblas_xerbla(0, routine, "MuBLIS backend returned status %d\n",          \
                   status_);                                                   \
  } while (0)

/* ---------------------------------------------------------------- GEMM -- */
#define DEF_GEMM(pfx, T)                                                       \
  void cblas_##pfx##gemm(const enum CBLAS_ORDER Order,                         \
                         const enum CBLAS_TRANSPOSE TransA,                    \
                         const enum CBLAS_TRANSPOSE TransB, const int M,       \
                         const int N, const int K, const T alpha, const T *A,  \
                         const int lda, const T *B, const int ldb,             \
                         const T beta, T *C, const int ldc) {                  \
    const char *routine = "cblas_" #pfx "gemm";                                \
    const int col = Order == CblasColMajor;                                    \
    REQUIRE(ok_order(Order), 1, "Illegal Order setting, %d\n", Order);         \
    REQUIRE(ok_trans(TransA), 2, "Illegal TransA setting, %d\n", TransA);      \
    REQUIRE(ok_trans(TransB), 3, "Illegal TransB setting, %d\n", TransB);      \
    REQUIRE(M >= 0, 4, "M must be nonnegative, %d\n", M);                      \
    REQUIRE(N >= 0, 5, "N must be nonnegative, %d\n", N);                      \
    REQUIRE(K >= 0, 6, "K must be nonnegative, %d\n", K);                      \
    /* rows of the stored A and B in column-major terms */                     \
    const int ra = col ? (TransA == CblasNoTrans ? M : K)                      \
                       : (TransA == CblasNoTrans ? K : M);                     \
    const int rb = col ? (TransB == CblasNoTrans ? K : N)                      \
                       : (TransB == CblasNoTrans ? N : K);                     \
    REQUIRE(lda >= at_least1(ra), 9, "lda must be at least %d, %d\n",         \
            at_least1(ra), lda);                                               \
    REQUIRE(ldb >= at_least1(rb), 11, "ldb must be at least %d, %d\n",        \
            at_least1(rb), ldb);                                               \
    REQUIRE(ldc >= at_least1(col ? M : N), 14, "ldc is too small, %d\n", ldc); \
    if (col)                                                                   \
      BACKEND(mublis_##pfx##gemm(to_trans(TransA, 0), to_trans(TransB, 0), M,  \
                                 N, K, alpha, A, 1, lda, B, 1, ldb, beta, C,   \
                                 1, ldc));                                     \
    else                                                                       \
      BACKEND(mublis_##pfx##gemm(to_trans(TransB, 0), to_trans(TransA, 0), N,  \
                                 M, K, alpha, B, 1, ldb, A, 1, lda, beta, C,   \
                                 1, ldc));                                     \
  }
DEF_GEMM(s, float)
DEF_GEMM(d, double)

/* ---------------------------------------------------------------- SYMM -- */
#define DEF_SYMM(pfx, T)                                                       \
  void cblas_##pfx##symm(const enum CBLAS_ORDER Order,                         \
                         const enum CBLAS_SIDE Side,                           \
                         const enum CBLAS_UPLO Uplo, const int M, const int N, \
                         const T alpha, const T *A, const int lda, const T *B, \
                         const int ldb, const T beta, T *C, const int ldc) {   \
    const char *routine = "cblas_" #pfx "symm";                                \
    const int col = Order == CblasColMajor;                                    \
    REQUIRE(ok_order(Order), 1, "Illegal Order setting, %d\n", Order);         \
    REQUIRE(ok_side(Side), 2, "Illegal Side setting, %d\n", Side);             \
    REQUIRE(ok_uplo(Uplo), 3, "Illegal Uplo setting, %d\n", Uplo);            \
    REQUIRE(M >= 0, 4, "M must be nonnegative, %d\n", M);                      \
    REQUIRE(N >= 0, 5, "N must be nonnegative, %d\n", N);                      \
    REQUIRE(lda >= at_least1(Side == CblasLeft ? M : N), 8,                    \
            "lda is too small, %d\n", lda);                                    \
    const int rbc = at_least1(col ? M : N);                                    \
    REQUIRE(ldb >= rbc, 10, "ldb must be at least %d, %d\n", rbc, ldb);        \
    REQUIRE(ldc >= rbc, 13, "ldc must be at least %d, %d\n", rbc, ldc);        \
    BACKEND(mublis_##pfx##symm(to_side(Side, !col), to_uplo(Uplo, !col),       \
                               col ? M : N, col ? N : M, alpha, A, 1, lda, B,  \
                               1, ldb, beta, C, 1, ldc));                      \
  }
DEF_SYMM(s, float)
DEF_SYMM(d, double)

/* ------------------------------------------------------- SYRK / SYR2K -- */
#define DEF_SYRK(pfx, T)                                                       \
  void cblas_##pfx##syrk(const enum CBLAS_ORDER Order,                         \
                         const enum CBLAS_UPLO Uplo,                           \
                         const enum CBLAS_TRANSPOSE Trans, const int N,        \
                         const int K, const T alpha, const T *A,               \
                         const int lda, const T beta, T *C, const int ldc) {   \
    const char *routine = "cblas_" #pfx "syrk";                                \
    const int col = Order == CblasColMajor;                                    \
    REQUIRE(ok_order(Order), 1, "Illegal Order setting, %d\n", Order);         \
    REQUIRE(ok_uplo(Uplo), 2, "Illegal Uplo setting, %d\n", Uplo);            \
    REQUIRE(ok_trans(Trans), 3, "Illegal Trans setting, %d\n", Trans);        \
    REQUIRE(N >= 0, 4, "N must be nonnegative, %d\n", N);                      \
    REQUIRE(K >= 0, 5, "K must be nonnegative, %d\n", K);                      \
    const int ra = at_least1((Trans == CblasNoTrans) == col ? N : K);          \
    REQUIRE(lda >= ra, 8, "lda must be at least %d, %d\n", ra, lda);           \
    REQUIRE(ldc >= at_least1(N), 11, "ldc is too small, %d\n", ldc);           \
    BACKEND(mublis_##pfx##syrk(to_uplo(Uplo, !col), to_trans(Trans, !col), N,  \
                               K, alpha, A, 1, lda, beta, C, 1, ldc));         \
  }
DEF_SYRK(s, float)
DEF_SYRK(d, double)

#define DEF_SYR2K(pfx, T)                                                      \
  void cblas_##pfx##syr2k(const enum CBLAS_ORDER Order,                        \
                          const enum CBLAS_UPLO Uplo,                          \
                          const enum CBLAS_TRANSPOSE Trans, const int N,       \
                          const int K, const T alpha, const T *A,              \
                          const int lda, const T *B, const int ldb,            \
                          const T beta, T *C, const int ldc) {                 \
    const char *routine = "cblas_" #pfx "syr2k";                               \
    const int col = Order == CblasColMajor;                                    \
    REQUIRE(ok_order(Order), 1, "Illegal Order setting, %d\n", Order);         \
    REQUIRE(ok_uplo(Uplo), 2, "Illegal Uplo setting, %d\n", Uplo);            \
    REQUIRE(ok_trans(Trans), 3, "Illegal Trans setting, %d\n", Trans);        \
    REQUIRE(N >= 0, 4, "N must be nonnegative, %d\n", N);                      \
    REQUIRE(K >= 0, 5, "K must be nonnegative, %d\n", K);                      \
    const int rab = at_least1((Trans == CblasNoTrans) == col ? N : K);         \
    REQUIRE(lda >= rab, 8, "lda must be at least %d, %d\n", rab, lda);         \
    REQUIRE(ldb >= rab, 10, "ldb must be at least %d, %d\n", rab, ldb);        \
    REQUIRE(ldc >= at_least1(N), 13, "ldc is too small, %d\n", ldc);           \
    BACKEND(mublis_##pfx##syr2k(to_uplo(Uplo, !col), to_trans(Trans, !col), N, \
                                K, alpha, A, 1, lda, B, 1, ldb, beta, C, 1,    \
                                ldc));                                         \
  }
DEF_SYR2K(s, float)
DEF_SYR2K(d, double)

/* -------------------------------------------------------- TRMM / TRSM -- */
#define DEF_TR(pfx, op, T)                                                     \
  void cblas_##pfx##op(const enum CBLAS_ORDER Order,                           \
                       const enum CBLAS_SIDE Side, const enum CBLAS_UPLO Uplo, \
                       const enum CBLAS_TRANSPOSE TransA,                      \
                       const enum CBLAS_DIAG Diag, const int M, const int N,   \
                       const T alpha, const T *A, const int lda, T *B,         \
                       const int ldb) {                                        \
    const char *routine = "cblas_" #pfx #op;                                   \
    const int col = Order == CblasColMajor;                                    \
    REQUIRE(ok_order(Order), 1, "Illegal Order setting, %d\n", Order);         \
    REQUIRE(ok_side(Side), 2, "Illegal Side setting, %d\n", Side);             \
    REQUIRE(ok_uplo(Uplo), 3, "Illegal Uplo setting, %d\n", Uplo);            \
    REQUIRE(ok_trans(TransA), 4, "Illegal TransA setting, %d\n", TransA);      \
    REQUIRE(ok_diag(Diag), 5, "Illegal Diag setting, %d\n", Diag);            \
    REQUIRE(M >= 0, 6, "M must be nonnegative, %d\n", M);                      \
    REQUIRE(N >= 0, 7, "N must be nonnegative, %d\n", N);                      \
    REQUIRE(lda >= at_least1(Side == CblasLeft ? M : N), 10,                   \
            "lda is too small, %d\n", lda);                                    \
    const int rb = at_least1(col ? M : N);                                     \
    REQUIRE(ldb >= rb, 12, "ldb must be at least %d, %d\n", rb, ldb);          \
    BACKEND(mublis_##pfx##op(                                                  \
        to_side(Side, !col), to_uplo(Uplo, !col), to_trans(TransA, 0),         \
        Diag == CblasUnit ? MUBLIS_UNIT_DIAG : MUBLIS_NON_UNIT_DIAG,           \
        col ? M : N, col ? N : M, alpha, A, 1, lda, B, 1, ldb));               \
  }
DEF_TR(s, trmm, float)
DEF_TR(d, trmm, double)
DEF_TR(s, trsm, float)
DEF_TR(d, trsm, double)
