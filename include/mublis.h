/*
 * mublis.h -- stride-based L3 API and the generic L3 driver.
 *
 * Same ABI as the reference's public header (include/mublis.h:8-197): enum
 * values, struct layouts and prototypes are identical, so code written
 * against MuBLIS links against this library unchanged.
 *
 * Element (i, j) of a matrix X is X[i*rs_x + j*cs_x].  A transpose flag is a
 * swap of (rs, cs).  Every routine returns 0 on success and non-zero on
 * failure (1: device/context failure, 2: out of device memory, 3: unsupported
 * argument, 4: a HIP runtime error).
 *
 * All routines execute on the GPU.  Host pointers are staged in and out;
 * device pointers are used in place.  Calls return after the result is
 * visible to the caller.
 */
#ifndef MUBLIS_H
#define MUBLIS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { MUBLIS_NO_TRANSPOSE = 0, MUBLIS_TRANSPOSE = 1 } mublis_trans_t;
typedef enum { MUBLIS_DENSE = 0, MUBLIS_LOWER = 1, MUBLIS_UPPER = 2 } mublis_uplo_t;
typedef enum { MUBLIS_LEFT = 0, MUBLIS_RIGHT = 1 } mublis_side_t;
typedef enum { MUBLIS_NON_UNIT_DIAG = 0, MUBLIS_UNIT_DIAG = 1 } mublis_diag_t;

/* Logical structure of an operand of the generic driver. */
typedef enum {
  MUBLIS_STRUC_GENERAL = 0,    /* every element is read as stored            */
  MUBLIS_STRUC_SYMMETRIC = 1,  /* only `uplo` is stored; the rest reflects   */
  MUBLIS_STRUC_TRIANGULAR = 2  /* outside `uplo` the operand is zero         */
} mublis_packm_struc_t;

/* Diagonal treatment of a TRIANGULAR operand. */
typedef enum {
  MUBLIS_PACKM_DIAG_NONUNIT = 0, /* read the diagonal                       */
  MUBLIS_PACKM_DIAG_UNIT = 1,    /* diagonal is 1 and never read            */
  MUBLIS_PACKM_DIAG_INVERT = 2   /* diagonal is replaced by its reciprocal  */
} mublis_packm_diag_t;

/* Index relations: LOWER means lhs <= rhs, UPPER means lhs >= rhs. */
typedef enum {
  MUBLIS_L3_ALL = 0,
  MUBLIS_L3_LOWER = 1,
  MUBLIS_L3_UPPER = 2
} mublis_l3_relation_t;

/* ji selects the written part of C (j column, i row); jp and pi describe
 * which (j, p) and (p, i) pairs carry non-zero terms. */
typedef struct {
  mublis_l3_relation_t jp;
  mublis_l3_relation_t ji;
  mublis_l3_relation_t pi;
} mublis_l3_domain_t;

typedef struct {
  mublis_packm_struc_t struc;
  mublis_uplo_t uplo;
  mublis_packm_diag_t diag;
} mublis_l3_operand_t;

typedef struct {
  mublis_l3_domain_t domain;
  mublis_l3_operand_t a;
  mublis_l3_operand_t b;
} mublis_l3_product_t;

/* C := alpha * A_view * B_view + beta * C on the elements allowed by ji. */
int mublis_l3_sdriver(int m, int n, int k, float alpha,
                      const float *a, int rs_a, int cs_a,
                      const float *b, int rs_b, int cs_b, float beta,
                      float *c, int rs_c, int cs_c,
                      const mublis_l3_product_t *product);
int mublis_l3_ddriver(int m, int n, int k, double alpha,
                      const double *a, int rs_a, int cs_a,
                      const double *b, int rs_b, int cs_b, double beta,
                      double *c, int rs_c, int cs_c,
                      const mublis_l3_product_t *product);

int mublis_sgemm(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n,
                 int k, float alpha, const float *a, int rs_a, int cs_a,
                 const float *b, int rs_b, int cs_b, float beta, float *c,
                 int rs_c, int cs_c);
int mublis_ssymm(mublis_side_t side, mublis_uplo_t uplo, int m, int n,
                 float alpha, const float *a, int rs_a, int cs_a,
                 const float *b, int rs_b, int cs_b, float beta, float *c,
                 int rs_c, int cs_c);
int mublis_ssyrk(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k,
                 float alpha, const float *a, int rs_a, int cs_a, float beta,
                 float *c, int rs_c, int cs_c);
int mublis_ssyr2k(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k,
                  float alpha, const float *a, int rs_a, int cs_a,
                  const float *b, int rs_b, int cs_b, float beta, float *c,
                  int rs_c, int cs_c);
int mublis_strmm(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                 mublis_diag_t diag, int m, int n, float alpha, const float *a,
                 int rs_a, int cs_a, float *b, int rs_b, int cs_b);
int mublis_strsm(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                 mublis_diag_t diag, int m, int n, float alpha, const float *a,
                 int rs_a, int cs_a, float *b, int rs_b, int cs_b);

int mublis_dgemm(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n,
                 int k, double alpha, const double *a, int rs_a, int cs_a,
                 const double *b, int rs_b, int cs_b, double beta, double *c,
                 int rs_c, int cs_c);
int mublis_dsymm(mublis_side_t side, mublis_uplo_t uplo, int m, int n,
                 double alpha, const double *a, int rs_a, int cs_a,
                 const double *b, int rs_b, int cs_b, double beta, double *c,
                 int rs_c, int cs_c);
int mublis_dsyrk(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k,
                 double alpha, const double *a, int rs_a, int cs_a, double beta,
                 double *c, int rs_c, int cs_c);
int mublis_dsyr2k(mublis_uplo_t uplo, mublis_trans_t trans, int n, int k,
                  double alpha, const double *a, int rs_a, int cs_a,
                  const double *b, int rs_b, int cs_b, double beta, double *c,
                  int rs_c, int cs_c);
int mublis_dtrmm(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                 mublis_diag_t diag, int m, int n, double alpha, const double *a,
                 int rs_a, int cs_a, double *b, int rs_b, int cs_b);
int mublis_dtrsm(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                 mublis_diag_t diag, int m, int n, double alpha, const double *a,
                 int rs_a, int cs_a, double *b, int rs_b, int cs_b);

#ifdef __cplusplus
}
#endif
#endif /* MUBLIS_H */
