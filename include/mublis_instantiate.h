/*
 * mublis_instantiate.h -- the target/config plugin seam.
 *
 * Same ABI as the reference (include/mublis_instantiate.h:31-247).  A target
 * supplies GEMM and GEMMTRSM micro-kernels plus register (mr, nr) and cache
 * (mc, kc, nc) block sizes per precision; a config's mublis_get_context()
 * picks one target at run time.
 *
 * The gfx950 target (mublis_amd/targets/gfx950) fills this context with
 * micro-kernels that run one MR x NR tile as a wavefront MFMA tile on the
 * GPU, and with block sizes that describe the device tiling (mr x nr = one
 * wave's tile, mc x nc = one workgroup's tile, kc = one LDS stage).
 */
#ifndef MUBLIS_INSTANTIATE_H
#define MUBLIS_INSTANTIATE_H

#ifdef __cplusplus
#define MUBLIS_RESTRICT __restrict__
extern "C" {
#else
#define MUBLIS_RESTRICT restrict
#endif

/* Optional hint: the C tile the frame visits next (may be NULL). */
typedef struct {
  const void *c_next;
} mublis_auxinfo_t;

/*
 * GEMM micro-kernel: C := alpha*A*B + beta*C on one full MR x NR tile.
 * A(i,p) = a[p*MR + i], B(p,j) = b[p*NR + j], C(i,j) = c[i*rs_c + j*cs_c].
 * beta == 0 means C is not read.
 */
#define MUBLIS_GEMM_UKR_TYPE(ctype, name)                                      \
  typedef void (*name)(int k, ctype alpha, const ctype *MUBLIS_RESTRICT a,     \
                       const ctype *MUBLIS_RESTRICT b, ctype beta,             \
                       ctype *MUBLIS_RESTRICT c, int rs_c, int cs_c,           \
                       const mublis_auxinfo_t *aux)
MUBLIS_GEMM_UKR_TYPE(float, mublis_sgemm_ukr_ft);
MUBLIS_GEMM_UKR_TYPE(double, mublis_dgemm_ukr_ft);
#undef MUBLIS_GEMM_UKR_TYPE

/*
 * GEMMTRSM micro-kernel: R := alpha*B11 - A1x*Bx1, then X solves A11*X = R.
 * a1x: MR x k packed as a1x[p*MR + i]; bx1: k x NR packed as bx1[p*NR + j];
 * a11: MR x MR column-major with the diagonal already inverted; b11: MR x NR
 * row-major, overwritten with X; X is also written to c11 with (rs_c, cs_c).
 * The lower kernel substitutes forward, the upper one backward.
 */
#define MUBLIS_GEMMTRSM_UKR_TYPE(ctype, name)                                  \
  typedef void (*name)(int k, ctype alpha, const ctype *MUBLIS_RESTRICT a1x,   \
                       const ctype *MUBLIS_RESTRICT a11,                       \
                       const ctype *MUBLIS_RESTRICT bx1,                       \
                       ctype *MUBLIS_RESTRICT b11, ctype *MUBLIS_RESTRICT c11, \
                       int rs_c, int cs_c, const mublis_auxinfo_t *aux)
MUBLIS_GEMMTRSM_UKR_TYPE(float, mublis_sgemmtrsml_ukr_ft);
MUBLIS_GEMMTRSM_UKR_TYPE(float, mublis_sgemmtrsmu_ukr_ft);
MUBLIS_GEMMTRSM_UKR_TYPE(double, mublis_dgemmtrsml_ukr_ft);
MUBLIS_GEMMTRSM_UKR_TYPE(double, mublis_dgemmtrsmu_ukr_ft);
#undef MUBLIS_GEMMTRSM_UKR_TYPE

/* Per-precision kernels and block sizes.  All sizes positive; mc % mr == 0,
 * nc % nr == 0, kc % mr == 0 and kc % nr == 0. */
typedef struct {
  mublis_sgemm_ukr_ft gemm_ukr;
  mublis_sgemmtrsml_ukr_ft gemmtrsml_ukr;
  mublis_sgemmtrsmu_ukr_ft gemmtrsmu_ukr;
  int mr, nr;
  int mc, kc, nc;
} mublis_scontext_t;

typedef struct {
  mublis_dgemm_ukr_ft gemm_ukr;
  mublis_dgemmtrsml_ukr_ft gemmtrsml_ukr;
  mublis_dgemmtrsmu_ukr_ft gemmtrsmu_ukr;
  int mr, nr;
  int mc, kc, nc;
} mublis_dcontext_t;

typedef struct {
  mublis_scontext_t s;
  mublis_dcontext_t d;
} mublis_context_t;

#define MUBLIS_GEMM_UKR_PROT_STAMP(ctype, name)                                \
  void name(int k, ctype alpha, const ctype *MUBLIS_RESTRICT a,                \
            const ctype *MUBLIS_RESTRICT b, ctype beta,                        \
            ctype *MUBLIS_RESTRICT c, int rs_c, int cs_c,                      \
            const mublis_auxinfo_t *aux);

#define MUBLIS_GEMMTRSM_UKR_PROT_STAMP(ctype, name)                            \
  void name(int k, ctype alpha, const ctype *MUBLIS_RESTRICT a1x,              \
            const ctype *MUBLIS_RESTRICT a11, const ctype *MUBLIS_RESTRICT bx1,\
            ctype *MUBLIS_RESTRICT b11, ctype *MUBLIS_RESTRICT c11, int rs_c,  \
            int cs_c, const mublis_auxinfo_t *aux);

/* Defines `const mublis_context_t context_name` from the two groups of
 * per-precision values (single first, then double). */
#define MUBLIS_CONTEXT_STAMP(context_name, s_gemm, s_trsml, s_trsmu, s_mr,     \
                             s_nr, s_mc, s_kc, s_nc, d_gemm, d_trsml, d_trsmu, \
                             d_mr, d_nr, d_mc, d_kc, d_nc)                     \
  const mublis_context_t context_name = {                                      \
    {(s_gemm), (s_trsml), (s_trsmu), (s_mr), (s_nr), (s_mc), (s_kc), (s_nc)},  \
    {(d_gemm), (d_trsml), (d_trsmu), (d_mr), (d_nr), (d_mc), (d_kc), (d_nc)}   \
  };

/* Every config defines this: write the selected context, return 0 on success. */
int mublis_get_context(mublis_context_t *context);

#ifdef __cplusplus
}
#endif
#endif /* MUBLIS_INSTANTIATE_H */
