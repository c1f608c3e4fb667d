/*
 * mublis_gfx950.h -- thin C ABI between the host C frame and the gfx950
 * (MI355X, CDNA4) HIP layer.
 *
 * The frame (cblas.h / mublis.h entry points) stages operands and calls the
 * device-level routines below.  Every pointer passed to a *_dev routine is a
 * device pointer; sizes are element counts, strides are element strides
 * (64-bit).  Routines enqueue work on `stream` (NULL: the calling thread's
 * library stream) and return without synchronising; the return value is
 * MUBLIS_OK or one of the MUBLIS_E* codes.
 *
 * Each routine names the reference interface whose work it replaces.
 */
#ifndef MUBLIS_GFX950_H
#define MUBLIS_GFX950_H

#include <stddef.h>
#include <stdint.h>

#include "mublis.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MUBLIS_OK = 0,
  MUBLIS_ECONTEXT = 1, /* no usable gfx950 device, or context invalid     */
  MUBLIS_ENOMEM = 2,   /* device or pinned allocation failed              */
  MUBLIS_EINVAL = 3,   /* argument the device path does not support       */
  MUBLIS_EHIP = 4      /* a HIP runtime call failed                       */
};

/* ------------------------------------------------------------- runtime -- */
/* Lazily selects the device and checks that it is gfx950.  Replaces the
 * first-call work of mublis_get_safe_context + mublis_pool_init
 * (frame/base/safe_dispatch.c:45-71, frame/base/pool.c:182-244). */
int mublis_rt_init(void);
/* Architecture name of the current device ("gfx950"), or "" before init. */
const char *mublis_rt_arch(void);
/* 1 if p is device (or managed) memory usable by kernels in place. */
int mublis_rt_is_device_ptr(const void *p);
/* The calling thread's stream (created on first use). */
void *mublis_rt_stream(void);
/* Device workspace from the thread-safe caching arena (replaces the
 * reference's mutex-guarded pack-buffer pool, frame/base/pool.c:266-318). */
void *mublis_rt_alloc(size_t bytes);
void mublis_rt_free(void *p);
int mublis_rt_h2d(void *dst, const void *src, size_t bytes, void *stream);
int mublis_rt_d2h(void *dst, const void *src, size_t bytes, void *stream);
/* A rows x cols host block (rows contiguous, ld elements between column
 * starts, elements of esz bytes) to or from a dense device block, through
 * the library's pool of pinned host blocks: host columns are gathered into
 * a pinned block by a team of copy threads while the previous block's DMA
 * runs.  h2d returns once every piece is queued on stream (the host memory
 * may be reused then); d2h returns once the data is in host memory.  The
 * host-pointer staging of the reference's pack-buffer pool
 * (frame/base/pool.c:120-318) for a device across PCIe; calls serialise on
 * the pool. */
int mublis_rt_h2d_2d(void *dev, const void *host, int64_t rows, int64_t cols,
                     int64_t ld, size_t esz, void *stream);
int mublis_rt_d2h_2d(void *host, const void *dev, int64_t rows, int64_t cols,
                     int64_t ld, size_t esz, void *stream);
int mublis_rt_sync(void *stream);
/* Last HIP error text seen by this thread (for diagnostics). */
const char *mublis_rt_last_error(void);

/* ------------------------------------------------------- generic L3 ------ */
/*
 * C := alpha * Aview * Bview + beta * C on the elements (i, j) allowed by ji
 * (0 all, 1 lower: j <= i, 2 upper: j >= i).  A is m x k, B is k x n.
 * When a2/b2 are non-NULL a second product alpha * A2view * B2view is added
 * in the same launch (SYR2K).  Any (rs, cs) pair is accepted; operands with
 * no unit stride are first packed to a dense buffer.
 * Replaces the generic driver's loop nest + packm + gemm micro-kernel
 * (frame/l3/l3_driver.c:126-239, frame/l1m/packm.c:18-84,
 * include/mublis_instantiate.h:53-66).
 */
int mublis_gfx950_dgemm_dev(int m, int n, int k, double alpha,
                            const double *a, int64_t rs_a, int64_t cs_a,
                            const double *b, int64_t rs_b, int64_t cs_b,
                            const double *a2, int64_t rs_a2, int64_t cs_a2,
                            const double *b2, int64_t rs_b2, int64_t cs_b2,
                            double beta, double *c, int64_t rs_c, int64_t cs_c,
                            int ji, void *stream);
int mublis_gfx950_sgemm_dev(int m, int n, int k, float alpha,
                            const float *a, int64_t rs_a, int64_t cs_a,
                            const float *b, int64_t rs_b, int64_t cs_b,
                            const float *a2, int64_t rs_a2, int64_t cs_a2,
                            const float *b2, int64_t rs_b2, int64_t cs_b2,
                            float beta, float *c, int64_t rs_c, int64_t cs_c,
                            int ji, void *stream);
/* C := alpha * Aview * Bview + beta * C with C in host memory (rs_c == 1,
 * cs_c >= m) and A, B in host or device memory (a host operand needs a unit
 * stride): host A, and host B whose columns are not contiguous, are sent
 * whole; C is then computed by column panels, each panel's B and C columns
 * sent, multiplied and brought back while the next panel is sent, so the
 * PCIe transfers overlap the GEMM kernels.  Returns once C is in host
 * memory.  k > 0 and m, n > 0.  The host-pointer form of cblas_?gemm
 * (frame/compat/cblas/cblas.c:188-200 -> frame/l3/gemm.c:14-24). */
int mublis_gfx950_dgemm_host(int m, int n, int k, double alpha, const double *a,
                             int64_t rs_a, int64_t cs_a, const double *b,
                             int64_t rs_b, int64_t cs_b, double beta, double *c,
                             int64_t rs_c, int64_t cs_c, void *stream);
int mublis_gfx950_sgemm_host(int m, int n, int k, float alpha, const float *a,
                             int64_t rs_a, int64_t cs_a, const float *b,
                             int64_t rs_b, int64_t cs_b, float beta, float *c,
                             int64_t rs_c, int64_t cs_c, void *stream);

/*
 * Materialise the view of an m x n operand into dense column-major `dst`
 * (leading dimension ld): struc 0 GENERAL, 1 SYMMETRIC (read the reflected
 * element outside `uplo`), 2 TRIANGULAR (zero outside `uplo`; diag 1 UNIT
 * writes 1 without reading, 2 INVERT writes 1/a).  diagoff = row - col offset
 * of the view's (0,0) from the diagonal.  mask (0 all, 1 lower, 2 upper)
 * then zeroes, without reading, the elements on the other side of that
 * diagonal (the driver's jp / pi domains).  Unreferenced storage is never
 * read, so NaN there cannot leak.  Replaces packm at whole-operand
 * granularity (frame/l1m/packm.c:18-84).  Codes outside 0..2 give
 * MUBLIS_EINVAL.
 */
int mublis_gfx950_dpack_dev(int m, int n, const double *src, int64_t rs,
                            int64_t cs, int struc, int uplo, int diag,
                            int diagoff, int mask, double *dst, int64_t ld,
                            void *stream);
int mublis_gfx950_spack_dev(int m, int n, const float *src, int64_t rs,
                            int64_t cs, int struc, int uplo, int diag,
                            int diagoff, int mask, float *dst, int64_t ld,
                            void *stream);

/*
 * In-place triangular solve, Left side only: B := alpha * A^{-1} * B, A m x m
 * lower (lower = 1) or upper triangular as viewed through (rs_a, cs_a); unit
 * = 1 means the diagonal is 1 and is never read.  Right-side and transposed
 * problems are mapped onto this form by the caller exactly as the reference
 * does (frame/l3/trsm.c:35-54).  The 64 x 64 diagonal blocks are inverted in
 * FP64 (also for float data); narrow B is solved by one pipelined GEMMTRSM
 * launch, wide B by a recursive schedule of MFMA GEMMs with GEMMTRSM leaves.
 * Synchronises the stream before returning (it releases its workspace).
 * Replaces the TRSM loop nest and the GEMMTRSM micro-kernels
 * (frame/l3/trsm.c:87-431, include/mublis_instantiate.h:95-111).
 */
int mublis_gfx950_dtrsm_dev(int lower, int unit, int m, int n, double alpha,
                            const double *a, int64_t rs_a, int64_t cs_a,
                            double *b, int64_t rs_b, int64_t cs_b, void *stream);
int mublis_gfx950_strsm_dev(int lower, int unit, int m, int n, float alpha,
                            const float *a, int64_t rs_a, int64_t cs_a,
                            float *b, int64_t rs_b, int64_t cs_b, void *stream);

/* Diagonal-block solve (m <= 64) whose stored diagonal already holds the
 * reciprocals, as the GEMMTRSM micro-kernel contract packs it
 * (include/mublis_instantiate.h:84-86, frame/l1m/packm.c:143-144). */
int mublis_gfx950_dtrsm_inverted_dev(int lower, int m, int n, const double *a,
                                     int64_t rs_a, int64_t cs_a, double *b,
                                     int64_t rs_b, int64_t cs_b, void *stream);
int mublis_gfx950_strsm_inverted_dev(int lower, int m, int n, const float *a,
                                     int64_t rs_a, int64_t cs_a, float *b,
                                     int64_t rs_b, int64_t cs_b, void *stream);

/* Left-side TRSM with caller workspace, fully asynchronous on `stream`: the
 * form the serving loop uses so its scheduler never waits on a solve.  At
 * least mublis_gfx950_trsm_ws_bytes(m) bytes of device memory are needed
 * (the 64-row diagonal-block inverses and the pipeline's flags).  With
 * mublis_gfx950_?trsm_sched_ws_bytes(m, n) bytes or more the solve runs the
 * same schedule as the plain mublis_gfx950_?trsm_dev call (GEMM leaves on
 * inverted 1024-row blocks where they apply) and gives the same bits; with
 * less it ends its recursion in pipelined or panel leaves instead. */
size_t mublis_gfx950_trsm_ws_bytes(int m);
size_t mublis_gfx950_dtrsm_sched_ws_bytes(int m, int n);
size_t mublis_gfx950_strsm_sched_ws_bytes(int m, int n);
/* Widest B (columns after the side transformation) that the left-side
 * solves hand to the pipelined GEMMTRSM kernel (one launch, block rows
 * chained by device-side flags) instead of the recursive schedule; set it
 * with max_n >= 0 (at most 4096), query with max_n < 0.  Values below 1024
 * also turn off the recursive schedule's GEMM leaves (inverted 1024-row
 * diagonal blocks), so 0 leaves the plain recursion with 64-row leaves.
 * Returns the previous value.  Tuning knob; the frame's calls share it. */
int mublis_gfx950_trsm_pipe_max_n(int max_n);
/* Polls a pipelined solve's wait makes before it gives up (default 2^20,
 * about a second; set with spins > 0, returns the previous value).  A wait
 * that gives up raises the solve's abort flag: every other wait then
 * returns, and the workgroups store NaN for their blocks of X, so the call
 * completes and the failure shows in B.  No counterpart in the reference;
 * tests use a tiny limit to exercise that path. */
int mublis_gfx950_trsm_pipe_spin_limit(int spins);
int mublis_gfx950_dtrsm_ws_dev(int lower, int unit, int m, int n, double alpha,
                               const double *a, int64_t rs_a, int64_t cs_a, double *b,
                               int64_t rs_b, int64_t cs_b, void *ws, size_t ws_bytes,
                               void *stream);
int mublis_gfx950_strsm_ws_dev(int lower, int unit, int m, int n, float alpha,
                               const float *a, int64_t rs_a, int64_t cs_a, float *b,
                               int64_t rs_b, int64_t cs_b, void *ws, size_t ws_bytes,
                               void *stream);

/* Allocation-free DTRSM: the solve of mublis_gfx950_dtrsm_ws_dev given
 * mublis_gfx950_trsm_ws_bytes(m) bytes (the same schedule and bits), with
 * that workspace taken from a static device array of the library instead
 * (8.5 MB per GPU, held until the process ends).  No counterpart in the
 * reference: for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Limits: m <= 16384, and A and B each need a unit stride (else
 * MUBLIS_EINVAL with mublis_rt_last_error() set); stream NULL is the null
 * stream.  Calls never allocate; calls that overlap on one GPU share the
 * array and race, so callers serialise them per device. */
int mublis_gfx950_dtrsm_noalloc_dev(int lower, int unit, int m, int n, double alpha,
                                    const double *a, int64_t rs_a, int64_t cs_a,
                                    double *b, int64_t rs_b, int64_t cs_b, void *stream);

/* Strided copy of an m x n matrix (dst may have other strides). */
int mublis_gfx950_dcopy_dev(int m, int n, const double *src, int64_t rs_s,
                            int64_t cs_s, double *dst, int64_t rs_d,
                            int64_t cs_d, void *stream);
int mublis_gfx950_scopy_dev(int m, int n, const float *src, int64_t rs_s,
                            int64_t cs_s, float *dst, int64_t rs_d,
                            int64_t cs_d, void *stream);
/* X := value everywhere (alpha == 0 paths of TRMM/TRSM, frame/l3/trsm.c:27-33). */
int mublis_gfx950_dfill_dev(int m, int n, double value, double *x, int64_t rs,
                            int64_t cs, void *stream);
int mublis_gfx950_sfill_dev(int m, int n, float value, float *x, int64_t rs,
                            int64_t cs, void *stream);

/* GEMM whose A (tri_a) or B (tri_b) operand is a materialised triangle:
 * tri_a 1 means A(i,p) = 0 for p > i, 2 for p < i; tri_b 1 means B(p,j) = 0
 * for p < j, 2 for p > j.  K slices that are entirely zero for a tile are
 * skipped, halving the work of TRMM-shaped products (frame/l3/trmm.c:51-82
 * expresses the same skipping through the jp/pi domain relations). */
int mublis_gfx950_dgemm_tri_dev(int m, int n, int k, double alpha, const double *a,
                                int64_t rs_a, int64_t cs_a, const double *b,
                                int64_t rs_b, int64_t cs_b, double beta, double *c,
                                int64_t rs_c, int64_t cs_c, int tri_a, int tri_b,
                                void *stream);
int mublis_gfx950_sgemm_tri_dev(int m, int n, int k, float alpha, const float *a,
                                int64_t rs_a, int64_t cs_a, const float *b,
                                int64_t rs_b, int64_t cs_b, float beta, float *c,
                                int64_t rs_c, int64_t cs_c, int tri_a, int tri_b,
                                void *stream);

/* ---------------------------------------------- precision extensions -- */
/*
 * SGEMM as a 3-term BF16 split product on BF16 MFMA with FP32 accumulation
 * (hi/mid/lo planes, six partial products).  Same contract as
 * mublis_gfx950_sgemm_dev with ji = 0; it replaces the FP32 micro-kernel
 * arithmetic of the GEMM path (include/mublis_instantiate.h:53-66) when the
 * caller selects it.
 */
int mublis_gfx950_sgemm_bf16x3_dev(int m, int n, int k, float alpha, const float *a,
                                   int64_t rs_a, int64_t cs_a, const float *b,
                                   int64_t rs_b, int64_t cs_b, float beta, float *c,
                                   int64_t rs_c, int64_t cs_c, void *stream);
/*
 * Left-side STRSM (same form as mublis_gfx950_strsm_dev) followed by `sweeps`
 * rounds of iterative refinement whose residual alpha*B - op(A)*X and update
 * X += D run in FP64.  Replaces the TRSM path (frame/l3/trsm.c:12-441) when
 * the caller asks for refinement.
 */
int mublis_gfx950_strsm_refined_dev(int lower, int unit, int m, int n, float alpha,
                                    const float *a, int64_t rs_a, int64_t cs_a,
                                    float *b, int64_t rs_b, int64_t cs_b, int sweeps,
                                    void *stream);

/* Host-callable forms with the conventions of include/mublis.h (host or
 * device pointers, synchronous, 0 on success).  No counterpart in the
 * reference: they add the precision variants beside mublis_sgemm and
 * mublis_strsm. */
int mublis_sgemm_bf16x3(mublis_trans_t trans_a, mublis_trans_t trans_b, int m, int n,
                        int k, float alpha, const float *a, int rs_a, int cs_a,
                        const float *b, int rs_b, int cs_b, float beta, float *c,
                        int rs_c, int cs_c);
int mublis_strsm_refined(mublis_side_t side, mublis_uplo_t uplo, mublis_trans_t trans_a,
                         mublis_diag_t diag, int m, int n, float alpha, const float *a,
                         int rs_a, int cs_a, float *b, int rs_b, int cs_b, int sweeps);
/* Arithmetic used by cblas_sgemm / mublis_sgemm for plain GEMM products:
 * 0 native FP32 MFMA (default), 1 BF16x3 split.  The initial value comes
 * from the environment variable MUBLIS_SGEMM_MODE ("native" or "bf16x3").
 * Returns the previous mode, or -1 for an unknown mode. */
int mublis_set_sgemm_mode(int mode);
int mublis_get_sgemm_mode(void);

/* ---------------------------------------------- batched extensions ---- */
/*
 * Grouped batched FP32 GEMM / SYRK / SYR2K / TRSM: many uneven problems in
 * one set of persistent launches, captured once in a hipGraph and replayed.
 * No counterpart in the reference (which is one call per problem,
 * frame/l3/l3_driver.c:49-249); each GEMM/SYRK/SYR2K problem's result is
 * bit-identical to the corresponding cblas_sgemm / cblas_ssyrk /
 * cblas_ssyr2k call on the same device operands.
 *
 * Operands are column-major DEVICE pointers.  GEMM: C := alpha*op(A)*op(B)
 * + beta*C with op from trans_a / trans_b (0 or 1), C m x n.  SYRK:
 * C := alpha*op(A)*op(A)' + beta*C on the `uplo` triangle (MUBLIS_LOWER /
 * MUBLIS_UPPER) of the n x n C, op(A) n x k (trans_a), b unused.  SYR2K:
 * C := alpha*(op(A)*op(B)' + op(B)*op(A)') + beta*C on the `uplo` triangle,
 * op(A) and op(B) n x k (trans_a applies to both).  TRSM: the m x n B at
 * c (ldc) is overwritten by alpha*inv(op(A))*B (side MUBLIS_LEFT, A m x m)
 * or alpha*B*inv(op(A)) (MUBLIS_RIGHT, A n x n), A triangular (`uplo`,
 * trans_a, `diag` MUBLIS_UNIT_DIAG or MUBLIS_NON_UNIT_DIAG); b, k, beta
 * unused.  The 64-row diagonal blocks are inverted in FP64 and the block
 * updates run on FP64 MFMA, as in cblas_strsm; problems whose triangle has
 * order up to 1024 run as 16- or 32-column panels inside the grouped launches,
 * larger ones (and alpha = 0) as their own solve in the same plan or
 * launch.  Problems of one batch must not write what another reads or
 * writes.  side and diag are ignored by the other ops (zero them).
 */
enum { MUBLIS_BATCH_GEMM = 0, MUBLIS_BATCH_SYRK = 1, MUBLIS_BATCH_SYR2K = 2,
       MUBLIS_BATCH_TRSM = 3 };
typedef struct {
  int op;
  int trans_a, trans_b;
  int uplo;
  int m, n, k;
  float alpha, beta;
  const float *a;
  int lda;
  const float *b;
  int ldb;
  float *c;
  int ldc;
  int side, diag; /* TRSM only */
} mublis_sbatch_op_t;
typedef struct mublis_grouped_plan_s *mublis_grouped_plan_t;
/* flags for plan creation: force one tile shape (0: 32, 1: 64, 2: 128) */
#define MUBLIS_GROUPED_FORCE_TILE(t) (((t) + 1) << 4)
/* flags for plan creation: force launch schedule variant v (0 or 1) of every
 * tile class (workgroups per CU, threads, K-slice depth; results are the
 * same under every schedule) */
#define MUBLIS_GROUPED_FORCE_SCHED(v) (((v) + 1) << 8)

/* Expand the problems into a cost-sorted tile work list (tile shapes from
 * the autotuner table, else a size heuristic), upload it and capture the
 * launch.  Allocates device memory; executing the plan does not. */
int mublis_sgrouped_plan_create(const mublis_sbatch_op_t *ops, int count, int flags,
                                mublis_grouped_plan_t *plan);
/* Replay the captured launch on `stream` (NULL: this thread's stream);
 * asynchronous. */
int mublis_grouped_plan_execute(mublis_grouped_plan_t plan, void *stream);
/* Algorithmic flops and minimum HBM bytes of the plan, work items, and how
 * many problems got each tile shape. */
int mublis_grouped_plan_info(mublis_grouped_plan_t plan, double *flops, double *bytes,
                             int *nitems, int *tile_hist3);
/* Release a plan whose executions have completed. */
int mublis_grouped_plan_destroy(mublis_grouped_plan_t plan);
/* Without a plan or graph: size of the device workspace (and of the pinned
 * host staging buffer) a launch of these ops needs, and an asynchronous
 * launch that stages the work list through `host_stage` (pinned, kept
 * untouched until the launch completes) into `dev_ws`.  Used by the
 * serving loop, which cannot stall its scheduler on a plan. */
size_t mublis_sgrouped_ws_bytes(const mublis_sbatch_op_t *ops, int count);
int mublis_sgrouped_launch_ws(const mublis_sbatch_op_t *ops, int count, void *host_stage,
                              void *dev_ws, size_t ws_bytes, void *stream);
/* One-shot: plan, execute, synchronise, destroy. */
int mublis_sgemm_grouped(const mublis_sbatch_op_t *ops, int count);
/* Autotuner: choose the launch schedule of each tile class, then the tile
 * shape of every GEMM bucket (floor log2 of m, n, k in [lo, hi], 4 <= lo <=
 * hi <= 12) and every SYRK bucket (log2 n, log2 k), each candidate timed on
 * a batch that fills the GPU (at least per_bucket problems), best of reps
 * replays; near-ties take the size heuristic's choice.  Writes the table at
 * `path` and loads it.  The environment variable MUBLIS_TUNE_TABLE names
 * the table loaded at first use; mublis_grouped_load_table switches tables
 * (NULL: heuristic only, schedule variant 0) and returns the number of
 * buckets read. */
int mublis_grouped_autotune(const char *path, int lo, int hi, int per_bucket, int reps);
int mublis_grouped_load_table(const char *path);

/* ------------------------------------------------------------------------
 * Serving: a request loop over device-resident matrices.
 *
 * No counterpart in the reference (its L3 calls are synchronous and
 * per-thread, frame/l3/l3_driver.c:49-256); the requests carry the same
 * arguments as the cblas_* calls they stand for (include/cblas.h:37-95),
 * column-major.  A server owns one HBM budget, reserved once at creation:
 * matrices (handles) and launch workspace are carved out of it.  Requests
 * are submitted from any thread, wait for admission (a free launch slot,
 * workspace within the budget, no conflict with earlier requests on the
 * same matrices), and a scheduler thread groups every admissible FP32
 * GEMM/SYRK/SYR2K request into one grouped persistent launch; TRSM and
 * DGEMM requests run as their own launches.  Launches are spread over
 * several streams, so new requests start while earlier ones still run.
 * Requests on one matrix keep their submission order.
 * ---------------------------------------------------------------------- */
typedef struct mublis_server_s *mublis_server_t;
typedef struct mublis_matrix_s *mublis_matrix_t;

enum { MUBLIS_F32 = 0, MUBLIS_F64 = 1 };
enum {
  MUBLIS_REQ_SGEMM = 0,  /* C := alpha*op(A)*op(B) + beta*C, C m x n, k inner  */
  MUBLIS_REQ_SSYRK = 1,  /* C := alpha*op(A)*op(A)' + beta*C, uplo of n x n C  */
  MUBLIS_REQ_SSYR2K = 2, /* C := alpha*(op(A)op(B)' + op(B)op(A)') + beta*C     */
  MUBLIS_REQ_DGEMM = 3,
  MUBLIS_REQ_STRSM = 4,  /* B := alpha*op(A)^-1*B (left) or alpha*B*op(A)^-1    */
  MUBLIS_REQ_DTRSM = 5
};

typedef struct {
  size_t hbm_budget;  /* bytes reserved for matrices and workspace (0: 1 GiB)   */
  int streams;        /* device streams launches are spread over (0: 4)         */
  int slots_per_stream; /* launches in flight per stream (0: 4)                */
  int max_group;      /* requests per grouped launch (0: 256)                   */
  double split_flop;  /* a GEMM-like request above this many flop gets its own
                         launch, so it does not hold small ones back (0: 2e10;
                         negative: never split)                                 */
} mublis_server_config_t;

typedef struct {
  int op;              /* MUBLIS_REQ_*                                          */
  int trans_a;         /* MUBLIS_NO_TRANSPOSE / MUBLIS_TRANSPOSE: op(A)         */
  int trans_b;         /* GEMM: op(B)                                           */
  int side;            /* TRSM: MUBLIS_LEFT / MUBLIS_RIGHT                       */
  int uplo;            /* SYRK/SYR2K: triangle of C; TRSM: triangle of A        */
  int diag;            /* TRSM: MUBLIS_NON_UNIT_DIAG / MUBLIS_UNIT_DIAG          */
  int m, n, k;         /* as in the cblas call (SYRK/SYR2K: n, k; TRSM: m, n)    */
  double alpha, beta;  /* converted to float for the S operations               */
  mublis_matrix_t a, b, c; /* operands (TRSM: A and B, B is overwritten)        */
} mublis_request_t;

typedef struct {
  uint64_t submitted, completed, failed;
  uint64_t launches;         /* device launches issued                         */
  uint64_t grouped_launches; /* of which grouped FP32 launches                 */
  uint64_t grouped_requests; /* requests that ran inside grouped launches      */
  uint64_t max_pending;      /* most requests waiting for admission at once    */
  uint64_t max_inflight;     /* most launches in flight at once                */
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion              */
  double flop;               /* algorithmic flop of completed requests         */
  double elapsed_s;          /* first submit -> last completion                */
  double tflops;             /* flop / elapsed_s / 1e12                        */
  size_t hbm_budget, hbm_used, hbm_peak; /* bytes                              */
} mublis_server_stats_t;

/* Create a server on the current device; reserves cfg->hbm_budget bytes
 * (NULL cfg: defaults).  MUBLIS_ENOMEM if the reservation fails. */
int mublis_server_create(const mublis_server_config_t *cfg, mublis_server_t *out);
/* Complete every submitted request, then release everything, handles
 * included. */
int mublis_server_destroy(mublis_server_t srv);
/* A rows x cols column-major matrix in the budget (leading dimension from
 * mublis_matrix_ld, padded to 64 bytes).  When it does not fit now: wait
 * for memory to be released if `wait` is non-zero, else MUBLIS_ENOMEM; a
 * matrix larger than the budget less the launch-workspace reserve is
 * MUBLIS_ENOMEM at once.  Contents are undefined until written. */
int mublis_matrix_create(mublis_server_t srv, int dtype, int rows, int cols, int wait,
                         mublis_matrix_t *out);
/* Copy from / to host memory (leading dimension ld >= rows); writes wait
 * for every submitted request on the matrix, reads for every writer. */
int mublis_matrix_write(mublis_server_t srv, mublis_matrix_t h, const void *host, int64_t ld);
int mublis_matrix_read(mublis_server_t srv, mublis_matrix_t h, void *host, int64_t ld);
/* Device address and leading dimension (elements), for callers that fill or
 * consume the matrix on the device themselves; such work must be complete
 * (or on the legacy default stream) before a request on the matrix is
 * submitted, and a consumer must wait for the request first. */
void *mublis_matrix_device_ptr(mublis_matrix_t h);
int64_t mublis_matrix_ld(mublis_matrix_t h);
/* Release the matrix once the requests submitted on it have completed. */
int mublis_matrix_destroy(mublis_server_t srv, mublis_matrix_t h);
/* Queue a request; returns at once with a ticket for mublis_wait.  Invalid
 * requests (shapes beyond the matrices, wrong precision, C aliasing an
 * input) are rejected here with MUBLIS_EINVAL. */
int mublis_submit(mublis_server_t srv, const mublis_request_t *req, uint64_t *ticket);
/* Replay a request trace from the calling thread: request i is submitted
 * (as by mublis_submit, ticket into tickets[i]) arrival_s[i] seconds after
 * the call, sleeping between arrivals; returns once the last is submitted.
 * Offsets must lie in [0, 86400] (else MUBLIS_EINVAL before any submission).
 * Stops at the first rejected request and returns its status; *submitted
 * counts the requests queued (their tickets are valid), *max_lag_ms the
 * largest delay of a submission behind its arrival time.  No counterpart in
 * the reference: a load generator for capacity and latency measurement whose
 * arrivals do not depend on the caller's own threads being scheduled on
 * time (an interpreter's client threads delayed arrivals by up to 30 ms). */
int mublis_server_replay(mublis_server_t srv, const mublis_request_t *reqs,
                         const double *arrival_s, int n, uint64_t *tickets, int *submitted,
                         double *max_lag_ms);
/* Block until the request completes; returns its status and (optionally)
 * its latency from submission to completion.  Each ticket is waited once. */
int mublis_wait(mublis_server_t srv, uint64_t ticket, double *latency_ms);
/* Counters and the latency distribution of the requests completed since the
 * server was created or mublis_server_reset_stats was called. */
int mublis_server_stats(mublis_server_t srv, mublis_server_stats_t *out);
int mublis_server_reset_stats(mublis_server_t srv);

#ifdef __cplusplus
}
#endif
#endif /* MUBLIS_GFX950_H */
