// Shared declarations of the gfx950 HIP layer (device-side helpers used by
// more than one translation unit).  Explicit instantiations for float and
// double live in pack.hip.
#ifndef MUBLIS_GFX950_COMMON_H
#define MUBLIS_GFX950_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mublis_gfx950.h"

// Set the calling thread's mublis_rt_last_error() text.
void gfx950_set_error(const char *msg);

// dst(i, j) := src(i, j) for an m x n matrix; any strides.
template <typename T>
int gfx950_copy(int m, int n, const T *src, int64_t rs_s, int64_t cs_s, T *dst,
                int64_t rs_d, int64_t cs_d, hipStream_t s);

// Materialise an operand view (GENERAL / SYMMETRIC / TRIANGULAR) densely.
template <typename T>
int gfx950_pack(int m, int n, const T *src, int64_t rs, int64_t cs, int struc,
                int uplo, int diag, int diagoff, int mask, T *dst, int64_t ld,
                hipStream_t s);

template <typename T>
int gfx950_fill(int m, int n, T value, T *x, int64_t rs, int64_t cs, hipStream_t s);

// GEMM entry shared by the TRSM/TRMM schedules (same contract as
// mublis_gfx950_?gemm_dev).
template <typename T>
int gfx950_gemm(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                int64_t cs_c, hipStream_t s);
// Same, except that rows lo <= i < hi of the result go to c2 (its row
// i - lo) instead of c; beta still reads them from c.  lo a multiple of 128,
// hi too unless hi >= m; k > 0 and alpha != 0.
template <typename T>
int gfx950_gemm_split_c(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                        const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                        int64_t cs_c, T *c2, int64_t rs_c2, int64_t cs_c2, int lo, int hi,
                        hipStream_t s);
// `count` problems of one shape in one launch; operand X of problem z starts
// at x + (z % bat) * bsx[0] + (z / bat) * bsx[1].  A and B need a unit
// stride.  tri_a / tri_b as for gfx950_gemm_tri_a.
template <typename T>
int gfx950_gemm_batched(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                        const int64_t bsa[2], const T *b, int64_t rs_b, int64_t cs_b,
                        const int64_t bsb[2], T beta, T *c, int64_t rs_c, int64_t cs_c,
                        const int64_t bsc[2], int count, int bat, int tri_a, int tri_b,
                        hipStream_t s);
// Same with a triangular A whose zero triangle's K slices are skipped
// (tri_a 1: A(i,p) = 0 for p > i, 2: for p < i).
template <typename T>
int gfx950_gemm_tri_a(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                      const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                      int64_t cs_c, int tri_a, hipStream_t s);

// Precision helpers (refined STRSM): FP32 <-> FP64 copies, FP32 triangular
// view materialised in FP64, and x64 += d; out := (float) x64.
int gfx950_convert_f2d(int m, int n, const float *src, int64_t rs_s, int64_t cs_s,
                       double *dst, int64_t rs_d, int64_t cs_d, hipStream_t s);
int gfx950_convert_d2f(int m, int n, const double *src, int64_t rs_s, int64_t cs_s,
                       float *dst, int64_t rs_d, int64_t cs_d, hipStream_t s);
int gfx950_pack_f2d(int m, int n, const float *src, int64_t rs, int64_t cs, int struc,
                    int uplo, int diag, double *dst, int64_t ld, hipStream_t s);
int gfx950_refine_update(int m, int n, const float *d, double *x64, float *out,
                         int64_t rs_o, int64_t cs_o, hipStream_t s);

// Left-side FP32 TRSM in caller workspace of ws_bytes (the full schedule
// with gfx950_strsm_keep_ws_bytes(m, n) bytes), asynchronous on s.  With
// *inv_ready false the solve inverts the triangle's diagonal blocks and sets
// it when they stay valid for more solves against the same triangle, n and
// triangle flags; with *inv_ready true it reuses them.
size_t gfx950_strsm_keep_ws_bytes(int m, int n);
int gfx950_strsm_keep(bool lower, bool unit, int m, int n, float alpha, const float *a,
                      int64_t rs_a, int64_t cs_a, float *b, int64_t rs_b, int64_t cs_b,
                      void *ws, size_t ws_bytes, bool *inv_ready, hipStream_t s);

// Grouped triangular solves (the batched interface's TRSM problems), each in
// the left form B := alpha * inv(A) * B: A m x m lower (lower = 1) or upper
// through (rs_a, cs_a), unit diagonal when unit = 1, B m x n through (rs_b,
// cs_b).  The problem's FP64 inverted 64-row diagonal blocks live `inv`
// doubles into the launch's inverse workspace (ceil(m / 64) * 64 * 64
// doubles per problem).
template <typename T>
struct TrsmGroupProb {
  int lower, unit, m, n;
  T alpha;
  const T *a;
  int64_t rs_a, cs_a;
  T *b;
  int64_t rs_b, cs_b;
  int64_t inv;
};
// Work item: problem, and its diagonal block (inverse launch) or column
// panel (solve launch).
struct TrsmGroupItem {
  int prob, idx;
};
// Panel classes by the triangle's order: up to kGroupTrsmWide, panels of
// kGroupTrsmWideCols columns; up to kGroupTrsmSmall, of kGroupTrsmCols
// columns (both two workgroups per CU); up to kGroupTrsmMax, of
// kGroupTrsmCols columns at one workgroup per CU.  Larger problems are
// solved by themselves.
constexpr int kGroupTrsmCols = 16;
constexpr int kGroupTrsmWideCols = 32;
constexpr int kGroupTrsmWide = 256;
constexpr int kGroupTrsmSmall = 512;
constexpr int kGroupTrsmMax = 1024;
constexpr int kGroupTrsmLists = 8;
// Launches over kGroupTrsmLists work lists stored back to back in `items`
// (device memory), counts[c] items each: the inverse items of lower, then
// of upper problems; then, for each panel class in the order above, the
// panels of its lower, then of its upper problems.
// Asynchronous on s, nothing allocated.  Given a second stream s2 and two
// events, the upper problems' launches run on s2, forked from s and joined
// back into s (concurrent graph branches when s is being captured).
template <typename T>
int gfx950_trsm_grouped(const TrsmGroupProb<T> *probs, const TrsmGroupItem *items,
                        const int *counts, double *inv, hipStream_t s,
                        hipStream_t s2 = nullptr, hipEvent_t fork = nullptr,
                        hipEvent_t join = nullptr);

#endif
