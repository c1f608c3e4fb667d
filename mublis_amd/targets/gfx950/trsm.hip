// gfx950 triangular solve: B := alpha * A^{-1} * B for a lower or upper
// triangular A on the left (the form the reference reduces every TRSM to,
// frame/l3/trsm.c:35-54).
//
// The reference walks kc-row diagonal blocks, solving each with the fused
// GEMMTRSM micro-kernel and pushing a rank-kc GEMM update through the rest
// of B (frame/l3/trsm.c:91-431).  On the GPU a rank-64 update of all of B
// per block would re-stream B from HBM m/64 times, so the schedule is
// recursive instead: split A in halves, solve the first, update the second
// half of B with ONE large MFMA GEMM (K = m/2), recurse.  All but O(m * NB * n)
// of the flops run in big GEMMs; the NB x NB diagonal blocks are solved by
// a leaf kernel in FP64 (also for float data), one lane per column of B,
// multiplying by reciprocal diagonals like the reference's GEMMTRSM
// (targets/reference/gemmtrsm_ukr.inc:29-40, frame/l1m/packm.c:143-144).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gfx950_common.h"

namespace {

constexpr int NB = 64;  // diagonal block (leaf) size = columns per leaf workgroup

// diag_mode: 0 read the diagonal and use its reciprocal, 1 unit (never read),
// 2 the stored diagonal is already the reciprocal (GEMMTRSM micro-kernel
// contract, include/mublis_instantiate.h:84-86).
//
// Prep: every NB x NB diagonal block of A in FP64, column-major, with zeros
// outside the triangle AND on the diagonal, followed by the NB reciprocal
// diagonal entries (1 for UNIT, never read) -- the device counterpart of
// packm_trsm_a (frame/l1m/packm.c:114-157), for all blocks in one launch.
// Only the stored triangle of A is read.
constexpr int BLK = NB * NB + NB;  // doubles per prepared block

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_prep(int m, const T *__restrict__ a,
                                                 int64_t rs_a, int64_t cs_a,
                                                 int diag_mode,
                                                 double *__restrict__ out) {
  const int blk = blockIdx.x;
  const int base = blk * NB;
  const int nb = min(NB, m - base);
  double *o = out + (int64_t)blk * BLK;
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int p = t / NB, i = t % NB;  // o[p*NB + i] = A(base+i, base+p)
    double v = 0.0;
    if (i < nb && p < nb && (LOWER ? (i > p) : (i < p)))
      v = (double)a[(int64_t)(base + i) * rs_a + (int64_t)(base + p) * cs_a];
    o[t] = v;
  }
  for (int i = threadIdx.x; i < NB; i += blockDim.x) {
    double v = 0.0;
    if (i < nb) {
      if (diag_mode == 1) v = 1.0;
      else {
        const int64_t g = base + i;
        const double d = (double)a[g * rs_a + g * cs_a];
        v = diag_mode == 2 ? d : 1.0 / d;
      }
    }
    o[NB * NB + i] = v;
  }
}

// Broadcast lane `src` of every group of 4 lanes to the whole group
// (DPP quad permute; `src` must be a constant after unrolling).
__device__ inline double quad_bcast(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  switch (src) {
    case 0:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0x00, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0x00, 0xF, 0xF, false);
      break;
    case 1:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0x55, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0x55, 0xF, 0xF, false);
      break;
    case 2:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0xAA, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0xAA, 0xF, 0xF, false);
      break;
    default:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0xFF, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0xFF, 0xF, 0xF, false);
      break;
  }
  return __hiloint2double(hi, lo);
}

// Leaf: solve one prepared NB x NB block against LC = 16 columns of B per
// workgroup of 64 lanes, in FP64.  Lane 4c + q owns column c and rows
// 16q .. 16q + 15.  The substitution runs column by column: x_p is finished
// by its owner lane (times the reciprocal diagonal), broadcast to the other
// three lanes of the column with a DPP quad permute, and every lane then
// updates its 16 rows with 16 independent FMAs (rows the triangle does not
// reach have a zero coefficient).  The B tile moves through LDS so global
// accesses run along B's unit stride.
constexpr int LC = 16;

template <typename T, bool LOWER>
__global__ __launch_bounds__(64) void trsm_leaf(int nb, int n, T alpha,
                                                const double *__restrict__ blk,
                                                T *__restrict__ b, int64_t rs_b,
                                                int64_t cs_b) {
  __shared__ __attribute__((aligned(16))) double as[NB * NB];
  __shared__ double dinv[NB];
  __shared__ double xs[NB][LC + 1];
  const int tid = threadIdx.x;
  {
    // 32 independent 16-byte loads per lane in flight, then the LDS stores
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 *src = reinterpret_cast<const d2 *>(blk);
    d2 *dst = reinterpret_cast<d2 *>(as);
    d2 tmp[NB * NB / 2 / 64];
#pragma unroll
    for (int u = 0; u < NB * NB / 2 / 64; ++u) tmp[u] = src[u * 64 + tid];
#pragma unroll
    for (int u = 0; u < NB * NB / 2 / 64; ++u) dst[u * 64 + tid] = tmp[u];
  }
  dinv[tid] = blk[NB * NB + tid];
  const int j0 = blockIdx.x * LC;
  if (rs_b == 1) {
#pragma unroll
    for (int c = 0; c < LC; ++c) {
      double v = 0.0;
      if (tid < nb && j0 + c < n) v = (double)b[(int64_t)tid + (int64_t)(j0 + c) * cs_b];
      xs[tid][c] = v;
    }
  } else {
#pragma unroll
    for (int it = 0; it < NB * LC / 64; ++it) {
      const int lin = it * 64 + tid, i = lin / LC, c = lin % LC;
      double v = 0.0;
      if (i < nb && j0 + c < n)
        v = (double)b[(int64_t)i * rs_b + (int64_t)(j0 + c) * cs_b];
      xs[i][c] = v;
    }
  }
  __syncthreads();
  const int c = tid >> 2, q = tid & 3;
  const double al = (double)alpha;
  double x[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) x[r] = al * xs[16 * q + r][c];
#pragma unroll
  for (int s = 0; s < NB; ++s) {
    const int p = LOWER ? s : NB - 1 - s;
    const int qp = p >> 4, rp = p & 15;
    const double xp = quad_bcast(x[rp] * dinv[p], qp);
    x[rp] = (q == qp) ? xp : x[rp];
    const double *colp = &as[p * NB + 16 * q];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = fma(-colp[r], xp, x[r]);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) xs[16 * q + r][c] = x[r];
  __syncthreads();
  if (rs_b == 1) {
#pragma unroll
    for (int cc = 0; cc < LC; ++cc)
      if (tid < nb && j0 + cc < n)
        b[(int64_t)tid + (int64_t)(j0 + cc) * cs_b] = (T)xs[tid][cc];
  } else {
#pragma unroll
    for (int it = 0; it < NB * LC / 64; ++it) {
      const int lin = it * 64 + tid, i = lin / LC, cc = lin % LC;
      if (i < nb && j0 + cc < n)
        b[(int64_t)i * rs_b + (int64_t)(j0 + cc) * cs_b] = (T)xs[i][cc];
    }
  }
}

template <typename T>
int prep(bool lower, int m, const T *a, int64_t rs_a, int64_t cs_a, int diag_mode,
         double *out, hipStream_t s) {
  const int blocks = (m + NB - 1) / NB;
  if (lower)
    hipLaunchKernelGGL((trsm_prep<T, true>), dim3(blocks), dim3(256), 0, s, m, a, rs_a,
                       cs_a, diag_mode, out);
  else
    hipLaunchKernelGGL((trsm_prep<T, false>), dim3(blocks), dim3(256), 0, s, m, a, rs_a,
                       cs_a, diag_mode, out);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

template <typename T>
int leaf(bool lower, int nb, int n, T alpha, const double *ad, T *b, int64_t rs_b,
         int64_t cs_b, hipStream_t s) {
  const int grid = (n + LC - 1) / LC;
  if (lower)
    hipLaunchKernelGGL((trsm_leaf<T, true>), dim3(grid), dim3(NB), 0, s, nb, n, alpha, ad,
                       b, rs_b, cs_b);
  else
    hipLaunchKernelGGL((trsm_leaf<T, false>), dim3(grid), dim3(NB), 0, s, nb, n, alpha,
                       ad, b, rs_b, cs_b);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

// `ad` holds the prepared diagonal blocks of this sub-problem (block 0 is
// its top-left NB x NB block).
template <typename T>
int trsm_rec(bool lower, const double *ad, int m, int n, T alpha, const T *a,
             int64_t rs_a, int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b,
             hipStream_t s) {
  if (m <= NB) return leaf<T>(lower, m, n, alpha, ad, b, rs_b, cs_b, s);
  // First part: a multiple of NB close to half, so GEMM K and M stay tile
  // friendly at every level.
  int m1 = (m / 2 + NB - 1) / NB * NB;
  if (m1 >= m) m1 = m - NB;
  const int m2 = m - m1;
  const T *a11 = a;
  const T *a21 = a + (int64_t)m1 * rs_a;
  const T *a12 = a + (int64_t)m1 * cs_a;
  const T *a22 = a + (int64_t)m1 * rs_a + (int64_t)m1 * cs_a;
  T *b1 = b, *b2 = b + (int64_t)m1 * rs_b;
  const double *ad1 = ad, *ad2 = ad + (int64_t)(m1 / NB) * BLK;
  int st;
  if (lower) {
    if ((st = trsm_rec<T>(lower, ad1, m1, n, alpha, a11, rs_a, cs_a, b1, rs_b, cs_b,
                          s)))
      return st;
    // B2 := alpha*B2 - A21 * X1 (the reference's GEMM with alpha -1 and
    // beta = alpha_eff, frame/l3/trsm.c:165-197, 241-249)
    if ((st = gfx950_gemm<T>(m2, n, m1, T(-1), a21, rs_a, cs_a, b1, rs_b, cs_b, alpha,
                             b2, rs_b, cs_b, s)))
      return st;
    return trsm_rec<T>(lower, ad2, m2, n, T(1), a22, rs_a, cs_a, b2, rs_b, cs_b, s);
  }
  if ((st = trsm_rec<T>(lower, ad2, m2, n, alpha, a22, rs_a, cs_a, b2, rs_b, cs_b,
                        s)))
    return st;
  if ((st = gfx950_gemm<T>(m1, n, m2, T(-1), a12, rs_a, cs_a, b2, rs_b, cs_b, alpha,
                           b1, rs_b, cs_b, s)))
    return st;
  return trsm_rec<T>(lower, ad1, m1, n, T(1), a11, rs_a, cs_a, b1, rs_b, cs_b, s);
}

// Prepare all diagonal blocks into `ad` (ws_doubles(m) doubles) and run the
// recursive schedule; asynchronous on s.
int64_t ws_doubles(int m) { return (int64_t)((m + NB - 1) / NB) * BLK; }

template <typename T>
int solve_ws(bool lower, int diag_mode, int m, int n, T alpha, const T *a, int64_t rs_a,
             int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, double *ad, hipStream_t s) {
  int st = prep<T>(lower, m, a, rs_a, cs_a, diag_mode, ad, s);
  if (st == MUBLIS_OK) st = trsm_rec<T>(lower, ad, m, n, alpha, a, rs_a, cs_a, b, rs_b,
                                        cs_b, s);
  return st;
}

// Same with workspace from the arena, released once the stream has drained.
template <typename T>
int solve(bool lower, int diag_mode, int m, int n, T alpha, const T *a, int64_t rs_a,
          int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, hipStream_t s) {
  double *ad = (double *)mublis_rt_alloc(sizeof(double) * (size_t)ws_doubles(m));
  if (!ad) return MUBLIS_ENOMEM;
  int st = solve_ws<T>(lower, diag_mode, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  mublis_rt_free(ad);
  return st;
}

template <typename T>
int trsm_dev(int lower, int unit, int m, int n, T alpha, const T *a, int64_t rs_a,
             int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, void *stream) {
  if (m < 0 || n < 0) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (alpha == T(0)) return gfx950_fill<T>(m, n, T(0), b, rs_b, cs_b, s);
  return solve<T>(lower != 0, unit ? 1 : 0, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, s);
}

}  // namespace

extern "C" int mublis_gfx950_dtrsm_inverted_dev(int lower, int m, int n,
                                                const double *a, int64_t rs_a,
                                                int64_t cs_a, double *b, int64_t rs_b,
                                                int64_t cs_b, void *stream) {
  if (m < 0 || n < 0 || m > NB) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  return solve<double>(lower != 0, 2, m, n, 1.0, a, rs_a, cs_a, b, rs_b, cs_b,
                       (hipStream_t)(stream ? stream : mublis_rt_stream()));
}
extern "C" int mublis_gfx950_strsm_inverted_dev(int lower, int m, int n,
                                                const float *a, int64_t rs_a,
                                                int64_t cs_a, float *b, int64_t rs_b,
                                                int64_t cs_b, void *stream) {
  if (m < 0 || n < 0 || m > NB) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  return solve<float>(lower != 0, 2, m, n, 1.0f, a, rs_a, cs_a, b, rs_b, cs_b,
                      (hipStream_t)(stream ? stream : mublis_rt_stream()));
}

extern "C" int mublis_gfx950_dtrsm_dev(int lower, int unit, int m, int n, double alpha,
                                       const double *a, int64_t rs_a, int64_t cs_a,
                                       double *b, int64_t rs_b, int64_t cs_b,
                                       void *stream) {
  return trsm_dev<double>(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                          stream);
}
extern "C" int mublis_gfx950_strsm_dev(int lower, int unit, int m, int n, float alpha,
                                       const float *a, int64_t rs_a, int64_t cs_a,
                                       float *b, int64_t rs_b, int64_t cs_b,
                                       void *stream) {
  return trsm_dev<float>(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                         stream);
}

extern "C" size_t mublis_gfx950_trsm_ws_bytes(int m) {
  return m <= 0 ? 0 : sizeof(double) * (size_t)ws_doubles(m);
}

template <typename T>
static int trsm_ws_dev(int lower, int unit, int m, int n, T alpha, const T *a, int64_t rs_a,
                       int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, void *ws,
                       size_t ws_bytes, void *stream) {
  if (m < 0 || n < 0) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (alpha == T(0)) return gfx950_fill<T>(m, n, T(0), b, rs_b, cs_b, s);
  if (!ws || ws_bytes < mublis_gfx950_trsm_ws_bytes(m)) return MUBLIS_EINVAL;
  return solve_ws<T>(lower != 0, unit ? 1 : 0, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                     (double *)ws, s);
}

extern "C" int mublis_gfx950_dtrsm_ws_dev(int lower, int unit, int m, int n, double alpha,
                                          const double *a, int64_t rs_a, int64_t cs_a,
                                          double *b, int64_t rs_b, int64_t cs_b, void *ws,
                                          size_t ws_bytes, void *stream) {
  return trsm_ws_dev<double>(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ws,
                             ws_bytes, stream);
}
extern "C" int mublis_gfx950_strsm_ws_dev(int lower, int unit, int m, int n, float alpha,
                                          const float *a, int64_t rs_a, int64_t cs_a, float *b,
                                          int64_t rs_b, int64_t cs_b, void *ws, size_t ws_bytes,
                                          void *stream) {
  return trsm_ws_dev<float>(lower, unit, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ws,
                            ws_bytes, stream);
}
