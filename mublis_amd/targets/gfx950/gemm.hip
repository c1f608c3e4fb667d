// gfx950 GEMM kernel family: the device form of MuBLIS's generic L3 driver.
//
// The reference's 5-loop BLIS nest (frame/l3/l3_driver.c:126-239) becomes:
//   jc / ic blocks     -> a grid of 128 x 128 output tiles, one workgroup
//                         each, mapped XCD by XCD so that the 8 L2s each see
//                         a compact group of tiles;
//   pc + packm         -> the K loop: 128 x BK slices of A and B are loaded
//                         from HBM with 16-byte vector loads and staged in
//                         LDS ([k][m] and [k][n], double-buffered);
//   jr / ir            -> 4 wavefronts in a 2 x 2 layout, each owning 64 x 64
//                         (FP64 128-tiles and full-depth 64-tiles: 8 in
//                         2 x 4, each 64 x 32 or 32 x 16);
//   micro-kernel p loop-> chains of v_mfma_f64_16x16x4_f64 (FP64) or
//                         v_mfma_f32_16x16x4_f32 (FP32), 16 (8) tiles per wave;
//   beta pre-pass, c_buf edge tiles, ji masks
//                      -> one predicated epilogue: C = alpha*acc + beta*C,
//                         beta == 0 never reads C, elements outside the ji
//                         relation are never written (l3_driver.c:88-121,
//                         192-233).
// SYRK/SYR2K launch only the tiles that touch the ji triangle.  SYR2K adds
// its second product (frame/l3/syr2k.c:47-68) as a second K segment in the
// same launch, so C is read and written once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>

#include "mublis_gfx950.h"
#include "gfx950_common.h"

#include "gemm_tile.h"

namespace {

using namespace gfx950;

// K slice depth and workgroups per CU of the one-problem kernel.  SHALLOW
// halves the slice, which halves the LDS ring and frees the operand
// registers for more resident workgroups: 64-tiles 20 KiB and 78 (FP64) / 57
// (FP32) VGPRs at 6 per CU instead of 4; FP32 128-tiles 36 KiB and 127 VGPRs
// at 4 per CU instead of 2 (FP64 128-tiles would spill at 3).  Measured on
// MI355X (TFLOP/s): SGEMM 8192 op(A) = A^T 126.3 -> 129.8, NN 123.5 ->
// 129.1, 4096 119.8 -> 125, 3072 x 3072 x 1024 74 -> 81, 8192 x 8192 x
// 1024 112 -> 121; DGEMM 2048 x 3072 x 1024 49.4 -> 54.4, 1536 x 4096 x
// 1536 52.0 -> 56.3; DSYRK 3072 47.0 -> 52.7; SSYRK 3072 83.4 -> 93.5.  But
// launches with a triangular operand (long-first walk, the TRSM leaves) ran
// 13 % slower, and grids that the deep kernel already holds in one round
// (64-tiles: at most four per CU; FP32 128-tiles: two per CU; the few long
// SYRK diagonal tiles, the smaller updates of a TRSM recursion) gain no
// residency and lose 1-5 % (FP64 1024 x 4096 x 1024, exactly four per CU:
// 161.7 us shallow, 153.5 us deep): those keep the deep slice.  The slice
// depth never changes the k order of the MFMAs.
template <typename T, int TILE>
bool use_shallow(const GemmArgs<T> &p, int64_t wgs) {
  if (p.lpt || TILE == 32) return false;
  if (TILE == 64) return wgs > 1024;
  return sizeof(T) == 4 && wgs > 512;
}
//
// FP64 128-tiles run 8 waves (512 threads, each wave 64 x 32) instead of 4:
// the same 2 workgroups per CU hold 4 waves per SIMD instead of 2, at 116-128
// VGPRs.  Measured on MI355X (TFLOP/s): DGEMM 8192 op(A) = A^T 63.4 -> 65.5,
// NN 63.0 -> 65.5, 4096 62.2 -> 64.5, DSYRK 8192 60.0 -> 63-64; DTRSM 8192 x
// 4096 5.22 -> 5.09 ms.  FP64 full-depth 64-tiles (the TRSM leaves and
// smaller levels, SYRK diagonal blocks) likewise run 8 waves of 32 x 16 at 4
// workgroups per CU, 8 waves per SIMD at 53-64 VGPRs: DTRSM 8192 x 4096 5.07
// -> 4.99 ms, 8192 x 1024 2.00 -> 1.94, 4096 x 4096 1.54 -> 1.51; DSYRK 4096
// / 2048 +2.6 %.  FP32 full-depth 64-tiles too (SSYRK 4096 / 2048 +3.6 %,
// SGEMM within noise).
template <typename T, int TILE, bool SHALLOW>
struct KernelShape {
  static_assert(!SHALLOW || TILE == 64 || sizeof(T) == 4, "FP64 128-tiles: deep only");
  static constexpr int BK = SHALLOW ? Mfma<T>::BK / 2 : Mfma<T>::BK;
  static constexpr int WG_PER_CU =
      SHALLOW ? (TILE == 64 ? 6 : 4) : Tile<TILE>::MIN_WG_PER_CU;
  static constexpr int NTH = ((sizeof(T) == 8 && TILE == 128) || (TILE == 64 && !SHALLOW)) ? 512 : NT;
  // __launch_bounds__'s second argument: minimum waves per SIMD
  static constexpr int WAVES_PER_SIMD = WG_PER_CU * NTH / 256;
};

template <typename T, int TILE, bool A_KC, bool B_NC, bool SHALLOW>
__global__ __launch_bounds__((KernelShape<T, TILE, SHALLOW>::NTH),
                             (KernelShape<T, TILE, SHALLOW>::WAVES_PER_SIMD)) void gemm_kernel(
    GemmArgs<T> p) {
  constexpr int BK = KernelShape<T, TILE, SHALLOW>::BK;
  __shared__ T lds[LdsSize<T, TILE, BK>::value];
  if (gridDim.y > 1) {
    const int64_t zi = blockIdx.y % p.bat, zo = blockIdx.y / p.bat;
    p.a += zi * p.bsa[0] + zo * p.bsa[1];
    p.b += zi * p.bsb[0] + zo * p.bsb[1];
    p.c += zi * p.bsc[0] + zo * p.bsc[1];
  }
  int tm, tn;
  if (p.lpt) {
    // batched: problems of at most 32 workgroups per XCD share rows of 32
    const int per_xcd = (int)gridDim.x / 8;
    const bool rev = gridDim.y > 1 && per_xcd <= 32 && 32 % per_xcd == 0 &&
                     ((int)blockIdx.y * per_xcd / 32) % 2 == 1;
    if (!map_tile_lpt(blockIdx.x, p.tiles_m, p.tiles_n, p.lpt, p.lpt_g, tm, tn, p.lpt_r1, rev))
      return;
  } else {
    map_tile(blockIdx.x, p.ntiles, p.tiles_m, p.tiles_n, p.tri, p.ji, tm, tn);
    if (tm >= p.tiles_m || tn >= p.tiles_n) return;  // mode 3 past a short last block
  }
  gemm_tile<T, TILE, A_KC, B_NC, BK, KernelShape<T, TILE, SHALLOW>::NTH>(p, tm, tn, lds);
}

// beta-only update on the ji part of C (alpha == 0 or k == 0): the
// reference's pre-pass for tiles that get no product
// (frame/l3/l3_driver.c:88-121).  beta == 0 stores zeros without reading.
template <typename T>
__global__ void scale_kernel(int m, int n, T beta, T *c, int64_t rs, int64_t cs,
                             int ji) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % m), j = (int)(t / m);
    if (ji == 1 && j > i) continue;
    if (ji == 2 && j < i) continue;
    T *cp = c + (int64_t)i * rs + (int64_t)j * cs;
    *cp = beta == T(0) ? T(0) : beta * *cp;
  }
}

template <typename T>
bool aligned16(const T *x, int64_t ld) {
  return ((uintptr_t)x % 16 == 0) && ((ld * (int64_t)sizeof(T)) % 16 == 0);
}

template <typename T, int TILE, bool SHALLOW>
void launch_layout(const GemmArgs<T> &p, bool akc, bool bnc, dim3 grid, hipStream_t s) {
  const dim3 block(KernelShape<T, TILE, SHALLOW>::NTH);
  if (akc && bnc)
    hipLaunchKernelGGL((gemm_kernel<T, TILE, true, true, SHALLOW>), grid, block, 0, s, p);
  else if (akc)
    hipLaunchKernelGGL((gemm_kernel<T, TILE, true, false, SHALLOW>), grid, block, 0, s, p);
  else if (bnc)
    hipLaunchKernelGGL((gemm_kernel<T, TILE, false, true, SHALLOW>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<T, TILE, false, false, SHALLOW>), grid, block, 0, s, p);
}

template <typename T, int TILE>
hipError_t launch(GemmArgs<T> p, bool akc, bool bnc, hipStream_t s, int nbatch = 1) {
  p.tiles_m = (p.m + TILE - 1) / TILE;
  p.tiles_n = (p.n + TILE - 1) / TILE;
  const int t = p.tiles_m;
  p.ntiles = p.tri == 1 ? t * (t + 1) / 2
           : p.tri == 2 ? t * (t - 1) / 2
           : p.tri == 3 ? 10 * ((t + 3) / 4)
                        : p.tiles_m * p.tiles_n;
  if (p.ntiles == 0) return hipSuccess;
  int wgs = p.ntiles;
  if (p.lpt) {
    const int lines = p.lpt <= 2 ? p.tiles_m : p.tiles_n;
    const int per = p.lpt <= 2 ? p.tiles_n : p.tiles_m;
    p.lpt_g = lpt_unit(lines, per);
    const int64_t units = (int64_t)lines * ((per + p.lpt_g - 1) / p.lpt_g);
    wgs = (int)(8 * ((units + 7) / 8) * p.lpt_g);
    // (long-first launches always take the deep slice)
    p.lpt_r1 = nbatch == 1 ? 32 * KernelShape<T, TILE, false>::WG_PER_CU : 0;
  }
  const dim3 grid(wgs, nbatch);
  if constexpr ((TILE == 64 || sizeof(T) == 4) && TILE != 32) {
    if (use_shallow<T, TILE>(p, (int64_t)wgs * nbatch)) {
      launch_layout<T, TILE, true>(p, akc, bnc, grid, s);
      return hipGetLastError();
    }
  }
  launch_layout<T, TILE, false>(p, akc, bnc, grid, s);
  return hipGetLastError();
}

// Pack an operand without a unit stride into a dense column-major buffer.
template <typename T>
int densify(int rows, int cols, const T *x, int64_t rs, int64_t cs, T **out,
            int64_t *ors, int64_t *ocs, void **owned, hipStream_t s);

template <typename T>
int gemm_dev(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
             const T *b, int64_t rs_b, int64_t cs_b, const T *a2, int64_t rs_a2,
             int64_t cs_a2, const T *b2, int64_t rs_b2, int64_t cs_b2, T beta,
             T *c, int64_t rs_c, int64_t cs_c, int ji, void *stream, int tri_a = 0,
             int tri_b = 0, T *c2 = nullptr, int64_t rs_c2 = 0, int64_t cs_c2 = 0,
             int c2_lo = 0, int c2_hi = 0) {
  if (m < 0 || n < 0 || k < 0) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (alpha == T(0) || k == 0) {
    const int64_t total = (int64_t)m * n;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(scale_kernel<T>, dim3(grid), dim3(256), 0, s, m, n, beta, c,
                       rs_c, cs_c, ji);
    return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
  }
  void *owned[4] = {nullptr, nullptr, nullptr, nullptr};
  int st = MUBLIS_OK;
  // Operands need one unit stride for the staged loads.
  T *da = nullptr, *db = nullptr, *da2 = nullptr, *db2 = nullptr;
  if (rs_a != 1 && cs_a != 1) {
    st = densify<T>(m, k, a, rs_a, cs_a, &da, &rs_a, &cs_a, &owned[0], s);
    a = da;
  }
  if (st == MUBLIS_OK && rs_b != 1 && cs_b != 1) {
    st = densify<T>(k, n, b, rs_b, cs_b, &db, &rs_b, &cs_b, &owned[1], s);
    b = db;
  }
  if (st == MUBLIS_OK && a2 && rs_a2 != 1 && cs_a2 != 1) {
    st = densify<T>(m, k, a2, rs_a2, cs_a2, &da2, &rs_a2, &cs_a2, &owned[2], s);
    a2 = da2;
  }
  if (st == MUBLIS_OK && b2 && rs_b2 != 1 && cs_b2 != 1) {
    st = densify<T>(k, n, b2, rs_b2, cs_b2, &db2, &rs_b2, &cs_b2, &owned[3], s);
    b2 = db2;
  }
  if (st == MUBLIS_OK) {
    // A: K-contiguous whenever cs_a == 1, else M-contiguous (rs_a == 1).
    // B: K-contiguous whenever rs_b == 1, else N-contiguous (cs_b == 1).
    // A second product is read with the same layout when one suits both
    // (always so for CBLAS SYR2K, whose operands share trans; strides of
    // extent-1 dimensions are normalised for it), else it is added in a
    // launch of its own with beta = 1.
    bool same2 = true;
    int la = -1, lb = -1;
    if (a2) {
      la = common_layout(m, k, rs_a, cs_a, &rs_a2, &cs_a2, true);
      lb = common_layout(k, n, rs_b, cs_b, &rs_b2, &cs_b2, false);
      same2 = la >= 0 && lb >= 0;
    }
    if (!same2 || !a2) {
      la = common_layout(m, k, rs_a, cs_a, nullptr, nullptr, true);
      lb = common_layout(k, n, rs_b, cs_b, nullptr, nullptr, false);
    }
    const bool akc = la == 1;
    const bool bnc = lb == 1;
    GemmArgs<T> p;
    p.m = m; p.n = n; p.k = k; p.alpha = alpha; p.beta = beta;
    p.a = a; p.rs_a = rs_a; p.cs_a = cs_a;
    p.b = b; p.rs_b = rs_b; p.cs_b = cs_b;
    p.a2 = same2 ? a2 : nullptr; p.rs_a2 = rs_a2; p.cs_a2 = cs_a2;
    p.b2 = same2 ? b2 : nullptr; p.rs_b2 = rs_b2; p.cs_b2 = cs_b2;
    p.c = c; p.rs_c = rs_c; p.cs_c = cs_c; p.ji = ji;
    p.c2 = c2; p.rs_c2 = rs_c2; p.cs_c2 = cs_c2;
    p.c2_lo = c2 ? c2_lo : 0;
    p.c2_hi = c2 ? c2_hi : 0;
    p.tri = (ji != 0 && m == n) ? 1 : 0;
    p.tri_a = a2 ? 0 : tri_a;
    p.tri_b = a2 ? 0 : tri_b;
    // A triangular operand makes tile costs uneven (K from BK to k): walk
    // the tiles long-first, balanced over the XCDs.
    p.lpt = 0;
    p.lpt_g = 1;
    p.lpt_r1 = 0;
    if (!p.tri) {
      if (p.tri_a) p.lpt = p.tri_a == 1 ? 1 : 2;
      else if (p.tri_b) p.lpt = p.tri_b == 2 ? 3 : 4;
    }
    p.vec_a = aligned16(a, akc ? rs_a : cs_a);
    p.vec_b = aligned16(b, bnc ? rs_b : cs_b);
    p.vec_a2 = a2 ? aligned16(a2, akc ? rs_a2 : cs_a2) : 0;
    p.vec_b2 = b2 ? aligned16(b2, bnc ? rs_b2 : cs_b2) : 0;
    // 128 x 128 tiles unless that grid would give fewer than two
    // workgroups per CU (256 CUs).  With uneven tiles (long-first walk) a
    // single round of 128 tiles lasts as long as its longest tile, so 128
    // tiles need at least four rounds there.
    const int64_t t128 = (int64_t)((m + 127) / 128) * ((n + 127) / 128);
    const int64_t n128 = p.tri ? ((m + 127) / 128) * ((m + 127) / 128 + 1) / 2 : t128;
    const int64_t need = p.lpt ? 2048 : 512;
    hipError_t e;
    if (p.tri && n128 >= need) {
      // The triangle's diagonal 128-tiles would compute twice what they
      // store and push the tile count past a round of 512: 128-tiles
      // strictly inside the triangle, then the diagonal blocks as ten
      // 32-tiles each (three 64-tiles each left most CUs idle: 192
      // workgroups at n = 8192).
      p.tri = 2;
      e = launch<T, 128>(p, akc, bnc, s);
      p.tri = 3;
      if (e == hipSuccess) e = launch<T, 32>(p, akc, bnc, s);
    } else {
      // 64-tile grids of at most 1.5 workgroups per CU take 32-tiles (4 per
      // CU, one to 1.5 rounds), the same MFMA steps per element.  Measured
      // on MI355X (TFLOP/s, 64- -> 32-tiles): SGEMM 1024 x 1024 x 4096 50.0
      // -> 56.4, 768 x 1500 x 2000 33.6 -> 40.4, 3000 x 500 x 3000 46.0 ->
      // 54.3; DGEMM the same shapes 33.6 -> 37.6, 23.2 -> 29.7, 31.2 ->
      // 36.0; STRSM 8192 x 1024 1.21 -> 1.08 ms (its 1024-row leaves and
      // updates), 4096 x 512 0.47 -> 0.36 ms, DTRSM 8192 x 1024 1.92 -> 1.71
      // ms.  With 400 or more 64-tiles (1280 x 1280 x 4096, 500 x 4000 x
      // 2500: 1.6-2 rounds of 32-tiles) they were 4-11 % slower.  The same
      // for a triangle of at most 384 64-tiles (SYRK/SYR2K n up to 1700),
      // but FP32 triangles only from K = 1025: DSYRK 512 / 2048 3.9 -> 5.3,
      // 1024 / 4096 17.6 -> 21.5, 1536 / 1536 23.7 -> 28.1; SSYRK 1024 /
      // 4096 27.8 -> 32.5, 1536 / 1536 32.6 -> 38.8, but 1024 / 1024 17.3 ->
      // 15.9 and 1700 / 512 23.8 -> 21.5.
      const int64_t t64 = (int64_t)((m + 63) / 64) * ((n + 63) / 64);
      const int64_t tri64 = (int64_t)((m + 63) / 64) * ((m + 63) / 64 + 1) / 2;
      const bool small32 = p.tri ? tri64 <= 384 && (sizeof(T) == 8 || k > 1024) : t64 <= 384;
      e = n128 >= need ? launch<T, 128>(p, akc, bnc, s)
          : small32    ? launch<T, 32>(p, akc, bnc, s)
                       : launch<T, 64>(p, akc, bnc, s);
    }
    if (e != hipSuccess) st = MUBLIS_EHIP;
    if (st == MUBLIS_OK && a2 && !same2) {
      // Layouts differ between the two products: add the second one in a
      // separate launch with beta = 1.
      st = gemm_dev<T>(m, n, k, alpha, a2, rs_a2, cs_a2, b2, rs_b2, cs_b2, nullptr, 0,
                       0, nullptr, 0, 0, T(1), c, rs_c, cs_c, ji, s);
    }
  }
  if (owned[0] || owned[1] || owned[2] || owned[3]) {
    // Workspace goes back to the arena only after the kernels that read it.
    if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
    for (void *o : owned) mublis_rt_free(o);
  }
  return st;
}

template <typename T>
int densify(int rows, int cols, const T *x, int64_t rs, int64_t cs, T **out,
            int64_t *ors, int64_t *ocs, void **owned, hipStream_t s) {
  T *d = (T *)mublis_rt_alloc(sizeof(T) * (size_t)std::max(rows, 1) *
                              (size_t)std::max(cols, 1));
  if (!d) return MUBLIS_ENOMEM;
  *owned = d;
  int st = gfx950_copy<T>(rows, cols, x, rs, cs, d, 1, std::max(rows, 1), s);
  *out = d;
  *ors = 1;
  *ocs = std::max(rows, 1);
  return st;
}

}  // namespace

extern "C" int mublis_gfx950_dgemm_dev(
    int m, int n, int k, double alpha, const double *a, int64_t rs_a, int64_t cs_a,
    const double *b, int64_t rs_b, int64_t cs_b, const double *a2, int64_t rs_a2,
    int64_t cs_a2, const double *b2, int64_t rs_b2, int64_t cs_b2, double beta,
    double *c, int64_t rs_c, int64_t cs_c, int ji, void *stream) {
  return gemm_dev<double>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, a2, rs_a2,
                          cs_a2, b2, rs_b2, cs_b2, beta, c, rs_c, cs_c, ji, stream);
}

extern "C" int mublis_gfx950_sgemm_dev(
    int m, int n, int k, float alpha, const float *a, int64_t rs_a, int64_t cs_a,
    const float *b, int64_t rs_b, int64_t cs_b, const float *a2, int64_t rs_a2,
    int64_t cs_a2, const float *b2, int64_t rs_b2, int64_t cs_b2, float beta,
    float *c, int64_t rs_c, int64_t cs_c, int ji, void *stream) {
  return gemm_dev<float>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, a2, rs_a2,
                         cs_a2, b2, rs_b2, cs_b2, beta, c, rs_c, cs_c, ji, stream);
}

template <typename T>
int gfx950_gemm(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                int64_t cs_c, hipStream_t s) {
  return gemm_dev<T>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                     nullptr, 0, 0, beta, c, rs_c, cs_c, 0, (void *)s);
}
template int gfx950_gemm<float>(int, int, int, float, const float *, int64_t, int64_t,
                                const float *, int64_t, int64_t, float, float *,
                                int64_t, int64_t, hipStream_t);
template int gfx950_gemm<double>(int, int, int, double, const double *, int64_t,
                                 int64_t, const double *, int64_t, int64_t, double,
                                 double *, int64_t, int64_t, hipStream_t);

template <typename T>
int gfx950_gemm_split_c(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                        const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                        int64_t cs_c, T *c2, int64_t rs_c2, int64_t cs_c2, int lo, int hi,
                        hipStream_t s) {
  // whole 128-row tiles (and so whole 64-row ones) on either side of both
  // bounds; the beta-only path (alpha == 0 or k == 0) has no split
  if (!c2 || k <= 0 || alpha == T(0) || lo < 0 || hi <= lo || lo % 128 != 0 ||
      (hi % 128 != 0 && hi < m))
    return MUBLIS_EINVAL;
  return gemm_dev<T>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0, nullptr, 0,
                     0, beta, c, rs_c, cs_c, 0, (void *)s, 0, 0, c2, rs_c2, cs_c2, lo, hi);
}
template int gfx950_gemm_split_c<float>(int, int, int, float, const float *, int64_t,
                                        int64_t, const float *, int64_t, int64_t, float,
                                        float *, int64_t, int64_t, float *, int64_t, int64_t,
                                        int, int, hipStream_t);
template int gfx950_gemm_split_c<double>(int, int, int, double, const double *, int64_t,
                                         int64_t, const double *, int64_t, int64_t, double,
                                         double *, int64_t, int64_t, double *, int64_t,
                                         int64_t, int, int, hipStream_t);

// Batched form for the TRSM inversion: `count` problems of one shape whose
// operands are offset per problem as GemmArgs::bat / bs* describe (strides
// {inner, outer} per operand, `bat` problems per outer step).  A and B need
// one unit stride each; no workspace, no densify.
template <typename T>
int gfx950_gemm_batched(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                        const int64_t bsa[2], const T *b, int64_t rs_b, int64_t cs_b,
                        const int64_t bsb[2], T beta, T *c, int64_t rs_c, int64_t cs_c,
                        const int64_t bsc[2], int count, int bat, int tri_a, int tri_b,
                        hipStream_t s) {
  if (m <= 0 || n <= 0 || count <= 0) return MUBLIS_OK;
  if (k <= 0 || bat <= 0 || count > 65535) return MUBLIS_EINVAL;
  if ((rs_a != 1 && cs_a != 1) || (rs_b != 1 && cs_b != 1)) return MUBLIS_EINVAL;
  const bool akc = (cs_a == 1), bnc = (rs_b != 1);
  GemmArgs<T> p;
  memset(&p, 0, sizeof p);
  p.m = m; p.n = n; p.k = k; p.alpha = alpha; p.beta = beta;
  p.a = a; p.rs_a = rs_a; p.cs_a = cs_a;
  p.b = b; p.rs_b = rs_b; p.cs_b = cs_b;
  p.c = c; p.rs_c = rs_c; p.cs_c = cs_c;
  p.tri_a = tri_a; p.tri_b = tri_b;
  p.lpt = tri_a ? (tri_a == 1 ? 1 : 2) : tri_b ? (tri_b == 2 ? 3 : 4) : 0;
  p.lpt_g = 1;
  // every problem's operands must keep the 16-byte alignment of the first
  auto al = [](const int64_t *bs) {
    return (bs[0] * (int64_t)sizeof(T)) % 16 == 0 && (bs[1] * (int64_t)sizeof(T)) % 16 == 0;
  };
  p.vec_a = aligned16(a, akc ? rs_a : cs_a) && al(bsa);
  p.vec_b = aligned16(b, bnc ? rs_b : cs_b) && al(bsb);
  p.bat = bat;
  for (int q = 0; q < 2; ++q) {
    p.bsa[q] = bsa[q];
    p.bsb[q] = bsb[q];
    p.bsc[q] = bsc[q];
  }
  // 32-tiles for batches of at most 384 64-tiles, as in gemm_dev (the TRSM
  // inversion's doubling levels up to h = 256)
  const int64_t t128 = (int64_t)((m + 127) / 128) * ((n + 127) / 128) * count;
  const int64_t t64 = (int64_t)((m + 63) / 64) * ((n + 63) / 64) * count;
  const hipError_t e = t128 >= 2048 ? launch<T, 128>(p, akc, bnc, s, count)
                       : t64 <= 384 ? launch<T, 32>(p, akc, bnc, s, count)
                                    : launch<T, 64>(p, akc, bnc, s, count);
  return e == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}
template int gfx950_gemm_batched<float>(int, int, int, float, const float *, int64_t, int64_t,
                                        const int64_t *, const float *, int64_t, int64_t,
                                        const int64_t *, float, float *, int64_t, int64_t,
                                        const int64_t *, int, int, int, int, hipStream_t);
template int gfx950_gemm_batched<double>(int, int, int, double, const double *, int64_t,
                                         int64_t, const int64_t *, const double *, int64_t,
                                         int64_t, const int64_t *, double, double *, int64_t,
                                         int64_t, const int64_t *, int, int, int, int,
                                         hipStream_t);

template <typename T>
int gfx950_gemm_tri_a(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                      const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                      int64_t cs_c, int tri_a, hipStream_t s) {
  return gemm_dev<T>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0, nullptr, 0,
                     0, beta, c, rs_c, cs_c, 0, (void *)s, tri_a, 0);
}
template int gfx950_gemm_tri_a<float>(int, int, int, float, const float *, int64_t, int64_t,
                                      const float *, int64_t, int64_t, float, float *,
                                      int64_t, int64_t, int, hipStream_t);
template int gfx950_gemm_tri_a<double>(int, int, int, double, const double *, int64_t,
                                       int64_t, const double *, int64_t, int64_t, double,
                                       double *, int64_t, int64_t, int, hipStream_t);

extern "C" int mublis_gfx950_dgemm_tri_dev(int m, int n, int k, double alpha,
                                           const double *a, int64_t rs_a, int64_t cs_a,
                                           const double *b, int64_t rs_b, int64_t cs_b,
                                           double beta, double *c, int64_t rs_c,
                                           int64_t cs_c, int tri_a, int tri_b,
                                           void *stream) {
  return gemm_dev<double>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                          nullptr, 0, 0, beta, c, rs_c, cs_c, 0, stream, tri_a, tri_b);
}
extern "C" int mublis_gfx950_sgemm_tri_dev(int m, int n, int k, float alpha,
                                           const float *a, int64_t rs_a, int64_t cs_a,
                                           const float *b, int64_t rs_b, int64_t cs_b,
                                           float beta, float *c, int64_t rs_c,
                                           int64_t cs_c, int tri_a, int tri_b,
                                           void *stream) {
  return gemm_dev<float>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                         nullptr, 0, 0, beta, c, rs_c, cs_c, 0, stream, tri_a, tri_b);
}
