// gfx950 GEMM kernel family: the device form of MuBLIS's generic L3 driver.
//
// The reference's 5-loop BLIS nest (frame/l3/l3_driver.c:126-239) becomes:
//   jc / ic blocks     -> a grid of 128 x 128 output tiles, one workgroup
//                         each, mapped XCD by XCD so that the 8 L2s each see
//                         a compact group of tiles;
//   pc + packm         -> the K loop: 128 x BK slices of A and B are loaded
//                         from HBM with 16-byte vector loads and staged in
//                         LDS ([k][m] and [k][n], double-buffered);
//   jr / ir            -> 4 wavefronts in a 2 x 2 layout, each owning 64 x 64;
//   micro-kernel p loop-> chains of v_mfma_f64_16x16x4_f64 (FP64) or
//                         v_mfma_f32_16x16x4_f32 (FP32), 16 tiles per wave;
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

#include "mublis_gfx950.h"
#include "gfx950_common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc_t;
  typedef d2 vec_t;  // 16-byte global vector
  static constexpr int BK = 16;
  __device__ static inline acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Output row of register r in lane group g (= lane / 16): the FP64 MFMA
  // interleaves rows across lane groups (measured on gfx950).
  __device__ static inline int row(int g, int r) { return g + 4 * r; }
};
template <> struct Mfma<float> {
  typedef f4 acc_t;
  typedef f4 vec_t;
  static constexpr int BK = 32;
  __device__ static inline acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // The FP32 MFMA keeps 4 consecutive rows per lane group.
  __device__ static inline int row(int g, int r) { return 4 * g + r; }
};

constexpr int NT = 256;       // threads per workgroup (4 waves, 2 x 2)
constexpr int LDS_PAD = 16;   // elements; shifts successive k rows by half the banks

// Workgroup tile TILE x TILE: 128 for large problems, 64 when the 128-tile
// grid would leave most CUs idle (the skinny GEMMs of the TRSM recursion,
// small batched problems).  Each wave owns (TILE/2)^2 = TI x TI MFMA tiles.
template <int TILE>
struct Tile {
  static constexpr int BM = TILE, BN = TILE, S = TILE + LDS_PAD;
  static constexpr int TI = TILE / 32;  // 16 x 16 MFMA tiles per wave and dimension
  static constexpr int MIN_WG_PER_CU = TILE == 128 ? 2 : 4;
};

template <typename T>
struct GemmArgs {
  int m, n, k;
  T alpha, beta;
  const T *a; int64_t rs_a, cs_a;
  const T *b; int64_t rs_b, cs_b;
  const T *a2; int64_t rs_a2, cs_a2;
  const T *b2; int64_t rs_b2, cs_b2;
  T *c; int64_t rs_c, cs_c;
  int ji;          // 0 all, 1 lower (j <= i), 2 upper (j >= i)
  int tiles_m, tiles_n;
  int tri;         // grid enumerates only tiles touching the ji triangle
  int ntiles;
  int vec_a, vec_b, vec_a2, vec_b2;  // 16-byte loads are aligned
  // Triangular operand structure, used only to skip K slices that are zero:
  // tri_a 1: A(i,p) = 0 for p > i, 2: for p < i; tri_b 1: B(p,j) = 0 for
  // p < j, 2: for p > j.  Values are still read as stored (the caller has
  // materialised the zeros), so rounding K bounds outward is harmless.
  int tri_a, tri_b;
};

// Map a hardware workgroup id to an output tile.  Consecutive hardware ids
// land on different XCDs (round robin over 8), so ids are first regrouped so
// that each XCD owns a contiguous range, and that range is walked in groups
// of 8 tile rows, so the workgroups resident on one XCD share A and B panels
// in its L2.
__device__ inline void map_tile(int w, int ntiles, int tiles_m, int tiles_n,
                                int tri, int ji, int &tm, int &tn) {
  const int q = ntiles / 8, r = ntiles % 8;
  const int xcd = w % 8, local = w / 8;
  const int id = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
  if (tri) {
    // Row-major walk over the lower triangle of tiles (tn <= tm); for an
    // upper ji the same walk is transposed.
    int t = (int)((sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((t + 1) * (t + 2) / 2 <= id) ++t;
    while (t * (t + 1) / 2 > id) --t;
    int a = t, b = id - t * (t + 1) / 2;
    if (ji == 2) { tm = b; tn = a; } else { tm = a; tn = b; }
    return;
  }
  const int GROUP = 8;
  const int per_group = GROUP * tiles_n;
  const int g = id / per_group;
  const int first_m = g * GROUP;
  const int gsize = min(tiles_m - first_m, GROUP);
  const int in = id - g * per_group;
  tm = first_m + in % gsize;
  tn = in / gsize;
}

// One operand slice (A: BM x BK, B: BK x BN) in registers.  `c` runs along the
// operand's unit-stride dimension, `o` along the other; V elements per vector.
template <typename T, int LC, int LO, int S>
struct Slice {
  static constexpr int V = 16 / sizeof(T);
  static constexpr int NV = LC * LO / (V * NT);
  T v[NV][V];

  __device__ inline void load(const T *__restrict__ x, int64_t ld, int c0, int o0,
                              int c_lim, int o_lim, bool vec_ok, int tid) {
    typedef typename Mfma<T>::vec_t vec_t;
    const bool full = vec_ok && (c0 + LC <= c_lim) && (o0 + LO <= o_lim);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = i * NT + tid;
      const int c = (lin % (LC / V)) * V;
      const int o = lin / (LC / V);
      const T *p = x + (int64_t)(o0 + o) * ld + (c0 + c);
      if (full) {
        vec_t t = *reinterpret_cast<const vec_t *>(p);
#pragma unroll
        for (int e = 0; e < V; ++e) v[i][e] = t[e];
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e) {
          T t = T(0);
          if ((c0 + c + e < c_lim) && (o0 + o < o_lim)) t = p[e];
          v[i][e] = t;
        }
      }
    }
  }

  // LDS tile is [k][mn] with row stride S.  When the unit-stride dimension
  // is k (CK), element (c, o) goes to [c][o]; otherwise to [o][c].
  template <bool CK>
  __device__ inline void store(T *lds, int tid) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = i * NT + tid;
      const int c = (lin % (LC / V)) * V;
      const int o = lin / (LC / V);
#pragma unroll
      for (int e = 0; e < V; ++e) {
        if (CK) lds[(c + e) * S + o] = v[i][e];
        else lds[o * S + c + e] = v[i][e];
      }
    }
  }
};

template <typename T, int TILE, bool A_KC, bool B_NC>
struct Operands {
  static constexpr int BK = Mfma<T>::BK;
  static constexpr int BM = Tile<TILE>::BM, BN = Tile<TILE>::BN, S = Tile<TILE>::S;
  // A is m x k: K-contiguous -> (c = k, o = m); else (c = m, o = k).
  typedef Slice<T, A_KC ? BK : BM, A_KC ? BM : BK, S> SA;
  // B is k x n: N-contiguous -> (c = n, o = k); else (c = k, o = n).
  typedef Slice<T, B_NC ? BN : BK, B_NC ? BK : BN, S> SB;
};

template <typename T, int TILE, bool A_KC, bool B_NC>
__device__ inline void load_slices(typename Operands<T, TILE, A_KC, B_NC>::SA &ra,
                                   typename Operands<T, TILE, A_KC, B_NC>::SB &rb,
                                   const T *a, int64_t lda, const T *b,
                                   int64_t ldb, int m0, int n0, int k0, int m,
                                   int n, int k, bool va, bool vb, int tid) {
  if (A_KC) ra.load(a, lda, k0, m0, k, m, va, tid);
  else ra.load(a, lda, m0, k0, m, k, va, tid);
  if (B_NC) rb.load(b, ldb, n0, k0, n, k, vb, tid);
  else rb.load(b, ldb, k0, n0, k, n, vb, tid);
}

template <typename T, int TILE, bool A_KC, bool B_NC>
__global__ __launch_bounds__(NT, Tile<TILE>::MIN_WG_PER_CU) void gemm_kernel(GemmArgs<T> p) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int BK = MF::BK;
  constexpr int BM = Tile<TILE>::BM, BN = Tile<TILE>::BN, S = Tile<TILE>::S;
  constexpr int TI = Tile<TILE>::TI, WT = TILE / 2;
  typedef Operands<T, TILE, A_KC, B_NC> Ops;
  __shared__ T As[2][BK * S];
  __shared__ T Bs[2][BK * S];

  int tm, tn;
  map_tile(blockIdx.x, p.ntiles, p.tiles_m, p.tiles_n, p.tri, p.ji, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WT, wn = (wave >> 1) * WT;
  const int lr = lane & 15, lk = lane >> 4;

  acc_t acc[TI][TI];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  typename Ops::SA ra;
  typename Ops::SB rb;

#pragma unroll 1
  for (int seg = 0; seg < 2; ++seg) {
    const T *a = seg ? p.a2 : p.a;
    const T *b = seg ? p.b2 : p.b;
    if (!a) break;
    const int64_t lda = seg ? (A_KC ? p.rs_a2 : p.cs_a2) : (A_KC ? p.rs_a : p.cs_a);
    const int64_t ldb = seg ? (B_NC ? p.rs_b2 : p.cs_b2) : (B_NC ? p.rs_b : p.cs_b);
    const bool va = seg ? p.vec_a2 : p.vec_a;
    const bool vb = seg ? p.vec_b2 : p.vec_b;
    int kbeg = 0, kend = p.k;
    if (p.tri_a == 1) kend = min(kend, m0 + BM);
    if (p.tri_a == 2) kbeg = max(kbeg, m0);
    if (p.tri_b == 1) kbeg = max(kbeg, n0);
    if (p.tri_b == 2) kend = min(kend, n0 + BN);
    const int kt0 = kbeg / BK;
    const int ktiles = kend > kbeg ? (kend + BK - 1) / BK - kt0 : 0;
    if (ktiles <= 0) continue;

    load_slices<T, TILE, A_KC, B_NC>(ra, rb, a, lda, b, ldb, m0, n0, kt0 * BK, p.m, p.n,
                                     p.k, va, vb, tid);
    __syncthreads();  // previous segment's readers are done with buffer 0
    ra.template store<A_KC>(As[0], tid);
    rb.template store<!B_NC>(Bs[0], tid);
    __syncthreads();

#pragma unroll 1
    for (int kt = 0; kt < ktiles; ++kt) {
      const int cur = kt & 1;
      const bool more = kt + 1 < ktiles;
      if (more)
        load_slices<T, TILE, A_KC, B_NC>(ra, rb, a, lda, b, ldb, m0, n0,
                                         (kt0 + kt + 1) * BK, p.m, p.n, p.k, va, vb, tid);
      const T *as = As[cur];
      const T *bs = Bs[cur];
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        T af[TI], bf[TI];
        const int row = (kk * 4 + lk) * S;
#pragma unroll
        for (int i = 0; i < TI; ++i) af[i] = as[row + wm + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TI; ++j) bf[j] = bs[row + wn + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TI; ++j) acc[i][j] = MF::step(af[i], bf[j], acc[i][j]);
      }
      if (more) {
        ra.template store<A_KC>(As[cur ^ 1], tid);
        rb.template store<!B_NC>(Bs[cur ^ 1], tid);
      }
      __syncthreads();
    }
  }

  // Epilogue.  MFMA 16x16 output layout: lane l, register r holds column
  // l%16 and row MF::row(l/16, r) of its 16 x 16 tile.
  const T alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < TI; ++i) {
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const int col = n0 + wn + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + MF::row(lk, r);
        if (row >= p.m || col >= p.n) continue;
        if (p.ji == 1 && col > row) continue;
        if (p.ji == 2 && col < row) continue;
        T *cp = p.c + (int64_t)row * p.rs_c + (int64_t)col * p.cs_c;
        const T prod = alpha * acc[i][j][r];
        if (beta == T(0)) *cp = prod;
        else if (beta == T(1)) *cp = *cp + prod;
        else *cp = prod + beta * *cp;
      }
    }
  }
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

template <typename T, int TILE>
hipError_t launch(GemmArgs<T> p, bool akc, bool bnc, hipStream_t s) {
  p.tiles_m = (p.m + TILE - 1) / TILE;
  p.tiles_n = (p.n + TILE - 1) / TILE;
  p.ntiles = p.tri ? p.tiles_m * (p.tiles_m + 1) / 2 : p.tiles_m * p.tiles_n;
  const dim3 grid(p.ntiles), block(NT);
  if (akc && bnc) hipLaunchKernelGGL((gemm_kernel<T, TILE, true, true>), grid, block, 0, s, p);
  else if (akc) hipLaunchKernelGGL((gemm_kernel<T, TILE, true, false>), grid, block, 0, s, p);
  else if (bnc) hipLaunchKernelGGL((gemm_kernel<T, TILE, false, true>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((gemm_kernel<T, TILE, false, false>), grid, block, 0, s, p);
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
             int tri_b = 0) {
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
    // When a dimension has extent 1 both readings address the same data.
    const bool akc = (cs_a == 1);
    const bool bnc = (rs_b != 1);
    bool same2 = true;
    if (a2) {
      const bool akc2 = akc ? (cs_a2 == 1) : (rs_a2 == 1);
      const bool bnc2 = bnc ? (cs_b2 == 1) : (rs_b2 == 1);
      same2 = akc2 && bnc2;
    }
    GemmArgs<T> p;
    p.m = m; p.n = n; p.k = k; p.alpha = alpha; p.beta = beta;
    p.a = a; p.rs_a = rs_a; p.cs_a = cs_a;
    p.b = b; p.rs_b = rs_b; p.cs_b = cs_b;
    p.a2 = same2 ? a2 : nullptr; p.rs_a2 = rs_a2; p.cs_a2 = cs_a2;
    p.b2 = same2 ? b2 : nullptr; p.rs_b2 = rs_b2; p.cs_b2 = cs_b2;
    p.c = c; p.rs_c = rs_c; p.cs_c = cs_c; p.ji = ji;
    p.tri = (ji != 0 && m == n) ? 1 : 0;
    p.tri_a = a2 ? 0 : tri_a;
    p.tri_b = a2 ? 0 : tri_b;
    p.vec_a = aligned16(a, akc ? rs_a : cs_a);
    p.vec_b = aligned16(b, bnc ? rs_b : cs_b);
    p.vec_a2 = a2 ? aligned16(a2, akc ? rs_a2 : cs_a2) : 0;
    p.vec_b2 = b2 ? aligned16(b2, bnc ? rs_b2 : cs_b2) : 0;
    // 128 x 128 tiles unless that grid would give fewer than two
    // workgroups per CU (256 CUs).
    const int64_t t128 = (int64_t)((m + 127) / 128) * ((n + 127) / 128);
    const int64_t n128 = p.tri ? ((m + 127) / 128) * ((m + 127) / 128 + 1) / 2 : t128;
    hipError_t e = n128 >= 512 ? launch<T, 128>(p, akc, bnc, s)
                               : launch<T, 64>(p, akc, bnc, s);
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
