// Device building blocks of the gfx950 GEMM family, shared by the one-problem
// kernel (gemm.hip) and the grouped persistent kernel (batched.hip): MFMA
// traits, tile traits, operand staging and the per-tile compute routine.
#ifndef MUBLIS_GFX950_GEMM_TILE_H
#define MUBLIS_GFX950_GEMM_TILE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gfx950 {


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
  // LDS column swizzle of k row k for K-contiguous operands (see Slice).
  __device__ static inline int swz(int k) { return k & 15; }
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
  __device__ static inline int swz(int k) { return k & 28; }
};

constexpr int NT = 256;       // threads per workgroup (4 waves, 2 x 2) unless stated
constexpr int LDS_PAD = 16;   // elements; shifts successive k rows by half the banks

// Workgroup tile TILE x TILE: 128 for large problems, 64 when the 128-tile
// grid would leave most CUs idle (the skinny GEMMs of the TRSM recursion,
// small batched problems).  A workgroup of NTH threads has 2 waves along M
// and NTH / 128 along N; each wave owns (TILE / 2) x (TILE / (NTH / 128)).
template <int TILE>
struct Tile {
  static constexpr int BM = TILE, BN = TILE, S = TILE + LDS_PAD;
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
  // Rows c2_lo <= i < c2_hi of the result are written to c2 (row i - c2_lo)
  // instead of c; beta still reads them from c.  Both bounds are multiples
  // of the tile height, so a tile writes to one of the two.  Used by the
  // TRSM recursion: the update that last writes a diagonal block's rows of
  // B writes them straight into the buffer its explicit-inverse leaf reads.
  T *c2; int64_t rs_c2, cs_c2;
  int c2_lo, c2_hi;
  int ji;          // 0 all, 1 lower (j <= i), 2 upper (j >= i)
  int tiles_m, tiles_n;
  // Grid over the ji triangle (SYRK/SYR2K): 0 every tile; 1 the tiles
  // touching the triangle; 2 only the tiles strictly inside it; 3 the tiles
  // of the diagonal blocks of four times the tile size on the triangle's
  // side (ten per block), which complete mode 2 of 128-tiles with 32-tiles.
  int tri;
  int ntiles;
  int vec_a, vec_b, vec_a2, vec_b2;  // 16-byte loads are aligned
  // Triangular operand structure, used only to skip K slices that are zero:
  // tri_a 1: A(i,p) = 0 for p > i, 2: for p < i; tri_b 1: B(p,j) = 0 for
  // p < j, 2: for p > j.  Values are still read as stored (the caller has
  // materialised the zeros), so rounding K bounds outward is harmless.
  int tri_a, tri_b;
  // Long-first walk for a triangular operand (map_tile_lpt): 0 off; 1 / 2
  // tile rows are the lines, longest at the high / low end; 3 / 4 the same
  // for tile columns.  lpt_g: tiles per dealt unit (a whole line, or 1).
  int lpt, lpt_g;
  // workgroups per XCD in one round, for the long-first walk's first-round
  // snake (map_tile_lpt's r1); 0 for batched launches
  int lpt_r1;
  // Batched launch (gridDim.y problems, the one-problem kernel only):
  // problem z starts at a + (z % bat) * bsa[0] + (z / bat) * bsa[1], and
  // likewise for b and c.
  int bat;
  int64_t bsa[2], bsb[2], bsc[2];
};

// Staging layout of an operand read through strides (r1, c1): 1 when it is
// read contiguous along its second index (c1 == 1), 0 along its first (r1 ==
// 1), -1 when neither stride is 1.  `prefer_second` breaks the tie when both
// are 1 (A prefers K = its second index, B prefers K = its first).  With a
// second operand of the same rows x cols shape read in the same tile loop
// through (*r2, *c2) (SYR2K's second product), the first operand's layout is
// kept when the second can be read that way too, else the other layout is
// taken when it suits both; a dimension of extent 1 allows either reading
// and its stride is then set to 1 so that the reading holds.  -1 when no
// layout suits both (nothing is changed then).
__host__ inline int common_layout(int rows, int cols, int64_t &r1, int64_t &c1, int64_t *r2,
                                  int64_t *c2, bool prefer_second) {
  const bool f1 = r1 == 1, s1 = c1 == 1;
  if (!f1 && !s1) return -1;
  const int own = (s1 && (prefer_second || !f1)) ? 1 : 0;
  if (!r2) return own;
  auto ok = [&](int lay, int64_t r, int64_t c) {
    return lay ? (c == 1 || cols <= 1) : (r == 1 || rows <= 1);
  };
  for (int lay : {own, 1 - own}) {
    if (!ok(lay, r1, c1) || !ok(lay, *r2, *c2)) continue;
    if (lay) {
      c1 = 1;
      *c2 = 1;
    } else {
      r1 = 1;
      *r2 = 1;
    }
    return lay;
  }
  return -1;
}

// Long-first walk: with a triangular operand, a tile's K extent grows
// linearly along one axis (a "line" of tiles shares one K), so one launch
// mixes tiles of K = BK .. k.  Units of g tiles -- whole lines when there
// are enough of them to balance (the tiles of a line then share A or B in
// one L2), else single tiles -- are dealt to the 8 XCDs in snake order of
// decreasing K (XCD x gets units x, 15 - x, 16 + x, ...), so every XCD
// receives the same total K, and each XCD starts its longest units first,
// which lets the short tiles fill the slots the long ones leave (longest
// processing time first).  The grid is 8 * ceil(units / 8) * g workgroups;
// returns false for the padding workgroups.
//
// r1 (0: off) is the number of workgroups an XCD holds at once, 32 CUs
// times the workgroups per CU.  Those first r1 of each XCD start together,
// dealt over its CUs breadth-first (local index l to CU l % 32), so in
// plain long-first order CU 0 would take the longest tile of every row of
// 32 and CU 31 the shortest.  Every other row of 32 is therefore reversed
// (a snake), pairing long tiles with short on every CU.  Measured on
// MI355X, the 1024 x 4096 x 1024 triangular product of the TRSM leaves (one
// round of 64-tiles): 97 -> 89.5 us; the depth-first pairing (a CU's
// slots taking consecutive indices) ran 133 us, which is how the dispatch
// order was told apart.
//
// rev reverses an XCD's order instead: a batched launch of small problems
// (the TRSM inversion's doubling GEMMs) lays several problems' workgroups
// side by side in one row of 32, so each CU takes the same-rank tile of two
// problems; reversing the problems of every other row pairs long with short
// there (the doubling's 8 x (512 x 512 x 512) level: 45.5 -> 42 us).
__host__ __device__ inline bool map_tile_lpt(int w, int tiles_m, int tiles_n, int lpt, int g,
                                             int &tm, int &tn, int r1 = 0, bool rev = false) {
  const bool rows = lpt <= 2;
  const int lines = rows ? tiles_m : tiles_n, per = rows ? tiles_n : tiles_m;
  const int upl = (per + g - 1) / g;  // units per line
  const int xcd = w % 8, local0 = w / 8;
  const int nx = (lines * upl + 7) / 8 * g;  // this XCD's workgroups
  int local = local0;
  if (rev) {
    local = nx - 1 - local0;
  } else if (local0 < r1) {
    const int nr = r1 < nx ? r1 : nx;  // first round
    const int q = local0 / 32, j = local0 - 32 * q, len = nr - 32 * q < 32 ? nr - 32 * q : 32;
    if (q & 1) local = 32 * q + (len - 1 - j);
  }
  const int k = local / g, oi = local - k * g;
  const int d = 8 * k + ((k & 1) ? 7 - xcd : xcd);  // unit rank in decreasing K
  const int dl = d / upl, o = (d - dl * upl) * g + oi;
  if (dl >= lines || o >= per) return false;
  const int line = (lpt == 1 || lpt == 3) ? lines - 1 - dl : dl;
  tm = rows ? line : o;
  tn = rows ? o : line;
  return true;
}

// Unit size for map_tile_lpt: whole lines once every XCD gets four or more.
__host__ __device__ inline int lpt_unit(int lines, int per) { return lines >= 32 ? per : 1; }

// Map a hardware workgroup id to an output tile.  Consecutive hardware ids
// land on different XCDs (round robin over 8), so ids are first regrouped so
// that each XCD owns a contiguous range, and that range is walked in groups
// of 8 tile rows, so the workgroups resident on one XCD share A and B panels
// in its L2.
__host__ __device__ inline void map_tile(int w, int ntiles, int tiles_m, int tiles_n,
                                         int tri, int ji, int &tm, int &tn) {
  const int q = ntiles / 8, r = ntiles % 8;
  const int xcd = w % 8, local = w / 8;
  const int id = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
  if (tri == 3) {
    // diagonal block d of four times the tile: its 10 tiles on the ji side,
    // row by row
    const int d = id / 10, r = id - 10 * d;
    const int row = (r >= 1) + (r >= 3) + (r >= 6), col = r - row * (row + 1) / 2;
    const int a = 4 * d + row, b = 4 * d + col;
    if (ji == 2) { tm = b; tn = a; } else { tm = a; tn = b; }
    return;
  }
  if (tri) {
    // Row-major walk over the lower triangle of tiles (tn <= tm, or tn < tm
    // for mode 2); for an upper ji the same walk is transposed.
    int t = (int)((sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((t + 1) * (t + 2) / 2 <= id) ++t;
    while (t * (t + 1) / 2 > id) --t;
    int a = t + (tri == 2), b = id - t * (t + 1) / 2;
    if (ji == 2) { tm = b; tn = a; } else { tm = a; tn = b; }
    return;
  }
  const int GROUP = 8;
  const int per_group = GROUP * tiles_n;
  const int g = id / per_group;
  const int first_m = g * GROUP;
  const int gsize = tiles_m - first_m < GROUP ? tiles_m - first_m : GROUP;
  const int in = id - g * per_group;
  tm = first_m + in % gsize;
  tn = in / gsize;
}

// Operands and C are in global memory.  A pointer that reaches a kernel
// inside a descriptor loaded from memory (the grouped kernel's problem
// list) is a generic pointer to the compiler, which then emits flat loads
// and stores: those count against the LDS counter as well, so every wait
// for an LDS access also waits for every load in flight.  Accesses go
// through these casts instead, which make them global loads and stores.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *gmem(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// One operand slice (A: BM x BK, B: BK x BN) in registers.  `c` runs along the
// operand's unit-stride dimension, `o` along the other; V elements per vector.
template <typename T, int LC, int LO, int S, int NTH>
struct Slice {
  static constexpr int V = 16 / sizeof(T);
  static constexpr int NV = LC * LO / (V * NTH);
  static_assert(NV >= 1 && NV * V * NTH == LC * LO, "every thread loads whole vectors");
  static_assert(LC % V == 0, "a vector stays within one row");
  T v[NV][V];

  // EV (edge vectors): in a slice that is not whole, a vector whose row is
  // in range and whose elements all are is still loaded whole, a vector
  // whose row is out of range is zero, and only the vector that straddles
  // the unit-stride edge goes element by element.  Without it the whole
  // slice goes element by element.
  template <bool EV = false>
  __device__ inline void load(const T *__restrict__ x, int64_t ld, int c0, int o0,
                              int c_lim, int o_lim, bool vec_ok, int tid) {
    typedef typename Mfma<T>::vec_t vec_t;
    const bool full = vec_ok && (c0 + LC <= c_lim) && (o0 + LO <= o_lim);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = i * NTH + tid;
      const int c = (lin % (LC / V)) * V;
      const int o = lin / (LC / V);
      const T *p = x + (int64_t)(o0 + o) * ld + (c0 + c);
      if (full) {
        // 16 bytes from an address aligned to the element only: gfx950's
        // global loads need no more than dword alignment
        vec_t t;
        __builtin_memcpy(&t, gmem(p), sizeof t);
#pragma unroll
        for (int e = 0; e < V; ++e) v[i][e] = t[e];
      } else if constexpr (EV) {
        if (o0 + o >= o_lim) {
#pragma unroll
          for (int e = 0; e < V; ++e) v[i][e] = T(0);
        } else if (vec_ok && c0 + c + V <= c_lim) {
          vec_t t;
          __builtin_memcpy(&t, gmem(p), sizeof t);
#pragma unroll
          for (int e = 0; e < V; ++e) v[i][e] = t[e];
        } else {
#pragma unroll
          for (int e = 0; e < V; ++e) v[i][e] = c0 + c + e < c_lim ? gmem(p)[e] : T(0);
        }
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e) {
          T t = T(0);
          if ((c0 + c + e < c_lim) && (o0 + o < o_lim)) t = gmem(p)[e];
          v[i][e] = t;
        }
      }
    }
  }

  // LDS tile is [k][mn] with row stride S.  When the unit-stride dimension
  // is k (CK), element (c, o) goes to [c][o ^ swz(c)]; otherwise to [o][c].
  // Without the swizzle the lanes of one CK store instruction write the
  // same column o of 8 different k rows, and with S = TILE + 16 those rows
  // start in the same bank: an 8-way conflict on every store (measured: 19
  // conflict cycles per LDS instruction in FP64).  XOR-ing o by the k row
  // spreads them over the banks; readers apply the same XOR, which only
  // permutes words inside each 16 (FP64) or 32 (FP32) column group, so
  // their 16 consecutive columns stay conflict-free.
  template <bool CK>
  __device__ inline void store(T *lds, int tid) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = i * NTH + tid;
      const int c = (lin % (LC / V)) * V;
      const int o = lin / (LC / V);
#pragma unroll
      for (int e = 0; e < V; ++e) {
        if (CK) lds[(c + e) * S + (o ^ Mfma<T>::swz(c + e))] = v[i][e];
        else lds[o * S + c + e] = v[i][e];
      }
    }
  }
};

template <typename T, int TILE, bool A_KC, bool B_NC, int BK, int NTH>
struct Operands {
  static constexpr int BM = Tile<TILE>::BM, BN = Tile<TILE>::BN, S = Tile<TILE>::S;
  // A is m x k: K-contiguous -> (c = k, o = m); else (c = m, o = k).
  typedef Slice<T, A_KC ? BK : BM, A_KC ? BM : BK, S, NTH> SA;
  // B is k x n: N-contiguous -> (c = n, o = k); else (c = k, o = n).
  typedef Slice<T, B_NC ? BN : BK, B_NC ? BK : BN, S, NTH> SB;
};

template <typename T, int TILE, bool A_KC, bool B_NC, int BK, int NTH, bool EV = false>
__device__ inline void load_slices(typename Operands<T, TILE, A_KC, B_NC, BK, NTH>::SA &ra,
                                   typename Operands<T, TILE, A_KC, B_NC, BK, NTH>::SB &rb,
                                   const T *a, int64_t lda, const T *b,
                                   int64_t ldb, int m0, int n0, int k0, int m,
                                   int n, int k, bool va, bool vb, int tid) {
  if (A_KC) ra.template load<EV>(a, lda, k0, m0, k, m, va, tid);
  else ra.template load<EV>(a, lda, m0, k0, m, k, va, tid);
  if (B_NC) rb.template load<EV>(b, ldb, n0, k0, n, k, vb, tid);
  else rb.template load<EV>(b, ldb, k0, n0, k, n, vb, tid);
}

// LDS elements one workgroup needs for a TILE x TILE tile of type T staged
// in K slices of BK.
template <typename T, int TILE, int BK = Mfma<T>::BK>
struct LdsSize {
  static constexpr int value = 4 * BK * Tile<TILE>::S;
};

// The MFMA steps of one staged slice (the K loop of gemm_tile below, as a
// function for the loop that loads two slices ahead).  Only the first nk of
// its BK / 4 steps run; the caller skips only steps whose k rows all lie
// past the end of K (exact zeros in both operands) and accounts for them.
template <typename T, int TILE, bool A_KC, bool B_NC, int BK, int NTH, int TIM, int TIN>
__device__ inline void mfma_slice(typename Mfma<T>::acc_t (&acc)[TIM][TIN], const T *as,
                                  const T *bs, int wm, int wn, int lr, int lk,
                                  int nk = BK / 4) {
  typedef Mfma<T> MF;
  constexpr int S = Tile<TILE>::S;
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int kk = 0; kk < BK / 4; ++kk) {
    if (kk >= nk) break;
    T af[TIM], bf[TIN];
    const int kr = kk * 4 + lk, row = kr * S;
    const int swa = A_KC ? MF::swz(kr) : 0, swb = B_NC ? 0 : MF::swz(kr);
    const int hia = A_KC ? MF::swz(4 * kk) & ~15 : 0;
    const int hib = B_NC ? 0 : MF::swz(4 * kk) & ~15;
    const int rowa = row + (lr ^ (swa & 15)), rowb = row + (lr ^ (swb & 15));
    constexpr bool SPLIT_B = A_KC || B_NC;
#pragma unroll
    for (int i = 0; i < TIM; ++i) af[i] = as[rowa + ((wm + i * 16) ^ hia)];
#pragma unroll
    for (int j = 0; j < TIN; ++j)
      bf[j] = SPLIT_B ? bs[rowb + ((wn + j * 16) ^ hib)] : bs[row + ((wn + j * 16 + lr) ^ swb)];
#pragma unroll
    for (int i = 0; i < TIM; ++i)
#pragma unroll
      for (int j = 0; j < TIN; ++j) acc[i][j] = MF::step(af[i], bf[j], acc[i][j]);
  }
  __builtin_amdgcn_s_setprio(0);
}

// Adds +0 to every accumulator: what a 4-deep MFMA step of exact zeros does.
template <typename T, int TIM, int TIN>
__device__ inline void zero_pad_fix(typename Mfma<T>::acc_t (&acc)[TIM][TIN]) {
#pragma unroll
  for (int i = 0; i < TIM; ++i)
#pragma unroll
    for (int j = 0; j < TIN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] += T(0);
}

// One TILE x TILE output tile (tm, tn) of the product described by p,
// computed by the calling workgroup of NTH threads with `lds` (LdsSize
// elements) as staging memory.  The caller synchronises before reusing lds
// (SMALL: the tile's last K slice ends with a barrier and nothing after it
// touches lds, so a tile that follows needs no barrier of its own).
//
// BK, the K depth of one staged slice, sets only how K is cut into slices:
// every depth runs the same MFMA steps in the same k order, and the steps of
// a slice past k add exact zeros.  Such a zero step leaves every
// accumulator as it was (denormals included, measured on gfx950) except -0,
// which becomes +0 (an accumulator reaches -0 only when the products of an
// element underflow).  So that this does not depend on the depth either,
// every K segment ends as if zero steps ran up to a multiple of the full
// depth Mfma<T>::BK: a shallower slice depth, or SMALL's last slice, which
// skips its zero steps, adds +0 to the accumulators once instead
// (zero_pad_fix).  Segments cut short by a triangular operand (tri_a,
// tri_b) run the same slices at every depth they are launched with.
//
// SMALL is the grouped kernel's mode for small tiles, whose short K loops
// leave the fixed costs of a tile a large share of its time; it changes only
// how data moves, never the arithmetic of an element:
//  * a column-major C (rs_c == 1) moves in 16-byte runs: the FP32 MFMA
//    output layout gives each lane 4 consecutive rows of one column, and the
//    item's C values are read as its last K slice begins;
//  * a wave whose region of the tile stores nothing skips its MFMAs;
//  * slices of edge tiles are loaded by whole vectors where they can
//    (Slice::load's EV);
//  * the first segment stages its first slice without a barrier of its
//    own: the previous tile's last K slice ended with one.
template <typename T, int TILE, bool A_KC, bool B_NC, int BK = Mfma<T>::BK, int NTH = NT,
          bool SMALL = false>
__device__ inline void gemm_tile(const GemmArgs<T> &p, int tm, int tn, T *lds) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  static_assert(BK % 4 == 0 && BK <= Mfma<T>::BK, "whole MFMA k steps, swizzle range");
  constexpr int BM = Tile<TILE>::BM, BN = Tile<TILE>::BN, S = Tile<TILE>::S;
  constexpr int WNW = NTH / 128, WTM = TILE / 2, WTN = TILE / WNW;
  constexpr int TIM = WTM / 16, TIN = WTN / 16;  // 16 x 16 MFMA tiles per wave
  static_assert(NTH % 128 == 0 && WTN >= 16, "2 x (NTH / 128) waves of whole MFMA tiles");
  typedef Operands<T, TILE, A_KC, B_NC, BK, NTH> Ops;
  T(*As)[BK * S] = reinterpret_cast<T(*)[BK * S]>(lds);
  T(*Bs)[BK * S] = reinterpret_cast<T(*)[BK * S]>(lds + 2 * BK * S);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WTM, wn = (wave >> 1) * WTN;
  const int lr = lane & 15, lk = lane >> 4;
  // SMALL: a wave whose whole region of the tile stores nothing (past the
  // edge of C, or on the unstored side of the ji triangle) skips its
  // fragment reads and MFMAs; it still loads and stages its share of every
  // slice.  The elements that are stored see the same MFMA steps either way.
  // Measured on MI355X, the batched replay of 4096 SGEMM + 1024 SSYRK: 1.093
  // -> 1.072 ms, the same bits; one-problem launches of SYRK and ragged GEMMs
  // showed no gain, so they keep the unconditional loop.
  const int wr0 = m0 + __builtin_amdgcn_readfirstlane(wm);
  const int wc0 = n0 + __builtin_amdgcn_readfirstlane(wn);
  const bool live = !SMALL || (wr0 < p.m && wc0 < p.n && !(p.ji == 1 && wc0 > wr0 + WTM - 1) &&
                               !(p.ji == 2 && wr0 > wc0 + WTN - 1));

  acc_t acc[TIM][TIN];
#pragma unroll
  for (int i = 0; i < TIM; ++i)
#pragma unroll
    for (int j = 0; j < TIN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  typename Ops::SA ra;
  typename Ops::SB rb;

  // Elements of C this tile stores.
  auto c_in = [&](int row, int col) {
    return row < p.m && col < p.n && !(p.ji == 1 && col > row) && !(p.ji == 2 && col < row);
  };
  // SMALL with a column-major C: the FP32 MFMA leaves each lane 4
  // consecutive rows of one column of each of its 16 x 16 tiles, a 16-byte
  // run of C.  Run u = (i, j) = (u / TIN, u % TIN) of a lane covers rows
  // run_row(u) .. + 3 of column run_col(u).  A run that C stores whole moves
  // as one vector; any other (the edge of C, the diagonal of a SYRK tile)
  // element by element, only where c_in holds.
  constexpr int CV = 16 / sizeof(T), CU = TIM * TIN;
  static_assert(!SMALL || (sizeof(T) == 4 && CV == 4), "FP32: a lane's 4 rows are one run");
  auto run_row = [&](int u) { return m0 + wm + 16 * (u / TIN) + 4 * lk; };
  auto run_col = [&](int u) { return n0 + wn + 16 * (u % TIN) + lr; };
  typedef typename MF::vec_t cvec_t;
  auto run_whole = [&](int row, int col) {
    return row + CV - 1 < p.m && col < p.n && !(p.ji == 1 && col > row) &&
           !(p.ji == 2 && col < row + CV - 1);
  };
  auto load_run = [&](int row, int col, T(&v)[CV]) {
    const T *cp = p.c + (int64_t)row + (int64_t)col * p.cs_c;
    if (run_whole(row, col)) {
      cvec_t t;
      __builtin_memcpy(&t, gmem(cp), sizeof t);
#pragma unroll
      for (int e = 0; e < CV; ++e) v[e] = t[e];
    } else {
#pragma unroll
      for (int e = 0; e < CV; ++e) v[e] = c_in(row + e, col) ? gmem(cp)[e] : T(0);
    }
  };
  // When a thread's share is at most 8 values they are read as the last K
  // slice begins -- its operand registers are free by then -- so the read
  // latency runs under that slice's MFMAs instead of after them.
  constexpr bool CPRE = SMALL && CU * CV <= 8;
  T cpre[CPRE ? CU : 1][CV];
  bool have_c = false;
  if constexpr (SMALL) {
    // One K segment (SYR2K runs two, each over its own operands); the first
    // segment stages its first slice without a barrier of its own (every
    // reader of lds passed the barrier that ends the previous tile's last
    // slice, and the epilogue does not use lds).  PRE: the last segment.
    auto segment = [&](auto PRE, int seg) {
      const T *a = seg ? p.a2 : p.a;
      const T *b = seg ? p.b2 : p.b;
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
      if (ktiles <= 0) return;
      load_slices<T, TILE, A_KC, B_NC, BK, NTH, true>(ra, rb, a, lda, b, ldb, m0, n0, kt0 * BK,
                                                      p.m, p.n, p.k, va, vb, tid);
      if (seg > 0) __syncthreads();  // the first segment's readers are done with buffer 0
      ra.template store<A_KC>(As[0], tid);
      rb.template store<!B_NC>(Bs[0], tid);
      __syncthreads();
#pragma unroll 1
      for (int kt = 0; kt + 1 < ktiles; ++kt) {
        const int cur = kt & 1;
        load_slices<T, TILE, A_KC, B_NC, BK, NTH, true>(ra, rb, a, lda, b, ldb, m0, n0,
                                                        (kt0 + kt + 1) * BK, p.m, p.n, p.k, va,
                                                        vb, tid);
        if (live) mfma_slice<T, TILE, A_KC, B_NC, BK, NTH>(acc, As[cur], Bs[cur], wm, wn, lr, lk);
        ra.template store<A_KC>(As[cur ^ 1], tid);
        rb.template store<!B_NC>(Bs[cur ^ 1], tid);
        __syncthreads();
      }
      if constexpr (decltype(PRE)::value && CPRE) {
        if (p.rs_c == 1 && p.beta != T(0)) {
#pragma unroll
          for (int u = 0; u < CU; ++u) load_run(run_row(u), run_col(u), cpre[u]);
          have_c = true;
        }
      }
      // The last slice runs only the 4-deep steps that reach into K (the
      // rest have exact zeros in both operands), then stands in for the
      // zero steps up to a multiple of the full depth (see above).
      const int klast = (kt0 + ktiles - 1) * BK;
      const int nk = min(BK / 4, (p.k - klast + 3) / 4);
      if (live) {
        mfma_slice<T, TILE, A_KC, B_NC, BK, NTH>(acc, As[(ktiles - 1) & 1], Bs[(ktiles - 1) & 1],
                                                 wm, wn, lr, lk, nk);
        if (klast + 4 * nk < (p.k + MF::BK - 1) / MF::BK * MF::BK) zero_pad_fix<T>(acc);
      }
      __syncthreads();
    };
    if (p.a2) segment(std::false_type{}, 0);
    segment(std::true_type{}, p.a2 ? 1 : 0);
  }

#pragma unroll 1
  for (int seg = 0; seg < (SMALL ? 0 : 2); ++seg) {
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

    // FP64 128-tiles with op(A) = A^T and column-major B (both operands
    // K-contiguous) load their slices two ahead, into two register sets: a
    // slice's loads then have two MFMA blocks to land instead of one.
    // Measured on MI355X, DGEMM 8192 op(A) = A^T: 67.2-67.5 -> 68.4-68.6
    // TFLOP/s, same bits (it spills 68 bytes per lane outside the MFMA
    // block); in the other layouts the second set spilled inside the loop
    // (no transpose: 65.5 -> 57.5), and B alone two ahead, or one set
    // requested before the barrier instead of after it, ran 4-5 % slower.
    constexpr bool AHEAD2 = sizeof(T) == 8 && TILE == 128 && NTH == 512 && A_KC && !B_NC && !SMALL;
    if constexpr (AHEAD2) {
      typename Ops::SA ra2[2];
      typename Ops::SB rb2[2];
      load_slices<T, TILE, A_KC, B_NC, BK, NTH, SMALL>(ra2[0], rb2[0], a, lda, b, ldb, m0, n0,
                                                       kt0 * BK, p.m, p.n, p.k, va, vb, tid);
      if (ktiles > 1)
        load_slices<T, TILE, A_KC, B_NC, BK, NTH, SMALL>(ra2[1], rb2[1], a, lda, b, ldb, m0, n0,
                                                         (kt0 + 1) * BK, p.m, p.n, p.k, va, vb, tid);
      __syncthreads();  // previous segment's readers are done with buffer 0
      ra2[0].template store<A_KC>(As[0], tid);
      rb2[0].template store<!B_NC>(Bs[0], tid);
      __syncthreads();
      // iteration kt (set q = kt & 1): slice kt + 2 into set q (its slice kt
      // is staged already), MFMAs on buffer q, slice kt + 1 from set q ^ 1
      // into buffer q ^ 1
      auto step = [&](auto Q, int kt) {
        constexpr int q = decltype(Q)::value;
        if (kt + 2 < ktiles)
          load_slices<T, TILE, A_KC, B_NC, BK, NTH, SMALL>(ra2[q], rb2[q], a, lda, b, ldb, m0, n0,
                                                           (kt0 + kt + 2) * BK, p.m, p.n, p.k, va,
                                                           vb, tid);
        mfma_slice<T, TILE, A_KC, B_NC, BK, NTH>(acc, As[q], Bs[q], wm, wn, lr, lk);
        if (kt + 1 < ktiles) {
          ra2[q ^ 1].template store<A_KC>(As[q ^ 1], tid);
          rb2[q ^ 1].template store<!B_NC>(Bs[q ^ 1], tid);
        }
        __syncthreads();
      };
#pragma unroll 1
      for (int kt = 0; kt < ktiles; kt += 2) {
        step(std::integral_constant<int, 0>{}, kt);
        if (kt + 1 < ktiles) step(std::integral_constant<int, 1>{}, kt + 1);
      }
      continue;
    }
    // a shallow depth stands in for the zero steps up to a multiple of the
    // full depth once the segment's slices are done (see above)
    const bool pad_fix = BK < MF::BK && !p.tri_a && !p.tri_b &&
                         (p.k + BK - 1) / BK * BK < (p.k + MF::BK - 1) / MF::BK * MF::BK;
    load_slices<T, TILE, A_KC, B_NC, BK, NTH>(ra, rb, a, lda, b, ldb, m0, n0, kt0 * BK, p.m, p.n,
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
        load_slices<T, TILE, A_KC, B_NC, BK, NTH>(ra, rb, a, lda, b, ldb, m0, n0,
                                                  (kt0 + kt + 1) * BK, p.m, p.n, p.k, va, vb, tid);
      const T *as = As[cur];
      const T *bs = Bs[cur];
      // A wave raises its priority for its MFMA block, so that of the waves
      // on a SIMD the one with MFMAs ready keeps issuing them while the
      // others run their loads and LDS stores (measured: FP64 NN +2.6 %, FP32
      // +1.3 %; FP64 with op(A) = A^T and column-major B -0.5 % at 2 waves
      // per SIMD, +3.1 % at 4).
      if (live) {
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
          T af[TIM], bf[TIN];
          const int kr = kk * 4 + lk, row = kr * S;
          const int swa = A_KC ? MF::swz(kr) : 0, swb = B_NC ? 0 : MF::swz(kr);
          // (x + lr) ^ s == (x ^ (s & ~15)) + (lr ^ (s & 15)) for x a multiple
          // of 16 and lr < 16; s & ~15 depends on kk alone (lk < 4 sets only
          // bits 0-1 of kr), so each tile's part is an immediate offset.  This
          // form saves address registers (it took the FP64 8-wave kernel with
          // op(A) = A^T out of spilling).  For B when only B is swizzled
          // (column-major A and B) the plain form measured faster.
          const int hia = A_KC ? MF::swz(4 * kk) & ~15 : 0;
          const int hib = B_NC ? 0 : MF::swz(4 * kk) & ~15;
          const int rowa = row + (lr ^ (swa & 15)), rowb = row + (lr ^ (swb & 15));
          constexpr bool SPLIT_B = A_KC || B_NC;
#pragma unroll
          for (int i = 0; i < TIM; ++i) af[i] = as[rowa + ((wm + i * 16) ^ hia)];
#pragma unroll
          for (int j = 0; j < TIN; ++j)
            bf[j] = SPLIT_B ? bs[rowb + ((wn + j * 16) ^ hib)] : bs[row + ((wn + j * 16 + lr) ^ swb)];
#pragma unroll
          for (int i = 0; i < TIM; ++i)
#pragma unroll
            for (int j = 0; j < TIN; ++j) acc[i][j] = MF::step(af[i], bf[j], acc[i][j]);
        }
        __builtin_amdgcn_s_setprio(0);
      }
      if (more) {
        ra.template store<A_KC>(As[cur ^ 1], tid);
        rb.template store<!B_NC>(Bs[cur ^ 1], tid);
      }
      __syncthreads();
    }
    if constexpr (BK < MF::BK) {
      if (pad_fix) zero_pad_fix<T>(acc);
    }
  }


  // Epilogue.  MFMA 16x16 output layout: lane l, register r holds column
  // l%16 and row MF::row(l/16, r) of its 16 x 16 tile.  C is read and written
  // through the same pointer, so the compiler must assume a store may alias
  // the next load: reading each element just before writing it would make
  // every element a load -> wait -> store round trip.  The C values of a row
  // of tiles are therefore all loaded before any of them is written (one
  // memory latency per row of tiles instead of one per element).
  if constexpr (SMALL) {
    if (p.rs_c == 1) {
      // Straight from the accumulators, in the runs described above.
      const T alpha = p.alpha, beta = p.beta;
      constexpr int CH = CU < 2 ? CU : 2;  // runs of C in flight per thread
#pragma unroll
      for (int u0 = 0; u0 < CU; u0 += CH) {
        T cv[CH][CV];
        if (beta != T(0)) {
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            if constexpr (CPRE) {
              if (have_c) {
#pragma unroll
                for (int e = 0; e < CV; ++e) cv[u][e] = cpre[u0 + u][e];
                continue;
              }
            }
            load_run(run_row(u0 + u), run_col(u0 + u), cv[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int i = (u0 + u) / TIN, j = (u0 + u) % TIN;
          const int row = run_row(u0 + u), col = run_col(u0 + u);
          T out[CV];
#pragma unroll
          for (int e = 0; e < CV; ++e) {
            const T prod = alpha * acc[i][j][e];
            out[e] = beta == T(0) ? prod : beta == T(1) ? cv[u][e] + prod : prod + beta * cv[u][e];
          }
          T *cp = p.c + (int64_t)row + (int64_t)col * p.cs_c;
          if (run_whole(row, col)) {
            cvec_t t;
#pragma unroll
            for (int e = 0; e < CV; ++e) t[e] = out[e];
            __builtin_memcpy(gmem(cp), &t, sizeof t);
          } else {
#pragma unroll
            for (int e = 0; e < CV; ++e)
              if (c_in(row + e, col)) gmem(cp)[e] = out[e];
          }
        }
      }
      return;
    }
  }
  const T alpha = p.alpha, beta = p.beta;
  auto c_ok = [&](int row, int col) {
    return row < p.m && col < p.n && !(p.ji == 1 && col > row) && !(p.ji == 2 && col < row);
  };
  T *cw = p.c;
  int64_t rs_w = p.rs_c, cs_w = p.cs_c;
  if (m0 >= p.c2_lo && m0 < p.c2_hi) {
    cw = p.c2 - (int64_t)p.c2_lo * p.rs_c2;
    rs_w = p.rs_c2;
    cs_w = p.cs_c2;
  }
  if constexpr (sizeof(T) == 4) {
    if (p.rs_c == 1 && rs_w == 1) {
      // FP32, column-major C: a lane's 4 rows of a 16 x 16 tile are one
      // 16-byte run, moved whole where C stores all of it (run_whole), as
      // in the SMALL epilogue; the C values of a row of tiles are still all
      // read before any is written.
#pragma unroll
      for (int i = 0; i < TIM; ++i) {
        const int row = m0 + wm + i * 16 + 4 * lk;
        T cv[TIN][CV];
        if (beta != T(0)) {
#pragma unroll
          for (int j = 0; j < TIN; ++j) load_run(row, n0 + wn + j * 16 + lr, cv[j]);
        }
#pragma unroll
        for (int j = 0; j < TIN; ++j) {
          const int col = n0 + wn + j * 16 + lr;
          T out[CV];
#pragma unroll
          for (int e = 0; e < CV; ++e) {
            const T prod = alpha * acc[i][j][e];
            out[e] = beta == T(0) ? prod : beta == T(1) ? cv[j][e] + prod : prod + beta * cv[j][e];
          }
          T *cp = cw + (int64_t)row + (int64_t)col * cs_w;
          if (run_whole(row, col)) {
            cvec_t t;
#pragma unroll
            for (int e = 0; e < CV; ++e) t[e] = out[e];
            __builtin_memcpy(gmem(cp), &t, sizeof t);
          } else {
#pragma unroll
            for (int e = 0; e < CV; ++e)
              if (c_ok(row + e, col)) gmem(cp)[e] = out[e];
          }
        }
      }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < TIM; ++i) {
    T cv[TIN][4];
    if (beta != T(0)) {
#pragma unroll
      for (int j = 0; j < TIN; ++j) {
        const int col = n0 + wn + j * 16 + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + i * 16 + MF::row(lk, r);
          cv[j][r] =
              c_ok(row, col) ? gmem(p.c)[(int64_t)row * p.rs_c + (int64_t)col * p.cs_c] : T(0);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < TIN; ++j) {
      const int col = n0 + wn + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + MF::row(lk, r);
        if (!c_ok(row, col)) continue;
        auto *cp = gmem(cw) + (int64_t)row * rs_w + (int64_t)col * cs_w;
        const T prod = alpha * acc[i][j][r];
        if (beta == T(0)) *cp = prod;
        else if (beta == T(1)) *cp = cv[j][r] + prod;
        else *cp = prod + beta * cv[j][r];
      }
    }
  }
}

}  // namespace gfx950

#endif
