// gfx950 triangular solve: B := alpha * A^{-1} * B for a lower or upper
// triangular A on the left (the form the reference reduces every TRSM to,
// frame/l3/trsm.c:35-54).
//
// The reference walks kc-row diagonal blocks, solving each with the fused
// GEMMTRSM micro-kernel and pushing a rank-kc GEMM update through the rest
// of B (frame/l3/trsm.c:91-431).  Two GPU schedules replace that loop:
//
// * wide B: recursive -- split A in halves, solve the first, update the
//   second half of B with ONE large MFMA GEMM (K = m/2), recurse.  All but
//   O(m * 512 * n) of the flops run in big GEMMs.  The 512-row diagonal
//   blocks of A are inverted up front (one launch of independent pipelined
//   solves against the identity), so each leaf is one more GEMM,
//   X_i = alpha * inv(A_ii) * B_i, with no chain of dependent block rows.
// * narrow B: one pipelined launch over block rows (below).
//
// Both start from the same prep launch, which inverts every NB x NB
// diagonal block in FP64 (also for float data), in parallel.  The
// reference's micro-kernel multiplies by reciprocal diagonals
// (targets/reference/gemmtrsm_ukr.inc:29-40, frame/l1m/packm.c:143-144);
// here the whole diagonal block is inverted once, so the dependent part of
// every leaf is one small matrix product instead of an NB-step substitution.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>

#include "gfx950_common.h"

namespace {

constexpr int NB = 64;  // diagonal block size
// Workspace per diagonal block, in doubles: its FP64 inverse plus room for
// the pipeline's per-block flags.
constexpr int BLK = NB * NB + NB;

// diag_mode: 0 read the diagonal, 1 unit (never read), 2 the stored diagonal
// is already the reciprocal (GEMMTRSM micro-kernel contract,
// include/mublis_instantiate.h:84-86).

// ---------------------------------------------------------------------------
// Pipelined GEMMTRSM for narrow right-hand sides.
//
// With few columns in B the recursive schedule is a chain of ~2 m / NB small
// launches, each too narrow to fill the GPU.  Here one launch solves the
// whole system: workgroup (j, c) owns block row j (in solve order) of
// column group c.  It accumulates the update sum_{i<j} A(j, i) X(i) block by
// block as soon as X(i) is published, then forms X(j) = inv(A(j, j)) *
// (alpha B(j) - update) and publishes it through a per-block flag -- the
// fused GEMM + triangular-solve step of the reference's GEMMTRSM
// micro-kernel (targets/reference/gemmtrsm_ukr.inc:1-45), pipelined across
// block rows instead of looped.  A workgroup only waits for flags of
// workgroups with lower linear ids, which the dispatcher starts first, so
// every wait ends; the waits are also bounded, and a timed-out wait raises
// an abort flag that releases all others.  Updates and the inverse product
// run in FP64 MFMA, also for float data; the inverted diagonal blocks are
// computed in FP64 by one prep launch that also clears the flags.
constexpr int PB = NB;  // block rows of the pipeline

// Inverse of diagonal block p of A into inv + p * PB * PB (row-major), by a
// workgroup of kInvThreads (four waves), by blocked doubling in the lower
// frame (an upper block is index-reversed, which makes it lower triangular):
//  1. wave w inverts 16 x 16 diagonal sub-block w by 16-step substitution,
//     one column per lane (the four 16-lane groups compute the same column);
//  2. waves 0 and 1 complete the two 32 x 32 diagonal inverses,
//     X21 = -inv(D1) * L21 * inv(D0), each two chains of four FP64 MFMAs;
//  3. the four waves complete the 64 x 64 inverse's lower-left 32 x 32
//     quadrant the same way, one 16 x 16 tile each.
// The same recurrence the recursive solve uses for its large inverses
// (inv_double below); a single 64-step substitution per column is a chain of
// dependent LDS reads that keeps one wave busy for ~30 us.  W is the LDS:
// the block's strict lower triangle in frame indices (W[r][c] = L(r, c)),
// overwritten in place by the inverse; its upper-right quadrant parks the
// phase-3 products that another wave needs.  rd is the reciprocal diagonal.
constexpr int kInvThreads = 256;

template <typename T, bool LOWER>
__device__ __forceinline__ void inv_prep_block(int p, int m, const T *__restrict__ a, int64_t rs_a,
                                               int64_t cs_a, int diag_mode, double *__restrict__ inv,
                                               double (*W)[PB + 1], double *rd) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int base = p * PB;
  const int nbr = min(PB, m - base);
  // block index -> frame index
  auto fr = [](int i) { return LOWER ? i : PB - 1 - i; };
  // lanes run along A's unit stride
  const bool cm = rs_a == 1;
  {
    // 16 elements per thread, all loads in flight before the LDS stores.  A
    // whole block is loaded without bounds checks and the triangle is
    // selected at the store; rows and columns past m stay 0 (identity rows).
    T v[16];
    if (nbr == PB) {
#pragma unroll
      for (int d = 0; d < 16; ++d) {
        const int q = wave + 4 * d, r = cm ? lane : q, c = cm ? q : lane;
        v[d] = a[(int64_t)(base + r) * rs_a + (int64_t)(base + c) * cs_a];
      }
    } else {
#pragma unroll
      for (int d = 0; d < 16; ++d) {
        const int q = wave + 4 * d, r = cm ? lane : q, c = cm ? q : lane;
        v[d] = (r < nbr && c < nbr) ? a[(int64_t)(base + r) * rs_a + (int64_t)(base + c) * cs_a]
                                    : T(0);
      }
    }
#pragma unroll
    for (int d = 0; d < 16; ++d) {
      const int q = wave + 4 * d, rf = fr(cm ? lane : q), cf = fr(cm ? q : lane);
      W[rf][cf] = rf > cf ? (double)v[d] : 0.0;
    }
  }
  if (tid < PB) {
    double v = 1.0;  // rows past m: identity
    if (tid < nbr && diag_mode != 1) {
      const int64_t g = base + tid;
      const double d = (double)a[g * rs_a + g * cs_a];
      v = diag_mode == 2 ? d : 1.0 / d;
    }
    rd[fr(tid)] = v;
  }
  __syncthreads();
  // 1. Right-looking substitution on sub-block w: once x_t is final it is
  // pushed into every later row; the triangle is read as LDS broadcasts.
  {
    const int o = 16 * wave, cc = lane & 15;
    double s[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) s[u] = u == cc ? 1.0 : 0.0;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const double xt = s[t] * rd[o + t];
      s[t] = xt;
#pragma unroll
      for (int u = t + 1; u < 16; ++u) s[u] = fma(-W[o + u][o + t], xt, s[u]);
    }
    // the wave's own reads of its sub-block precede these stores
    if (lane < 16) {
#pragma unroll
      for (int u = 0; u < 16; ++u) W[o + u][o + cc] = s[u];
    }
  }
  __syncthreads();
  // FP64 MFMA 16x16x4: lane l supplies A(l & 15, k0 + (l >> 4)) and
  // B(k0 + (l >> 4), l & 15); accumulator register r holds D((l >> 4) + 4 r,
  // l & 15), which is also the B operand of a following product's k slice r.
  const int lr = lane & 15, lk = lane >> 4;
  const d4 z = {0.0, 0.0, 0.0, 0.0};
  // 2. The two 32 x 32 diagonal blocks.
  if (wave < 2) {
    const int o = 32 * wave;
    d4 t = z;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      t = __builtin_amdgcn_mfma_f64_16x16x4f64(W[o + 16 + lr][o + 4 * kk + lk],
                                               W[o + 4 * kk + lk][o + lr], t, 0, 0, 0);
    d4 x = z;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      x = __builtin_amdgcn_mfma_f64_16x16x4f64(W[o + 16 + lr][o + 16 + 4 * kk + lk], -t[kk], x,
                                               0, 0, 0);
    // overwrites L21, which only this wave read
#pragma unroll
    for (int r = 0; r < 4; ++r) W[o + 16 + lk + 4 * r][o + lr] = x[r];
  }
  __syncthreads();
  // 3. Lower-left quadrant: T = C * inv(A), X21 = -inv(B) * T, where A and
  // B are the 32 x 32 diagonal blocks and C the quadrant of L.  Wave w owns
  // tile (i, n) = (w >> 1, w & 1); inv(A) tile (k, n) is 0 for k < n.
  {
    const int i = wave >> 1, n = wave & 1;
    d4 t = z;
    for (int kt = n; kt < 2; ++kt) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
        t = __builtin_amdgcn_mfma_f64_16x16x4f64(W[32 + 16 * i + lr][16 * kt + 4 * kk + lk],
                                                 W[16 * kt + 4 * kk + lk][16 * n + lr], t, 0, 0,
                                                 0);
    }
    // wave (1, n) also needs T(0, n)
    if (i == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) W[lk + 4 * r][32 + 16 * n + lr] = t[r];
    }
    __syncthreads();  // also: every read of C precedes the stores below
    d4 x = z;
    if (i == 1) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
        x = __builtin_amdgcn_mfma_f64_16x16x4f64(W[48 + lr][32 + 4 * kk + lk],
                                                 -W[4 * kk + lk][32 + 16 * n + lr], x, 0, 0, 0);
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      x = __builtin_amdgcn_mfma_f64_16x16x4f64(W[32 + 16 * i + lr][32 + 16 * i + 4 * kk + lk],
                                               -t[kk], x, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) W[32 + 16 * i + lk + 4 * r][16 * n + lr] = x[r];
  }
  __syncthreads();
  // Row-major inverse in block indices; the frame's upper triangle is 0.
  double *o = inv + (int64_t)p * PB * PB;
#pragma unroll
  for (int d = 0; d < 16; ++d) {
    const int r = wave + 4 * d, rf = fr(r), cf = fr(lane);
    o[r * PB + lane] = rf >= cf ? W[rf][cf] : 0.0;
  }
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(kInvThreads) void trsm_inv_prep(int m, const T *__restrict__ a,
                                                             int64_t rs_a, int64_t cs_a,
                                                             int diag_mode,
                                                             double *__restrict__ inv,
                                                             int *__restrict__ flags, int nflags) {
  __shared__ double W[PB][PB + 1];
  __shared__ double rd[PB];
  const int p = blockIdx.x;
  for (int i = p * kInvThreads + (int)threadIdx.x; i < nflags; i += gridDim.x * kInvThreads)
    flags[i] = 0;
  inv_prep_block<T, LOWER>(p, m, a, rs_a, cs_a, diag_mode, inv, W, rd);
}

// Spin until *f == epoch (or the abort flag is up; after max_spins polls the
// wait gives up and raises it).  The loads are relaxed: an acquiring load
// invalidates the L1 and the XCD's L2 on every iteration, which would evict
// everyone's operands while hundreds of workgroups wait.  Callers order
// their data loads after the wait with one acquire fence.
__device__ inline void wait_flag(int *f, int epoch, int *abort_flag, int max_spins) {
  int spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    __builtin_amdgcn_s_sleep(2);
    if (++spins > max_spins) {
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
}

// Registers <-> block row `blk` x columns c0 .. c0+CW of B, along B's unit
// stride (zero outside the matrix).  Element e = it * 256 + tid of the tile
// is (r, cc) = (e % PB, e / PB) when rs_b == 1, else (e / CW, e % CW).
// Loads keep the operand's own type; values are widened to FP64 where they
// are used (the LDS store, the MFMA operand).  Tiles inside the matrix load
// without bounds checks: a float load under a per-element check has its
// widening folded into the check's branch, together with a wait for the
// load, which serialises the tile's loads.
template <int CW>
__device__ inline void tile_rc(int e, bool colmajor, int &r, int &cc) {
  if (colmajor) { r = e % PB; cc = e / PB; }
  else { r = e / CW; cc = e % CW; }
}

template <typename T, int CW>
__device__ inline void load_tile(T (&v)[PB * CW / 256], const T *b, int64_t rs_b,
                                 int64_t cs_b, int m, int n, int blk, int c0) {
  const bool cm = rs_b == 1;
  if ((blk + 1) * PB <= m && c0 + CW <= n) {
#pragma unroll
    for (int it = 0; it < PB * CW / 256; ++it) {
      int r, cc;
      tile_rc<CW>(it * 256 + (int)threadIdx.x, cm, r, cc);
      v[it] = b[(int64_t)(blk * PB + r) * rs_b + (int64_t)(c0 + cc) * cs_b];
    }
    return;
  }
#pragma unroll
  for (int it = 0; it < PB * CW / 256; ++it) {
    int r, cc;
    tile_rc<CW>(it * 256 + (int)threadIdx.x, cm, r, cc);
    const int gr = blk * PB + r, gc = c0 + cc;
    v[it] = (gr < m && gc < n) ? b[(int64_t)gr * rs_b + (int64_t)gc * cs_b] : T(0);
  }
}

template <int CW, typename T>
__device__ inline void tile_to_lds(const T (&v)[PB * CW / 256], double (*xs)[CW + 1], bool cm) {
#pragma unroll
  for (int it = 0; it < PB * CW / 256; ++it) {
    int r, cc;
    tile_rc<CW>(it * 256 + (int)threadIdx.x, cm, r, cc);
    xs[r][cc] = (double)v[it];
  }
}

// This lane's 16 values of A(p rows, block q columns) as the MFMA A operand
// (widened at the MFMA); `inside`: block row p and block column q lie inside
// the matrix (no bounds checks).
template <typename T>
__device__ inline void load_ablk(T (&av)[16], const T *a, int64_t rs_a, int64_t cs_a,
                                 int m, int arow, int q, int lk, bool inside) {
  if (inside) {
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
      av[kk] = a[(int64_t)arow * rs_a + (int64_t)(q * PB + 4 * kk + lk) * cs_a];
    return;
  }
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int col = q * PB + 4 * kk + lk;
    av[kk] = (arow < m && col < m) ? a[(int64_t)arow * rs_a + (int64_t)col * cs_a] : T(0);
  }
}

// Independent solves in one launch (blockIdx.y = solve z): solve z covers
// rows z * m_blk .. of the m-row problem, its A, B, inverted blocks and
// flags start `a`, `b`, `inv`, `flags` elements after solve z - 1's.  A
// single solve is m_blk = m with zero strides.  `eye`: every solve's
// right-hand side is the identity.
struct PipeBatch {
  int m_blk;
  int64_t a, b, inv;
  int flags;
  int eye;
};

// LDS of a pipeline workgroup: two 64 x (CW + 1) X buffers, at least one
// 64 x 65 array (the folded product's staging).
constexpr int pipe_lds_doubles(int cw) {
  return 2 * PB * (cw + 1) > PB * (PB + 1) ? 2 * PB * (cw + 1) : PB * (PB + 1);
}

template <typename T, bool LOWER, int CW>
__global__ __launch_bounds__(256) void trsm_pipe(int m, int n, T alpha,
                                                 const T *__restrict__ a, int64_t rs_a,
                                                 int64_t cs_a, T *b, int64_t rs_b,
                                                 int64_t cs_b, const double *__restrict__ inv,
                                                 int *flags, int cg, int epoch, int max_spins,
                                                 PipeBatch bt) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int TN = CW / 16;          // 16-column MFMA tiles per wave
  constexpr int TV = PB * CW / 256;    // tile elements per thread
  // Folding the last dependency (below) needs M = inv(A(p, p)) A(p, q),
  // which depends on no X: it is formed when the workgroup starts, staged
  // through the LDS of the two X buffers viewed as one 64 x 65 array (sized
  // for it with 16-column groups).
  __shared__ double lds[pipe_lds_doubles(CW)];
  double(*xs)[PB][CW + 1] = reinterpret_cast<double(*)[PB][CW + 1]>(lds);
  double(*ab)[PB + 1] = reinterpret_cast<double(*)[PB + 1]>(lds);
  __shared__ int s_front;
  {
    const int z = blockIdx.y;
    m = min(bt.m_blk, m - z * bt.m_blk);
    a += z * bt.a;
    b += z * bt.b;
    inv += z * bt.inv;
    flags += z * bt.flags;
  }
  const int nbk = (m + PB - 1) / PB;
  const int j = blockIdx.x / cg, cgi = blockIdx.x % cg;
  // Block rows past a short last solve: nothing to do, and no workgroup
  // waits for them (waits go to lower block rows only).
  if (j >= nbk) return;
  // Identity right-hand side: block q of X is zero in the columns c0 ..
  // c0 + CW - 1 when all its rows lie above them (lower: 64 q + 63 < c0) or
  // below them (upper: 64 q > c0 + CW - 1).  Such workgroups have nothing
  // to store (B holds those zeros already) and no workgroup waits for them;
  // the others start their updates at the first nonzero block.
  int i0 = 0;
  if (bt.eye) {
    i0 = LOWER ? (cgi * CW) / PB : max(0, nbk - 1 - (cgi * CW + CW - 1) / PB);
    if (j < i0) return;
  }
  const int p = LOWER ? j : nbk - 1 - j;  // physical block row
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int c0 = cgi * CW;
  const bool cm = rs_b == 1;
  const int arow = p * PB + 16 * wave + lr;  // row of A / inv this lane feeds
  const bool prow = (p + 1) * PB <= m;       // block row p is whole
  int *abort_flag = flags;  // flags[0]; block (j, c) publishes flags[1 + j * cg + c]
  flags += 1;
  // independent of the other block rows: this block's right-hand side and
  // inverse rows, fetched before any wait
  T breg[TV];
  double ireg[16];
  load_tile<T, CW>(breg, b, rs_b, cs_b, m, n, p, c0);
  {
    const double *ip = inv + (int64_t)p * PB * PB + (16 * wave + lr) * PB;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) ireg[kk] = ip[4 * kk + lk];
  }
  // Every dependency is accumulated as it is published -- all but the last,
  // X(j - 1), which is folded in after the wait as one product
  // with M = inv(A(p, p)) A(p, q(j - 1)), formed here, before any wait: A(p,
  // q) through LDS as the B operand, M back through LDS into the A-operand
  // layout, negated.
  const bool fold = j > i0;
  const int jend = fold ? j - 1 : j;
  const int ql = LOWER ? j - 1 : nbk - j;  // physical block of X(j - 1)
  double mreg[16];
  if (fold) {
    {
      T at[PB * PB / 256];
      load_tile<T, PB>(at, a, rs_a, cs_a, m, m, p, ql * PB);
      tile_to_lds<PB>(at, ab, rs_a == 1);
    }
    __syncthreads();
    d4 mm[PB / 16];
#pragma unroll
    for (int t = 0; t < PB / 16; ++t) mm[t] = d4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
      for (int t = 0; t < PB / 16; ++t)
        mm[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ireg[kk], ab[4 * kk + lk][16 * t + lr],
                                                     mm[t], 0, 0, 0);
    __syncthreads();  // every wave has read A(p, ql)
#pragma unroll
    for (int t = 0; t < PB / 16; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) ab[16 * wave + lk + 4 * r][16 * t + lr] = mm[t][r];
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) mreg[kk] = -ab[16 * wave + lr][4 * kk + lk];
    __syncthreads();  // the X buffers are free again
  }
  d4 acc[TN];
#pragma unroll
  for (int t = 0; t < TN; ++t) acc[t] = d4{0, 0, 0, 0};
  T av[16], xr[TV];
  int i = i0;
  while (i < jend) {
    // A(p, q(i)) before the wait; then the run of published blocks i .. f-1
    {
      const int q = LOWER ? i : nbk - 1 - i;
      load_ablk<T>(av, a, rs_a, cs_a, m, arow, q, lk, prow && (q + 1) * PB <= m);
    }
    if (tid == 0) {
      wait_flag(&flags[i * cg + cgi], epoch, abort_flag, max_spins);
      int f = i + 1;
      while (f < jend &&
             __hip_atomic_load(&flags[f * cg + cgi], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == epoch)
        ++f;
      s_front = f;
    }
    __syncthreads();
    // acquire: the flags read above were published with release stores
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int f = s_front;
    load_tile<T, CW>(xr, b, rs_b, cs_b, m, n, LOWER ? i : nbk - 1 - i, c0);
    for (; i < f; ++i) {
      const int buf = i & 1;
      tile_to_lds<CW>(xr, xs[buf], cm);
      double ac[16];
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) ac[kk] = (double)av[kk];
      __syncthreads();
      if (i + 1 < f) {  // next block of the run, in flight during the MFMAs
        const int q1 = LOWER ? i + 1 : nbk - 2 - i;
        load_tile<T, CW>(xr, b, rs_b, cs_b, m, n, q1, c0);
        load_ablk<T>(av, a, rs_a, cs_a, m, arow, q1, lk, prow && (q1 + 1) * PB <= m);
      }
#pragma unroll
      for (int kk = 0; kk < 16; ++kk)
#pragma unroll
        for (int t = 0; t < TN; ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[kk], xs[buf][4 * kk + lk][16 * t + lr],
                                                        acc[t], 0, 0, 0);
    }
    __syncthreads();  // the next run reuses both buffers
  }
  // Y = inv(A(p, p)) (alpha * B(p) - update) (FP64 MFMA output row of
  // register r in lane group g is g + 4 r); R goes through xs[0]
  tile_to_lds<CW>(breg, xs[0], cm);
  __syncthreads();
  const double al = (double)alpha;
#pragma unroll
  for (int t = 0; t < TN; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double &x = xs[0][16 * wave + lk + 4 * r][16 * t + lr];
      x = al * x - acc[t][r];
    }
  __syncthreads();
  d4 out[TN];
#pragma unroll
  for (int t = 0; t < TN; ++t) out[t] = d4{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 16; ++kk)
#pragma unroll
    for (int t = 0; t < TN; ++t)
      out[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ireg[kk], xs[0][4 * kk + lk][16 * t + lr],
                                                    out[t], 0, 0, 0);
  if (fold) {
    // X(p) = Y - M X(ql): wait for X(ql), the only step after the wait
    if (tid == 0) wait_flag(&flags[(j - 1) * cg + cgi], epoch, abort_flag, max_spins);
    __syncthreads();  // also: every wave is past its Y MFMAs on xs[0]
    // acquire: X(ql) was published with a release store
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    load_tile<T, CW>(xr, b, rs_b, cs_b, m, n, ql, c0);
    tile_to_lds<CW>(xr, xs[0], cm);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
      for (int t = 0; t < TN; ++t)
        out[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(mreg[kk], xs[0][4 * kk + lk][16 * t + lr],
                                                      out[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < TN; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) xs[1][16 * wave + lk + 4 * r][16 * t + lr] = out[t][r];
  __syncthreads();
  // store X(p) along B's unit stride, then publish it.  After a timed-out
  // wait (abort flag up) X is not valid: NaN is stored instead, so that the
  // failure shows in the result rather than as plausible wrong values.
  const bool aborted =
      __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
#pragma unroll
  for (int it = 0; it < TV; ++it) {
    int r, cc;
    tile_rc<CW>(it * 256 + tid, cm, r, cc);
    const int gr = p * PB + r, gc = c0 + cc;
    if (gr < m && gc < n)
      b[(int64_t)gr * rs_b + (int64_t)gc * cs_b] =
          aborted ? T(__builtin_nan("")) : (T)xs[1][r][cc];
  }
  // Every wave's stores are complete at the barrier; thread 0's release
  // store then publishes them all (release is cumulative over the
  // workgroup's synchronisation), one L2 write-back per block.
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(&flags[j * cg + cgi], epoch, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Recursive-schedule leaf: X = inv(A(0,0)) * alpha * B for the nb <= NB
// rows of a sub-problem, CW = 64 columns of B per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void trsm_inv_apply(int nb, int n, T alpha,
                                                      const double *__restrict__ inv, T *b,
                                                      int64_t rs_b, int64_t cs_b) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int CW = 64, TN = CW / 16, TV = PB * CW / 256;
  __shared__ double xs[PB][CW + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int c0 = blockIdx.x * CW;
  const bool cm = rs_b == 1;
  double ireg[16], v[TV];
  const double *ip = inv + (16 * wave + lr) * PB;
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) ireg[kk] = ip[4 * kk + lk];
  {
    T raw[TV];
    load_tile<T, CW>(raw, b, rs_b, cs_b, nb, n, 0, c0);
    const double al = (double)alpha;
#pragma unroll
    for (int it = 0; it < TV; ++it) v[it] = al * (double)raw[it];
  }
  tile_to_lds<CW>(v, xs, cm);
  __syncthreads();
  d4 out[TN];
#pragma unroll
  for (int t = 0; t < TN; ++t) out[t] = d4{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 16; ++kk)
#pragma unroll
    for (int t = 0; t < TN; ++t)
      out[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ireg[kk], xs[4 * kk + lk][16 * t + lr],
                                                    out[t], 0, 0, 0);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < TN; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) xs[16 * wave + lk + 4 * r][16 * t + lr] = out[t][r];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < TV; ++it) {
    int r, cc;
    tile_rc<CW>(it * 256 + tid, cm, r, cc);
    const int gc = c0 + cc;
    if (r < nb && gc < n) b[(int64_t)r * rs_b + (int64_t)gc * cs_b] = (T)xs[r][cc];
  }
}

// Recursive-schedule leaf for wide B: the sub-problem's (at most
// kPanelRows) rows times PW = 16 columns of B per workgroup.  The workgroup
// walks the block rows in solve order itself, keeping the solved rows in
// LDS: X(j) = inv(A(j,j)) * (alpha B(j) - sum_{i<j} A(j,i) X(i)), the
// reference's GEMMTRSM step (targets/reference/gemmtrsm_ukr.inc:1-45) with
// a panel of B per workgroup.  No workgroup waits for another; parallelism
// comes from the n / 16 column panels.
constexpr int kPanelRows = 512;
constexpr int PW = 16;

// Columns c0 .. c0 + W - 1 of the solve, by one workgroup (each wave 16 rows
// of a block row times W columns); xs holds at least m rows.
template <typename T, bool LOWER, int W = PW>
__device__ __forceinline__ void panel_solve(int m, int n, T alpha, const T *__restrict__ a,
                                            int64_t rs_a, int64_t cs_a, T *b, int64_t rs_b,
                                            int64_t cs_b, const double *__restrict__ inv,
                                            int c0, double (*xs)[W + 1]) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int TV = PB * W / 256;  // tile elements per thread
  constexpr int TN = W / 16;        // 16-column MFMA tiles per wave
  const int nbk = (m + PB - 1) / PB;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const bool cm = rs_b == 1;
  const double al = (double)alpha;
  for (int j = 0; j < nbk; ++j) {
    const int p = LOWER ? j : nbk - 1 - j;
    const int arow = p * PB + 16 * wave + lr;
    T breg[TV];
    double ireg[16];
    load_tile<T, W>(breg, b, rs_b, cs_b, m, n, p, c0);
    {
      const double *ip = inv + (int64_t)p * PB * PB + (16 * wave + lr) * PB;
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) ireg[kk] = ip[4 * kk + lk];
    }
    d4 acc[TN];
#pragma unroll
    for (int t = 0; t < TN; ++t) acc[t] = d4{0, 0, 0, 0};
    T av[16], an[16];
    const bool prow = (p + 1) * PB <= m;
    if (j > 0) {
      const int q = LOWER ? 0 : nbk - 1;
      load_ablk<T>(av, a, rs_a, cs_a, m, arow, q, lk, prow && (q + 1) * PB <= m);
    }
    for (int i = 0; i < j; ++i) {
      const int q = LOWER ? i : nbk - 1 - i;
      if (i + 1 < j) {
        const int q1 = LOWER ? i + 1 : nbk - 2 - i;
        load_ablk<T>(an, a, rs_a, cs_a, m, arow, q1, lk, prow && (q1 + 1) * PB <= m);
      }
#pragma unroll
      for (int kk = 0; kk < 16; ++kk)
#pragma unroll
        for (int t = 0; t < TN; ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              (double)av[kk], xs[q * PB + 4 * kk + lk][16 * t + lr], acc[t], 0, 0, 0);
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) av[kk] = an[kk];
    }
    // R = alpha B(p) - update, staged in block p's rows of xs
#pragma unroll
    for (int it = 0; it < TV; ++it) {
      int r, cc;
      tile_rc<W>(it * 256 + tid, cm, r, cc);
      xs[p * PB + r][cc] = al * (double)breg[it];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TN; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) xs[p * PB + 16 * wave + lk + 4 * r][16 * t + lr] -= acc[t][r];
    __syncthreads();
    d4 out[TN];
#pragma unroll
    for (int t = 0; t < TN; ++t) out[t] = d4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
      for (int t = 0; t < TN; ++t)
        out[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
            ireg[kk], xs[p * PB + 4 * kk + lk][16 * t + lr], out[t], 0, 0, 0);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TN; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) xs[p * PB + 16 * wave + lk + 4 * r][16 * t + lr] = out[t][r];
    __syncthreads();
  }
  for (int j = 0; j < nbk; ++j) {
#pragma unroll
    for (int it = 0; it < TV; ++it) {
      int r, cc;
      tile_rc<W>(it * 256 + tid, cm, r, cc);
      const int gr = j * PB + r, gc = c0 + cc;
      if (gr < m && gc < n) b[(int64_t)gr * rs_b + (int64_t)gc * cs_b] = (T)xs[j * PB + r][cc];
    }
  }
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_panel(int m, int n, T alpha,
                                                  const T *__restrict__ a, int64_t rs_a,
                                                  int64_t cs_a, T *b, int64_t rs_b,
                                                  int64_t cs_b,
                                                  const double *__restrict__ inv) {
  __shared__ double xs[kPanelRows][PW + 1];
  panel_solve<T, LOWER>(m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, inv, blockIdx.x * PW, xs);
}

// Grouped solves: one workgroup per item of a work list of independent
// items (no item waits for another).  The inverse launch inverts one
// diagonal block per item; the solve launches run one column panel per
// item, the same step as trsm_panel, with LDS for ROWS rows of X.  (A
// persistent loop over the items made these register-heavy bodies spill.)
static_assert(kGroupTrsmCols == PW, "grouped panels are the column panels");
static_assert(kGroupTrsmSmall == kPanelRows, "small grouped panels fit trsm_panel's LDS");
static_assert(kGroupTrsmLists == 8, "two inverse lists and three panel classes");

// One launch per triangle orientation (both in one kernel spill).
template <typename T, bool LOWER>
__global__ __launch_bounds__(kInvThreads) void grouped_inv_prep(const TrsmGroupProb<T> *__restrict__ probs,
                                                                const TrsmGroupItem *__restrict__ items,
                                                                double *inv) {
  __shared__ double W[PB][PB + 1];
  __shared__ double rd[PB];
  const TrsmGroupItem w = items[blockIdx.x];
  const TrsmGroupProb<T> &d = probs[w.prob];
  inv_prep_block<T, LOWER>(w.idx, d.m, d.a, d.rs_a, d.cs_a, d.unit, inv + d.inv, W, rd);
}

// Two workgroups per CU where LDS allows (the 32-column lower body needs the
// bound to fit 256 registers; without it, one per CU).
template <typename T, bool LOWER, int ROWS, int W>
__global__ __launch_bounds__(256, ROWS <= kGroupTrsmSmall ? 2 : 1) void grouped_panels(const TrsmGroupProb<T> *__restrict__ probs,
                                                      const TrsmGroupItem *__restrict__ items,
                                                      const double *inv) {
  __shared__ double xs[ROWS][W + 1];
  const TrsmGroupItem w = items[blockIdx.x];
  const TrsmGroupProb<T> &d = probs[w.prob];
  panel_solve<T, LOWER, W>(d.m, d.n, d.alpha, d.a, d.rs_a, d.cs_a, d.b, d.rs_b, d.cs_b,
                           inv + d.inv, w.idx * W, xs);
}


// Widest B solved by the pipelined kernel, at most 4096; 0 disables it.
// Below that cap a whole solve takes the pipeline for B up to 128 columns,
// up to 2048 columns and m * n = 2^22, or up to 768 columns and m * n =
// 2^23; beyond, the recursive schedule's large GEMMs fill the GPU better
// than one chain of block rows (measured on MI355X, FP64: pipelined 8192 x
// 512 1.38 ms vs 1.62, 8192 x 768 1.74 vs 1.89, 4096 x 1024 0.71 vs 0.79,
// 2048 x 2048 0.41 vs 0.46, 6000 x 700 1.17 vs 1.48; recursive 8192 x 1024
// 2.03 ms vs 2.18, 4096 x 2048 1.04 vs 1.18, 1024 x 4096 0.22 vs 0.26).
// Float solves wider than 128 columns with m * n from 2^21 prefer the
// recursion's GEMM leaves whenever those apply (float, leaves vs pipelined:
// 8192 x 512 1.00 ms vs 1.33, 4096 x 512 0.47 vs 0.49; below: 8192 x 192
// 0.80 pipelined vs 0.93, 2048 x 512 0.24 vs 0.27).  The recursive
// schedule in turn ends in GEMM leaves (B from 256 columns for double, from
// 64 for float; see use_inv_leaves), else hands its diagonal sub-problems of
// at most kPipeLeaf rows to the pipeline (narrower B) or to column-panel
// workgroups (wider B, when the solve brings no room for the GEMM leaves'
// inverses).
// Both knobs are set from any thread (mublis_gfx950_trsm_pipe_max_n /
// _spin_limit, used by tests to force schedules) and read by solves on
// others: they are atomics, and a solve reads them once, at its outermost
// entry, into a per-thread snapshot that everything it enqueues uses (the
// KnobScope below), so one solve never mixes two settings.
std::atomic<int> g_pipe_max_n_set{4096};
// Polls a pipeline wait makes before it gives up (about a second; the
// dispatch order makes every wait end long before, this only bounds a
// fault).
std::atomic<int> g_pipe_spins_set{1 << 20};
struct PipeKnobs {
  int max_n, spins, depth;
};
thread_local PipeKnobs t_knobs = {4096, 1 << 20, 0};
struct KnobScope {
  KnobScope() {
    if (t_knobs.depth++ == 0) {
      t_knobs.max_n = g_pipe_max_n_set.load(std::memory_order_relaxed);
      t_knobs.spins = g_pipe_spins_set.load(std::memory_order_relaxed);
    }
  }
  ~KnobScope() { --t_knobs.depth; }
  KnobScope(const KnobScope &) = delete;
  KnobScope &operator=(const KnobScope &) = delete;
};

constexpr int kPipeLeaf = 512;
// From this many columns up, the pipelined leaves of at most kPanelRows
// rows are column panels (n / 16 independent workgroups) instead.
constexpr int kPanelMinN = 2048;

bool use_pipe(int m, int n) {
  const int64_t mn = (int64_t)m * n;
  return m > NB && n <= t_knobs.max_n &&
         (n <= 128 || (n <= 2048 && mn <= (int64_t(1) << 22)) ||
          (n <= 768 && mn <= (int64_t(1) << 23)));
}

// Column-group width of the pipelined leaves of the recursive schedule
// (they share the whole solve's flags, laid out for this width).
int pipe_cw(int n) { return n <= 32 ? 16 : (n <= 128 ? 32 : 64); }

// Column-group width of a whole pipelined solve: the narrowest of 16 (B up
// to 128 columns), 32 (up to 1024) and 64 whose grid stays within 1024
// workgroups, except that 16 also needs more than 256 of them (or B of at
// most 32 columns).  Narrower groups shorten each step and add parallelism
// until the A blocks every group reloads outweigh it (measured on MI355X,
// FP64, one workgroup per CU: 8192 x 128 0.74 ms at 16 columns vs 0.80 at
// 32; 4096 x 64 0.34 at 16 vs 0.32 at 32; 8192 x 256 0.92 at 32 vs 1.17 at
// 64; 4096 x 512 0.51 at 32 vs 0.60; 8192 x 512 1.38 at 64 vs 1.45 at 32;
// 2048 x 2048 0.41 at 64 vs 0.49).
int pipe_cw_whole(int m, int n) {
  const int64_t nbk = (m + NB - 1) / NB, g16 = nbk * ((n + 15) / 16);
  if (n <= 32 || (n <= 128 && g16 > 256 && g16 <= 1024)) return 16;
  if (n <= 1024 && nbk * ((n + 31) / 32) <= 1024) return 32;
  return 64;
}

// Workspace of a solve: the inverted diagonal blocks (NB * NB doubles each,
// row-major), then the pipeline's flags: flags[0] is the abort flag,
// flags[1 + j * cg + c] the publication epoch of block row j, column group c.
// The flags fit in the NB doubles per block that follow the inverses while
// 1 + nbk * cg <= nbk * 2 * NB.
int64_t ws_doubles(int m) { return (int64_t)((m + NB - 1) / NB) * BLK; }

bool flags_fit(int m, int n) {
  const int64_t nbk = (m + NB - 1) / NB, cg = (n + pipe_cw(n) - 1) / pipe_cw(n);
  return 1 + nbk * cg <= nbk * 2 * NB;
}

template <typename T>
int prep(bool lower, int m, const T *a, int64_t rs_a, int64_t cs_a, int diag_mode,
         double *inv, int *flags, int nflags, hipStream_t s) {
  const int blocks = (m + NB - 1) / NB;
  if (lower)
    hipLaunchKernelGGL((trsm_inv_prep<T, true>), dim3(blocks), dim3(kInvThreads), 0, s, m, a, rs_a,
                       cs_a, diag_mode, inv, flags, nflags);
  else
    hipLaunchKernelGGL((trsm_inv_prep<T, false>), dim3(blocks), dim3(kInvThreads), 0, s, m, a, rs_a,
                       cs_a, diag_mode, inv, flags, nflags);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

template <typename T>
int leaf(int nb, int n, T alpha, const double *inv, T *b, int64_t rs_b, int64_t cs_b,
         hipStream_t s) {
  hipLaunchKernelGGL((trsm_inv_apply<T>), dim3((n + 63) / 64), dim3(256), 0, s, nb, n, alpha,
                     inv, b, rs_b, cs_b);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

constexpr size_t kLdsOneWg = 96 * 1024;

template <typename T, bool LOWER, int CW>
int pipe_launch(int m, int n, T alpha, const T *a, int64_t rs_a, int64_t cs_a, T *b,
                int64_t rs_b, int64_t cs_b, const double *inv, int *flags, int epoch,
                const PipeBatch &bt, hipStream_t s) {
  const int nbk = (bt.m_blk + PB - 1) / PB, cg = (n + CW - 1) / CW;
  const int nz = (m + bt.m_blk - 1) / bt.m_blk;
  // A single solve holds one workgroup per CU: a second workgroup of the
  // same chain on a CU slowed it (float 8192 x 128, 32-column groups: 1.17
  // ms at two per CU, 0.82 ms at one).  Dynamic LDS past half of the CU's
  // 160 KB keeps the second one off.  Batches of independent solves keep
  // the kernel's own occupancy.
  const size_t lds = nz == 1 ? kLdsOneWg - sizeof(double) * pipe_lds_doubles(CW) : 0;
  hipLaunchKernelGGL((trsm_pipe<T, LOWER, CW>), dim3(nbk * cg, nz), dim3(256), lds, s, m, n,
                     alpha, a, rs_a, cs_a, b, rs_b, cs_b, inv, flags, cg, epoch, t_knobs.spins,
                     bt);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

// One pipelined solve of an m-row (sub-)problem whose inverted blocks start
// at `inv`; flags must hold `epoch - 1` or less (prep clears them to 0).
// With `bt` given, the independent solves it describes instead.
template <typename T>
int pipe(bool lower, int m, int n, T alpha, const T *a, int64_t rs_a, int64_t cs_a, T *b,
         int64_t rs_b, int64_t cs_b, const double *inv, int *flags, int epoch,
         hipStream_t s, const PipeBatch *batch = nullptr, int cw = 0) {
  if (!cw) cw = pipe_cw(n);
  const PipeBatch bt = batch ? *batch : PipeBatch{m, 0, 0, 0, 0, 0};
#define MUBLIS_PIPE(L, W)                                                                   \
  return pipe_launch<T, L, W>(m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, inv, flags, epoch, \
                              bt, s)
  if (lower) {
    if (cw == 16) MUBLIS_PIPE(true, 16);
    if (cw == 32) MUBLIS_PIPE(true, 32);
    MUBLIS_PIPE(true, 64);
  }
  if (cw == 16) MUBLIS_PIPE(false, 16);
  if (cw == 32) MUBLIS_PIPE(false, 32);
  MUBLIS_PIPE(false, 64);
#undef MUBLIS_PIPE
}

// Wide right-hand sides: every kInvLeaf-row diagonal block of A is
// inverted whole (by doubling the NB-block inverses with batched GEMMs, or
// by one launch of independent pipelined solves against the identity when A
// has no unit stride), and the recursion, split at multiples of kInvLeaf,
// ends in leaves X = alpha * inv(A_ii) * B_i -- a copy of B_i and one GEMM
// that skips the inverse's zero triangle.  No leaf then waits on a chain of
// block rows.  Explicitly inverted diagonal blocks are the usual GPU TRSM
// form; their forward error grows with the blocks' conditioning, not with m.
// Block rows (both types): the leaves then hold the fewest small GEMMs,
// and inverting 1024-row blocks by doubling costs a few batched GEMMs.
template <typename T>
constexpr int kInvLeaf = 1024;
// Narrowest B for GEMM leaves.  Float solves take them whenever they
// apply (their pipeline runs the block updates in FP64: 4096 x 256 solves in
// 1.00 ms pipelined, 0.48 ms with GEMM leaves); double solves keep the
// pipeline where it is at least as fast (use_pipe).
template <typename T>
constexpr int kInvMinN = sizeof(T) == 8 ? 256 : 64;
// Column group of the identity solves: narrow groups keep the pipeline's
// LDS small, so all inversions run at once instead of in waves.
constexpr int kEyeCw = 32;

template <typename T>
bool use_inv_leaves(int m, int n) {
  return m > kInvLeaf<T> && n >= kInvMinN<T> && t_knobs.max_n >= kInvLeaf<T>;
}

template <typename T>
int64_t inv_leaf_elems(int m) {
  return (int64_t)((m + kInvLeaf<T> - 1) / kInvLeaf<T>) * kInvLeaf<T> * kInvLeaf<T>;
}

// nz = ceil(m / kInvLeaf) identity blocks, kInvLeaf x kInvLeaf column-major
// each; rows past m stay zero.
template <typename T>
__global__ void eye_blocks(int m, T *x, int64_t total) {
  constexpr int64_t KK = (int64_t)kInvLeaf<T> * kInvLeaf<T>;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = t / KK, e = t - z * KK;
    const int i = (int)(e % kInvLeaf<T>), j = (int)(e / kInvLeaf<T>);
    x[t] = (i == j && z * kInvLeaf<T> + i < m) ? T(1) : T(0);
  }
}

// iv := the inverses of A's kInvLeaf-row diagonal blocks (block z at
// iv + z * kInvLeaf^2, ld kInvLeaf), from the NB-block inverses in `ad`.
// Uses pipeline epoch 1 of `flags`.
template <typename T>
int inv_blocks(bool lower, int m, const T *a, int64_t rs_a, int64_t cs_a, const double *ad,
               int *flags, T *iv, hipStream_t s) {
  const int64_t total = inv_leaf_elems<T>(m);
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(eye_blocks<T>, dim3(grid), dim3(256), 0, s, m, iv, total);
  if (hipGetLastError() != hipSuccess) return MUBLIS_EHIP;
  const int cg = kInvLeaf<T> / kEyeCw;
  const PipeBatch bt{kInvLeaf<T>, (int64_t)kInvLeaf<T> * (rs_a + cs_a),
                     (int64_t)kInvLeaf<T> * kInvLeaf<T>,
                     (int64_t)(kInvLeaf<T> / NB) * NB * NB,
                     1 + (kInvLeaf<T> / NB) * cg, 1};
  return pipe<T>(lower, m, kInvLeaf<T>, T(1), a, rs_a, cs_a, iv, 1, kInvLeaf<T>, ad, flags, 1,
                 s, &bt, kEyeCw);
}

// iv := the block diagonal of the NB-block inverses in `ad` (row-major NB x NB
// each) in inv_blocks' layout, zero elsewhere.
template <typename T>
__global__ void inv_scatter(int m, const double *__restrict__ ad, T *__restrict__ iv,
                            int64_t total) {
  constexpr int64_t L = kInvLeaf<T>;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = t / (L * L), e = t - z * L * L;
    const int64_t r = z * L + e % L, c = z * L + e / L;  // global row, column
    T v = T(0);
    if (r < m && c < m && r / NB == c / NB)
      v = (T)ad[(r / NB) * NB * NB + (r % NB) * NB + c % NB];
    iv[t] = v;
  }
}

// The same inverses by doubling the NB-block ones: for h = NB, 2 NB, ..
// kInvLeaf / 2, every aligned 2h-block [A11 0; A21 A22] (lower) gets
//   inv21 = -inv22 * (A21 * inv11)
// and every [A11 A12; 0 A22] (upper) inv12 = -inv11 * (A12 * inv22) -- two
// batched GEMMs per level over all such blocks at once, each skipping its
// triangular operand's zero triangle, instead of a chain of dependent block
// rows.  The last pair of a level may be short (m not a multiple of 2h):
// two exact-size GEMMs.  `tmp` holds m * kInvLeaf / 4 elements.  A needs a
// unit stride.
template <typename T>
int inv_double(bool lower, int m, const T *a, int64_t rs_a, int64_t cs_a, const double *ad,
               T *iv, T *tmp, hipStream_t s) {
  constexpr int L = kInvLeaf<T>;
  const int64_t total = inv_leaf_elems<T>(m);
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(inv_scatter<T>, dim3(grid), dim3(256), 0, s, m, ad, iv, total);
  if (hipGetLastError() != hipSuccess) return MUBLIS_EHIP;
  const int64_t da = rs_a + cs_a;  // one diagonal step in A
  int st = MUBLIS_OK;
  for (int h = NB; h < L && st == MUBLIS_OK; h *= 2) {
    const int per = L / (2 * h);  // pairs per L-block
    const int full = m / (2 * h);
    const int64_t sA[2] = {2 * h * da, (int64_t)L * da};
    const int64_t sI[2] = {(int64_t)2 * h * (1 + L), (int64_t)L * L};
    const int64_t sT[2] = {(int64_t)h * h, (int64_t)per * h * h};
    const int64_t d1 = (int64_t)h * (1 + L);  // second diagonal half of a pair
    if (full > 0) {
      if (lower) {
        st = gfx950_gemm_batched<T>(h, h, h, T(1), a + (int64_t)h * rs_a, rs_a, cs_a, sA, iv,
                                    1, L, sI, T(0), tmp, 1, h, sT, full, per, 0, 1, s);
        if (st == MUBLIS_OK)
          st = gfx950_gemm_batched<T>(h, h, h, T(-1), iv + d1, 1, L, sI, tmp, 1, h, sT, T(0),
                                      iv + h, 1, L, sI, full, per, 1, 0, s);
      } else {
        st = gfx950_gemm_batched<T>(h, h, h, T(1), a + (int64_t)h * cs_a, rs_a, cs_a, sA,
                                    iv + d1, 1, L, sI, T(0), tmp, 1, h, sT, full, per, 0, 2, s);
        if (st == MUBLIS_OK)
          st = gfx950_gemm_batched<T>(h, h, h, T(-1), iv, 1, L, sI, tmp, 1, h, sT, T(0),
                                      iv + (int64_t)h * L, 1, L, sI, full, per, 2, 0, s);
      }
    }
    const int o = full * 2 * h, h2 = m - o - h;  // a short last pair
    if (st == MUBLIS_OK && h2 > 0) {
      T *ivz = iv + (int64_t)(o / L) * L * L;
      const int64_t r0 = o % L;
      const T *a0 = a + (int64_t)o * da;
      if (lower) {
        st = gfx950_gemm<T>(h2, h, h, T(1), a0 + (int64_t)h * rs_a, rs_a, cs_a,
                            ivz + r0 * (1 + L), 1, L, T(0), tmp, 1, h2, s);
        if (st == MUBLIS_OK)
          st = gfx950_gemm<T>(h2, h, h2, T(-1), ivz + r0 * (1 + L) + d1, 1, L, tmp, 1, h2,
                              T(0), ivz + r0 * (1 + L) + h, 1, L, s);
      } else {
        st = gfx950_gemm<T>(h, h2, h2, T(1), a0 + (int64_t)h * cs_a, rs_a, cs_a,
                            ivz + r0 * (1 + L) + d1, 1, L, T(0), tmp, 1, h, s);
        if (st == MUBLIS_OK)
          st = gfx950_gemm<T>(h, h2, h, T(-1), ivz + r0 * (1 + L), 1, L, tmp, 1, h, T(0),
                              ivz + r0 * (1 + L) + (int64_t)h * L, 1, L, s);
      }
    }
  }
  return st;
}

// Flags for inv_blocks: one abort flag and (kInvLeaf / NB) x cg block
// flags per diagonal block.
template <typename T>
int64_t inv_flag_count(int m) {
  const int cg = kInvLeaf<T> / kEyeCw;
  return (int64_t)((m + kInvLeaf<T> - 1) / kInvLeaf<T>) * (1 + (kInvLeaf<T> / NB) * cg);
}

// How the recursion's diagonal sub-problems end: GEMM leaves when `w` (a
// kInvLeaf x n buffer for the leaf's right-hand side) is set, else
// pipelined or column-panel leaves when `flags` is set, else NB-row inverse
// products.
//  * w_ready: the update before the next GEMM leaf has written that leaf's
//    rows of B into w already (gfx950_gemm_split_c), so the leaf copies
//    nothing;
template <typename T>
struct Leaves {
  int *flags;
  int epoch;
  T *w;
  bool w_ready = false;
};

// Recursive schedule.  `ad` holds the NB-block inverses of this sub-problem
// (block 0 is its top-left NB x NB block), `iv` its kInvLeaf-block inverses
// for GEMM leaves.
template <typename T>
int trsm_rec(bool lower, const double *ad, const T *iv, int m, int n, T alpha, const T *a,
             int64_t rs_a, int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, Leaves<T> &lv,
             hipStream_t s) {
  if (lv.w && m <= kInvLeaf<T>) {
    int st = MUBLIS_OK;
    if (!lv.w_ready) st = gfx950_copy<T>(m, n, b, rs_b, cs_b, lv.w, 1, m, s);
    lv.w_ready = false;
    if (st) return st;
    return gfx950_gemm_tri_a<T>(m, n, m, alpha, iv, 1, kInvLeaf<T>, lv.w, 1, m, T(0), b, rs_b,
                                cs_b, lower ? 1 : 2, s);
  }
  if (m <= NB) return leaf<T>(m, n, alpha, ad, b, rs_b, cs_b, s);
  if (!lv.w && lv.flags && m <= kPanelRows && n >= kPanelMinN) {
    if (lower)
      hipLaunchKernelGGL((trsm_panel<T, true>), dim3((n + PW - 1) / PW), dim3(256), 0, s, m, n,
                         alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad);
    else
      hipLaunchKernelGGL((trsm_panel<T, false>), dim3((n + PW - 1) / PW), dim3(256), 0, s, m,
                         n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad);
    return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
  }
  if (!lv.w && lv.flags && m <= kPipeLeaf)
    return pipe<T>(lower, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad, lv.flags, ++lv.epoch,
                   s);
  // First part: a multiple of NB (of kInvLeaf with GEMM leaves) close to
  // half, so GEMM K and M stay tile friendly at every level.
  const int unit = lv.w ? kInvLeaf<T> : NB;
  int m1 = (m / 2 + unit - 1) / unit * unit;
  if (m1 >= m) m1 = m - NB;
  const int m2 = m - m1;
  const T *a11 = a;
  const T *a21 = a + (int64_t)m1 * rs_a;
  const T *a12 = a + (int64_t)m1 * cs_a;
  const T *a22 = a + (int64_t)m1 * rs_a + (int64_t)m1 * cs_a;
  T *b1 = b, *b2 = b + (int64_t)m1 * rs_b;
  const double *ad1 = ad, *ad2 = ad + (int64_t)(m1 / NB) * NB * NB;
  constexpr int64_t KL = kInvLeaf<T>;
  const T *iv1 = iv, *iv2 = iv ? iv + (int64_t)(m1 / KL) * KL * KL : iv;
  // With GEMM leaves the update writes the rows of the next leaf (the first
  // kInvLeaf-block of B2 for lower, the last of B1 for upper: the splits sit
  // at multiples of kInvLeaf) straight into the leaf's buffer.
  auto update = [&](int mu, int ku, const T *au, const T *xu, T *bu, int lo, int hi) {
    if (lv.w) {
      const int st = gfx950_gemm_split_c<T>(mu, n, ku, T(-1), au, rs_a, cs_a, xu, rs_b, cs_b,
                                            alpha, bu, rs_b, cs_b, lv.w, 1, hi - lo, lo, hi,
                                            s);
      lv.w_ready = st == MUBLIS_OK;
      return st;
    }
    return gfx950_gemm<T>(mu, n, ku, T(-1), au, rs_a, cs_a, xu, rs_b, cs_b, alpha, bu, rs_b,
                          cs_b, s);
  };
  int st;
  if (lower) {
    if ((st = trsm_rec<T>(lower, ad1, iv1, m1, n, alpha, a11, rs_a, cs_a, b1, rs_b, cs_b, lv,
                          s)))
      return st;
    // B2 := alpha*B2 - A21 * X1 (the reference's GEMM with alpha -1 and
    // beta = alpha_eff, frame/l3/trsm.c:165-197, 241-249)
    if ((st = update(m2, m1, a21, b1, b2, 0, std::min(m2, (int)KL))))
      return st;
    return trsm_rec<T>(lower, ad2, iv2, m2, n, T(1), a22, rs_a, cs_a, b2, rs_b, cs_b, lv, s);
  }
  if ((st = trsm_rec<T>(lower, ad2, iv2, m2, n, alpha, a22, rs_a, cs_a, b2, rs_b, cs_b, lv,
                        s)))
    return st;
  if ((st = update(m1, m2, a12, b2, b1, m1 - (int)KL, m1)))
    return st;
  return trsm_rec<T>(lower, ad1, iv1, m1, n, T(1), a11, rs_a, cs_a, b1, rs_b, cs_b, lv, s);
}

// Invert the diagonal blocks into `ad` (ws_doubles(m) doubles) and run the
// pipelined or the recursive schedule; asynchronous on s.  With `iv`
// (inv_leaf_elems<T>(m) elements) and `w` (kInvLeaf<T> * n elements) given and
// use_inv_leaves<T>(m, n), the recursion ends in GEMM leaves.
template <typename T>
int solve_ws(bool lower, int diag_mode, int m, int n, T alpha, const T *a, int64_t rs_a,
             int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, double *ad, T *iv, T *w, T *tmp,
             hipStream_t s, bool *inv_ready = nullptr) {
  KnobScope knobs;
  const int nbk = (m + NB - 1) / NB;
  const bool piped = m > NB && n <= t_knobs.max_n && flags_fit(m, n);
  const bool gemm_ok = iv && w && use_inv_leaves<T>(m, n);
  const bool f32_leaves = sizeof(T) == 4 && gemm_ok && n > 128 &&
                          (int64_t)m * n >= (int64_t(1) << 21);
  const bool whole = piped && use_pipe(m, n) && !f32_leaves;
  const bool gemm_leaves = !whole && gemm_ok;
  int *flags = (piped || gemm_leaves) ? reinterpret_cast<int *>(ad + (int64_t)nbk * NB * NB)
                                      : nullptr;
  const int cw = whole ? pipe_cw_whole(m, n) : pipe_cw(n);
  const int cg = (n + cw - 1) / cw;
  int64_t nflags = piped ? 1 + (int64_t)nbk * cg : 0;
  if (gemm_leaves) nflags = std::max(nflags, inv_flag_count<T>(m));
  if (gemm_leaves && inv_ready && *inv_ready) {
    // the inverses of this triangle are in iv from an earlier solve: GEMM
    // leaves read nothing else
    Leaves<T> lv{flags, 1, w};
    return trsm_rec<T>(lower, ad, iv, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, lv, s);
  }
  int st = prep<T>(lower, m, a, rs_a, cs_a, diag_mode, ad, flags, (int)nflags, s);
  if (st) return st;
  if (whole)
    return pipe<T>(lower, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad, flags, 1, s, nullptr,
                   cw);
  Leaves<T> lv{flags, 0, nullptr};
  if (gemm_leaves) {
    if (tmp && (rs_a == 1 || cs_a == 1))
      st = inv_double<T>(lower, m, a, rs_a, cs_a, ad, iv, tmp, s);
    else
      st = inv_blocks<T>(lower, m, a, rs_a, cs_a, ad, flags, iv, s);
    if (st) return st;
    lv.epoch = 1;
    lv.w = w;
    if (inv_ready) *inv_ready = true;
  }
  return trsm_rec<T>(lower, ad, gemm_leaves ? iv : nullptr, m, n, alpha, a, rs_a, cs_a, b,
                     rs_b, cs_b, lv, s);
}

// Same with workspace from the arena, released once the stream has drained.
template <typename T>
int solve(bool lower, int diag_mode, int m, int n, T alpha, const T *a, int64_t rs_a,
          int64_t cs_a, T *b, int64_t rs_b, int64_t cs_b, hipStream_t s) {
  KnobScope knobs;
  double *ad = (double *)mublis_rt_alloc(sizeof(double) * (size_t)ws_doubles(m));
  if (!ad) return MUBLIS_ENOMEM;
  T *iv = nullptr, *w = nullptr, *tmp = nullptr;
  if (use_inv_leaves<T>(m, n)) {
    // without room for them, the pipelined / panel leaves instead
    iv = (T *)mublis_rt_alloc(sizeof(T) * (size_t)inv_leaf_elems<T>(m));
    w = (T *)mublis_rt_alloc(sizeof(T) * (size_t)kInvLeaf<T> * (size_t)n);
    // doubling workspace; without it the inverses come from identity solves
    tmp = (T *)mublis_rt_alloc(sizeof(T) * (size_t)m * kInvLeaf<T> / 4);
  }
  int st = solve_ws<T>(lower, diag_mode, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad, iv, w,
                       tmp, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  mublis_rt_free(ad);
  mublis_rt_free(iv);
  mublis_rt_free(w);
  mublis_rt_free(tmp);
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

template <typename T>
int gfx950_trsm_grouped(const TrsmGroupProb<T> *probs, const TrsmGroupItem *items,
                        const int *counts, double *inv, hipStream_t s, hipStream_t s2,
                        hipEvent_t fork, hipEvent_t join) {
  // The inverse kernel keeps a 64-double column per lane in registers (one
  // wavefront per SIMD); the panels hold 256 x 32, 512 x 16 or 1024 x 16
  // doubles of X in LDS, two, two and one workgroups per CU.  Lists run
  // longest items first.  Lower and upper problems are independent: with a
  // second stream they run as two concurrent branches.
  const bool two = s2 && fork && join;
  if (two && (hipEventRecord(fork, s) != hipSuccess || hipStreamWaitEvent(s2, fork, 0) != hipSuccess))
    return MUBLIS_EHIP;
  hipStream_t su = two ? s2 : s;
  const TrsmGroupItem *it[kGroupTrsmLists];
  it[0] = items;
  for (int c = 1; c < kGroupTrsmLists; ++c) it[c] = it[c - 1] + counts[c - 1];
  if (counts[0] > 0)
    hipLaunchKernelGGL((grouped_inv_prep<T, true>), dim3(counts[0]), dim3(kInvThreads), 0, s, probs,
                       it[0], inv);
  if (counts[1] > 0)
    hipLaunchKernelGGL((grouped_inv_prep<T, false>), dim3(counts[1]), dim3(kInvThreads), 0, su, probs,
                       it[1], inv);
  if (counts[2] > 0)
    hipLaunchKernelGGL((grouped_panels<T, true, kGroupTrsmWide, kGroupTrsmWideCols>),
                       dim3(counts[2]), dim3(256), 0, s, probs, it[2], inv);
  if (counts[3] > 0)
    hipLaunchKernelGGL((grouped_panels<T, false, kGroupTrsmWide, kGroupTrsmWideCols>),
                       dim3(counts[3]), dim3(256), 0, su, probs, it[3], inv);
  if (counts[4] > 0)
    hipLaunchKernelGGL((grouped_panels<T, true, kGroupTrsmSmall, PW>), dim3(counts[4]),
                       dim3(256), 0, s, probs, it[4], inv);
  if (counts[5] > 0)
    hipLaunchKernelGGL((grouped_panels<T, false, kGroupTrsmSmall, PW>), dim3(counts[5]),
                       dim3(256), 0, su, probs, it[5], inv);
  if (counts[6] > 0)
    hipLaunchKernelGGL((grouped_panels<T, true, kGroupTrsmMax, PW>), dim3(counts[6]),
                       dim3(256), 0, s, probs, it[6], inv);
  if (counts[7] > 0)
    hipLaunchKernelGGL((grouped_panels<T, false, kGroupTrsmMax, PW>), dim3(counts[7]),
                       dim3(256), 0, su, probs, it[7], inv);
  if (hipGetLastError() != hipSuccess) return MUBLIS_EHIP;
  if (two && (hipEventRecord(join, s2) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess))
    return MUBLIS_EHIP;
  return MUBLIS_OK;
}
// Instantiated for the batched interface's FP32 problems.
template int gfx950_trsm_grouped<float>(const TrsmGroupProb<float> *, const TrsmGroupItem *,
                                        const int *, double *, hipStream_t, hipStream_t,
                                        hipEvent_t, hipEvent_t);

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

// Workspace of the full schedule, which is the one a call with arena
// workspace runs (solve): the 64-row inverses and flags, then, when the
// solve takes GEMM leaves, the 1024-row inverses, the leaf copy of B and the
// doubling's scratch, each 256-byte aligned.  With at least this much
// workspace a workspace call runs exactly the schedule of the plain call and
// gives its bits; with less (at least mublis_gfx950_trsm_ws_bytes(m)) it
// runs the schedule that needs only the 64-row inverses.
static size_t ws_part(size_t bytes) { return (bytes + 255) / 256 * 256; }

template <typename T>
size_t sched_ws_bytes(int m, int n) {
  KnobScope knobs;
  size_t bytes = ws_part(sizeof(double) * (size_t)ws_doubles(m));
  if (use_inv_leaves<T>(m, n))
    bytes += ws_part(sizeof(T) * (size_t)inv_leaf_elems<T>(m)) +
             ws_part(sizeof(T) * (size_t)kInvLeaf<T> * (size_t)n) +
             ws_part(sizeof(T) * (size_t)m * kInvLeaf<T> / 4);
  return bytes;
}

// Carve `ws` as sched_ws_bytes<T>(m, n) lays it out; iv / w / tmp stay null
// when the solve takes no GEMM leaves or ws_bytes is too small for them.
template <typename T>
void carve_ws(void *ws, size_t ws_bytes, int m, int n, double **ad, T **iv, T **w, T **tmp) {
  char *p = (char *)ws;
  *ad = (double *)p;
  *iv = *w = *tmp = nullptr;
  if (!use_inv_leaves<T>(m, n) || ws_bytes < sched_ws_bytes<T>(m, n)) return;
  p += ws_part(sizeof(double) * (size_t)ws_doubles(m));
  *iv = (T *)p;
  p += ws_part(sizeof(T) * (size_t)inv_leaf_elems<T>(m));
  *w = (T *)p;
  p += ws_part(sizeof(T) * (size_t)kInvLeaf<T> * (size_t)n);
  *tmp = (T *)p;
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
  KnobScope knobs;
  double *ad;
  T *iv, *w, *tmp;
  carve_ws<T>(ws, ws_bytes, m, n, &ad, &iv, &w, &tmp);
  return solve_ws<T>(lower != 0, unit ? 1 : 0, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad,
                     iv, w, tmp, s);
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

// FP32 solves against one triangle that keep its inverted diagonal blocks
// in caller workspace between calls (the refined STRSM's three solves).
size_t gfx950_strsm_keep_ws_bytes(int m, int n) { return sched_ws_bytes<float>(m, n); }

int gfx950_strsm_keep(bool lower, bool unit, int m, int n, float alpha, const float *a,
                      int64_t rs_a, int64_t cs_a, float *b, int64_t rs_b, int64_t cs_b,
                      void *ws, size_t ws_bytes, bool *inv_ready, hipStream_t s) {
  if (m < 0 || n < 0 || !ws) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  if (alpha == 0.0f) return gfx950_fill<float>(m, n, 0.0f, b, rs_b, cs_b, s);
  if (ws_bytes < mublis_gfx950_trsm_ws_bytes(m)) return MUBLIS_EINVAL;
  KnobScope knobs;
  double *ad;
  float *iv, *w, *tmp;
  carve_ws<float>(ws, ws_bytes, m, n, &ad, &iv, &w, &tmp);
  return solve_ws<float>(lower, unit ? 1 : 0, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, ad,
                         iv, w, tmp, s, inv_ready);
}

extern "C" size_t mublis_gfx950_dtrsm_sched_ws_bytes(int m, int n) {
  return m <= 0 || n <= 0 ? 0 : sched_ws_bytes<double>(m, n);
}
extern "C" size_t mublis_gfx950_strsm_sched_ws_bytes(int m, int n) {
  return m <= 0 || n <= 0 ? 0 : sched_ws_bytes<float>(m, n);
}

// ------------------------------------------------- allocation-free DTRSM --
// mublis_gfx950_dtrsm_noalloc_dev takes the workspace of the workspace
// entry's schedule (the 64-row diagonal-block inverses and the pipeline's
// flags; no GEMM leaves) from a file-scope device array instead of an
// allocation: kNoAllocMaxM / 64 blocks of BLK doubles (8.5 MB for orders up
// to 16384), held by every GPU the library runs on until the process ends.
// Each solve rewrites what it reads (the inverses, and the flags, which the
// inversion kernel clears), so nothing carries over from the last call; two
// calls that overlap on one GPU share the array and race.
constexpr int kNoAllocMaxM = 16384;
constexpr int64_t kScratchDoubles = (int64_t)(kNoAllocMaxM / NB) * BLK;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
// The array's device address, one per device: in host code its name is a
// host-side stand-in, not a device address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Resolve (once per device, with that device current) the address of this
// device's copy of `scratch`.
double *scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_dev[dev] = (double *)ptr;
  }
  return g_scratch_dev[dev];
}
}  // namespace

extern "C" int mublis_gfx950_dtrsm_noalloc_dev(int lower, int unit, int m, int n,
                                               double alpha, const double *a, int64_t rs_a,
                                               int64_t cs_a, double *b, int64_t rs_b,
                                               int64_t cs_b, void *stream) {
  if (m < 0 || n < 0) return MUBLIS_EINVAL;
  if (m > kNoAllocMaxM) {
    gfx950_set_error("mublis_gfx950_dtrsm_noalloc_dev: order m above 16384, the size of the "
                     "static workspace");
    return MUBLIS_EINVAL;
  }
  // an operand without a unit stride would need a packed copy (an allocation)
  if ((rs_a != 1 && cs_a != 1) || (rs_b != 1 && cs_b != 1)) {
    gfx950_set_error("mublis_gfx950_dtrsm_noalloc_dev: A and B each need a unit stride");
    return MUBLIS_EINVAL;
  }
  if (m == 0 || n == 0) return MUBLIS_OK;
  int st = mublis_rt_init();
  if (st) return st;
  // NULL: the null stream (creating the thread's own stream would allocate)
  hipStream_t s = (hipStream_t)stream;
  if (alpha == 0.0) return gfx950_fill<double>(m, n, 0.0, b, rs_b, cs_b, s);
  double *ws = scratch_address();
  if (!ws) {
    gfx950_set_error("mublis_gfx950_dtrsm_noalloc_dev: no device address for the static "
                     "workspace");
    return MUBLIS_EHIP;
  }
  return solve_ws<double>(lower != 0, unit ? 1 : 0, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                          ws, nullptr, nullptr, nullptr, s);
}

extern "C" int mublis_gfx950_trsm_pipe_spin_limit(int spins) {
  return spins > 0 ? g_pipe_spins_set.exchange(spins, std::memory_order_relaxed)
                   : g_pipe_spins_set.load(std::memory_order_relaxed);
}

extern "C" int mublis_gfx950_trsm_pipe_max_n(int max_n) {
  return max_n >= 0
             ? g_pipe_max_n_set.exchange(std::min(max_n, 4096), std::memory_order_relaxed)
             : g_pipe_max_n_set.load(std::memory_order_relaxed);
}
