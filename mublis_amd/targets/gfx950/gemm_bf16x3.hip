// SGEMM as a 3-term BF16 split product on BF16 MFMA with FP32 accumulation.
//
// Every FP32 operand element is split exactly as x = hi + mid + lo with hi,
// mid and lo BF16 (8 significant bits each, round to nearest even; the
// remainders x - hi and x - hi - mid are exact in FP32; values past the
// largest finite BF16 round hi toward zero: bf16_split.h).  Operands holding
// an Inf or NaN take the native FP32 MFMA GEMM instead (the split kernels
// flag them on the device).  Then
//   A*B ~= Ah*Bh + (Ah*Bm + Am*Bh) + (Ah*Bl + Al*Bh + Am*Bm)
// keeps every partial product down to 2^-16 relative; the three dropped
// terms are below FP32 rounding.  Products of two BF16 values are exact in
// FP32.  Ah*Bh accumulates in its own FP32 accumulator and the five smaller
// products in a second one (2^8 smaller, so its rounding is negligible), so
// the result rounds like a native FP32 GEMM.
//
// Layout: a split pass writes each operand as three K-contiguous BF16 planes
// ([row][ldp], rows padded to the tile, K to 64, zero filled; ldp = K + 64
// so that rows never share L2 sets at one k offset), so the GEMM kernel
// streams aligned 16-byte vectors with no bounds checks.  The GEMM kernel:
// 256 x 128 output tile per workgroup, one workgroup per CU, 8 waves of
// 64 x 64 (two per SIMD), each a 2 x 2 grid of v_mfma_f32_32x32x16_bf16
// tiles (6 MFMAs per tile per k = 16), operands streamed by direct-to-LDS
// loads into a 4-stage LDS ring, the XCD-aware tile order of the FP64/FP32
// kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "bf16_split.h"
#include "gfx950_common.h"
#include "mublis_gfx950.h"

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef unsigned short u16;
typedef u16 u16x8 __attribute__((ext_vector_type(8)));

constexpr int XTM = 256;   // output tile rows
constexpr int XTN = 128;   // output tile columns
constexpr int XBK = 16;    // k per stage (one MFMA k step)
constexpr int kWaves = 8;  // waves per workgroup (4 or 8)

using gfx950::bf16_split3;

// Split an m x k FP32 view (strides rs, cs) into three K-contiguous BF16
// planes of mp x kp.  A 64 x 64 tile goes through LDS so that reads run
// along the source's unit stride and writes along k.
// Any Inf or NaN in the view sets *nonfinite (plain stores of 1; the caller
// clears it first).
__global__ __launch_bounds__(256) void split3_kernel(int m, int k, const float *__restrict__ x,
                                                     int64_t rs, int64_t cs, int mp, int kp,
                                                     int ldp,
                                                     u16 *__restrict__ hi, u16 *__restrict__ mid,
                                                     u16 *__restrict__ lo, int *nonfinite) {
  __shared__ float t[64][65];
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int tid = threadIdx.x;
  const bool rows_contig = rs == 1;
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int lin = it * 256 + tid;
    const int a = lin & 63, b = lin >> 6;  // a runs along the unit stride
    const int r = rows_contig ? a : b, c = rows_contig ? b : a;
    float v = 0.0f;
    if (r0 + r < m && k0 + c < k) v = x[(int64_t)(r0 + r) * rs + (int64_t)(k0 + c) * cs];
    t[r][c] = v;
  }
  __syncthreads();
  // each thread writes 16 consecutive k of one row into each plane
  const int r = tid >> 2, c0 = (tid & 3) * 16;
  if (r0 + r >= mp) return;
  u16x8 vh[2], vm[2], vl[2];
  bool bad = false;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    u16 h, md, l;
    const float v = t[r][c0 + e];
    bad |= (__float_as_uint(v) & 0x7F800000u) == 0x7F800000u;
    bf16_split3(v, h, md, l);
    vh[e / 8][e % 8] = h;
    vm[e / 8][e % 8] = md;
    vl[e / 8][e % 8] = l;
  }
  if (bad) *nonfinite = 1;
  const int64_t off = (int64_t)(r0 + r) * ldp + k0 + c0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (k0 + c0 + 8 * h < kp) {
      *reinterpret_cast<u16x8 *>(hi + off + 8 * h) = vh[h];
      *reinterpret_cast<u16x8 *>(mid + off + 8 * h) = vm[h];
      *reinterpret_cast<u16x8 *>(lo + off + 8 * h) = vl[h];
    }
  }
}

struct XArgs {
  int m, n, kp, ldp, mp, np;
  float alpha, beta;
  const u16 *a[3];  // [mp][ldp], k < kp used
  const u16 *b[3];  // [np][ldp]
  float *c;
  int64_t rs_c, cs_c;
  int tiles_m, tiles_n, ntiles;
};

__device__ inline void map_tile_x(int w, int ntiles, int tiles_m, int tiles_n, int &tm,
                                  int &tn) {
  const int q = ntiles / 8, r = ntiles % 8;
  const int xcd = w % 8, local = w / 8;
  const int id = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
  const int G = 8, per = G * tiles_n, g = id / per, first = g * G;
  const int gs = min(tiles_m - first, G), in = id - g * per;
  tm = first + in % gs;
  tn = in / gs;
}

// The operand stages stream from global memory
// straight into a 4-slot LDS ring (global_load_lds_dwordx4: lane i of a
// wave instruction writes bytes [16 i, 16 i + 16) of a 1 KiB block at the
// M0 base), three stages ahead of the MFMAs and with no registers held for
// them.  A stage is 36 blocks of 32 rows x 16 k (A: 3 planes x 8 blocks, B:
// 3 planes x 4); wave w issues blocks w, w + 4, ...  Inside a block, row r's
// two 16-byte halves are stored swapped when bit 3 of r is set, so the 16
// lanes of a fragment-read pass (16 rows, one half) hit 16 distinct 16-byte
// bank groups.  The four slots are separate LDS arrays and the loop is
// unrolled by four, so every slot index is a compile-time constant.
constexpr int XRING = 4;
constexpr int XSTAGE_U16 = 36 * 512;  // one stage: 36 KiB

template <int SLOT>
__device__ inline u16 *ring_slot(u16 *r0, u16 *r1, u16 *r2, u16 *r3) {
  return SLOT == 0 ? r0 : SLOT == 1 ? r1 : SLOT == 2 ? r2 : r3;
}

// One 16-byte-per-lane global -> LDS transfer to the wave-uniform LDS byte
// address `lds` (lane i lands at lds + 16 i).  Issued from asm so that the
// compiler, which cannot tell which ring slot a DMA writes, does not drain
// every DMA in flight before each LDS read; the kernel orders them itself
// (s_waitcnt vmcnt + barrier).  The instruction has no destination
// registers.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"  // m0 is reserved; the clobber is still honoured
__device__ inline void dma16(const u16 *g, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(g), "s"(lds)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// NW = 4: waves of 128 x 64 (4 x 2 MFMA tiles, two accumulators: 256
// AGPRs, one wave per SIMD).  NW = 8: waves of 64 x 64 (2 x 2 tiles, 128
// AGPRs), two waves per SIMD, so one wave's MFMAs cover the other's LDS
// reads and barrier; waves 0-3 stream the operands either way.
template <int NW>
__global__ __launch_bounds__(64 * NW, 1) void gemm_bf16x3_kernel(XArgs p) {
  constexpr int MW = NW / 2;        // waves along M
  constexpr int WM = XTM / MW;      // rows per wave
  constexpr int TI = WM / 32;       // 32-row MFMA tiles per wave
  __shared__ __attribute__((aligned(1024))) u16 R0[XSTAGE_U16];
  __shared__ __attribute__((aligned(1024))) u16 R1[XSTAGE_U16];
  __shared__ __attribute__((aligned(1024))) u16 R2[XSTAGE_U16];
  __shared__ __attribute__((aligned(1024))) u16 R3[XSTAGE_U16];
  int tm, tn;
  map_tile_x(blockIdx.x, p.ntiles, p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * XTM, n0 = tn * XTN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave % MW) * WM, wn = (wave / MW) * 64;
  const bool streams = wave < 4;

  // This lane's source element offset (u16) for each of the wave's 9 blocks
  // of a stage, at k = 0; block b = wave + 4 j.
  const int lrow = lane >> 1, lhalf = (lane & 1) ^ ((lane >> 4) & 1);
  const u16 *src[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int sw = wave & 3, b = sw + 4 * j;
    if (j < 6) {  // A plane b / 8, rows 32 (b % 8) ..
      src[j] = p.a[b >> 3] + (int64_t)(m0 + 32 * (b & 7) + lrow) * p.ldp + 8 * lhalf;
    } else {      // B plane j - 6, rows 32 wave ..
      src[j] = p.b[j - 6] + (int64_t)(n0 + 32 * sw + lrow) * p.ldp + 8 * lhalf;
    }
  }
  auto issue = [&](u16 *slot, int stage) {
    if (!streams) return;
    const int64_t k0 = (int64_t)stage * XBK;
    const uint32_t base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) u16 *)slot;
#pragma unroll
    for (int j = 0; j < 9; ++j) dma16(src[j] + k0, base + (wave + 4 * j) * 1024);
  };

  f16v big[TI][2], small[TI][2];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      big[i][j] = f16v{};
      small[i][j] = f16v{};
    }

  const int stages = p.kp / XBK;  // a multiple of XRING
  issue(R0, 0);
  issue(R1, min(1, stages - 1));
  issue(R2, min(2, stages - 1));

  // fragment addresses (u16) inside a slot: row r, half h at
  // r * 16 + 8 * (h ^ ((r >> 3) & 1)); B starts after the 24 A blocks
  const int fr = lane & 31, fh = lane >> 5;
  const int fsw = 8 * (fh ^ ((fr >> 3) & 1));
  auto stage_compute = [&](const u16 *slot) {
    bf8 bf[3][2];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bf[s][j] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8 *>(
                                               slot + 24 * 512 + s * 2048 +
                                               (wn + j * 32 + fr) * 16 + fsw));
#pragma unroll
    for (int pa = 0; pa < 3; ++pa) {
      bf8 af[TI];
#pragma unroll
      for (int i = 0; i < TI; ++i)
        af[i] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8 *>(
                                            slot + pa * 4096 + (wm + i * 32 + fr) * 16 + fsw));
#pragma unroll
      for (int pb = 0; pb < 3 - pa; ++pb)
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if (pa == 0 && pb == 0)
              big[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[0][j], big[i][j], 0,
                                                                  0, 0);
            else
              small[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[pb][j],
                                                                    small[i][j], 0, 0, 0);
          }
    }
  };
  // Stage s: wait for this wave's blocks of s (the 18 of s + 1 and s + 2 may
  // still be in flight), barrier (every wave's blocks of s have landed, and
  // every wave is done with slot s - 1), refill slot s - 1 with stage s + 3
  // (clamped: past the end it re-reads the last stage into a slot that is
  // never read again), compute stage s.
#define X3_STEP(U)                                                              \
  {                                                                             \
    asm volatile("s_waitcnt vmcnt(18)" ::: "memory");                           \
    __builtin_amdgcn_s_barrier();                                               \
    issue(ring_slot<(U + 3) % XRING>(R0, R1, R2, R3), min(s4 + U + 3, stages - 1)); \
    stage_compute(ring_slot<U>(R0, R1, R2, R3));                                \
  }
#pragma unroll 1
  for (int s4 = 0; s4 < stages; s4 += XRING) {
    X3_STEP(0)
    X3_STEP(1)
    X3_STEP(2)
    X3_STEP(3)
  }
#undef X3_STEP
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

  const float alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + j * 32 + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
        if (row >= p.m || col >= p.n) continue;
        float *cp = p.c + (int64_t)row * p.rs_c + (int64_t)col * p.cs_c;
        const float prod = alpha * (big[i][j][r] + small[i][j][r]);
        if (beta == 0.0f) *cp = prod;
        else if (beta == 1.0f) *cp = *cp + prod;
        else *cp = prod + beta * *cp;
      }
    }
}

int split3(int rows, int k, const float *x, int64_t rs, int64_t cs, int rp, int kp, int ldp,
           u16 *planes, int *nonfinite, hipStream_t s) {
  const size_t plane = (size_t)rp * ldp;
  dim3 grid((kp + 63) / 64, (rp + 63) / 64);
  hipLaunchKernelGGL(split3_kernel, grid, dim3(256), 0, s, rows, k, x, rs, cs, rp, kp, ldp,
                     planes, planes + plane, planes + 2 * plane, nonfinite);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

}  // namespace

extern "C" int mublis_gfx950_sgemm_bf16x3_dev(int m, int n, int k, float alpha,
                                              const float *a, int64_t rs_a, int64_t cs_a,
                                              const float *b, int64_t rs_b, int64_t cs_b,
                                              float beta, float *c, int64_t rs_c,
                                              int64_t cs_c, void *stream) {
  if (m < 0 || n < 0 || k < 0) return MUBLIS_EINVAL;
  if (m == 0 || n == 0) return MUBLIS_OK;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (alpha == 0.0f || k == 0)
    return mublis_gfx950_sgemm_dev(m, n, 0, 0.0f, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0,
                                   0, nullptr, 0, 0, beta, c, rs_c, cs_c, 0, s);
  const int mp = (m + XTM - 1) / XTM * XTM, np = (n + XTN - 1) / XTN * XTN;
  const int kp = (k + 4 * XBK - 1) / (4 * XBK) * (4 * XBK);  // whole rounds of the ring
  // Plane rows are 128 bytes longer than K: with a power-of-two row stride
  // the rows a workgroup streams at one k offset fall into a few L2 sets and
  // evict each other (at K = 16384 the L2 hit rate halved and HBM traffic
  // grew twentyfold).
  const int ldp = kp + 64;
  const size_t a_bytes = sizeof(u16) * 3 * (size_t)mp * ldp;
  u16 *pa = (u16 *)mublis_rt_alloc(a_bytes + 256);  // + the non-finite flag
  u16 *pb = (u16 *)mublis_rt_alloc(sizeof(u16) * 3 * (size_t)np * ldp);
  int *flag = pa ? (int *)((char *)pa + a_bytes) : nullptr;
  int st = (pa && pb) ? MUBLIS_OK : MUBLIS_ENOMEM;
  if (st == MUBLIS_OK && hipMemsetAsync(flag, 0, sizeof(int), s) != hipSuccess)
    st = MUBLIS_EHIP;
  if (st == MUBLIS_OK) st = split3(m, k, a, rs_a, cs_a, mp, kp, ldp, pa, flag, s);
  // B is split as its transpose: row j of the plane holds B(:, j).
  if (st == MUBLIS_OK) st = split3(n, k, b, cs_b, rs_b, np, kp, ldp, pb, flag, s);
  // An Inf or NaN operand element runs the native FP32 MFMA GEMM instead: its
  // products with the other operand's zero remainder parts would be Inf * 0 =
  // NaN where FP32 gives Inf.  (This call synchronises before it returns
  // anyway, to release the planes, so reading the flag costs one round trip.)
  int nonfinite = 0;
  if (st == MUBLIS_OK &&
      (hipMemcpyAsync(&nonfinite, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = MUBLIS_EHIP;
  if (st == MUBLIS_OK && nonfinite) {
    mublis_rt_free(pa);
    mublis_rt_free(pb);
    st = mublis_gfx950_sgemm_dev(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                                 nullptr, 0, 0, beta, c, rs_c, cs_c, 0, s);
    if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
    return st;
  }
  if (st == MUBLIS_OK) {
    XArgs p;
    p.m = m; p.n = n; p.kp = kp; p.ldp = ldp; p.mp = mp; p.np = np;
    p.alpha = alpha; p.beta = beta;
    const size_t pa_plane = (size_t)mp * ldp, pb_plane = (size_t)np * ldp;
    for (int q = 0; q < 3; ++q) {
      p.a[q] = pa + q * pa_plane;
      p.b[q] = pb + q * pb_plane;
    }
    p.c = c; p.rs_c = rs_c; p.cs_c = cs_c;
    p.tiles_m = mp / XTM; p.tiles_n = np / XTN; p.ntiles = p.tiles_m * p.tiles_n;
    hipLaunchKernelGGL(gemm_bf16x3_kernel<kWaves>, dim3(p.ntiles), dim3(64 * kWaves), 0, s, p);
    if (hipGetLastError() != hipSuccess) st = MUBLIS_EHIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  mublis_rt_free(pa);
  mublis_rt_free(pb);
  return st;
}
