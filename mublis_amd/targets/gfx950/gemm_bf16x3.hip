// SGEMM as a 3-term BF16 split product on BF16 MFMA with FP32 accumulation.
//
// Every FP32 operand element is split exactly as x = hi + mid + lo with hi,
// mid and lo BF16 (8 significant bits each, round to nearest even; the
// remainders x - hi and x - hi - mid are exact in FP32).  Then
//   A*B ~= Ah*Bh + (Ah*Bm + Am*Bh) + (Ah*Bl + Al*Bh + Am*Bm)
// keeps every partial product down to 2^-16 relative; the three dropped
// terms are below FP32 rounding.  Products of two BF16 values are exact in
// FP32.  Ah*Bh accumulates in its own FP32 accumulator and the five smaller
// products in a second one (2^8 smaller, so its rounding is negligible), so
// the result rounds like a native FP32 GEMM.
//
// Layout: a split pass writes each operand as three K-contiguous BF16 planes
// ([row][kp], rows padded to the tile, K to 32, zero filled), so the GEMM
// kernel streams aligned 16-byte vectors with no bounds checks.  The GEMM
// kernel: 256 x 128 output tile per workgroup, one workgroup per CU, 4 waves
// of 128 x 64, each a 4 x 2 grid of v_mfma_f32_32x32x16_bf16 tiles (6 MFMAs
// per tile per k = 16: 48 MFMAs for 18 LDS fragment reads per lane), LDS
// double-buffered, the XCD-aware tile order of the FP64/FP32 kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

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
constexpr int XLD = 24;    // LDS row stride in BF16 (48 B: conflict-free b128 reads)
constexpr int XNT = 256;

__device__ inline u16 bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  const uint32_t bias = 0x7FFFu + ((u >> 16) & 1u);
  return (u16)((u + bias) >> 16);
}
__device__ inline float bf16_to_f(u16 h) { return __uint_as_float((uint32_t)h << 16); }

// Split an m x k FP32 view (strides rs, cs) into three K-contiguous BF16
// planes of mp x kp.  A 64 x 64 tile goes through LDS so that reads run
// along the source's unit stride and writes along k.
__global__ __launch_bounds__(256) void split3_kernel(int m, int k, const float *__restrict__ x,
                                                     int64_t rs, int64_t cs, int mp, int kp,
                                                     int ldp,
                                                     u16 *__restrict__ hi, u16 *__restrict__ mid,
                                                     u16 *__restrict__ lo) {
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
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const float v = t[r][c0 + e];
    const u16 h = bf16_rne(v);
    const float r1 = v - bf16_to_f(h);
    const u16 md = bf16_rne(r1);
    const float r2 = r1 - bf16_to_f(md);
    const u16 l = bf16_rne(r2);
    vh[e / 8][e % 8] = h;
    vm[e / 8][e % 8] = md;
    vl[e / 8][e % 8] = l;
  }
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

__global__ __launch_bounds__(XNT, 1) void gemm_bf16x3_kernel(XArgs p) {
  __shared__ __attribute__((aligned(16))) u16 As[2][3][XTM * XLD];
  __shared__ __attribute__((aligned(16))) u16 Bs[2][3][XTN * XLD];
  int tm, tn;
  map_tile_x(blockIdx.x, p.ntiles, p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * XTM, n0 = tn * XTN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 128, wn = (wave >> 1) * 64;
  // loaders: A row tid (both 8-element halves of the 16-k stage, 3 planes);
  // B row tid / 2, half tid % 2 (3 planes)
  const int brow = tid >> 1, bhalf = (tid & 1) * 8;

  f16v big[4][2], small[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      big[i][j] = f16v{};
      small[i][j] = f16v{};
    }

  // The loads of stage st + 1 are issued while stage st is computed and
  // land in the other LDS slot at the end of the iteration.  (One register
  // set: with the 256 accumulators in AGPRs the architectural VGPRs are the
  // scarce resource, and fragments read early hide LDS latency better than
  // a second prefetch set.)
  u16x8 ra[3][2], rb[3];
  const int64_t aoff = (int64_t)(m0 + tid) * p.ldp;
  const int64_t boff = (int64_t)(n0 + brow) * p.ldp + bhalf;
  const int stages = p.kp / XBK;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    ra[s][0] = *reinterpret_cast<const u16x8 *>(p.a[s] + aoff);
    ra[s][1] = *reinterpret_cast<const u16x8 *>(p.a[s] + aoff + 8);
    rb[s] = *reinterpret_cast<const u16x8 *>(p.b[s] + boff);
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    *reinterpret_cast<u16x8 *>(&As[0][s][tid * XLD]) = ra[s][0];
    *reinterpret_cast<u16x8 *>(&As[0][s][tid * XLD + 8]) = ra[s][1];
    *reinterpret_cast<u16x8 *>(&Bs[0][s][brow * XLD + bhalf]) = rb[s];
  }
  __syncthreads();

  const int fr = lane & 31, fk = (lane >> 5) * 8;
  // kp is a multiple of 2 * XBK, so stages come in pairs (the LDS slot is a
  // compile-time constant) and no branch interrupts the pipeline (branches
  // make the compiler's wait counts conservative).  Loads are issued
  // unconditionally: past the end they re-read the last stage, in bounds,
  // into a slot that is never read again.
#pragma unroll 1
  for (int st = 0; st < stages; st += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int cur = h;
      {
        const int64_t k0 = (int64_t)min(st + h + 1, stages - 1) * XBK;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          ra[s][0] = *reinterpret_cast<const u16x8 *>(p.a[s] + aoff + k0);
          ra[s][1] = *reinterpret_cast<const u16x8 *>(p.a[s] + aoff + k0 + 8);
          rb[s] = *reinterpret_cast<const u16x8 *>(p.b[s] + boff + k0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      bf8 bf[3][2];
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          bf[s][j] = __builtin_bit_cast(
              bf8, *reinterpret_cast<const u16x8 *>(&Bs[cur][s][(wn + j * 32 + fr) * XLD + fk]));
      // Products grouped by A plane, so one plane of A fragments is live at
      // a time: hi x (hi -> big; mid, lo -> small), mid x (hi, mid), lo x hi.
      // Consecutive MFMAs update independent accumulators.
#pragma unroll
      for (int pa = 0; pa < 3; ++pa) {
        bf8 af[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          af[i] = __builtin_bit_cast(
              bf8, *reinterpret_cast<const u16x8 *>(&As[cur][pa][(wm + i * 32 + fr) * XLD + fk]));
#pragma unroll
        for (int pb = 0; pb < 3 - pa; ++pb)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              if (pa == 0 && pb == 0)
                big[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[0][j], big[i][j],
                                                                    0, 0, 0);
              else
                small[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[pb][j],
                                                                      small[i][j], 0, 0, 0);
            }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        *reinterpret_cast<u16x8 *>(&As[cur ^ 1][s][tid * XLD]) = ra[s][0];
        *reinterpret_cast<u16x8 *>(&As[cur ^ 1][s][tid * XLD + 8]) = ra[s][1];
        *reinterpret_cast<u16x8 *>(&Bs[cur ^ 1][s][brow * XLD + bhalf]) = rb[s];
      }
      __syncthreads();
    }
  }

  // 32 x 32 MFMA output: lane l, register r holds column l%32 and row
  // 8*(r/4) + 4*(l/32) + r%4 of its tile (measured on gfx950).
  const float alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < 4; ++i)
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
           u16 *planes, hipStream_t s) {
  const size_t plane = (size_t)rp * ldp;
  dim3 grid((kp + 63) / 64, (rp + 63) / 64);
  hipLaunchKernelGGL(split3_kernel, grid, dim3(256), 0, s, rows, k, x, rs, cs, rp, kp, ldp,
                     planes, planes + plane, planes + 2 * plane);
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
  const int kp = (k + 2 * XBK - 1) / (2 * XBK) * (2 * XBK);  // whole stage pairs
  // Plane rows are 128 bytes longer than K: with a power-of-two row stride
  // the rows a workgroup streams at one k offset fall into a few L2 sets and
  // evict each other (at K = 16384 the L2 hit rate halved and HBM traffic
  // grew twentyfold).
  const int ldp = kp + 64;
  u16 *pa = (u16 *)mublis_rt_alloc(sizeof(u16) * 3 * (size_t)mp * ldp);
  u16 *pb = (u16 *)mublis_rt_alloc(sizeof(u16) * 3 * (size_t)np * ldp);
  int st = (pa && pb) ? MUBLIS_OK : MUBLIS_ENOMEM;
  if (st == MUBLIS_OK) st = split3(m, k, a, rs_a, cs_a, mp, kp, ldp, pa, s);
  // B is split as its transpose: row j of the plane holds B(:, j).
  if (st == MUBLIS_OK) st = split3(n, k, b, cs_b, rs_b, np, kp, ldp, pb, s);
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
    hipLaunchKernelGGL(gemm_bf16x3_kernel, dim3(p.ntiles), dim3(XNT), 0, s, p);
    if (hipGetLastError() != hipSuccess) st = MUBLIS_EHIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  mublis_rt_free(pa);
  mublis_rt_free(pb);
  return st;
}
