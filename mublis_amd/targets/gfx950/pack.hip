// Operand materialisation, strided copies and fills for the gfx950 layer.
//
// The reference packs every MC x KC / KC x NC block into micro-panels and
// applies the operand's structure while packing (frame/l1m/packm.c:18-84).
// On the GPU the GEMM kernel stages plain strided tiles itself, so packing
// survives only where structure must be applied: a SYMMETRIC or TRIANGULAR
// operand is materialised once, densely, and then multiplied as GENERAL.
// Values outside the stored triangle and a unit diagonal are never loaded,
// so NaN in unreferenced storage cannot reach the product.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gfx950_common.h"

namespace {

inline int grid_for(int64_t total) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 1 << 16));
}

// packm's stored-triangle test (frame/l1m/packm.c:5-16).
__device__ inline bool stored(int uplo, int64_t delta) {
  return uplo == 1 ? delta >= 0 : (uplo == 2 ? delta <= 0 : true);
}

// `mask` (0 all, 1 lower, 2 upper) keeps only the elements on its side of
// the same diagonal and writes zero elsewhere without reading: the driver's
// jp / pi read domains applied per element.
template <typename T, typename D = T>
__global__ void pack_kernel(int m, int n, const T *__restrict__ src, int64_t rs,
                            int64_t cs, int struc, int uplo, int diag, int diagoff,
                            int mask, D *__restrict__ dst, int64_t ld) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % m, j = t / m;
    const int64_t d = diagoff + i - j;
    D v;
    if (!stored(mask, d)) {
      v = D(0);
    } else if (struc == 1) {  // SYMMETRIC: reflect when outside the stored triangle
      v = (D)(stored(uplo, d) ? src[i * rs + j * cs]
                              : src[(j - diagoff) * rs + (i + diagoff) * cs]);
    } else if (struc == 2) {  // TRIANGULAR
      if (!stored(uplo, d)) v = D(0);
      else if (d == 0 && diag == 1) v = D(1);
      else {
        v = (D)src[i * rs + j * cs];
        if (d == 0 && diag == 2) v = D(1) / v;
      }
    } else {
      v = (D)src[i * rs + j * cs];
    }
    dst[i + j * ld] = v;
  }
}

template <typename T, typename D = T>
__global__ void copy_kernel(int m, int n, const T *__restrict__ src, int64_t rs_s,
                            int64_t cs_s, D *__restrict__ dst, int64_t rs_d,
                            int64_t cs_d) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % m, j = t / m;
    dst[i * rs_d + j * cs_d] = (D)src[i * rs_s + j * cs_s];
  }
}

// Refinement update: x64 += d (dense m x n), then out := (float) x64.
__global__ void refine_update_kernel(int m, int n, const float *__restrict__ d,
                                     double *__restrict__ x64, float *__restrict__ out,
                                     int64_t rs_o, int64_t cs_o) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % m, j = t / m;
    const double v = x64[t] + (double)d[t];
    x64[t] = v;
    out[i * rs_o + j * cs_o] = (float)v;
  }
}

template <typename T>
__global__ void fill_kernel(int m, int n, T value, T *x, int64_t rs, int64_t cs) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % m, j = t / m;
    x[i * rs + j * cs] = value;
  }
}

}  // namespace

template <typename T>
int gfx950_copy(int m, int n, const T *src, int64_t rs_s, int64_t cs_s, T *dst,
                int64_t rs_d, int64_t cs_d, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL(copy_kernel<T>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     m, n, src, rs_s, cs_s, dst, rs_d, cs_d);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

template <typename T>
int gfx950_pack(int m, int n, const T *src, int64_t rs, int64_t cs, int struc,
                int uplo, int diag, int diagoff, int mask, T *dst, int64_t ld,
                hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  if (struc < 0 || struc > 2 || uplo < 0 || uplo > 2 || diag < 0 || diag > 2 || mask < 0 ||
      mask > 2)
    return MUBLIS_EINVAL;
  hipLaunchKernelGGL(pack_kernel<T>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     m, n, src, rs, cs, struc, uplo, diag, diagoff, mask, dst, ld);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

template <typename T>
int gfx950_fill(int m, int n, T value, T *x, int64_t rs, int64_t cs, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     m, n, value, x, rs, cs);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

int gfx950_convert_f2d(int m, int n, const float *src, int64_t rs_s, int64_t cs_s,
                       double *dst, int64_t rs_d, int64_t cs_d, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL((copy_kernel<float, double>), dim3(grid_for((int64_t)m * n)),
                     dim3(256), 0, s, m, n, src, rs_s, cs_s, dst, rs_d, cs_d);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}
int gfx950_convert_d2f(int m, int n, const double *src, int64_t rs_s, int64_t cs_s,
                       float *dst, int64_t rs_d, int64_t cs_d, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL((copy_kernel<double, float>), dim3(grid_for((int64_t)m * n)),
                     dim3(256), 0, s, m, n, src, rs_s, cs_s, dst, rs_d, cs_d);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}
int gfx950_pack_f2d(int m, int n, const float *src, int64_t rs, int64_t cs, int struc,
                    int uplo, int diag, double *dst, int64_t ld, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL((pack_kernel<float, double>), dim3(grid_for((int64_t)m * n)),
                     dim3(256), 0, s, m, n, src, rs, cs, struc, uplo, diag, 0, 0, dst, ld);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}
int gfx950_refine_update(int m, int n, const float *d, double *x64, float *out,
                         int64_t rs_o, int64_t cs_o, hipStream_t s) {
  if (m <= 0 || n <= 0) return MUBLIS_OK;
  hipLaunchKernelGGL(refine_update_kernel, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     m, n, d, x64, out, rs_o, cs_o);
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

template int gfx950_copy<float>(int, int, const float *, int64_t, int64_t, float *,
                                int64_t, int64_t, hipStream_t);
template int gfx950_copy<double>(int, int, const double *, int64_t, int64_t, double *,
                                 int64_t, int64_t, hipStream_t);
template int gfx950_pack<float>(int, int, const float *, int64_t, int64_t, int, int,
                                int, int, int, float *, int64_t, hipStream_t);
template int gfx950_pack<double>(int, int, const double *, int64_t, int64_t, int, int,
                                 int, int, int, double *, int64_t, hipStream_t);
template int gfx950_fill<float>(int, int, float, float *, int64_t, int64_t,
                                hipStream_t);
template int gfx950_fill<double>(int, int, double, double *, int64_t, int64_t,
                                 hipStream_t);

static hipStream_t pick(void *stream) {
  return (hipStream_t)(stream ? stream : mublis_rt_stream());
}

extern "C" {
int mublis_gfx950_dpack_dev(int m, int n, const double *src, int64_t rs, int64_t cs,
                            int struc, int uplo, int diag, int diagoff, int mask,
                            double *dst, int64_t ld, void *stream) {
  return gfx950_pack<double>(m, n, src, rs, cs, struc, uplo, diag, diagoff, mask, dst, ld,
                             pick(stream));
}
int mublis_gfx950_spack_dev(int m, int n, const float *src, int64_t rs, int64_t cs,
                            int struc, int uplo, int diag, int diagoff, int mask,
                            float *dst, int64_t ld, void *stream) {
  return gfx950_pack<float>(m, n, src, rs, cs, struc, uplo, diag, diagoff, mask, dst, ld,
                            pick(stream));
}
int mublis_gfx950_dcopy_dev(int m, int n, const double *src, int64_t rs_s,
                            int64_t cs_s, double *dst, int64_t rs_d, int64_t cs_d,
                            void *stream) {
  return gfx950_copy<double>(m, n, src, rs_s, cs_s, dst, rs_d, cs_d, pick(stream));
}
int mublis_gfx950_scopy_dev(int m, int n, const float *src, int64_t rs_s,
                            int64_t cs_s, float *dst, int64_t rs_d, int64_t cs_d,
                            void *stream) {
  return gfx950_copy<float>(m, n, src, rs_s, cs_s, dst, rs_d, cs_d, pick(stream));
}
int mublis_gfx950_dfill_dev(int m, int n, double value, double *x, int64_t rs,
                            int64_t cs, void *stream) {
  return gfx950_fill<double>(m, n, value, x, rs, cs, pick(stream));
}
int mublis_gfx950_sfill_dev(int m, int n, float value, float *x, int64_t rs,
                            int64_t cs, void *stream) {
  return gfx950_fill<float>(m, n, value, x, rs, cs, pick(stream));
}
}
