// Exact three-term BF16 split of an FP32 value, x = hi + mid + lo, for the
// BF16x3 SGEMM (gemm_bf16x3.hip).  Host and device: the host build checks
// the split (tests/native/split_check.cpp).
//
//  * finite x: hi = x rounded to BF16 (nearest, ties to even); mid = the
//    remainder x - hi rounded the same way; lo = what is left, which has at
//    most 8 significant bits and so is exact.  Every remainder is exact in
//    FP32 (hi and mid are the leading bits of what they are subtracted
//    from).  The split is exact for |x| >= 2^-110; below that lo would need
//    bits under the smallest BF16 subnormal (2^-133), which are rounded
//    away (error at most 2^-134).  When rounding x to nearest would overflow BF16 (|x| above the
//    largest finite BF16, 0x7F7F = 3.3895e38) hi is rounded toward zero
//    instead, so hi stays finite and the split stays exact.
//  * x = +-Inf: hi = x, mid = lo = 0, so products give what FP32 gives
//    (+-Inf, or NaN for Inf * 0), never Inf - Inf from the remainders.
//  * x = NaN: hi = a quiet NaN, mid = lo = 0.
#ifndef MUBLIS_GFX950_BF16_SPLIT_H
#define MUBLIS_GFX950_BF16_SPLIT_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfx950 {

__host__ __device__ inline uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
__host__ __device__ inline float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline uint16_t bf16_rne(float x) {
  const uint32_t u = f2u(x);
  const uint32_t bias = 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)((u + bias) >> 16);
}
__host__ __device__ inline float bf16_to_f(uint16_t h) { return u2f((uint32_t)h << 16); }

__host__ __device__ inline void bf16_split3(float x, uint16_t &hi, uint16_t &mid,
                                            uint16_t &lo) {
  const uint32_t u = f2u(x);
  if ((u & 0x7F800000u) == 0x7F800000u) {  // Inf or NaN
    hi = (uint16_t)(u >> 16);
    if (u & 0x007FFFFFu) hi |= 0x0040u;     // NaN stays NaN (quiet)
    mid = lo = 0;
    return;
  }
  hi = bf16_rne(x);
  if ((hi & 0x7F80u) == 0x7F80u) hi = (uint16_t)(u >> 16);  // would overflow: toward zero
  const float r1 = x - bf16_to_f(hi);
  mid = bf16_rne(r1);
  const float r2 = r1 - bf16_to_f(mid);
  lo = bf16_rne(r2);
}

}  // namespace gfx950

#endif
