// Host check of the BF16x3 split (mublis_amd/targets/gfx950/bf16_split.h),
// compiled from the header the kernel uses:
//  * finite x: hi + mid + lo == x exactly (summed in double) for |x| >=
//    2^-110 and within 2^-134 below, each part a
//    finite BF16, |hi| <= |x| when hi is rounded toward zero, and hi is x
//    rounded to nearest BF16 whenever that is finite;
//  * +-Inf: hi = x, mid = lo = 0;  NaN: hi NaN, mid = lo = 0.
// Covered: every FP32 bit pattern with a stride of 61 (about 70 million),
// every pattern of the top binade (the overflow-prone range near FLT_MAX,
// 2 x 2^23) and of the lowest two (subnormals and the smallest normals).
// A planted split without the overflow guard must fail the same check.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bf16_split.h"

using namespace gfx950;

static long failures = 0, checked = 0;

static bool bf16_finite(uint16_t h) { return (h & 0x7F80u) != 0x7F80u; }

template <typename Split>
static bool check_one(uint32_t u, Split split) {
  const float x = u2f(u);
  uint16_t h, m, l;
  split(x, h, m, l);
  if (std::isnan(x)) return std::isnan(bf16_to_f(h)) && m == 0 && l == 0;
  if (std::isinf(x)) return bf16_to_f(h) == x && m == 0 && l == 0;
  if (!bf16_finite(h) || !bf16_finite(m) || !bf16_finite(l)) return false;
  const double sum = (double)bf16_to_f(h) + (double)bf16_to_f(m) + (double)bf16_to_f(l);
  if (std::fabs((double)x) >= std::ldexp(1.0, -110) ? sum != (double)x
                                                     : std::fabs(sum - x) > std::ldexp(1.0, -134))
    return false;
  const uint16_t rne = bf16_rne(x);
  if (bf16_finite(rne) && h != rne) return false;
  if (!bf16_finite(rne) && std::fabs(bf16_to_f(h)) > std::fabs(x)) return false;
  return true;
}

// The split without the guards: plain round-to-nearest for every part.
static void unguarded(float x, uint16_t &h, uint16_t &m, uint16_t &l) {
  h = bf16_rne(x);
  const float r1 = x - bf16_to_f(h);
  m = bf16_rne(r1);
  l = bf16_rne(r1 - bf16_to_f(m));
}

int main() {
  auto run = [](uint32_t u) {
    ++checked;
    if (!check_one(u, bf16_split3) && failures++ < 10)
      std::printf("FAIL 0x%08x (%g)\n", u, (double)u2f(u));
  };
  for (uint64_t u = 0; u <= 0xFFFFFFFFull; u += 61) run((uint32_t)u);
  for (uint32_t sgn = 0; sgn < 2; ++sgn)
    for (uint32_t mant = 0; mant < (1u << 23); ++mant) {
      run((sgn << 31) | (0xFEu << 23) | mant);  // top finite binade
      run((sgn << 31) | (0x00u << 23) | mant);  // subnormals
      run((sgn << 31) | (0x01u << 23) | mant);  // smallest normals
    }
  const uint32_t specials[] = {0x7F800000u, 0xFF800000u, 0x7FC00000u, 0x7F800001u,
                               0xFFFFFFFFu, 0x7F7FFFFFu, 0xFF7FFFFFu, 0x7F7F8000u,
                               0x7F7F7FFFu, 0x00000001u, 0x80000000u, 0x00000000u};
  for (uint32_t u : specials) run(u);
  // the planted unguarded split fails on the top binade and on Inf
  long planted = 0;
  for (uint32_t u : specials) planted += !check_one(u, unguarded);
  if (planted < 3) {
    std::printf("FAIL planted split passed (%ld specials failed)\n", planted);
    ++failures;
  }
  std::printf("checked %ld values, %ld failures\n", checked, failures);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
