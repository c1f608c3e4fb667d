// Host check of the GEMM kernels' workgroup -> tile maps (gemm_tile.h),
// compiled from the same header the kernels use:
//  * map_tile, full grid and triangle modes 1..3: every workgroup lands on a
//    distinct tile inside the grid, and modes 2 + 3 together cover exactly
//    the 32-tiles touching the ji triangle, each once;
//  * map_tile_lpt, lines along rows or columns, long end high or low: a
//    bijection onto the tiles (padding workgroups excepted), every XCD
//    (workgroup id mod 8) receives the same total K up to one dealt unit,
//    and each XCD meets its tiles in non-increasing K -- after its first
//    round when the first-round snake is on (r1 > 0): then no later tile is
//    longer than a first-round one, and within the first round, dealt over
//    the XCD's 32 CUs breadth-first (its local index l to CU l % 32), every
//    CU gets about the same K, where the plain order fails that bound;
//  * the batched reversal (rev): bijections per problem, and the first
//    round balanced over the CUs where the plain order pairs equal ranks.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <utility>
#include <vector>

#include "gemm_tile.h"

using gfx950::lpt_unit;
using gfx950::map_tile;
using gfx950::map_tile_lpt;

static int failures = 0;
#define CHECK(c, ...)                       \
  do {                                      \
    if (!(c)) {                             \
      if (failures++ < 20) {                \
        std::printf("FAIL %s: ", #c);       \
        std::printf(__VA_ARGS__);           \
        std::printf("\n");                  \
      }                                     \
    }                                       \
  } while (0)

static void full_grid(int tm_, int tn_) {
  std::set<std::pair<int, int>> seen;
  const int nt = tm_ * tn_;
  for (int w = 0; w < nt; ++w) {
    int a, b;
    map_tile(w, nt, tm_, tn_, 0, 0, a, b);
    CHECK(a >= 0 && a < tm_ && b >= 0 && b < tn_, "full %dx%d w %d -> %d,%d", tm_, tn_, w, a, b);
    CHECK(seen.insert({a, b}).second, "full %dx%d duplicate %d,%d", tm_, tn_, a, b);
  }
  CHECK((int)seen.size() == nt, "full %dx%d covers %zu", tm_, tn_, seen.size());
}

static void triangle(int t, int ji) {
  std::set<std::pair<int, int>> seen;
  const int nt = t * (t + 1) / 2;
  for (int w = 0; w < nt; ++w) {
    int a, b;
    map_tile(w, nt, t, t, 1, ji, a, b);
    const bool in = ji == 1 ? b <= a : a <= b;
    CHECK(in && a < t && b < t, "tri %d ji %d w %d -> %d,%d", t, ji, w, a, b);
    CHECK(seen.insert({a, b}).second, "tri %d duplicate", t);
  }
  CHECK((int)seen.size() == nt, "tri %d covers %zu", t, seen.size());
}

// Modes 2 (128-tiles strictly inside) + 3 (diagonal 128-blocks as 32-tiles)
// against the 32-tiles of an n x n triangle.
static void split_triangle(int n, int ji) {
  const int t128 = (n + 127) / 128, t32 = (n + 31) / 32;
  std::vector<int> hits(t32 * t32, 0);
  const int n2 = t128 * (t128 - 1) / 2;
  for (int w = 0; w < n2; ++w) {
    int a, b;
    map_tile(w, n2, t128, t128, 2, ji, a, b);
    for (int i = 4 * a; i < 4 * a + 4 && i < t32; ++i)
      for (int j = 4 * b; j < 4 * b + 4 && j < t32; ++j) hits[i * t32 + j]++;
  }
  const int n3 = 10 * ((t32 + 3) / 4);
  for (int w = 0; w < n3; ++w) {
    int a, b;
    map_tile(w, n3, t32, t32, 3, ji, a, b);
    if (a >= t32 || b >= t32) continue;  // the kernel's bounds exit
    CHECK(a / 4 == b / 4, "diag n %d -> %d,%d off the diagonal block", n, a, b);
    hits[a * t32 + b]++;
  }
  for (int i = 0; i < t32; ++i)
    for (int j = 0; j < t32; ++j) {
      const bool want = ji == 1 ? j <= i : i <= j;
      CHECK(hits[i * t32 + j] == (want ? 1 : 0), "split n %d ji %d tile %d,%d hit %d", n, ji,
            i, j, hits[i * t32 + j]);
    }
}

static void lpt(int tm_, int tn_, int mode, int r1 = 0) {
  const bool rows = mode <= 2;
  const int lines = rows ? tm_ : tn_, per = rows ? tn_ : tm_;
  const int g = lpt_unit(lines, per);
  const long units = (long)lines * ((per + g - 1) / g);
  const int wgs = (int)(8 * ((units + 7) / 8) * g);
  // K of a tile: its line's rank from the long end, plus one
  auto kof = [&](int a, int b) {
    const int line = rows ? a : b;
    return (mode == 1 || mode == 3) ? line + 1 : lines - line;
  };
  std::set<std::pair<int, int>> seen;
  long sum[8] = {0};
  int last[8], first_min[8];
  // first-round K per (XCD, CU), this walk and the plain one (r1 = 0)
  std::vector<long> cu(8 * 32, 0), cu_plain(8 * 32, 0);
  for (int x = 0; x < 8; ++x) last[x] = first_min[x] = 1 << 30;
  for (int w = 0; w < wgs; ++w) {
    int a, b;
    const int x = w % 8, l = w / 8;
    int pa, pb;
    if (l < r1 && map_tile_lpt(w, tm_, tn_, mode, g, pa, pb))
      cu_plain[x * 32 + l % 32] += kof(pa, pb);
    if (!map_tile_lpt(w, tm_, tn_, mode, g, a, b, r1)) continue;
    CHECK(a >= 0 && a < tm_ && b >= 0 && b < tn_, "lpt %dx%d m%d w %d -> %d,%d", tm_, tn_,
          mode, w, a, b);
    CHECK(seen.insert({a, b}).second, "lpt %dx%d m%d duplicate", tm_, tn_, mode);
    const int k = kof(a, b);
    if (l < r1) {
      cu[x * 32 + l % 32] += k;
      first_min[x] = k < first_min[x] ? k : first_min[x];
    } else {
      CHECK(k <= last[x] && k <= first_min[x], "lpt %dx%d m%d XCD %d K rises %d -> %d", tm_,
            tn_, mode, x, last[x] < first_min[x] ? last[x] : first_min[x], k);
      last[x] = k;
    }
    sum[x] += k;
  }
  if (r1 > 0) {
    // Checked where an XCD's first round is at least two whole rows of 32.
    // Row q of the plain order hands CU c the (32q + c)-th longest tile; the
    // snake pairs row q's c-th with row q+1's (31 - c)-th, which cancels K's
    // decrease along the rows up to rounding (2 per pair); an unpaired last
    // row adds its own K range.
    const int nx = wgs / 8, nr = r1 < nx ? r1 : nx;
    if (nr >= 64 && nr % 32 == 0) {
      const int nrow = nr / 32;
      for (int x = 0; x < 8; ++x) {
        long range_last = 0;
        {
          int pa, pb, kf = 0, kl = 0;
          const int q = nrow - 1;
          if (map_tile_lpt(x + 8 * (32 * q), tm_, tn_, mode, g, pa, pb)) kf = kof(pa, pb);
          if (map_tile_lpt(x + 8 * (32 * q + 31), tm_, tn_, mode, g, pa, pb)) kl = kof(pa, pb);
          range_last = kf - kl;
        }
        // the largest K step between neighbouring first-round workgroups
        long step = 1;
        for (int l = 0; l + 1 < nr; ++l) {
          int pa, pb, qa, qb;
          if (!map_tile_lpt(x + 8 * l, tm_, tn_, mode, g, pa, pb) ||
              !map_tile_lpt(x + 8 * (l + 1), tm_, tn_, mode, g, qa, qb))
            continue;
          const long d = (long)kof(pa, pb) - kof(qa, qb);
          step = d > step ? d : step;
        }
        const long bound = (2L * (nrow / 2) + 2) * step + ((nrow & 1) ? range_last : 0);
        long lo = 1L << 40, hi = -1, plo = 1L << 40, phi = -1;
        for (int c = 0; c < 32; ++c) {
          const long v = cu[x * 32 + c], pv = cu_plain[x * 32 + c];
          lo = v < lo ? v : lo;
          hi = v > hi ? v : hi;
          plo = pv < plo ? pv : plo;
          phi = pv > phi ? pv : phi;
        }
        CHECK(hi - lo <= bound, "lpt %dx%d m%d r1 %d XCD %d first-round CU K spread %ld > %ld",
              tm_, tn_, mode, r1, x, hi - lo, bound);
        if (tm_ == 16 && tn_ == 64 && mode <= 2 && r1 == 128) {
          // the TRSM leaf grid: the plain order fails the bound by far
          CHECK(phi - plo > bound && 2 * (hi - lo) < phi - plo,
                "lpt 16x64 XCD %d spread %ld, plain %ld, bound %ld", x, hi - lo, phi - plo,
                bound);
          if (x == 0 && mode == 1)
            std::printf("16x64 first round: CU K spread %ld (plain order %ld, bound %ld)\n",
                        hi - lo, phi - plo, bound);
        }
      }
    }
  }
  CHECK((int)seen.size() == tm_ * tn_, "lpt %dx%d m%d covers %zu", tm_, tn_, mode, seen.size());
  long lo = sum[0], hi = sum[0];
  for (int x = 1; x < 8; ++x) {
    lo = sum[x] < lo ? sum[x] : lo;
    hi = sum[x] > hi ? sum[x] : hi;
  }
  // Snake dealing: two consecutive rounds give every XCD a pair of units
  // whose K ranks sum to the same value, so with K linear in the rank the
  // first-order differences cancel; round-robin dealing would leave XCD 0
  // ahead by ~7/8 of the longest unit's K.  Few units: at most one unit.
  const long unit_k = (long)lines * g;
  const long bound = units >= 64 ? unit_k / 4 + 8L * g : unit_k;
  CHECK(hi - lo <= bound, "lpt %dx%d m%d XCD K spread %ld > %ld", tm_, tn_, mode, hi - lo,
        bound);
}

// A batched launch of nprob problems, each a tm_ x tn_ long-first grid of
// P <= 32 workgroups per XCD: with the problems of odd rows of 32 reversed
// (rev), every problem's walk is still a bijection, and over the first round
// the CUs' K (breadth-first: problem z's XCD-local index l at CU (z P + l) %
// 32) is balanced where the plain order pairs equal ranks.
static void batched_rev(int tm_, int tn_, int mode, int nprob) {
  const bool rows = mode <= 2;
  const int lines = rows ? tm_ : tn_, per = rows ? tn_ : tm_;
  const int g = lpt_unit(lines, per);
  const long units = (long)lines * ((per + g - 1) / g);
  const int wgs = (int)(8 * ((units + 7) / 8) * g), P = wgs / 8;
  if (P > 32 || 32 % P) return;
  auto kof = [&](int a, int b) {
    const int line = rows ? a : b;
    return (mode == 1 || mode == 3) ? line + 1 : lines - line;
  };
  std::vector<long> cu(8 * 32, 0), cu_plain(8 * 32, 0);
  for (int z = 0; z < nprob; ++z) {
    const bool rev = (z * P / 32) % 2 == 1;
    std::set<std::pair<int, int>> seen;
    for (int w = 0; w < wgs; ++w) {
      int a, b, pa, pb;
      const int x = w % 8, c = (z * P + w / 8) % 32;
      if (map_tile_lpt(w, tm_, tn_, mode, g, pa, pb)) cu_plain[x * 32 + c] += kof(pa, pb);
      if (!map_tile_lpt(w, tm_, tn_, mode, g, a, b, 0, rev)) continue;
      CHECK(seen.insert({a, b}).second, "batched %dx%d m%d z %d duplicate", tm_, tn_, mode, z);
      cu[x * 32 + c] += kof(a, b);
    }
    CHECK((int)seen.size() == tm_ * tn_, "batched %dx%d m%d z %d covers %zu", tm_, tn_, mode, z,
          seen.size());
  }
  for (int x = 0; x < 8; ++x) {
    long lo = 1L << 40, hi = -1, plo = 1L << 40, phi = -1;
    for (int c = 0; c < 32; ++c) {
      lo = cu[x * 32 + c] < lo ? cu[x * 32 + c] : lo;
      hi = cu[x * 32 + c] > hi ? cu[x * 32 + c] : hi;
      plo = cu_plain[x * 32 + c] < plo ? cu_plain[x * 32 + c] : plo;
      phi = cu_plain[x * 32 + c] > phi ? cu_plain[x * 32 + c] : phi;
    }
    if (tm_ == 8 && tn_ == 8 && mode == 1 && nprob == 8) {
      // the doubling's 512-level: CU K 16 vs 2 in plain order
      CHECK(hi - lo <= 2 && phi - plo >= 8, "batched 8x8 XCD %d spread %ld, plain %ld", x,
            hi - lo, phi - plo);
      if (x == 0) std::printf("8 x (8x8) batched: CU K spread %ld (plain %ld)\n", hi - lo, phi - plo);
    }
  }
}

int main() {
  for (int a = 1; a <= 40; ++a)
    for (int b : {1, 2, 3, 7, 8, 9, 16, 33}) full_grid(a, b);
  for (int t = 1; t <= 70; ++t) {
    triangle(t, 1);
    triangle(t, 2);
  }
  for (int n : {1, 63, 64, 65, 127, 128, 129, 200, 255, 256, 1000, 4100, 8192})
    for (int ji : {1, 2}) split_triangle(n, ji);
  for (int a : {1, 2, 5, 8, 9, 16, 31, 32, 33, 64, 128, 129})
    for (int b : {1, 3, 8, 16, 17, 64})
      for (int m = 1; m <= 4; ++m)
        for (int r1 : {0, 64, 128}) lpt(a, b, m, r1);  // 16 x 64, r1 128: the TRSM leaves
  for (int t : {1, 2, 4, 8, 16})
    for (int m = 1; m <= 4; ++m)
      for (int np : {2, 8, 16, 64}) batched_rev(t, t, m, np);  // 8 x 8, 8 problems: doubling
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
