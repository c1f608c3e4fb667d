// Grouped persistent batched L3 kernel for gfx950 (FP32 GEMM, SYRK, SYR2K
// and TRSM).
//
// Thousands of uneven problems run as ONE launch: the host expands every
// problem into output tiles (SYRK: only the tiles that touch its ji
// triangle), tags each with a tile shape chosen by the autotuner's table and
// the problem's operand layout, and sorts the work list by estimated cost
// (largest first).  Persistent grids walk the list cyclically; each item
// runs the same per-tile MFMA routine as the one-problem kernel
// (gemm_tile.h), so a problem's result is bit-identical
// to what cblas_sgemm / cblas_ssyrk compute for it.  The 128-tile items and
// the smaller ones are two launches (different LDS and occupancy), captured
// once in a hipGraph and replayed.  TRSM problems run as independent
// column panels after one launch that inverts all their diagonal blocks
// (trsm.hip, gfx950_trsm_grouped), in the same graph.
//
// The reference has no batched interface; this is the batched addition on
// top of its per-call driver (frame/l3/l3_driver.c:49-249).
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gemm_tile.h"
#include "gfx950_common.h"
#include "mublis_gfx950.h"

using namespace gfx950;

namespace {

struct GroupItem {
  int prob, tm, tn, kind;  // kind = tile index * 4 + layout (akc | bnc << 1)
};

// The device descriptor of one grouped problem: what the per-tile routine
// reads, with 32-bit leading dimensions.  The unit strides follow from the
// item's layout and C is column-major, so every layout case of run_tile
// builds its GemmArgs with those strides as constants -- the full GemmArgs
// (64-bit strides of every operand) held in scalar registers had the
// grouped kernels spilling SGPRs to VGPR lanes.
struct GroupProb {
  const float *a, *b, *a2, *b2;
  float *c;
  int m, n, k, lda, ldb, lda2, ldb2, ldc;
  float alpha, beta;
  int ji, vec;  // vec bits: 1 a, 2 b, 4 a2, 8 b2 (whole 16-byte vectors may be loaded)
};

template <bool AKC, bool BNC>
__device__ inline GemmArgs<float> args_of(const GroupProb &g) {
  GemmArgs<float> q{};
  q.m = g.m; q.n = g.n; q.k = g.k; q.alpha = g.alpha; q.beta = g.beta;
  q.a = g.a; q.rs_a = AKC ? g.lda : 1; q.cs_a = AKC ? 1 : g.lda;
  q.b = g.b; q.rs_b = BNC ? g.ldb : 1; q.cs_b = BNC ? 1 : g.ldb;
  q.a2 = g.a2; q.rs_a2 = AKC ? g.lda2 : 1; q.cs_a2 = AKC ? 1 : g.lda2;
  q.b2 = g.b2; q.rs_b2 = BNC ? g.ldb2 : 1; q.cs_b2 = BNC ? 1 : g.ldb2;
  q.c = g.c; q.rs_c = 1; q.cs_c = g.ldc; q.ji = g.ji;
  q.vec_a = g.vec & 1; q.vec_b = (g.vec >> 1) & 1; q.vec_a2 = (g.vec >> 2) & 1;
  q.vec_b2 = (g.vec >> 3) & 1;
  return q;
}

constexpr int kTiles[3] = {32, 64, 128};

// Launch schedules of each tile class: workgroups per CU, threads per
// workgroup, K-slice depth.  No schedule changes the MFMA order of an
// element (gemm_tile: neither the slice depth nor the wave layout does), so
// results do not depend on it; the autotuner picks one per class.
struct Sched {
  int wg_per_cu, nth, bk;
};
constexpr int kSchedVariants = 2;
constexpr Sched kScheds[3][kSchedVariants] = {
    {{6, 256, 32}, {12, 128, 16}},  // 32-tiles: 4 waves of 16 x 16 / 2 of 16 x 32
    {{3, 512, 32}, {4, 256, 32}},   // 64-tiles: 8 waves of 32 x 16 / 4 of 32 x 32
    {{2, 256, 32}, {2, 512, 32}},   // 128-tiles: 4 waves of 64 x 64 / 8 of 64 x 32
};


// SMALL (gemm_tile's mode for small tiles) for the 64- and 32-tile items.
template <int TILE, int BK, int NTH>
__device__ inline void run_tile(const GroupProb &g, int tm, int tn, int layout, float *lds) {
  constexpr bool SM = TILE <= 64;
  switch (layout) {
    case 0: gemm_tile<float, TILE, false, false, BK, NTH, SM>(args_of<false, false>(g), tm, tn, lds); break;
    case 1: gemm_tile<float, TILE, true, false, BK, NTH, SM>(args_of<true, false>(g), tm, tn, lds); break;
    case 2: gemm_tile<float, TILE, false, true, BK, NTH, SM>(args_of<false, true>(g), tm, tn, lds); break;
    default: gemm_tile<float, TILE, true, true, BK, NTH, SM>(args_of<true, true>(g), tm, tn, lds); break;
  }
}

// MAXT: the tile shape of every item of the launch, which sizes LDS (and
// registers) and so the workgroups per CU: 128-tile items run 2 per CU,
// 64-tile items 3 (of 8 waves: 6 per SIMD) and 32-tile items 6, so the many
// small items of a batch of small problems hide more of their short K
// loops' latency.  NTH threads per workgroup; __launch_bounds__ takes the
// minimum waves per SIMD.  (64-tile items at 4 per CU of 4 waves: 1.597 ms
// per replay; 3 per CU of 8 waves: 1.589; 4 of 8 waves spill: 1.643.)
template <int MAXT, int WG_PER_CU, int BK, int NTH = NT>
__global__ __launch_bounds__(NTH, WG_PER_CU * NTH / 256) void grouped_kernel(
    const GroupProb *__restrict__ probs, const GroupItem *__restrict__ items, int nitems) {
  __shared__ float lds[LdsSize<float, MAXT, BK>::value];
  int it = blockIdx.x;
  if (it >= nitems) return;
  GroupItem w = items[it];
  for (;;) {
    // the next item's entry is read while this one runs
    const int nx = it + gridDim.x;
    const GroupItem wn = items[nx < nitems ? nx : it];
    const GroupProb g = probs[w.prob];
    // every item of a launch has the launch's tile shape (kind >> 2)
    // No barrier between items: a tile's last K slice ends with one, after
    // which neither epilogue touches lds, and a 128-tile's first slice
    // waits at a barrier of its own before it restages lds.  A wave that
    // finishes its part of the epilogue starts loading the next item's
    // first slice while the others still store.
    run_tile<MAXT, BK, NTH>(g, w.tm, w.tn, w.kind & 3, lds);
    if (nx >= nitems) break;
    it = nx;
    w = wn;
  }
}

// ------------------------------------------------------------ tune table --
// Tile shape per size bucket, separately for GEMM problems (bucket (floor
// log2 m, floor log2 n, floor log2 k)) and SYRK / SYR2K problems (bucket
// (floor log2 n, floor log2 k)), each log clamped to [4, 13]; value = index
// into kTiles.  And a launch schedule per tile class (index into kScheds).
std::mutex g_tab_mu;
bool g_tab_loaded = false;
std::map<std::tuple<int, int, int>, int> g_tab;
std::map<std::pair<int, int>, int> g_stab;
int g_sched[3] = {0, 0, 0};

int lg(int v) {
  int l = 0;
  while ((2 << l) <= v) ++l;
  return std::min(std::max(l, 4), 13);
}

int heuristic_tile(int m, int n, int k) {
  (void)k;
  const int mn = std::min(m, n);
  if (mn >= 384) return 2;
  if (mn >= 96) return 1;
  return 0;
}

// Table lines: "gemm lm ln lk tile", "syrk ln lk tile", "sched class
// variant" (a trailing "~" marks a near-tie; '#' starts a comment).  The
// first table format, "lm ln lk tile", reads as gemm lines.
void load_table_locked(const char *path) {
  g_tab.clear();
  g_stab.clear();
  for (int c = 0; c < 3; ++c) g_sched[c] = 0;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    int a, b, c, t;
    if (line[0] == '#') continue;
    if (sscanf(line, "gemm %d %d %d %d", &a, &b, &c, &t) == 4 ||
        sscanf(line, "%d %d %d %d", &a, &b, &c, &t) == 4) {
      if (t >= 0 && t < 3) g_tab[std::make_tuple(a, b, c)] = t;
    } else if (sscanf(line, "syrk %d %d %d", &a, &b, &t) == 3) {
      if (t >= 0 && t < 3) g_stab[std::make_pair(a, b)] = t;
    } else if (sscanf(line, "sched %d %d", &c, &t) == 2) {
      if (c >= 0 && c < 3 && t >= 0 && t < kSchedVariants) g_sched[c] = t;
    }
  }
  fclose(f);
}

// $MUBLIS_TUNE_TABLE, else tune/grouped_gfx950.txt beside the library's
// lib/ directory (the table the autotuner wrote for this device family).
std::string default_table_path() {
  const char *e = getenv("MUBLIS_TUNE_TABLE");
  if (e && *e) return e;
  Dl_info info;
  if (dladdr((void *)&heuristic_tile, &info) && info.dli_fname) {
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tune/grouped_gfx950.txt";
  }
  return "";
}

void ensure_table_locked() {
  if (!g_tab_loaded) {
    const std::string p = default_table_path();
    if (!p.empty()) load_table_locked(p.c_str());
    g_tab_loaded = true;
  }
}

int choose_tile(int m, int n, int k, bool sym) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  ensure_table_locked();
  if (sym) {
    auto it = g_stab.find(std::make_pair(lg(n), lg(k)));
    return it != g_stab.end() ? it->second : heuristic_tile(n, n, k);
  }
  auto it = g_tab.find(std::make_tuple(lg(m), lg(n), lg(k)));
  return it != g_tab.end() ? it->second : heuristic_tile(m, n, k);
}

void current_scheds(int out[3]) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  ensure_table_locked();
  for (int c = 0; c < 3; ++c) out[c] = g_sched[c];
}

// Staged slices load 16-byte vectors wherever a whole vector is in range;
// gfx950's global loads need only element (dword) alignment for that.
bool vec_ok_f(const float *x, int64_t ld) {
  (void)ld;
  return (uintptr_t)x % sizeof(float) == 0;
}

}  // namespace

struct mublis_grouped_plan_s {
  GroupProb *d_probs = nullptr;
  GroupItem *d_items = nullptr;
  int nitems = 0, ncls[3] = {0, 0, 0};
  void *d_trsm = nullptr;  // TRSM descriptors, work lists, inverses, solo workspace
  hipStream_t side = nullptr;  // second branch of the captured TRSM launches
  hipEvent_t fork = nullptr, join = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double flops = 0, bytes = 0;
  int tile_hist[3] = {0, 0, 0};
};

extern "C" {

int mublis_grouped_load_table(const char *path) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (path) {
    load_table_locked(path);
  } else {
    g_tab.clear();
    g_stab.clear();
    for (int c = 0; c < 3; ++c) g_sched[c] = 0;
  }
  g_tab_loaded = true;
  return (int)(g_tab.size() + g_stab.size());
}

}  // extern "C"

namespace {

// A TRSM problem solved by itself (triangle of order above kGroupTrsmMax,
// or alpha = 0), in the left form, with its workspace offset.
struct SoloTrsm {
  int lower, unit, m, n;
  float alpha;
  const float *a;
  int64_t rs_a, cs_a;
  float *b;
  int64_t rs_b, cs_b;
  size_t ws_off, ws_bytes;
};

struct Work {
  int sched[3] = {0, 0, 0};  // launch schedule of each tile class
  std::vector<GemmArgs<float>> probs;
  std::vector<GroupProb> gprobs;  // the device form of probs
  std::vector<GroupItem> items;  // by tile shape (128, 64, 32), each by cost
  int ncls[3] = {0, 0, 0};       // items per tile shape (index 0: 32, 1: 64, 2: 128)
  double flops = 0, bytes = 0;
  int hist[3] = {0, 0, 0};
  // TRSM: descriptors, the work lists of gfx950_trsm_grouped back to back,
  // FP64 inverse workspace, and the problems solved by themselves
  std::vector<TrsmGroupProb<float>> tprobs;
  std::vector<TrsmGroupItem> titems;
  int tcounts[kGroupTrsmLists] = {};
  int64_t inv_doubles = 0;
  std::vector<SoloTrsm> solo;
  size_t solo_bytes = 0;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Left form of a TRSM op as the frame maps it (frame/l3/trsm.c:35-54):
// Right side solves the transposed system (B's strides and dimensions
// swapped, op(A) transposed); a transposed A is a stride swap that turns
// the triangle over.
int add_trsm(const mublis_sbatch_op_t &o, Work &w,
             std::vector<std::vector<TrsmGroupItem>> &lists) {
  const bool left = o.side == MUBLIS_LEFT;
  if (o.m < 0 || o.n < 0 || (o.side != MUBLIS_LEFT && o.side != MUBLIS_RIGHT) ||
      (o.uplo != MUBLIS_LOWER && o.uplo != MUBLIS_UPPER) || (o.trans_a & ~1) ||
      (o.diag & ~1))
    return MUBLIS_EINVAL;
  const int order = left ? o.m : o.n;
  if (o.ldc < std::max(1, o.m) || o.lda < std::max(1, order)) return MUBLIS_EINVAL;
  int64_t rsa = 1, csa = o.lda, rsb = 1, csb = o.ldc;
  int mm = o.m, nn = o.n;
  bool trans = o.trans_a != 0, lower = o.uplo == MUBLIS_LOWER;
  if (!left) {
    std::swap(mm, nn);
    std::swap(rsb, csb);
    trans = !trans;
  }
  if (trans) {
    std::swap(rsa, csa);
    lower = !lower;
  }
  w.flops += (double)mm * mm * nn;
  w.bytes += 4.0 * ((double)mm * (mm + 1) / 2 + 2.0 * mm * nn);
  if (mm == 0 || nn == 0) return MUBLIS_OK;
  if (mm > kGroupTrsmMax || o.alpha == 0.0f) {
    // the full schedule's workspace: the bits of the single cblas_strsm call
    const size_t ws = o.alpha == 0.0f ? 0 : align256(mublis_gfx950_strsm_sched_ws_bytes(mm, nn));
    w.solo.push_back({lower, o.diag, mm, nn, o.alpha, o.a, rsa, csa, o.c, rsb, csb,
                      w.solo_bytes, ws});
    w.solo_bytes += ws;
    return MUBLIS_OK;
  }
  const int pid = (int)w.tprobs.size();
  w.tprobs.push_back({lower, o.diag, mm, nn, o.alpha, o.a, rsa, csa, o.c, rsb, csb,
                      w.inv_doubles});
  const int nbk = (mm + 63) / 64;
  w.inv_doubles += (int64_t)nbk * 64 * 64;
  for (int p = 0; p < nbk; ++p) lists[lower ? 0 : 1].push_back({pid, p});
  const int cls = (mm <= kGroupTrsmWide ? 2 : mm <= kGroupTrsmSmall ? 4 : 6) + (lower ? 0 : 1);
  const int cols = mm <= kGroupTrsmWide ? kGroupTrsmWideCols : kGroupTrsmCols;
  for (int c = 0; c < (nn + cols - 1) / cols; ++c) lists[cls].push_back({pid, c});
  return MUBLIS_OK;
}

// Expand ops into problems and a cost-sorted tile list.
int build_work(const mublis_sbatch_op_t *ops, int count, int flags, Work &w) {
  std::vector<GroupItem> items;
  std::vector<double> cost;
  std::vector<std::vector<TrsmGroupItem>> tlists(kGroupTrsmLists);
  w.probs.reserve(count);
  const int forced = (flags >> 4) & 3;  // MUBLIS_GROUPED_FORCE_TILE(t) stores t + 1
  const int forced_s = (flags >> 8) & 3;  // MUBLIS_GROUPED_FORCE_SCHED(v) stores v + 1
  current_scheds(w.sched);
  if (forced_s)
    for (int c = 0; c < 3; ++c) w.sched[c] = std::min(forced_s - 1, kSchedVariants - 1);
  for (int i = 0; i < count; ++i) {
    const mublis_sbatch_op_t &o = ops[i];
    if (o.op == MUBLIS_BATCH_TRSM) {
      const int st = add_trsm(o, w, tlists);
      if (st) return st;
      continue;
    }
    const bool syrk = o.op == MUBLIS_BATCH_SYRK, syr2k = o.op == MUBLIS_BATCH_SYR2K;
    const bool sym = syrk || syr2k;
    const int m = sym ? o.n : o.m, n = o.n, k = o.k;
    if (m < 0 || n < 0 || k < 0 || (!sym && o.op != MUBLIS_BATCH_GEMM)) return MUBLIS_EINVAL;
    if ((o.trans_a & ~1) || (!sym && (o.trans_b & ~1)) ||
        (sym && o.uplo != MUBLIS_LOWER && o.uplo != MUBLIS_UPPER))
      return MUBLIS_EINVAL;
    // The CBLAS leading-dimension rules (frame/compat/cblas/cblas.c:150-171,
    // 306-325, 380-404, column-major): every stored column of every operand
    // fits its ld, so no tile reads or writes outside the caller's storage.
    const int rows_a = o.trans_a ? k : m;
    const int rows_b = sym ? rows_a : (o.trans_b ? n : k);
    if (o.lda < std::max(1, rows_a) || o.ldc < std::max(1, m) ||
        ((syr2k || !sym) && o.ldb < std::max(1, rows_b)))
      return MUBLIS_EINVAL;
    GemmArgs<float> p;
    memset(&p, 0, sizeof p);
    // alpha == 0 reads neither A nor B (frame/l3/l3_driver.c:123-124): K = 0
    p.m = m; p.n = n; p.k = o.alpha == 0.0f ? 0 : k;
    p.alpha = o.alpha; p.beta = o.beta;
    // column-major operands; transposes are stride swaps (frame/l3/gemm.c:14-24)
    int64_t rsa = 1, csa = o.lda, rsb = 1, csb = o.ldb;
    if (o.trans_a) std::swap(rsa, csa);
    if (sym) {
      if (syr2k && o.trans_a) std::swap(rsb, csb);
      p.a = o.a; p.rs_a = rsa; p.cs_a = csa;
      if (syrk) {
        p.b = o.a; p.rs_b = csa; p.cs_b = rsa;  // B = A' (frame/l3/syrk.c:43-51)
      } else {
        // A*B' and B*A' as the two K segments (frame/l3/syr2k.c:47-68)
        p.b = o.b; p.rs_b = csb; p.cs_b = rsb;
        p.a2 = o.b; p.rs_a2 = rsb; p.cs_a2 = csb;
        p.b2 = o.a; p.rs_b2 = csa; p.cs_b2 = rsa;
      }
      p.ji = o.uplo == MUBLIS_UPPER ? 2 : 1;
    } else {
      if (o.trans_b) std::swap(rsb, csb);
      p.a = o.a; p.rs_a = rsa; p.cs_a = csa;
      p.b = o.b; p.rs_b = rsb; p.cs_b = csb;
      p.ji = 0;
    }
    p.c = o.c; p.rs_c = 1; p.cs_c = o.ldc;
    // the same layout rule as the one-problem kernel (gemm.hip), so both
    // stage the same tiles; SYR2K's second product shares the first's
    const bool pair = syr2k && p.k > 0;
    const int la = common_layout(m, k, p.rs_a, p.cs_a, pair ? &p.rs_a2 : nullptr,
                                 pair ? &p.cs_a2 : nullptr, true);
    const int lb = common_layout(k, n, p.rs_b, p.cs_b, pair ? &p.rs_b2 : nullptr,
                                 pair ? &p.cs_b2 : nullptr, false);
    if (la < 0 || lb < 0) return MUBLIS_EINVAL;
    const bool akc = la == 1, bnc = lb == 1;
    p.vec_a = vec_ok_f(p.a, akc ? p.rs_a : p.cs_a);
    p.vec_b = vec_ok_f(p.b, bnc ? p.rs_b : p.cs_b);
    if (syr2k) {
      p.vec_a2 = vec_ok_f(p.a2, akc ? p.rs_a2 : p.cs_a2);
      p.vec_b2 = vec_ok_f(p.b2, bnc ? p.rs_b2 : p.cs_b2);
    }
    if (p.k == 0) p.a2 = p.b2 = nullptr;
    const int tix = forced ? forced - 1 : choose_tile(m, n, k, sym);
    const int T = kTiles[tix];
    w.hist[tix]++;
    const int layout = (akc ? 1 : 0) | (bnc ? 2 : 0);
    const int tmn = (m + T - 1) / T, tnn = (n + T - 1) / T;
    const int pid = (int)w.probs.size();
    w.probs.push_back(p);
    GroupProb g;
    g.a = p.a; g.b = p.b; g.a2 = p.a2; g.b2 = p.b2; g.c = p.c;
    g.m = p.m; g.n = p.n; g.k = p.k;
    g.lda = (int)(akc ? p.rs_a : p.cs_a);
    g.ldb = (int)(bnc ? p.rs_b : p.cs_b);
    g.lda2 = p.a2 ? (int)(akc ? p.rs_a2 : p.cs_a2) : 0;
    g.ldb2 = p.b2 ? (int)(bnc ? p.rs_b2 : p.cs_b2) : 0;
    g.ldc = o.ldc;
    g.alpha = p.alpha; g.beta = p.beta; g.ji = p.ji;
    g.vec = (p.vec_a ? 1 : 0) | (p.vec_b ? 2 : 0) | (p.vec_a2 ? 4 : 0) | (p.vec_b2 ? 8 : 0);
    w.gprobs.push_back(g);
    const double tri = n * (n + 1) / 2.0;
    w.flops += syrk ? (double)n * (n + 1) * k : syr2k ? 2.0 * n * (n + 1) * k
                                                      : 2.0 * m * n * k;
    w.bytes += 4.0 * ((syr2k ? 2.0 : 1.0) * m * k + (sym ? 0.0 : (double)k * n) +
                      (o.beta != 0.0f ? 2.0 : 1.0) * (sym ? tri : (double)m * n));
    if (m == 0 || n == 0) continue;
    for (int tn = 0; tn < tnn; ++tn)
      for (int tm = 0; tm < tmn; ++tm) {
        if (sym && p.ji == 1 && tn > tm) continue;  // tile entirely above
        if (sym && p.ji == 2 && tm > tn) continue;  // tile entirely below
        items.push_back({pid, tm, tn, tix * 4 + layout});
        cost.push_back((double)T * T * std::max(k, 1) * (syr2k ? 2 : 1) + 64.0 * T * T);
      }
  }
  // One launch per tile shape, largest shape first.  Within a launch:
  // whole problems are dealt to the 8 XCDs, largest first onto the least
  // loaded one (LPT), and the per-XCD lists are interleaved so that list
  // position p runs on XCD p % 8 (workgroup p % grid, grids being multiples
  // of 8, and hardware workgroup ids dealt to XCDs round robin): the tiles of
  // one problem then share that XCD's L2 instead of re-reading their A and B
  // panels on eight.  Each list runs largest items first, so the cyclic walk
  // ends with small items.
  w.items.clear();
  w.items.reserve(items.size());
  for (int c = 2; c >= 0; --c) {
    std::vector<int> probs_c;  // problems of this class, by cost of their tiles
    std::vector<double> pcost(w.probs.size(), 0.0);
    std::vector<std::vector<int>> ptiles(w.probs.size());
    for (size_t i = 0; i < items.size(); ++i) {
      if ((items[i].kind >> 2) != c) continue;
      const int pr = items[i].prob;
      if (ptiles[pr].empty()) probs_c.push_back(pr);
      ptiles[pr].push_back((int)i);
      pcost[pr] += cost[i];
    }
    std::stable_sort(probs_c.begin(), probs_c.end(), [&](int x, int y) {
      return cost[ptiles[x][0]] > cost[ptiles[y][0]];
    });
    std::vector<int> lists[8];
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int pr : probs_c) {
      const int x = (int)(std::min_element(load, load + 8) - load);
      load[x] += pcost[pr];
      lists[x].insert(lists[x].end(), ptiles[pr].begin(), ptiles[pr].end());
    }
    size_t pos[8] = {0, 0, 0, 0, 0, 0, 0, 0}, left = 0;
    for (int x = 0; x < 8; ++x) left += lists[x].size();
    w.ncls[c] = (int)left;
    for (size_t p = 0; left > 0; ++p) {
      int x = (int)(p % 8);
      if (pos[x] == lists[x].size()) {  // this XCD's list is done: take from the longest
        size_t best = 0;
        for (int y = 0; y < 8; ++y)
          if (lists[y].size() - pos[y] > best) {
            best = lists[y].size() - pos[y];
            x = y;
          }
      }
      w.items.push_back(items[lists[x][pos[x]++]]);
      --left;
    }
  }
  // TRSM panels run longest first (a panel's work grows with the square of
  // its problem's order)
  for (int c = 2; c < kGroupTrsmLists; ++c)
    std::stable_sort(tlists[c].begin(), tlists[c].end(),
                     [&](const TrsmGroupItem &x, const TrsmGroupItem &y) {
                       return w.tprobs[x.prob].m > w.tprobs[y.prob].m;
                     });
  for (int c = 0; c < kGroupTrsmLists; ++c) {
    w.tcounts[c] = (int)tlists[c].size();
    w.titems.insert(w.titems.end(), tlists[c].begin(), tlists[c].end());
  }
  return MUBLIS_OK;
}

// Device bytes of a work list's TRSM part: descriptors, items, inverses,
// solo workspace (each 256-byte aligned, in that order).
size_t trsm_bytes(const Work &w) {
  return align256(sizeof(TrsmGroupProb<float>) * w.tprobs.size()) +
         align256(sizeof(TrsmGroupItem) * w.titems.size()) +
         align256(sizeof(double) * (size_t)w.inv_doubles) + w.solo_bytes;
}
// Bytes of the TRSM part that come from the host (descriptors and items).
size_t trsm_host_bytes(const Work &w) {
  return align256(sizeof(TrsmGroupProb<float>) * w.tprobs.size()) +
         align256(sizeof(TrsmGroupItem) * w.titems.size());
}
void trsm_stage(const Work &w, char *h) {
  memcpy(h, w.tprobs.data(), sizeof(TrsmGroupProb<float>) * w.tprobs.size());
  memcpy(h + align256(sizeof(TrsmGroupProb<float>) * w.tprobs.size()), w.titems.data(),
         sizeof(TrsmGroupItem) * w.titems.size());
}
// The TRSM launches of a work list whose TRSM part sits at dev (laid out as
// trsm_bytes describes), on s (upper-triangle launches on s2 between the
// fork and join events, when given).
int launch_trsm(const Work &w, char *dev, hipStream_t s, hipStream_t s2 = nullptr,
                hipEvent_t fork = nullptr, hipEvent_t join = nullptr) {
  const size_t pb = align256(sizeof(TrsmGroupProb<float>) * w.tprobs.size());
  const size_t ib = align256(sizeof(TrsmGroupItem) * w.titems.size());
  const size_t vb = align256(sizeof(double) * (size_t)w.inv_doubles);
  int st = MUBLIS_OK;
  if (!w.titems.empty())
    st = gfx950_trsm_grouped<float>((const TrsmGroupProb<float> *)dev,
                                    (const TrsmGroupItem *)(dev + pb), w.tcounts,
                                    (double *)(dev + pb + ib), s, s2, fork, join);
  char *solo_ws = dev + pb + ib + vb;
  for (size_t i = 0; i < w.solo.size() && !st; ++i) {
    const SoloTrsm &t = w.solo[i];
    st = mublis_gfx950_strsm_ws_dev(t.lower, t.unit, t.m, t.n, t.alpha, t.a, t.rs_a, t.cs_a,
                                    t.b, t.rs_b, t.cs_b, t.ws_bytes ? solo_ws + t.ws_off : nullptr,
                                    t.ws_bytes, s);
  }
  return st;
}

// Persistent grid: per_cu workgroups per CU, at most one per item, rounded
// up to a multiple of the 8 XCDs (the work lists are laid out for that).
int persistent_grid(int nitems, int per_cu) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  const int g = std::max(1, std::min(nitems, per_cu * cus));
  return (g + 7) / 8 * 8;
}

template <int TILE, int W, int BK, int NTH>
void launch_class(const GroupProb *probs, const GroupItem *items, int n, hipStream_t s) {
  hipLaunchKernelGGL((grouped_kernel<TILE, W, BK, NTH>), dim3(persistent_grid(n, W)),
                     dim3(NTH), 0, s, probs, items, n);
}

// The launches of a work list, back to back on one stream (any may be
// empty): the 128-tile items, the 64-tile items, the 32-tile items, each
// class with its schedule (kScheds).  Measured on the batched workload,
// back-to-back launches beat parallel graph branches (1.86 vs 2.69 ms for
// two classes) and one launch for all items at 2 workgroups per CU (2.27 ms).
int launch_items(const GroupProb *probs, const GroupItem *items, const int *ncls,
                 const int *sched, hipStream_t s) {
  const int n128 = ncls[2], n64 = ncls[1], n32 = ncls[0];
#define MUBLIS_CLASS(T, C, V, P, N)                                                   \
  launch_class<T, kScheds[C][V].wg_per_cu, kScheds[C][V].bk, kScheds[C][V].nth>(probs, P, N, s)
  if (n128 > 0) {
    if (sched[2] == 1) MUBLIS_CLASS(128, 2, 1, items, n128);
    else MUBLIS_CLASS(128, 2, 0, items, n128);
  }
  if (n64 > 0) {
    if (sched[1] == 1) MUBLIS_CLASS(64, 1, 1, items + n128, n64);
    else MUBLIS_CLASS(64, 1, 0, items + n128, n64);
  }
  if (n32 > 0) {
    if (sched[0] == 1) MUBLIS_CLASS(32, 0, 1, items + n128 + n64, n32);
    else MUBLIS_CLASS(32, 0, 0, items + n128 + n64, n32);
  }
#undef MUBLIS_CLASS
  return hipGetLastError() == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

// ------------------------------------------------------------ autotuner --
// Every candidate (tile shape of a bucket, schedule of a class) is timed on
// a batch that fills the GPU -- at least kTuneTiles tiles at the 32-tile
// shape -- after a warm-up replay, as the best of `reps` replays: throughput
// inside a large batch, which is what the grouped workloads run, not the
// latency of a few problems.  Every candidate within kTieMargin of the
// fastest is a near-tie; among near-ties the heuristic's choice (for
// schedules, variant 0) wins, else the lowest-numbered one, so repeated
// tuning settles on the same table wherever the timings cannot tell the
// choices apart (run-to-run noise of a batch is a few percent).
constexpr int kTuneTiles = 6144;
constexpr double kTieMargin = 0.03;

// Index chosen among n timings: the preferred candidate if it is within the
// margin of the fastest, else the lowest-numbered candidate that is.  *tie
// is set when more than one candidate is within the margin.
inline int tune_pick(const double *ms, int n, int preferred, bool *tie) {
  int best = 0;
  for (int i = 1; i < n; ++i)
    if (ms[i] < ms[best]) best = i;
  const double lim = ms[best] * (1 + kTieMargin);
  int pick = -1, within = 0;
  for (int i = 0; i < n; ++i)
    if (ms[i] <= lim) {
      ++within;
      if (pick < 0) pick = i;
    }
  if (ms[preferred] <= lim) pick = preferred;
  *tie = within > 1;
  return pick;
}

struct Tuner {
  float *buf = nullptr;
  size_t cap = 0;  // floats
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int reps = 1;
  uint64_t seed = 314159;

  ~Tuner() {
    if (buf) (void)hipFree(buf);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  int init(int r) {
    reps = r;
    s = (hipStream_t)mublis_rt_stream();
    if (!s || hipEventCreate(&e0) || hipEventCreate(&e1)) return MUBLIS_EHIP;
    return MUBLIS_OK;
  }
  int uniform(int lo, int hi) {  // in [lo, hi)
    seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    return lo + (int)((seed >> 33) % (uint64_t)(hi - lo));
  }
  // device memory for `floats` operand elements (zero-filled; values do not
  // change the timing of the MFMA path)
  int reserve(size_t floats) {
    if (floats <= cap) return MUBLIS_OK;
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, floats * sizeof(float)) != hipSuccess) return MUBLIS_ENOMEM;
    if (hipMemset(buf, 0, floats * sizeof(float)) != hipSuccess) return MUBLIS_EHIP;
    cap = floats;
    return MUBLIS_OK;
  }
  // point the ops (sizes set) at disjoint column-major operands in buf
  int place(std::vector<mublis_sbatch_op_t> &ops) {
    size_t total = 0;
    for (const auto &o : ops)
      total += (size_t)o.m * o.k + (o.op == MUBLIS_BATCH_GEMM ? (size_t)o.k * o.n : 0) +
               (size_t)o.m * o.n + 64;
    int st = reserve(total);
    if (st) return st;
    size_t off = 0;
    for (auto &o : ops) {
      o.a = buf + off;
      off += (size_t)o.m * o.k + 16;
      if (o.op == MUBLIS_BATCH_GEMM) {
        o.b = buf + off;
        off += (size_t)o.k * o.n + 16;
      }
      o.c = buf + off;
      off += (size_t)o.m * o.n + 32;
    }
    return MUBLIS_OK;
  }
  static mublis_sbatch_op_t op_of(bool sym, int m, int n, int k) {
    mublis_sbatch_op_t o;
    memset(&o, 0, sizeof o);
    o.op = sym ? MUBLIS_BATCH_SYRK : MUBLIS_BATCH_GEMM;
    o.uplo = MUBLIS_LOWER;
    o.m = sym ? n : m;
    o.n = n;
    o.k = k;
    o.alpha = 1.0f;
    o.beta = 1.0f;
    o.lda = o.m;
    o.ldb = o.k;
    o.ldc = o.m;
    return o;
  }
  static int tiles32(const mublis_sbatch_op_t &o) {
    const int tm = (o.m + 31) / 32, tn = (o.n + 31) / 32;
    return o.op == MUBLIS_BATCH_GEMM ? tm * tn : tm * (tm + 1) / 2;
  }
  // problems of one bucket: each size uniform in [2^l, 2^(l+1))
  int batch_bucket(std::vector<mublis_sbatch_op_t> &ops, bool sym, int lm, int ln, int lk,
                   int min_count) {
    int tiles = 0;
    while ((int)ops.size() < min_count || tiles < kTuneTiles) {
      const int n = uniform(1 << ln, 2 << ln), k = uniform(1 << lk, 2 << lk);
      const int m = sym ? n : uniform(1 << lm, 2 << lm);
      ops.push_back(op_of(sym, m, n, k));
      tiles += tiles32(ops.back());
    }
    return place(ops);
  }
  // GEMM problems with m, n in [lo, hi) and k in [16, 512)
  int batch_range(std::vector<mublis_sbatch_op_t> &ops, int lo, int hi, int cls,
                  int min_count) {
    int tiles = 0;
    const int need = kTuneTiles << (2 * cls);  // as many tiles of the class's own shape
    while ((int)ops.size() < min_count || tiles < need) {
      ops.push_back(op_of(false, uniform(lo, hi), uniform(lo, hi), uniform(16, 512)));
      tiles += tiles32(ops.back());
    }
    return place(ops);
  }
  // Best of `reps` replays of each of n plans of the same ops, created with
  // flags[i], into ms[i].  The replays go round-robin over the candidates
  // (after one warm-up replay each), so a drift of the clocks over the run
  // does not favour the candidates timed first or last.
  int time(const std::vector<mublis_sbatch_op_t> &ops, const int *flags, int n, double *ms) {
    std::vector<mublis_grouped_plan_t> plans(n, nullptr);
    int st = MUBLIS_OK;
    for (int i = 0; i < n && !st; ++i) {
      st = mublis_sgrouped_plan_create(ops.data(), (int)ops.size(), flags[i], &plans[i]);
      if (!st) st = mublis_grouped_plan_execute(plans[i], s);  // warm-up
      ms[i] = 1e30;
    }
    for (int r = 0; r < reps && !st; ++r)
      for (int i = 0; i < n && !st; ++i) {
        float t = 0;
        if (hipEventRecord(e0, s) != hipSuccess) st = MUBLIS_EHIP;
        if (!st) st = mublis_grouped_plan_execute(plans[i], s);
        if (!st && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t, e0, e1) != hipSuccess))
          st = MUBLIS_EHIP;
        if (!st) ms[i] = std::min(ms[i], (double)t);
      }
    if (hipStreamSynchronize(s) != hipSuccess && !st) st = MUBLIS_EHIP;
    for (auto p : plans) mublis_grouped_plan_destroy(p);
    return st;
  }
};

}  // namespace

extern "C" {

int mublis_sgrouped_plan_create(const mublis_sbatch_op_t *ops, int count, int flags,
                                mublis_grouped_plan_t *out) {
  if (!ops || count < 0 || !out) return MUBLIS_EINVAL;
  int st = mublis_rt_init();
  if (st) return st;
  Work w;
  if ((st = build_work(ops, count, flags, w))) return st;
  auto *plan = new mublis_grouped_plan_s();
  plan->flops = w.flops;
  plan->bytes = w.bytes;
  for (int i = 0; i < 3; ++i) plan->tile_hist[i] = w.hist[i];
  const std::vector<GroupProb> &probs = w.gprobs;
  const std::vector<GroupItem> &sorted = w.items;
  plan->nitems = (int)sorted.size() + (int)w.solo.size();
  for (int c = 2; c < kGroupTrsmLists; ++c) plan->nitems += w.tcounts[c];
  for (int c = 0; c < 3; ++c) plan->ncls[c] = w.ncls[c];
  st = MUBLIS_OK;
  if (!probs.empty() && hipMalloc(&plan->d_probs, sizeof(GroupProb) * probs.size()))
    st = MUBLIS_ENOMEM;
  if (!st && !sorted.empty() && hipMalloc(&plan->d_items, sizeof(GroupItem) * sorted.size()))
    st = MUBLIS_ENOMEM;
  if (!st && !probs.empty() &&
      hipMemcpy(plan->d_probs, probs.data(), sizeof(GroupProb) * probs.size(),
                hipMemcpyHostToDevice))
    st = MUBLIS_EHIP;
  if (!st && !sorted.empty() &&
      hipMemcpy(plan->d_items, sorted.data(), sizeof(GroupItem) * sorted.size(),
                hipMemcpyHostToDevice))
    st = MUBLIS_EHIP;
  const size_t tb = trsm_bytes(w);
  if (!st && tb > 0 && hipMalloc(&plan->d_trsm, tb)) st = MUBLIS_ENOMEM;
  if (!st && trsm_host_bytes(w) > 0) {
    std::vector<char> h(trsm_host_bytes(w));
    trsm_stage(w, h.data());
    if (hipMemcpy(plan->d_trsm, h.data(), h.size(), hipMemcpyHostToDevice)) st = MUBLIS_EHIP;
  }
  // Capture the launches once; executions replay the graph.
  if (!st && hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking)) st = MUBLIS_EHIP;
  if (!st && w.tcounts[0] > 0 && w.tcounts[1] > 0 &&  // lower and upper problems
      (hipStreamCreateWithFlags(&plan->side, hipStreamNonBlocking) ||
       hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming) ||
       hipEventCreateWithFlags(&plan->join, hipEventDisableTiming)))
    st = MUBLIS_EHIP;
  if (!st && hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal)) st = MUBLIS_EHIP;
  if (!st) {
    int l = launch_items(plan->d_probs, plan->d_items, plan->ncls, w.sched, plan->cap);
    if (!l) l = launch_trsm(w, (char *)plan->d_trsm, plan->cap, plan->side, plan->fork,
                            plan->join);
    hipError_t e2 = hipStreamEndCapture(plan->cap, &plan->graph);
    if (l || e2) st = l ? l : MUBLIS_EHIP;
  }
  if (!st && hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0))
    st = MUBLIS_EHIP;
  if (st) {
    mublis_grouped_plan_destroy(plan);
    return st;
  }
  *out = plan;
  return MUBLIS_OK;
}

size_t mublis_sgrouped_ws_bytes(const mublis_sbatch_op_t *ops, int count) {
  Work w;
  if (!ops || count < 0 || build_work(ops, count, 0, w)) return 0;
  return align256(sizeof(GroupProb) * w.gprobs.size()) +
         align256(sizeof(GroupItem) * w.items.size()) + trsm_bytes(w) + 256;
}

int mublis_sgrouped_launch_ws(const mublis_sbatch_op_t *ops, int count, void *host_stage,
                              void *dev_ws, size_t ws_bytes, void *stream) {
  if (!ops || count < 0 || !host_stage || !dev_ws) return MUBLIS_EINVAL;
  Work w;
  int st = build_work(ops, count, 0, w);
  if (st) return st;
  const size_t pb = align256(sizeof(GroupProb) * w.gprobs.size());
  const size_t ib = align256(sizeof(GroupItem) * w.items.size());
  const size_t hb = trsm_host_bytes(w);
  if (pb + ib + trsm_bytes(w) > ws_bytes) return MUBLIS_EINVAL;
  if (w.items.empty() && w.titems.empty() && w.solo.empty()) return MUBLIS_OK;
  // host_stage mirrors dev_ws: GEMM descriptors and items, then the TRSM
  // descriptors and items (the inverses and solo workspace stay on device)
  char *h = (char *)host_stage;
  memcpy(h, w.gprobs.data(), sizeof(GroupProb) * w.gprobs.size());
  memcpy(h + pb, w.items.data(), sizeof(GroupItem) * w.items.size());
  trsm_stage(w, h + pb + ib);
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  if (hipMemcpyAsync(dev_ws, h, pb + ib + hb, hipMemcpyHostToDevice, s)) return MUBLIS_EHIP;
  st = launch_items((const GroupProb *)dev_ws, (const GroupItem *)((char *)dev_ws + pb),
                    w.ncls, w.sched, s);
  if (!st) st = launch_trsm(w, (char *)dev_ws + pb + ib, s);
  return st;
}

int mublis_grouped_plan_execute(mublis_grouped_plan_t plan, void *stream) {
  if (!plan) return MUBLIS_EINVAL;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  return hipGraphLaunch(plan->exec, s) == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
}

int mublis_grouped_plan_info(mublis_grouped_plan_t plan, double *flops, double *bytes,
                             int *nitems, int *tile_hist3) {
  if (!plan) return MUBLIS_EINVAL;
  if (flops) *flops = plan->flops;
  if (bytes) *bytes = plan->bytes;
  if (nitems) *nitems = plan->nitems;
  if (tile_hist3)
    for (int i = 0; i < 3; ++i) tile_hist3[i] = plan->tile_hist[i];
  return MUBLIS_OK;
}

int mublis_grouped_plan_destroy(mublis_grouped_plan_t plan) {
  if (!plan) return MUBLIS_OK;
  if (plan->cap) (void)hipStreamSynchronize(plan->cap);
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->cap) (void)hipStreamDestroy(plan->cap);
  if (plan->side) (void)hipStreamDestroy(plan->side);
  if (plan->fork) (void)hipEventDestroy(plan->fork);
  if (plan->join) (void)hipEventDestroy(plan->join);
  if (plan->d_probs) (void)hipFree(plan->d_probs);
  if (plan->d_items) (void)hipFree(plan->d_items);
  if (plan->d_trsm) (void)hipFree(plan->d_trsm);
  delete plan;
  return MUBLIS_OK;
}

int mublis_sgemm_grouped(const mublis_sbatch_op_t *ops, int count) {
  mublis_grouped_plan_t plan = nullptr;
  int st = mublis_sgrouped_plan_create(ops, count, 0, &plan);
  if (st) return st;
  void *s = mublis_rt_stream();
  st = mublis_grouped_plan_execute(plan, s);
  int sy = mublis_rt_sync(s);
  if (!st) st = sy;
  mublis_grouped_plan_destroy(plan);
  return st;
}

/*
 * Autotuner: for every size bucket (floor log2 of m, n, k over [lo, hi]),
 * time a grouped launch of `per_bucket` random problems of that bucket with
 * each tile shape forced, and keep the fastest.  Writes "lm ln lk tile"
 * lines to `path` and loads the table.  Operands are device buffers owned by
 * the tuner; values are irrelevant to timing.
 */
int mublis_grouped_autotune(const char *path, int lo, int hi, int per_bucket, int reps) {
  int st = mublis_rt_init();
  if (st) return st;
  if (!path || lo < 4 || hi > 12 || lo > hi || per_bucket < 1 || reps < 1) return MUBLIS_EINVAL;
  Tuner t;
  if ((st = t.init(reps))) return st;
  FILE *f = fopen(path, "w");
  if (!f) return MUBLIS_EINVAL;
  fprintf(f, "# gfx950 grouped-kernel table (tiles 32 64 128)\n");
  fprintf(f, "# sched class variant | gemm lm ln lk tile | syrk ln lk tile; buckets are "
             "floor(log2) of the sizes\n");
  fprintf(f, "# each candidate: a batch of >= %d tiles (>= %d problems), best of %d replays; "
             "candidates within %.0f%% of the\n# fastest are near-ties (marked ~): the "
             "heuristic's (schedule variant 0) if among them, else the first\n",
          kTuneTiles, per_bucket, reps, 100 * kTieMargin);
  // 1. the schedule of each tile class, on a batch of problems of that
  //    class's typical sizes with the tile shape forced
  static const int kSchedRange[3][2] = {{16, 128}, {64, 512}, {256, 1024}};
  int sched[3] = {0, 0, 0};
  for (int c = 0; c < 3 && !st; ++c) {
    std::vector<mublis_sbatch_op_t> ops;
    st = t.batch_range(ops, kSchedRange[c][0], kSchedRange[c][1], c, per_bucket);
    double ms[kSchedVariants];
    int fl[kSchedVariants];
    for (int v = 0; v < kSchedVariants; ++v)
      fl[v] = MUBLIS_GROUPED_FORCE_TILE(c) | MUBLIS_GROUPED_FORCE_SCHED(v);
    if (!st) st = t.time(ops, fl, kSchedVariants, ms);
    if (st) break;
    bool tie = false;
    sched[c] = tune_pick(ms, kSchedVariants, 0, &tie);
    fprintf(f, "sched %d %d%s\n", c, sched[c], tie ? " ~" : "");
  }
  // 2. the tile shape of each GEMM bucket, then of each SYRK bucket, with
  //    those schedules
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    for (int c = 0; c < 3; ++c) g_sched[c] = sched[c];
    g_tab_loaded = true;
  }
  for (int sym = 0; sym < 2 && !st; ++sym)
    for (int lm = sym ? 0 : lo; lm <= (sym ? 0 : hi) && !st; ++lm)
      for (int ln = lo; ln <= hi && !st; ++ln)
        for (int lk = lo; lk <= hi && !st; ++lk) {
          std::vector<mublis_sbatch_op_t> ops;
          st = t.batch_bucket(ops, sym != 0, lm, ln, lk, per_bucket);
          double ms[3];
          const int fl[3] = {MUBLIS_GROUPED_FORCE_TILE(0), MUBLIS_GROUPED_FORCE_TILE(1),
                             MUBLIS_GROUPED_FORCE_TILE(2)};
          if (!st) st = t.time(ops, fl, 3, ms);
          if (st) break;
          const int h = sym ? heuristic_tile(1 << ln, 1 << ln, 1 << lk)
                            : heuristic_tile(1 << lm, 1 << ln, 1 << lk);
          bool tie = false;
          const int pick = tune_pick(ms, 3, h, &tie);
          if (sym) fprintf(f, "syrk %d %d %d%s\n", ln, lk, pick, tie ? " ~" : "");
          else fprintf(f, "gemm %d %d %d %d%s\n", lm, ln, lk, pick, tie ? " ~" : "");
        }
  fclose(f);
  if (!st) mublis_grouped_load_table(path);
  return st;
}

}  // extern "C"
