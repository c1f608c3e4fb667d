// Randomised check of the serving loop's budget allocator
// (mublis_amd/runtime/budget_arena.h) against a brute-force model that
// tracks every 256-byte unit.  Built with the host sanitizers by
// tests/test_serve_host.py; exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "budget_arena.h"

using mublis_serve::BudgetArena;

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main(int argc, char **argv) {
  const unsigned seed = argc > 1 ? (unsigned)std::atoi(argv[1]) : 1u;
  const size_t units = 4096;
  const size_t cap = units * BudgetArena::kAlign;
  BudgetArena arena(cap + 100);  // capacity rounds down to the alignment
  CHECK(arena.capacity() == cap && arena.largest_free() == cap);
  std::vector<int> owner(units, -1);  // model: which live block holds a unit
  struct Blk { size_t off, len; };
  std::vector<Blk> live;
  std::mt19937_64 rng(seed);
  for (int step = 0; step < 200000; ++step) {
    const bool do_alloc = live.empty() || (rng() % 100) < 55;
    if (do_alloc) {
      const size_t bytes = (rng() % 4 == 0) ? rng() % 64 : rng() % (cap / 16);
      const size_t need = BudgetArena::round(bytes);
      // model: is there a free run of need/256 units?
      size_t run = 0, best = 0;
      for (size_t u = 0; u < units; ++u) {
        run = owner[u] < 0 ? run + 1 : 0;
        if (run > best) best = run;
      }
      const size_t off = arena.alloc(bytes);
      CHECK(arena.largest_free() <= best * BudgetArena::kAlign);
      if (off == BudgetArena::kFail) {
        CHECK(best * BudgetArena::kAlign < need);
        continue;
      }
      CHECK(off % BudgetArena::kAlign == 0 && off + need <= cap);
      for (size_t u = off / BudgetArena::kAlign; u < (off + need) / BudgetArena::kAlign; ++u) {
        CHECK(owner[u] < 0);  // no overlap with a live block
        owner[u] = (int)live.size();
      }
      live.push_back({off, need});
    } else {
      const size_t i = rng() % live.size();
      const Blk b = live[i];
      CHECK(arena.release(b.off));
      CHECK(!arena.release(b.off));  // a second release is refused
      for (size_t u = b.off / BudgetArena::kAlign; u < (b.off + b.len) / BudgetArena::kAlign; ++u)
        owner[u] = -1;
      live[i] = live.back();
      live.pop_back();
      for (size_t j = 0; j < live.size(); ++j)
        for (size_t u = live[j].off / BudgetArena::kAlign;
             u < (live[j].off + live[j].len) / BudgetArena::kAlign; ++u)
          owner[u] = (int)j;
    }
    size_t used = 0;
    for (const Blk &b : live) used += b.len;
    CHECK(arena.used() == used && arena.live_blocks() == live.size());
    CHECK(arena.peak() >= used);
    // coalescing: free blocks are never adjacent, so there are at most
    // live + 1 of them
    CHECK(arena.free_blocks() <= live.size() + 1);
  }
  for (const Blk &b : live) CHECK(arena.release(b.off));
  CHECK(arena.used() == 0 && arena.free_blocks() == 1 && arena.largest_free() == cap);
  std::printf("ok seed %u\n", seed);
  return 0;
}
