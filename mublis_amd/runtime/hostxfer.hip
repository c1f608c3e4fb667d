// Host-pointer operands: pooled pinned staging, and a GEMM that streams its
// column panels so that the transfers overlap the kernels.
//
// The reference computes on the caller's host memory directly and keeps a
// process-lifetime pool of pack buffers (frame/base/pool.c:120-318).  Here
// host operands cross PCIe, and the pool holds pinned (page-locked) blocks:
//   * kBlocks blocks of kBlock bytes, allocated with hipHostMalloc and
//     touched once by the CPU when the pool is built (the first CPU write to
//     fresh pinned memory faults its pages in at a few GB/s; later writes
//     run at memory speed), each with an event marking its last transfer;
//   * a team of copy threads (the caller among them) gathers host columns
//     into a block and scatters a block back (one thread moves about 30
//     GB/s, four about 85, so the team keeps up with the ~57 GB/s link; a
//     pageable hipMemcpy moved 31 GB/s host to device and 18 back);
//   * a block is reused once its event has completed, so the gather of one
//     piece overlaps the DMA of the one before.
// mublis_rt_h2d_2d / mublis_rt_d2h_2d move a rows x cols block (rows
// contiguous, host leading dimension ld) to or from a dense device block;
// the frame stages large operands with them.  mublis_gfx950_?gemm_host runs
// C := alpha * A * B + beta * C with C (and any of A, B) in host memory:
// A, and B when its columns are not contiguous, are sent whole; then for
// each column panel of C the panel's B and C columns are sent on an upload
// stream, the panel's GEMM runs on the caller's stream, and its C columns
// come back on a download stream while the next panel is sent and computed.
// One transfer or host GEMM runs at a time (the pool's mutex).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "mublis_gfx950.h"

void gfx950_set_error(const char *msg);

namespace {

constexpr size_t kBlock = size_t(16) << 20;  // bytes per pinned block
constexpr int kBlocks = 16;                   // 256 MiB of pinned memory
constexpr int kCopyThreads = 3;               // plus the calling thread
constexpr size_t kSeg = size_t(1) << 20;      // bytes per copy task
constexpr int kPanelBlocks = 4;               // most blocks one panel's C (or B) takes

// ------------------------------------------------------------ copy team --
struct Seg {
  char *dst;
  const char *src;
  size_t n;
};

class CopyTeam {
 public:
  ~CopyTeam() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  // Runs every copy of segs[0 .. n) across the team; returns when all are done.
  void run(const std::vector<Seg> &segs) {
    size_t total = 0;
    for (const Seg &s : segs) total += s.n;
    if (total < 4 * kSeg) {
      for (const Seg &s : segs) memcpy(s.dst, s.src, s.n);
      return;
    }
    start();
    {
      std::lock_guard<std::mutex> lk(mu_);
      segs_ = &segs;
      next_.store(0);
      active_ = (int)th_.size();
      ++gen_;
    }
    cv_.notify_all();
    work(segs);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return active_ == 0; });
    segs_ = nullptr;
  }

 private:
  void start() {
    if (!th_.empty()) return;
    for (int i = 0; i < kCopyThreads; ++i) th_.emplace_back([this] { loop(); });
  }
  void work(const std::vector<Seg> &segs) {
    for (;;) {
      const size_t i = next_.fetch_add(1);
      if (i >= segs.size()) return;
      memcpy(segs[i].dst, segs[i].src, segs[i].n);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      const std::vector<Seg> *segs;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        segs = segs_;
      }
      work(*segs);
      std::lock_guard<std::mutex> lk(mu_);
      if (--active_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::vector<Seg> *segs_ = nullptr;
  std::atomic<size_t> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// ---------------------------------------------------------- pinned pool --
struct Pool {
  std::mutex mu;  // held for a whole transfer or host GEMM
  bool built = false;
  char *blk[kBlocks] = {};
  hipEvent_t ev[kBlocks] = {};
  bool pending[kBlocks] = {};  // an event was recorded and not waited for yet
  bool held[kBlocks] = {};     // holds data a scatter has still to take
  int next = 0;
  hipStream_t up = nullptr, dn = nullptr;
  CopyTeam team;
};
Pool g_pool;

int fail(hipError_t e, const char *what) {
  char buf[160];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  gfx950_set_error(buf);
  return MUBLIS_EHIP;
}

int build_pool() {
  Pool &g = g_pool;
  if (g.built) return MUBLIS_OK;
  hipError_t e = hipSuccess;
  for (int i = 0; i < kBlocks && e == hipSuccess; ++i) {
    if (!g.blk[i]) e = hipHostMalloc((void **)&g.blk[i], kBlock, hipHostMallocDefault);
    if (e == hipSuccess && !g.ev[i]) e = hipEventCreateWithFlags(&g.ev[i], hipEventDisableTiming);
  }
  if (e == hipSuccess && !g.up) e = hipStreamCreateWithFlags(&g.up, hipStreamNonBlocking);
  if (e == hipSuccess && !g.dn) e = hipStreamCreateWithFlags(&g.dn, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(e, "pinned staging pool");
  // fault every page in once, spread over the team
  static const char zero[kSeg] = {};
  std::vector<Seg> segs;
  for (int i = 0; i < kBlocks; ++i)
    for (size_t o = 0; o < kBlock; o += kSeg) segs.push_back({g.blk[i] + o, zero, kSeg});
  g.team.run(segs);
  g.built = true;
  return MUBLIS_OK;
}

// A block no scatter still needs, its last transfer complete; -1 on error.
int acquire() {
  Pool &g = g_pool;
  for (int tries = 0; tries < kBlocks; ++tries) {
    const int b = g.next;
    g.next = (g.next + 1) % kBlocks;
    if (g.held[b]) continue;
    if (g.pending[b]) {
      const hipError_t e = hipEventSynchronize(g.ev[b]);
      if (e != hipSuccess) return fail(e, "pinned block wait"), -1;
      g.pending[b] = false;
    }
    return b;
  }
  gfx950_set_error("pinned staging pool: every block is held");
  return -1;
}

// Pieces of a rows x cols block that fit one pinned block: whole columns
// when a column fits, else row ranges of one column.
struct Piece {
  int64_t c0, nc, r0, nr;
};
void pieces(int64_t rows, int64_t cols, size_t esz, std::vector<Piece> &out) {
  out.clear();
  if (rows <= 0 || cols <= 0) return;
  const int64_t per_blk = (int64_t)(kBlock / esz);
  if (rows <= per_blk) {
    const int64_t per = per_blk / rows;
    for (int64_t c0 = 0; c0 < cols; c0 += per) out.push_back({c0, std::min(per, cols - c0), 0, rows});
  } else {
    for (int64_t c = 0; c < cols; ++c)
      for (int64_t r0 = 0; r0 < rows; r0 += per_blk)
        out.push_back({c, 1, r0, std::min(per_blk, rows - r0)});
  }
}

// Host columns of piece p (block columns c0.., rows r0..) <-> pinned block.
void column_segs(const Piece &p, char *pinned, char *host, int64_t ld, size_t esz, bool to_pinned,
                 std::vector<Seg> &segs) {
  segs.clear();
  const size_t colb = (size_t)p.nr * esz;
  for (int64_t c = 0; c < p.nc; ++c) {
    char *h = host + ((p.c0 + c) * ld + p.r0) * (int64_t)esz;
    char *q = pinned + c * colb;
    for (size_t o = 0; o < colb; o += kSeg) {
      const size_t n = std::min(kSeg, colb - o);
      if (to_pinned) segs.push_back({q + o, h + o, n});
      else segs.push_back({h + o, q + o, n});
    }
  }
}

// Queue the upload of a rows x cols host block (leading dimension ld) into
// the dense device block dev, piece by piece, on stream s.
int h2d(char *dev, const char *host, int64_t rows, int64_t cols, int64_t ld, size_t esz,
        hipStream_t s) {
  Pool &g = g_pool;
  std::vector<Piece> ps;
  std::vector<Seg> segs;
  pieces(rows, cols, esz, ps);
  for (const Piece &p : ps) {
    const int b = acquire();
    if (b < 0) return MUBLIS_EHIP;
    column_segs(p, g.blk[b], const_cast<char *>(host), ld, esz, true, segs);
    g.team.run(segs);
    const size_t bytes = (size_t)p.nc * p.nr * esz;
    hipError_t e = hipMemcpyAsync(dev + (p.c0 * rows + p.r0) * (int64_t)esz, g.blk[b], bytes,
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(g.ev[b], s);
    if (e != hipSuccess) return fail(e, "pinned upload");
    g.pending[b] = true;
  }
  return MUBLIS_OK;
}

// A download queued into pinned blocks, scattered to the host by take().
struct Down {
  std::vector<Piece> ps;
  std::vector<int> blocks;
  char *host = nullptr;
  int64_t ld = 0;
  size_t esz = 0;
};

int queue_d2h(Down &d, char *host, const char *dev, int64_t rows, int64_t cols, int64_t ld,
              size_t esz, hipStream_t s) {
  Pool &g = g_pool;
  d.host = host;
  d.ld = ld;
  d.esz = esz;
  d.blocks.clear();
  pieces(rows, cols, esz, d.ps);
  for (const Piece &p : d.ps) {
    const int b = acquire();
    if (b < 0) return MUBLIS_EHIP;
    const size_t bytes = (size_t)p.nc * p.nr * esz;
    hipError_t e = hipMemcpyAsync(g.blk[b], dev + (p.c0 * rows + p.r0) * (int64_t)esz, bytes,
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(g.ev[b], s);
    if (e != hipSuccess) return fail(e, "pinned download");
    g.pending[b] = true;
    g.held[b] = true;
    d.blocks.push_back(b);
  }
  return MUBLIS_OK;
}

int take(Down &d) {
  Pool &g = g_pool;
  std::vector<Seg> segs;
  int st = MUBLIS_OK;
  for (size_t i = 0; i < d.blocks.size(); ++i) {
    const int b = d.blocks[i];
    if (st == MUBLIS_OK) {
      const hipError_t e = hipEventSynchronize(g.ev[b]);
      if (e != hipSuccess) st = fail(e, "pinned download wait");
    }
    g.pending[b] = false;
    if (st == MUBLIS_OK) {
      column_segs(d.ps[i], g.blk[b], d.host, d.ld, d.esz, false, segs);
      g.team.run(segs);
    }
    g.held[b] = false;
  }
  d.blocks.clear();
  return st;
}

// Releases whatever a failed call left held.
void drop(Down &d) {
  for (int b : d.blocks) g_pool.held[b] = false;
  d.blocks.clear();
}

// A view with strides (rs, cs) as a host block with contiguous rows: rows /
// cols / ld of the block, and the strides of its dense device copy.  A
// single column or row with a non-unit stride is a block of one-element
// columns.
bool as_block(int m, int n, int64_t rs, int64_t cs, int64_t *rows, int64_t *cols, int64_t *ld,
              int64_t *drs, int64_t *dcs) {
  if (rs == 1) {
    *rows = m; *cols = n; *ld = n == 1 ? m : cs; *drs = 1; *dcs = m;
  } else if (cs == 1) {
    *rows = n; *cols = m; *ld = m == 1 ? n : rs; *drs = n; *dcs = 1;
  } else if (n == 1) {
    *rows = 1; *cols = m; *ld = rs; *drs = 1; *dcs = m;
  } else if (m == 1) {
    *rows = 1; *cols = n; *ld = cs; *drs = 1; *dcs = 1;
  } else {
    return false;
  }
  return *ld >= *rows;
}

template <typename T>
int gemm_dev_call(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
                  const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
                  int64_t cs_c, hipStream_t s);
template <>
int gemm_dev_call<double>(int m, int n, int k, double alpha, const double *a, int64_t rs_a,
                          int64_t cs_a, const double *b, int64_t rs_b, int64_t cs_b,
                          double beta, double *c, int64_t rs_c, int64_t cs_c, hipStream_t s) {
  return mublis_gfx950_dgemm_dev(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                                 nullptr, 0, 0, beta, c, rs_c, cs_c, 0, s);
}
template <>
int gemm_dev_call<float>(int m, int n, int k, float alpha, const float *a, int64_t rs_a,
                         int64_t cs_a, const float *b, int64_t rs_b, int64_t cs_b, float beta,
                         float *c, int64_t rs_c, int64_t cs_c, hipStream_t s) {
  return mublis_gfx950_sgemm_dev(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, nullptr, 0, 0,
                                 nullptr, 0, 0, beta, c, rs_c, cs_c, 0, s);
}

template <typename T>
int gemm_host(int m, int n, int k, T alpha, const T *a, int64_t rs_a, int64_t cs_a,
              const T *b, int64_t rs_b, int64_t cs_b, T beta, T *c, int64_t rs_c,
              int64_t cs_c, void *stream) {
  constexpr size_t E = sizeof(T);
  if (m <= 0 || n <= 0 || k <= 0 || !a || !b || !c || rs_c != 1 || cs_c < m)
    return MUBLIS_EINVAL;
  hipStream_t s = (hipStream_t)(stream ? stream : mublis_rt_stream());
  const bool ha = !mublis_rt_is_device_ptr(a), hb = !mublis_rt_is_device_ptr(b);
  int64_t ar, ac, ald, adrs, adcs, br, bc, bld, bdrs, bdcs;
  if (ha && !as_block(m, k, rs_a, cs_a, &ar, &ac, &ald, &adrs, &adcs)) return MUBLIS_EINVAL;
  if (hb && !as_block(k, n, rs_b, cs_b, &br, &bc, &bld, &bdrs, &bdcs)) return MUBLIS_EINVAL;
  // B's columns are contiguous when rs_b == 1: its panels travel with C's
  const bool b_panels = hb && rs_b == 1;
  // panel width: a multiple of 128 columns (whole GEMM tiles), C's and B's
  // panel each within kPanelBlocks blocks, about 8 panels when n allows
  const int64_t tall = std::max<int64_t>(m, b_panels ? k : 1);
  int64_t jw = (int64_t)(kPanelBlocks * kBlock / E) / tall / 128 * 128;
  if (jw < 128) return MUBLIS_EINVAL;  // columns too tall for the pool
  jw = std::min<int64_t>(jw, std::max<int64_t>(128, (n / 8 + 127) / 128 * 128));
  std::lock_guard<std::mutex> lk(g_pool.mu);
  int st = build_pool();
  if (st) return st;
  Pool &g = g_pool;
  T *da = nullptr, *db = nullptr, *dc = nullptr;
  int64_t ra = rs_a, ca = cs_a, rb = rs_b, cb = cs_b;
  const T *pa = a, *pb = b;
  if (ha) da = (T *)mublis_rt_alloc(E * (size_t)m * k);
  if (hb) db = (T *)mublis_rt_alloc(E * (size_t)k * n);
  dc = (T *)mublis_rt_alloc(E * (size_t)m * n);
  hipEvent_t start = nullptr, up_done = nullptr, k_done = nullptr;
  std::vector<Down> downs(2);
  hipError_t e = hipSuccess;
  if ((ha && !da) || (hb && !db) || !dc) st = MUBLIS_ENOMEM;
  if (st == MUBLIS_OK) {
    e = hipEventCreateWithFlags(&start, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&up_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&k_done, hipEventDisableTiming);
    // the copy streams start after what the caller queued on s
    if (e == hipSuccess) e = hipEventRecord(start, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(g.up, start, 0);
    if (e == hipSuccess) e = hipStreamWaitEvent(g.dn, start, 0);
    if (e != hipSuccess) st = fail(e, "host GEMM events");
  }
  if (st == MUBLIS_OK && ha) {
    st = h2d((char *)da, (const char *)a, ar, ac, ald, E, g.up);
    pa = da; ra = adrs; ca = adcs;
  }
  if (st == MUBLIS_OK && hb) {
    if (!b_panels) st = h2d((char *)db, (const char *)b, br, bc, bld, E, g.up);
    pb = db; rb = bdrs; cb = bdcs;
  }
  int prev = -1;
  for (int64_t j0 = 0, p = 0; j0 < n && st == MUBLIS_OK; j0 += jw, ++p) {
    const int64_t w = std::min<int64_t>(jw, n - j0);
    if (b_panels)
      st = h2d((char *)(db + j0 * k), (const char *)(b + j0 * bld), k, w, bld, E, g.up);
    if (st == MUBLIS_OK && beta != T(0))
      st = h2d((char *)(dc + j0 * m), (const char *)(c + j0 * cs_c), m, w, cs_c, E, g.up);
    if (st != MUBLIS_OK) break;
    e = hipEventRecord(up_done, g.up);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, up_done, 0);
    if (e != hipSuccess) {
      st = fail(e, "host GEMM upload");
      break;
    }
    st = gemm_dev_call<T>(m, (int)w, k, alpha, pa, ra, ca, pb + j0 * cb, rb, cb, beta,
                          dc + j0 * m, 1, m, s);
    if (st != MUBLIS_OK) break;
    e = hipEventRecord(k_done, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(g.dn, k_done, 0);
    if (e != hipSuccess) {
      st = fail(e, "host GEMM kernel event");
      break;
    }
    const int cur = (int)(p & 1);
    st = queue_d2h(downs[cur], (char *)(c + j0 * cs_c), (const char *)(dc + j0 * m), m, w, cs_c,
                   E, g.dn);
    if (st == MUBLIS_OK && prev >= 0) st = take(downs[prev]);
    prev = cur;
  }
  if (st == MUBLIS_OK && prev >= 0) st = take(downs[prev]);
  for (Down &d : downs) drop(d);
  // nothing may still read the device buffers or pinned blocks
  if (hipStreamSynchronize(g.up) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  if (hipStreamSynchronize(g.dn) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  if (hipStreamSynchronize(s) != hipSuccess && st == MUBLIS_OK) st = MUBLIS_EHIP;
  for (int i = 0; i < kBlocks; ++i) g.pending[i] = false;
  if (start) (void)hipEventDestroy(start);
  if (up_done) (void)hipEventDestroy(up_done);
  if (k_done) (void)hipEventDestroy(k_done);
  mublis_rt_free(da);
  mublis_rt_free(db);
  mublis_rt_free(dc);
  return st;
}

}  // namespace

extern "C" int mublis_rt_h2d_2d(void *dev, const void *host, int64_t rows, int64_t cols,
                                int64_t ld, size_t esz, void *stream) {
  if (rows <= 0 || cols <= 0) return MUBLIS_OK;
  if (!dev || !host || esz == 0 || ld < rows) return MUBLIS_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_pool.mu);
  int st = build_pool();
  if (st == MUBLIS_OK) st = h2d((char *)dev, (const char *)host, rows, cols, ld, esz, s);
  return st;
}

extern "C" int mublis_rt_d2h_2d(void *host, const void *dev, int64_t rows, int64_t cols,
                                int64_t ld, size_t esz, void *stream) {
  if (rows <= 0 || cols <= 0) return MUBLIS_OK;
  if (!dev || !host || esz == 0 || ld < rows) return MUBLIS_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_pool.mu);
  int st = build_pool();
  if (st) return st;
  // pieces in flight at once: half the pool
  std::vector<Piece> ps;
  pieces(rows, cols, esz, ps);
  const size_t per = kBlocks / 2;
  for (size_t i0 = 0; i0 < ps.size() && st == MUBLIS_OK; i0 += per) {
    Down d;
    d.host = (char *)host;
    d.ld = ld;
    d.esz = esz;
    for (size_t i = i0; i < std::min(ps.size(), i0 + per) && st == MUBLIS_OK; ++i) {
      const Piece &p = ps[i];
      const int b = acquire();
      if (b < 0) {
        st = MUBLIS_EHIP;
        break;
      }
      const size_t bytes = (size_t)p.nc * p.nr * esz;
      hipError_t e = hipMemcpyAsync(g_pool.blk[b], (const char *)dev + (p.c0 * rows + p.r0) * (int64_t)esz,
                                    bytes, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipEventRecord(g_pool.ev[b], s);
      if (e != hipSuccess) {
        st = fail(e, "pinned download");
        break;
      }
      g_pool.pending[b] = true;
      g_pool.held[b] = true;
      d.ps.push_back(p);
      d.blocks.push_back(b);
    }
    if (st == MUBLIS_OK) st = take(d);
    drop(d);
  }
  return st;
}

extern "C" int mublis_gfx950_dgemm_host(int m, int n, int k, double alpha, const double *a,
                                        int64_t rs_a, int64_t cs_a, const double *b,
                                        int64_t rs_b, int64_t cs_b, double beta, double *c,
                                        int64_t rs_c, int64_t cs_c, void *stream) {
  return gemm_host<double>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, cs_c,
                           stream);
}

extern "C" int mublis_gfx950_sgemm_host(int m, int n, int k, float alpha, const float *a,
                                        int64_t rs_a, int64_t cs_a, const float *b,
                                        int64_t rs_b, int64_t cs_b, float beta, float *c,
                                        int64_t rs_c, int64_t cs_c, void *stream) {
  return gemm_host<float>(m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, cs_c,
                          stream);
}
