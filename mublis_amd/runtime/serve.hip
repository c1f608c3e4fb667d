// Serving loop: device-resident matrices by handle, admission against one
// HBM budget, same-precision grouping into one persistent launch, and
// continuous batching over several streams.
//
// No counterpart in the reference: its L3 entry points are synchronous
// per-thread calls (frame/l3/l3_driver.c:49-256).  Requests carry the
// arguments of the cblas_* calls they stand for (include/cblas.h:37-95).
//
// Threads: callers submit, wait and move data; one scheduler thread per
// server admits requests, issues launches and detects completions.  Every
// shared field is guarded by the server mutex; the scheduler drops the
// mutex only while it issues the launches it has just admitted, whose slots
// and matrices are already marked busy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "budget_arena.h"
#include "mublis_gfx950.h"

using mublis_serve::BudgetArena;
using Clock = std::chrono::steady_clock;

struct mublis_server_s;

struct mublis_matrix_s {
  mublis_server_s *srv = nullptr;
  int dtype = 0, rows = 0, cols = 0;
  int64_t ld = 0;
  size_t off = 0;
  char *ptr = nullptr;
  int users = 0;         // submitted requests not yet completed
  int writers = 0;       // of which write the matrix
  int busy = 0;          // launched requests not yet completed
  int busy_writers = 0;  // of which write the matrix
  int host_reading = 0, host_writing = 0;  // host copies in progress
  bool dead = false;     // destroyed by the caller, freed when users == 0
  size_t elem() const { return dtype == MUBLIS_F64 ? 8 : 4; }
};

namespace {

struct Req {
  uint64_t ticket = 0;
  mublis_request_t r{};
  Clock::time_point t_submit, t_done;
  int status = MUBLIS_OK;
  bool done = false;
  double flop = 0;
  mublis_matrix_s *rd[2] = {nullptr, nullptr};
  int nrd = 0;
  mublis_matrix_s *wr = nullptr;
  bool grouped = false;  // FP32 GEMM/SYRK/SYR2K
  bool selected = false;
};

struct Slot {
  int stream = 0;
  void *stage = nullptr;  // pinned host staging of a grouped work list
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<Req *> reqs;
  bool grouped = false;
  int ws_arena = -1;  // 0: workspace arena, 1: matrix arena, -1: none
  size_t ws_off = 0, ws_bytes = 0;
  int launch_status = MUBLIS_OK;
};

constexpr size_t kStageBytes = 8u << 20;

double req_flop(const mublis_request_t &r) {
  const double m = r.m, n = r.n, k = r.k;
  switch (r.op) {
    case MUBLIS_REQ_SGEMM:
    case MUBLIS_REQ_DGEMM: return 2.0 * m * n * k;
    case MUBLIS_REQ_SSYRK: return n * (n + 1) * k;
    case MUBLIS_REQ_SSYR2K: return 2.0 * n * (n + 1) * k;
    default: return r.side == MUBLIS_RIGHT ? m * n * n : m * m * n;
  }
}

bool fits(const mublis_matrix_s *h, int rows, int cols) {
  return h && rows <= h->rows && cols <= h->cols;
}

}  // namespace

struct mublis_server_s {
  int device = 0;
  mublis_server_config_t cfg{};
  char *base = nullptr;
  BudgetArena mat_arena, ws_arena;  // matrices | launch workspace
  size_t ws_base = 0;               // offset of the workspace arena in base
  std::vector<hipStream_t> streams;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_mem;
  std::deque<Req *> pending;
  std::unordered_map<uint64_t, Req *> reqs;  // submitted, not yet waited
  std::unordered_set<mublis_matrix_s *> matrices;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread sched;

  // statistics since the last reset
  mublis_server_stats_t st{};
  std::vector<double> lat_ms;
  bool any_submit = false;
  Clock::time_point first_submit, last_done;
  int inflight = 0;

  void free_matrix_locked(mublis_matrix_s *h) {
    mat_arena.release(h->off);
    matrices.erase(h);
    delete h;
    cv_mem.notify_all();
  }

  char *ws_ptr(const Slot &s) const {
    return base + (s.ws_arena == 0 ? ws_base : 0) + s.ws_off;
  }

  // Reserve launch workspace: the workspace arena first, then the matrix
  // arena.  Returns false when neither has a block that large.
  bool ws_alloc(size_t bytes, Slot &s) {
    s.ws_bytes = bytes;
    if (bytes == 0) { s.ws_arena = -1; return true; }
    size_t off = ws_arena.alloc(bytes);
    if (off != BudgetArena::kFail) { s.ws_arena = 0; s.ws_off = off; return true; }
    off = mat_arena.alloc(bytes);
    if (off != BudgetArena::kFail) { s.ws_arena = 1; s.ws_off = off; return true; }
    return false;
  }
  void ws_release(Slot &s) {
    if (s.ws_arena == 0) ws_arena.release(s.ws_off);
    if (s.ws_arena == 1) { mat_arena.release(s.ws_off); cv_mem.notify_all(); }
    s.ws_arena = -1;
  }

  void mark(Req *r, int d) {
    for (int i = 0; i < r->nrd; ++i) r->rd[i]->busy += d;
    r->wr->busy += d;
    r->wr->busy_writers += d;
  }

  using MatSet = std::unordered_set<const mublis_matrix_s *>;

  // A request may start when nothing launched or left waiting before it
  // writes what it reads, or touches what it writes.
  static bool conflicts(const Req *r, const MatSet &pw, const MatSet &pr) {
    const mublis_matrix_s *w = r->wr;
    if (w->busy || w->host_reading || w->host_writing || pw.count(w) || pr.count(w))
      return true;
    for (int i = 0; i < r->nrd; ++i) {
      const mublis_matrix_s *h = r->rd[i];
      if (h->busy_writers || h->host_writing || pw.count(h)) return true;
    }
    return false;
  }

  bool freeable(const mublis_matrix_s *h) const {
    return h->dead && h->users == 0 && !h->host_reading && !h->host_writing;
  }

  void complete_locked(Slot &s, int status) {
    const auto now = Clock::now();
    for (Req *r : s.reqs) {
      r->t_done = now;
      r->status = status;
      r->done = true;
      mark(r, -1);
      for (int i = 0; i < r->nrd; ++i) r->rd[i]->users--;
      r->wr->users--;
      r->wr->writers--;
      mublis_matrix_s *ms[3] = {r->wr, r->rd[0], r->rd[1]};
      for (mublis_matrix_s *h : ms)
        if (h && matrices.count(h) && freeable(h)) free_matrix_locked(h);
      if (status == MUBLIS_OK) {
        st.completed++;
        st.flop += r->flop;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
      } else {
        st.failed++;
      }
    }
    last_done = now;
    s.reqs.clear();
    ws_release(s);
    s.busy = false;
    inflight--;
    cv_done.notify_all();
  }

  // Completed launches; true if any.
  bool poll_locked() {
    bool any = false;
    for (Slot &s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) continue;
      complete_locked(s, q == hipSuccess ? s.launch_status : MUBLIS_EHIP);
      any = true;
    }
    return any;
  }

  // Admit what can start now into free slots.  Returns the slots to launch.
  std::vector<Slot *> admit_locked() {
    std::vector<Slot *> out;
    std::vector<Slot *> free_slots;
    {
      // spread over streams: prefer the stream with the fewest launches
      std::vector<int> load(streams.size(), 0);
      for (Slot &s : slots)
        if (s.busy) load[s.stream]++;
      for (Slot &s : slots)
        if (!s.busy) free_slots.push_back(&s);
      std::stable_sort(free_slots.begin(), free_slots.end(),
                       [&](Slot *a, Slot *b) { return load[a->stream] < load[b->stream]; });
    }
    if (free_slots.empty() || pending.empty()) return out;
    size_t next_slot = 0;
    MatSet pw, pr;  // touched by requests left waiting
    auto skip = [&](Req *r) {
      pw.insert(r->wr);
      for (int i = 0; i < r->nrd; ++i) pr.insert(r->rd[i]);
    };
    std::vector<Req *> group;
    Slot *group_slot = nullptr;
    const bool split = cfg.split_flop >= 0;
    for (Req *r : pending) {
      if (conflicts(r, pw, pr)) { skip(r); continue; }
      const bool joins = r->grouped && !(split && r->flop > cfg.split_flop);
      if (joins && group_slot && (int)group.size() < cfg.max_group) {
        group.push_back(r);
        mark(r, +1);
        r->selected = true;
        continue;
      }
      if (next_slot >= free_slots.size() || (joins && group_slot)) { skip(r); continue; }
      Slot *s = free_slots[next_slot];
      if (joins) {
        next_slot++;
        group_slot = s;
        group.push_back(r);
        mark(r, +1);
        r->selected = true;
        continue;
      }
      // a launch of its own: TRSM (with workspace), DGEMM, or a large FP32
      // request on the single-problem kernel
      // A TRSM gets the full schedule's workspace, so it runs the schedule of
      // the plain call (the same bits, alone or in any mix).
      size_t ws = 0;
      if (r->r.op == MUBLIS_REQ_STRSM || r->r.op == MUBLIS_REQ_DTRSM) {
        const bool right = r->r.side == MUBLIS_RIGHT;
        const int mm = right ? r->r.n : r->r.m, nn = right ? r->r.m : r->r.n;
        ws = r->r.op == MUBLIS_REQ_DTRSM ? mublis_gfx950_dtrsm_sched_ws_bytes(mm, nn)
                                         : mublis_gfx950_strsm_sched_ws_bytes(mm, nn);
      }
      if (!ws_alloc(ws, *s)) {
        if (inflight == 0 && out.empty() && ws_arena.used() == 0) {
          // can never fit: fail it rather than wait forever
          s->reqs = {r};
          r->selected = true;
          mark(r, +1);
          s->busy = true;
          inflight++;
          complete_locked(*s, MUBLIS_ENOMEM);
          continue;
        }
        skip(r);
        continue;
      }
      next_slot++;
      s->reqs = {r};
      s->grouped = false;
      mark(r, +1);
      r->selected = true;
      s->busy = true;
      inflight++;
      out.push_back(s);
    }
    if (group_slot) {
      // fit the work list into the slot's staging buffer and the budget,
      // dropping the tail of the group (it stays pending, in order)
      std::vector<mublis_sbatch_op_t> ops;
      for (Req *r : group) ops.push_back(to_op(r));
      while (!group.empty()) {
        const size_t ws = mublis_sgrouped_ws_bytes(ops.data(), (int)ops.size());
        if (ws <= kStageBytes && ws_alloc(ws, *group_slot)) break;
        if (group.size() == 1 && (ws > kStageBytes || (inflight == 0 && out.empty()))) {
          // cannot run even alone: fail it rather than wait forever
          group_slot->reqs = group;
          group_slot->busy = true;
          inflight++;
          complete_locked(*group_slot, MUBLIS_ENOMEM);
          group.clear();
          break;
        }
        const size_t keep = group.size() / 2;
        for (size_t i = keep; i < group.size(); ++i) {
          mark(group[i], -1);
          group[i]->selected = false;
        }
        group.resize(keep);
        ops.resize(keep);
      }
      if (!group.empty()) {
        group_slot->reqs = group;
        group_slot->grouped = true;
        group_slot->busy = true;
        inflight++;
        out.push_back(group_slot);
      }
    }
    std::deque<Req *> rest;
    for (Req *r : pending)
      if (!r->selected) rest.push_back(r);
    pending.swap(rest);
    return out;
  }

  static mublis_sbatch_op_t to_op(const Req *q) {
    const mublis_request_t &r = q->r;
    mublis_sbatch_op_t o{};
    o.op = r.op == MUBLIS_REQ_SGEMM ? MUBLIS_BATCH_GEMM
         : r.op == MUBLIS_REQ_SSYRK ? MUBLIS_BATCH_SYRK : MUBLIS_BATCH_SYR2K;
    o.trans_a = r.trans_a;
    o.trans_b = r.trans_b;
    o.uplo = r.uplo;
    o.m = r.op == MUBLIS_REQ_SGEMM ? r.m : r.n;
    o.n = r.n;
    o.k = r.k;
    o.alpha = (float)r.alpha;
    o.beta = (float)r.beta;
    o.a = (const float *)r.a->ptr;
    o.lda = (int)r.a->ld;
    if (r.b) { o.b = (const float *)r.b->ptr; o.ldb = (int)r.b->ld; }
    o.c = (float *)r.c->ptr;
    o.ldc = (int)r.c->ld;
    return o;
  }

  // Issue one admitted launch (scheduler thread, mutex not held).
  int launch(Slot &s) {
    hipStream_t stream = streams[s.stream];
    int status = MUBLIS_OK;
    if (s.grouped) {
      std::vector<mublis_sbatch_op_t> ops;
      for (Req *r : s.reqs) ops.push_back(to_op(r));
      status = mublis_sgrouped_launch_ws(ops.data(), (int)ops.size(), s.stage, ws_ptr(s),
                                         s.ws_bytes, stream);
    } else {
      const mublis_request_t &r = s.reqs[0]->r;
      if (s.reqs[0]->grouped) {
        // op(A) is m x k; SYRK pairs it with its own transpose, SYR2K adds
        // the swapped product as a second K segment (frame/l3/syrk.c:43-51,
        // frame/l3/syr2k.c:47-68)
        const bool sym = r.op != MUBLIS_REQ_SGEMM;
        const int m = sym ? r.n : r.m;
        const float *a = (const float *)r.a->ptr, *b = r.b ? (const float *)r.b->ptr : nullptr;
        int64_t rsa = 1, csa = r.a->ld, rsb = 1, csb = r.b ? r.b->ld : 0;
        if (r.trans_a) std::swap(rsa, csa);
        const float *a2 = nullptr, *b2 = nullptr;
        int64_t rsa2 = 0, csa2 = 0, rsb2 = 0, csb2 = 0;
        int ji = 0;
        if (!sym) {
          if (r.trans_b) std::swap(rsb, csb);
        } else {
          ji = r.uplo == MUBLIS_UPPER ? 2 : 1;
          if (r.op == MUBLIS_REQ_SSYRK) {
            b = a; rsb = csa; csb = rsa;
          } else {
            if (r.trans_a) std::swap(rsb, csb);
            a2 = b; rsa2 = rsb; csa2 = csb;
            b2 = a; rsb2 = csa; csb2 = rsa;
            std::swap(rsb, csb);
          }
        }
        status = mublis_gfx950_sgemm_dev(m, r.n, r.k, (float)r.alpha, a, rsa, csa, b, rsb, csb,
                                         a2, rsa2, csa2, b2, rsb2, csb2, (float)r.beta,
                                         (float *)r.c->ptr, 1, r.c->ld, ji, stream);
      } else if (r.op == MUBLIS_REQ_DGEMM) {
        int64_t rsa = 1, csa = r.a->ld, rsb = 1, csb = r.b->ld;
        if (r.trans_a) std::swap(rsa, csa);
        if (r.trans_b) std::swap(rsb, csb);
        status = mublis_gfx950_dgemm_dev(r.m, r.n, r.k, r.alpha, (const double *)r.a->ptr, rsa,
                                         csa, (const double *)r.b->ptr, rsb, csb, nullptr, 0,
                                         0, nullptr, 0, 0, r.beta, (double *)r.c->ptr, 1,
                                         r.c->ld, 0, stream);
      } else {
        // Right side: X*op(A) = alpha*B  <=>  op(A)'*X' = alpha*B'
        // (frame/l3/trsm.c:84-109); a transposed triangle flips uplo.
        const bool right = r.side == MUBLIS_RIGHT;
        const bool eff_t = (r.trans_a != 0) != right;
        const bool lower = (r.uplo == MUBLIS_LOWER) != eff_t;
        const int64_t la = r.a->ld, lb = r.b->ld;
        const int64_t rsa = eff_t ? la : 1, csa = eff_t ? 1 : la;
        const int64_t rsb = right ? lb : 1, csb = right ? 1 : lb;
        const int mm = right ? r.n : r.m, nn = right ? r.m : r.n;
        const int unit = r.diag == MUBLIS_UNIT_DIAG;
        if (r.op == MUBLIS_REQ_DTRSM)
          status = mublis_gfx950_dtrsm_ws_dev(lower, unit, mm, nn, r.alpha,
                                              (const double *)r.a->ptr, rsa, csa,
                                              (double *)r.b->ptr, rsb, csb, ws_ptr(s),
                                              s.ws_bytes, stream);
        else
          status = mublis_gfx950_strsm_ws_dev(lower, unit, mm, nn, (float)r.alpha,
                                              (const float *)r.a->ptr, rsa, csa,
                                              (float *)r.b->ptr, rsb, csb, ws_ptr(s),
                                              s.ws_bytes, stream);
      }
    }
    s.launch_status = status;
    if (hipEventRecord(s.ev, stream) != hipSuccess) return MUBLIS_EHIP;
    return MUBLIS_OK;
  }

  void loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      const bool progressed = poll_locked();
      std::vector<Slot *> todo = admit_locked();
      if ((int)st.max_inflight < inflight) st.max_inflight = inflight;
      if (!todo.empty()) {
        st.launches += todo.size();
        for (Slot *s : todo)
          if (s->grouped) {
            st.grouped_launches++;
            st.grouped_requests += s->reqs.size();
          }
        lk.unlock();
        std::vector<int> fail(todo.size(), MUBLIS_OK);
        for (size_t i = 0; i < todo.size(); ++i) fail[i] = launch(*todo[i]);
        lk.lock();
        for (size_t i = 0; i < todo.size(); ++i)
          if (fail[i] != MUBLIS_OK) complete_locked(*todo[i], fail[i]);
        continue;
      }
      if (stop && pending.empty() && inflight == 0) break;
      if (inflight == 0)
        cv_sched.wait(lk);
      else if (!progressed)
        cv_sched.wait_for(lk, std::chrono::microseconds(20));
    }
  }
};

namespace {

int validate(const mublis_request_t &r, Req &q) {
  const bool f64 = r.op == MUBLIS_REQ_DGEMM || r.op == MUBLIS_REQ_DTRSM;
  const int dt = f64 ? MUBLIS_F64 : MUBLIS_F32;
  if (r.m < 0 || r.n < 0 || r.k < 0) return MUBLIS_EINVAL;
  auto ok_dt = [&](const mublis_matrix_s *h) { return h && !h->dead && h->dtype == dt; };
  switch (r.op) {
    case MUBLIS_REQ_SGEMM:
    case MUBLIS_REQ_DGEMM: {
      if (!ok_dt(r.a) || !ok_dt(r.b) || !ok_dt(r.c)) return MUBLIS_EINVAL;
      if (!fits(r.c, r.m, r.n)) return MUBLIS_EINVAL;
      if (!(r.trans_a ? fits(r.a, r.k, r.m) : fits(r.a, r.m, r.k))) return MUBLIS_EINVAL;
      if (!(r.trans_b ? fits(r.b, r.n, r.k) : fits(r.b, r.k, r.n))) return MUBLIS_EINVAL;
      if (r.c == r.a || r.c == r.b) return MUBLIS_EINVAL;
      q.rd[0] = r.a;
      q.rd[1] = r.b == r.a ? nullptr : r.b;
      q.nrd = q.rd[1] ? 2 : 1;
      q.wr = r.c;
      q.grouped = r.op == MUBLIS_REQ_SGEMM;
      return MUBLIS_OK;
    }
    case MUBLIS_REQ_SSYRK:
    case MUBLIS_REQ_SSYR2K: {
      const bool two = r.op == MUBLIS_REQ_SSYR2K;
      if (!ok_dt(r.a) || !ok_dt(r.c) || (two && !ok_dt(r.b))) return MUBLIS_EINVAL;
      if (r.uplo != MUBLIS_LOWER && r.uplo != MUBLIS_UPPER) return MUBLIS_EINVAL;
      if (!fits(r.c, r.n, r.n)) return MUBLIS_EINVAL;
      auto op_fits = [&](const mublis_matrix_s *h) {
        return r.trans_a ? fits(h, r.k, r.n) : fits(h, r.n, r.k);
      };
      if (!op_fits(r.a) || (two && !op_fits(r.b))) return MUBLIS_EINVAL;
      if (r.c == r.a || (two && r.c == r.b)) return MUBLIS_EINVAL;
      q.rd[0] = r.a;
      q.rd[1] = two && r.b != r.a ? r.b : nullptr;
      q.nrd = q.rd[1] ? 2 : 1;
      q.wr = r.c;
      q.grouped = true;
      return MUBLIS_OK;
    }
    case MUBLIS_REQ_STRSM:
    case MUBLIS_REQ_DTRSM: {
      if (!ok_dt(r.a) || !ok_dt(r.b) || r.a == r.b) return MUBLIS_EINVAL;
      if (r.uplo != MUBLIS_LOWER && r.uplo != MUBLIS_UPPER) return MUBLIS_EINVAL;
      if (r.side != MUBLIS_LEFT && r.side != MUBLIS_RIGHT) return MUBLIS_EINVAL;
      const int na = r.side == MUBLIS_RIGHT ? r.n : r.m;
      if (!fits(r.a, na, na) || !fits(r.b, r.m, r.n)) return MUBLIS_EINVAL;
      q.rd[0] = r.a;
      q.nrd = 1;
      q.wr = r.b;
      q.grouped = false;
      return MUBLIS_OK;
    }
    default: return MUBLIS_EINVAL;
  }
}

}  // namespace

extern "C" {

int mublis_server_create(const mublis_server_config_t *cfg, mublis_server_t *out) {
  if (!out) return MUBLIS_EINVAL;
  *out = nullptr;
  int st = mublis_rt_init();
  if (st) return st;
  auto *s = new mublis_server_s();
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.hbm_budget) s->cfg.hbm_budget = size_t(1) << 30;
  if (s->cfg.streams <= 0) s->cfg.streams = 4;
  if (s->cfg.slots_per_stream <= 0) s->cfg.slots_per_stream = 4;
  if (s->cfg.max_group <= 0) s->cfg.max_group = 256;
  if (s->cfg.split_flop == 0) s->cfg.split_flop = 2e10;
  (void)hipGetDevice(&s->device);
  // workspace reserve: an eighth of the budget, at most 256 MiB
  const size_t budget = s->cfg.hbm_budget / BudgetArena::kAlign * BudgetArena::kAlign;
  const size_t reserve =
      std::min<size_t>(budget / 8, size_t(256) << 20) / BudgetArena::kAlign * BudgetArena::kAlign;
  st = MUBLIS_OK;
  if (budget < (size_t(4) << 20) || hipMalloc(&s->base, budget) != hipSuccess) {
    s->base = nullptr;
    st = MUBLIS_ENOMEM;
  }
  if (!st) {
    s->mat_arena.reset(budget - reserve);
    s->ws_arena.reset(reserve);
    s->ws_base = budget - reserve;
    s->st.hbm_budget = budget;
  }
  for (int i = 0; !st && i < s->cfg.streams; ++i) {
    hipStream_t h;
    if (hipStreamCreateWithFlags(&h, hipStreamDefault) != hipSuccess) st = MUBLIS_EHIP;
    else s->streams.push_back(h);
  }
  const int nslots = s->cfg.streams * s->cfg.slots_per_stream;
  s->slots.resize(st ? 0 : nslots);
  for (int i = 0; !st && i < nslots; ++i) {
    Slot &sl = s->slots[i];
    sl.stream = i % s->cfg.streams;
    if (hipHostMalloc(&sl.stage, kStageBytes, hipHostMallocDefault) != hipSuccess) {
      sl.stage = nullptr;
      st = MUBLIS_ENOMEM;
    } else if (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
      sl.ev = nullptr;
      st = MUBLIS_EHIP;
    }
  }
  if (st) {
    for (Slot &sl : s->slots) {
      if (sl.stage) (void)hipHostFree(sl.stage);
      if (sl.ev) (void)hipEventDestroy(sl.ev);
    }
    for (hipStream_t h : s->streams) (void)hipStreamDestroy(h);
    if (s->base) (void)hipFree(s->base);
    delete s;
    return st;
  }
  s->sched = std::thread([s] { s->loop(); });
  *out = s;
  return MUBLIS_OK;
}

int mublis_server_destroy(mublis_server_t s) {
  if (!s) return MUBLIS_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  s->sched.join();
  for (auto &kv : s->reqs) delete kv.second;
  for (mublis_matrix_s *h : s->matrices) delete h;
  for (Slot &sl : s->slots) {
    (void)hipHostFree(sl.stage);
    (void)hipEventDestroy(sl.ev);
  }
  for (hipStream_t h : s->streams) (void)hipStreamDestroy(h);
  const int st = hipFree(s->base) == hipSuccess ? MUBLIS_OK : MUBLIS_EHIP;
  delete s;
  return st;
}

int mublis_matrix_create(mublis_server_t s, int dtype, int rows, int cols, int wait,
                         mublis_matrix_t *out) {
  if (!s || !out || rows < 0 || cols < 0 || (dtype != MUBLIS_F32 && dtype != MUBLIS_F64))
    return MUBLIS_EINVAL;
  *out = nullptr;
  const size_t es = dtype == MUBLIS_F64 ? 8 : 4;
  const int64_t per = 64 / es;  // leading dimension padded to 64 bytes
  const int64_t ld = std::max<int64_t>(per, (rows + per - 1) / per * per);
  const size_t bytes = (size_t)ld * (size_t)std::max(cols, 1) * es;
  std::unique_lock<std::mutex> lk(s->mu);
  if (BudgetArena::round(bytes) > s->mat_arena.capacity()) return MUBLIS_ENOMEM;
  size_t off;
  while ((off = s->mat_arena.alloc(bytes)) == BudgetArena::kFail) {
    if (!wait) return MUBLIS_ENOMEM;
    s->cv_mem.wait(lk);
  }
  auto *h = new mublis_matrix_s();
  h->srv = s;
  h->dtype = dtype;
  h->rows = rows;
  h->cols = cols;
  h->ld = ld;
  h->off = off;
  h->ptr = s->base + off;
  s->matrices.insert(h);
  *out = h;
  return MUBLIS_OK;
}

static int host_copy(mublis_server_t s, mublis_matrix_t h, const void *src, void *dst,
                     int64_t ld, bool to_device) {
  if (!s || !h || h->srv != s || (!src && !dst) || ld < std::max(1, h->rows))
    return MUBLIS_EINVAL;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    if (h->dead) return MUBLIS_EINVAL;
    if (to_device) {
      s->cv_done.wait(lk, [&] { return h->users == 0 && !h->host_reading && !h->host_writing; });
      h->host_writing++;
    } else {
      s->cv_done.wait(lk, [&] { return h->writers == 0 && !h->host_writing; });
      h->host_reading++;
    }
  }
  hipStream_t stream = (hipStream_t)mublis_rt_stream();
  const size_t es = h->elem();
  int st = MUBLIS_OK;
  if (h->rows && h->cols) {
    hipError_t e =
        to_device ? hipMemcpy2DAsync(h->ptr, h->ld * es, src, ld * es, h->rows * es, h->cols,
                                     hipMemcpyHostToDevice, stream)
                  : hipMemcpy2DAsync(dst, ld * es, h->ptr, h->ld * es, h->rows * es, h->cols,
                                     hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) st = MUBLIS_EHIP;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (to_device) h->host_writing--;
    else h->host_reading--;
    if (s->freeable(h)) s->free_matrix_locked(h);
  }
  s->cv_done.notify_all();
  s->cv_sched.notify_all();
  return st;
}

int mublis_matrix_write(mublis_server_t s, mublis_matrix_t h, const void *host, int64_t ld) {
  return host ? host_copy(s, h, host, nullptr, ld, true) : MUBLIS_EINVAL;
}

int mublis_matrix_read(mublis_server_t s, mublis_matrix_t h, void *host, int64_t ld) {
  return host ? host_copy(s, h, nullptr, host, ld, false) : MUBLIS_EINVAL;
}

void *mublis_matrix_device_ptr(mublis_matrix_t h) { return h ? h->ptr : nullptr; }

int64_t mublis_matrix_ld(mublis_matrix_t h) { return h ? h->ld : 0; }

int mublis_matrix_destroy(mublis_server_t s, mublis_matrix_t h) {
  if (!s || !h) return MUBLIS_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->matrices.count(h) || h->dead) return MUBLIS_EINVAL;
  h->dead = true;
  if (s->freeable(h)) s->free_matrix_locked(h);
  return MUBLIS_OK;
}

int mublis_submit(mublis_server_t s, const mublis_request_t *req, uint64_t *ticket) {
  if (!s || !req || !ticket) return MUBLIS_EINVAL;
  auto *q = new Req();
  q->r = *req;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const mublis_matrix_s *hs[3] = {req->a, req->b, req->c};
    for (const mublis_matrix_s *h : hs)
      if (h && !s->matrices.count((mublis_matrix_s *)h)) {
        delete q;
        return MUBLIS_EINVAL;
      }
    const int st = validate(*req, *q);
    if (st) {
      delete q;
      return st;
    }
    q->flop = req_flop(*req);
    q->ticket = s->next_ticket++;
    q->t_submit = Clock::now();
    if (!s->any_submit) {
      s->any_submit = true;
      s->first_submit = q->t_submit;
    }
    for (int i = 0; i < q->nrd; ++i) q->rd[i]->users++;
    q->wr->users++;
    q->wr->writers++;
    s->pending.push_back(q);
    s->reqs[q->ticket] = q;
    s->st.submitted++;
    if (s->pending.size() > s->st.max_pending) s->st.max_pending = s->pending.size();
    *ticket = q->ticket;
  }
  s->cv_sched.notify_one();
  return MUBLIS_OK;
}

int mublis_server_replay(mublis_server_t s, const mublis_request_t *reqs,
                         const double *arrival_s, int n, uint64_t *tickets, int *submitted,
                         double *max_lag_ms) {
  if (submitted) *submitted = 0;
  if (max_lag_ms) *max_lag_ms = 0;
  if (!s || n < 0 || (n > 0 && (!reqs || !arrival_s || !tickets))) return MUBLIS_EINVAL;
  // offsets must be finite, from 0 to a day (the clock arithmetic below
  // would overflow on larger ones)
  for (int i = 0; i < n; ++i)
    if (!(arrival_s[i] >= 0.0 && arrival_s[i] <= 86400.0)) return MUBLIS_EINVAL;
  using SClock = std::chrono::steady_clock;
  const SClock::time_point start = SClock::now();
  double lag = 0;
  for (int i = 0; i < n; ++i) {
    const SClock::time_point t =
        start + std::chrono::duration_cast<SClock::duration>(
                    std::chrono::duration<double>(arrival_s[i]));
    // sleep to within 50 us of the arrival (the scheduler's wake-up
    // granularity), then yield until it
    const SClock::time_point near = t - std::chrono::microseconds(50);
    if (SClock::now() < near) std::this_thread::sleep_until(near);
    while (SClock::now() < t) std::this_thread::yield();
    const int st = mublis_submit(s, &reqs[i], &tickets[i]);
    if (st) return st;
    if (submitted) *submitted = i + 1;
    lag = std::max(lag, std::chrono::duration<double, std::milli>(SClock::now() - t).count());
  }
  if (max_lag_ms) *max_lag_ms = lag;
  return MUBLIS_OK;
}

int mublis_wait(mublis_server_t s, uint64_t ticket, double *latency_ms) {
  if (!s) return MUBLIS_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MUBLIS_EINVAL;
  Req *q = it->second;
  s->cv_done.wait(lk, [&] { return q->done; });
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
  const int st = q->status;
  s->reqs.erase(ticket);
  delete q;
  return st;
}

int mublis_server_stats(mublis_server_t s, mublis_server_stats_t *out) {
  if (!s || !out) return MUBLIS_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  std::vector<double> v = s->lat_ms;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      const double x = p * (v.size() - 1);
      const size_t i = (size_t)x;
      const double f = x - i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_ms = sum / v.size();
  }
  if (s->any_submit && out->completed)
    out->elapsed_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
  out->tflops = out->elapsed_s > 0 ? out->flop / out->elapsed_s / 1e12 : 0;
  out->hbm_used = s->mat_arena.used() + s->ws_arena.used();
  out->hbm_peak = s->mat_arena.peak() + s->ws_arena.peak();
  return MUBLIS_OK;
}

int mublis_server_reset_stats(mublis_server_t s) {
  if (!s) return MUBLIS_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  const size_t budget = s->st.hbm_budget;
  s->st = mublis_server_stats_t{};
  s->st.hbm_budget = budget;
  s->lat_ms.clear();
  s->any_submit = false;
  s->mat_arena.reset_peak();
  s->ws_arena.reset_peak();
  return MUBLIS_OK;
}

}  // extern "C"
