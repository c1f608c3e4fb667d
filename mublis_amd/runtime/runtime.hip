// Host runtime for the gfx950 L3 engine: device selection, per-thread
// streams, a caching device-memory arena and staging copies.
//
// The reference keeps a process-lifetime, mutex-guarded freelist of CPU pack
// buffers and a pthread_once-cached context (frame/base/pool.c:120-318,
// frame/base/safe_dispatch.c:45-71).  On the GPU the same roles are:
//   * init once per process, cache success or failure;
//   * one HIP stream per calling thread (default priority, blocking with
//     respect to the legacy default stream), recycled through a free list
//     when the thread exits, never destroyed;
//   * a caching arena of device blocks keyed by rounded size, under one
//     mutex, never returned to HIP except to retry a failed hipMalloc.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "mublis_gfx950.h"

namespace {

std::once_flag g_init_once;
int g_init_status = MUBLIS_ECONTEXT;
int g_device = 0;
char g_arch[64] = "";

thread_local char t_last_error[256] = "";

void record(hipError_t e, const char *what) {
  snprintf(t_last_error, sizeof t_last_error, "%s: %s", what,
           hipGetErrorString(e));
}

void do_init() {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    record(e == hipSuccess ? hipErrorNoDevice : e, "hipGetDeviceCount");
    g_init_status = MUBLIS_ECONTEXT;
    return;
  }
  e = hipGetDevice(&g_device);
  if (e != hipSuccess) {
    record(e, "hipGetDevice");
    g_init_status = MUBLIS_ECONTEXT;
    return;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, g_device);
  if (e != hipSuccess) {
    record(e, "hipGetDeviceProperties");
    g_init_status = MUBLIS_ECONTEXT;
    return;
  }
  // gcnArchName looks like "gfx950:sramecc+:xnack-".
  size_t n = strcspn(prop.gcnArchName, ":");
  if (n >= sizeof g_arch) n = sizeof g_arch - 1;
  memcpy(g_arch, prop.gcnArchName, n);
  g_arch[n] = 0;
  if (strcmp(g_arch, "gfx950") != 0) {
    snprintf(t_last_error, sizeof t_last_error,
             "device %d is %s; this library is built for gfx950", g_device,
             g_arch);
    g_init_status = MUBLIS_ECONTEXT;
    return;
  }
  g_init_status = MUBLIS_OK;
}

// ----------------------------------------------------------- streams ----
std::mutex g_stream_mu;
std::vector<hipStream_t> g_free_streams;

struct ThreadStream {
  hipStream_t s = nullptr;
  ~ThreadStream() {
    if (s) {
      std::lock_guard<std::mutex> lk(g_stream_mu);
      g_free_streams.push_back(s);
    }
  }
};
thread_local ThreadStream t_stream;

// ------------------------------------------------------------- arena ----
std::mutex g_arena_mu;
std::multimap<size_t, void *> g_free_blocks;  // size -> block
std::map<void *, size_t> g_block_size;         // live + cached blocks

size_t round_size(size_t b) {
  if (b < 256) return 256;
  if (b <= (size_t(1) << 20)) {  // powers of two up to 1 MiB
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
  }
  const size_t g = size_t(2) << 20;  // 2 MiB granules above
  return (b + g - 1) / g * g;
}

}  // namespace

extern "C" {

int mublis_rt_init(void) {
  std::call_once(g_init_once, do_init);
  if (g_init_status == MUBLIS_OK && hipSetDevice(g_device) != hipSuccess)
    return MUBLIS_EHIP;
  return g_init_status;
}

const char *mublis_rt_arch(void) { return g_arch; }

const char *mublis_rt_last_error(void) { return t_last_error; }

}  // extern "C"

// The calling thread's mublis_rt_last_error() text, for argument errors
// found inside the HIP layer.
void gfx950_set_error(const char *msg) {
  snprintf(t_last_error, sizeof t_last_error, "%s", msg);
}

extern "C" {

int mublis_rt_is_device_ptr(const void *p) {
  if (!p) return 0;
  hipPointerAttribute_t attr;
  memset(&attr, 0, sizeof attr);
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: clear and stage
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void *mublis_rt_stream(void) {
  if (!t_stream.s) {
    {
      std::lock_guard<std::mutex> lk(g_stream_mu);
      if (!g_free_streams.empty()) {
        t_stream.s = g_free_streams.back();
        g_free_streams.pop_back();
      }
    }
    if (!t_stream.s) {
      // A blocking stream: work the caller queued on the legacy default
      // stream (how most callers produce device operands) completes before
      // this stream's kernels start, which the CBLAS ABI -- it has no stream
      // argument -- needs for device pointers.
      hipError_t e = hipStreamCreateWithFlags(&t_stream.s, hipStreamDefault);
      if (e != hipSuccess) {
        record(e, "hipStreamCreateWithFlags");
        t_stream.s = nullptr;
      }
    }
  }
  return t_stream.s;
}

void *mublis_rt_alloc(size_t bytes) {
  size_t sz = round_size(bytes);
  std::lock_guard<std::mutex> lk(g_arena_mu);
  auto it = g_free_blocks.lower_bound(sz);
  // Reuse a cached block of the same size, or one at most 2x larger.
  if (it != g_free_blocks.end() && it->first <= 2 * sz) {
    void *p = it->second;
    g_free_blocks.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // Give cached blocks back to HIP and retry once.
    for (auto &kv : g_free_blocks) {
      (void)hipFree(kv.second);
      g_block_size.erase(kv.second);
    }
    g_free_blocks.clear();
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      record(e, "hipMalloc");
      return nullptr;
    }
  }
  g_block_size[p] = sz;
  return p;
}

void mublis_rt_free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_arena_mu);
  auto it = g_block_size.find(p);
  if (it == g_block_size.end()) return;
  g_free_blocks.emplace(it->second, p);
}

int mublis_rt_h2d(void *dst, const void *src, size_t bytes, void *stream) {
  if (!bytes) return MUBLIS_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                (hipStream_t)stream);
  if (e != hipSuccess) {
    record(e, "hipMemcpyAsync H2D");
    return MUBLIS_EHIP;
  }
  return MUBLIS_OK;
}

int mublis_rt_d2h(void *dst, const void *src, size_t bytes, void *stream) {
  if (!bytes) return MUBLIS_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                                (hipStream_t)stream);
  if (e != hipSuccess) {
    record(e, "hipMemcpyAsync D2H");
    return MUBLIS_EHIP;
  }
  return MUBLIS_OK;
}

int mublis_rt_sync(void *stream) {
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) {
    record(e, "hipStreamSynchronize");
    return MUBLIS_EHIP;
  }
  return MUBLIS_OK;
}

}  // extern "C"
