// Internal declarations shared by the HIP translation units of libxs_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "xs_hip.h"

namespace xs {

// Thread-local error text behind xs_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define XS_HIP_CHECK(call)                                                     \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::xs::fail(XS_ERR_HIP, std::string(#call) + ": " +               \
                                        hipGetErrorString(e_));                \
  } while (0)

#define XS_TRY(call)                                                           \
  do {                                                                         \
    int r_ = (call);                                                           \
    if (r_ != XS_OK) return r_;                                                \
  } while (0)

inline unsigned ntypes_of(unsigned n_el) { return n_el * (n_el + 1) / 2; }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Event pair around one kernel launch, resolved lazily by xs_profile_read.
struct ProfRecord {
  std::string name;
  hipEvent_t start, stop;
};

}  // namespace xs

struct xs_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  size_t max_lds = 0;

  // Device workspace carved per call (grown on demand, never shrunk).
  void *ws = nullptr;
  size_t ws_size = 0;
  // Pinned host staging for small uploads (item lists, parameter tables).
  void *pin = nullptr;
  size_t pin_size = 0;

  // Cached device-side work list of the histogram kernel.
  std::vector<uint32_t> hist_items_key;
  void *hist_items_dev = nullptr;
  size_t hist_items_cap = 0;
  size_t hist_items_n = 0;

  bool profiling = false;
  std::vector<xs::ProfRecord> prof_pending;
  std::map<std::string, std::pair<double, unsigned long long>> prof_done;
  std::vector<hipEvent_t> event_pool;
};

namespace xs {

// Make the context's device current for this thread.
int activate(xs_ctx *ctx);
// Ensure ctx->ws holds at least `bytes` (synchronises the stream if it grows).
int ensure_workspace(xs_ctx *ctx, size_t bytes);
// Ensure ctx->pin holds at least `bytes` (synchronises the stream if it grows).
int ensure_pinned(xs_ctx *ctx, size_t bytes);

// Profiling: wrap a launch with events when collection is on.
int prof_begin(xs_ctx *ctx, const char *name, hipEvent_t *start);
int prof_end(xs_ctx *ctx, const char *name, hipEvent_t start);

// Bump allocator over the workspace for one call.
struct Carver {
  char *base;
  size_t used = 0;
  explicit Carver(void *b) : base(static_cast<char *>(b)) {}
  template <class T>
  T *take(size_t count) {
    used = align_up(used, 256);
    T *p = reinterpret_cast<T *>(base + used);
    used += count * sizeof(T);
    return p;
  }
  static size_t need(std::initializer_list<size_t> sizes) {
    size_t s = 0;
    for (size_t b : sizes) s = align_up(s, 256) + b;
    return align_up(s, 256);
  }
};

}  // namespace xs
