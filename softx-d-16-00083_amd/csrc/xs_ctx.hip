// Context, workspace, error reporting and event profiling of libxs_hip.
#include <cstring>

#include "xs_internal.h"

namespace xs {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

int activate(xs_ctx *ctx) {
  if (!ctx) return fail(XS_ERR_ARG, "null xs_ctx");
  XS_HIP_CHECK(hipSetDevice(ctx->device));
  return XS_OK;
}

// Grow a device buffer (after draining the stream that may still read it).
static int ensure_device(xs_ctx *ctx, void **buf, size_t *cap, size_t bytes, size_t gran,
                         const char *what) {
  if (bytes <= *cap) return XS_OK;
  size_t want = align_up(bytes + bytes / 4, gran);
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (*buf) XS_HIP_CHECK(hipFree(*buf));
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc(buf, want) != hipSuccess) {
    *buf = nullptr;
    (void)hipGetLastError();
    return fail(XS_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " " + what + " bytes failed");
  }
  *cap = want;
  return XS_OK;
}

int ensure_workspace(xs_ctx *ctx, size_t bytes) {
  return ensure_device(ctx, &ctx->ws, &ctx->ws_size, bytes, size_t(1) << 20, "workspace");
}

int ensure_aux(xs_ctx *ctx, size_t bytes) {
  return ensure_device(ctx, &ctx->aux, &ctx->aux_size, bytes, size_t(1) << 16, "table");
}

int ensure_matrix(xs_ctx *ctx, size_t bytes) {
  return ensure_device(ctx, &ctx->mat, &ctx->mat_size, bytes, size_t(1) << 20, "matrix");
}

int ensure_io(xs_ctx *ctx, size_t bytes) {
  return ensure_device(ctx, &ctx->io, &ctx->io_size, bytes, size_t(1) << 20, "host-copy");
}

int ensure_pinned(xs_ctx *ctx, size_t bytes) {
  if (bytes <= ctx->pin_size) return XS_OK;
  size_t want = align_up(bytes + bytes / 4, size_t(1) << 16);
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->pin) XS_HIP_CHECK(hipHostFree(ctx->pin));
  ctx->pin = nullptr;
  ctx->pin_size = 0;
  if (hipHostMalloc(&ctx->pin, want, hipHostMallocDefault) != hipSuccess) {
    ctx->pin = nullptr;
    (void)hipGetLastError();
    return fail(XS_ERR_NOMEM, "hipHostMalloc of " + std::to_string(want) + " bytes failed");
  }
  ctx->pin_size = want;
  return XS_OK;
}

static int get_event(xs_ctx *ctx, hipEvent_t *ev) {
  if (!ctx->event_pool.empty()) {
    *ev = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return XS_OK;
  }
  XS_HIP_CHECK(hipEventCreate(ev));
  return XS_OK;
}

int prof_begin(xs_ctx *ctx, const char *name, hipEvent_t *start) {
  (void)name;
  *start = nullptr;
  if (!ctx->profiling) return XS_OK;
  XS_TRY(get_event(ctx, start));
  XS_HIP_CHECK(hipEventRecord(*start, ctx->stream));
  return XS_OK;
}

int prof_end(xs_ctx *ctx, const char *name, hipEvent_t start) {
  if (!ctx->profiling || !start) return XS_OK;
  hipEvent_t stop;
  XS_TRY(get_event(ctx, &stop));
  XS_HIP_CHECK(hipEventRecord(stop, ctx->stream));
  ctx->prof_pending.push_back(ProfRecord{name, start, stop});
  return XS_OK;
}

static int prof_resolve(xs_ctx *ctx) {
  for (auto &r : ctx->prof_pending) {
    XS_HIP_CHECK(hipEventSynchronize(r.stop));
    float ms = 0.f;
    XS_HIP_CHECK(hipEventElapsedTime(&ms, r.start, r.stop));
    auto &slot = ctx->prof_done[r.name];
    slot.first += ms;
    slot.second += 1;
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->prof_pending.clear();
  return XS_OK;
}

}  // namespace xs

using namespace xs;

extern "C" {

const char *xs_last_error(void) { return g_error.c_str(); }

int xs_device_count(int *count) {
  if (!count) return fail(XS_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    (void)hipGetLastError();
    return fail(XS_ERR_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return XS_OK;
}

int xs_init(int device, xs_ctx **out) {
  if (!out) return fail(XS_ERR_ARG, "null output pointer");
  *out = nullptr;
  int n = 0;
  XS_TRY(xs_device_count(&n));
  if (device < 0 || device >= n)
    return fail(XS_ERR_DEVICE, "device " + std::to_string(device) + " not in [0, " +
                                   std::to_string(n) + ")");
  XS_HIP_CHECK(hipSetDevice(device));
  auto *ctx = new xs_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return fail(XS_ERR_DEVICE, "hipGetDeviceProperties failed");
  }
  ctx->cu_count = prop.multiProcessorCount;
  ctx->max_lds = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return fail(XS_ERR_HIP, "hipStreamCreateWithFlags failed");
  }
  ctx->stream = ctx->own_stream;
  // the allocation-free entry point's scratch address on this device; a
  // failure here is reported by that entry point alone (it asks again)
  if (debye_scratch_init(device) != XS_OK) set_error("");
  *out = ctx;
  return XS_OK;
}

int xs_free(xs_ctx *ctx) {
  if (!ctx) return XS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  while (!ctx->plans.empty()) destroy_batch(*ctx->plans.begin());
  while (!ctx->atom_sets.empty()) destroy_atoms(*ctx->atom_sets.begin());
  for (auto &r : ctx->prof_pending) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->aux) (void)hipFree(ctx->aux);
  if (ctx->mat) (void)hipFree(ctx->mat);
  if (ctx->io) (void)hipFree(ctx->io);
  if (ctx->quot_check) (void)hipFree(ctx->quot_check);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return XS_OK;
}

int xs_set_stream(xs_ctx *ctx, void *stream) {
  XS_TRY(activate(ctx));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  return XS_OK;
}

int xs_synchronize(xs_ctx *ctx) {
  XS_TRY(activate(ctx));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return XS_OK;
}

int xs_profile(xs_ctx *ctx, int enable) {
  XS_TRY(activate(ctx));
  XS_TRY(prof_resolve(ctx));
  ctx->prof_done.clear();
  ctx->profiling = enable != 0;
  return XS_OK;
}

int xs_profile_read(xs_ctx *ctx, const char *kernel, double *total_ms,
                    unsigned long long *count) {
  XS_TRY(activate(ctx));
  if (!kernel) return fail(XS_ERR_ARG, "null kernel name");
  XS_TRY(prof_resolve(ctx));
  auto it = ctx->prof_done.find(kernel);
  if (total_ms) *total_ms = it == ctx->prof_done.end() ? 0.0 : it->second.first;
  if (count) *count = it == ctx->prof_done.end() ? 0ull : it->second.second;
  return XS_OK;
}

}  // extern "C"
