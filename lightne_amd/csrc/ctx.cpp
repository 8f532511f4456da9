// Context, error reporting and scratch management for liblightne_amd.
#include "lightne_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace

int ctx_fail(lightne_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  std::lock_guard<std::mutex> g(g_err_mu);
  g_err = buf;
  return code;
}

int global_fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> g(g_err_mu);
  g_err = buf;
  return code;
}

namespace {
// frees the cached buffers not used by the call in progress; false if none
bool evict_stale(lightne_ctx* ctx) {
  if (ctx->depth == 0) return false;
  bool any = false;
  for (auto it = ctx->bufs.begin(); it != ctx->bufs.end();) {
    if (it->second.epoch != ctx->epoch) {
      if (it->second.p) (void)hipFree(it->second.p);
      it = ctx->bufs.erase(it);
      any = true;
    } else {
      ++it;
    }
  }
  return any;
}
}  // namespace

int ctx_malloc(lightne_ctx* ctx, void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) return hipSuccess;
  (void)hipGetLastError();
  *p = nullptr;
  if (!evict_stale(ctx)) return e;
  e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
  }
  return e;
}

void* ctx_buf(lightne_ctx* ctx, const char* name, size_t bytes) {
  auto& b = ctx->bufs[name];
  b.epoch = ctx->epoch;
  if (b.bytes >= bytes && b.p) return b.p;
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes ? bytes : 256;
  void* q = nullptr;
  if (ctx_malloc(ctx, &q, want) != hipSuccess) {
    ctx->bufs.erase(name);
    ctx_fail(ctx, LIGHTNE_ERR_OOM, "out of device memory allocating %zu bytes for %s", want, name);
    return nullptr;
  }
  // evict_stale may have rehashed the map: look the entry up again
  auto& b2 = ctx->bufs[name];
  b2.p = q;
  b2.bytes = want;
  b2.epoch = ctx->epoch;
  return q;
}

int WsCarve::commit(lightne_ctx* ctx, size_t reserve) {
  const size_t need = std::max(end(), reserve);
  char* w = (char*)ctx_buf(ctx, "ws", need);
  if (!w) return LIGHTNE_ERR_OOM;
  size_t o = base;
  for (auto& r : req) {
    *r.first = w + o;
    o += (r.second + 255) & ~(size_t)255;
  }
  return LIGHTNE_OK;
}

int dev_alloc(lightne_ctx* ctx, DevMem& m, size_t bytes, const char* what) {
  if (m.p) {
    (void)hipFree(m.p);
    m.p = nullptr;
  }
  if (ctx_malloc(ctx, &m.p, bytes) != hipSuccess) {
    m.p = nullptr;
    return ctx_fail(ctx, LIGHTNE_ERR_OOM, "out of device memory allocating %zu bytes for %s", bytes, what);
  }
  return LIGHTNE_OK;
}

double ev_seconds(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0;
  }
  return ms * 1e-3;
}

extern "C" {

const char* lightne_version(void) { return "lightne_amd 0.1 (gfx950)"; }

int lightne_ctx_create(int device, void* stream, lightne_ctx** out) {
  if (!out) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return global_fail(LIGHTNE_ERR_NO_DEVICE, "lightne_ctx_create: no HIP device visible");
  }
  if (device < 0 || device >= ndev)
    return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_ctx_create: device %d out of range [0,%d)", device, ndev);
  if (hipSetDevice(device) != hipSuccess)
    return global_fail(LIGHTNE_ERR_HIP, "lightne_ctx_create: hipSetDevice(%d) failed", device);
  auto* ctx = new lightne_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
  }
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      lightne_ctx_destroy(ctx);
      return global_fail(LIGHTNE_ERR_HIP, "lightne_ctx_create: hipEventCreate failed");
    }
  }
  // the allocation-free entry points' scratch array: its address on this device
  if (dense_scratch_setup(ctx) != LIGHTNE_OK) {
    std::string msg = ctx->err;
    lightne_ctx_destroy(ctx);
    return global_fail(LIGHTNE_ERR_HIP, "lightne_ctx_create: %s", msg.c_str());
  }
  *out = ctx;
  return LIGHTNE_OK;
}

int lightne_ctx_destroy(lightne_ctx* ctx) {
  if (!ctx) return LIGHTNE_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  wide_handle_free(ctx);
  delete ctx;
  return LIGHTNE_OK;
}

int lightne_ctx_release(lightne_ctx* ctx) {
  if (!ctx) return global_fail(LIGHTNE_ERR_BAD_ARG, "ctx is NULL");
  LN_HIP(ctx, hipSetDevice(ctx->device));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  ctx->bufs.clear();
  return LIGHTNE_OK;
}

int lightne_ctx_synchronize(lightne_ctx* ctx) {
  if (!ctx) return global_fail(LIGHTNE_ERR_BAD_ARG, "ctx is NULL");
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

const char* lightne_last_error(const lightne_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

const char* lightne_global_error(void) {
  static thread_local std::string copy;
  std::lock_guard<std::mutex> g(g_err_mu);
  copy = g_err;
  return copy.c_str();
}

void lightne_host_free(void* p) { std::free(p); }

}  // extern "C"
