// context.hip - contexts, workspace, memory and timing plumbing, error
// reporting, and the device-side synthetic data generators.
#include <cstdarg>
#include <cstdlib>
#include <mutex>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {

static thread_local char g_last_error[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
  return code;
}

void die(const char* where, int status) {
  std::fprintf(stderr, "%s: %s (%s)\n", where, rfx_strerror(status), g_last_error);
  std::abort();
}

int ensure_workspace(rfx_ctx* ctx, uint64_t bytes) {
  if (bytes <= ctx->ws_bytes) return RFX_OK;
  RFX_HIP(hipSetDevice(ctx->device));
  if (ctx->ws) {
    // Work already queued may still read the old workspace.
    RFX_HIP(hipStreamSynchronize(ctx->stream));
    RFX_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  uint64_t want = align_up(bytes + bytes / 8, 1ull << 20);
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    return fail(RFX_ENOMEM, "workspace allocation of %llu bytes failed",
                static_cast<unsigned long long>(want));
  }
  ctx->ws = p;
  ctx->ws_bytes = want;
  return RFX_OK;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    p[i] = v;
}

int fill_u32(rfx_ctx* ctx, uint32_t* p, uint64_t words, uint32_t value) {
  if (words == 0) return RFX_OK;
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_fill_u32<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(p, words, value);
  RFX_LAUNCHED();
  return RFX_OK;
}

// splitmix64 (Steele, Lea, Flood 2014): the generator behind every
// synthetic input.
__device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_gen_u32(uint32_t* out, uint64_t n, uint64_t seed) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    out[i] = static_cast<uint32_t>(splitmix64(seed, i) >> 32);
}
__global__ void k_gen_u64(uint64_t* out, uint64_t n, uint64_t seed) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    out[i] = splitmix64(seed, i);
}
__global__ void k_iota_u32(uint32_t* out, uint64_t n) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    out[i] = static_cast<uint32_t>(i);
}
__global__ void k_ones_u32(uint32_t* out, uint64_t n) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    out[i] = 1u;
}

static uint32_t grid_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b == 0) b = 1;
  return static_cast<uint32_t>(b);
}

}  // namespace rfx

using namespace rfx;

extern "C" {

const char* rfx_strerror(int status) {
  switch (status) {
    case RFX_OK: return "ok";
    case RFX_EINVAL: return "invalid argument";
    case RFX_ENOMEM: return "out of device memory";
    case RFX_EHIP: return "HIP runtime error";
    case RFX_ERANGE: return "size out of range";
    case RFX_ENOTSUP: return "not supported";
    case RFX_EBUDGET: return "over HBM budget";
    default: return "unknown error";
  }
}

const char* rfx_last_error(void) { return g_last_error; }

const char* rfx_version(void) { return "rfx 0.1 gfx950"; }

int rfx_ctx_create(rfx_ctx** out, int device, void* stream) {
  if (!out) return fail(RFX_EINVAL, "rfx_ctx_create: null out pointer");
  RFX_HIP(hipSetDevice(device));
  rfx_ctx* c = new rfx_ctx;
  c->device = device;
  c->stream = static_cast<hipStream_t>(stream);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) ==
          hipSuccess &&
      cus > 0)
    c->device_cus = cus;
  *out = c;
  return RFX_OK;
}

int rfx_ctx_destroy(rfx_ctx* ctx) {
  if (!ctx) return RFX_OK;
  if (ctx->ws) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->ws);
  }
  delete ctx;
  return RFX_OK;
}

int rfx_ctx_set_stream(rfx_ctx* ctx, void* stream) {
  if (!ctx) return fail(RFX_EINVAL, "null context");
  ctx->stream = static_cast<hipStream_t>(stream);
  return RFX_OK;
}

void* rfx_ctx_stream(rfx_ctx* ctx) { return ctx ? ctx->stream : nullptr; }

int rfx_ctx_reserve(rfx_ctx* ctx, uint64_t bytes) {
  if (!ctx) return fail(RFX_EINVAL, "null context");
  return ensure_workspace(ctx, bytes);
}

uint64_t rfx_ctx_workspace_bytes(rfx_ctx* ctx) { return ctx ? ctx->ws_bytes : 0; }

int rfx_ctx_synchronize(rfx_ctx* ctx) {
  if (!ctx) return fail(RFX_EINVAL, "null context");
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  return RFX_OK;
}

rfx_ctx* rfx_default_ctx(void) {
  static std::once_flag once;
  static rfx_ctx* ctx = nullptr;
  std::call_once(once, [] {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (rfx_ctx_create(&ctx, dev, nullptr) != RFX_OK) ctx = nullptr;
  });
  return ctx;
}

int rfx_device_count(int* count) {
  if (!count) return fail(RFX_EINVAL, "null count");
  RFX_HIP(hipGetDeviceCount(count));
  return RFX_OK;
}
int rfx_set_device(int device) {
  RFX_HIP(hipSetDevice(device));
  return RFX_OK;
}
int rfx_malloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(RFX_EINVAL, "null pointer");
  *ptr = nullptr;
  if (bytes == 0) bytes = 1;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(RFX_ENOMEM, "hipMalloc(%llu) failed", static_cast<unsigned long long>(bytes));
  }
  return RFX_OK;
}
int rfx_free(void* ptr) {
  if (ptr) RFX_HIP(hipFree(ptr));
  return RFX_OK;
}
int rfx_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
  if (bytes) RFX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return RFX_OK;
}
int rfx_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
  if (bytes) RFX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return RFX_OK;
}
int rfx_memcpy_dtod(void* dst, const void* src, uint64_t bytes) {
  if (bytes) RFX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return RFX_OK;
}
int rfx_memset(void* dst, int value, uint64_t bytes) {
  if (bytes) RFX_HIP(hipMemset(dst, value, bytes));
  return RFX_OK;
}
int rfx_device_synchronize(void) {
  RFX_HIP(hipDeviceSynchronize());
  return RFX_OK;
}
int rfx_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  RFX_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return RFX_OK;
}

int rfx_event_create(void** ev) {
  if (!ev) return fail(RFX_EINVAL, "null event pointer");
  hipEvent_t e;
  RFX_HIP(hipEventCreate(&e));
  *ev = e;
  return RFX_OK;
}
int rfx_event_destroy(void* ev) {
  if (ev) RFX_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return RFX_OK;
}
int rfx_event_record(void* ev, rfx_ctx* ctx) {
  RFX_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), ctx ? ctx->stream : nullptr));
  return RFX_OK;
}
int rfx_event_elapsed(float* ms, void* start, void* stop) {
  RFX_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  RFX_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                              static_cast<hipEvent_t>(stop)));
  return RFX_OK;
}

int rfx_gen_u32(rfx_ctx* ctx, uint32_t* out, uint64_t n, uint64_t seed) {
  if (!ctx || (!out && n)) return fail(RFX_EINVAL, "rfx_gen_u32: null argument");
  if (!n) return RFX_OK;
  k_gen_u32<<<grid_for(n), 256, 0, ctx->stream>>>(out, n, seed);
  RFX_LAUNCHED();
  return RFX_OK;
}
int rfx_gen_u64(rfx_ctx* ctx, uint64_t* out, uint64_t n, uint64_t seed) {
  if (!ctx || (!out && n)) return fail(RFX_EINVAL, "rfx_gen_u64: null argument");
  if (!n) return RFX_OK;
  k_gen_u64<<<grid_for(n), 256, 0, ctx->stream>>>(out, n, seed);
  RFX_LAUNCHED();
  return RFX_OK;
}
int rfx_gen_iota_u32(rfx_ctx* ctx, uint32_t* out, uint64_t n) {
  if (!ctx || (!out && n)) return fail(RFX_EINVAL, "rfx_gen_iota_u32: null argument");
  if (!n) return RFX_OK;
  k_iota_u32<<<grid_for(n), 256, 0, ctx->stream>>>(out, n);
  RFX_LAUNCHED();
  return RFX_OK;
}
int rfx_generate_ones(rfx_ctx* ctx, uint32_t* out, uint64_t n) {
  if (!ctx || (!out && n)) return fail(RFX_EINVAL, "rfx_generate_ones: null argument");
  if (!n) return RFX_OK;
  k_ones_u32<<<grid_for(n), 256, 0, ctx->stream>>>(out, n);
  RFX_LAUNCHED();
  return RFX_OK;
}

}  // extern "C"
