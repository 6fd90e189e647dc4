// context.hip - contexts, workspace, memory and timing plumbing, error
// reporting, and the device-side synthetic data generators.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
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
  if (ctx->capturing)
    return fail(RFX_EINVAL, "workspace must grow to %llu bytes during graph capture: reserve first",
                static_cast<unsigned long long>(bytes));
  if (ctx->live_graphs > 0)
    return fail(RFX_EINVAL,
                "workspace must grow to %llu bytes while %d captured graph(s) of this context "
                "still use it: destroy them first, or reserve before capturing",
                static_cast<unsigned long long>(bytes), ctx->live_graphs);
  if (ctx->ws_borrowed)
    return fail(RFX_EBUDGET, "operator needs %llu workspace bytes, the borrowed region holds %llu",
                static_cast<unsigned long long>(bytes),
                static_cast<unsigned long long>(ctx->ws_bytes));
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

static hipEvent_t prof_event(rfx_ctx* ctx) {
  if (!ctx->prof_pool.empty()) {
    hipEvent_t e = ctx->prof_pool.back();
    ctx->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void prof_begin(rfx_ctx* ctx, const char* tag) {
  rfx_prof_rec r{tag, prof_event(ctx), prof_event(ctx)};
  (void)hipEventRecord(r.start, ctx->stream);
  ctx->prof_pending.push_back(r);
}

void prof_end(rfx_ctx* ctx) {
  if (ctx->prof_pending.empty()) return;
  (void)hipEventRecord(ctx->prof_pending.back().stop, ctx->stream);
}

// Fold finished launches into the per-tag totals.
static int prof_collect(rfx_ctx* ctx) {
  for (auto& r : ctx->prof_pending) {
    float ms = 0.f;
    RFX_HIP(hipEventSynchronize(r.stop));
    RFX_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    auto& st = ctx->prof_stats[r.tag];
    st.first += 1;
    st.second += ms;
    ctx->prof_pool.push_back(r.start);
    ctx->prof_pool.push_back(r.stop);
  }
  ctx->prof_pending.clear();
  return RFX_OK;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride)
    p[i] = v;
}

int ctx_cells(rfx_ctx* ctx, uint64_t** cells) {
  if (!ctx->cells) {
    if (ctx->capturing) return fail(RFX_EINVAL, "context cells must be allocated before capture");
    RFX_HIP(hipSetDevice(ctx->device));
    void* p = nullptr;
    if (hipMalloc(&p, kCtxCells * sizeof(uint64_t)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(RFX_ENOMEM, "allocating the context's scalar cells failed");
    }
    ctx->cells = static_cast<uint64_t*>(p);
  }
  *cells = ctx->cells;
  return RFX_OK;
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

// Zipf(s) over [1, n_max] by rejection-inversion (Hormann & Derflinger,
// "Rejection-inversion to generate variates from monotone discrete
// distributions", 1996).  Element i tries draws i*64 + 0, 1, ... of its
// stream until one is accepted.
struct ZipfParams {
  double q, s_const, h_x1, h_n;
  uint64_t n_max;
};
__device__ __forceinline__ double zipf_H(double x, double q) { return (exp(q * log(x)) - 1.0) / q; }
__device__ __forceinline__ double zipf_Hinv(double y, double q) { return exp(log1p(q * y) / q); }

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_gen_zipf(uint64_t* out, uint64_t n, uint64_t seed, ZipfParams zp) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const double s = 1.0 - zp.q;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    double k = 1.0;
    for (int a = 0; a < 64; ++a) {
      const double u = static_cast<double>(splitmix64(seed, i * 64 + a) >> 11) * (1.0 / 9007199254740992.0);
      const double ur = zp.h_n + u * (zp.h_x1 - zp.h_n);
      const double x = zipf_Hinv(ur, zp.q);
      k = floor(x + 0.5);
      if (k < 1.0) k = 1.0;
      if (k > static_cast<double>(zp.n_max)) k = static_cast<double>(zp.n_max);
      if (k - x <= zp.s_const || ur >= zipf_H(k + 0.5, zp.q) - exp(-s * log(k))) break;
    }
    out[i] = mix64(static_cast<uint64_t>(k));
  }
}

__global__ void k_gen_lineitem(int32_t* shipdate, uint8_t* flag, uint8_t* status, int32_t* qty,
                               double* price, double* disc, double* tax, uint64_t n,
                               uint64_t seed) {
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    const uint64_t a = splitmix64(seed, i), b = splitmix64(seed + 1, i);
    shipdate[i] = static_cast<int32_t>((a >> 32) % 2526);
    flag[i] = static_cast<uint8_t>(((a >> 16) & 0xFFFF) % 3);
    status[i] = static_cast<uint8_t>(((a >> 8) & 0xFF) % 2);
    const int32_t q = static_cast<int32_t>((a & 0xFF) % 50 + 1);
    qty[i] = q;
    const uint64_t cents = ((b >> 32) % 110001) + 90000;
    price[i] = static_cast<double>(q) * (static_cast<double>(cents) / 100.0);
    disc[i] = static_cast<double>(((b >> 16) & 0xFFFF) % 11) / 100.0;
    tax[i] = static_cast<double>((b & 0xFFFF) % 9) / 100.0;
  }
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
  forget_graphs_of(ctx);
  (void)prof_collect(ctx);
  for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
  if (ctx->ws && !ctx->ws_borrowed) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->ws);
  }
  if (ctx->cells) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->cells);
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

// Tuning knobs: name, field, allowed values (0 = the built-in default).
static int* tune_field(rfx_ctx* ctx, const char* knob, const int** allowed, int* nallowed) {
  static const int sort_ipt[] = {0, 8, 12, 16};
  static const int tiles[] = {0, 2048, 4096};
  static const int mf_modes[] = {0, 1, 2};
  if (!std::strcmp(knob, "sort.ipt")) {
    *allowed = sort_ipt;
    *nallowed = 4;
    return &ctx->tune_sort_ipt;
  }
  if (!std::strcmp(knob, "sort.digit")) {
    static const int digits[] = {0, 8, 9, 10};
    *allowed = digits;
    *nallowed = 4;
    return &ctx->tune_sort_digit;
  }
  if (!std::strcmp(knob, "sort.status")) {
    static const int widths[] = {0, 64};
    *allowed = widths;
    *nallowed = 2;
    return &ctx->tune_sort_status;
  }
  if (!std::strcmp(knob, "join.fb_tile")) {
    *allowed = tiles;
    *nallowed = 3;
    return &ctx->tune_fb_tile;
  }
  if (!std::strcmp(knob, "join.ex_tile")) {
    *allowed = tiles;
    *nallowed = 3;
    return &ctx->tune_ex_tile;
  }
  if (!std::strcmp(knob, "select.vt")) {
    static const int vts[] = {0, 8, 16, 24, 32};
    *allowed = vts;
    *nallowed = 5;
    return &ctx->tune_sel_vt;
  }
  if (!std::strcmp(knob, "mapfilter.mode")) {
    *allowed = mf_modes;
    *nallowed = 3;
    return &ctx->tune_mf_mode;
  }
  return nullptr;
}

int rfx_ctx_tune(rfx_ctx* ctx, const char* knob, int value) {
  if (!ctx || !knob) return fail(RFX_EINVAL, "rfx_ctx_tune: null argument");
  const int* allowed = nullptr;
  int na = 0;
  int* f = tune_field(ctx, knob, &allowed, &na);
  if (!f) return fail(RFX_EINVAL, "rfx_ctx_tune: unknown knob '%s'", knob);
  for (int i = 0; i < na; ++i)
    if (allowed[i] == value) {
      *f = value;
      return RFX_OK;
    }
  return fail(RFX_EINVAL, "rfx_ctx_tune: %s cannot be %d", knob, value);
}

int rfx_ctx_tuning(rfx_ctx* ctx, const char* knob, int* value) {
  if (!ctx || !knob || !value) return fail(RFX_EINVAL, "rfx_ctx_tuning: null argument");
  const int* allowed = nullptr;
  int na = 0;
  int* f = tune_field(ctx, knob, &allowed, &na);
  if (!f) return fail(RFX_EINVAL, "rfx_ctx_tuning: unknown knob '%s'", knob);
  *value = *f;
  return RFX_OK;
}

int rfx_ctx_synchronize(rfx_ctx* ctx) {
  if (!ctx) return fail(RFX_EINVAL, "null context");
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  return RFX_OK;
}

int rfx_ctx_profile(rfx_ctx* ctx, int enable) {
  if (!ctx) return fail(RFX_EINVAL, "null context");
  RFX_TRY(prof_collect(ctx));
  ctx->prof_stats.clear();
  ctx->profiling = enable != 0;
  return RFX_OK;
}

int rfx_ctx_profile_read(rfx_ctx* ctx, const char* tag, uint64_t* launches, double* total_ms) {
  if (!ctx || !tag) return fail(RFX_EINVAL, "null argument");
  RFX_TRY(prof_collect(ctx));
  auto it = ctx->prof_stats.find(tag);
  if (launches) *launches = it == ctx->prof_stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == ctx->prof_stats.end() ? 0.0 : it->second.second;
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
int rfx_gen_zipf_u64(rfx_ctx* ctx, uint64_t* out, uint64_t n, uint64_t seed, double s,
                     uint64_t n_max) {
  if (!ctx || (!out && n)) return fail(RFX_EINVAL, "rfx_gen_zipf_u64: null argument");
  if (!(s > 1.0) || n_max < 1) return fail(RFX_EINVAL, "rfx_gen_zipf_u64: s must exceed 1");
  if (!n) return RFX_OK;
  ZipfParams zp;
  zp.q = 1.0 - s;
  auto H = [&](double x) { return (std::exp(zp.q * std::log(x)) - 1.0) / zp.q; };
  auto Hinv = [&](double y) { return std::exp(std::log1p(zp.q * y) / zp.q); };
  zp.h_x1 = H(1.5) - 1.0;
  zp.h_n = H(static_cast<double>(n_max) + 0.5);
  zp.s_const = 2.0 - Hinv(H(2.5) - std::exp(-s * std::log(2.0)));
  zp.n_max = n_max;
  k_gen_zipf<<<grid_for(n), 256, 0, ctx->stream>>>(out, n, seed, zp);
  RFX_LAUNCHED();
  return RFX_OK;
}

int rfx_gen_lineitem(rfx_ctx* ctx, int32_t* shipdate, uint8_t* returnflag, uint8_t* linestatus,
                     int32_t* qty, double* price, double* disc, double* tax, uint64_t n,
                     uint64_t seed) {
  if (!ctx || (n && (!shipdate || !returnflag || !linestatus || !qty || !price || !disc || !tax)))
    return fail(RFX_EINVAL, "rfx_gen_lineitem: null argument");
  if (!n) return RFX_OK;
  k_gen_lineitem<<<grid_for(n), 256, 0, ctx->stream>>>(shipdate, returnflag, linestatus, qty,
                                                        price, disc, tax, n, seed);
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
