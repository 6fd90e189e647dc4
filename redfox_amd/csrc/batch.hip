// batch.hip - support for running many RA query instances as one plan:
// streams and HIP-graph capture of a context's work, segmented (instance-
// prefixed) keys, and the synthetic fact/dimension generator of the
// batched join-aggregate workload.
#include <map>
#include <mutex>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {

// Instantiated graphs and the context whose workspace they captured: while
// a context owns live graphs its workspace may neither grow nor be freed
// (the graphs hold its addresses).
static std::mutex g_graph_mu;
static std::map<void*, rfx_ctx*> g_graph_owner;

void forget_graphs_of(rfx_ctx* ctx) {
  std::lock_guard<std::mutex> lk(g_graph_mu);
  for (auto it = g_graph_owner.begin(); it != g_graph_owner.end();)
    it = it->second == ctx ? g_graph_owner.erase(it) : std::next(it);
}

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Segment of row i: the last s with offsets[s] <= i (offsets ascending,
// offsets[0] = 0, nseg + 1 entries).
__device__ __forceinline__ uint32_t segment_of(const uint64_t* offsets, uint32_t nseg, uint64_t i) {
  uint32_t lo = 0, hi = nseg;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

constexpr uint32_t kSegLds = 4096;  // segment offsets searched in LDS

__global__ __launch_bounds__(256) void k_segment_keys(uint32_t* __restrict__ out,
                                                      const uint32_t* __restrict__ key,
                                                      const uint64_t* __restrict__ offsets,
                                                      uint32_t nseg, uint64_t n,
                                                      uint32_t key_bits) {
  __shared__ uint64_t s_off[kSegLds + 1];
  const bool lds = nseg <= kSegLds;
  if (lds) {
    for (uint32_t i = threadIdx.x; i <= nseg; i += blockDim.x) s_off[i] = offsets[i];
    __syncthreads();
  }
  const uint64_t* off = lds ? s_off : offsets;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    const uint32_t s = segment_of(off, nseg, i);
    out[i] = (s << key_bits) | key[i];
  }
}

__device__ __forceinline__ double zH(double x, double q) { return (exp(q * log(x)) - 1.0) / q; }
__device__ __forceinline__ double zHinv(double y, double q) { return exp(log1p(q * y) / q); }

// Fact rows of every instance: fk = perm_s[rank - 1] with rank ~ Zipf(s)
// over [1, d_s] (rejection-inversion, Hormann & Derflinger 1996), measure =
// U{0..100000} / 100.
__global__ void k_gen_fk_batch(uint32_t* __restrict__ fk, double* __restrict__ measure,
                               const uint64_t* __restrict__ row_off,
                               const uint64_t* __restrict__ dim_off,
                               const uint32_t* __restrict__ perm, uint32_t nseg, uint64_t n,
                               uint64_t seed, double s) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const double q = 1.0 - s;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    const uint32_t b = segment_of(row_off, nseg, i);
    const uint64_t d = dim_off[b + 1] - dim_off[b];
    const double h_x1 = zH(1.5, q) - 1.0;
    const double h_n = zH(static_cast<double>(d) + 0.5, q);
    const double s_const = 2.0 - zHinv(zH(2.5, q) - exp(-s * log(2.0)), q);
    double k = 1.0;
    for (int a = 0; a < 64; ++a) {
      const double u = static_cast<double>(mix(seed, i * 64 + a) >> 11) * (1.0 / 9007199254740992.0);
      const double ur = h_n + u * (h_x1 - h_n);
      const double x = zHinv(ur, q);
      k = floor(x + 0.5);
      if (k < 1.0) k = 1.0;
      if (k > static_cast<double>(d)) k = static_cast<double>(d);
      if (k - x <= s_const || ur >= zH(k + 0.5, q) - exp(-s * log(k))) break;
    }
    fk[i] = perm[dim_off[b] + static_cast<uint64_t>(k) - 1];
    measure[i] = static_cast<double>(mix(seed + 1, i) % 100001) / 100.0;
  }
}

uint32_t grid_of(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  return static_cast<uint32_t>(b ? b : 1);
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" {

int rfx_stream_create(void** stream) {
  if (!stream) return fail(RFX_EINVAL, "rfx_stream_create: null pointer");
  hipStream_t s = nullptr;
  RFX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return RFX_OK;
}

int rfx_stream_destroy(void* stream) {
  if (stream) RFX_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return RFX_OK;
}

int rfx_stream_synchronize(void* stream) {
  RFX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RFX_OK;
}

int rfx_graph_begin(rfx_ctx* ctx) {
  if (!ctx || !ctx->stream) return fail(RFX_EINVAL, "rfx_graph_begin: needs a context with its own stream");
  if (ctx->profiling) return fail(RFX_EINVAL, "rfx_graph_begin: disable profiling first");
  RFX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  return RFX_OK;
}

int rfx_graph_end(rfx_ctx* ctx, void** graph_exec) {
  if (!ctx || !graph_exec) return fail(RFX_EINVAL, "rfx_graph_end: null argument");
  hipGraph_t g = nullptr;
  ctx->capturing = false;
  RFX_HIP(hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess)
    return fail(RFX_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(err));
  *graph_exec = e;
  {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    g_graph_owner[e] = ctx;
  }
  ctx->live_graphs += 1;
  return RFX_OK;
}

int rfx_graph_launch(rfx_ctx* ctx, void* graph_exec) {
  if (!ctx || !graph_exec) return fail(RFX_EINVAL, "rfx_graph_launch: null argument");
  RFX_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), ctx->stream));
  return RFX_OK;
}

int rfx_graph_destroy(void* graph_exec) {
  if (!graph_exec) return RFX_OK;
  {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graph_owner.find(graph_exec);
    if (it != g_graph_owner.end()) {
      it->second->live_graphs -= 1;
      g_graph_owner.erase(it);
    }
  }
  RFX_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return RFX_OK;
}

int rfx_copy(rfx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(RFX_EINVAL, "rfx_copy: null argument");
  if (bytes) RFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return RFX_OK;
}

int rfx_segment_keys(rfx_ctx* ctx, uint32_t* out, const uint32_t* key,
                     const uint64_t* seg_offsets, uint32_t nseg, uint64_t n, uint32_t key_bits) {
  if (!ctx || (n && (!out || !key || !seg_offsets)) || nseg == 0)
    return fail(RFX_EINVAL, "rfx_segment_keys: null argument");
  uint32_t segbits = 0;
  while ((1u << segbits) < nseg) ++segbits;
  if (key_bits + segbits > 32)
    return fail(RFX_ERANGE, "rfx_segment_keys: %u key bits + %u segment bits exceed 32", key_bits,
                segbits);
  if (!n) return RFX_OK;
  k_segment_keys<<<grid_of(n), 256, 0, ctx->stream>>>(out, key, seg_offsets, nseg, n, key_bits);
  RFX_LAUNCHED();
  return RFX_OK;
}

int rfx_gen_fk_batch(rfx_ctx* ctx, uint32_t* fk, double* measure, const uint64_t* row_off,
                     const uint64_t* dim_off, const uint32_t* perm, uint32_t nseg, uint64_t n,
                     uint64_t seed, double s) {
  if (!ctx || (n && (!fk || !measure || !row_off || !dim_off || !perm)) || nseg == 0)
    return fail(RFX_EINVAL, "rfx_gen_fk_batch: null argument");
  if (!(s > 1.0)) return fail(RFX_EINVAL, "rfx_gen_fk_batch: s must exceed 1");
  if (!n) return RFX_OK;
  k_gen_fk_batch<<<grid_of(n), 256, 0, ctx->stream>>>(fk, measure, row_off, dim_off, perm, nseg, n,
                                                      seed, s);
  RFX_LAUNCHED();
  return RFX_OK;
}

}  // extern "C"
