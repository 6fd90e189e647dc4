// strings.hip - sorted search over string keys (find_bounds_string,
// ModernGPUJoin.cu:213-312) for gfx950.
//
// Keys are device addresses of NUL-terminated strings, each side sorted by
// compare_string2 (ModernGPUJoin.cu:72-97): bytes compared as signed char
// up to the first NUL of either, a proper prefix ordered first.  One
// thread per left key finds the first right key not below it and the first
// right key above it by binary search (every probe a string comparison in
// global memory); the counts are then turned into their exclusive prefix
// sums by a tile scan with decoupled look-back, whose last tile writes the
// total.  The outputs are those of rfx_join_find_bounds, so rfx_join_main
// expands them into row pairs.
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr uint64_t kLow7 = 0x7F7F7F7F7F7F7F7Full;

// 0x80 in every byte of v that is zero (exact: no false flags).
__device__ __forceinline__ uint64_t zero_bytes(uint64_t v) {
  return ~(((v & kLow7) + kLow7) | v | kLow7);
}
// 0x80 in every byte of v that is not zero.
__device__ __forceinline__ uint64_t nonzero_bytes(uint64_t v) {
  return (((v & kLow7) + kLow7) | v) & ~kLow7;
}

// The 8 bytes at address p (little-endian), read as aligned words: the
// word after the first is read only when the first holds no NUL at or
// after p, so every word read holds a byte of the string (an aligned word
// does not cross a page).  Bytes after a NUL are don't-cares.
__device__ __forceinline__ uint64_t load8(uint64_t p) {
  const uint64_t* w = reinterpret_cast<const uint64_t*>(p & ~7ull);
  const uint32_t sh = static_cast<uint32_t>(p & 7) * 8;
  uint64_t v = w[0] >> sh;
  if (sh && !(zero_bytes(v | (~0ull << (64 - sh))))) v |= w[1] << (64 - sh);
  return v;
}

// compare_string2: a < b, 8 bytes per step.  A null address reads as the
// empty string.
__device__ __forceinline__ bool str_less(uint64_t pa, uint64_t pb) {
  if (!pa || !pb) return !pa && pb && *reinterpret_cast<const uint8_t*>(pb) != 0;
  for (;; pa += 8, pb += 8) {
    const uint64_t va = load8(pa), vb = load8(pb);
    const uint64_t m = zero_bytes(va) | zero_bytes(vb) | nonzero_bytes(va ^ vb);
    if (m) {
      const uint32_t sh = static_cast<uint32_t>(__builtin_ctzll(m)) & ~7u;
      const uint8_t x = static_cast<uint8_t>(va >> sh), y = static_cast<uint8_t>(vb >> sh);
      if (x == 0 || y == 0) return x == 0 && y != 0;
      return static_cast<int8_t>(x) < static_cast<int8_t>(y);
    }
  }
}

constexpr int kSNT = 256;

__global__ __launch_bounds__(kSNT) void k_str_bounds(const uint64_t* __restrict__ left, uint64_t nl,
                                                     const uint64_t* __restrict__ right,
                                                     uint64_t nr, int32_t* __restrict__ lower,
                                                     int32_t* __restrict__ count) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSNT + threadIdx.x;
  if (i >= nl) return;
  const uint64_t key = left[i];
  uint64_t lo = 0, hi = nr;  // first right key not below key
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (str_less(right[mid], key)) lo = mid + 1;
    else hi = mid;
  }
  uint64_t ub_lo = lo, ub_hi = nr;  // first right key above key
  while (ub_lo < ub_hi) {
    const uint64_t mid = (ub_lo + ub_hi) >> 1;
    if (str_less(key, right[mid])) ub_hi = mid;
    else ub_lo = mid + 1;
  }
  lower[i] = static_cast<int32_t>(lo);
  count[i] = static_cast<int32_t>(ub_lo - lo);
}

// count[] -> exclusive prefix sums in place, tiles of 256 x 8; *total gets
// the sum.
constexpr int kCVT = 8;
constexpr int kCTile = kSNT * kCVT;

__global__ __launch_bounds__(kSNT) void k_count_scan(int32_t* __restrict__ count, uint64_t n,
                                                     uint32_t ntiles, uint64_t* __restrict__ total,
                                                     uint64_t* __restrict__ status,
                                                     uint32_t* __restrict__ ticket) {
  __shared__ uint64_t s_scan[kSNT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t at = static_cast<uint64_t>(tile) * kCTile + static_cast<uint64_t>(tid) * kCVT;
  uint64_t v[kCVT], sum = 0;
#pragma unroll
  for (int j = 0; j < kCVT; ++j) {
    v[j] = at + j < n ? static_cast<uint32_t>(count[at + j]) : 0u;
    sum += v[j];
  }
  uint64_t tile_sum = 0;
  const uint64_t excl = block_excl_scan<kSNT>(sum, s_scan, &tile_sum);
  if (w == 0) {
    if (tile == 0) {
      if (lane == 0) {
        status_store64(status, kFlagIncl64 | tile_sum);
        s_prefix = 0;
      }
    } else {
      if (lane == 0) status_store64(status + tile, kFlagAgg64 | tile_sum);
      const uint64_t prev = wave_lookback64(status, tile);
      if (lane == 0) {
        status_store64(status + tile, kFlagIncl64 | (prev + tile_sum));
        s_prefix = prev;
      }
    }
  }
  __syncthreads();
  uint64_t run = s_prefix + excl;
#pragma unroll
  for (int j = 0; j < kCVT; ++j) {
    if (at + j < n) count[at + j] = static_cast<int32_t>(run);
    run += v[j];
  }
  if (tile == ntiles - 1 && tid == 0) *total = s_prefix + tile_sum;
}

// The order-preserving code of one byte of a string: the terminator 0,
// then signed chars -128..-1, 1..127 as 1..255 (compare_string2's order).
__device__ __forceinline__ uint64_t str_code(uint8_t b) {
  return b == 0 ? 0u : static_cast<uint64_t>((b ^ 0x80u) + (b >= 0x80u ? 1u : 0u));
}

// The longest string (bytes before the NUL; null addresses are empty)
// into *longest, which starts at 0.
__global__ __launch_bounds__(kSNT) void k_str_len(const uint64_t* __restrict__ addr, uint64_t n,
                                                  uint64_t* __restrict__ longest) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSNT + threadIdx.x;
  uint64_t k = 0;
  if (i < n) {
    const uint8_t* s = reinterpret_cast<const uint8_t*>(addr[i]);
    if (s)
      while (s[k]) ++k;
  }
  // One atomic per wave.
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint64_t o = __shfl_xor(k, d, kWave);
    k = o > k ? o : k;
  }
  if (lane_id() == 0 && k) atomicMax(reinterpret_cast<unsigned long long*>(longest), k);
}

// Chunk c of every string as a big-endian key of 8 codes (codes past the
// end are 0).  With rows, element i is the string of row rows[i].
__global__ __launch_bounds__(kSNT) void k_str_chunk(const uint64_t* __restrict__ addr,
                                                    const uint32_t* __restrict__ rows, uint64_t n,
                                                    uint64_t chunk, uint64_t* __restrict__ key) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSNT + threadIdx.x;
  if (i >= n) return;
  const uint8_t* s = reinterpret_cast<const uint8_t*>(addr[rows ? rows[i] : i]);
  uint64_t k = 0;
  bool ended = s == nullptr;
  for (uint64_t q = 0; q < 8 * chunk && !ended; ++q) ended = s[q] == 0;
  for (int b = 0; b < 8; ++b) {
    const uint8_t c = ended ? 0 : s[8 * chunk + b];
    ended = ended || c == 0;
    k = (k << 8) | str_code(c);
  }
  key[i] = k;
}

// Run heads of sorted strings: 1 where a string differs from the one
// before it (sorted, so where the one before is less), 0 elsewhere.
__global__ __launch_bounds__(kSNT) void k_str_heads(const uint64_t* __restrict__ addr, uint64_t n,
                                                    int32_t* __restrict__ head) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSNT + threadIdx.x;
  if (i >= n) return;
  head[i] = i == 0 || str_less(addr[i - 1], addr[i]) ? 1 : 0;
}

// After the exclusive scan of the heads: group of row i = heads up to and
// including i, less one (the scan of row i + 1, or the total for the last).
__global__ __launch_bounds__(kSNT) void k_str_groups(const int32_t* __restrict__ excl, uint64_t n,
                                                     const uint64_t* __restrict__ total,
                                                     uint32_t* __restrict__ group) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSNT + threadIdx.x;
  if (i >= n) return;
  const uint64_t incl = i + 1 < n ? static_cast<uint32_t>(excl[i + 1]) : *total;
  group[i] = static_cast<uint32_t>(incl - 1);
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_string_groups(rfx_ctx* ctx, const uint64_t* sorted, uint64_t n, uint32_t* group,
                                 int32_t* scratch, uint64_t* n_groups) {
  if (!ctx || !n_groups || (n && (!sorted || !group || !scratch)))
    return fail(RFX_EINVAL, "rfx_string_groups: null argument");
  if (n >= (1ull << 31)) return fail(RFX_ERANGE, "rfx_string_groups: more than 2^31 - 1 rows");
  if (n == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(n_groups), 2, 0);
  const uint32_t grid = static_cast<uint32_t>((n + kSNT - 1) / kSNT);
  k_str_heads<<<grid, kSNT, 0, ctx->stream>>>(sorted, n, scratch);
  RFX_LAUNCHED();
  const uint64_t ntiles = (n + kCTile - 1) / kCTile;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* status = w.take<uint64_t>(ntiles);
  uint32_t* ticket = w.take<uint32_t>(1);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  k_count_scan<<<static_cast<uint32_t>(ntiles), kSNT, 0, ctx->stream>>>(
      scratch, n, static_cast<uint32_t>(ntiles), n_groups, status, ticket);
  RFX_LAUNCHED();
  k_str_groups<<<grid, kSNT, 0, ctx->stream>>>(scratch, n, n_groups, group);
  RFX_LAUNCHED();
  return RFX_OK;
}

extern "C" int rfx_sort_strings(rfx_ctx* ctx, uint64_t* addresses, uint64_t* scratch, uint64_t n) {
  if (!ctx || (n && (!addresses || !scratch)))
    return fail(RFX_EINVAL, "rfx_sort_strings: null argument");
  if (n < 2) return RFX_OK;
  if (n >= (1ull << 40)) return fail(RFX_ERANGE, "rfx_sort_strings: too many strings");
  const uint32_t grid = static_cast<uint32_t>((n + kSNT - 1) / kSNT);
  // The longest string fixes the number of passes (scratch[n] holds it).
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(scratch + n), 2, 0));
  k_str_len<<<grid, kSNT, 0, ctx->stream>>>(addresses, n, scratch + n);
  RFX_LAUNCHED();
  uint64_t longest = 0;
  RFX_HIP(hipMemcpyAsync(&longest, scratch + n, 8, hipMemcpyDeviceToHost, ctx->stream));
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  // LSD over 8-byte chunks, last chunk first: each pass a stable sort of
  // (chunk key, address) pairs.
  for (uint64_t chunk = (longest + 7) / 8; chunk-- > 0;) {
    k_str_chunk<<<grid, kSNT, 0, ctx->stream>>>(addresses, nullptr, n, chunk, scratch);
    RFX_LAUNCHED();
    RFX_TRY(rfx_sort_pairs(ctx, scratch, addresses, n, 8, 64, 8));
  }
  return RFX_OK;
}

extern "C" int rfx_sort_strings_rows(rfx_ctx* ctx, const uint64_t* addresses, uint32_t* rows,
                                     uint64_t* scratch, uint64_t n) {
  if (!ctx || (n && (!addresses || !rows || !scratch)))
    return fail(RFX_EINVAL, "rfx_sort_strings_rows: null argument");
  if (n >= (1ull << 32)) return fail(RFX_ERANGE, "rfx_sort_strings_rows: too many rows");
  if (n < 2) return RFX_OK;
  const uint32_t grid = static_cast<uint32_t>((n + kSNT - 1) / kSNT);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(scratch + n), 2, 0));
  k_str_len<<<grid, kSNT, 0, ctx->stream>>>(addresses, n, scratch + n);
  RFX_LAUNCHED();
  uint64_t longest = 0;
  RFX_HIP(hipMemcpyAsync(&longest, scratch + n, 8, hipMemcpyDeviceToHost, ctx->stream));
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  // As rfx_sort_strings, carrying row numbers: each pass keys the rows in
  // their current order by the chunk of their own string.
  for (uint64_t chunk = (longest + 7) / 8; chunk-- > 0;) {
    k_str_chunk<<<grid, kSNT, 0, ctx->stream>>>(addresses, rows, n, chunk, scratch);
    RFX_LAUNCHED();
    RFX_TRY(rfx_sort_pairs(ctx, scratch, rows, n, 8, 64, 4));
  }
  return RFX_OK;
}

extern "C" int rfx_join_find_bounds_string(rfx_ctx* ctx, int32_t* lower_bound, int32_t* left_count,
                                           uint64_t* total, const uint64_t* left, uint64_t nl,
                                           const uint64_t* right, uint64_t nr) {
  if (!ctx || !total || (nl && (!lower_bound || !left_count || !left)) || (nr && !right))
    return fail(RFX_EINVAL, "rfx_join_find_bounds_string: null argument");
  if (nl >= (1ull << 31) || nr >= (1ull << 31))
    return fail(RFX_ERANGE, "rfx_join_find_bounds_string: more than 2^31 - 1 keys");
  if (nl == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(total), 2, 0);
  const uint64_t ntiles = (nl + kCTile - 1) / kCTile;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* status = w.take<uint64_t>(ntiles);
  uint32_t* ticket = w.take<uint32_t>(1);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  {
    RFX_PROF(ctx, "join.find_bounds_string");
    k_str_bounds<<<static_cast<uint32_t>((nl + kSNT - 1) / kSNT), kSNT, 0, ctx->stream>>>(
        left, nl, right, nr, lower_bound, left_count);
    RFX_LAUNCHED();
  }
  k_count_scan<<<static_cast<uint32_t>(ntiles), kSNT, 0, ctx->stream>>>(
      left_count, nl, static_cast<uint32_t>(ntiles), total, status, ticket);
  RFX_LAUNCHED();
  return RFX_OK;
}
