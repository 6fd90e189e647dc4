// rfx_device.h - CDNA4 (gfx950, wave64) device building blocks shared by
// the operator kernels: element types, wave-level ballot / prefix helpers,
// block scans and the decoupled look-back protocol.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rfx {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Elements.  Keys/values are 1, 2, 4, 8 or 16 bytes; 16-byte elements are two
// little-endian u64 words compared high word first.
// ---------------------------------------------------------------------------
struct alignas(16) u128 {
  uint64_t lo, hi;
};
struct Empty {};

template <int B> struct Elem;
template <> struct Elem<0> { using T = Empty; };
template <> struct Elem<1> { using T = uint8_t; };
template <> struct Elem<2> { using T = uint16_t; };
template <> struct Elem<4> { using T = uint32_t; };
template <> struct Elem<8> { using T = uint64_t; };
template <> struct Elem<16> { using T = u128; };

__device__ __forceinline__ bool lt(uint8_t a, uint8_t b) { return a < b; }
__device__ __forceinline__ bool lt(uint16_t a, uint16_t b) { return a < b; }
__device__ __forceinline__ bool lt(uint32_t a, uint32_t b) { return a < b; }
__device__ __forceinline__ bool lt(uint64_t a, uint64_t b) { return a < b; }
__device__ __forceinline__ bool lt(const u128& a, const u128& b) {
  return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}
__device__ __forceinline__ bool eq(uint8_t a, uint8_t b) { return a == b; }
__device__ __forceinline__ bool eq(uint16_t a, uint16_t b) { return a == b; }
__device__ __forceinline__ bool eq(uint32_t a, uint32_t b) { return a == b; }
__device__ __forceinline__ bool eq(uint64_t a, uint64_t b) { return a == b; }
__device__ __forceinline__ bool eq(const u128& a, const u128& b) {
  return a.hi == b.hi && a.lo == b.lo;
}

// Radix digit of `k` at bit `shift` (a multiple of 8) masked by dmask.
template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, uint32_t shift, uint32_t dmask) {
  return static_cast<uint32_t>(static_cast<uint64_t>(k) >> shift) & dmask;
}
// Keys of at most 32 bits: one 32-bit shift (shift < 32 for them).
template <>
__device__ __forceinline__ uint32_t digit_of<uint32_t>(uint32_t k, uint32_t shift, uint32_t dmask) {
  return (k >> shift) & dmask;
}
template <>
__device__ __forceinline__ uint32_t digit_of<uint16_t>(uint16_t k, uint32_t shift, uint32_t dmask) {
  return (static_cast<uint32_t>(k) >> shift) & dmask;
}
template <>
__device__ __forceinline__ uint32_t digit_of<uint8_t>(uint8_t k, uint32_t shift, uint32_t dmask) {
  return (static_cast<uint32_t>(k) >> shift) & dmask;
}
// 64-bit keys: the 32-bit half holding the digit (shift is a multiple of
// 8, so a digit never straddles the halves).
template <>
__device__ __forceinline__ uint32_t digit_of<uint64_t>(uint64_t k, uint32_t shift, uint32_t dmask) {
  const uint32_t h = shift >= 32 ? static_cast<uint32_t>(k >> 32) : static_cast<uint32_t>(k);
  return (h >> (shift & 31)) & dmask;
}
template <>
__device__ __forceinline__ uint32_t digit_of<u128>(u128 k, uint32_t shift, uint32_t dmask) {
  uint64_t w = shift >= 64 ? (k.hi >> (shift - 64)) : (k.lo >> shift);
  return static_cast<uint32_t>(w) & dmask;
}

// 128-bit helpers used by tuple composition.
__device__ __forceinline__ u128 to128(uint8_t v) { return {v, 0}; }
__device__ __forceinline__ u128 to128(uint16_t v) { return {v, 0}; }
__device__ __forceinline__ u128 to128(uint32_t v) { return {v, 0}; }
__device__ __forceinline__ u128 to128(uint64_t v) { return {v, 0}; }
__device__ __forceinline__ u128 to128(u128 v) { return v; }
__device__ __forceinline__ u128 shl128(u128 v, uint32_t s) {
  if (s == 0) return v;
  if (s >= 128) return {0, 0};
  if (s >= 64) return {0, v.lo << (s - 64)};
  return {v.lo << s, (v.hi << s) | (v.lo >> (64 - s))};
}
__device__ __forceinline__ u128 shr128(u128 v, uint32_t s) {
  if (s == 0) return v;
  if (s >= 128) return {0, 0};
  if (s >= 64) return {v.hi >> (s - 64), 0};
  return {(v.lo >> s) | (v.hi << (64 - s)), v.hi >> s};
}
__device__ __forceinline__ u128 or128(u128 a, u128 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ u128 mask128(u128 v, uint32_t bits) {
  if (bits >= 128) return v;
  if (bits >= 64) return {v.lo, bits > 64 ? (v.hi & (~0ull >> (128 - bits))) : 0};
  return {bits ? (v.lo & (~0ull >> (64 - bits))) : 0, 0};
}
template <class T> __device__ __forceinline__ T from128(u128 v) {
  return static_cast<T>(v.lo);
}
template <> __device__ __forceinline__ u128 from128<u128>(u128 v) { return v; }

// ---------------------------------------------------------------------------
// Runs of equal keys in a sorted LDS array.
// ---------------------------------------------------------------------------
// Copies of x immediately before index i in the LDS array s (gallop then
// binary search; 0 when s[i - 1] != x).
template <class K>
__device__ __forceinline__ uint32_t lds_run_back(const K* s, uint32_t i, K x) {
  if (i == 0 || !eq(s[i - 1], x)) return 0;
  uint32_t hi = i - 1, step = 1, lo;  // invariant: s[hi] == x
  while (true) {
    if (hi < step) {
      lo = 0;
      if (eq(s[0], x)) return i;
      break;
    }
    lo = hi - step;
    if (!eq(s[lo], x)) break;
    hi = lo;
    step <<= 1;
  }
  // s[lo] != x, s[hi] == x: first equal index in (lo, hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (eq(s[mid], x)) hi = mid;
    else lo = mid;
  }
  return i - hi;
}

// Copies of x from index i forward in s[0, n) (0 when s[i] != x).
template <class K>
__device__ __forceinline__ uint32_t lds_run_fwd(const K* s, uint32_t i, uint32_t n, K x) {
  if (i >= n || !eq(s[i], x)) return 0;
  uint32_t lo = i, step = 1, hi;  // invariant: s[lo] == x
  while (true) {
    hi = lo + step;
    if (hi >= n) {
      if (eq(s[n - 1], x)) return n - i;
      hi = n - 1;
      break;
    }
    if (!eq(s[hi], x)) break;
    lo = hi;
    step <<= 1;
  }
  // s[lo] == x, s[hi] != x: last equal index in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (eq(s[mid], x)) lo = mid;
    else hi = mid;
  }
  return lo + 1 - i;
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// popcount(mask & lanes-below-me), two VALU ops.
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Inclusive prefix sum across the wave (shuffle-based).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T y = __shfl_up(x, d, kWave);
    if (lane >= static_cast<uint32_t>(d)) x += y;
  }
  return x;
}
// This thread's wave within its block.  The same for every lane, and said
// so with readfirstlane: values derived from it (wave offsets, row bases)
// stay in scalar registers instead of being recomputed per lane.
__device__ __forceinline__ uint32_t wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// Block-wide exclusive scan of one value per thread.  `smem` holds at least
// NT/64 + 1 elements.  Returns the exclusive prefix; *total gets the sum.
template <int NT, class T>
__device__ __forceinline__ T block_excl_scan(T x, T* smem, T* total) {
  constexpr int NW = NT / kWave;
  const uint32_t lane = lane_id();
  const uint32_t w = wave_id();
  T incl = wave_incl_scan(x);
  if (lane == kWave - 1) smem[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      T c = smem[i];
      smem[i] = run;
      run += c;
    }
    smem[NW] = run;
  }
  __syncthreads();
  T r = smem[w] + incl - x;
  *total = smem[NW];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Decoupled look-back (single-pass prefix across tiles).  Each tile owns a
// status word per lane of the look-back: the top two bits say whether the
// word holds nothing yet (0), the tile's own aggregate (1) or the inclusive
// prefix through the tile (2); the low 30 bits hold the count.  Status words
// are zeroed before the launch; tiles are numbered by an atomic ticket, so a
// tile only waits on tiles already running.  Loads and stores are relaxed
// agent-scope atomics: the value travels in the same word as its flag.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagIncl = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1;

__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile `tile` for the status column at `col` with
// `stride` words between consecutive tiles.
__device__ __forceinline__ uint32_t lookback(const uint32_t* status, uint32_t tile,
                                             uint32_t stride, uint32_t col) {
  uint32_t excl = 0;
  int64_t k = static_cast<int64_t>(tile) - 1;
  while (k >= 0) {
    uint32_t v = status_load(status + static_cast<uint64_t>(k) * stride + col);
    uint32_t f = v & ~kCountMask;
    if (f == 0) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += v & kCountMask;
    if (f == kFlagIncl) break;
    --k;
  }
  return excl;
}

// The same walk, polling W predecessors per round trip: tiles k, k-1, ...,
// k-W+1 are loaded together and consumed in order up to the first word not
// yet published (poll again from there) or the first inclusive prefix
// (done).  Tiles before tile 0 count as an inclusive prefix of 0.
template <int W>
__device__ __forceinline__ uint32_t lookback_window(const uint32_t* status, uint32_t tile,
                                                    uint32_t stride, uint32_t col) {
  uint32_t excl = 0;
  int64_t k = static_cast<int64_t>(tile) - 1;
  while (true) {
    uint32_t v[W];
#pragma unroll
    for (int i = 0; i < W; ++i)
      v[i] = (k - i >= 0) ? status_load(status + static_cast<uint64_t>(k - i) * stride + col)
                          : kFlagIncl;
    bool stop = false, fin = false;
    int used = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const uint32_t f = v[i] & ~kCountMask;
      const bool take = !stop && f != 0;
      excl += take ? (v[i] & kCountMask) : 0u;
      used += take;
      fin |= take && f == kFlagIncl;
      stop |= (f == 0) || (take && f == kFlagIncl);
    }
    if (fin) return excl;
    k -= used;
    if (used < W) __builtin_amdgcn_s_sleep(1);
  }
}

// 64-bit variant for prefixes that may exceed 2^30 (compaction offsets):
// status is a u64 word, flag in bits 62-63, value in the low 62 bits.
constexpr uint64_t kFlagAgg64 = 1ull << 62;
constexpr uint64_t kFlagIncl64 = 2ull << 62;
constexpr uint64_t kCountMask64 = (1ull << 62) - 1;

__device__ __forceinline__ void status_store64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load64(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-cooperative look-back over 64 predecessors at a time: lane l reads
// tile (base - 1 - l).  Call from one full wave; every lane returns the
// exclusive prefix of `tile`.
__device__ __forceinline__ uint64_t wave_lookback64(const uint64_t* status, uint32_t tile) {
  const uint32_t lane = lane_id();
  uint64_t excl = 0;
  int64_t base = static_cast<int64_t>(tile);
  while (base > 0) {
    const int64_t k = base - 1 - static_cast<int64_t>(lane);
    uint64_t v = 0;
    bool have = k < 0;  // lanes past tile 0 count as an empty inclusive prefix
    while (true) {
      if (!have) {
        v = status_load64(status + k);
        have = (v >> 62) != 0;
      }
      const uint64_t ready = ballot(have);
      const bool incl = k < 0 || (v >> 62) == 2;
      const uint64_t incl_mask = ballot(have && incl);
      // Only the lanes up to the nearest inclusive word are needed, and
      // they must all have published.
      const uint32_t nr = ready == ~0ull ? 64u : static_cast<uint32_t>(__builtin_ctzll(~ready));
      const uint32_t fi = incl_mask ? static_cast<uint32_t>(__builtin_ctzll(incl_mask)) : 64u;
      if (fi < nr) {
        excl += wave_sum((lane <= fi && k >= 0) ? (v & kCountMask64) : 0ull);
        return excl;
      }
      if (nr == 64) {
        excl += wave_sum(v & kCountMask64);  // 64 aggregates
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    base -= kWave;
  }
  return excl;
}

}  // namespace rfx
