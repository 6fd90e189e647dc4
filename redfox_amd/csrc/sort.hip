// sort.hip - stable LSD radix sort for gfx950: 8-bit digits (9 or 10 bits
// when a context is tuned to them), one upfront histogram pass over the
// keys for all digit positions, then one "onesweep" kernel per digit that
// ranks a tile in registers/LDS, obtains its global digit offsets by
// decoupled look-back and scatters through an LDS reorder buffer.  Wider
// digits take fewer passes over the data (26-bit keys: 3 instead of 4;
// 20-bit: 2 instead of 3) at the cost of more bins per tile: more ranking
// ballots and look-back words, shorter runs per bin in the scatter.
//
// Replaces redfox::sort_pair / sort_key (b40cSort.cu:58-1021: b40c LSD
// radix, 4-5 bit digits, upsweep + spine + downsweep per digit, pong
// buffers cudaMalloc'd per call) and the mergesort variants
// (ModernGPUSort.cu:238-574).  Both are stable, so for the same key bits
// they produce exactly the order this sort produces.
//
// Everything after the launch is decided on the device: a digit position
// whose histogram has one full bucket is an identity permutation and is
// skipped; which of the two buffers holds the data is tracked in a device
// word, and a final kernel copies back only if it is the pong buffer.  No
// host round trip, so a sort can be captured into a HIP graph.
#include <type_traits>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

// Digit width DB: 1 << DB bins.  8 by default; 9 and 10 for the
// sort.digit knob.
template <int DB>
constexpr int kBinsOf = 1 << DB;

#ifndef RFX_SORT_NT
#define RFX_SORT_NT 512
#endif
#ifndef RFX_SORT_IPT
#define RFX_SORT_IPT 16
#endif
#ifndef RFX_SORT_MINW
#define RFX_SORT_MINW 1
#endif
#ifndef RFX_SORT_LBW
#define RFX_SORT_LBW 4
#endif
// Tile shape: NT threads x IPT keys per thread.  LDS per block is about
// kMax*TILE (reorder) + TILE (digits) + 8 KB (wave counters): 16 keys per
// thread for <= 4-byte elements, 14 for 8-byte (2 blocks per CU), 8 for
// 16-byte.  (12 and 16 measure within 2% of each other on 2^28 u32 pairs,
// in either order from box to box; the autotuner can pick 8, 12 or 16.)
template <int KB, int VB>
struct SortTune {
  static constexpr int kMax = KB > VB ? KB : VB;
  static constexpr int NT = RFX_SORT_NT;
  static constexpr int IPT = kMax >= 16 ? RFX_SORT_IPT / 2
                             : kMax >= 8 ? (RFX_SORT_IPT * 7) / 8 : RFX_SORT_IPT;
  static constexpr int TILE = NT * IPT;
};

// Look-back status words: 32 bits (2 flag bits, a 30-bit count) for up to
// 2^30 - 1 rows, 64 bits (flags in bits 62-63) beyond, up to 2^31 - 1.
template <class S>
struct StatusWord;
template <>
struct StatusWord<uint32_t> {
  static constexpr uint32_t kAgg = kFlagAgg, kIncl = kFlagIncl, kMask = kCountMask;
};
template <>
struct StatusWord<uint64_t> {
  static constexpr uint64_t kAgg = kFlagAgg64, kIncl = kFlagIncl64, kMask = kCountMask64;
};

template <class S>
__device__ __forceinline__ void sw_store(S* p, S v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class S>
__device__ __forceinline__ S sw_load(const S* p) {
  return __hip_atomic_load(const_cast<S*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of column `col` of tile `tile` (stride words between
// tiles), polling W predecessors per round trip: they are consumed in
// order up to the first word not yet published (polled again from there)
// or the first inclusive prefix.  Tiles before tile 0 count as an
// inclusive prefix of 0.
template <int W, class S>
__device__ __forceinline__ uint32_t sort_lookback(const S* status, uint32_t tile, uint32_t stride,
                                                  uint32_t col) {
  using SW = StatusWord<S>;
  uint32_t excl = 0;
  int64_t k = static_cast<int64_t>(tile) - 1;
  while (true) {
    S v[W];
#pragma unroll
    for (int i = 0; i < W; ++i)
      v[i] = (k - i >= 0) ? sw_load(status + static_cast<uint64_t>(k - i) * stride + col) : SW::kIncl;
    bool stop = false, fin = false;
    int used = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const S f = v[i] & ~SW::kMask;
      const bool take = !stop && f != 0;
      excl += take ? static_cast<uint32_t>(v[i] & SW::kMask) : 0u;
      used += take;
      fin |= take && f == SW::kIncl;
      stop |= (f == 0) || (take && f == SW::kIncl);
    }
    if (fin) return excl;
    k -= used;
    if (used < W) __builtin_amdgcn_s_sleep(1);
  }
}

struct SortLayout {
  uint32_t* hist;    // [P][bins]
  uint32_t* base;    // [P][bins] exclusive global digit offsets
  uint32_t* ctrs;    // [P][2]   tile ticket, finished tiles
  uint32_t* skip;    // [P]
  uint32_t* sel;     // [1]      0: data in caller's buffers, 1: in pong
  void* status;      // [P][tiles][bins] status words of S
  void* pong_keys;
  void* pong_vals;
  uint64_t zero_words;  // words from hist to the end of status
};

static void carve_sort(Carve& c, uint64_t n, int kb, int vb, int passes, uint64_t tiles,
                       uint32_t bins, SortLayout* L, bool wide = false) {
  // The zeroed block comes first so it is one contiguous range.
  const uint64_t status_words = static_cast<uint64_t>(passes) * tiles * bins;
  L->hist = c.take<uint32_t>(static_cast<uint64_t>(passes) * bins);
  L->base = c.take<uint32_t>(static_cast<uint64_t>(passes) * bins);
  L->ctrs = c.take<uint32_t>(static_cast<uint64_t>(passes) * 2);
  L->skip = c.take<uint32_t>(static_cast<uint64_t>(passes));
  L->sel = c.take<uint32_t>(1);
  L->status = wide ? static_cast<void*>(c.take<uint64_t>(status_words))
                   : static_cast<void*>(c.take<uint32_t>(status_words));
  L->zero_words = (c.off - (reinterpret_cast<uintptr_t>(L->hist) - c.base)) / 4;
  L->pong_keys = c.take<char>(n * kb);
  L->pong_vals = vb ? c.take<char>(n * vb) : nullptr;
}

template <int DB>
__device__ __forceinline__ uint32_t pass_mask(int pass, uint32_t key_bits) {
  uint32_t width = key_bits - static_cast<uint32_t>(DB * pass);
  if (width > DB) width = DB;
  return (1u << width) - 1u;
}

// Digit of key k at `shift`.  8-bit digits never straddle a 32-bit half of
// a 64-bit key (digit_of's fast form); wider ones may, and take the
// 64-bit shift.
template <int DB, class K>
__device__ __forceinline__ uint32_t dig(K k, uint32_t shift, uint32_t m) {
  if constexpr (DB == 8 || sizeof(K) <= 4) {
    return digit_of(k, shift, m);
  } else {
    static_assert(sizeof(K) == 8, "digits wider than 8 bits are for keys of up to 8 bytes");
    return static_cast<uint32_t>(k >> shift) & m;
  }
}

// Histogram of every digit position in one read of the keys.  With an
// aligned key pointer each thread reads 16 bytes at a time.
template <int KB, bool VEC, int DB>
__global__ __launch_bounds__(256) void k_radix_hist(const typename Elem<KB>::T* __restrict__ keys,
                                                    uint64_t n, int passes, uint32_t key_bits,
                                                    uint32_t* __restrict__ hist) {
  using K = typename Elem<KB>::T;
  constexpr int kBins = kBinsOf<DB>;
  constexpr int PER = (VEC && KB < 16) ? 16 / KB : 1;
  __shared__ uint32_t lh[(8 * KB + DB - 1) / DB * kBins];
  for (int i = threadIdx.x; i < passes * kBins; i += 256) lh[i] = 0;
  __syncthreads();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  const uint64_t gtid = blockIdx.x * 256ull + threadIdx.x;
  auto count = [&](K k) {
    for (int p = 0; p < passes; ++p)
      atomicAdd(&lh[p * kBins + dig<DB>(k, DB * p, pass_mask<DB>(p, key_bits))], 1u);
  };
  uint64_t done = 0;
  if constexpr (PER > 1) {
    // A 16-byte vector whose keys share a digit adds once; when every lane's
    // vector shares one digit (sorted or low-cardinality keys), the first
    // lane adds for the whole wave instead of 64 lanes hitting one word.
    struct alignas(16) Vec { K k[PER]; };
    const Vec* kv = reinterpret_cast<const Vec*>(keys);
    const uint64_t nvec = n / PER;
    for (uint64_t v = gtid; v < nvec; v += stride) {
      const Vec x = kv[v];
      const uint64_t active = ballot(true);
      for (int p = 0; p < passes; ++p) {
        const uint32_t sh = DB * p, m = pass_mask<DB>(p, key_bits);
        const uint32_t d = dig<DB>(x.k[0], sh, m);
        bool same = true;
#pragma unroll
        for (int e = 1; e < PER; ++e) same &= dig<DB>(x.k[e], sh, m) == d;
        if (ballot(same) == active) {
          const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
          if (ballot(d != d0) == 0) {
            if (mbcnt(active) == 0)
              atomicAdd(&lh[p * kBins + d], static_cast<uint32_t>(PER * __popcll(active)));
            continue;
          }
        }
        if (same) {
          atomicAdd(&lh[p * kBins + d], static_cast<uint32_t>(PER));
        } else {
#pragma unroll
          for (int e = 0; e < PER; ++e) atomicAdd(&lh[p * kBins + dig<DB>(x.k[e], sh, m)], 1u);
        }
      }
    }
    done = nvec * PER;
  }
  for (uint64_t i = done + gtid; i < n; i += stride) count(keys[i]);
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kBins; i += 256)
    if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// Exclusive scan of each digit histogram (one thread per bin); flags
// identity passes.
template <int DB>
__global__ __launch_bounds__(kBinsOf<DB>) void k_radix_scan(const uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ base,
                                                            uint32_t* __restrict__ skip,
                                                            uint64_t n) {
  constexpr int kBins = kBinsOf<DB>;
  __shared__ uint32_t tmp[kBins / kWave + 1];
  const int p = blockIdx.x;
  const uint32_t c = hist[p * kBins + threadIdx.x];
  uint32_t total;
  uint32_t e = block_excl_scan<kBins>(c, tmp, &total);
  base[p * kBins + threadIdx.x] = e;
  if (c == n) skip[p] = 1u;  // every key has the same digit here
}

template <int KB, int VB, int NT, int IPT, int DB, class S>
__global__ __launch_bounds__(NT, RFX_SORT_MINW) void k_onesweep(
    typename Elem<KB>::T* __restrict__ k0, typename Elem<KB>::T* __restrict__ k1,
    typename Elem<VB>::T* __restrict__ v0, typename Elem<VB>::T* __restrict__ v1, uint64_t n,
    int pass, uint32_t key_bits, const uint32_t* __restrict__ base, uint32_t* __restrict__ ctrs,
    const uint32_t* __restrict__ skip, uint32_t* __restrict__ sel, S* __restrict__ status,
    uint32_t ntiles) {
  using K = typename Elem<KB>::T;
  using V = typename Elem<VB>::T;
  constexpr int NW = NT / kWave;
  constexpr int TILE = NT * IPT;
  constexpr int RB = (KB > VB ? KB : VB) * TILE;
  constexpr int kBins = kBinsOf<DB>;
  constexpr int BPT = (kBins + NT - 1) / NT;  // bins per thread in the per-bin steps
  // Wave counters and tile positions stay below TILE: 16 bits hold them,
  // which keeps the wider digits' counters within LDS.
  using WC = typename std::conditional<(DB > 8), uint16_t, uint32_t>::type;
  using DG = typename std::conditional<(DB > 8), uint16_t, uint8_t>::type;
  static_assert(TILE <= 65535 || DB == 8, "16-bit wave counters");

  if (skip[pass]) return;
  const uint32_t cur = *sel;
  const K* __restrict__ kin = cur ? k1 : k0;
  K* __restrict__ kout = cur ? k0 : k1;
  const V* __restrict__ vin = cur ? v1 : v0;
  V* __restrict__ vout = cur ? v0 : v1;

  __shared__ uint32_t s_tile;
  __shared__ WC wcnt[NW][kBins];
  __shared__ uint32_t gdelta[kBins];
  __shared__ uint32_t scan_tmp[NW + 1];
  __shared__ __attribute__((aligned(16))) unsigned char rbuf[RB];
  __shared__ DG sdig[VB > 0 ? TILE : 1];  // digit of each reordered row

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t w = wave_id();
  const uint32_t shift = static_cast<uint32_t>(DB * pass);
  const uint32_t dmask = pass_mask<DB>(pass, key_bits);

  if (tid == 0) s_tile = atomicAdd(&ctrs[pass * 2], 1u);
  for (uint32_t i = tid; i < NW * kBins; i += NT) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t tile_base = static_cast<uint64_t>(tile) * TILE;
  const uint64_t wave_base = tile_base + static_cast<uint64_t>(w) * kWave * IPT;
  using SW = StatusWord<S>;
  S* st = status + (static_cast<uint64_t>(pass) * ntiles) * kBins;

  // Every tile but the last is full: its copy of the body has no per-key
  // bounds checks.
  auto body = [&](auto full_tag) {
    constexpr bool FULL = decltype(full_tag)::value;
    // Keys only: values are read after the keys are ranked, which keeps
    // the ranking phase's register footprint small.
    K keys[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint64_t idx = wave_base + static_cast<uint64_t>(j) * kWave + lane;
      if (FULL) keys[j] = kin[idx];
      else keys[j] = idx < n ? kin[idx] : K{};
    }

    // Stable rank within the wave: lanes sharing a digit are found by one
    // ballot per digit bit; the highest such lane bumps the wave counter.
    // (An LDS atomic per group leader, issued for all keys before any
    // result is used, measured slower on MI355X: 1.67 against 1.45 ms per
    // pass.)
    uint32_t pos[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint64_t idx = wave_base + static_cast<uint64_t>(j) * kWave + lane;
      const bool valid = FULL || idx < n;
      const uint32_t d = dig<DB>(keys[j], shift, dmask);
      // acc collects, per lane, the lanes whose digit differs in some bit:
      // x is the lane's bit b as 0 or all ones, and acc | (bb ^ x) is one
      // bitop3 per 32-bit half (table 0xDE with acc in the middle operand
      // is the same under either operand order of the truth table).  Bits
      // above a narrow last digit are zero everywhere and change nothing.
      uint32_t acc_lo = 0, acc_hi = 0;
#pragma unroll
      for (int b = 0; b < DB; ++b) {
        const uint32_t x = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(d), b, 1));
        // The ballot of x itself (a plain ballot(x != 0) is rewritten into
        // a second shift-and-compare of d); 33 is the not-equal predicate.
        // Lanes outside exec give 0, as a ballot does.
        const uint64_t bb = __builtin_amdgcn_uicmp(x, 0u, 33);
        acc_lo = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(bb), acc_lo, x, 0xDE);
        acc_hi = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(bb >> 32), acc_hi, x, 0xDE);
      }
      const uint64_t m =
          (FULL ? ~0ull : ballot(valid)) & ~((static_cast<uint64_t>(acc_hi) << 32) | acc_lo);
      const uint32_t below = mbcnt(m);
      const uint32_t cnt = __popcll(m);
      uint32_t prior = 0;
      if (valid) {
        prior = wcnt[w][d];
        if (below == cnt - 1) wcnt[w][d] = static_cast<WC>(prior + cnt);
      }
      pos[j] = prior + below;
    }
    __syncthreads();

    // Per-digit tile totals and wave prefixes; publish the tile aggregate.
    // wcnt[w][d] becomes the tile position of wave w's first key of digit
    // d: the digit's start in the tile plus the waves before w.
    // Thread t handles bins t, t + NT, ... (one bin, or none, for 8 and 9
    // bits).
    uint32_t total_d[BPT], bs[BPT];
    uint32_t wpre[BPT][NW];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const uint32_t bin = tid + q * NT;
      total_d[q] = 0;
      if (bin < kBins) {
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
          wpre[q][ww] = run;
          run += wcnt[ww][bin];
        }
        total_d[q] = run;
        sw_store<S>(st + static_cast<uint64_t>(tile) * kBins + bin,
                    (tile == 0 ? SW::kIncl : SW::kAgg) | static_cast<S>(run));
      }
    }
    {
      uint32_t carry = 0;
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        uint32_t all;
        bs[q] = carry + block_excl_scan<NT>(total_d[q], scan_tmp, &all);
        carry += all;
      }
    }
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const uint32_t bin = tid + q * NT;
      if (bin < kBins) {
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) wcnt[ww][bin] = static_cast<WC>(bs[q] + wpre[q][ww]);
      }
    }
    __syncthreads();

    // Work that does not need the global offsets goes before the look-back
    // so it overlaps the wait: the LDS reorder of the keys and the value
    // loads.
    K* kbuf = reinterpret_cast<K*>(rbuf);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint64_t idx = wave_base + static_cast<uint64_t>(j) * kWave + lane;
      if (FULL || idx < n) {
        const uint32_t d = dig<DB>(keys[j], shift, dmask);
        pos[j] += wcnt[w][d];
        kbuf[pos[j]] = keys[j];
      }
    }
    V vals[IPT];
    if constexpr (VB > 0) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint64_t idx = wave_base + static_cast<uint64_t>(j) * kWave + lane;
        if (FULL || idx < n) vals[j] = vin[idx];
      }
    }
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const uint32_t bin = tid + q * NT;
      if (bin < kBins) {
        uint32_t excl = 0;
        if (tile > 0) {
          excl = sort_lookback<RFX_SORT_LBW>(st, tile, kBins, bin);
          sw_store<S>(st + static_cast<uint64_t>(tile) * kBins + bin,
                      SW::kIncl | static_cast<S>(excl + total_d[q]));
        }
        gdelta[bin] = base[pass * kBins + bin] + excl - bs[q];
      }
    }
    __syncthreads();

    const uint32_t tile_n =
        FULL ? static_cast<uint32_t>(TILE)
             : static_cast<uint32_t>((n - tile_base) < TILE ? (n - tile_base) : TILE);
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const uint32_t i = k * NT + tid;
      if (FULL || i < tile_n) {
        const K key = kbuf[i];
        const uint32_t d = dig<DB>(key, shift, dmask);
        if constexpr (VB > 0) sdig[i] = static_cast<DG>(d);
        const uint32_t dst = gdelta[d] + i;
        if (dst < n) kout[dst] = key;  // always true when ranks are right
      }
    }
    if constexpr (VB > 0) {
      __syncthreads();  // every key is out of kbuf
      V* vbuf = reinterpret_cast<V*>(rbuf);
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint64_t idx = wave_base + static_cast<uint64_t>(j) * kWave + lane;
        if (FULL || idx < n) vbuf[pos[j]] = vals[j];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < IPT; ++k) {
        const uint32_t i = k * NT + tid;
        if (FULL || i < tile_n) {
          const uint32_t dst = gdelta[sdig[i]] + i;
          if (dst < n) vout[dst] = vbuf[i];
        }
      }
    }
  };
  if (tile_base + TILE <= n) body(std::true_type{});
  else body(std::false_type{});

  // The last tile to finish flips the buffer selector.  Every block read
  // *sel before arriving here, so nobody reads the new value this pass.
  if (tid == 0) {
    const uint32_t prev = atomicAdd(&ctrs[pass * 2 + 1], 1u);
    if (prev == ntiles - 1)
      __hip_atomic_store(sel, cur ^ 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int KB, int VB>
__global__ __launch_bounds__(256) void k_copyback(typename Elem<KB>::T* __restrict__ k0,
                                                  const typename Elem<KB>::T* __restrict__ k1,
                                                  typename Elem<VB>::T* __restrict__ v0,
                                                  const typename Elem<VB>::T* __restrict__ v1,
                                                  uint64_t n, const uint32_t* __restrict__ sel) {
  if (*sel == 0) return;
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    k0[i] = k1[i];
    if constexpr (VB > 0) v0[i] = v1[i];
  }
}

// Tunable tile shapes: the autotuner may pick 8, 12 or 16 keys per thread for
// 4- and 8-byte keys with 0-, 4- or 8-byte values (other shapes keep the
// default).  The smallest tile needs the most look-back status words.
template <int KB, int VB>
constexpr bool kSortTunable = (KB == 4 || KB == 8) && (VB == 0 || VB == 4 || VB == 8);
constexpr int kSortMinIPT = 8;

uint64_t sort_ws_tile(uint64_t n, int kb, int vb, int key_bits, uint64_t tile, int db,
                      bool wide = false) {
  const int passes = (key_bits + db - 1) / db;
  const uint64_t tiles = (n + tile - 1) / tile;
  Carve s(nullptr);
  SortLayout L;
  carve_sort(s, n, kb, vb, passes, tiles, 1u << db, &L, wide);
  return s.off;
}

// Sorts of more than 2^30 - 1 rows take 64-bit look-back status words, at
// the default tile and 8-bit digits.
constexpr uint64_t kSortMaxRows = (1ull << 31) - 1;

// Digit width chosen for key_bits when the context does not set one.  An
// odd number of passes leaves the data in the pong buffer and costs a copy
// back (about 0.87 ms for 2^28 u32 pairs, a full read and write), so 17-20
// bit keys take two passes of 9 or 10 bits instead of three of 8: 2^28
// pairs, 18 bits 3.36 ms against 5.04, 20 bits 4.04 against 4.90.  Wider
// keys keep 8 bits (26 bits: 5.51 ms against 5.65 with 9; 32 bits: 5.56
// against 6.29).
constexpr int auto_digit(int key_bits) {
  return key_bits >= 17 && key_bits <= 18 ? 9 : key_bits >= 19 && key_bits <= 20 ? 10 : 8;
}

// Workspace for any tile shape the context may be tuned to, at the digit
// width chosen for key_bits (a context tuned to other digit widths grows
// its own workspace on first use).
template <int KB, int VB>
uint64_t sort_ws_bytes(uint64_t n, int key_bits) {
  using Tn = SortTune<KB, VB>;
  const int ipt = kSortTunable<KB, VB> && kSortMinIPT < Tn::IPT ? kSortMinIPT : Tn::IPT;
  if (n > kCountMask)
    return sort_ws_tile(n, KB, VB, key_bits, static_cast<uint64_t>(Tn::NT) * Tn::IPT, 8, true);
  uint64_t ws = sort_ws_tile(n, KB, VB, key_bits, static_cast<uint64_t>(Tn::NT) * ipt, 8);
  if (kSortTunable<KB, VB> && auto_digit(key_bits) != 8) {
    const int wide_ipt = Tn::IPT < 12 ? Tn::IPT : 12;
    const uint64_t w2 = sort_ws_tile(n, KB, VB, key_bits,
                                     static_cast<uint64_t>(Tn::NT) * wide_ipt,
                                     auto_digit(key_bits));
    if (w2 > ws) ws = w2;
  }
  return ws;
}

template <int KB, int VB, int IPT, int DB = 8, class S = uint32_t>
int sort_impl(rfx_ctx* ctx, void* keys, void* vals, uint64_t n, int key_bits) {
  using K = typename Elem<KB>::T;
  using V = typename Elem<VB>::T;
  constexpr int NT = SortTune<KB, VB>::NT;
  constexpr int TILE = NT * IPT;
  constexpr bool kWide = sizeof(S) == 8;
  if (n < 2) return RFX_OK;
  if (n > (kWide ? kSortMaxRows : static_cast<uint64_t>(kCountMask)))
    return fail(RFX_ERANGE, "sort: %llu rows exceeds 2^31-1", static_cast<unsigned long long>(n));
  constexpr uint32_t kBins = kBinsOf<DB>;
  const int passes = (key_bits + DB - 1) / DB;
  const uint64_t tiles = (n + TILE - 1) / TILE;
  RFX_TRY(ensure_workspace(ctx, sort_ws_tile(n, KB, VB, key_bits, TILE, DB, kWide)));
  Carve c(ctx->ws);
  SortLayout L;
  carve_sort(c, n, KB, VB, passes, tiles, kBins, &L, kWide);
  RFX_TRY(fill_u32(ctx, L.hist, L.zero_words, 0u));

  uint64_t hb = (n + 256 * 16 - 1) / (256 * 16);
  uint64_t hmax = static_cast<uint64_t>(ctx->device_cus) * 8;
  if (hb > hmax) hb = hmax;
  if (hb == 0) hb = 1;
  {
    RFX_PROF(ctx, "sort.hist");
    if (reinterpret_cast<uintptr_t>(keys) % 16 == 0)
      k_radix_hist<KB, true, DB><<<static_cast<uint32_t>(hb), 256, 0, ctx->stream>>>(
          static_cast<const K*>(keys), n, passes, static_cast<uint32_t>(key_bits), L.hist);
    else
      k_radix_hist<KB, false, DB><<<static_cast<uint32_t>(hb), 256, 0, ctx->stream>>>(
          static_cast<const K*>(keys), n, passes, static_cast<uint32_t>(key_bits), L.hist);
    RFX_LAUNCHED();
  }
  k_radix_scan<DB><<<passes, kBins, 0, ctx->stream>>>(L.hist, L.base, L.skip, n);
  RFX_LAUNCHED();
  for (int p = 0; p < passes; ++p) {
    RFX_PROF(ctx, "sort.onesweep");
    k_onesweep<KB, VB, NT, IPT, DB, S><<<static_cast<uint32_t>(tiles), NT, 0, ctx->stream>>>(
        static_cast<K*>(keys), static_cast<K*>(L.pong_keys), static_cast<V*>(vals),
        static_cast<V*>(L.pong_vals), n, p, static_cast<uint32_t>(key_bits), L.base, L.ctrs,
        L.skip, L.sel, static_cast<S*>(L.status), static_cast<uint32_t>(tiles));
    RFX_LAUNCHED();
  }
  uint64_t cb = (n + 255) / 256;
  if (cb > 8192) cb = 8192;
  k_copyback<KB, VB><<<static_cast<uint32_t>(cb), 256, 0, ctx->stream>>>(
      static_cast<K*>(keys), static_cast<const K*>(L.pong_keys), static_cast<V*>(vals),
      static_cast<const V*>(L.pong_vals), n, L.sel);
  RFX_LAUNCHED();
  return RFX_OK;
}

template <int KB, int VB>
int sort_tuned(rfx_ctx* ctx, void* keys, void* vals, uint64_t n, int key_bits) {
  // More than 2^30 - 1 rows (or the sort.status knob at 64): 64-bit status
  // words, default tile, 8-bit digits.
  if (n > kCountMask || ctx->tune_sort_status == 64)
    return sort_impl<KB, VB, SortTune<KB, VB>::IPT, 8, uint64_t>(ctx, keys, vals, n, key_bits);
  if constexpr (kSortTunable<KB, VB>) {
    // Wider digits: the default tile, or 12 keys per thread.  Knob value 0
    // takes the digit width chosen for key_bits.
    constexpr int D = SortTune<KB, VB>::IPT;
    const int digit = ctx->tune_sort_digit ? ctx->tune_sort_digit : auto_digit(key_bits);
    if (digit == 9)
      return ctx->tune_sort_ipt == 12 ? sort_impl<KB, VB, 12, 9>(ctx, keys, vals, n, key_bits)
                                      : sort_impl<KB, VB, D, 9>(ctx, keys, vals, n, key_bits);
    if (digit == 10)
      return ctx->tune_sort_ipt == 12 ? sort_impl<KB, VB, 12, 10>(ctx, keys, vals, n, key_bits)
                                      : sort_impl<KB, VB, D, 10>(ctx, keys, vals, n, key_bits);
    if (ctx->tune_sort_ipt == 8) return sort_impl<KB, VB, 8>(ctx, keys, vals, n, key_bits);
    if (ctx->tune_sort_ipt == 12) return sort_impl<KB, VB, 12>(ctx, keys, vals, n, key_bits);
    if (ctx->tune_sort_ipt == 16) return sort_impl<KB, VB, 16>(ctx, keys, vals, n, key_bits);
  }
  return sort_impl<KB, VB, SortTune<KB, VB>::IPT>(ctx, keys, vals, n, key_bits);
}

template <int KB>
int sort_dispatch_v(rfx_ctx* ctx, void* keys, void* vals, uint64_t n, int key_bits, int vb) {
  switch (vb) {
    case 0: return sort_tuned<KB, 0>(ctx, keys, nullptr, n, key_bits);
    case 1: return sort_tuned<KB, 1>(ctx, keys, vals, n, key_bits);
    case 2: return sort_tuned<KB, 2>(ctx, keys, vals, n, key_bits);
    case 4: return sort_tuned<KB, 4>(ctx, keys, vals, n, key_bits);
    case 8: return sort_tuned<KB, 8>(ctx, keys, vals, n, key_bits);
    case 16: return sort_tuned<KB, 16>(ctx, keys, vals, n, key_bits);
    default: return fail(RFX_EINVAL, "sort: value width %d", vb);
  }
}

template <int KB>
uint64_t ws_dispatch_v(uint64_t n, int key_bits, int vb) {
  switch (vb) {
    case 0: return sort_ws_bytes<KB, 0>(n, key_bits);
    case 1: return sort_ws_bytes<KB, 1>(n, key_bits);
    case 2: return sort_ws_bytes<KB, 2>(n, key_bits);
    case 4: return sort_ws_bytes<KB, 4>(n, key_bits);
    case 8: return sort_ws_bytes<KB, 8>(n, key_bits);
    case 16: return sort_ws_bytes<KB, 16>(n, key_bits);
    default: return 0;
  }
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_sort_pairs(rfx_ctx* ctx, void* keys, void* vals, uint64_t n, int key_bytes,
                              int key_bits, int val_bytes) {
  if (!ctx) return fail(RFX_EINVAL, "rfx_sort_pairs: null context");
  if (!valid_width(key_bytes)) return fail(RFX_EINVAL, "rfx_sort_pairs: key width %d", key_bytes);
  if (val_bytes != 0 && !valid_width(val_bytes))
    return fail(RFX_EINVAL, "rfx_sort_pairs: value width %d", val_bytes);
  if (key_bits <= 0 || key_bits > 8 * key_bytes)
    return fail(RFX_EINVAL, "rfx_sort_pairs: key_bits %d for %d-byte keys", key_bits, key_bytes);
  if (n && !keys) return fail(RFX_EINVAL, "rfx_sort_pairs: null keys");
  if (n && val_bytes && !vals) return fail(RFX_EINVAL, "rfx_sort_pairs: null values");
  switch (key_bytes) {
    case 1: return sort_dispatch_v<1>(ctx, keys, vals, n, key_bits, val_bytes);
    case 2: return sort_dispatch_v<2>(ctx, keys, vals, n, key_bits, val_bytes);
    case 4: return sort_dispatch_v<4>(ctx, keys, vals, n, key_bits, val_bytes);
    case 8: return sort_dispatch_v<8>(ctx, keys, vals, n, key_bits, val_bytes);
    case 16: return sort_dispatch_v<16>(ctx, keys, vals, n, key_bits, val_bytes);
  }
  return fail(RFX_EINVAL, "rfx_sort_pairs: key width %d", key_bytes);
}

extern "C" uint64_t rfx_sort_workspace(uint64_t n, int key_bytes, int key_bits, int val_bytes) {
  if (!valid_width(key_bytes) || key_bits <= 0 || key_bits > 8 * key_bytes) return 0;
  switch (key_bytes) {
    case 1: return ws_dispatch_v<1>(n, key_bits, val_bytes);
    case 2: return ws_dispatch_v<2>(n, key_bits, val_bytes);
    case 4: return ws_dispatch_v<4>(n, key_bits, val_bytes);
    case 8: return ws_dispatch_v<8>(n, key_bits, val_bytes);
    case 16: return ws_dispatch_v<16>(n, key_bits, val_bytes);
  }
  return 0;
}
