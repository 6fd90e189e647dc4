// join.hip - sort-merge join core for gfx950.
//
//  find_bounds   one pass over the merge path of the sorted left and right
//                key columns: each left key's lower bound in the right
//                column, its equal-key count (galloping from the lower
//                bound, mostly inside the tile's LDS copy of the right
//                keys), and the exclusive scan of the counts across tiles by
//                decoupled look-back.  Replaces find_bounds_{8..128}
//                (ModernGPUJoin.cu:109-663: SortedSearch + SortedEquality-
//                Count + Scan + an H2D copy of the total).
//  expand        load-balanced expansion of the scanned counts into
//                (left, right) index pairs in left-major, right-minor order:
//                a merge path over the output positions and the row starts.
//                Replaces redfox::join (ModernGPUJoin.cu:665-733 ->
//                KernelLeftJoin, join.cuh:45-99).
//  gather        builds output tuples from the index pairs (gather_*,
//                ModernGPUJoin.h:43-200); expand_gather fuses the two and
//                never writes the index columns.
#include <mutex>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {


// ---- element access by runtime width -------------------------------------
__device__ __forceinline__ u128 load_w(const void* base, uint64_t i, int w) {
  switch (w) {
    case 1: return {static_cast<const uint8_t*>(base)[i], 0};
    case 2: return {static_cast<const uint16_t*>(base)[i], 0};
    case 4: return {static_cast<const uint32_t*>(base)[i], 0};
    case 8: return {static_cast<const uint64_t*>(base)[i], 0};
    default: return static_cast<const u128*>(base)[i];
  }
}
__device__ __forceinline__ void store_w(void* base, uint64_t i, int w, u128 v) {
  switch (w) {
    case 1: static_cast<uint8_t*>(base)[i] = static_cast<uint8_t>(v.lo); break;
    case 2: static_cast<uint16_t*>(base)[i] = static_cast<uint16_t>(v.lo); break;
    case 4: static_cast<uint32_t*>(base)[i] = static_cast<uint32_t>(v.lo); break;
    case 8: static_cast<uint64_t*>(base)[i] = v.lo; break;
    default: static_cast<u128*>(base)[i] = v; break;
  }
}

// Number of A elements among the first `diag` merged elements when A wins
// ties (A[i] <= B[j] puts A[i] first).
template <class K>
__device__ __forceinline__ uint64_t merge_path_lower(const K* a, uint64_t na, const K* b,
                                                     uint64_t nb, uint64_t diag) {
  uint64_t lo = diag > nb ? diag - nb : 0;
  uint64_t hi = diag < na ? diag : na;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (!lt(b[diag - 1 - mid], a[mid])) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// find_bounds
// ---------------------------------------------------------------------------
#ifndef RFX_FB_NV
#define RFX_FB_NV 4096
#endif
// find_bounds tile: merged (left + right) elements per block.
constexpr int kFbNV = RFX_FB_NV;
constexpr int kFbNT = 256;

template <class K>
__global__ __launch_bounds__(256) void k_fb_partition(const K* __restrict__ a, uint64_t na,
                                                      const K* __restrict__ b, uint64_t nb,
                                                      uint64_t ntiles, uint32_t nv,
                                                      uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles) return;
  uint64_t diag = t * nv;
  if (diag > na + nb) diag = na + nb;
  split[t] = merge_path_lower(a, na, b, nb, diag);
}

// The largest key of each width: the sentinel that follows a tile's right
// keys in LDS, so that fixed-step probes past them stay in the array.
template <class K>
__device__ __forceinline__ K max_key() {
  if constexpr (sizeof(K) == 16) return K{~0ull, ~0ull};
  else return static_cast<K>(~static_cast<K>(0));
}

// LDS index of count i: one pad word per 32, so that a thread reading its
// VT consecutive counts and lanes writing thread-strided counts both stay
// (nearly) free of bank conflicts.
__device__ __forceinline__ uint32_t cnt_slot(uint32_t i) { return i + (i >> 5); }

// Sub-tiles per block.  Blocks are ordered by an atomic ticket, so that the
// decoupled look-back only ever waits on blocks already running.  Device-
// scope atomics on one word complete one after another, though: with one
// ticket per 4096-key tile they bounded the kernel (2^24 x 2^26 keys: 0.25
// ms with the searches and the look-back taken out, 0.09 ms with the ticket
// taken out as well).  A block of kFbSub x kFbNT threads therefore works on
// kFbSub consecutive tiles at once, one per kFbNT-thread group, under one
// ticket and one status word.
constexpr int kFbSub = 4;

// One tile of the merge path per thread group: its left keys A[a0, a1) and
// right keys B[b0, b1).  With A winning ties, every B[j] before b0 is <
// A[a0] and every B[j] from b1 on is >= A[a1 - 1], so each left key's lower
// bound lies in [b0, b1].  The tile's right keys are followed in LDS by the
// next NT right keys (B[b1, b1 + NT)), loaded with the rest, so the equal
// run of a left key that crosses the tile's end - the last key of nearly
// every tile - is measured in LDS too; only a run longer than that
// continues in global memory (a gallop from b1 + NT).  Lower and upper
// bounds come from fixed-step binary searches in LDS, four left keys per
// thread in lockstep (neighbouring lanes probe the same words).  The counts
// are then scanned in left-key order within the tile, the block's tiles are
// added up in order, and the block takes its prefix from the blocks before
// it by decoupled look-back.
template <class K, int NT, int NV>
__global__ __launch_bounds__(NT * kFbSub, 8) void k_find_bounds(
    const K* __restrict__ a, uint64_t na, const K* __restrict__ b, uint64_t nb,
    const uint64_t* __restrict__ split, uint32_t ntiles, int32_t* __restrict__ lower_bound,
    int32_t* __restrict__ left_count, uint64_t* __restrict__ total,
    uint64_t* __restrict__ status, uint32_t* __restrict__ ticket) {
  constexpr int VT = NV / NT;
  static_assert(VT % 4 == 0, "left keys are searched four per thread at a time");
  constexpr int EXT = NT;
  constexpr int NWS = NT / kWave;  // waves per tile group
  constexpr int kRunWin = 31;      // run window: fixed steps 16, 8, 4, 2, 1
  constexpr int KEY_BYTES = (NV + EXT + kRunWin + 1) * static_cast<int>(sizeof(K));
  constexpr int CNT_BYTES = (NV + NV / 32) * 4;
  constexpr int SUB_BYTES = ((KEY_BYTES > CNT_BYTES ? KEY_BYTES : CNT_BYTES) + 15) / 16 * 16;
  // The keys are dead once the searches are done; the counts reuse them.
  __shared__ __attribute__((aligned(16))) unsigned char smem[kFbSub * SUB_BYTES];
  __shared__ uint64_t s_wave[kFbSub * NWS];
  __shared__ uint64_t s_sub[kFbSub];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;

  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t blk = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  // The tile group is the same for a whole wave (NT is a multiple of 64):
  // readfirstlane lets the compiler keep the tile's bounds, and everything
  // derived from them, in scalar registers.
  const uint32_t sub = __builtin_amdgcn_readfirstlane(threadIdx.x / NT);
  const uint32_t tid = threadIdx.x % NT;
  const uint32_t tile = blk * kFbSub + sub;
  K* keys = reinterpret_cast<K*>(smem + sub * SUB_BYTES);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(smem + sub * SUB_BYTES);

  // A group past the last tile holds an empty tile.
  uint64_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
  if (tile < ntiles) {
    const uint64_t d0 = static_cast<uint64_t>(tile) * NV;
    const uint64_t d1 = (d0 + NV < na + nb) ? d0 + NV : na + nb;
    a0 = split[tile];
    a1 = split[tile + 1];
    b0 = d0 - a0;
    b1 = d1 - a1;
  }
  const uint32_t nat = static_cast<uint32_t>(a1 - a0), nbt = static_cast<uint32_t>(b1 - b0);
  const uint32_t ntot = nat + nbt;
  const uint32_t next = tile < ntiles ? static_cast<uint32_t>(nb - b1 < EXT ? nb - b1 : EXT) : 0u;

  // The tile's A keys, its B keys, then the B keys past its end: every load
  // in flight before the first LDS write.  Merged index i reads A[a0 + i]
  // below nat and B[b0 + i - nat] from there.  Row j of the tile (indices
  // j*NT .. j*NT + NT - 1) lies wholly in A or wholly in B but for one row,
  // so the side is a uniform branch and the address a scalar base plus the
  // thread's offset.
  const K* __restrict__ at = a + a0;
  const K* __restrict__ bt = b + b0;
  const K* __restrict__ bx = b + b1;
  {
    K v[VT];
    K ev{};
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t i0 = j * NT, i = i0 + tid;
      v[j] = K{};
      if (i0 + NT <= nat) {
        v[j] = at[i];
      } else if (i0 >= nat && i0 + NT <= ntot) {
        v[j] = bt[i - nat];
      } else if (i < ntot) {
        v[j] = i < nat ? at[i] : bt[i - nat];
      }
    }
    if (tid < next) ev = bx[tid];
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t i = tid + j * NT;
      if (i < ntot) keys[i] = v[j];
    }
    if (tid < next) keys[ntot + tid] = ev;
    // Sentinels after the right keys: the run window's probes never leave
    // the array, and a count is clamped to the keys actually there.
    if (tid <= static_cast<uint32_t>(kRunWin)) keys[ntot + next + tid] = max_key<K>();
  }
  __syncthreads();
  const K* sa = keys;
  const K* sb = keys + nat;
  const uint32_t nbx = nbt + next;  // right keys in LDS
  int32_t* __restrict__ lbt = lower_bound + a0;

  // Left key i = tid + j * NT; JA (uniform over the group) bounds j.  The
  // lower bound by Shar's fixed-step binary search of the tile's right keys
  // (one probe at the largest power of two P <= nbt picks the window [0, P)
  // or [nbt - P, nbt), then steps P/2 .. 1 and a last probe); the equal run
  // after it by a fixed-step search of the next kRunWin right keys, widened
  // to all of them only for a wave in which some run is longer.  Positions
  // are byte offsets into the right keys and the steps compile-time, so a
  // probe is one LDS read at a constant offset, a compare and a select.
  const uint32_t JA = (nat + NT - 1) / NT;
  const uint32_t top_lo = nbt ? 1u << (31 - __builtin_clz(nbt)) : 0u;
  constexpr uint32_t SZ = sizeof(K);
  constexpr int kLogNV = __builtin_ctz(NV);
  uint32_t cnt[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) cnt[j] = 0;
#pragma unroll
  for (int jb = 0; jb < VT; jb += 4) {
    if (static_cast<uint32_t>(jb) < JA) {
      K kv[4];
      uint32_t lo[4], up[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t i = tid + (jb + q) * NT;
        kv[q] = sa[i < nat ? i : nat - 1];
        lo[q] = 0;
      }
      const char* sbb = reinterpret_cast<const char*>(sb);
      uint32_t off[4] = {0, 0, 0, 0};
      if (top_lo) {
        const K pivot = sb[top_lo - 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) off[q] = lt(pivot, kv[q]) ? (nbt - top_lo) * SZ : 0u;
#pragma unroll
        for (int sh = kLogNV - 1; sh >= 0; --sh) {
          const uint32_t step = 1u << sh;
          if (step < top_lo) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const K x = *reinterpret_cast<const K*>(sbb + off[q] + (step - 1) * SZ);
              off[q] = lt(x, kv[q]) ? off[q] + step * SZ : off[q];
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const K x = *reinterpret_cast<const K*>(sbb + off[q]);
          off[q] = lt(x, kv[q]) ? off[q] + SZ : off[q];
        }
      }
      // sb[lo, up) == key throughout: sb[lo..] >= key, so the upper bound
      // is the end of the run of equal keys.
      bool longer = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        lo[q] = off[q] / SZ;
        uint32_t u = off[q];
#pragma unroll
        for (int sh = 4; sh >= 0; --sh) {
          const uint32_t step = 1u << sh;
          const K x = *reinterpret_cast<const K*>(sbb + u + (step - 1) * SZ);
          u = lt(kv[q], x) ? u : u + step * SZ;
        }
        const uint32_t uw = u / SZ;
        up[q] = uw < nbx ? uw : nbx;
        longer |= uw == lo[q] + kRunWin && uw < nbx;
      }
      if (ballot(longer)) {
        // Some run fills the window: finish those from its end.
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (up[q] == lo[q] + kRunWin && up[q] < nbx) {
            uint32_t u = 0;
            const uint32_t rem = nbx - up[q];
            const uint32_t top = 1u << (31 - __builtin_clz(rem));
            for (uint32_t step = top; step; step >>= 1)
              if (u + step <= rem && !lt(kv[q], sb[up[q] + u + step - 1])) u += step;
            up[q] += u;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t i = tid + (jb + q) * NT;
        if (i < nat) {
          uint64_t c = up[q] - lo[q];
          // A run reaching past the keys in LDS continues in global memory.
          const uint64_t bxe = b1 + next;
          if (up[q] == nbx && bxe < nb && eq(b[bxe], kv[q])) {
            uint64_t l = bxe + 1, st = 1, h;
            while (true) {  // invariant: b[l - 1] == key
              h = l + st - 1;
              if (h >= nb) { h = nb; break; }
              if (!eq(b[h], kv[q])) break;
              l = h + 1;
              st <<= 1;
            }
            while (l < h) {  // first index in [l, h) whose key differs
              const uint64_t mid = (l + h) >> 1;
              if (eq(b[mid], kv[q])) l = mid + 1;
              else h = mid;
            }
            c += l - bxe;
          }
          lbt[i] = static_cast<int32_t>(b0 + lo[q]);
          cnt[jb + q] = static_cast<uint32_t>(c);
        }
      }
    }
  }
  __syncthreads();  // every search is done with the keys
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint32_t i = tid + j * NT;
    if (i < nat) s_cnt[cnt_slot(i)] = cnt[j];
  }
  __syncthreads();

  // Exclusive scan of the counts in left-key order within the tile: thread
  // t of the group owns elements [t*VT, t*VT + VT).  They are read again
  // afterwards rather than held in registers through the look-back.
  const uint32_t c0 = tid * VT;
  uint64_t my_sum = 0;
#pragma unroll
  for (int j = 0; j < VT; ++j) my_sum += (c0 + j < nat) ? s_cnt[cnt_slot(c0 + j)] : 0u;
  const uint32_t w = wave_id();
  const uint64_t incl = wave_incl_scan(my_sum);
  if (lane_id() == kWave - 1) s_wave[w] = incl;
  __syncthreads();
  uint64_t run = incl - my_sum;  // within the tile: waves before this one
  uint64_t sub_sum = 0;
#pragma unroll
  for (int i = 0; i < NWS; ++i) {
    const uint64_t x = s_wave[sub * NWS + i];
    run += (static_cast<uint32_t>(i) < w % NWS) ? x : 0ull;
    sub_sum += x;
  }
  if (tid == 0) s_sub[sub] = sub_sum;
  __syncthreads();
  uint64_t before = 0, blk_sum = 0;  // tiles of this block before this one; all of them
#pragma unroll
  for (int i = 0; i < kFbSub; ++i) {
    before += (static_cast<uint32_t>(i) < sub) ? s_sub[i] : 0ull;
    blk_sum += s_sub[i];
  }

  // Block prefix by decoupled look-back (wave 0).
  if (threadIdx.x < kWave) {
    if (blk == 0) {
      if (threadIdx.x == 0) {
        status_store64(status, kFlagIncl64 | blk_sum);
        s_prefix = 0;
      }
    } else {
      if (threadIdx.x == 0) status_store64(status + blk, kFlagAgg64 | blk_sum);
      const uint64_t excl = wave_lookback64(status, blk);
      if (threadIdx.x == 0) {
        status_store64(status + blk, kFlagIncl64 | (excl + blk_sum));
        s_prefix = excl;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    if (c0 + j < nat) {
      const uint32_t c = s_cnt[cnt_slot(c0 + j)];
      s_cnt[cnt_slot(c0 + j)] = static_cast<uint32_t>(run);  // tile-local prefix
      run += c;
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix + before;
  int32_t* __restrict__ lct = left_count + a0;
  for (uint32_t i = tid; i < nat; i += NT)
    lct[i] = static_cast<int32_t>(prefix + s_cnt[cnt_slot(i)]);
  if (tile == ntiles - 1 && tid == 0) *total = prefix + sub_sum;
}

// ---------------------------------------------------------------------------
// Expansion: merge path of output positions g (A = 0, 1, 2, ...) against row
// starts (B = left_count, non-decreasing).  A row start s comes before
// output g when s <= g, so output g belongs to the last row started.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t merge_path_rows(const int32_t* starts, uint64_t nrows,
                                                    uint64_t total, uint64_t diag) {
  uint64_t lo = diag > nrows ? diag - nrows : 0;
  uint64_t hi = diag < total ? diag : total;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    uint64_t s = static_cast<uint32_t>(starts[diag - 1 - mid]);
    if (mid < s) lo = mid + 1;  // output mid precedes the row start
    else hi = mid;
  }
  return lo;  // number of outputs among the first diag merged items
}

__global__ __launch_bounds__(256) void k_expand_partition(const int32_t* __restrict__ starts,
                                                          uint64_t nrows,
                                                          const uint64_t* __restrict__ total_p,
                                                          uint64_t total_cap, uint64_t ntiles_cap,
                                                          uint32_t nv,
                                                          uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles_cap) return;
  // Never expand past the caller's capacity, whatever the device says.
  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  uint64_t diag = t * nv;
  if (diag > total + nrows) diag = total + nrows;
  split[t] = merge_path_rows(starts, nrows, total, diag);
}

// Output g of the tile gets (row, right index).  Mode 0 writes the index
// columns; mode 1 gathers the output tuple directly.
struct GatherSpec {
  void* out;
  int out_bytes;
  const void* lkey;
  int lkey_bytes;
  const void* lval;
  int lval_bytes;
  const void* rval;
  int rval_bytes;
  uint32_t key_shift;
  uint32_t lval_shift;
};

// A source element held in S: uint64_t when every source is at most 8
// bytes wide, u128 otherwise (fewer registers for the common case).
template <class S>
__device__ __forceinline__ S load_src(const void* base, uint64_t i, int w);
template <>
__device__ __forceinline__ uint64_t load_src<uint64_t>(const void* base, uint64_t i, int w) {
  switch (w) {
    case 1: return static_cast<const uint8_t*>(base)[i];
    case 2: return static_cast<const uint16_t*>(base)[i];
    case 4: return static_cast<const uint32_t*>(base)[i];
    default: return static_cast<const uint64_t*>(base)[i];
  }
}
template <>
__device__ __forceinline__ u128 load_src<u128>(const void* base, uint64_t i, int w) {
  return load_w(base, i, w);
}

// The output tuple of one joined row from its sources, already loaded:
// (key << key_shift) | (left value << lval_shift) | right value.
template <class S>
__device__ __forceinline__ u128 compose_v(const GatherSpec& g, S k, S lv, S rv) {
  u128 o = shl128(to128(k), g.key_shift);
  if (g.lval) o = or128(o, shl128(to128(lv), g.lval_shift));
  if (g.rval) o = or128(o, to128(rv));
  return mask128(o, 8u * static_cast<uint32_t>(g.out_bytes));
}

// Mode 2: the output row is projected straight into up to two columns.
// Field f takes bits [in_shift, in_shift + bits) of source src (0 left key,
// 1 left value, 2 right value) to out_shift of output column col.
struct EmitField {
  uint8_t col, src, in_shift, bits, out_shift;
};
struct EmitSpec {
  const void* src[3];
  int src_bytes[3];
  void* out[2];
  int out_bytes[2];
  int nf;
  uint32_t used;  // bit s: some field reads source s
  EmitField f[8];
};

// The projected output row of sources v0 (left key), v1 (left value) and
// v2 (right value), already loaded.
template <class S>
__device__ __forceinline__ void emit_v(const EmitSpec& e, uint64_t g, S s0, S s1, S s2) {
  const u128 zero{0, 0};
  const u128 v0 = to128(s0), v1 = to128(s1), v2 = to128(s2);
  u128 o0 = zero, o1 = zero;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k < e.nf) {
      const EmitField f = e.f[k];
      const u128 v = f.src == 0 ? v0 : f.src == 1 ? v1 : v2;
      const u128 x = shl128(mask128(shr128(v, f.in_shift), f.bits), f.out_shift);
      if (f.col == 0) o0 = or128(o0, x);
      else o1 = or128(o1, x);
    }
  }
  store_w(e.out[0], g, e.out_bytes[0], o0);
  if (e.out[1]) store_w(e.out[1], g, e.out_bytes[1], o1);
}

#ifndef RFX_EX_NV
#define RFX_EX_NV 4096
#endif
// expand tile: merged (output + row) items per block.
constexpr int kExNV = RFX_EX_NV;
constexpr int kExNT = 512;

// One tile of the merge path of outputs against row starts: outputs
// [g0, g1) and rows [r0, r1), plus row r0 - 1 (already open at the tile's
// start) in LDS slot 0.  Output g belongs to slot #{tile rows j : start_j
// <= g}, found by fixed-step binary searches of the slots' starts that a
// thread runs in lockstep for all its outputs; its right index is the
// row's lower bound plus g - start.  Outputs are thread-strided, so every
// store is coalesced.  The gathering modes load the sources of kExBatch
// outputs before storing any of them: a store may alias a later load as
// far as the compiler knows, so loads interleaved with stores would wait
// for each other one memory latency at a time.
constexpr int kExBatch = 4;
template <int MODE, int NT, int NV, class S>
__global__ __launch_bounds__(NT) void k_expand(
    const int32_t* __restrict__ lower_bound, const int32_t* __restrict__ starts, uint64_t nrows,
    const uint64_t* __restrict__ total_p, uint64_t total_cap, const uint64_t* __restrict__ split,
    int32_t* __restrict__ left_idx, int32_t* __restrict__ right_idx, GatherSpec gs, EmitSpec es) {
  constexpr int VT = NV / NT;
  __shared__ uint32_t s_start[NV + 1];
  __shared__ uint32_t s_lb[NV + 1];

  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  const uint32_t tile = blockIdx.x;
  const uint64_t d0 = static_cast<uint64_t>(tile) * NV;
  if (d0 >= total + nrows) return;
  const uint64_t d1 = (d0 + NV < total + nrows) ? d0 + NV : total + nrows;
  const uint64_t g0 = split[tile], g1 = split[tile + 1];
  const uint64_t r0 = d0 - g0, r1 = d1 - g1;
  const uint32_t ng = static_cast<uint32_t>(g1 - g0), nr = static_cast<uint32_t>(r1 - r0);
  const uint32_t tid = threadIdx.x;

  for (uint32_t i = tid; i < nr + 1; i += NT) {
    const int64_t r = static_cast<int64_t>(r0) - 1 + i;
    s_start[i] = r >= 0 ? static_cast<uint32_t>(starts[r]) : 0u;
    s_lb[i] = r >= 0 ? static_cast<uint32_t>(lower_bound[r]) : 0u;
  }
  __syncthreads();

  // Slot of output g: how many of the tile's rows (starts s_start[1..nr],
  // ascending) start at or before g.  Shar's fixed-step search as in
  // find_bounds: one probe at the largest power of two P <= nr picks the
  // window, then steps P/2 .. 1 at compile-time offsets from a byte
  // position, and a last probe.
  const uint32_t JA = (ng + NT - 1) / NT;
  const uint32_t top = nr ? 1u << (31 - __builtin_clz(nr)) : 0u;
  constexpr int kLogNV = __builtin_ctz(NV);
  const char* sab = reinterpret_cast<const char*>(s_start + 1);
  uint32_t slot[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) slot[j] = 0;
  if (top) {
    const uint32_t pivot = s_start[top];
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t g = static_cast<uint32_t>(g0) + tid + j * NT;
      slot[j] = pivot <= g ? (nr - top) * 4u : 0u;
    }
#pragma unroll
    for (int sh = kLogNV - 1; sh >= 0; --sh) {
      const uint32_t step = 1u << sh;
      if (step < top) {
#pragma unroll
        for (int j = 0; j < VT; ++j) {
          if (static_cast<uint32_t>(j) < JA) {
            const uint32_t g = static_cast<uint32_t>(g0) + tid + j * NT;
            const uint32_t x = *reinterpret_cast<const uint32_t*>(sab + slot[j] + (step - 1) * 4);
            slot[j] = x <= g ? slot[j] + step * 4 : slot[j];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t g = static_cast<uint32_t>(g0) + tid + j * NT;
      const uint32_t x = *reinterpret_cast<const uint32_t*>(sab + slot[j]);
      slot[j] = (x <= g ? slot[j] + 4 : slot[j]) / 4;
    }
  }
  static_assert(VT % kExBatch == 0, "outputs are gathered in batches");
#pragma unroll
  for (int jb = 0; jb < VT; jb += kExBatch) {
    if (static_cast<uint32_t>(jb) >= JA) break;
    uint32_t gq[kExBatch], rowq[kExBatch], rixq[kExBatch];
    bool okq[kExBatch];
#pragma unroll
    for (int q = 0; q < kExBatch; ++q) {
      const uint32_t i = tid + (jb + q) * NT;
      okq[q] = static_cast<uint32_t>(jb + q) < JA && i < ng;
      const uint32_t sl = slot[jb + q];
      gq[q] = static_cast<uint32_t>(g0) + i;
      rowq[q] = static_cast<uint32_t>(r0 - 1 + sl);
      rixq[q] = okq[q] ? s_lb[sl] + (gq[q] - s_start[sl]) : 0u;
    }
    if constexpr (MODE == 0) {
#pragma unroll
      for (int q = 0; q < kExBatch; ++q)
        if (okq[q]) {
          left_idx[gq[q]] = static_cast<int32_t>(rowq[q]);
          right_idx[gq[q]] = static_cast<int32_t>(rixq[q]);
        }
    } else {
      S x0[kExBatch], x1[kExBatch], x2[kExBatch];
      const void* p0 = MODE == 1 ? gs.lkey : es.src[0];
      const void* p1 = MODE == 1 ? gs.lval : es.src[1];
      const void* p2 = MODE == 1 ? gs.rval : es.src[2];
      const int w0 = MODE == 1 ? gs.lkey_bytes : es.src_bytes[0];
      const int w1 = MODE == 1 ? gs.lval_bytes : es.src_bytes[1];
      const int w2 = MODE == 1 ? gs.rval_bytes : es.src_bytes[2];
      const bool u0 = MODE == 1 ? true : (es.used & 1) != 0;
      const bool u1 = MODE == 1 ? gs.lval != nullptr : (es.used & 2) != 0;
      const bool u2 = MODE == 1 ? gs.rval != nullptr : (es.used & 4) != 0;
#pragma unroll
      for (int q = 0; q < kExBatch; ++q) {
        x0[q] = x1[q] = x2[q] = S{};
        if (okq[q]) {
          if (u0) x0[q] = load_src<S>(p0, rowq[q], w0);
          if (u1) x1[q] = load_src<S>(p1, rowq[q], w1);
          if (u2) x2[q] = load_src<S>(p2, rixq[q], w2);
        }
      }
#pragma unroll
      for (int q = 0; q < kExBatch; ++q) {
        if (!okq[q]) continue;
        if constexpr (MODE == 2) emit_v(es, gq[q], x0[q], x1[q], x2[q]);
        else store_w(gs.out, gq[q], gs.out_bytes, compose_v(gs, x0[q], x1[q], x2[q]));
      }
    }
  }
}

// Grid-stride gather, kExBatch outputs per thread at a time: their index
// pairs and sources are all loaded before any output is stored (a store
// may alias a later load as far as the compiler knows).
template <class S>
__global__ __launch_bounds__(256) void k_gather(GatherSpec gs, const int32_t* __restrict__ li,
                                                const int32_t* __restrict__ ri,
                                                const uint64_t* __restrict__ total_p,
                                                uint64_t total_cap) {
  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t g0 = blockIdx.x * 256ull + threadIdx.x; g0 < total; g0 += kExBatch * stride) {
    uint32_t lq[kExBatch], rq[kExBatch];
#pragma unroll
    for (int q = 0; q < kExBatch; ++q) {
      const uint64_t g = g0 + q * stride;
      lq[q] = g < total ? static_cast<uint32_t>(li[g]) : 0u;
      rq[q] = g < total ? static_cast<uint32_t>(ri[g]) : 0u;
    }
    S x0[kExBatch], x1[kExBatch], x2[kExBatch];
#pragma unroll
    for (int q = 0; q < kExBatch; ++q) {
      x0[q] = x1[q] = x2[q] = S{};
      if (g0 + q * stride < total) {
        x0[q] = load_src<S>(gs.lkey, lq[q], gs.lkey_bytes);
        if (gs.lval) x1[q] = load_src<S>(gs.lval, lq[q], gs.lval_bytes);
        if (gs.rval) x2[q] = load_src<S>(gs.rval, rq[q], gs.rval_bytes);
      }
    }
#pragma unroll
    for (int q = 0; q < kExBatch; ++q) {
      const uint64_t g = g0 + q * stride;
      if (g < total) store_w(gs.out, g, gs.out_bytes, compose_v(gs, x0[q], x1[q], x2[q]));
    }
  }
}

// Scratch of one find_bounds launch over `ntiles` tiles: the merge-path
// splits, the look-back status words (one per block) and the block ticket,
// each aligned to 256 bytes.
constexpr uint64_t fb_blocks(uint64_t ntiles) { return (ntiles + kFbSub - 1) / kFbSub; }
constexpr uint64_t fb_scratch_bytes(uint64_t ntiles) {
  return align_up(align_up(8 * (ntiles + 1), 256) + 8 * fb_blocks(ntiles), 256) + 4;
}

// `scratch` is the device memory the launch uses for fb_scratch_bytes():
// nullptr takes it from the context's workspace (which may grow), anything
// else is the caller's and is not profiled (profiling creates events).
template <class K, int NV>
int find_bounds_nv(rfx_ctx* ctx, int32_t* lb, int32_t* lc, uint64_t* total, const void* left,
                   uint64_t nl, const void* right, uint64_t nr, void* scratch) {
  if (nl == 0 || nr == 0) {
    // No matches: every count is zero and every lower bound is 0 (or nr
    // when the right side is empty... which is also 0).
    if (nl) {
      RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(lb), nl, 0));
      RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(lc), nl, 0));
    }
    RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(total), 2, 0));
    return RFX_OK;
  }
  const uint64_t ntiles = (nl + nr + NV - 1) / NV;
  const bool own = scratch == nullptr;
  if (own) {
    RFX_TRY(ensure_workspace(ctx, fb_scratch_bytes(ntiles)));
    scratch = ctx->ws;
  }
  Carve w(scratch);
  uint64_t* split = w.take<uint64_t>(ntiles + 1);
  const uint64_t nblk = fb_blocks(ntiles);
  uint64_t* status = w.take<uint64_t>(nblk);
  uint32_t* ticket = w.take<uint32_t>(1);
  // status and ticket are adjacent after alignment; zero both.
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  k_fb_partition<K><<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, ntiles, NV, split);
  RFX_LAUNCHED();
  const bool prof = own && ctx->profiling;
  if (prof) prof_begin(ctx, "join.find_bounds");
  k_find_bounds<K, kFbNT, NV><<<static_cast<uint32_t>(nblk), kFbNT * kFbSub, 0, ctx->stream>>>(
      static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, split,
      static_cast<uint32_t>(ntiles), lb, lc, total, status, ticket);
  if (prof) prof_end(ctx);
  RFX_LAUNCHED();
  return RFX_OK;
}

// Tile of 2048 or 4096 merged keys (tunable for keys of up to 8 bytes;
// 16-byte keys always take 2048).
template <class K>
int find_bounds_impl(rfx_ctx* ctx, int32_t* lb, int32_t* lc, uint64_t* total, const void* left,
                     uint64_t nl, const void* right, uint64_t nr, void* scratch = nullptr) {
  if constexpr (sizeof(K) <= 8) {
    const int nv = scratch ? kFbNV : ctx->tune_fb_tile ? ctx->tune_fb_tile : kFbNV;
    if (nv == 2048) return find_bounds_nv<K, 2048>(ctx, lb, lc, total, left, nl, right, nr, scratch);
    if (nv == 4096) return find_bounds_nv<K, 4096>(ctx, lb, lc, total, left, nl, right, nr, scratch);
    return find_bounds_nv<K, kFbNV>(ctx, lb, lc, total, left, nl, right, nr, scratch);
  } else {
    return find_bounds_nv<K, 2048>(ctx, lb, lc, total, left, nl, right, nr, scratch);
  }
}

// ---------------------------------------------------------------------------
// find_bounds without allocation: the scratch of one launch lives in this
// file-scope device array, one copy per device, large enough for
// kFbStaticTiles tiles (2^28 merged keys of up to 8 bytes at the default
// tile, 2^27 of 16 bytes).  Host code never uses the array's name as an
// address: each device's copy is looked up once with hipGetSymbolAddress
// and that pointer is what the kernels get.
// ---------------------------------------------------------------------------
constexpr uint64_t kFbStaticTiles = 65536;
constexpr uint64_t kFbStaticBytes = fb_scratch_bytes(kFbStaticTiles);
__device__ __align__(256) unsigned char scratch[kFbStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch` (looked up on the first call per device,
// with that device current).
int static_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(RFX_EINVAL, "find_bounds scratch: device %d", device);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    RFX_HIP(hipGetDevice(&cur));
    if (cur != device) RFX_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) RFX_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(RFX_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                  hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return RFX_OK;
}

template <int NV>
int expand_nv(rfx_ctx* ctx, int mode, int32_t* li, int32_t* ri, const GatherSpec& gs,
              const uint64_t* total, uint64_t total_cap, const int32_t* lb, const int32_t* lc,
              uint64_t nl, const EmitSpec& es) {
  if (nl == 0 || total_cap == 0) return RFX_OK;
  const uint64_t ntiles = (total_cap + nl + NV - 1) / NV;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles + 1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* split = w.take<uint64_t>(ntiles + 1);
  k_expand_partition<<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      lc, nl, total, total_cap, ntiles, NV, split);
  RFX_LAUNCHED();
  RFX_PROF(ctx, mode == 0 ? "join.expand" : mode == 1 ? "join.expand_gather" : "join.expand_project");
  // Sources wider than 8 bytes take the 128-bit register form.
  const bool wide = mode == 1 ? (gs.lkey_bytes == 16 || (gs.lval && gs.lval_bytes == 16) ||
                                 (gs.rval && gs.rval_bytes == 16))
                              : (es.src_bytes[0] == 16 || es.src_bytes[1] == 16 ||
                                 es.src_bytes[2] == 16);
  const uint32_t grid = static_cast<uint32_t>(ntiles);
  if (mode == 0)
    k_expand<0, kExNT, NV, uint64_t><<<grid, kExNT, 0, ctx->stream>>>(lb, lc, nl, total, total_cap,
                                                                      split, li, ri, gs, es);
  else if (mode == 1 && !wide)
    k_expand<1, kExNT, NV, uint64_t><<<grid, kExNT, 0, ctx->stream>>>(lb, lc, nl, total, total_cap,
                                                                      split, li, ri, gs, es);
  else if (mode == 1)
    k_expand<1, kExNT, NV, u128><<<grid, kExNT, 0, ctx->stream>>>(lb, lc, nl, total, total_cap,
                                                                  split, li, ri, gs, es);
  else if (!wide)
    k_expand<2, kExNT, NV, uint64_t><<<grid, kExNT, 0, ctx->stream>>>(lb, lc, nl, total, total_cap,
                                                                      split, li, ri, gs, es);
  else
    k_expand<2, kExNT, NV, u128><<<grid, kExNT, 0, ctx->stream>>>(lb, lc, nl, total, total_cap,
                                                                  split, li, ri, gs, es);
  RFX_LAUNCHED();
  return RFX_OK;
}

int expand_impl(rfx_ctx* ctx, int mode, int32_t* li, int32_t* ri, const GatherSpec& gs,
                const uint64_t* total, uint64_t total_cap, const int32_t* lb, const int32_t* lc,
                uint64_t nl, const EmitSpec& es = EmitSpec{}) {
  const int nv = ctx->tune_ex_tile ? ctx->tune_ex_tile : kExNV;
  if (nv == 2048) return expand_nv<2048>(ctx, mode, li, ri, gs, total, total_cap, lb, lc, nl, es);
  if (nv == 4096) return expand_nv<4096>(ctx, mode, li, ri, gs, total, total_cap, lb, lc, nl, es);
  return expand_nv<kExNV>(ctx, mode, li, ri, gs, total, total_cap, lb, lc, nl, es);
}

int check_gather(const GatherSpec& g) {
  if (!valid_width(g.out_bytes) || !valid_width(g.lkey_bytes))
    return fail(RFX_EINVAL, "join gather: out/key width %d/%d", g.out_bytes, g.lkey_bytes);
  if (g.lval && !valid_width(g.lval_bytes))
    return fail(RFX_EINVAL, "join gather: left value width %d", g.lval_bytes);
  if (g.rval && !valid_width(g.rval_bytes))
    return fail(RFX_EINVAL, "join gather: right value width %d", g.rval_bytes);
  if (g.key_shift >= 128 || g.lval_shift >= 128)
    return fail(RFX_EINVAL, "join gather: shift out of range");
  if (!g.out || !g.lkey) return fail(RFX_EINVAL, "join gather: null pointer");
  return RFX_OK;
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_join_find_bounds(rfx_ctx* ctx, int32_t* lower_bound, int32_t* left_count,
                                    uint64_t* total, const void* left, uint64_t nl,
                                    const void* right, uint64_t nr, int key_bytes) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_find_bounds: null argument");
  if (nl && (!lower_bound || !left_count || !left))
    return fail(RFX_EINVAL, "rfx_join_find_bounds: null column");
  if (nr && !right) return fail(RFX_EINVAL, "rfx_join_find_bounds: null right keys");
  if (nl + nr >= (1ull << 31))
    return fail(RFX_ERANGE, "rfx_join_find_bounds: %llu rows exceed the int32 index range",
                static_cast<unsigned long long>(nl + nr));
  switch (key_bytes) {
    case 1: return find_bounds_impl<uint8_t>(ctx, lower_bound, left_count, total, left, nl, right, nr);
    case 2: return find_bounds_impl<uint16_t>(ctx, lower_bound, left_count, total, left, nl, right, nr);
    case 4: return find_bounds_impl<uint32_t>(ctx, lower_bound, left_count, total, left, nl, right, nr);
    case 8: return find_bounds_impl<uint64_t>(ctx, lower_bound, left_count, total, left, nl, right, nr);
    case 16: return find_bounds_impl<u128>(ctx, lower_bound, left_count, total, left, nl, right, nr);
  }
  return fail(RFX_EINVAL, "rfx_join_find_bounds: key width %d", key_bytes);
}

extern "C" int rfx_join_find_bounds_noalloc(rfx_ctx* ctx, int32_t* lower_bound,
                                            int32_t* left_count, uint64_t* total,
                                            const void* left, uint64_t nl, const void* right,
                                            uint64_t nr, int key_bytes) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_find_bounds_noalloc: null argument");
  if (nl && (!lower_bound || !left_count || !left))
    return fail(RFX_EINVAL, "rfx_join_find_bounds_noalloc: null column");
  if (nr && !right) return fail(RFX_EINVAL, "rfx_join_find_bounds_noalloc: null right keys");
  if (!valid_width(key_bytes))
    return fail(RFX_EINVAL, "rfx_join_find_bounds_noalloc: key width %d", key_bytes);
  const uint64_t tile = key_bytes == 16 ? 2048 : kFbNV;
  if (nl + nr > kFbStaticTiles * tile)
    return fail(RFX_ERANGE,
                "rfx_join_find_bounds_noalloc: %llu keys exceed its static scratch "
                "(at most %llu merged keys of %d bytes)",
                static_cast<unsigned long long>(nl + nr),
                static_cast<unsigned long long>(kFbStaticTiles * tile), key_bytes);
  void* s = nullptr;
  RFX_TRY(static_scratch(ctx->device, &s));
  switch (key_bytes) {
    case 1: return find_bounds_impl<uint8_t>(ctx, lower_bound, left_count, total, left, nl, right, nr, s);
    case 2: return find_bounds_impl<uint16_t>(ctx, lower_bound, left_count, total, left, nl, right, nr, s);
    case 4: return find_bounds_impl<uint32_t>(ctx, lower_bound, left_count, total, left, nl, right, nr, s);
    case 8: return find_bounds_impl<uint64_t>(ctx, lower_bound, left_count, total, left, nl, right, nr, s);
    default: return find_bounds_impl<u128>(ctx, lower_bound, left_count, total, left, nl, right, nr, s);
  }
}

extern "C" int rfx_join_main(rfx_ctx* ctx, int32_t* left_idx, int32_t* right_idx,
                             const uint64_t* total, uint64_t total_cap,
                             const int32_t* lower_bound, const int32_t* left_count, uint64_t nl) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_main: null argument");
  if (total_cap && (!left_idx || !right_idx || !lower_bound || !left_count))
    return fail(RFX_EINVAL, "rfx_join_main: null column");
  if (total_cap + nl >= (1ull << 31))
    return fail(RFX_ERANGE, "rfx_join_main: output exceeds the int32 index range");
  GatherSpec gs{};
  return expand_impl(ctx, 0, left_idx, right_idx, gs, total, total_cap, lower_bound, left_count, nl);
}

extern "C" int rfx_join_gather(rfx_ctx* ctx, void* out, int out_bytes, const void* lkey,
                               int lkey_bytes, const void* lval, int lval_bytes,
                               const void* rval, int rval_bytes, const int32_t* left_idx,
                               const int32_t* right_idx, const uint64_t* total,
                               uint64_t total_cap, uint32_t key_shift, uint32_t lval_shift) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_gather: null argument");
  if (total_cap == 0) return RFX_OK;
  GatherSpec gs{out, out_bytes, lkey, lkey_bytes, lval, lval_bytes, rval, rval_bytes,
                key_shift, lval_shift};
  RFX_TRY(check_gather(gs));
  if (!left_idx || !right_idx) return fail(RFX_EINVAL, "rfx_join_gather: null index column");
  uint64_t blocks = (total_cap + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(ctx->device_cus) * 16;
  if (blocks > cap) blocks = cap;
  RFX_PROF(ctx, "join.gather");
  const bool wide = gs.lkey_bytes == 16 || (gs.lval && gs.lval_bytes == 16) ||
                    (gs.rval && gs.rval_bytes == 16);
  if (wide)
    k_gather<u128><<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(gs, left_idx, right_idx,
                                                                           total, total_cap);
  else
    k_gather<uint64_t><<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(
        gs, left_idx, right_idx, total, total_cap);
  RFX_LAUNCHED();
  return RFX_OK;
}

extern "C" int rfx_join_expand_gather(rfx_ctx* ctx, void* out, int out_bytes, const void* lkey,
                                      int lkey_bytes, const void* lval, int lval_bytes,
                                      const void* rval, int rval_bytes,
                                      const int32_t* lower_bound, const int32_t* left_count,
                                      uint64_t nl, const uint64_t* total, uint64_t total_cap,
                                      uint32_t key_shift, uint32_t lval_shift) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_expand_gather: null argument");
  if (total_cap == 0 || nl == 0) return RFX_OK;
  GatherSpec gs{out, out_bytes, lkey, lkey_bytes, lval, lval_bytes, rval, rval_bytes,
                key_shift, lval_shift};
  RFX_TRY(check_gather(gs));
  if (!lower_bound || !left_count) return fail(RFX_EINVAL, "rfx_join_expand_gather: null column");
  if (total_cap + nl >= (1ull << 31))
    return fail(RFX_ERANGE, "rfx_join_expand_gather: output exceeds the int32 index range");
  return expand_impl(ctx, 1, nullptr, nullptr, gs, total, total_cap, lower_bound, left_count, nl);
}

extern "C" int rfx_join_expand_project(rfx_ctx* ctx, const int32_t* lower_bound,
                                       const int32_t* left_count, uint64_t nl,
                                       const uint64_t* total, uint64_t total_cap,
                                       const void* lkey, int lkey_bytes, const void* lval,
                                       int lval_bytes, const void* rval, int rval_bytes,
                                       void* out0, int out0_bytes, void* out1, int out1_bytes,
                                       int nfields, const uint32_t* fields) {
  if (!ctx || !total) return fail(RFX_EINVAL, "rfx_join_expand_project: null argument");
  if (total_cap == 0 || nl == 0) return RFX_OK;
  if (!lower_bound || !left_count || !out0 || !fields)
    return fail(RFX_EINVAL, "rfx_join_expand_project: null column");
  if (nfields < 1 || nfields > 8) return fail(RFX_EINVAL, "rfx_join_expand_project: %d fields", nfields);
  if (!valid_width(out0_bytes) || (out1 && !valid_width(out1_bytes)))
    return fail(RFX_EINVAL, "rfx_join_expand_project: output widths");
  if (total_cap + nl >= (1ull << 31))
    return fail(RFX_ERANGE, "rfx_join_expand_project: output exceeds the int32 index range");
  EmitSpec es{};
  es.src[0] = lkey; es.src[1] = lval; es.src[2] = rval;
  es.src_bytes[0] = lkey_bytes; es.src_bytes[1] = lval_bytes; es.src_bytes[2] = rval_bytes;
  es.out[0] = out0; es.out[1] = out1;
  es.out_bytes[0] = out0_bytes; es.out_bytes[1] = out1 ? out1_bytes : 0;
  es.nf = nfields;
  es.used = 0;
  for (int k = 0; k < nfields; ++k) {
    const uint32_t* f = fields + 5 * k;
    const uint32_t col = f[0], src = f[1], in_shift = f[2], bits = f[3], out_shift = f[4];
    if (col > 1 || (col == 1 && !out1) || src > 2 || !es.src[src] || !valid_width(es.src_bytes[src]))
      return fail(RFX_EINVAL, "rfx_join_expand_project: field %d source/column", k);
    if (bits == 0 || bits > 128 || in_shift + bits > 8u * es.src_bytes[src] ||
        out_shift + bits > 8u * es.out_bytes[col])
      return fail(RFX_EINVAL, "rfx_join_expand_project: field %d bits out of range", k);
    es.used |= 1u << src;
    es.f[k] = EmitField{static_cast<uint8_t>(col), static_cast<uint8_t>(src),
                        static_cast<uint8_t>(in_shift), static_cast<uint8_t>(bits),
                        static_cast<uint8_t>(out_shift)};
  }
  GatherSpec gs{};
  return expand_impl(ctx, 2, nullptr, nullptr, gs, total, total_cap, lower_bound, left_count, nl,
                     es);
}
