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
constexpr int kFbNT = 512;

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

// One tile of the merge path: its left keys A[a0, a1) and right keys
// B[b0, b1).  With A winning ties, every B[j] before b0 is < A[a0] and every
// B[j] from b1 on is >= A[a1 - 1], so each left key's lower bound lies in
// [b0, b1] and is found by a binary search of the tile's right keys in LDS
// (one left key per lane: neighbouring lanes probe the same words), and so
// is its upper bound; the equal-key count is their difference, extended by
// a gallop through global memory only for a run that crosses the tile's
// end.
template <class K, int NT, int NV>
__global__ __launch_bounds__(NT) void k_find_bounds(
    const K* __restrict__ a, uint64_t na, const K* __restrict__ b, uint64_t nb,
    const uint64_t* __restrict__ split, uint32_t ntiles, int32_t* __restrict__ lower_bound,
    int32_t* __restrict__ left_count, uint64_t* __restrict__ total,
    uint64_t* __restrict__ status, uint32_t* __restrict__ ticket) {
  constexpr int VT = NV / NT;
  __shared__ K keys[NV];
  __shared__ uint32_t s_cnt[NV];
  __shared__ uint64_t scan_tmp[NT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;

  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t d0 = static_cast<uint64_t>(tile) * NV;
  const uint64_t d1 = (d0 + NV < na + nb) ? d0 + NV : na + nb;
  const uint64_t a0 = split[tile], a1 = split[tile + 1];
  const uint64_t b0 = d0 - a0, b1 = d1 - a1;
  const uint32_t nat = static_cast<uint32_t>(a1 - a0), nbt = static_cast<uint32_t>(b1 - b0);

  // Tile's A keys then B keys, contiguous in LDS; all loads in flight
  // before the first LDS write.
  if constexpr (sizeof(K) <= 8) {
    K v[VT];
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t i = tid + j * NT;
      if (i < nat) v[j] = a[a0 + i];
      else if (i < nat + nbt) v[j] = b[b0 + (i - nat)];
    }
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint32_t i = tid + j * NT;
      if (i < nat + nbt) keys[i] = v[j];
    }
  } else {
    for (uint32_t i = tid; i < nat + nbt; i += NT)
      keys[i] = i < nat ? a[a0 + i] : b[b0 + (i - nat)];
  }
  __syncthreads();
  const K* sa = keys;
  const K* sb = keys + nat;

  // Lower and upper bounds of each left key in the tile's right keys by
  // fixed-step binary searches: all of a thread's searches advance in
  // lockstep, so their LDS reads overlap.  JA (block-uniform) bounds the
  // left keys per thread.
  const uint32_t JA = (nat + NT - 1) / NT;
  uint32_t top = 0;
  if (nbt) top = 1u << (31 - __builtin_clz(nbt));
  K kv[VT];
  uint32_t lo[VT], up[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    lo[j] = up[j] = 0;
    if (static_cast<uint32_t>(j) < JA) {
      const uint32_t i = tid + j * NT;
      kv[j] = sa[i < nat ? i : nat - 1];
    }
  }
  for (uint32_t step = top; step; step >>= 1) {
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      if (static_cast<uint32_t>(j) < JA) {
        if (lo[j] + step <= nbt && lt(sb[lo[j] + step - 1], kv[j])) lo[j] += step;
        if (up[j] + step <= nbt && !lt(kv[j], sb[up[j] + step - 1])) up[j] += step;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint32_t i = tid + j * NT;
    if (static_cast<uint32_t>(j) < JA && i < nat) {
      const K key = kv[j];
      uint64_t cnt = up[j] - lo[j];
      // A run of equal right keys that reaches the tile's end continues
      // in global memory from b1.
      if (up[j] == nbt && b1 < nb && eq(b[b1], key)) {
        uint64_t l = b1 + 1, step = 1, h;
        while (true) {  // invariant: b[l - 1] == key
          h = l + step - 1;
          if (h >= nb) { h = nb; break; }
          if (!eq(b[h], key)) break;
          l = h + 1;
          step <<= 1;
        }
        while (l < h) {  // first index in [l, h) whose key differs
          const uint64_t mid = (l + h) >> 1;
          if (eq(b[mid], key)) l = mid + 1;
          else h = mid;
        }
        cnt += l - b1;
      }
      lower_bound[a0 + i] = static_cast<int32_t>(b0 + lo[j]);
      s_cnt[i] = static_cast<uint32_t>(cnt);
    }
  }
  __syncthreads();

  // Exclusive scan of the counts in left-key order: thread t owns
  // elements [t*VT, t*VT + VT).
  const uint32_t c0 = tid * VT;
  uint32_t cv[VT];
  uint64_t my_sum = 0;
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    cv[j] = (c0 + j < nat) ? s_cnt[c0 + j] : 0u;
    my_sum += cv[j];
  }
  uint64_t tile_sum;
  uint64_t run = block_excl_scan<NT>(my_sum, scan_tmp, &tile_sum);

  // Tile prefix by decoupled look-back (wave 0).
  if (tid < kWave) {
    if (tile == 0) {
      if (tid == 0) status_store64(status, kFlagIncl64 | tile_sum);
      if (tid == 0) s_prefix = 0;
    } else {
      if (tid == 0) status_store64(status + tile, kFlagAgg64 | tile_sum);
      uint64_t excl = wave_lookback64(status, tile);
      if (tid == 0) {
        status_store64(status + tile, kFlagIncl64 | (excl + tile_sum));
        s_prefix = excl;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    if (c0 + j < nat) s_cnt[c0 + j] = static_cast<uint32_t>(run);  // tile-local prefix
    run += cv[j];
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  for (uint32_t i = tid; i < nat; i += NT)
    left_count[a0 + i] = static_cast<int32_t>(prefix + s_cnt[i]);
  if (tile == ntiles - 1 && tid == 0) *total = prefix + tile_sum;
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

__device__ __forceinline__ u128 compose(const GatherSpec& g, uint64_t li, uint64_t ri) {
  u128 o = shl128(load_w(g.lkey, li, g.lkey_bytes), g.key_shift);
  if (g.lval) o = or128(o, shl128(load_w(g.lval, li, g.lval_bytes), g.lval_shift));
  if (g.rval) o = or128(o, load_w(g.rval, ri, g.rval_bytes));
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

__device__ __forceinline__ void emit(const EmitSpec& e, uint64_t g, uint64_t li, uint64_t ri) {
  const u128 zero{0, 0};
  const u128 v0 = (e.used & 1) ? load_w(e.src[0], li, e.src_bytes[0]) : zero;
  const u128 v1 = (e.used & 2) ? load_w(e.src[1], li, e.src_bytes[1]) : zero;
  const u128 v2 = (e.used & 4) ? load_w(e.src[2], ri, e.src_bytes[2]) : zero;
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
// store is coalesced.
template <int MODE, int NT, int NV>
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

  const uint32_t JA = (ng + NT - 1) / NT;
  uint32_t top = 0;
  if (nr) top = 1u << (31 - __builtin_clz(nr));
  uint32_t slot[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) slot[j] = 0;
  for (uint32_t step = top; step; step >>= 1) {
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      if (static_cast<uint32_t>(j) < JA) {
        const uint32_t g = static_cast<uint32_t>(g0) + tid + j * NT;
        if (slot[j] + step <= nr && s_start[slot[j] + step] <= g) slot[j] += step;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint32_t i = tid + j * NT;
    if (static_cast<uint32_t>(j) < JA && i < ng) {
      const uint32_t g = static_cast<uint32_t>(g0) + i;
      const uint32_t sl = slot[j];
      const uint32_t row = static_cast<uint32_t>(r0 - 1 + sl);
      const uint32_t rix = s_lb[sl] + (g - s_start[sl]);
      if constexpr (MODE == 0) {
        left_idx[g] = static_cast<int32_t>(row);
        right_idx[g] = static_cast<int32_t>(rix);
      } else if constexpr (MODE == 2) {
        emit(es, g, row, rix);
      } else {
        store_w(gs.out, g, gs.out_bytes, compose(gs, row, rix));
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_gather(GatherSpec gs, const int32_t* __restrict__ li,
                                                const int32_t* __restrict__ ri,
                                                const uint64_t* __restrict__ total_p,
                                                uint64_t total_cap) {
  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < total; g += stride)
    store_w(gs.out, g, gs.out_bytes,
            compose(gs, static_cast<uint32_t>(li[g]), static_cast<uint32_t>(ri[g])));
}

template <class K, int NV>
int find_bounds_nv(rfx_ctx* ctx, int32_t* lb, int32_t* lc, uint64_t* total, const void* left,
                   uint64_t nl, const void* right, uint64_t nr) {
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
  Carve c(nullptr);
  c.take<uint64_t>(ntiles + 1);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* split = w.take<uint64_t>(ntiles + 1);
  uint64_t* status = w.take<uint64_t>(ntiles);
  uint32_t* ticket = w.take<uint32_t>(1);
  // status and ticket are adjacent after alignment; zero both.
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  k_fb_partition<K><<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, ntiles, NV, split);
  RFX_LAUNCHED();
  {
    RFX_PROF(ctx, "join.find_bounds");
    k_find_bounds<K, kFbNT, NV><<<static_cast<uint32_t>(ntiles), kFbNT, 0, ctx->stream>>>(
        static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, split,
        static_cast<uint32_t>(ntiles), lb, lc, total, status, ticket);
    RFX_LAUNCHED();
  }
  return RFX_OK;
}

// Tile of 2048 or 4096 merged keys (tunable for keys of up to 8 bytes;
// 16-byte keys always take 2048).
template <class K>
int find_bounds_impl(rfx_ctx* ctx, int32_t* lb, int32_t* lc, uint64_t* total, const void* left,
                     uint64_t nl, const void* right, uint64_t nr) {
  if constexpr (sizeof(K) <= 8) {
    const int nv = ctx->tune_fb_tile ? ctx->tune_fb_tile : kFbNV;
    if (nv == 2048) return find_bounds_nv<K, 2048>(ctx, lb, lc, total, left, nl, right, nr);
    if (nv == 4096) return find_bounds_nv<K, 4096>(ctx, lb, lc, total, left, nl, right, nr);
    return find_bounds_nv<K, kFbNV>(ctx, lb, lc, total, left, nl, right, nr);
  } else {
    return find_bounds_nv<K, 2048>(ctx, lb, lc, total, left, nl, right, nr);
  }
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
  if (mode == 0)
    k_expand<0, kExNT, NV><<<static_cast<uint32_t>(ntiles), kExNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
  else if (mode == 1)
    k_expand<1, kExNT, NV><<<static_cast<uint32_t>(ntiles), kExNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
  else
    k_expand<2, kExNT, NV><<<static_cast<uint32_t>(ntiles), kExNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
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
  k_gather<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(gs, left_idx, right_idx, total,
                                                                   total_cap);
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
