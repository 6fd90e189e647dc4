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

constexpr int kJoinNT = 256;
constexpr int kJoinVT = 8;
constexpr int kJoinNV = kJoinNT * kJoinVT;

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
template <class K>
__global__ __launch_bounds__(256) void k_fb_partition(const K* __restrict__ a, uint64_t na,
                                                      const K* __restrict__ b, uint64_t nb,
                                                      uint64_t ntiles,
                                                      uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles) return;
  uint64_t diag = t * kJoinNV;
  if (diag > na + nb) diag = na + nb;
  split[t] = merge_path_lower(a, na, b, nb, diag);
}

template <class K>
__global__ __launch_bounds__(kJoinNT) void k_find_bounds(
    const K* __restrict__ a, uint64_t na, const K* __restrict__ b, uint64_t nb,
    const uint64_t* __restrict__ split, uint32_t ntiles, int32_t* __restrict__ lower_bound,
    int32_t* __restrict__ left_count, uint64_t* __restrict__ total,
    uint64_t* __restrict__ status, uint32_t* __restrict__ ticket) {
  __shared__ K keys[kJoinNV];
  __shared__ int32_t s_lb[kJoinNV];
  __shared__ int32_t s_cnt[kJoinNV];
  __shared__ uint64_t scan_tmp[kJoinNT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;

  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t d0 = static_cast<uint64_t>(tile) * kJoinNV;
  const uint64_t d1 = (d0 + kJoinNV < na + nb) ? d0 + kJoinNV : na + nb;
  const uint64_t a0 = split[tile], a1 = split[tile + 1];
  const uint64_t b0 = d0 - a0, b1 = d1 - a1;
  const uint32_t nat = static_cast<uint32_t>(a1 - a0), nbt = static_cast<uint32_t>(b1 - b0);

  // Tile's A keys then B keys, contiguous in LDS.
  for (uint32_t i = tid; i < nat + nbt; i += kJoinNT)
    keys[i] = i < nat ? a[a0 + i] : b[b0 + (i - nat)];
  __syncthreads();
  const K* sa = keys;
  const K* sb = keys + nat;

  // Thread's slice of the tile's merge path.
  const uint32_t diag = tid * kJoinVT;
  uint32_t ai = 0, bj = 0;
  if (diag < nat + nbt) {
    ai = static_cast<uint32_t>(merge_path_lower(sa, nat, sb, nbt, diag));
    bj = diag - ai;
  }
  uint64_t my_sum = 0;
  const uint32_t a_begin = ai;
#pragma unroll
  for (int s = 0; s < kJoinVT; ++s) {
    if (diag + s >= nat + nbt) break;
    const bool take_a = ai < nat && (bj >= nbt || !lt(sb[bj], sa[ai]));
    if (take_a) {
      const K key = sa[ai];
      const uint64_t lb = b0 + bj;
      // Gallop for the first right key > key, starting at lb.
      uint64_t cnt = 0;
      auto rk = [&](uint64_t idx) -> K { return idx < b1 ? sb[idx - b0] : b[idx]; };
      if (lb < nb && eq(rk(lb), key)) {
        uint64_t lo = lb + 1, step = 1, hi = lb + 1;
        while (true) {  // invariant: rk(lo - 1) == key
          hi = lo + step - 1;
          if (hi >= nb) { hi = nb; break; }
          if (!eq(rk(hi), key)) break;
          lo = hi + 1;
          step <<= 1;
        }
        while (lo < hi) {  // first index in [lo, hi) whose key differs
          uint64_t mid = (lo + hi) >> 1;
          if (eq(rk(mid), key)) lo = mid + 1;
          else hi = mid;
        }
        cnt = lo - lb;
      }
      s_lb[ai] = static_cast<int32_t>(lb);
      s_cnt[ai] = static_cast<int32_t>(cnt);
      my_sum += cnt;
      ++ai;
    } else {
      ++bj;
    }
  }
  const uint32_t a_end = ai;

  uint64_t tile_sum;
  uint64_t my_prefix = block_excl_scan<kJoinNT>(my_sum, scan_tmp, &tile_sum);

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
  // Exclusive prefix of each A element: overwrite its count in LDS.
  uint64_t run = my_prefix;
  for (uint32_t i = a_begin; i < a_end; ++i) {
    const uint64_t c = static_cast<uint64_t>(static_cast<uint32_t>(s_cnt[i]));
    s_cnt[i] = static_cast<int32_t>(run);  // tile-local prefix for now
    run += c;
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  for (uint32_t i = tid; i < nat; i += kJoinNT) {
    lower_bound[a0 + i] = s_lb[i];
    left_count[a0 + i] = static_cast<int32_t>(prefix + static_cast<uint32_t>(s_cnt[i]));
  }
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
                                                          uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles_cap) return;
  // Never expand past the caller's capacity, whatever the device says.
  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  uint64_t diag = t * kJoinNV;
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
  EmitField f[8];
};

__device__ __forceinline__ void emit(const EmitSpec& e, uint64_t g, uint64_t li, uint64_t ri) {
  u128 o[2] = {{0, 0}, {0, 0}};
  for (int k = 0; k < e.nf; ++k) {
    const EmitField f = e.f[k];
    const u128 v = load_w(e.src[f.src], f.src == 2 ? ri : li, e.src_bytes[f.src]);
    o[f.col] = or128(o[f.col], shl128(mask128(shr128(v, f.in_shift), f.bits), f.out_shift));
  }
  store_w(e.out[0], g, e.out_bytes[0], o[0]);
  if (e.out[1]) store_w(e.out[1], g, e.out_bytes[1], o[1]);
}

template <int MODE>
__global__ __launch_bounds__(kJoinNT) void k_expand(
    const int32_t* __restrict__ lower_bound, const int32_t* __restrict__ starts, uint64_t nrows,
    const uint64_t* __restrict__ total_p, uint64_t total_cap, const uint64_t* __restrict__ split,
    int32_t* __restrict__ left_idx, int32_t* __restrict__ right_idx, GatherSpec gs, EmitSpec es) {
  __shared__ int32_t s_start[kJoinNV + 1];
  __shared__ int32_t s_lb[kJoinNV + 1];
  __shared__ int32_t s_li[kJoinNV];
  __shared__ int32_t s_ri[kJoinNV];

  const uint64_t total = *total_p < total_cap ? *total_p : total_cap;
  const uint32_t tile = blockIdx.x;
  const uint64_t d0 = static_cast<uint64_t>(tile) * kJoinNV;
  if (d0 >= total + nrows) return;
  const uint64_t d1 = (d0 + kJoinNV < total + nrows) ? d0 + kJoinNV : total + nrows;
  const uint64_t g0 = split[tile], g1 = split[tile + 1];
  const uint64_t r0 = d0 - g0, r1 = d1 - g1;
  const uint32_t ng = static_cast<uint32_t>(g1 - g0), nr = static_cast<uint32_t>(r1 - r0);
  const uint32_t tid = threadIdx.x;

  // Rows r0-1 .. r1-1: slot 0 holds the row already open at tile start.
  for (uint32_t i = tid; i < nr + 1; i += kJoinNT) {
    int64_t r = static_cast<int64_t>(r0) - 1 + i;
    s_start[i] = r >= 0 ? starts[r] : 0;
    s_lb[i] = r >= 0 ? lower_bound[r] : 0;
  }
  __syncthreads();

  const uint32_t diag = tid * kJoinVT;
  if (diag < ng + nr) {
    // Local merge path over outputs g0.. and row starts s_start[1..nr].
    uint32_t lo = diag > nr ? diag - nr : 0;
    uint32_t hi = diag < ng ? diag : ng;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      uint64_t s = static_cast<uint32_t>(s_start[1 + diag - 1 - mid]);
      if (g0 + mid < s) lo = mid + 1;
      else hi = mid;
    }
    uint32_t gi = lo, rj = diag - lo;
#pragma unroll
    for (int s = 0; s < kJoinVT; ++s) {
      if (diag + s >= ng + nr) break;
      const bool take_row =
          rj < nr && (gi >= ng || static_cast<uint64_t>(static_cast<uint32_t>(s_start[1 + rj])) <=
                                      g0 + gi);
      if (take_row) {
        ++rj;
      } else {
        // Open row is local slot rj (row r0 - 1 + rj).
        const uint64_t g = g0 + gi;
        const int32_t row = static_cast<int32_t>(r0 - 1 + rj);
        const int32_t rix = static_cast<int32_t>(
            static_cast<uint32_t>(s_lb[rj]) + (g - static_cast<uint32_t>(s_start[rj])));
        s_li[gi] = row;
        s_ri[gi] = rix;
        ++gi;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < ng; i += kJoinNT) {
    if constexpr (MODE == 0) {
      left_idx[g0 + i] = s_li[i];
      right_idx[g0 + i] = s_ri[i];
    } else if constexpr (MODE == 2) {
      emit(es, g0 + i, static_cast<uint32_t>(s_li[i]), static_cast<uint32_t>(s_ri[i]));
    } else {
      store_w(gs.out, g0 + i, gs.out_bytes,
              compose(gs, static_cast<uint32_t>(s_li[i]), static_cast<uint32_t>(s_ri[i])));
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

template <class K>
int find_bounds_impl(rfx_ctx* ctx, int32_t* lb, int32_t* lc, uint64_t* total, const void* left,
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
  const uint64_t ntiles = (nl + nr + kJoinNV - 1) / kJoinNV;
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
      static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, ntiles, split);
  RFX_LAUNCHED();
  {
    RFX_PROF(ctx, "join.find_bounds");
    k_find_bounds<K><<<static_cast<uint32_t>(ntiles), kJoinNT, 0, ctx->stream>>>(
        static_cast<const K*>(left), nl, static_cast<const K*>(right), nr, split,
        static_cast<uint32_t>(ntiles), lb, lc, total, status, ticket);
    RFX_LAUNCHED();
  }
  return RFX_OK;
}

int expand_impl(rfx_ctx* ctx, int mode, int32_t* li, int32_t* ri, const GatherSpec& gs,
                const uint64_t* total, uint64_t total_cap, const int32_t* lb, const int32_t* lc,
                uint64_t nl, const EmitSpec& es = EmitSpec{}) {
  if (nl == 0 || total_cap == 0) return RFX_OK;
  const uint64_t ntiles = (total_cap + nl + kJoinNV - 1) / kJoinNV;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles + 1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* split = w.take<uint64_t>(ntiles + 1);
  k_expand_partition<<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      lc, nl, total, total_cap, ntiles, split);
  RFX_LAUNCHED();
  RFX_PROF(ctx, mode == 0 ? "join.expand" : mode == 1 ? "join.expand_gather" : "join.expand_project");
  if (mode == 0)
    k_expand<0><<<static_cast<uint32_t>(ntiles), kJoinNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
  else if (mode == 1)
    k_expand<1><<<static_cast<uint32_t>(ntiles), kJoinNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
  else
    k_expand<2><<<static_cast<uint32_t>(ntiles), kJoinNT, 0, ctx->stream>>>(
        lb, lc, nl, total, total_cap, split, li, ri, gs, es);
  RFX_LAUNCHED();
  return RFX_OK;
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
  for (int k = 0; k < nfields; ++k) {
    const uint32_t* f = fields + 5 * k;
    const uint32_t col = f[0], src = f[1], in_shift = f[2], bits = f[3], out_shift = f[4];
    if (col > 1 || (col == 1 && !out1) || src > 2 || !es.src[src] || !valid_width(es.src_bytes[src]))
      return fail(RFX_EINVAL, "rfx_join_expand_project: field %d source/column", k);
    if (bits == 0 || bits > 128 || in_shift + bits > 8u * es.src_bytes[src] ||
        out_shift + bits > 8u * es.out_bytes[col])
      return fail(RFX_EINVAL, "rfx_join_expand_project: field %d bits out of range", k);
    es.f[k] = EmitField{static_cast<uint8_t>(col), static_cast<uint8_t>(src),
                        static_cast<uint8_t>(in_shift), static_cast<uint8_t>(bits),
                        static_cast<uint8_t>(out_shift)};
  }
  GatherSpec gs{};
  return expand_impl(ctx, 2, nullptr, nullptr, gs, total, total_cap, lower_bound, left_count, nl,
                     es);
}
