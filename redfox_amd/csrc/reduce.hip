// reduce.hip - reduce-by-key over runs of equal adjacent keys, and single
// reductions, for gfx950.
//
// Replaces redfox::total_*_double / total_64_32 / min_32_double / count_*
// (Reduce.cu:175-1532: thrust::reduce_by_key with D2H/H2D size copies) and
// total_double / count / max_double (Reduce.cu:976-1136: thrust::reduce).
//
// Reduce-by-key reads the rows once.  Tiles of TILE consecutive rows
// (boundaries at multiples of TILE from row 0), one per block, numbered by
// an atomic ticket: a thread folds its VT consecutive rows, a fixed
// shuffle/LDS tree scans the threads with the segmented operator
// (h1, s1) . (h2, s2) = (h1 + h2, h2 ? s2 : s1 op s2), a decoupled
// look-back over head counts (integers, exact in any order) gives the
// tile's first run index, and the tile writes every run that ends in it,
// with its key, value and row count (end - head + 1).  The run open at the
// tile start is written with the tile's part only.  The tile aggregates
// (run heads, open-run value, last head) are then scanned in two levels by
// fixed trees: groups of kSpNT consecutive tiles, one block each, and one
// block over the group totals, each thread folding a fixed chunk of groups.
// A patch kernel completes each tile's open run with its carry-in.  Every
// FP64 operation's order is a function of the row count and the data
// alone, never of scheduling.
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kBNT = 512;            // threads per reduce-by-key tile
constexpr uint32_t kRChunks = 1024;  // single reductions: fixed chunk count
constexpr uint32_t kSpNT = 1024;     // tile aggregates per spine group (one block)

struct OpSum {
  template <class V> __device__ static V apply(V a, V b) { return a + b; }
  template <class V> __device__ static V identity() { return V(0); }
};
struct OpMin {
  template <class V> __device__ static V apply(V a, V b) { return b < a ? b : a; }
  template <class V> __device__ static V identity();
};
template <> __device__ double OpMin::identity<double>() { return __builtin_huge_val(); }
template <> __device__ uint32_t OpMin::identity<uint32_t>() { return 0xFFFFFFFFu; }
struct OpMax {
  template <class V> __device__ static V apply(V a, V b) { return b > a ? b : a; }
  template <class V> __device__ static V identity();
};
template <> __device__ double OpMax::identity<double>() { return -__builtin_huge_val(); }
template <> __device__ uint32_t OpMax::identity<uint32_t>() { return 0u; }
template <> __device__ uint64_t OpMax::identity<uint64_t>() { return 0ull; }

template <class V>
struct Seg {
  uint32_t h;   // run heads in the span
  V s;          // value of the span's open (last) run
  uint64_t lh;  // 1 + row of the span's last head, 0 if none
};

template <class Op, class V>
__device__ __forceinline__ Seg<V> seg_combine(const Seg<V>& a, const Seg<V>& b) {
  return {a.h + b.h, b.h ? b.s : Op::apply(a.s, b.s), b.lh ? b.lh : a.lh};
}

template <class V>
__device__ __forceinline__ Seg<V> seg_shfl_up(const Seg<V>& x, uint32_t d) {
  Seg<V> y;
  y.h = __shfl_up(x.h, d, kWave);
  y.s = __shfl_up(x.s, d, kWave);
  y.lh = __shfl_up(x.lh, d, kWave);
  return y;
}

struct Sizes {
  const uint64_t* n_ptr;  // rows = *n_ptr / n_div, at most n_cap
  uint64_t n_div;
  uint64_t n_cap;
  uint64_t* out_a;        // *out_a = runs * out_a_mul
  uint64_t out_a_mul;
  uint64_t* out_b;        // optional second size (Harmony value bytes)
  uint64_t out_b_mul;
};

__device__ __forceinline__ uint64_t rows_of(const Sizes& z, uint64_t raw) {
  const uint64_t n = raw / z.n_div;
  return n < z.n_cap ? n : z.n_cap;
}
__device__ __forceinline__ uint64_t rows_of(const Sizes& z) { return rows_of(z, *z.n_ptr); }

#ifndef RFX_RBK_VT
#define RFX_RBK_VT 24
#endif
// Rows per thread (halved for 16-byte keys).  On 2^28 u32 keys with f64
// values, 512 x 24 rows measured best among 256-1024 threads x 16-32 rows.
template <class K>
struct RbkTune {
  static constexpr int VT = sizeof(K) >= 16 ? RFX_RBK_VT / 2 : RFX_RBK_VT;
  static constexpr int TILE = kBNT * VT;
  static_assert(VT <= 32, "head and end masks hold 32 rows");
};

// Rows first .. first + VT - 1 of p into out (a thread's consecutive rows);
// rows at or past n are not read.  VEC: 16-byte loads (the column is
// 16-byte aligned and a thread's rows fill whole 16-byte words).
template <class T, int VT, bool VEC>
__device__ __forceinline__ void load_rows(const T* __restrict__ p, uint64_t first, uint64_t n,
                                          T (&out)[VT]) {
  if constexpr (VEC && (VT * sizeof(T)) % 16 == 0) {
    if (first + VT <= n) {
      const uint4* q = reinterpret_cast<const uint4*>(p + first);
      constexpr int W = VT * sizeof(T) / 16;
      uint4 w[W];
#pragma unroll
      for (int c = 0; c < W; ++c) w[c] = q[c];
      __builtin_memcpy(out, w, sizeof(w));
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < VT; ++j)
    if (first + j < n) out[j] = p[first + j];
}

// One tile's rows, heads and run ends, and the thread's aggregate.
template <class K, class V, class Op, bool HAS_V, bool VEC, int VT>
struct TileRows {
  K k[VT];
  V v[VT];
  uint32_t hmask = 0, emask = 0;
  Seg<V> mine{0, V{}, 0};
  bool any = false;

  __device__ __forceinline__ void load(const K* __restrict__ keys, const V* __restrict__ vals,
                                       uint64_t n, uint64_t base, uint64_t first, K* s_first,
                                       K* s_last) {
    constexpr int TILE = kBNT * VT;
    const uint32_t tid = threadIdx.x;
    // The keys just outside the tile go out with the rows (first and last
    // wave), not after the barrier: one round trip less before the scan.
    const uint32_t wv = wave_id();
    K out_prev{}, out_next{};
    if (wv == 0 && base) out_prev = keys[base - 1];
    if (wv == kBNT / kWave - 1 && base + TILE < n) out_next = keys[base + TILE];
    load_rows<K, VT, VEC>(keys, first, n, k);
    if constexpr (HAS_V) load_rows<V, VT, VEC>(vals, first, n, v);
    s_first[tid] = k[0];
    s_last[tid] = k[VT - 1];
    __syncthreads();
    const K prev = tid ? s_last[tid - 1] : out_prev;
    const K next = tid + 1 < kBNT ? s_first[tid + 1] : out_next;
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const uint64_t row = first + j;
      if (row < n) {
        const K pk = j ? k[j - 1] : prev;
        const K nk = j + 1 < VT ? k[j + 1] : next;
        const bool head = row == 0 || !eq(pk, k[j]);
        const bool end = row + 1 >= n || !eq(k[j], nk);
        hmask |= static_cast<uint32_t>(head) << j;
        emask |= static_cast<uint32_t>(end) << j;
        V x{};
        if constexpr (HAS_V) x = v[j];
        if (head) {
          ++mine.h;
          mine.s = x;
          mine.lh = row + 1;
        } else {
          mine.s = any ? Op::apply(mine.s, x) : x;
        }
        any = true;
      }
    }
  }
};

// Block scan of the thread aggregates: *before = combination of the
// threads before this one (have_before false for thread 0), returns the
// tile total.  Threads past the last row hold h = 0 and the identity.
template <class V, class Op>
__device__ __forceinline__ Seg<V> tile_scan(const Seg<V>& mine, Seg<V>* s_wave, Seg<V>* before,
                                            bool* have_before) {
  constexpr int NW = kBNT / kWave;
  const uint32_t lane = lane_id(), wv = wave_id();
  Seg<V> incl = mine;
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    const Seg<V> y = seg_shfl_up(incl, d);
    if (lane >= d) incl = seg_combine<Op>(y, incl);
  }
  if (lane == kWave - 1) s_wave[wv] = incl;
  const Seg<V> excl_w = seg_shfl_up(incl, 1);
  __syncthreads();
  Seg<V> b{0, V{}, 0};
  bool hb = false;
  for (uint32_t i = 0; i < wv; ++i) {
    b = hb ? seg_combine<Op>(b, s_wave[i]) : s_wave[i];
    hb = true;
  }
  if (lane) {
    b = hb ? seg_combine<Op>(b, excl_w) : excl_w;
    hb = true;
  }
  *before = b;
  *have_before = hb;
  Seg<V> total = s_wave[0];
  for (int i = 1; i < NW; ++i) total = seg_combine<Op>(total, s_wave[i]);
  return total;
}

template <class V, class Op>
struct SegE {
  Seg<V> s;
  bool e;  // empty (identity)
};
template <class V, class Op>
__device__ __forceinline__ SegE<V, Op> sege_combine(const SegE<V, Op>& a, const SegE<V, Op>& b) {
  if (a.e) return b;
  if (b.e) return a;
  return {seg_combine<Op>(a.s, b.s), false};
}
template <class V, class Op>
__device__ __forceinline__ SegE<V, Op> sege_shfl_up(const SegE<V, Op>& x, uint32_t d) {
  SegE<V, Op> y;
  y.s = seg_shfl_up(x.s, d);
  y.e = __shfl_up(static_cast<int>(x.e), d, kWave) != 0;
  return y;
}

// Exclusive scan of one value per thread over a block of NT threads by a
// fixed tree: shuffles inside each wave, then the wave totals in order.
// Returns the combination of the threads before this one (empty for
// thread 0) and sets *total to the block's.
template <int NT, class V, class Op>
__device__ __forceinline__ SegE<V, Op> block_scan_excl(SegE<V, Op> incl, SegE<V, Op>* s_wave,
                                                       SegE<V, Op>* total) {
  constexpr int NW = NT / kWave;
  const uint32_t lane = lane_id(), wv = wave_id();
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    const SegE<V, Op> y = sege_shfl_up(incl, d);
    if (lane >= d) incl = sege_combine(y, incl);
  }
  if (lane == kWave - 1) s_wave[wv] = incl;
  const SegE<V, Op> excl_w = sege_shfl_up(incl, 1);
  __syncthreads();
  SegE<V, Op> before{{0, V{}, 0}, true};
  for (uint32_t i = 0; i < wv; ++i) before = sege_combine(before, s_wave[i]);
  if (lane) before = sege_combine(before, excl_w);
  SegE<V, Op> t = s_wave[0];
  for (int i = 1; i < NW; ++i) t = sege_combine(t, s_wave[i]);
  *total = t;
  return before;
}

// Group pass: block b takes tiles [b*kSpNT, b*kSpNT + kSpNT), writes to
// ex[t] the combination of the tiles before t in the group (not written for
// the group's first tile, whose carry is the group's alone; ex may be agg),
// and writes the group total to grp[b].
template <class V, class Op>
__global__ __launch_bounds__(kSpNT) void k_rbk_spine_group(const Seg<V>* agg, Seg<V>* ex,
                                                           Sizes z, uint32_t tile_rows,
                                                           Seg<V>* __restrict__ grp) {
  __shared__ SegE<V, Op> s_wave[kSpNT / kWave];
  const uint64_t n = rows_of(z);
  const uint64_t T = (n + tile_rows - 1) / tile_rows;
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kSpNT;
  if (b0 >= T) return;  // uniform over the block
  const uint64_t t = b0 + threadIdx.x;
  SegE<V, Op> mine{{0, V{}, 0}, true};
  if (t < T) mine = {agg[t], false};
  SegE<V, Op> total;
  const SegE<V, Op> before = block_scan_excl<kSpNT, V, Op>(mine, s_wave, &total);
  if (t < T && !before.e) ex[t] = before.s;
  if (threadIdx.x == 0) grp[blockIdx.x] = total.s;
}

// One block of kRChunks threads: exclusive segmented scan of the group
// totals in place (thread t folds groups [t*C, t*C + C) in order, the block
// scan combines the threads; C depends on n only).
template <class V, class Op>
__global__ __launch_bounds__(kRChunks) void k_rbk_spine(Seg<V>* __restrict__ grp, Sizes z,
                                                        uint32_t tile_rows) {
  __shared__ SegE<V, Op> s_wave[kRChunks / kWave];
  const uint32_t tid = threadIdx.x;
  const uint64_t n = rows_of(z);
  const uint64_t T = ((n + tile_rows - 1) / tile_rows + kSpNT - 1) / kSpNT;
  const uint64_t C = (T + kRChunks - 1) / kRChunks;
  const uint64_t t0 = tid * C, t1 = t0 + C < T ? t0 + C : T;
  SegE<V, Op> acc{{0, V{}, 0}, true};
  for (uint64_t t = t0; t < t1; ++t) acc = sege_combine(acc, SegE<V, Op>{grp[t], false});
  SegE<V, Op> total;
  SegE<V, Op> carry = block_scan_excl<kRChunks, V, Op>(acc, s_wave, &total);
  for (uint64_t t = t0; t < t1; ++t) {
    const Seg<V> a = grp[t];
    grp[t] = carry.e ? Seg<V>{0, V{}, 0} : carry.s;
    carry = sege_combine(carry, SegE<V, Op>{a, false});
  }
}

// ---- one pass over the rows, then a patch of one run per tile -------------
// The rows are read once.  Tiles take numbers from a ticket; a tile's run
// indices come from a decoupled look-back over head counts (integer sums,
// exact in any order).  Every run that ends in the tile is written; the run
// open at the tile start (if it ends here) is written with the tile's own
// part only: its value from the tile's first row and its row count from
// the tile start.  The tile aggregates go to `agg`; the two small spine
// kernels above scan them by a fixed tree, and k_rbk_patch then combines
// the carry-in value with that run's tile part and adds the rows before
// the tile.  So no tile waits for another tile's values, and every FP64
// value is a fixed function of the data.
template <class K, class V, class Op, int MODE, bool VEC>
__global__ __launch_bounds__(kBNT) void k_rbk_onepass(
    const K* __restrict__ keys, const V* __restrict__ vals, Sizes z, uint64_t* st,
    Seg<V>* __restrict__ agg, K* __restrict__ out_keys, V* __restrict__ out_vals,
    uint32_t* __restrict__ out_counts, uint64_t out_cap) {
  constexpr int VT = RbkTune<K>::VT;
  constexpr int TILE = RbkTune<K>::TILE;
  constexpr bool HAS_V = MODE & 1;
  constexpr bool HAS_C = MODE & 2;
  __shared__ K s_first[kBNT];
  __shared__ K s_last[kBNT];
  __shared__ Seg<V> s_wave[kBNT / kWave];
  __shared__ uint64_t s_hx;
  __shared__ uint32_t s_tile;
  // st[0]: ticket (low half); st[1]: run count, from the last tile; st[2 + t]:
  // head-count word of tile t.
  uint64_t* hst = st + 2;
  // The row count is read while the ticket is in flight.
  const uint64_t n_raw = *z.n_ptr;
  if (threadIdx.x == 0) s_tile = atomicAdd(reinterpret_cast<uint32_t*>(st), 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t n = rows_of(z, n_raw);
  const uint64_t base = static_cast<uint64_t>(tile) * TILE;
  if (base >= n) return;  // uniform over the block
  const uint64_t first = base + static_cast<uint64_t>(threadIdx.x) * VT;
  TileRows<K, V, Op, HAS_V, VEC, VT> t;
  t.load(keys, vals, n, base, first, s_first, s_last);
  Seg<V> before;
  bool hb;
  // Threads past the last row carry the operator's identity, so the tile
  // aggregate is exact for MIN and MAX as well.
  Seg<V> mine = t.mine;
  if (!t.any) mine.s = Op::template identity<V>();
  const Seg<V> total = tile_scan<V, Op>(mine, s_wave, &before, &hb);
  if (threadIdx.x < kWave) {
    if (threadIdx.x == 0) {
      agg[tile] = total;
      status_store64(hst + tile, (tile ? kFlagAgg64 : kFlagIncl64) | total.h);
    }
    if (tile) {
      const uint64_t hx = wave_lookback64(hst, tile);
      if (threadIdx.x == 0) {
        status_store64(hst + tile, kFlagIncl64 | (hx + total.h));
        s_hx = hx;
        if (base + TILE >= n) st[1] = hx + total.h;
      }
    } else if (threadIdx.x == 0) {
      s_hx = 0;
      if (base + TILE >= n) st[1] = total.h;
    }
  }
  __syncthreads();
  if (!t.any) return;
  // Entering the thread: the runs before it (heads before the tile, then
  // in the tile); the open run's value and head row within the tile only.
  uint64_t run = s_hx + (hb ? before.h : 0u) - 1;
  V val = hb ? before.s : V{};
  bool open = hb;
  uint64_t hrow = (hb && before.lh) ? before.lh - 1 : base;
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint64_t row = first + j;
    if (row < n) {
      V x{};
      if constexpr (HAS_V) x = t.v[j];
      if ((t.hmask >> j) & 1) {
        ++run;
        val = x;
        hrow = row;
        open = true;
      } else {
        val = open ? Op::apply(val, x) : x;
        open = true;
      }
      if (((t.emask >> j) & 1) && run < out_cap) {
        out_keys[run] = t.k[j];
        if constexpr (HAS_V) out_vals[run] = val;
        if constexpr (HAS_C) out_counts[run] = static_cast<uint32_t>(row - hrow + 1);
      }
    }
  }
}

// One thread per tile: a tile whose first row continues a run that ends in
// the tile holds that run's tile part; add the carry-in (the spine's
// exclusive prefix: group carry, then the tiles before it in its group).
template <class K, class V, class Op, int MODE>
__global__ __launch_bounds__(256) void k_rbk_patch(const K* __restrict__ keys, Sizes z,
                                                   uint32_t tile_rows,
                                                   const Seg<V>* __restrict__ agg,
                                                   const Seg<V>* __restrict__ ex,
                                                   const Seg<V>* __restrict__ grp,
                                                   V* __restrict__ out_vals,
                                                   uint32_t* __restrict__ out_counts,
                                                   uint64_t out_cap) {
  const uint64_t n = rows_of(z);
  const uint64_t T = (n + tile_rows - 1) / tile_rows;
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t == 0 || t >= T) return;
  const uint64_t base = t * tile_rows;
  if (!eq(keys[base - 1], keys[base])) return;  // the tile starts a run
  // With no head in the tile, the open run ends here only at the last row
  // or where the next tile starts a run.
  const uint64_t next = base + tile_rows;
  if (agg[t].h == 0 && next < n && eq(keys[next - 1], keys[next])) return;
  const uint64_t g = t / kSpNT, l = t % kSpNT;
  Seg<V> cin;
  if (g == 0) cin = ex[t];
  else if (l == 0) cin = grp[g];
  else cin = seg_combine<Op>(grp[g], ex[t]);
  const uint64_t r = static_cast<uint64_t>(cin.h) - 1;
  if (r >= out_cap) return;
  if constexpr (MODE & 1) out_vals[r] = Op::apply(cin.s, out_vals[r]);
  if constexpr (MODE & 2) out_counts[r] += static_cast<uint32_t>(base - (cin.lh - 1));
}

// Sizes after the pass: the size variables may alias the row count, so
// they are written only once every tile has read it.
__global__ void k_rbk_sizes(Sizes z, const uint64_t* __restrict__ runs, uint64_t n_cap) {
  if (threadIdx.x == 0) {
    const uint64_t r = n_cap ? *runs : 0;
    *z.out_a = r * z.out_a_mul;
    if (z.out_b) *z.out_b = r * z.out_b_mul;
  }
}

// ---- single reductions ---------------------------------------------------
template <class V, class Op>
__global__ __launch_bounds__(256) void k_reduce_up(const V* __restrict__ v, uint64_t n,
                                                   V* __restrict__ part) {
  __shared__ V sm[256];
  const uint64_t per = (n + kRChunks - 1) / kRChunks;  // fixed by n
  const uint64_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  V acc = Op::template identity<V>();
  for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) acc = Op::apply(acc, v[i]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < static_cast<uint32_t>(d))
      sm[threadIdx.x] = Op::apply(sm[threadIdx.x], sm[threadIdx.x + d]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

template <class V, class Op>
__global__ __launch_bounds__(kRChunks) void k_reduce_final(const V* __restrict__ part,
                                                           V* __restrict__ out) {
  __shared__ V sm[kRChunks];
  sm[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int d = kRChunks / 2; d > 0; d >>= 1) {
    if (threadIdx.x < static_cast<uint32_t>(d))
      sm[threadIdx.x] = Op::apply(sm[threadIdx.x], sm[threadIdx.x + d]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

template <class K, class V, class Op, int MODE>
int rbk_impl(rfx_ctx* ctx, void* out_keys, void* out_vals, uint32_t* out_counts,
             uint64_t out_cap, const void* keys, const void* vals, const Sizes& z) {
  if (z.n_cap == 0) {
    RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_a), 2, 0));
    if (z.out_b) RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_b), 2, 0));
    return RFX_OK;
  }
  constexpr int TILE = RbkTune<K>::TILE;
  const uint64_t ntiles = (z.n_cap + TILE - 1) / TILE;
  if (ntiles > 0xFFFFFFFFull) return fail(RFX_EINVAL, "reduce_by_key: %llu tiles",
                                          static_cast<unsigned long long>(ntiles));
  const uint64_t ngroups = (ntiles + kSpNT - 1) / kSpNT;
  Carve c(nullptr);
  c.take<uint64_t>(2 + ntiles);
  c.take<Seg<V>>(ntiles);
  c.take<Seg<V>>(ntiles);
  c.take<Seg<V>>(ngroups);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve cw(ctx->ws);
  uint64_t* st = cw.take<uint64_t>(2 + ntiles);
  Seg<V>* agg = cw.take<Seg<V>>(ntiles);
  Seg<V>* ex = cw.take<Seg<V>>(ntiles);
  Seg<V>* grp = cw.take<Seg<V>>(ngroups);
  uint64_t* runs = st + 1;  // zero unless a tile ran
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(st), 2 * (2 + ntiles), 0));
  const bool vec = (reinterpret_cast<uintptr_t>(keys) % 16 == 0) &&
                   (!(MODE & 1) || reinterpret_cast<uintptr_t>(vals) % 16 == 0);
  const K* kp = static_cast<const K*>(keys);
  const V* vp = static_cast<const V*>(vals);
  K* ok = static_cast<K*>(out_keys);
  V* ov = static_cast<V*>(out_vals);
  const uint32_t g = static_cast<uint32_t>(ntiles);
  {
    RFX_PROF(ctx, "rbk.pass");
    if (vec)
      k_rbk_onepass<K, V, Op, MODE, true><<<g, kBNT, 0, ctx->stream>>>(kp, vp, z, st, agg, ok, ov,
                                                                       out_counts, out_cap);
    else
      k_rbk_onepass<K, V, Op, MODE, false><<<g, kBNT, 0, ctx->stream>>>(kp, vp, z, st, agg, ok, ov,
                                                                        out_counts, out_cap);
    RFX_LAUNCHED();
  }
  // Carry-ins: only runs that cross tiles need them, so one tile needs none,
  // and one group needs no scan of group totals.
  if (ntiles > 1) {
    k_rbk_spine_group<V, Op><<<static_cast<uint32_t>(ngroups), kSpNT, 0, ctx->stream>>>(
        agg, ex, z, TILE, grp);
    RFX_LAUNCHED();
    if (ngroups > 1) {
      k_rbk_spine<V, Op><<<1, kRChunks, 0, ctx->stream>>>(grp, z, TILE);
      RFX_LAUNCHED();
    }
    k_rbk_patch<K, V, Op, MODE><<<static_cast<uint32_t>((ntiles + 255) / 256), 256, 0,
                                  ctx->stream>>>(kp, z, TILE, agg, ex, grp, ov, out_counts,
                                                 out_cap);
    RFX_LAUNCHED();
  }
  k_rbk_sizes<<<1, 64, 0, ctx->stream>>>(z, runs, z.n_cap);
  RFX_LAUNCHED();
  return RFX_OK;
}

template <class V, class Op, int MODE>
int rbk_keys(rfx_ctx* ctx, int kb, void* ok, void* ov, uint32_t* oc, uint64_t cap, const void* k,
             const void* v, const Sizes& z) {
  switch (kb) {
    case 1: return rbk_impl<uint8_t, V, Op, MODE>(ctx, ok, ov, oc, cap, k, v, z);
    case 2: return rbk_impl<uint16_t, V, Op, MODE>(ctx, ok, ov, oc, cap, k, v, z);
    case 4: return rbk_impl<uint32_t, V, Op, MODE>(ctx, ok, ov, oc, cap, k, v, z);
    case 8: return rbk_impl<uint64_t, V, Op, MODE>(ctx, ok, ov, oc, cap, k, v, z);
    case 16: return rbk_impl<u128, V, Op, MODE>(ctx, ok, ov, oc, cap, k, v, z);
  }
  return fail(RFX_EINVAL, "reduce_by_key: key width %d", kb);
}

template <class V, class Op>
int single_impl(rfx_ctx* ctx, V* result, const V* v, uint64_t n) {
  Carve w0(nullptr);
  w0.take<V>(kRChunks);
  RFX_TRY(ensure_workspace(ctx, w0.off));
  Carve w(ctx->ws);
  V* part = w.take<V>(kRChunks);
  RFX_PROF(ctx, "reduce.single");
  k_reduce_up<V, Op><<<kRChunks, 256, 0, ctx->stream>>>(v, n, part);
  RFX_LAUNCHED();
  k_reduce_final<V, Op><<<1, kRChunks, 0, ctx->stream>>>(part, result);
  RFX_LAUNCHED();
  return RFX_OK;
}

}  // namespace

// Shared with harmony.cpp: reduce-by-key with byte-size variables.
// val_kind: 0 f64 (op), 1 u32 sum, 2 count (out_vals receives u32 counts),
// 3 f64 (op) with counts in out_counts as well.
int reduce_by_key_sized(rfx_ctx* ctx, int op, int val_kind, int kb, void* out_keys,
                        void* out_vals, const void* keys, const void* vals,
                        const uint64_t* n_ptr, uint64_t n_div, uint64_t n_cap, uint64_t* out_a,
                        uint64_t out_a_mul, uint64_t* out_b, uint64_t out_b_mul,
                        uint32_t* out_counts, uint64_t out_cap) {
  Sizes z{n_ptr, n_div, n_cap, out_a, out_a_mul, out_b, out_b_mul};
  if (!ctx || !n_ptr || !out_a) return fail(RFX_EINVAL, "reduce_by_key: null argument");
  if (n_cap && (!keys || !out_keys || !out_vals || (val_kind != 2 && !vals) ||
                (val_kind == 3 && !out_counts)))
    return fail(RFX_EINVAL, "reduce_by_key: null column");
  if (n_cap >= (1ull << 31)) return fail(RFX_ERANGE, "reduce_by_key: too many rows");
  if (n_cap) {
    // No output may overlap an input (as for thrust::reduce_by_key): tiles
    // read their neighbours' keys, and the patch reads keys again after
    // every run is written.  The inputs span n_cap rows; the outputs span
    // min(n_cap, out_cap) runs when the caller states out_cap, else only
    // their first row (their true extent depends on the run count).
    const uint64_t vb = val_kind == 1 || val_kind == 2 ? 4 : 8;
    const uint64_t runs = out_cap == ~0ull ? 1 : out_cap < n_cap ? out_cap : n_cap;
    struct Span { const void* p; uint64_t bytes; };
    const Span outs[3] = {{out_keys, runs * kb}, {out_vals, runs * vb},
                          {val_kind == 3 ? out_counts : nullptr, runs * 4}};
    const Span ins[2] = {{keys, n_cap * kb}, {val_kind == 2 ? nullptr : vals, n_cap * vb}};
    for (const Span& o : outs)
      for (const Span& i : ins) {
        if (!o.p || !i.p || !o.bytes) continue;
        const uintptr_t o0 = reinterpret_cast<uintptr_t>(o.p), i0 = reinterpret_cast<uintptr_t>(i.p);
        if (o0 < i0 + i.bytes && i0 < o0 + o.bytes)
          return fail(RFX_EINVAL, "reduce_by_key: an output overlaps an input");
      }
  }
  if (val_kind == 0 || val_kind == 3) {  // f64
    if (val_kind == 3) {
      switch (op) {
        case RFX_AGG_SUM: return rbk_keys<double, OpSum, 3>(ctx, kb, out_keys, out_vals, out_counts, out_cap, keys, vals, z);
        case RFX_AGG_MIN: return rbk_keys<double, OpMin, 3>(ctx, kb, out_keys, out_vals, out_counts, out_cap, keys, vals, z);
        case RFX_AGG_MAX: return rbk_keys<double, OpMax, 3>(ctx, kb, out_keys, out_vals, out_counts, out_cap, keys, vals, z);
      }
      return fail(RFX_EINVAL, "reduce_by_key: op %d", op);
    }
    switch (op) {
      case RFX_AGG_SUM: return rbk_keys<double, OpSum, 1>(ctx, kb, out_keys, out_vals, nullptr, out_cap, keys, vals, z);
      case RFX_AGG_MIN: return rbk_keys<double, OpMin, 1>(ctx, kb, out_keys, out_vals, nullptr, out_cap, keys, vals, z);
      case RFX_AGG_MAX: return rbk_keys<double, OpMax, 1>(ctx, kb, out_keys, out_vals, nullptr, out_cap, keys, vals, z);
    }
    return fail(RFX_EINVAL, "reduce_by_key: op %d", op);
  }
  if (val_kind == 1) {  // u32 sum
    if (op != RFX_AGG_SUM) return fail(RFX_ENOTSUP, "reduce_by_key: u32 supports sum only");
    return rbk_keys<uint32_t, OpSum, 1>(ctx, kb, out_keys, out_vals, nullptr, out_cap, keys, vals, z);
  }
  return rbk_keys<uint32_t, OpSum, 2>(ctx, kb, out_keys, nullptr, static_cast<uint32_t*>(out_vals),
                                      out_cap, keys, nullptr, z);
}

int single_reduce(rfx_ctx* ctx, int kind, void* result, const void* v, uint64_t n) {
  if (!ctx || !result || (n && !v)) return fail(RFX_EINVAL, "single reduce: null argument");
  switch (kind) {
    case 0: return single_impl<double, OpSum>(ctx, static_cast<double*>(result), static_cast<const double*>(v), n);
    case 1: return single_impl<uint32_t, OpSum>(ctx, static_cast<uint32_t*>(result), static_cast<const uint32_t*>(v), n);
    case 2: return single_impl<uint64_t, OpMax>(ctx, static_cast<uint64_t*>(result), static_cast<const uint64_t*>(v), n);
  }
  return fail(RFX_EINVAL, "single reduce: kind %d", kind);
}

}  // namespace rfx

using namespace rfx;

extern "C" int rfx_reduce_by_key_f64(rfx_ctx* ctx, void* out_keys, double* out_vals,
                                     uint64_t* n_out, const void* keys, const double* vals,
                                     const uint64_t* n, uint64_t n_cap, int key_bytes, int op) {
  return reduce_by_key_sized(ctx, op, 0, key_bytes, out_keys, out_vals, keys, vals, n, 1, n_cap,
                             n_out, 1, nullptr, 0, nullptr, ~0ull);
}

extern "C" int rfx_reduce_count_by_key_f64(rfx_ctx* ctx, void* out_keys, double* out_vals,
                                           uint32_t* out_counts, uint64_t out_cap,
                                           uint64_t* n_out, const void* keys, const double* vals,
                                           const uint64_t* n, uint64_t n_cap, int key_bytes,
                                           int op) {
  return reduce_by_key_sized(ctx, op, 3, key_bytes, out_keys, out_vals, keys, vals, n, 1, n_cap,
                             n_out, 1, nullptr, 0, out_counts, out_cap);
}

extern "C" int rfx_reduce_by_key_u32(rfx_ctx* ctx, void* out_keys, uint32_t* out_vals,
                                     uint64_t* n_out, const void* keys, const uint32_t* vals,
                                     const uint64_t* n, uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 1, key_bytes, out_keys, out_vals, keys, vals, n, 1,
                             n_cap, n_out, 1, nullptr, 0, nullptr, ~0ull);
}

extern "C" int rfx_count_by_key(rfx_ctx* ctx, void* out_keys, uint32_t* out_counts,
                                uint64_t* n_out, const void* keys, const uint64_t* n,
                                uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 2, key_bytes, out_keys, out_counts, keys, nullptr,
                             n, 1, n_cap, n_out, 1, nullptr, 0, nullptr, ~0ull);
}

extern "C" int rfx_total_f64(rfx_ctx* ctx, double* result, const double* v, uint64_t n) {
  return single_reduce(ctx, 0, result, v, n);
}
extern "C" int rfx_total_u32(rfx_ctx* ctx, uint32_t* result, const uint32_t* v, uint64_t n) {
  return single_reduce(ctx, 1, result, v, n);
}
extern "C" int rfx_max_u64(rfx_ctx* ctx, uint64_t* result, const uint64_t* v, uint64_t n) {
  return single_reduce(ctx, 2, result, v, n);
}
