// reduce.hip - reduce-by-key over runs of equal adjacent keys, and single
// reductions, for gfx950.
//
// Replaces redfox::total_*_double / total_64_32 / min_32_double / count_*
// (Reduce.cu:175-1532: thrust::reduce_by_key with D2H/H2D size copies) and
// total_double / count / max_double (Reduce.cu:976-1136: thrust::reduce).
//
// Reduce-by-key makes three launches that never wait on each other's
// progress.  Tiles of TILE consecutive rows (boundaries at multiples of
// TILE from row 0), one per block: a thread folds its VT consecutive rows,
// a fixed shuffle/LDS tree scans the threads with the segmented operator
// (h1, s1) . (h2, s2) = (h1 + h2, h2 ? s2 : s1 op s2), and the upsweep
// writes the tile's aggregate (run heads, open-run value, last head).  One
// block scans the tile aggregates, each thread folding a fixed chunk of
// tiles and a fixed tree combining the chunks; the downsweep reads each
// tile again with its carry-in and writes every run that ends in it, with
// its key, value and row count (end - head + 1).  Every addition's order
// is a function of the row count and the data alone, never of scheduling.
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kBNT = 256;            // threads per reduce-by-key tile
constexpr uint32_t kRChunks = 1024;  // single reductions: fixed chunk count

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

__device__ __forceinline__ uint64_t rows_of(const Sizes& z) {
  uint64_t n = *z.n_ptr / z.n_div;
  return n < z.n_cap ? n : z.n_cap;
}

#ifndef RFX_RBK_VT
#define RFX_RBK_VT 16
#endif
// Rows per thread (halved for 16-byte keys).
template <class K>
struct RbkTune {
  static constexpr int VT = sizeof(K) >= 16 ? RFX_RBK_VT / 2 : RFX_RBK_VT;
  static constexpr int TILE = kBNT * VT;
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
    load_rows<K, VT, VEC>(keys, first, n, k);
    if constexpr (HAS_V) load_rows<V, VT, VEC>(vals, first, n, v);
    s_first[tid] = k[0];
    s_last[tid] = k[VT - 1];
    __syncthreads();
    K prev{}, next{};
    if (tid) prev = s_last[tid - 1];
    else if (base) prev = keys[base - 1];
    if (tid + 1 < kBNT) next = s_first[tid + 1];
    else if (base + TILE < n) next = keys[base + TILE];
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
// tile total.  Threads past the last row hold h = 0.
template <class V, class Op>
__device__ __forceinline__ Seg<V> tile_scan(const Seg<V>& mine, Seg<V>* s_wave, Seg<V>* before,
                                            bool* have_before) {
  constexpr int NW = kBNT / kWave;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid / kWave;
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

template <class K, class V, class Op, int MODE, bool VEC>
__global__ __launch_bounds__(kBNT) void k_rbk_up(const K* __restrict__ keys,
                                                 const V* __restrict__ vals, Sizes z,
                                                 Seg<V>* __restrict__ agg) {
  constexpr int VT = RbkTune<K>::VT;
  constexpr int TILE = RbkTune<K>::TILE;
  __shared__ K s_first[kBNT];
  __shared__ K s_last[kBNT];
  __shared__ Seg<V> s_wave[kBNT / kWave];
  const uint64_t n = rows_of(z);
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * TILE;
  if (base >= n) return;
  TileRows<K, V, Op, (MODE & 1) != 0, VEC, VT> t;
  t.load(keys, vals, n, base, base + static_cast<uint64_t>(threadIdx.x) * VT, s_first, s_last);
  Seg<V> before;
  bool hb;
  const Seg<V> total = tile_scan<V, Op>(t.mine, s_wave, &before, &hb);
  if (threadIdx.x == 0) agg[blockIdx.x] = total;
}

// One block of kRChunks threads: exclusive segmented scan of the tile
// aggregates in place (thread t folds tiles [t*C, t*C + C) in order, a
// fixed Kogge-Stone tree combines the threads; C depends on n only), and
// the run count.
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
__global__ __launch_bounds__(kRChunks) void k_rbk_spine(Seg<V>* __restrict__ agg, Sizes z,
                                                        uint32_t tile_rows,
                                                        uint64_t* __restrict__ runs) {
  __shared__ SegE<V, Op> sm[kRChunks];
  const uint32_t tid = threadIdx.x;
  const uint64_t n = rows_of(z);
  const uint64_t T = (n + tile_rows - 1) / tile_rows;
  const uint64_t C = (T + kRChunks - 1) / kRChunks;
  const uint64_t t0 = tid * C, t1 = t0 + C < T ? t0 + C : T;
  SegE<V, Op> acc{{0, V{}, 0}, true};
  for (uint64_t t = t0; t < t1; ++t) acc = sege_combine(acc, SegE<V, Op>{agg[t], false});
  sm[tid] = acc;
  __syncthreads();
  for (uint32_t d = 1; d < kRChunks; d <<= 1) {
    SegE<V, Op> y = acc;
    if (tid >= d) y = sege_combine(sm[tid - d], acc);
    __syncthreads();
    acc = y;
    sm[tid] = acc;
    __syncthreads();
  }
  SegE<V, Op> carry = tid ? sm[tid - 1] : SegE<V, Op>{{0, V{}, 0}, true};
  for (uint64_t t = t0; t < t1; ++t) {
    const Seg<V> a = agg[t];
    agg[t] = carry.e ? Seg<V>{0, V{}, 0} : carry.s;
    carry = sege_combine(carry, SegE<V, Op>{a, false});
  }
  if (tid == kRChunks - 1) *runs = acc.e ? 0 : acc.s.h;
}

template <class K, class V, class Op, int MODE, bool VEC>
__global__ __launch_bounds__(kBNT) void k_rbk_down(const K* __restrict__ keys,
                                                   const V* __restrict__ vals, Sizes z,
                                                   const Seg<V>* __restrict__ carry,
                                                   K* __restrict__ out_keys,
                                                   V* __restrict__ out_vals,
                                                   uint32_t* __restrict__ out_counts,
                                                   uint64_t out_cap) {
  constexpr int VT = RbkTune<K>::VT;
  constexpr int TILE = RbkTune<K>::TILE;
  constexpr bool HAS_V = MODE & 1;
  constexpr bool HAS_C = MODE & 2;
  __shared__ K s_first[kBNT];
  __shared__ K s_last[kBNT];
  __shared__ Seg<V> s_wave[kBNT / kWave];
  const uint64_t n = rows_of(z);
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * TILE;
  if (base >= n) return;
  const uint64_t first = base + static_cast<uint64_t>(threadIdx.x) * VT;
  TileRows<K, V, Op, HAS_V, VEC, VT> t;
  t.load(keys, vals, n, base, first, s_first, s_last);
  Seg<V> before;
  bool hb;
  (void)tile_scan<V, Op>(t.mine, s_wave, &before, &hb);
  if (!t.any) return;
  // State entering the thread: the tile's carry-in, then the threads
  // before it in the tile.
  const Seg<V> cin = carry[blockIdx.x];  // h = heads before the tile
  const bool tile_open = base != 0;      // a run is open at every tile start but the first
  Seg<V> in;
  bool open;
  if (hb) {
    in = tile_open ? seg_combine<Op>(cin, before) : before;
    open = true;
  } else {
    in = cin;
    open = tile_open;
  }
  // Index of the run open on entry, in 64 bits: with no run open it is
  // -1 and the first head makes it 0.
  uint64_t run = static_cast<uint64_t>(in.h) - 1;
  V val = in.s;
  uint64_t hrow = in.lh - 1;
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
  Carve c(nullptr);
  c.take<Seg<V>>(ntiles);
  c.take<uint64_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve cw(ctx->ws);
  Seg<V>* agg = cw.take<Seg<V>>(ntiles);
  uint64_t* runs = cw.take<uint64_t>(1);
  const bool vec = (reinterpret_cast<uintptr_t>(keys) % 16 == 0) &&
                   (!(MODE & 1) || reinterpret_cast<uintptr_t>(vals) % 16 == 0);
  const K* kp = static_cast<const K*>(keys);
  const V* vp = static_cast<const V*>(vals);
  const uint32_t g = static_cast<uint32_t>(ntiles);
  {
    RFX_PROF(ctx, "rbk.up");
    if (vec) k_rbk_up<K, V, Op, MODE, true><<<g, kBNT, 0, ctx->stream>>>(kp, vp, z, agg);
    else k_rbk_up<K, V, Op, MODE, false><<<g, kBNT, 0, ctx->stream>>>(kp, vp, z, agg);
    RFX_LAUNCHED();
  }
  k_rbk_spine<V, Op><<<1, kRChunks, 0, ctx->stream>>>(agg, z, TILE, runs);
  RFX_LAUNCHED();
  {
    RFX_PROF(ctx, "rbk.down");
    if (vec)
      k_rbk_down<K, V, Op, MODE, true><<<g, kBNT, 0, ctx->stream>>>(
          kp, vp, z, agg, static_cast<K*>(out_keys), static_cast<V*>(out_vals), out_counts, out_cap);
    else
      k_rbk_down<K, V, Op, MODE, false><<<g, kBNT, 0, ctx->stream>>>(
          kp, vp, z, agg, static_cast<K*>(out_keys), static_cast<V*>(out_vals), out_counts, out_cap);
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
