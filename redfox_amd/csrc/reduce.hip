// reduce.hip - reduce-by-key over runs of equal adjacent keys, and single
// reductions, for gfx950.
//
// Replaces redfox::total_*_double / total_64_32 / min_32_double / count_*
// (Reduce.cu:175-1532: thrust::reduce_by_key with D2H/H2D size copies) and
// total_double / count / max_double (Reduce.cu:976-1136: thrust::reduce).
//
// Floating-point results must not depend on scheduling, so the order of
// every addition is fixed by the row count alone: the rows are cut into
// G = 1024 chunks of whole 2048-row tiles (chunk length depends on n only);
// an upsweep folds each chunk into (heads, open-segment value) in tile
// order, one block scans the chunk aggregates, and a downsweep replays each
// chunk with its carry-in and writes one row per run.  The segmented
// operator is (h1, s1) . (h2, s2) = (h1 + h2, h2 ? s2 : s1 op s2).
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kRNT = 256;
constexpr int kRVT = 8;
constexpr int kRTile = kRNT * kRVT;
constexpr uint32_t kRChunks = 1024;  // fixed chunk count: order depends on n only

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
  uint32_t h;  // heads (run starts) in the span
  V s;         // value of the span's open (last) run
};

template <class Op, class V>
__device__ __forceinline__ Seg<V> seg_combine(Seg<V> a, Seg<V> b) {
  return {a.h + b.h, b.h ? b.s : Op::apply(a.s, b.s)};
}

// Inclusive segmented scan over the block's threads, Kogge-Stone through
// LDS (a fixed combination tree).
template <class Op, class V>
__device__ __forceinline__ Seg<V> block_seg_scan_incl(Seg<V> x, Seg<V>* sm) {
  const uint32_t tid = threadIdx.x;
  sm[tid] = x;
  __syncthreads();
  for (uint32_t d = 1; d < kRNT; d <<= 1) {
    Seg<V> y = x;
    if (tid >= d) y = seg_combine<Op>(sm[tid - d], x);
    __syncthreads();
    x = y;
    sm[tid] = x;
    __syncthreads();
  }
  return x;
}

// Values: explicit column (VB > 0) or implicit ones (count).
template <class V, bool ONES>
__device__ __forceinline__ V load_val(const V* vals, uint64_t i) {
  if constexpr (ONES) return V(1);
  else return vals[i];
}

struct Sizes {
  const uint64_t* n_ptr;  // rows = *n_ptr / n_div, at most n_cap
  uint64_t n_div;
  uint64_t n_cap;
  uint64_t* out_a;        // *out_a = runs * out_a_mul
  uint64_t out_a_mul;
  uint64_t* out_b;        // optional second size (Harmony value bytes)
  uint64_t out_b_mul;
  uint64_t* n_saved;      // workspace copy of the row count: the size
                          // variables may be in/out and alias n_ptr
};

__device__ __forceinline__ uint64_t rows_of(const Sizes& z) {
  uint64_t n = *z.n_ptr / z.n_div;
  return n < z.n_cap ? n : z.n_cap;
}
__device__ __forceinline__ uint64_t chunk_tiles(uint64_t n) {
  uint64_t tiles = (n + kRTile - 1) / kRTile;
  return (tiles + kRChunks - 1) / kRChunks;
}

// One tile: loads rows [base, base + kRTile) into LDS and computes each
// thread's (heads, open value) over its kRVT consecutive rows.
template <class K, class V, class Op, bool ONES>
__device__ __forceinline__ Seg<V> load_tile(const K* __restrict__ keys, const V* __restrict__ vals,
                                            uint64_t n, uint64_t base, K* sk, V* sv,
                                            uint8_t* head, bool* next_head) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kRTile; i += kRNT) {
    const uint64_t idx = base + i;
    if (idx < n) {
      sk[i] = keys[idx];
      sv[i] = load_val<V, ONES>(vals, idx);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < kRTile; i += kRNT) {
    const uint64_t idx = base + i;
    if (idx < n) {
      bool h;
      if (idx == 0) h = true;
      else if (i == 0) h = !eq(keys[idx - 1], sk[0]);
      else h = !eq(sk[i - 1], sk[i]);
      head[i] = h;
    }
  }
  if (tid == 0) {
    const uint64_t nx = base + kRTile;
    *next_head = nx >= n ? true : !eq(keys[nx - 1], keys[nx]);
  }
  __syncthreads();
  Seg<V> t{0, Op::template identity<V>()};
  for (int j = 0; j < kRVT; ++j) {
    const uint32_t i = tid * kRVT + j;
    if (base + i >= n) break;
    if (head[i]) t = {t.h + 1, sv[i]};
    else t.s = Op::apply(t.s, sv[i]);
  }
  return t;
}

template <class K, class V, class Op, bool ONES>
__global__ __launch_bounds__(kRNT) void k_rbk_up(const K* __restrict__ keys,
                                                 const V* __restrict__ vals, Sizes z,
                                                 Seg<V>* __restrict__ chunk_agg) {
  __shared__ K sk[kRTile];
  __shared__ V sv[kRTile];
  __shared__ uint8_t head[kRTile];
  __shared__ Seg<V> sm[kRNT];
  __shared__ bool next_head;
  const uint64_t n = rows_of(z);
  const uint64_t ct = chunk_tiles(n);
  const uint64_t c = blockIdx.x;
  if (n == 0 || c * ct * kRTile >= n) return;
  Seg<V> agg{0, Op::template identity<V>()};
  for (uint64_t t = 0; t < ct; ++t) {
    const uint64_t base = (c * ct + t) * kRTile;
    if (base >= n) break;
    Seg<V> mine = load_tile<K, V, Op, ONES>(keys, vals, n, base, sk, sv, head, &next_head);
    Seg<V> incl = block_seg_scan_incl<Op>(mine, sm);
    if (threadIdx.x == kRNT - 1) agg = seg_combine<Op>(agg, incl);
    __syncthreads();
  }
  if (threadIdx.x == kRNT - 1) chunk_agg[c] = agg;
}

// One block: exclusive segmented scan of the chunk aggregates; writes the
// run count.
template <class V, class Op>
__global__ __launch_bounds__(kRChunks) void k_rbk_spine(Seg<V>* __restrict__ chunk_agg, Sizes z) {
  __shared__ Seg<V> sm[kRChunks];
  __shared__ uint32_t tot;
  const uint64_t n = rows_of(z);
  const uint64_t ct = chunk_tiles(n);
  const uint64_t nch = n ? (n + ct * kRTile - 1) / (ct * kRTile) : 0;
  const uint32_t tid = threadIdx.x;
  Seg<V> x = tid < nch ? chunk_agg[tid] : Seg<V>{0, Op::template identity<V>()};
  sm[tid] = x;
  __syncthreads();
  for (uint32_t d = 1; d < kRChunks; d <<= 1) {
    Seg<V> y = x;
    if (tid >= d) y = seg_combine<Op>(sm[tid - d], x);
    __syncthreads();
    x = y;
    sm[tid] = x;
    __syncthreads();
  }
  Seg<V> excl = tid ? sm[tid - 1] : Seg<V>{0, Op::template identity<V>()};
  if (tid == kRChunks - 1) tot = x.h;
  __syncthreads();
  if (tid < nch) chunk_agg[tid] = excl;
  if (tid == 0) {
    *z.n_saved = n;  // every thread has read *n_ptr before the barriers above
    *z.out_a = static_cast<uint64_t>(tot) * z.out_a_mul;
    if (z.out_b) *z.out_b = static_cast<uint64_t>(tot) * z.out_b_mul;
  }
}

template <class K, class V, class Op, bool ONES>
__global__ __launch_bounds__(kRNT) void k_rbk_down(const K* __restrict__ keys,
                                                   const V* __restrict__ vals, Sizes z,
                                                   const Seg<V>* __restrict__ chunk_carry,
                                                   K* __restrict__ out_keys,
                                                   V* __restrict__ out_vals) {
  __shared__ K sk[kRTile];
  __shared__ V sv[kRTile];
  __shared__ uint8_t head[kRTile];
  __shared__ Seg<V> sm[kRNT];
  __shared__ bool next_head;
  const uint64_t n = *z.n_saved;
  const uint64_t ct = chunk_tiles(n);
  const uint64_t c = blockIdx.x;
  if (n == 0 || c * ct * kRTile >= n) return;
  const uint32_t tid = threadIdx.x;
  Seg<V> carry = chunk_carry[c];  // heads before the chunk, open value
  for (uint64_t t = 0; t < ct; ++t) {
    const uint64_t base = (c * ct + t) * kRTile;
    if (base >= n) break;
    Seg<V> mine = load_tile<K, V, Op, ONES>(keys, vals, n, base, sk, sv, head, &next_head);
    Seg<V> incl = block_seg_scan_incl<Op>(mine, sm);
    Seg<V> before = tid ? sm[tid - 1] : Seg<V>{0, Op::template identity<V>()};
    Seg<V> in = seg_combine<Op>(carry, before);  // state entering this thread
    // Replay the thread's rows.
    uint64_t run = static_cast<uint64_t>(in.h) - 1;  // index of the open run
    V acc = in.s;
    for (int j = 0; j < kRVT; ++j) {
      const uint32_t i = tid * kRVT + j;
      const uint64_t idx = base + i;
      if (idx >= n) break;
      if (head[i]) {
        ++run;
        acc = sv[i];
      } else {
        acc = Op::apply(acc, sv[i]);
      }
      const bool last = (i + 1 < kRTile) ? (idx + 1 >= n || head[i + 1]) : next_head;
      if (last) {
        out_keys[run] = sk[i];
        out_vals[run] = acc;
      }
    }
    const Seg<V> tile_all = sm[kRNT - 1];
    __syncthreads();
    carry = seg_combine<Op>(carry, tile_all);
    (void)incl;
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

template <class K, class V, class Op, bool ONES>
int rbk_impl(rfx_ctx* ctx, void* out_keys, void* out_vals, const void* keys, const void* vals,
             const Sizes& z) {
  if (z.n_cap == 0) {
    RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_a), 2, 0));
    if (z.out_b) RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_b), 2, 0));
    return RFX_OK;
  }
  Carve w0(nullptr);
  w0.take<Seg<V>>(kRChunks);
  w0.take<uint64_t>(1);
  RFX_TRY(ensure_workspace(ctx, w0.off));
  Carve w(ctx->ws);
  Seg<V>* agg = w.take<Seg<V>>(kRChunks);
  Sizes zz = z;
  zz.n_saved = w.take<uint64_t>(1);
  const uint64_t tiles_cap = (z.n_cap + kRTile - 1) / kRTile;
  const uint32_t grid = static_cast<uint32_t>(tiles_cap < kRChunks ? tiles_cap : kRChunks);
  {
    RFX_PROF(ctx, "rbk.up");
    k_rbk_up<K, V, Op, ONES><<<grid, kRNT, 0, ctx->stream>>>(
        static_cast<const K*>(keys), static_cast<const V*>(vals), zz, agg);
    RFX_LAUNCHED();
  }
  k_rbk_spine<V, Op><<<1, kRChunks, 0, ctx->stream>>>(agg, zz);
  RFX_LAUNCHED();
  RFX_PROF(ctx, "rbk.down");
  k_rbk_down<K, V, Op, ONES><<<grid, kRNT, 0, ctx->stream>>>(
      static_cast<const K*>(keys), static_cast<const V*>(vals), zz, agg, static_cast<K*>(out_keys),
      static_cast<V*>(out_vals));
  RFX_LAUNCHED();
  return RFX_OK;
}

template <class V, class Op, bool ONES>
int rbk_keys(rfx_ctx* ctx, int kb, void* ok, void* ov, const void* k, const void* v,
             const Sizes& z) {
  switch (kb) {
    case 1: return rbk_impl<uint8_t, V, Op, ONES>(ctx, ok, ov, k, v, z);
    case 2: return rbk_impl<uint16_t, V, Op, ONES>(ctx, ok, ov, k, v, z);
    case 4: return rbk_impl<uint32_t, V, Op, ONES>(ctx, ok, ov, k, v, z);
    case 8: return rbk_impl<uint64_t, V, Op, ONES>(ctx, ok, ov, k, v, z);
    case 16: return rbk_impl<u128, V, Op, ONES>(ctx, ok, ov, k, v, z);
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
int reduce_by_key_sized(rfx_ctx* ctx, int op, int val_kind, int kb, void* out_keys,
                        void* out_vals, const void* keys, const void* vals,
                        const uint64_t* n_ptr, uint64_t n_div, uint64_t n_cap, uint64_t* out_a,
                        uint64_t out_a_mul, uint64_t* out_b, uint64_t out_b_mul) {
  Sizes z{n_ptr, n_div, n_cap, out_a, out_a_mul, out_b, out_b_mul, nullptr};
  if (!ctx || !n_ptr || !out_a) return fail(RFX_EINVAL, "reduce_by_key: null argument");
  if (n_cap && (!keys || !out_keys || !out_vals || (val_kind != 2 && !vals)))
    return fail(RFX_EINVAL, "reduce_by_key: null column");
  if (n_cap >= (1ull << 31)) return fail(RFX_ERANGE, "reduce_by_key: too many rows");
  if (val_kind == 0) {  // f64
    switch (op) {
      case RFX_AGG_SUM: return rbk_keys<double, OpSum, false>(ctx, kb, out_keys, out_vals, keys, vals, z);
      case RFX_AGG_MIN: return rbk_keys<double, OpMin, false>(ctx, kb, out_keys, out_vals, keys, vals, z);
      case RFX_AGG_MAX: return rbk_keys<double, OpMax, false>(ctx, kb, out_keys, out_vals, keys, vals, z);
    }
    return fail(RFX_EINVAL, "reduce_by_key: op %d", op);
  }
  if (val_kind == 1) {  // u32 sum
    if (op != RFX_AGG_SUM) return fail(RFX_ENOTSUP, "reduce_by_key: u32 supports sum only");
    return rbk_keys<uint32_t, OpSum, false>(ctx, kb, out_keys, out_vals, keys, vals, z);
  }
  return rbk_keys<uint32_t, OpSum, true>(ctx, kb, out_keys, out_vals, keys, nullptr, z);
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
                             n_out, 1, nullptr, 0);
}

extern "C" int rfx_reduce_by_key_u32(rfx_ctx* ctx, void* out_keys, uint32_t* out_vals,
                                     uint64_t* n_out, const void* keys, const uint32_t* vals,
                                     const uint64_t* n, uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 1, key_bytes, out_keys, out_vals, keys, vals, n, 1,
                             n_cap, n_out, 1, nullptr, 0);
}

extern "C" int rfx_count_by_key(rfx_ctx* ctx, void* out_keys, uint32_t* out_counts,
                                uint64_t* n_out, const void* keys, const uint64_t* n,
                                uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 2, key_bytes, out_keys, out_counts, keys, nullptr,
                             n, 1, n_cap, n_out, 1, nullptr, 0);
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
