// reduce.hip - reduce-by-key over runs of equal adjacent keys, and single
// reductions, for gfx950.
//
// Replaces redfox::total_*_double / total_64_32 / min_32_double / count_*
// (Reduce.cu:175-1532: thrust::reduce_by_key with D2H/H2D size copies) and
// total_double / count / max_double (Reduce.cu:976-1136: thrust::reduce).
//
// Reduce-by-key is one pass over the rows.  Tiles of TILE consecutive rows
// (boundaries at multiples of TILE from row 0) take tickets in order; a
// thread folds its VT consecutive rows, a fixed shuffle/LDS tree scans the
// threads with the segmented operator (h1, s1) . (h2, s2) = (h1 + h2, h2 ?
// s2 : s1 op s2), and the tile publishes its run-head count, the value of
// its trailing (open) run and the position of its last head.  The run
// index base comes from a decoupled look-back over the head counts
// (integers, exact in any order).  A run that began in earlier tiles is
// finished by the tile where it ends: it walks back to the tile holding the
// run's head and folds their trailing values right to left, p_s op (p_s+1
// op (... op p_e-1)), then applies its own part.  Every addition's order is
// therefore a function of the row count and the data alone, never of
// scheduling, and each run is written once, with its key, value and row
// count (end - head + 1).
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

template <class K>
struct RbkTune {
  static constexpr int VT = sizeof(K) >= 16 ? 8 : 16;
  static constexpr int TILE = kBNT * VT;
};

template <class V>
struct RbkWork {
  uint64_t* status;    // per tile: look-back word over head counts
  V* piece;            // per tile: value of its trailing run
  uint64_t* lasthead;  // per tile: 1 + row of its last head, 0 if none
  uint32_t* ticket;
  uint64_t* runs;      // total run count (written by the last tile)
};

// Rows first .. first + VT - 1 of p into out; rows at or past n are not
// read.  VEC: 16-byte loads (the column is 16-byte aligned and a thread's
// rows fill whole 16-byte words).
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

template <class K>
__device__ __forceinline__ K prev_of_tile(const K* keys, uint64_t base) {
  return keys[base - 1];
}

// MODE bit 0: values (Op over vals); bit 1: row counts.
template <class K, class V, class Op, int MODE, bool VEC>
__global__ __launch_bounds__(kBNT) void k_rbk(const K* __restrict__ keys,
                                              const V* __restrict__ vals, Sizes z, RbkWork<V> w,
                                              K* __restrict__ out_keys, V* __restrict__ out_vals,
                                              uint32_t* __restrict__ out_counts) {
  constexpr int VT = RbkTune<K>::VT;
  constexpr int TILE = RbkTune<K>::TILE;
  constexpr int NW = kBNT / kWave;
  constexpr bool HAS_V = MODE & 1;
  constexpr bool HAS_C = MODE & 2;
  __shared__ K s_first[kBNT];
  __shared__ K s_last[kBNT];
  __shared__ Seg<V> s_wave[NW];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_excl;
  __shared__ V s_carry;
  __shared__ uint64_t s_carry_head;

  const uint64_t n = rows_of(z);
  const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid / kWave;
  if (tid == 0) s_tile = atomicAdd(w.ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t base = static_cast<uint64_t>(tile) * TILE;
  if (base >= n) return;
  const uint64_t first = base + static_cast<uint64_t>(tid) * VT;

  K k[VT];
  V v[VT];
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

  // Heads and run ends of the thread's rows; the thread's aggregate.
  uint32_t hmask = 0, emask = 0;
  Seg<V> mine{0, V{}, 0};
  bool any = false;
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

  // Block scan of the thread aggregates (threads past the last row hold
  // h = 0 and never feed a result).
  Seg<V> incl = mine;
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    const Seg<V> y = seg_shfl_up(incl, d);
    if (lane >= d) incl = seg_combine<Op>(y, incl);
  }
  if (lane == kWave - 1) s_wave[wv] = incl;
  Seg<V> excl_w = seg_shfl_up(incl, 1);  // exclusive within the wave (lane 0: none)
  __syncthreads();
  Seg<V> before{0, V{}, 0};
  bool have_before = false;
  for (uint32_t i = 0; i < wv; ++i) {
    before = have_before ? seg_combine<Op>(before, s_wave[i]) : s_wave[i];
    have_before = true;
  }
  if (lane) {
    before = have_before ? seg_combine<Op>(before, excl_w) : excl_w;
    have_before = true;
  }
  Seg<V> total = s_wave[0];
  for (int i = 1; i < NW; ++i) total = seg_combine<Op>(total, s_wave[i]);
  // Does the tile's first row continue a run from an earlier tile?
  const bool continues = base != 0 && eq(prev_of_tile(keys, base), s_first[0]);

  // Publish, then the look-back (wave 0) and the carry walk (wave 1).
  if (tid == 0) {
    w.piece[tile] = total.s;
    w.lasthead[tile] = total.lh;
    __threadfence();
    status_store64(w.status + tile, (tile == 0 ? kFlagIncl64 : kFlagAgg64) | total.h);
  }
  if (wv == 0) {
    uint64_t ex = 0;
    if (tile) {
      ex = wave_lookback64(w.status, tile);
      if (tid == 0) status_store64(w.status + tile, kFlagIncl64 | (ex + total.h));
    }
    if (tid == 0) {
      s_excl = ex;
      if (base + TILE >= n) *w.runs = ex + total.h;
    }
  } else if (wv == 1 && continues) {
    // Walk back to the tile holding the open run's head, folding the
    // trailing values right to left in walk order.
    V acc{};
    bool have = false;
    uint64_t head_row = 0;
    int64_t t0 = static_cast<int64_t>(tile) - 1;
    bool done = false;
    while (!done) {
      const int64_t kt = t0 - static_cast<int64_t>(lane);
      bool pub = kt < 0;  // tile 0 always holds a head, so lanes past it are never folded
      while (true) {
        if (!pub) pub = (status_load64(w.status + kt) >> 62) != 0;
        const uint64_t ready = ballot(pub);
        const uint32_t prefix = ready == ~0ull ? 64u : static_cast<uint32_t>(__builtin_ctzll(~ready));
        if (prefix == 0) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        __threadfence();
        const bool in = lane < prefix && kt >= 0;
        const uint64_t lh = in ? w.lasthead[kt] : 0;
        const V pc = in ? w.piece[kt] : V{};
        const uint64_t hm = ballot(in && lh != 0);
        if (!hm && prefix < 64) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const uint32_t stop = hm ? static_cast<uint32_t>(__builtin_ctzll(hm)) : 63u;
        for (uint32_t j = 0; j <= stop; ++j) {
          const V pj = __shfl(pc, j, kWave);
          acc = have ? Op::apply(pj, acc) : pj;
          have = true;
        }
        if (hm) {
          head_row = __shfl(lh, stop, kWave) - 1;
          done = true;
        } else {
          t0 -= kWave;
        }
        break;
      }
    }
    if (lane == 0) {
      s_carry = acc;
      s_carry_head = head_row;
    }
  }
  __syncthreads();

  // Replay: write every run that ends in the thread's rows.
  if (!any) return;
  uint64_t run = s_excl + before.h - 1;  // index of the run open on entry
  V val{};
  uint64_t hrow = 0;
  bool open = false;
  if (have_before && before.h) {  // the open run began earlier in this tile
    val = before.s;
    hrow = before.lh - 1;
    open = true;
  } else if (continues) {  // it began in an earlier tile
    val = have_before ? Op::apply(s_carry, before.s) : s_carry;
    hrow = s_carry_head;
    open = true;
  }
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint64_t row = first + j;
    if (row < n) {
      V x{};
      if constexpr (HAS_V) x = v[j];
      if ((hmask >> j) & 1) {
        ++run;
        val = x;
        hrow = row;
        open = true;
      } else {
        val = open ? Op::apply(val, x) : x;
        open = true;
      }
      if ((emask >> j) & 1) {
        out_keys[run] = k[j];
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
             const void* keys, const void* vals, const Sizes& z) {
  if (z.n_cap == 0) {
    RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_a), 2, 0));
    if (z.out_b) RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(z.out_b), 2, 0));
    return RFX_OK;
  }
  constexpr int TILE = RbkTune<K>::TILE;
  const uint64_t ntiles = (z.n_cap + TILE - 1) / TILE;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  c.take<uint64_t>(1);
  c.take<V>(ntiles);
  c.take<uint64_t>(ntiles);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve cw(ctx->ws);
  RbkWork<V> w;
  w.status = cw.take<uint64_t>(ntiles);
  w.ticket = cw.take<uint32_t>(1);
  w.runs = cw.take<uint64_t>(1);
  w.piece = cw.take<V>(ntiles);
  w.lasthead = cw.take<uint64_t>(ntiles);
  // status, ticket and runs are contiguous after alignment; zero them.
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(w.status),
                   (reinterpret_cast<char*>(w.runs + 1) - reinterpret_cast<char*>(w.status)) / 4,
                   0));
  const bool vec = (reinterpret_cast<uintptr_t>(keys) % 16 == 0) &&
                   (!(MODE & 1) || reinterpret_cast<uintptr_t>(vals) % 16 == 0);
  {
    RFX_PROF(ctx, "rbk");
    if (vec)
      k_rbk<K, V, Op, MODE, true><<<static_cast<uint32_t>(ntiles), kBNT, 0, ctx->stream>>>(
          static_cast<const K*>(keys), static_cast<const V*>(vals), z, w,
          static_cast<K*>(out_keys), static_cast<V*>(out_vals), out_counts);
    else
      k_rbk<K, V, Op, MODE, false><<<static_cast<uint32_t>(ntiles), kBNT, 0, ctx->stream>>>(
          static_cast<const K*>(keys), static_cast<const V*>(vals), z, w,
          static_cast<K*>(out_keys), static_cast<V*>(out_vals), out_counts);
    RFX_LAUNCHED();
  }
  k_rbk_sizes<<<1, 64, 0, ctx->stream>>>(z, w.runs, z.n_cap);
  RFX_LAUNCHED();
  return RFX_OK;
}

template <class V, class Op, int MODE>
int rbk_keys(rfx_ctx* ctx, int kb, void* ok, void* ov, uint32_t* oc, const void* k,
             const void* v, const Sizes& z) {
  switch (kb) {
    case 1: return rbk_impl<uint8_t, V, Op, MODE>(ctx, ok, ov, oc, k, v, z);
    case 2: return rbk_impl<uint16_t, V, Op, MODE>(ctx, ok, ov, oc, k, v, z);
    case 4: return rbk_impl<uint32_t, V, Op, MODE>(ctx, ok, ov, oc, k, v, z);
    case 8: return rbk_impl<uint64_t, V, Op, MODE>(ctx, ok, ov, oc, k, v, z);
    case 16: return rbk_impl<u128, V, Op, MODE>(ctx, ok, ov, oc, k, v, z);
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
                        uint32_t* out_counts) {
  Sizes z{n_ptr, n_div, n_cap, out_a, out_a_mul, out_b, out_b_mul};
  if (!ctx || !n_ptr || !out_a) return fail(RFX_EINVAL, "reduce_by_key: null argument");
  if (n_cap && (!keys || !out_keys || !out_vals || (val_kind != 2 && !vals) ||
                (val_kind == 3 && !out_counts)))
    return fail(RFX_EINVAL, "reduce_by_key: null column");
  if (n_cap >= (1ull << 31)) return fail(RFX_ERANGE, "reduce_by_key: too many rows");
  if (val_kind == 0 || val_kind == 3) {  // f64
    if (val_kind == 3) {
      switch (op) {
        case RFX_AGG_SUM: return rbk_keys<double, OpSum, 3>(ctx, kb, out_keys, out_vals, out_counts, keys, vals, z);
        case RFX_AGG_MIN: return rbk_keys<double, OpMin, 3>(ctx, kb, out_keys, out_vals, out_counts, keys, vals, z);
        case RFX_AGG_MAX: return rbk_keys<double, OpMax, 3>(ctx, kb, out_keys, out_vals, out_counts, keys, vals, z);
      }
      return fail(RFX_EINVAL, "reduce_by_key: op %d", op);
    }
    switch (op) {
      case RFX_AGG_SUM: return rbk_keys<double, OpSum, 1>(ctx, kb, out_keys, out_vals, nullptr, keys, vals, z);
      case RFX_AGG_MIN: return rbk_keys<double, OpMin, 1>(ctx, kb, out_keys, out_vals, nullptr, keys, vals, z);
      case RFX_AGG_MAX: return rbk_keys<double, OpMax, 1>(ctx, kb, out_keys, out_vals, nullptr, keys, vals, z);
    }
    return fail(RFX_EINVAL, "reduce_by_key: op %d", op);
  }
  if (val_kind == 1) {  // u32 sum
    if (op != RFX_AGG_SUM) return fail(RFX_ENOTSUP, "reduce_by_key: u32 supports sum only");
    return rbk_keys<uint32_t, OpSum, 1>(ctx, kb, out_keys, out_vals, nullptr, keys, vals, z);
  }
  return rbk_keys<uint32_t, OpSum, 2>(ctx, kb, out_keys, nullptr, static_cast<uint32_t*>(out_vals),
                                      keys, nullptr, z);
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
                             n_out, 1, nullptr, 0, nullptr);
}

extern "C" int rfx_reduce_count_by_key_f64(rfx_ctx* ctx, void* out_keys, double* out_vals,
                                           uint32_t* out_counts, uint64_t* n_out,
                                           const void* keys, const double* vals,
                                           const uint64_t* n, uint64_t n_cap, int key_bytes,
                                           int op) {
  return reduce_by_key_sized(ctx, op, 3, key_bytes, out_keys, out_vals, keys, vals, n, 1, n_cap,
                             n_out, 1, nullptr, 0, out_counts);
}

extern "C" int rfx_reduce_by_key_u32(rfx_ctx* ctx, void* out_keys, uint32_t* out_vals,
                                     uint64_t* n_out, const void* keys, const uint32_t* vals,
                                     const uint64_t* n, uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 1, key_bytes, out_keys, out_vals, keys, vals, n, 1,
                             n_cap, n_out, 1, nullptr, 0, nullptr);
}

extern "C" int rfx_count_by_key(rfx_ctx* ctx, void* out_keys, uint32_t* out_counts,
                                uint64_t* n_out, const void* keys, const uint64_t* n,
                                uint64_t n_cap, int key_bytes) {
  return reduce_by_key_sized(ctx, RFX_AGG_SUM, 2, key_bytes, out_keys, out_counts, keys, nullptr,
                             n, 1, n_cap, n_out, 1, nullptr, 0, nullptr);
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
