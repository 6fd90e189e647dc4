// setops.hip - multiset union / difference / intersection of sorted inputs
// and unique, for gfx950.
//
// Replaces redfox::set_union / difference / unique (Union.cu:47-148,
// Difference.cu:35-143, Unique.cu:36-114: thrust::set_union /
// set_difference / unique followed by an H2D copy of the output size) and
// supplies the intersection the reference declares but never implemented
// (RelationalAlgebraKernel.cpp:919-922).
//
// One pass: tiles of the merge path of A and B (A wins ties) are loaded to
// LDS.  Each thread finds its diagonal's merge-path point with one search
// and then merges its items serially; the merge pointer into the other
// input gives each item's bound there.  From its rank among equal values
// in its own input and the count of equal values in the other input an
// item is kept or not (union keeps all of A and the copies of B beyond
// A's count; difference keeps the copies of A beyond B's count;
// intersection the copies of A within B's count).  The kept items are
// scanned in merged order and compacted, the tile offset coming from a
// decoupled look-back.  Runs of equal values that cross a tile edge are
// finished by galloping in global memory, which starts only when the
// neighbour just outside the tile is equal.  (Per-item binary searches of
// the tile, run in lockstep, measured 2.72 ms against 1.59 ms for the
// first merge-path version on a 2^27 + 2^27 u64 union: 22 LDS reads per
// item against about one.)
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kSNT = 256;
constexpr int kSVT = 8;
constexpr int kSNV = kSNT * kSVT;

template <class K>
__device__ __forceinline__ uint64_t mp_lower(const K* a, uint64_t na, const K* b, uint64_t nb,
                                             uint64_t diag) {
  uint64_t lo = diag > nb ? diag - nb : 0;
  uint64_t hi = diag < na ? diag : na;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (!lt(b[diag - 1 - mid], a[mid])) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Element shuffles across the wave (lane - delta / lane + delta; a lane
// with no such neighbour gets its own value).
template <class K>
__device__ __forceinline__ K shfl_up_elem(K v, int delta) {
  if constexpr (sizeof(K) <= 4) {
    return static_cast<K>(__shfl_up(static_cast<uint32_t>(v), delta, kWave));
  } else if constexpr (sizeof(K) == 8) {
    const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), delta, kWave);
    const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), delta, kWave);
    return (static_cast<uint64_t>(hi) << 32) | lo;
  } else {
    return K{shfl_up_elem<uint64_t>(v.lo, delta), shfl_up_elem<uint64_t>(v.hi, delta)};
  }
}
template <class K>
__device__ __forceinline__ K shfl_down_elem(K v, int delta) {
  if constexpr (sizeof(K) <= 4) {
    return static_cast<K>(__shfl_down(static_cast<uint32_t>(v), delta, kWave));
  } else if constexpr (sizeof(K) == 8) {
    const uint32_t lo = __shfl_down(static_cast<uint32_t>(v), delta, kWave);
    const uint32_t hi = __shfl_down(static_cast<uint32_t>(v >> 32), delta, kWave);
    return (static_cast<uint64_t>(hi) << 32) | lo;
  } else {
    return K{shfl_down_elem<uint64_t>(v.lo, delta), shfl_down_elem<uint64_t>(v.hi, delta)};
  }
}

template <class K>
__global__ __launch_bounds__(256) void k_set_partition(const K* __restrict__ a, uint64_t na,
                                                       const K* __restrict__ b, uint64_t nb,
                                                       uint64_t ntiles, uint32_t nv,
                                                       uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles) return;
  uint64_t diag = t * nv;
  if (diag > na + nb) diag = na + nb;
  split[t] = mp_lower(a, na, b, nb, diag);
}

// A view of one input: LDS for the tile's window [w0, w1), global elsewhere.
template <class K>
struct Side {
  const K* g;
  const K* s;
  uint64_t n, w0, w1;
  __device__ __forceinline__ K at(uint64_t i) const { return (i >= w0 && i < w1) ? s[i - w0] : g[i]; }
  // Number of elements equal to x from `from` forward (x == at(from) or 0).
  __device__ uint64_t run_fwd(uint64_t from, K x) const {
    if (from >= n || !eq(at(from), x)) return 0;
    uint64_t lo = from + 1, step = 1, hi;
    while (true) {
      hi = lo + step - 1;
      if (hi >= n) { hi = n; break; }
      if (!eq(at(hi), x)) break;
      lo = hi + 1;
      step <<= 1;
    }
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (eq(at(mid), x)) lo = mid + 1;
      else hi = mid;
    }
    return lo - from;
  }
  // Number of elements equal to x strictly before `upto` (x == at(upto-1) or 0).
  __device__ uint64_t run_bwd(uint64_t upto, K x) const {
    if (upto == 0 || !eq(at(upto - 1), x)) return 0;
    // find smallest i with at(i) == x, i < upto
    uint64_t hi = upto - 1, step = 1, lo;
    while (true) {  // invariant: at(hi) == x
      if (hi < step) { lo = 0; break; }
      lo = hi - step;
      if (!eq(at(lo), x)) { lo += 1; break; }
      hi = lo;
      step <<= 1;
    }
    // at(lo .. hi) == x region boundary lies in [lo, hi]; find first equal
    uint64_t l = lo, h = hi;  // at(h) == x
    while (l < h) {
      uint64_t mid = (l + h) >> 1;
      if (eq(at(mid), x)) h = mid;
      else l = mid + 1;
    }
    return upto - l;
  }
};

// Set-op tile: 512 threads x 12 merged items (6 for 16-byte keys); the
// one LDS array holds the tile's keys and then its kept items.  On 2^27 +
// 2^27 u64 sets this measured best among 256-1024 threads x 4-16 items.
template <class K>
struct SetTune {
  static constexpr int NT = 512;
  static constexpr int NV = sizeof(K) <= 8 ? 6144 : 3072;
};

// Fixed-step lower (LE = false) or upper (LE = true) bounds of x[j] in the
// sorted LDS array s[0, n), all of a thread's searches in lockstep.
template <class K, int VT, bool LE>
__device__ __forceinline__ void bounds(const K* s, uint32_t n, const K (&x)[VT], uint32_t jn,
                                       uint32_t (&pos)[VT]) {
  uint32_t top = n ? 1u << (31 - __builtin_clz(n)) : 0u;
#pragma unroll
  for (int j = 0; j < VT; ++j) pos[j] = 0;
  for (uint32_t step = top; step; step >>= 1) {
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      if (static_cast<uint32_t>(j) < jn && pos[j] + step <= n) {
        const K v = s[pos[j] + step - 1];
        if (LE ? !lt(x[j], v) : lt(v, x[j])) pos[j] += step;
      }
    }
  }
}

template <class K, int OP, int NT, int NV>
__global__ __launch_bounds__(NT) void k_set_op(
    const K* __restrict__ a, uint64_t na, const K* __restrict__ b, uint64_t nb,
    const uint64_t* __restrict__ split, uint32_t ntiles, K* __restrict__ out,
    uint64_t* __restrict__ n_out, uint64_t n_out_mul, uint64_t* __restrict__ status,
    uint32_t* __restrict__ ticket) {
  constexpr int VT = NV / NT;
  __shared__ K keys[NV];  // the tile's A keys, then its B keys; later the kept items
  __shared__ uint32_t scan_tmp[NT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  __shared__ K s_edge[4];  // a[a0 - 1], b[b0 - 1], b[b1], a[a1] where they exist
  __shared__ K s_wlast[NT / kWave], s_wfirst[NT / kWave];
  __shared__ bool s_wlast_a[NT / kWave];
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t d0 = static_cast<uint64_t>(tile) * NV;
  const uint64_t d1 = (d0 + NV < na + nb) ? d0 + NV : na + nb;
  const uint64_t a0 = split[tile], a1 = split[tile + 1];
  const uint64_t b0 = d0 - a0, b1 = d1 - a1;
  const uint32_t nat = static_cast<uint32_t>(a1 - a0), nbt = static_cast<uint32_t>(b1 - b0);
  const uint32_t nm = nat + nbt;
  // The neighbours just outside the tile, loaded with it: a run of equal
  // values crosses a tile edge only where one of them equals the element,
  // so the global gallops below start only then.  They reach LDS after the
  // tile's loads are issued, so that no wave waits on them alone.
  K edge = K{};
  if (tid == 0 && a0 > 0) edge = a[a0 - 1];
  if (tid == 1 && b0 > 0) edge = b[b0 - 1];
  if (tid == 2 && b1 < nb) edge = b[b1];
  if (tid == 3 && a1 < na) edge = a[a1];
  if constexpr (sizeof(K) <= 8) {
    // Each item's address is A's or B's, chosen without a branch, so every
    // load of a full tile is in flight before the first wait (a branch per
    // item made the compiler wait on the first load: 1.18 against 1.09 ms
    // on a 2^27 + 2^27 u64 union).
    K v[VT];
    const K* pa = a + a0;
    const K* pb = b + b0 - nat;
    if (nm == static_cast<uint32_t>(NV)) {
#pragma unroll
      for (int j = 0; j < VT; ++j) {
        const uint32_t i = tid + j * NT;
        v[j] = (i < nat ? pa : pb)[i];
      }
#pragma unroll
      for (int j = 0; j < VT; ++j) keys[tid + j * NT] = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < VT; ++j) {
        const uint32_t i = tid + j * NT;
        if (i < nm) v[j] = (i < nat ? pa : pb)[i];
      }
#pragma unroll
      for (int j = 0; j < VT; ++j) {
        const uint32_t i = tid + j * NT;
        if (i < nm) keys[i] = v[j];
      }
    }
  } else {
    for (uint32_t i = tid; i < nm; i += NT) keys[i] = i < nat ? a[a0 + i] : b[b0 + (i - nat)];
  }
  if (tid < 4) s_edge[tid] = edge;  // slots whose neighbour does not exist are never read
  __syncthreads();
  const K* sa = keys;
  const K* sb = keys + nat;
  const Side<K> GA{a, nullptr, na, 0, 0};  // global views for runs past the tile
  const Side<K> GB{b, nullptr, nb, 0, 0};

  // Thread t merges the tile's items [t*VT, t*VT + VT): one merge-path
  // search on its diagonal (A wins ties), then VT serial steps.  The
  // search takes fixed power-of-two steps from the largest below NV: the
  // first i in [lo, hi) with sb[diag-1-i] < sa[i], else hi.
  const uint32_t diag = tid * VT < nm ? tid * VT : nm;
  uint32_t ai0, bi0;
  {
    uint32_t pos = diag > nbt ? diag - nbt : 0u;
    const uint32_t hi = diag < nat ? diag : nat;
    constexpr uint32_t kTop = 1u << (31 - __builtin_clz(static_cast<uint32_t>(NV)));
#pragma unroll
    for (uint32_t step = kTop; step; step >>= 1) {
      if (pos + step <= hi) {
        const uint32_t m = pos + step - 1;
        if (!lt(sb[diag - 1 - m], sa[m])) pos += step;
      }
    }
    ai0 = pos;
    bi0 = diag - pos;
  }
  const uint32_t nk = nm - diag < static_cast<uint32_t>(VT) ? nm - diag : static_cast<uint32_t>(VT);

  // Fast path, for tiles in which no value has two copies in one input
  // (checked below): the merged items alone decide.  Equal items of A
  // and B are adjacent in merged order, A's first, so a B item is in A
  // exactly when the item before it equals it, and an A item is in B
  // exactly when the item after it does.
  // The head of each side stays in registers (one LDS read a step), and
  // all VT steps run: steps past nk read clamped slots, and their items
  // and amask bits are never used.
  K mv[VT];
  uint32_t amask = 0;
  {
    uint32_t ai = ai0, bi = bi0;
    K va = keys[ai < nat ? ai : 0u];
    K vb = keys[nat + bi < nm ? nat + bi : 0u];
#pragma unroll
    for (int k = 0; k < VT; ++k) {
      const bool take_a = bi >= nbt || (ai < nat && !lt(vb, va));
      mv[k] = take_a ? va : vb;
      amask |= static_cast<uint32_t>(take_a) << k;
      ai += take_a;
      bi += !take_a;
      const uint32_t nx = take_a ? ai : nat + bi;
      const K x = keys[nx < nm ? nx : 0u];
      va = take_a ? x : va;
      vb = take_a ? vb : x;
    }
  }
  // Neighbours across threads: the item before my first (and its source)
  // and the item after my last.
  K my_last = mv[0];
#pragma unroll
  for (int k = 1; k < VT; ++k)
    if (static_cast<uint32_t>(k) < nk) my_last = mv[k];
  const bool last_a = nk && ((amask >> (nk - 1)) & 1u);
  const uint32_t lane = lane_id(), w = wave_id();
  if (lane == kWave - 1) { s_wlast[w] = my_last; s_wlast_a[w] = last_a; }
  if (lane == 0) s_wfirst[w] = mv[0];
  bool dup = false;
  if (tid == 0) {
    // A value with copies on both sides of a tile edge.
    if (nat && a0 > 0 && eq(s_edge[0], sa[0])) dup = true;
    if (nbt && b0 > 0 && eq(s_edge[1], sb[0])) dup = true;
    if (nat && a1 < na && eq(s_edge[3], sa[nat - 1])) dup = true;
    if (nbt && b1 < nb && eq(s_edge[2], sb[nbt - 1])) dup = true;
  }
  __syncthreads();
  K prev = shfl_up_elem(my_last, 1);
  bool prev_a = __shfl_up(static_cast<int>(last_a), 1, kWave) != 0;
  K next = shfl_down_elem(mv[0], 1);
  if (lane == 0 && w > 0) { prev = s_wlast[w - 1]; prev_a = s_wlast_a[w - 1]; }
  if (lane == kWave - 1 && w + 1 < NT / kWave) next = s_wfirst[w + 1];
  const bool has_prev = diag > 0 || a0 > 0;        // an item before my first
  if (diag == 0 && a0 > 0) { prev = s_edge[0]; prev_a = true; }
  const bool last_item = diag + nk == nm;          // my last item ends the tile
  const bool has_next = !last_item || b1 < nb;
  if (last_item) next = s_edge[2];
  // Two equal items of one input are adjacent in merged order.
  if (nk) {
    if (diag > 0 && eq(prev, mv[0]) && prev_a == (amask & 1u)) dup = true;
#pragma unroll
    for (int k = 1; k < VT; ++k)
      if (static_cast<uint32_t>(k) < nk && eq(mv[k], mv[k - 1]) &&
          ((amask >> k) & 1u) == ((amask >> (k - 1)) & 1u))
        dup = true;
  }
  uint32_t kmask = 0;
  if (!__syncthreads_or(dup)) {
#pragma unroll
    for (int k = 0; k < VT; ++k) {
      if (static_cast<uint32_t>(k) >= nk) continue;
      const bool is_a = (amask >> k) & 1u;
      const K before = k == 0 ? prev : mv[k > 0 ? k - 1 : 0];
      const bool has_before = k > 0 || has_prev;
      const K after = k + 1 < static_cast<int>(nk) ? mv[k + 1 < VT ? k + 1 : k] : next;
      const bool has_after = k + 1 < static_cast<int>(nk) || has_next;
      bool keep;
      if constexpr (OP == RFX_SET_UNION) keep = is_a || !(has_before && eq(before, mv[k]));
      else if constexpr (OP == RFX_SET_INTERSECTION) keep = is_a && has_after && eq(after, mv[k]);
      else keep = is_a && !(has_after && eq(after, mv[k]));
      kmask |= static_cast<uint32_t>(keep) << k;
    }
  } else {
    // General path: multiset counts from equal runs, in LDS and, for runs
    // crossing a tile edge, in global memory.
    uint32_t ai = ai0, bi = bi0;
    for (int k = 0; k < VT; ++k) {
      if (static_cast<uint32_t>(k) >= nk) break;
      const bool take_a = bi >= nbt || (ai < nat && !lt(sb[bi], sa[ai]));
      bool keep;
      if (take_a) {
        const K x = sa[ai];
        keep = true;
        if constexpr (OP != RFX_SET_UNION) {
          // Copies of x in A before this one, and in B from bi on (equal
          // runs that cross the tile edge continue in global memory).
          const uint32_t before = lds_run_back(sa, ai, x);
          uint64_t rank = before;
          if (before == ai && a0 > 0 && eq(s_edge[0], x)) rank += GA.run_bwd(a0, x);
          const uint32_t in_b = lds_run_fwd(sb, bi, nbt, x);
          uint64_t cnt = in_b;
          if (bi + in_b == nbt && b1 < nb && eq(s_edge[2], x)) cnt += GB.run_fwd(b1, x);
          keep = (OP == RFX_SET_DIFFERENCE) ? rank >= cnt : rank < cnt;
        }
        ++ai;
      } else {
        const K y = sb[bi];
        keep = false;
        if constexpr (OP == RFX_SET_UNION) {
          // Copies of y in B before this one and in A up to ai (none of
          // A's copies lie at or past a1: they would be taken before y).
          const uint32_t before = lds_run_back(sb, bi, y);
          uint64_t rank = before;
          if (before == bi && b0 > 0 && eq(s_edge[1], y)) rank += GB.run_bwd(b0, y);
          const uint32_t in_a = lds_run_back(sa, ai, y);
          uint64_t cnt = in_a;
          if (in_a == ai && a0 > 0 && eq(s_edge[0], y)) cnt += GA.run_bwd(a0, y);
          keep = rank >= cnt;
        }
        ++bi;
      }
      kmask |= static_cast<uint32_t>(keep) << k;
    }
  }

  // Kept items in merged order: scan of the per-thread counts, then the
  // kept items packed into the key array (the scan's barriers come after
  // every thread's last key read).
  uint32_t tile_kept;
  uint32_t off = block_excl_scan<NT>(static_cast<uint32_t>(__popc(kmask)), scan_tmp, &tile_kept);
  // The tile's aggregate goes out before the packing, for the tiles after
  // it to read sooner.
  if (tid == 0 && tile != 0) status_store64(status + tile, kFlagAgg64 | tile_kept);
  K* mval = keys;
#pragma unroll
  for (int k = 0; k < VT; ++k)
    if ((kmask >> k) & 1u) mval[off++] = mv[k];
  if (tid < kWave) {
    if (tile == 0) {
      if (tid == 0) {
        status_store64(status, kFlagIncl64 | tile_kept);
        s_prefix = 0;
      }
    } else {
      const uint64_t excl = wave_lookback64(status, tile);
      if (tid == 0) {
        status_store64(status + tile, kFlagIncl64 | (excl + tile_kept));
        s_prefix = excl;
      }
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  for (uint32_t i = tid; i < tile_kept; i += NT) out[prefix + i] = mval[i];
  if (tile == ntiles - 1 && tid == 0) *n_out = (prefix + tile_kept) * n_out_mul;
}

// unique: keep element i when i == 0 or it differs from element i - 1.
template <class K>
__global__ __launch_bounds__(kSNT) void k_unique(const K* __restrict__ in, uint64_t n,
                                                 uint32_t ntiles, K* __restrict__ out,
                                                 uint64_t* __restrict__ n_out, uint64_t n_out_mul,
                                                 uint64_t* __restrict__ status,
                                                 uint32_t* __restrict__ ticket) {
  __shared__ K keys[kSNV + 1];
  __shared__ K kept[kSNV];
  __shared__ uint32_t scan_tmp[kSNT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t base = static_cast<uint64_t>(tile) * kSNV;
  const uint32_t cnt = static_cast<uint32_t>((n - base) < kSNV ? (n - base) : kSNV);
  // keys[0] = in[base - 1] (read only when base > 0), keys[1 + i] = in[base + i].
  // The loads are unconditional (slots past cnt read the tile's last row)
  // and all issued before the first wait.
  K v[kSVT];
#pragma unroll
  for (int s = 0; s < kSVT; ++s) {
    const uint32_t i = s * kSNT + tid;
    v[s] = in[base + (i < cnt ? i : cnt - 1)];
  }
  K before = K{};
  if (tid == 0 && base > 0) before = in[base - 1];
#pragma unroll
  for (int s = 0; s < kSVT; ++s) {
    const uint32_t i = s * kSNT + tid;
    if (i < cnt) keys[1 + i] = v[s];
  }
  if (tid == 0) keys[0] = before;
  __syncthreads();
  uint32_t keepmask = 0;
#pragma unroll
  for (int s = 0; s < kSVT; ++s) {
    const uint32_t i = tid * kSVT + s;
    if (i >= cnt) break;
    const bool keep = (base + i == 0) || !eq(keys[i], keys[i + 1]);
    keepmask |= static_cast<uint32_t>(keep) << s;
  }
  uint32_t tile_kept;
  uint32_t my_off = block_excl_scan<kSNT>(static_cast<uint32_t>(__popc(keepmask)), scan_tmp,
                                          &tile_kept);
#pragma unroll
  for (int s = 0; s < kSVT; ++s)
    if ((keepmask >> s) & 1u) kept[my_off++] = keys[tid * kSVT + s + 1];
  if (tid < kWave) {
    if (tile == 0) {
      if (tid == 0) {
        status_store64(status, kFlagIncl64 | tile_kept);
        s_prefix = 0;
      }
    } else {
      if (tid == 0) status_store64(status + tile, kFlagAgg64 | tile_kept);
      const uint64_t excl = wave_lookback64(status, tile);
      if (tid == 0) {
        status_store64(status + tile, kFlagIncl64 | (excl + tile_kept));
        s_prefix = excl;
      }
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  for (uint32_t i = tid; i < tile_kept; i += kSNT) out[prefix + i] = kept[i];
  if (tile == ntiles - 1 && tid == 0) *n_out = (prefix + tile_kept) * n_out_mul;
}

struct LbWork {
  uint64_t* split;
  uint64_t* status;
  uint32_t* ticket;
};

int lb_workspace(rfx_ctx* ctx, uint64_t ntiles, uint64_t extra_bytes, LbWork* w, void** extra) {
  Carve c(nullptr);
  c.take<uint64_t>(ntiles + 1);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  c.take<char>(extra_bytes);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve v(ctx->ws);
  w->split = v.take<uint64_t>(ntiles + 1);
  w->status = v.take<uint64_t>(ntiles);
  w->ticket = v.take<uint32_t>(1);
  if (extra) *extra = v.take<char>(extra_bytes);
  return fill_u32(ctx, reinterpret_cast<uint32_t*>(w->status),
                  (reinterpret_cast<char*>(w->ticket + 1) - reinterpret_cast<char*>(w->status)) / 4,
                  0);
}

template <class K>
int set_op_impl(rfx_ctx* ctx, int op, void* out, uint64_t* n_out, uint64_t mul, const void* a,
                uint64_t na, const void* b, uint64_t nb) {
  if (na + nb == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(n_out), 2, 0);
  constexpr int NT = SetTune<K>::NT, NV = SetTune<K>::NV;
  const uint64_t ntiles = (na + nb + NV - 1) / NV;
  LbWork w;
  RFX_TRY(lb_workspace(ctx, ntiles, 0, &w, nullptr));
  const K* A = static_cast<const K*>(a);
  const K* B = static_cast<const K*>(b);
  k_set_partition<K><<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      A, na, B, nb, ntiles, NV, w.split);
  RFX_LAUNCHED();
  const uint32_t g = static_cast<uint32_t>(ntiles);
  K* O = static_cast<K*>(out);
  RFX_PROF(ctx, "setop");
  switch (op) {
    case RFX_SET_UNION:
      k_set_op<K, RFX_SET_UNION, NT, NV><<<g, NT, 0, ctx->stream>>>(A, na, B, nb, w.split, g, O, n_out,
                                                             mul, w.status, w.ticket);
      break;
    case RFX_SET_DIFFERENCE:
      k_set_op<K, RFX_SET_DIFFERENCE, NT, NV><<<g, NT, 0, ctx->stream>>>(A, na, B, nb, w.split, g, O,
                                                                  n_out, mul, w.status, w.ticket);
      break;
    default:
      k_set_op<K, RFX_SET_INTERSECTION, NT, NV><<<g, NT, 0, ctx->stream>>>(
          A, na, B, nb, w.split, g, O, n_out, mul, w.status, w.ticket);
      break;
  }
  RFX_LAUNCHED();
  return RFX_OK;
}

template <class K>
int unique_impl(rfx_ctx* ctx, void* out, uint64_t* n_out, uint64_t mul, const void* in,
                uint64_t n) {
  if (n == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(n_out), 2, 0);
  const uint64_t ntiles = (n + kSNV - 1) / kSNV;
  const bool in_place = out == in;
  LbWork w;
  void* tmp = nullptr;
  RFX_TRY(lb_workspace(ctx, ntiles, in_place ? n * sizeof(K) : 0, &w, &tmp));
  K* dst = static_cast<K*>(in_place ? tmp : out);
  RFX_PROF(ctx, "unique");
  k_unique<K><<<static_cast<uint32_t>(ntiles), kSNT, 0, ctx->stream>>>(
      static_cast<const K*>(in), n, static_cast<uint32_t>(ntiles), dst, n_out, mul, w.status,
      w.ticket);
  RFX_LAUNCHED();
  if (in_place)
    RFX_HIP(hipMemcpyAsync(out, tmp, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
  return RFX_OK;
}

}  // namespace

int set_op_sized(rfx_ctx* ctx, int op, void* out, uint64_t* n_out, uint64_t mul, const void* a,
                 uint64_t na, const void* b, uint64_t nb, int eb) {
  if (!ctx || !n_out) return fail(RFX_EINVAL, "set op: null argument");
  if (op < RFX_SET_UNION || op > RFX_SET_INTERSECTION) return fail(RFX_EINVAL, "set op %d", op);
  if ((na && !a) || (nb && !b) || (na + nb && !out)) return fail(RFX_EINVAL, "set op: null column");
  if (na + nb >= (1ull << 40)) return fail(RFX_ERANGE, "set op: too many elements");
  switch (eb) {
    case 1: return set_op_impl<uint8_t>(ctx, op, out, n_out, mul, a, na, b, nb);
    case 2: return set_op_impl<uint16_t>(ctx, op, out, n_out, mul, a, na, b, nb);
    case 4: return set_op_impl<uint32_t>(ctx, op, out, n_out, mul, a, na, b, nb);
    case 8: return set_op_impl<uint64_t>(ctx, op, out, n_out, mul, a, na, b, nb);
    case 16: return set_op_impl<u128>(ctx, op, out, n_out, mul, a, na, b, nb);
  }
  return fail(RFX_EINVAL, "set op: element width %d", eb);
}

int unique_sized(rfx_ctx* ctx, void* out, uint64_t* n_out, uint64_t mul, const void* in,
                 uint64_t n, int eb) {
  if (!ctx || !n_out || (n && (!in || !out))) return fail(RFX_EINVAL, "unique: null argument");
  switch (eb) {
    case 1: return unique_impl<uint8_t>(ctx, out, n_out, mul, in, n);
    case 2: return unique_impl<uint16_t>(ctx, out, n_out, mul, in, n);
    case 4: return unique_impl<uint32_t>(ctx, out, n_out, mul, in, n);
    case 8: return unique_impl<uint64_t>(ctx, out, n_out, mul, in, n);
    case 16: return unique_impl<u128>(ctx, out, n_out, mul, in, n);
  }
  return fail(RFX_EINVAL, "unique: element width %d", eb);
}

}  // namespace rfx

using namespace rfx;

extern "C" int rfx_set_op(rfx_ctx* ctx, int op, void* out, uint64_t* n_out, const void* a,
                          uint64_t na, const void* b, uint64_t nb, int elem_bytes) {
  return set_op_sized(ctx, op, out, n_out, 1, a, na, b, nb, elem_bytes);
}

extern "C" int rfx_unique(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in, uint64_t n,
                          int elem_bytes) {
  return unique_sized(ctx, out, n_out, 1, in, n, elem_bytes);
}
