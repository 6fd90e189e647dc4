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
// LDS; every element decides from its rank among equal values in its own
// input and the count of equal values in the other input whether it is
// kept (union keeps all of A and the copies of B beyond A's count;
// difference keeps the copies of A beyond B's count; intersection the
// copies of A within B's count); kept elements are compacted in merged
// order, the tile offset coming from a decoupled look-back.
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

template <class K>
__global__ __launch_bounds__(256) void k_set_partition(const K* __restrict__ a, uint64_t na,
                                                       const K* __restrict__ b, uint64_t nb,
                                                       uint64_t ntiles,
                                                       uint64_t* __restrict__ split) {
  uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t > ntiles) return;
  uint64_t diag = t * kSNV;
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

template <class K, int OP>
__global__ __launch_bounds__(kSNT) void k_set_op(
    const K* __restrict__ a, uint64_t na, const K* __restrict__ b, uint64_t nb,
    const uint64_t* __restrict__ split, uint32_t ntiles, K* __restrict__ out,
    uint64_t* __restrict__ n_out, uint64_t n_out_mul, uint64_t* __restrict__ status,
    uint32_t* __restrict__ ticket) {
  __shared__ K keys[kSNV];
  __shared__ K kept[kSNV];
  __shared__ uint32_t scan_tmp[kSNT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t d0 = static_cast<uint64_t>(tile) * kSNV;
  const uint64_t d1 = (d0 + kSNV < na + nb) ? d0 + kSNV : na + nb;
  const uint64_t a0 = split[tile], a1 = split[tile + 1];
  const uint64_t b0 = d0 - a0, b1 = d1 - a1;
  const uint32_t nat = static_cast<uint32_t>(a1 - a0), nbt = static_cast<uint32_t>(b1 - b0);
  for (uint32_t i = tid; i < nat + nbt; i += kSNT)
    keys[i] = i < nat ? a[a0 + i] : b[b0 + (i - nat)];
  __syncthreads();
  const Side<K> A{a, keys, na, a0, a1};
  const Side<K> B{b, keys + nat, nb, b0, b1};

  const uint32_t diag = tid * kSVT;
  uint32_t ai = 0, bj = 0;
  if (diag < nat + nbt) {
    ai = static_cast<uint32_t>(mp_lower(keys, nat, keys + nat, nbt, diag));
    bj = diag - ai;
  }
  K elem[kSVT];
  uint32_t keepmask = 0;
#pragma unroll
  for (int s = 0; s < kSVT; ++s) {
    if (diag + s >= nat + nbt) break;
    const bool take_a = ai < nat && (bj >= nbt || !lt(keys[nat + bj], keys[ai]));
    if (take_a) {
      const K x = keys[ai];
      const uint64_t gi = a0 + ai;
      bool keep = true;
      if (OP != RFX_SET_UNION) {
        const uint64_t rank = A.run_bwd(gi, x);         // equal copies before it in A
        const uint64_t cnt_b = B.run_fwd(b0 + bj, x);   // B's copies (all >= position)
        keep = (OP == RFX_SET_DIFFERENCE) ? rank >= cnt_b : rank < cnt_b;
      }
      elem[s] = x;
      keepmask |= static_cast<uint32_t>(keep) << s;
      ++ai;
    } else {
      const K y = keys[nat + bj];
      const uint64_t gj = b0 + bj;
      if (OP == RFX_SET_UNION) {
        const uint64_t rank = B.run_bwd(gj, y);
        const uint64_t cnt_a = A.run_bwd(a0 + ai, y);  // A's copies precede it (A wins ties)
        elem[s] = y;
        keepmask |= static_cast<uint32_t>(rank >= cnt_a) << s;
      }
      ++bj;
    }
  }
  uint32_t tile_kept;
  uint32_t my_off = block_excl_scan<kSNT>(static_cast<uint32_t>(__popc(keepmask)), scan_tmp,
                                          &tile_kept);
#pragma unroll
  for (int s = 0; s < kSVT; ++s)
    if ((keepmask >> s) & 1u) kept[my_off++] = elem[s];
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
  const uint32_t tile = s_tile;
  const uint64_t base = static_cast<uint64_t>(tile) * kSNV;
  const uint32_t cnt = static_cast<uint32_t>((n - base) < kSNV ? (n - base) : kSNV);
  for (uint32_t i = tid; i < cnt + 1; i += kSNT) {
    if (i == 0) {
      if (base > 0) keys[0] = in[base - 1];
    } else {
      keys[i] = in[base + i - 1];
    }
  }
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
  const uint64_t ntiles = (na + nb + kSNV - 1) / kSNV;
  LbWork w;
  RFX_TRY(lb_workspace(ctx, ntiles, 0, &w, nullptr));
  const K* A = static_cast<const K*>(a);
  const K* B = static_cast<const K*>(b);
  k_set_partition<K><<<static_cast<uint32_t>((ntiles + 1 + 255) / 256), 256, 0, ctx->stream>>>(
      A, na, B, nb, ntiles, w.split);
  RFX_LAUNCHED();
  const uint32_t g = static_cast<uint32_t>(ntiles);
  K* O = static_cast<K*>(out);
  RFX_PROF(ctx, "setop");
  switch (op) {
    case RFX_SET_UNION:
      k_set_op<K, RFX_SET_UNION><<<g, kSNT, 0, ctx->stream>>>(A, na, B, nb, w.split, g, O, n_out,
                                                             mul, w.status, w.ticket);
      break;
    case RFX_SET_DIFFERENCE:
      k_set_op<K, RFX_SET_DIFFERENCE><<<g, kSNT, 0, ctx->stream>>>(A, na, B, nb, w.split, g, O,
                                                                  n_out, mul, w.status, w.ticket);
      break;
    default:
      k_set_op<K, RFX_SET_INTERSECTION><<<g, kSNT, 0, ctx->stream>>>(
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
