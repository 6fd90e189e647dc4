// select.hip - stable filter, field permutation, split and merge of packed
// tuples for gfx950.
//
// select replaces the reference's three-launch form (select_field_constant
// per-CTA compaction into a temp buffer with a CTA histogram, Select.h:30-114;
// a 1-CTA histogram scan, Scan.h:18; getResultSize + gather,
// Select.h:1013-1130) with one pass that streams rows through registers:
// each wave ballots the predicate over 64 consecutive rows at a time, the
// per-(wave, step) counts are scanned in LDS, the tile takes its output
// offset from a decoupled look-back, and each kept row is stored straight
// to its place (its rank inside the ballot by mbcnt).  Rows are read once
// and written once, in input order.
// project, split and merge are the streaming kernels of Project.h:58-130
// and Split_Merge.h:303-531.
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

#ifndef RFX_SEL_NT
#define RFX_SEL_NT 512
#endif
#ifndef RFX_SEL_VT16
#define RFX_SEL_VT16 24
#endif
constexpr int kFNT = RFX_SEL_NT;
constexpr int kFNW = kFNT / kWave;
// Rows per thread by default, from the autotuner's sweep over the same
// 2 GiB of tuples with half kept (ms for 8 / 16 / 24 / 32 rows): 16-byte
// tuples 0.74 / 0.68 / 0.63 / 0.74, 8-byte 1.06 / 0.84 / 0.72 / 0.78,
// 4-byte 1.71 / 1.24 / 1.05 / 0.96.  The select.vt knob picks 8, 16, 24 or
// 32 instead.
template <class K>
constexpr int kFVT = sizeof(K) >= 16 ? RFX_SEL_VT16 : sizeof(K) == 8 ? 24 : 32;

template <class K>
__device__ __forceinline__ u128 widen(K v) { return to128(v); }

__device__ __forceinline__ int64_t sext(uint64_t v, uint32_t bits) {
  if (bits == 0 || bits >= 64) return static_cast<int64_t>(v);
  const uint64_t m = 1ull << (bits - 1);
  return static_cast<int64_t>((v ^ m) - m);
}

// Comparison functors of Comparisons.h:115-300; the double forms keep the
// reference's 1e-6 margin.
__device__ __forceinline__ bool compare(uint64_t f, uint32_t bits, int op, uint64_t k) {
  switch (op) {
    case RFX_CMP_EQ: return f == k;
    case RFX_CMP_NE: return f != k;
    case RFX_CMP_LT: return f < k;
    case RFX_CMP_LE: return f <= k;
    case RFX_CMP_GT: return f > k;
    case RFX_CMP_GE: return f >= k;
    case RFX_CMP_SLT: return sext(f, bits) < sext(k, bits);
    case RFX_CMP_SLE: return sext(f, bits) <= sext(k, bits);
    case RFX_CMP_SGT: return sext(f, bits) > sext(k, bits);
    case RFX_CMP_SGE: return sext(f, bits) >= sext(k, bits);
    default: break;
  }
  const double x = __longlong_as_double(static_cast<long long>(f));
  const double y = __longlong_as_double(static_cast<long long>(k));
  switch (op) {
    case RFX_CMP_DLT: return (y - x) > 1e-6;
    case RFX_CMP_DLE: return (y - x) > -1e-6;
    case RFX_CMP_DGT: return (x - y) > 1e-6;
    case RFX_CMP_DGE: return (x - y) > -1e-6;
    default: return false;
  }
}

struct Pred {
  uint32_t shift, bits;
  int op;
  uint64_t k;
};

template <class K, int VT>
__global__ __launch_bounds__(kFNT) void k_select(const K* __restrict__ in, uint64_t n,
                                                 uint32_t ntiles, Pred p, K* __restrict__ out,
                                                 uint64_t* __restrict__ n_out,
                                                 uint64_t* __restrict__ status,
                                                 uint32_t* __restrict__ ticket) {
  constexpr int TILE = kFNT * VT;
  constexpr int kE = VT * kFNW, P = kE > kWave ? kE / kWave : 1;
  static_assert(kE <= kWave || kE % kWave == 0, "whole runs of step counts per lane");
  __shared__ uint32_t s_cnt[VT * kFNW + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t base = static_cast<uint64_t>(tile) * TILE;
  // Step j of wave w covers rows base + j*kFNT + w*64 .. +63.
  K r[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint64_t i = base + static_cast<uint64_t>(j) * kFNT + tid;
    if (i < n) r[j] = in[i];
  }
  uint32_t keep = 0;
  uint64_t m[VT];
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    const uint64_t i = base + static_cast<uint64_t>(j) * kFNT + tid;
    bool k = false;
    if (i < n) {
      const uint64_t f = mask128(shr128(widen(r[j]), p.shift), p.bits).lo;
      k = compare(f, p.bits, p.op, p.k);
    }
    keep |= static_cast<uint32_t>(k) << j;
    m[j] = ballot(k);
    if (lane == 0) s_cnt[j * kFNW + w] = static_cast<uint32_t>(__popcll(m[j]));
  }
  __syncthreads();
  // Exclusive scan of the (step, wave) counts in row order (wave 0).
  if (w == 0) {
    // Lane l scans entries [l*P, l*P + P) (entry j*kFNW + w: row order).
    uint32_t c[P], sum = 0;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const uint32_t e = lane * P + q;
      c[q] = e < kE ? s_cnt[e] : 0u;
      sum += c[q];
    }
    const uint32_t incl = wave_incl_scan(sum);
    uint32_t run = incl - sum;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const uint32_t e = lane * P + q;
      if (e < kE) s_cnt[e] = run;
      run += c[q];
    }
    const uint32_t tile_kept = __shfl(incl, kWave - 1, kWave);
    if (lane == 0) s_cnt[kE] = tile_kept;
    if (tile == 0) {
      if (lane == 0) {
        status_store64(status, kFlagIncl64 | tile_kept);
        s_prefix = 0;
      }
    } else {
      if (lane == 0) status_store64(status + tile, kFlagAgg64 | tile_kept);
      const uint64_t excl = wave_lookback64(status, tile);
      if (lane == 0) {
        status_store64(status + tile, kFlagIncl64 | (excl + tile_kept));
        s_prefix = excl;
      }
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
#pragma unroll
  for (int j = 0; j < VT; ++j)
    if ((keep >> j) & 1u) out[prefix + s_cnt[j * kFNW + w] + mbcnt(m[j])] = r[j];
  if (tile == ntiles - 1 && tid == 0) *n_out = prefix + s_cnt[kE];
}

__device__ __forceinline__ u128 load_any(const void* base, uint64_t i, int w) {
  switch (w) {
    case 1: return {static_cast<const uint8_t*>(base)[i], 0};
    case 2: return {static_cast<const uint16_t*>(base)[i], 0};
    case 4: return {static_cast<const uint32_t*>(base)[i], 0};
    case 8: return {static_cast<const uint64_t*>(base)[i], 0};
    default: return static_cast<const u128*>(base)[i];
  }
}
__device__ __forceinline__ void store_any(void* base, uint64_t i, int w, u128 v) {
  switch (w) {
    case 1: static_cast<uint8_t*>(base)[i] = static_cast<uint8_t>(v.lo); break;
    case 2: static_cast<uint16_t*>(base)[i] = static_cast<uint16_t>(v.lo); break;
    case 4: static_cast<uint32_t*>(base)[i] = static_cast<uint32_t>(v.lo); break;
    case 8: static_cast<uint64_t*>(base)[i] = v.lo; break;
    default: static_cast<u128*>(base)[i] = v; break;
  }
}

struct ProjSpec {
  int nf;
  uint32_t in_shift[16], bits[16], out_shift[16];
};

// n_dev (may be null): a device row count; the rows processed are then
// min(*n_dev, n), n being the capacity the grid was sized for.
__global__ __launch_bounds__(256) void k_project(void* out, int ob, const void* in, int ib,
                                                 uint64_t n, const uint64_t* n_dev, ProjSpec ps) {
  if (n_dev && *n_dev < n) n = *n_dev;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const u128 t = load_any(in, i, ib);
    u128 o{0, 0};
    for (int f = 0; f < ps.nf; ++f)
      o = or128(o, shl128(mask128(shr128(t, ps.in_shift[f]), ps.bits[f]), ps.out_shift[f]));
    store_any(out, i, ob, mask128(o, 8u * ob));
  }
}

__global__ __launch_bounds__(256) void k_split(void* key, int kb, void* val, int vb,
                                               const void* in, int ib, uint64_t n,
                                               uint32_t shift) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const u128 t = load_any(in, i, ib);
    store_any(key, i, kb, mask128(shr128(t, shift), 8u * kb));
    if (val) store_any(val, i, vb, mask128(mask128(t, shift), 8u * vb));
  }
}

__global__ __launch_bounds__(256) void k_merge(void* out, int ob, const void* key, int kb,
                                               const void* val, int vb, uint64_t n,
                                               uint32_t shift) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    u128 o = shl128(load_any(key, i, kb), shift);
    o = or128(o, load_any(val, i, vb));
    store_any(out, i, ob, mask128(o, 8u * ob));
  }
}

uint32_t stream_grid(rfx_ctx* ctx, uint64_t n) {
  uint64_t b = (n + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(ctx->device_cus) * 16;
  if (b > cap) b = cap;
  return static_cast<uint32_t>(b ? b : 1);
}

template <class K, int VT>
int select_vt(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in, uint64_t n, Pred p) {
  constexpr uint64_t TILE = static_cast<uint64_t>(kFNT) * VT;
  const uint64_t ntiles = (n + TILE - 1) / TILE;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* status = w.take<uint64_t>(ntiles);
  uint32_t* ticket = w.take<uint32_t>(1);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  RFX_PROF(ctx, "select");
  k_select<K, VT><<<static_cast<uint32_t>(ntiles), kFNT, 0, ctx->stream>>>(
      static_cast<const K*>(in), n, static_cast<uint32_t>(ntiles), p, static_cast<K*>(out), n_out,
      status, ticket);
  RFX_LAUNCHED();
  return RFX_OK;
}

// Rows per thread: the built-in kFVT<K>, or the context's select.vt knob.
template <class K>
int select_impl(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in, uint64_t n, Pred p) {
  if (n == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(n_out), 2, 0);
  switch (ctx->tune_sel_vt) {
    case 8: return select_vt<K, 8>(ctx, out, n_out, in, n, p);
    case 16: return select_vt<K, 16>(ctx, out, n_out, in, n, p);
    case 24: return select_vt<K, 24>(ctx, out, n_out, in, n, p);
    case 32: return select_vt<K, 32>(ctx, out, n_out, in, n, p);
  }
  return select_vt<K, kFVT<K>>(ctx, out, n_out, in, n, p);
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_select(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in, uint64_t n,
                          int tuple_bytes, uint32_t field_shift, uint32_t field_bits, int op,
                          uint64_t constant) {
  if (!ctx || !n_out || (n && (!in || !out))) return fail(RFX_EINVAL, "rfx_select: null argument");
  if (op < RFX_CMP_EQ || op > RFX_CMP_DGE) return fail(RFX_EINVAL, "rfx_select: op %d", op);
  if (field_bits == 0 || field_bits > 64 || field_shift + field_bits > 8u * tuple_bytes)
    return fail(RFX_EINVAL, "rfx_select: field [%u, +%u) outside a %d-byte tuple", field_shift,
                field_bits, tuple_bytes);
  if (op >= RFX_CMP_DLT && field_bits != 64)
    return fail(RFX_EINVAL, "rfx_select: double comparison needs a 64-bit field");
  if (n >= (1ull << 40)) return fail(RFX_ERANGE, "rfx_select: too many rows");
  Pred p{field_shift, field_bits, op, constant};
  switch (tuple_bytes) {
    case 1: return select_impl<uint8_t>(ctx, out, n_out, in, n, p);
    case 2: return select_impl<uint16_t>(ctx, out, n_out, in, n, p);
    case 4: return select_impl<uint32_t>(ctx, out, n_out, in, n, p);
    case 8: return select_impl<uint64_t>(ctx, out, n_out, in, n, p);
    case 16: return select_impl<u128>(ctx, out, n_out, in, n, p);
  }
  return fail(RFX_EINVAL, "rfx_select: tuple width %d", tuple_bytes);
}

namespace {

int project_impl(rfx_ctx* ctx, void* out, int out_bytes, const void* in, int in_bytes, uint64_t n,
                 const uint64_t* n_dev, int nfields, const uint32_t* in_shift,
                 const uint32_t* bits, const uint32_t* out_shift) {
  if (!ctx || (n && (!out || !in))) return fail(RFX_EINVAL, "rfx_project: null argument");
  if (!valid_width(out_bytes) || !valid_width(in_bytes))
    return fail(RFX_EINVAL, "rfx_project: widths %d/%d", out_bytes, in_bytes);
  if (nfields < 1 || nfields > 16 || !in_shift || !bits || !out_shift)
    return fail(RFX_EINVAL, "rfx_project: %d fields", nfields);
  ProjSpec ps{};
  ps.nf = nfields;
  for (int f = 0; f < nfields; ++f) {
    if (bits[f] == 0 || bits[f] > 128 || in_shift[f] + bits[f] > 8u * in_bytes ||
        out_shift[f] + bits[f] > 8u * out_bytes)
      return fail(RFX_EINVAL, "rfx_project: field %d out of range", f);
    ps.in_shift[f] = in_shift[f];
    ps.bits[f] = bits[f];
    ps.out_shift[f] = out_shift[f];
  }
  if (!n) return RFX_OK;
  k_project<<<stream_grid(ctx, n), 256, 0, ctx->stream>>>(out, out_bytes, in, in_bytes, n, n_dev,
                                                          ps);
  RFX_LAUNCHED();
  return RFX_OK;
}

}  // namespace

extern "C" int rfx_project(rfx_ctx* ctx, void* out, int out_bytes, const void* in, int in_bytes,
                           uint64_t n, int nfields, const uint32_t* in_shift,
                           const uint32_t* bits, const uint32_t* out_shift) {
  return project_impl(ctx, out, out_bytes, in, in_bytes, n, nullptr, nfields, in_shift, bits,
                      out_shift);
}

extern "C" int rfx_project_sized(rfx_ctx* ctx, void* out, int out_bytes, const void* in,
                                 int in_bytes, const uint64_t* n, uint64_t n_cap, int nfields,
                                 const uint32_t* in_shift, const uint32_t* bits,
                                 const uint32_t* out_shift) {
  if (n_cap && !n) return fail(RFX_EINVAL, "rfx_project_sized: null row count");
  return project_impl(ctx, out, out_bytes, in, in_bytes, n_cap, n, nfields, in_shift, bits,
                      out_shift);
}

extern "C" int rfx_split(rfx_ctx* ctx, void* key, int key_bytes, void* val, int val_bytes,
                         const void* in, int in_bytes, uint64_t n, uint32_t key_shift) {
  if (!ctx || (n && (!key || !in))) return fail(RFX_EINVAL, "rfx_split: null argument");
  if (!valid_width(key_bytes) || !valid_width(in_bytes) || (val && !valid_width(val_bytes)))
    return fail(RFX_EINVAL, "rfx_split: widths");
  if (key_shift >= 8u * in_bytes) return fail(RFX_EINVAL, "rfx_split: shift %u", key_shift);
  if (!n) return RFX_OK;
  k_split<<<stream_grid(ctx, n), 256, 0, ctx->stream>>>(key, key_bytes, val, val_bytes, in,
                                                         in_bytes, n, key_shift);
  RFX_LAUNCHED();
  return RFX_OK;
}

extern "C" int rfx_merge(rfx_ctx* ctx, void* out, int out_bytes, const void* key, int key_bytes,
                         const void* val, int val_bytes, uint64_t n, uint32_t key_shift) {
  if (!ctx || (n && (!out || !key || !val))) return fail(RFX_EINVAL, "rfx_merge: null argument");
  if (!valid_width(out_bytes) || !valid_width(key_bytes) || !valid_width(val_bytes))
    return fail(RFX_EINVAL, "rfx_merge: widths");
  if (key_shift >= 8u * out_bytes) return fail(RFX_EINVAL, "rfx_merge: shift %u", key_shift);
  if (!n) return RFX_OK;
  k_merge<<<stream_grid(ctx, n), 256, 0, ctx->stream>>>(out, out_bytes, key, key_bytes, val,
                                                         val_bytes, n, key_shift);
  RFX_LAUNCHED();
  return RFX_OK;
}
