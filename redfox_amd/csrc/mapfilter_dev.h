// mapfilter_dev.h - MAPFILTER device code shared by the interpreter kernel
// (mapfilter.hip) and the kernels compiled at run time for one program:
// the semantics of every op, and the two-phase tile (filter, ballot
// compaction with look-back, outputs for the kept rows).
//
// Arithmetic: integers in two's-complement 64 bits, C division and
// remainder (truncating; a zero divisor gives 0); doubles in IEEE binary64
// with no contraction into fused multiply-adds.  The date operations
// restate Date.h:17-118 (seconds since 1970 for the TPC-H years,
// wrapdate's month arithmetic included).
#pragma once

#include "rfx_device.h"
#include "rfx_ra.h"

#pragma clang fp contract(off)

namespace rfx {

constexpr int kMNT = 512;
constexpr int kMNW = kMNT / kWave;
constexpr int kMVT = 16;
constexpr int kMTile = kMNT * kMVT;

__device__ __forceinline__ double as_f(uint64_t v) {
  return __longlong_as_double(static_cast<long long>(v));
}
__device__ __forceinline__ uint64_t as_u(double d) {
  return static_cast<uint64_t>(__double_as_longlong(d));
}

// Element i of a column of type t at p: integers widened (zero- or sign-),
// floats to double bits.  Packed tuple columns yield the field [shift,
// shift + bits) named by imm (bits 0-7 shift, 8-15 bits, 16-17: 0
// unsigned, 1 signed, 2 double).
__device__ __forceinline__ uint64_t load_elem(const void* p, int t, uint64_t i, uint64_t imm) {
  switch (t) {
    case RFX_MF_U8: return static_cast<const uint8_t*>(p)[i];
    case RFX_MF_U16: return static_cast<const uint16_t*>(p)[i];
    case RFX_MF_U32: return static_cast<const uint32_t*>(p)[i];
    case RFX_MF_U64: return static_cast<const uint64_t*>(p)[i];
    case RFX_MF_I8: return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int8_t*>(p)[i]));
    case RFX_MF_I16: return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int16_t*>(p)[i]));
    case RFX_MF_I32: return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(p)[i]));
    case RFX_MF_I64: return static_cast<const uint64_t*>(p)[i];
    case RFX_MF_F32: return as_u(static_cast<double>(static_cast<const float*>(p)[i]));
    case RFX_MF_F64: return static_cast<const uint64_t*>(p)[i];
    default: break;
  }
  u128 x;
  switch (t) {
    case RFX_MF_TUPLE1: x = {static_cast<const uint8_t*>(p)[i], 0}; break;
    case RFX_MF_TUPLE2: x = {static_cast<const uint16_t*>(p)[i], 0}; break;
    case RFX_MF_TUPLE4: x = {static_cast<const uint32_t*>(p)[i], 0}; break;
    case RFX_MF_TUPLE8: x = {static_cast<const uint64_t*>(p)[i], 0}; break;
    default: x = static_cast<const u128*>(p)[i]; break;
  }
  const uint32_t shift = imm & 0xFF, bits = (imm >> 8) & 0xFF, kind = (imm >> 16) & 3;
  const uint64_t f = mask128(shr128(x, shift), bits).lo;
  if (kind == 1 && bits < 64) {
    const uint64_t m = 1ull << (bits - 1);
    return (f ^ m) - m;
  }
  return f;  // unsigned, or the bits of a double when kind == 2
}

// LDX: element i of a column of `rows` rows, or 0 past its end.
__device__ __forceinline__ uint64_t load_at(const void* p, int t, uint64_t rows, uint64_t i,
                                            uint64_t imm) {
  return i < rows ? load_elem(p, t, i, imm) : 0;
}

// Output element: integers keep their low bits, F32 rounds the double.
__device__ __forceinline__ void store_out(void* p, int t, uint64_t pos, uint64_t v) {
  switch (t) {
    case RFX_MF_U8: case RFX_MF_I8: static_cast<uint8_t*>(p)[pos] = static_cast<uint8_t>(v); break;
    case RFX_MF_U16: case RFX_MF_I16: static_cast<uint16_t*>(p)[pos] = static_cast<uint16_t>(v); break;
    case RFX_MF_U32: case RFX_MF_I32: static_cast<uint32_t*>(p)[pos] = static_cast<uint32_t>(v); break;
    case RFX_MF_F32: static_cast<float*>(p)[pos] = static_cast<float>(as_f(v)); break;
    default: static_cast<uint64_t*>(p)[pos] = v; break;  // U64, I64, F64
  }
}

__device__ __forceinline__ uint64_t mf_idiv(uint64_t a, uint64_t b) {
  const int64_t sa = static_cast<int64_t>(a), sb = static_cast<int64_t>(b);
  return sb == 0 ? 0 : (sb == -1 ? 0 - a : static_cast<uint64_t>(sa / sb));
}
__device__ __forceinline__ uint64_t mf_imod(uint64_t a, uint64_t b) {
  const int64_t sa = static_cast<int64_t>(a), sb = static_cast<int64_t>(b);
  return (sb == 0 || sb == -1) ? 0 : static_cast<uint64_t>(sa % sb);
}
__device__ __forceinline__ uint64_t mf_f2i(uint64_t a) {
  const double d = as_f(a);
  return !(d == d) ? 0
         : d >= 9223372036854775807.0 ? static_cast<uint64_t>(INT64_MAX)
         : d <= -9223372036854775808.0 ? static_cast<uint64_t>(INT64_MIN)
                                       : static_cast<uint64_t>(static_cast<int64_t>(d));
}

// Date.h: seconds since 1970-01-01 <-> (year, month, day) for 1992-1998.
__device__ inline void int2date(uint64_t seconds, int& year, int& month, int& day) {
  uint64_t days = seconds / 86400 - 8035;  // days since 1992-01-01
  int y = 1992;
  // Year lengths from 1992 on; the walk stops at the first that does not
  // fit (six steps at most, as Date.h:37-88 does).
  const uint32_t len[6] = {366, 365, 365, 365, 366, 365};
  for (int k = 0; k < 6 && days >= len[k]; ++k) {
    days -= len[k];
    ++y;
  }
  uint32_t start[13] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334, 365};
  if (y == 1992 || y == 1996)
    for (int k = 2; k < 13; ++k) ++start[k];
  int m = 0;
  for (int k = 1; k < 13; ++k) {
    if (days < start[k] && days >= start[k - 1]) {
      m = k;
      days -= start[k - 1];
      break;
    }
  }
  year = y;
  month = m;
  day = static_cast<int>(days) + 1;
}

__device__ inline uint64_t date2int(int year, int month, int day) {
  // Cumulative days before each month (non-leap); leap days only through
  // the 1992 and 1996 adjustments below, as Date.h:92-118 counts them.
  const uint32_t before[13] = {0, 0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
  uint64_t days = static_cast<uint64_t>(static_cast<int64_t>(year - 1970) * 365);
  days += before[month < 1 ? 0 : month > 12 ? 12 : month];
  days += static_cast<uint64_t>(static_cast<int64_t>(day - 1));
  if (year == 1992 && month <= 2) days += 5;
  if (year == 1992 && month > 2) days += 6;
  if (year > 1992 && year < 1996) days += 6;
  if (year == 1996 && month <= 2) days += 6;
  if (year == 1996 && month > 2) days += 7;
  if (year > 1996) days += 7;
  return days * 86400;
}

__device__ __forceinline__ void wrapdate(int& year, int& month) {
  if (month > 12) {
    year += (month - 1) / 12;
    month = month % 12;
  }
}

// SCHUNK: chunk k of the NUL-terminated string at address p as 8
// order-preserving byte codes, first byte most significant (the terminator
// and everything after it 0; signed chars -128..-1, 1..127 as 1..255).
__device__ inline uint64_t mf_schunk(uint64_t p, uint64_t k) {
  const uint8_t* s = reinterpret_cast<const uint8_t*>(p);
  bool ended = s == nullptr;
  for (uint64_t q = 0; q < 8 * k && !ended; ++q) ended = s[q] == 0;
  uint64_t v = 0;
  for (int b = 0; b < 8; ++b) {
    const uint8_t c = ended ? 0 : s[8 * k + b];
    ended = ended || c == 0;
    v = (v << 8) | (c == 0 ? 0u : static_cast<uint64_t>((c ^ 0x80u) + (c >= 0x80u ? 1u : 0u)));
  }
  return v;
}

// LIKE with '%' wildcards (Comparisons.h:12-372): the pattern's head (before
// its first '%') must start the string, its tail (after its last '%') must
// end it, and each middle word is then found leftmost from where the last
// one ended (a search that is not kept out of the tail, as there).  Null
// addresses read as the empty string.
__device__ inline uint64_t mf_like(uint64_t ps, uint64_t pp) {
  const uint8_t empty = 0;
  const uint8_t* s = ps ? reinterpret_cast<const uint8_t*>(ps) : &empty;
  const uint8_t* p = pp ? reinterpret_cast<const uint8_t*>(pp) : &empty;
  uint32_t s0 = 0, p0 = 0;
  if (p[0] != '%') {  // head
    uint32_t i = 0;
    while (s[i] && p[i] && p[i] != '%') {
      if (s[i] != p[i]) return 0;
      ++i;
    }
    if (p[i] && p[i] != '%') return 0;
    s0 = p0 = i;
  }
  if (p[p0] == 0) return s[s0] == 0;
  uint32_t sl = 0, pl = 0;
  while (s[sl]) ++sl;
  while (p[pl]) ++pl;
  uint32_t p1 = pl;
  if (p[pl - 1] != '%') {  // tail, compared from the ends
    uint32_t i = sl, j = pl;
    while (i > s0 && j > p0) {
      --i;
      --j;
      if (p[j] == '%') break;
      if (s[i] != p[j]) return 0;
    }
    if (p[j] != '%') return 0;
    p1 = j + 1;
  }
  uint32_t i = s0, j = p0;
  while (j < p1) {  // middle words, each found leftmost from i
    while (p[j] == '%') ++j;
    if (j >= p1) break;
    const uint32_t w0 = j;
    while (j < p1 && p[j] != '%') ++j;
    const uint32_t wl = j - w0;
    bool found = false;
    for (uint32_t a = i; s[a] && !found; ++a) {
      uint32_t k = 0;
      while (k < wl && s[a + k] && s[a + k] == p[w0 + k]) ++k;
      if (k == wl) {
        i = a + wl;
        found = true;
      }
    }
    if (!found) return 0;
  }
  return j == p1;
}

// The value of one op that is neither a column load (LDC) nor a store
// (OUT), from its operand registers a and b and its immediate.  With a
// constant code the switch folds to the one operation.
__device__ __forceinline__ uint64_t mf_apply(uint32_t code, uint64_t a, uint64_t b, uint64_t imm) {
  const int64_t sa = static_cast<int64_t>(a), sb = static_cast<int64_t>(b);
  switch (code) {
    case RFX_MF_LDI: return imm;
    case RFX_MF_MOV: return a;
    case RFX_MF_IADD: return a + b;
    case RFX_MF_ISUB: return a - b;
    case RFX_MF_IMUL: return a * b;
    case RFX_MF_IDIV: return mf_idiv(a, b);
    case RFX_MF_IMOD: return mf_imod(a, b);
    case RFX_MF_FADD: return as_u(as_f(a) + as_f(b));
    case RFX_MF_FSUB: return as_u(as_f(a) - as_f(b));
    case RFX_MF_FMUL: return as_u(as_f(a) * as_f(b));
    case RFX_MF_FDIV: return as_u(as_f(a) / as_f(b));
    case RFX_MF_I2F: return as_u(static_cast<double>(sa));
    case RFX_MF_U2F: return as_u(static_cast<double>(a));
    case RFX_MF_F2I: return mf_f2i(a);
    case RFX_MF_ILT: return sa < sb;
    case RFX_MF_ILE: return sa <= sb;
    case RFX_MF_IEQ: return a == b;
    case RFX_MF_INE: return a != b;
    case RFX_MF_ULT: return a < b;
    case RFX_MF_ULE: return a <= b;
    case RFX_MF_FLT: return as_f(a) < as_f(b);
    case RFX_MF_FLE: return as_f(a) <= as_f(b);
    case RFX_MF_FEQ: return as_f(a) == as_f(b);
    case RFX_MF_FNE: return as_f(a) != as_f(b);
    case RFX_MF_AND: return static_cast<uint64_t>((a != 0) & (b != 0));
    case RFX_MF_OR: return static_cast<uint64_t>((a != 0) | (b != 0));
    case RFX_MF_NOT: return a == 0;
    case RFX_MF_SCHUNK: return mf_schunk(a, imm);
    case RFX_MF_LIKE: return mf_like(a, b);
    case RFX_MF_ADD_MONTH:
    case RFX_MF_ADD_YEAR:
    case RFX_MF_PART_YEAR: {
      int y, m, d;
      int2date(a, y, m, d);
      if (code == RFX_MF_PART_YEAR) return static_cast<uint64_t>(y);
      if (code == RFX_MF_ADD_MONTH) {
        m += static_cast<int>(static_cast<int64_t>(imm));
        wrapdate(y, m);
      } else {
        y += static_cast<int>(static_cast<int64_t>(imm));
      }
      return date2int(y, m, d);
    }
    default: return 0;
  }
}

// Column and output pointers (and types, which the interpreter reads and
// a compiled program has as constants).
struct MfIn {
  const void* p[RFX_MF_MAX_COLS];
  int32_t t[RFX_MF_MAX_COLS];
  uint64_t rows[RFX_MF_MAX_COLS];  // bound of LDX's row index
};
struct MfOut {
  void* p[RFX_MF_MAX_OUTS];
  int32_t t[RFX_MF_MAX_OUTS];
};

// One tile of kMTile rows, in two phases.
//
// Filter: filter(r) says whether row r is kept.  It runs on every row in
// batches of UF steps of kMNT rows, rows past the end reading the last row
// (so a batch is straight-line code whose loads are all in flight at
// once) and counting as dropped.  Each wave ballots the flags of its 64
// rows per step, the (step, wave) counts are scanned in LDS, the kept rows'
// tile offsets are compacted into LDS, and the tile's output offset comes
// from a decoupled look-back.
//
// Outputs: compute(r) returns row r's values, store(r, pos, v) writes them
// at position pos.  Thread t of step j takes the (kMNT j + t)-th kept row, so
// the phase reads kept rows only, every step but the tile's last is dense,
// and stores go to consecutive positions.  A batch computes UO rows (the
// last kept row standing in past the tile's count) before it stores any.
template <int UF, int UO, class Filter, class Compute, class Store>
__device__ __forceinline__ void mapfilter_tile(Filter filter, Compute compute, Store store,
                                               uint64_t n, uint32_t ntiles, uint64_t* n_kept,
                                               uint64_t* status, uint32_t* ticket) {
  static_assert(kMVT % UF == 0 && kMVT % UO == 0, "batches divide the tile");
  __shared__ uint32_t s_cnt[kMVT * kMNW + 1];
  __shared__ uint16_t s_rows[kMTile];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t base = static_cast<uint64_t>(tile) * kMTile;
  uint32_t keep = 0;
#pragma unroll 1
  for (int j0 = 0; j0 < kMVT; j0 += UF) {
#pragma unroll
    for (int u = 0; u < UF; ++u) {
      const uint64_t r = base + static_cast<uint64_t>(j0 + u) * kMNT + tid;
      keep |= static_cast<uint32_t>(filter(r < n ? r : n - 1) & (r < n)) << (j0 + u);
    }
  }
  uint64_t ball[kMVT];
#pragma unroll
  for (int j = 0; j < kMVT; ++j) ball[j] = ballot((keep >> j) & 1u);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < kMVT; ++j) s_cnt[j * kMNW + w] = static_cast<uint32_t>(__popcll(ball[j]));
  }
  __syncthreads();
  if (w == 0) {
    // Lane l scans entries [l*P, l*P + P) (entry j*kMNW + w: row order).
    constexpr int kE = kMVT * kMNW, P = kE > kWave ? kE / kWave : 1;
    static_assert(kE <= kWave || kE % kWave == 0, "whole runs per lane");
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
    if (lane == kWave - 1) s_cnt[kE] = incl;
  }
  __syncthreads();
  const uint32_t kept = s_cnt[kMVT * kMNW];
#pragma unroll
  for (int j = 0; j < kMVT; ++j) {
    if ((ball[j] >> lane) & 1u)
      s_rows[s_cnt[j * kMNW + w] + mbcnt(ball[j])] = static_cast<uint16_t>(j * kMNT + tid);
  }
  if (w == 0) {
    if (tile == 0) {
      if (lane == 0) {
        status_store64(status, kFlagIncl64 | kept);
        s_prefix = 0;
      }
    } else {
      if (lane == 0) status_store64(status + tile, kFlagAgg64 | kept);
      const uint64_t excl = wave_lookback64(status, tile);
      if (lane == 0) {
        status_store64(status + tile, kFlagIncl64 | (excl + kept));
        s_prefix = excl;
      }
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  if (tile == ntiles - 1 && tid == 0) *n_kept = prefix + kept;
#pragma unroll 1
  for (uint32_t i0 = 0; i0 < kept; i0 += UO * kMNT) {
    decltype(compute(base)) v[UO];
#pragma unroll
    for (int u = 0; u < UO; ++u) {
      const uint32_t i = i0 + u * kMNT + tid;
      v[u] = compute(base + s_rows[i < kept ? i : kept - 1]);
    }
#pragma unroll
    for (int u = 0; u < UO; ++u) {
      const uint32_t i = i0 + u * kMNT + tid;
      if (i < kept) store(base + s_rows[i], prefix + i, v[u]);
    }
  }
}

}  // namespace rfx
