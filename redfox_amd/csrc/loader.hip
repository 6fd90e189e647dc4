// loader.hip - delimited text (TPC-H .tbl rows) to typed columns, on the
// device.  Restates the reference's loaders (TC_data_process.cu:20-100:
// getline, strtok on the delimiter, atoi per field, one packed tuple per
// row) as two HBM passes over the text once it is resident:
//
//  1. row ends: tiles of 512 threads x 64 bytes count the newlines in each
//     thread's 64 bytes, scan the counts across the block and take the
//     tile's offset from a decoupled look-back, then store the position of
//     every newline in text order;
//  2. fields: a block of 256 rows stages their text (one contiguous span)
//     in LDS with 16-byte loads; one thread per row then walks its bytes,
//     splitting at the delimiter, and converts field f by the type spec[f]
//     straight into column f.
//
// Field conversions (C semantics): integers as atoi/atol (leading spaces,
// a sign, digits up to the first other byte; wrapping), decimals as the
// digits before and after one '.' (the value mantissa / 10^k, one IEEE
// division, so the result is the correctly rounded one while the mantissa
// stays below 2^53), dates YYYY-MM-DD as seconds since 1970 in Date.h's
// calendar, a char as its first byte.  A '\r' before the newline is not
// part of the last field.
#include "mapfilter_dev.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kLNT = 512;
constexpr int kLBytes = 64;
constexpr int kLTile = kLNT * kLBytes;
constexpr int kMaxFields = RFX_TBL_MAX_COLS;

__global__ __launch_bounds__(kLNT) void k_tbl_rows(const uint8_t* __restrict__ text,
                                                   uint64_t bytes, bool aligned,
                                                   uint64_t* __restrict__ ends,
                                                   uint64_t cap, uint32_t ntiles,
                                                   uint64_t* __restrict__ n_newlines,
                                                   uint64_t* __restrict__ status,
                                                   uint32_t* __restrict__ ticket) {
  __shared__ uint32_t s_scan[kLNT / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);  // block-uniform: scalar from here
  const uint64_t at = static_cast<uint64_t>(tile) * kLTile + static_cast<uint64_t>(tid) * kLBytes;
  // The thread's 64 bytes as sixteen words; bytes past the end are not
  // newlines.
  uint32_t word[kLBytes / 4];
  if (aligned && at + kLBytes <= bytes) {
#pragma unroll
    for (int j = 0; j < kLBytes / 16; ++j) {
      const uint4 v = reinterpret_cast<const uint4*>(text + at)[j];
      word[4 * j] = v.x;
      word[4 * j + 1] = v.y;
      word[4 * j + 2] = v.z;
      word[4 * j + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kLBytes / 4; ++j) word[j] = 0;
    for (int b = 0; b < kLBytes; ++b)
      if (at + b < bytes) word[b >> 2] |= static_cast<uint32_t>(text[at + b]) << (8 * (b & 3));
  }
  uint64_t mask = 0;  // bit b: byte b is a newline
#pragma unroll
  for (int b = 0; b < kLBytes; ++b) {
    const uint32_t c = (word[b >> 2] >> (8 * (b & 3))) & 0xFF;
    mask |= static_cast<uint64_t>(c == '\n' && at + b < bytes) << b;
  }
  uint32_t total = 0;
  const uint32_t excl = block_excl_scan<kLNT>(static_cast<uint32_t>(__popcll(mask)), s_scan, &total);
  if (w == 0) {
    if (tile == 0) {
      if (lane == 0) {
        status_store64(status, kFlagIncl64 | total);
        s_prefix = 0;
      }
    } else {
      if (lane == 0) status_store64(status + tile, kFlagAgg64 | total);
      const uint64_t prev = wave_lookback64(status, tile);
      if (lane == 0) {
        status_store64(status + tile, kFlagIncl64 | (prev + total));
        s_prefix = prev;
      }
    }
  }
  __syncthreads();
  uint64_t k = s_prefix + excl;
  while (mask) {
    const int b = __builtin_ctzll(mask);
    mask &= mask - 1;
    if (k < cap) ends[k] = at + b;
    ++k;
  }
  if (tile == ntiles - 1 && tid == 0) *n_newlines = s_prefix + total;
}

struct TblCols {
  void* p[kMaxFields];
  uint8_t t[kMaxFields];
};

__device__ __forceinline__ void tbl_store(const TblCols& c, int f, uint64_t row, int64_t iv,
                                          double dv) {
  void* p = c.p[f];
  switch (c.t[f]) {
    case RFX_TBL_I32: case RFX_TBL_U32: static_cast<uint32_t*>(p)[row] = static_cast<uint32_t>(iv); break;
    case RFX_TBL_I64: case RFX_TBL_DATE: static_cast<int64_t*>(p)[row] = iv; break;
    case RFX_TBL_U8: case RFX_TBL_CHAR: static_cast<uint8_t*>(p)[row] = static_cast<uint8_t>(iv); break;
    case RFX_TBL_F64: static_cast<double*>(p)[row] = dv; break;
    case RFX_TBL_STR: static_cast<int64_t*>(p)[row] = iv; break;
    default: break;
  }
}

__device__ const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                      1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                      1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

constexpr int kFNT = 256;            // rows per block
constexpr int kStage = 48 * 1024;     // staged text bytes per block

// Parse row r ([pos, end) of the text) into the columns; at(q) gives byte
// q and scan(q, e) the first delimiter in [q, e) (e if none).  Each field's
// end is found first and its value parsed from its own bytes only, as
// splitting the row at the delimiters and then reading each field would
// (so a space, sign, point or digit delimiter never joins two fields).
template <class At, class Scan>
__device__ __forceinline__ void parse_row(At at, Scan scan, uint8_t* text, uint64_t bytes,
                                          uint64_t r, uint64_t pos, uint64_t end, int n_fields,
                                          const TblCols& cols, uint64_t* short_rows) {
  if (end > pos && at(end - 1) == '\r') --end;
  int f = 0;
  while (f < n_fields) {
    const int t = cols.t[f];
    const uint64_t fe = scan(pos, end);
    int64_t iv = 0;
    double dv = 0.0;
    if (t == RFX_TBL_CHAR) {
      iv = pos < fe ? at(pos) : 0;
    } else if (t == RFX_TBL_DATE) {
      // YYYY-MM-DD: each '-'-separated part read as atoi reads it
      // (leading spaces, a '+', digits up to the first other byte).
      int part[3] = {0, 0, 0}, k = 0, state = 0;  // 0 leading, 1 digits, 2 done
      for (uint64_t q = pos; q < fe && k < 3; ++q) {
        const uint8_t c = at(q);
        if (c == '-') {
          ++k;
          state = 0;
        } else if (c >= '0' && c <= '9' && state < 2) {
          part[k] = part[k] * 10 + (c - '0');
          state = 1;
        } else if (!(state == 0 && (c == ' ' || c == '+'))) {
          state = 2;
        }
      }
      iv = static_cast<int64_t>(date2int(part[0], part[1], part[2]));
    } else if (t != RFX_TBL_SKIP && t != RFX_TBL_STR) {
      uint64_t q = pos;
      while (q < fe && at(q) == ' ') ++q;
      bool neg = false;
      if (q < fe && (at(q) == '-' || at(q) == '+')) neg = at(q++) == '-';
      uint64_t m = 0;
      int frac = -1;  // digits after the point
      for (; q < fe; ++q) {
        const uint8_t c = at(q);
        if (c >= '0' && c <= '9') {
          m = m * 10 + (c - '0');
          if (frac >= 0) ++frac;
        } else if (c == '.' && t == RFX_TBL_F64 && frac < 0) {
          frac = 0;
        } else {
          break;
        }
      }
      iv = neg ? -static_cast<int64_t>(m) : static_cast<int64_t>(m);
      if (t == RFX_TBL_F64) {
        const double mag = static_cast<double>(m) / kPow10[frac > 0 ? (frac < 22 ? frac : 22) : 0];
        dv = neg ? -mag : mag;
      }
    } else if (t == RFX_TBL_STR) {
      // The field's address; its end (delimiter, '\r' or newline) becomes
      // its terminating NUL.  A field running to the end of the text has
      // nowhere to end: it reads as the null pointer (the empty string).
      iv = fe < bytes ? static_cast<int64_t>(reinterpret_cast<uintptr_t>(text + pos)) : 0;
      if (fe < bytes) text[fe] = 0;
    }
    tbl_store(cols, f, r, iv, dv);
    ++f;
    pos = fe;
    if (pos >= end) break;
    ++pos;  // past the delimiter
  }
  if (f < n_fields) {
    atomicAdd(reinterpret_cast<unsigned long long*>(short_rows), 1ull);
    for (; f < n_fields; ++f) tbl_store(cols, f, r, 0, 0.0);  // missing fields read 0
  }
}

// Rows [256 b, 256 b + 256) of block b.  Their text is one contiguous span:
// when it fits, the block stages it in LDS with 16-byte loads and every
// thread parses its row from there; otherwise from global memory.
__global__ __launch_bounds__(kFNT) void k_tbl_fields(uint8_t* __restrict__ text,
                                                     uint64_t bytes, bool aligned,
                                                     const uint64_t* __restrict__ ends,
                                                     uint64_t n_newlines, uint64_t rows,
                                                     uint8_t delim, int n_fields, TblCols cols,
                                                     uint64_t* __restrict__ short_rows) {
  __shared__ uint4 s_text[kStage / 16];
  __shared__ uint64_t s_span[2];
  const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * kFNT;
  if (threadIdx.x == 0) {
    const uint64_t last = r0 + kFNT - 1 < rows ? r0 + kFNT - 1 : rows - 1;
    s_span[0] = r0 == 0 ? 0 : ends[r0 - 1] + 1;
    s_span[1] = last < n_newlines ? ends[last] : bytes;
  }
  __syncthreads();
  const uint64_t base = s_span[0] & ~15ull, stop = s_span[1];
  const bool staged = aligned && stop - base <= static_cast<uint64_t>(kStage);
  if (staged) {
    const uint32_t nvec = static_cast<uint32_t>((stop - base + 15) / 16);
    for (uint32_t v = threadIdx.x; v < nvec; v += kFNT) {
      const uint64_t off = base + 16ull * v;
      if (off + 16 <= bytes) {
        s_text[v] = *reinterpret_cast<const uint4*>(text + off);
      } else {
        uint32_t w[4] = {0, 0, 0, 0};
        for (int b = 0; off + b < bytes && b < 16; ++b)
          w[b >> 2] |= static_cast<uint32_t>(text[off + b]) << (8 * (b & 3));
        s_text[v] = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
  }
  __syncthreads();
  const uint64_t r = r0 + threadIdx.x;
  if (r >= rows) return;
  const uint64_t pos = r == 0 ? 0 : ends[r - 1] + 1;
  const uint64_t end = r < n_newlines ? ends[r] : bytes;
  if (staged) {
    const uint8_t* lds = reinterpret_cast<const uint8_t*>(s_text);
    // The delimiter search reads 8 aligned bytes of LDS at a time: the
    // lowest flagged byte of (x - 0x01..01) & ~x & 0x80..80 is x's first
    // zero byte, x being the word XOR the delimiter in every byte.
    const uint64_t pat = 0x0101010101010101ull * delim;
    auto scan = [&](uint64_t q, uint64_t e) -> uint64_t {
      uint32_t i = static_cast<uint32_t>(q - base);
      const uint32_t ie = static_cast<uint32_t>(e - base);
      while (i < ie && (i & 7u)) {
        if (lds[i] == delim) return base + i;
        ++i;
      }
      for (; i + 8 <= ie; i += 8) {
        const uint64_t x = *reinterpret_cast<const uint64_t*>(lds + i) ^ pat;
        const uint64_t z = (x - 0x0101010101010101ull) & ~x & 0x8080808080808080ull;
        if (z) return base + i + (static_cast<uint32_t>(__builtin_ctzll(z)) >> 3);
      }
      while (i < ie && lds[i] != delim) ++i;
      return base + i;
    };
    parse_row([&](uint64_t q) -> uint8_t { return lds[q - base]; }, scan, text, bytes, r, pos,
              end, n_fields, cols, short_rows);
  } else {
    auto scan = [&](uint64_t q, uint64_t e) -> uint64_t {
      while (q < e && text[q] != delim) ++q;
      return q;
    };
    parse_row([&](uint64_t q) -> uint8_t { return text[q]; }, scan, text, bytes, r, pos, end,
              n_fields, cols, short_rows);
  }
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_tbl_parse(rfx_ctx* ctx, void* text, uint64_t bytes, char delim,
                             const int* types, int n_fields, void* const* columns,
                             uint64_t capacity, uint64_t* rows, uint64_t* short_rows) {
  if (!ctx || !rows || !short_rows || (bytes && !text) || (n_fields && (!types || !columns)))
    return fail(RFX_EINVAL, "rfx_tbl_parse: null argument");
  if (n_fields < 0 || n_fields > kMaxFields)
    return fail(RFX_EINVAL, "rfx_tbl_parse: %d fields (at most %d)", n_fields, kMaxFields);
  if (delim == '\n') return fail(RFX_EINVAL, "rfx_tbl_parse: the delimiter cannot be a newline");
  if (bytes >= (1ull << 40)) return fail(RFX_ERANGE, "rfx_tbl_parse: text too large");
  TblCols cols{};
  for (int f = 0; f < n_fields; ++f) {
    if (types[f] < RFX_TBL_SKIP || types[f] > RFX_TBL_STR)
      return fail(RFX_EINVAL, "rfx_tbl_parse: field %d type %d", f, types[f]);
    if (types[f] != RFX_TBL_SKIP && capacity && !columns[f])
      return fail(RFX_EINVAL, "rfx_tbl_parse: field %d has no column", f);
    cols.p[f] = types[f] == RFX_TBL_SKIP ? nullptr : columns[f];
    cols.t[f] = static_cast<uint8_t>(types[f]);
  }
  // Row ends are kept for the rows that are parsed (the count covers all).
  const uint64_t ntiles = bytes ? (bytes + kLTile - 1) / kLTile : 0;
  const uint64_t ecap = capacity < bytes ? capacity + 1 : (bytes ? bytes : 1);
  Carve c(nullptr);
  c.take<uint64_t>(ecap);
  c.take<uint64_t>(ntiles ? ntiles : 1);
  c.take<uint32_t>(1);
  c.take<uint64_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* ends = w.take<uint64_t>(ecap);
  uint64_t* status = w.take<uint64_t>(ntiles ? ntiles : 1);
  uint32_t* ticket = w.take<uint32_t>(1);
  uint64_t* n_nl = w.take<uint64_t>(1);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(n_nl + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(short_rows), 2, 0));
  uint64_t newlines = 0;
  if (ntiles) {
    RFX_PROF(ctx, "tbl.rows");
    k_tbl_rows<<<static_cast<uint32_t>(ntiles), kLNT, 0, ctx->stream>>>(
        static_cast<const uint8_t*>(text), bytes, reinterpret_cast<uintptr_t>(text) % 16 == 0, ends,
        ecap, static_cast<uint32_t>(ntiles), n_nl, status, ticket);
    RFX_LAUNCHED();
  }
  // The row count sizes the field pass: read it back.
  RFX_HIP(hipMemcpyAsync(&newlines, n_nl, 8, hipMemcpyDeviceToHost, ctx->stream));
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  if (!ntiles) newlines = 0;
  uint8_t last = 0;
  if (bytes) {
    RFX_HIP(hipMemcpyAsync(&last, static_cast<const uint8_t*>(text) + bytes - 1, 1,
                           hipMemcpyDeviceToHost, ctx->stream));
    RFX_HIP(hipStreamSynchronize(ctx->stream));
  }
  const uint64_t n = newlines + (bytes && last != '\n' ? 1 : 0);
  RFX_HIP(hipMemcpyAsync(rows, &n, 8, hipMemcpyHostToDevice, ctx->stream));
  const uint64_t todo = n < capacity ? n : capacity;
  if (todo && n_fields) {
    RFX_PROF(ctx, "tbl.fields");
    k_tbl_fields<<<static_cast<uint32_t>((todo + kFNT - 1) / kFNT), kFNT, 0, ctx->stream>>>(
        static_cast<uint8_t*>(text), bytes, reinterpret_cast<uintptr_t>(text) % 16 == 0,
        ends, newlines, todo, static_cast<uint8_t>(delim), n_fields, cols, short_rows);
    RFX_LAUNCHED();
  }
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  return RFX_OK;
}
