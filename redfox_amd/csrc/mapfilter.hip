// mapfilter.hip - a fused filter / project / arithmetic pass over columns
// (the RA MAPFILTER operator) for gfx950.
//
// Replaces the reference's per-instance generated kernels for arithmetic
// and selection (set_arith / set_convert, RelationalAlgebraKernel.cpp:
// 2001-2215: one new field computed from a C expression over the source
// tuple's fields; set_select, :2301-2593: a comparison filter with
// per-CTA compaction) with one kernel that interprets a small register
// program.  Every lane runs the same program, so its opcode branches are
// uniform and cost no divergence; the register file is indexed with
// uniform indices (VGPR indexing mode, no scratch).
//
// One pass per tile of 256 x 8 rows, in two phases.  The filter program
// runs on every row; each wave ballots the keep flags of 64 consecutive
// rows per step, the (step, wave) counts are scanned in LDS, and the tile
// takes its output offset from a decoupled look-back.  The output program
// then runs on the kept rows only and stores each result straight to its
// place, so the output keeps input order.
//
// Arithmetic: integers in two's-complement 64 bits, C division and
// remainder (truncating; a zero divisor gives 0); doubles in IEEE binary64
// with no contraction into fused multiply-adds, so results equal a serial
// evaluation of the same expression.  The date operations restate
// Date.h:17-118 (seconds since 1970 for the TPC-H years, wrapdate's month
// arithmetic included).
#include "rfx_device.h"
#include "rfx_internal.h"

#pragma clang fp contract(off)

namespace rfx {
namespace {

constexpr int kMNT = 256;
constexpr int kMNW = kMNT / kWave;
constexpr int kMVT = 8;
constexpr int kMTile = kMNT * kMVT;

struct MfIn {
  const void* p[RFX_MF_MAX_COLS];
  int32_t t[RFX_MF_MAX_COLS];
};
struct MfOut {
  void* p[RFX_MF_MAX_OUTS];
  int32_t t[RFX_MF_MAX_OUTS];
};

__device__ __forceinline__ double as_f(uint64_t v) {
  return __longlong_as_double(static_cast<long long>(v));
}
__device__ __forceinline__ uint64_t as_u(double d) {
  return static_cast<uint64_t>(__double_as_longlong(d));
}

// Element i of a column of type t at p: integers widened (zero- or sign-), floats to double bits.  Packed
// tuple columns yield the field [shift, shift + bits) named by the op's
// immediate (bits 0-7 shift, 8-15 bits, 16-17: 0 unsigned, 1 signed,
// 2 double).
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

__device__ __forceinline__ void store_out(const MfOut& out, uint32_t k, uint64_t pos, uint64_t v) {
  void* p = out.p[k];
  switch (out.t[k]) {
    case RFX_MF_U8: case RFX_MF_I8: static_cast<uint8_t*>(p)[pos] = static_cast<uint8_t>(v); break;
    case RFX_MF_U16: case RFX_MF_I16: static_cast<uint16_t*>(p)[pos] = static_cast<uint16_t>(v); break;
    case RFX_MF_U32: case RFX_MF_I32: static_cast<uint32_t*>(p)[pos] = static_cast<uint32_t>(v); break;
    case RFX_MF_F32: static_cast<float*>(p)[pos] = static_cast<float>(as_f(v)); break;
    default: static_cast<uint64_t*>(p)[pos] = v; break;  // U64, I64, F64
  }
}

// Date.h: seconds since 1970-01-01 <-> (year, month, day) for 1992-1998.
__device__ void int2date(uint64_t seconds, int& year, int& month, int& day) {
  uint64_t days = seconds / 86400 - 8035;  // days since 1992-01-01
  int y = 1992;
  // Year lengths from 1992 on; the walk stops at the first that does not fit
  // (six steps at most, as Date.h:37-88 does).
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

__device__ uint64_t date2int(int year, int month, int day) {
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

// Run ops [begin, end) for row r.  OUT ops store to position `pos`.
// Returns the keep register.
__device__ __forceinline__ uint64_t run_ops(const rfx_mf_program& prog, uint32_t begin,
                                            uint32_t end, const MfIn& in, const MfOut& out,
                                            uint64_t r, uint64_t pos) {
  uint64_t R[RFX_MF_REGS];
#pragma unroll
  for (int k = 0; k < RFX_MF_REGS; ++k) R[k] = 0;
  for (uint32_t pc = begin; pc < end; ++pc) {
    const rfx_mf_op op = prog.ops[pc];
    const uint64_t a = R[op.a], b = R[op.b];
    const int64_t sa = static_cast<int64_t>(a), sb = static_cast<int64_t>(b);
    uint64_t v = 0;
    switch (op.code) {
      case RFX_MF_LDC: v = load_elem(in.p[op.a], in.t[op.a], r, op.imm); break;
      case RFX_MF_LDI: v = op.imm; break;
      case RFX_MF_MOV: v = a; break;
      case RFX_MF_IADD: v = a + b; break;
      case RFX_MF_ISUB: v = a - b; break;
      case RFX_MF_IMUL: v = a * b; break;
      case RFX_MF_IDIV:
        v = sb == 0 ? 0 : (sb == -1 ? 0 - a : static_cast<uint64_t>(sa / sb));
        break;
      case RFX_MF_IMOD: v = (sb == 0 || sb == -1) ? 0 : static_cast<uint64_t>(sa % sb); break;
      case RFX_MF_FADD: v = as_u(as_f(a) + as_f(b)); break;
      case RFX_MF_FSUB: v = as_u(as_f(a) - as_f(b)); break;
      case RFX_MF_FMUL: v = as_u(as_f(a) * as_f(b)); break;
      case RFX_MF_FDIV: v = as_u(as_f(a) / as_f(b)); break;
      case RFX_MF_I2F: v = as_u(static_cast<double>(sa)); break;
      case RFX_MF_U2F: v = as_u(static_cast<double>(a)); break;
      case RFX_MF_F2I: {
        const double d = as_f(a);
        v = !(d == d) ? 0
            : d >= 9223372036854775807.0 ? static_cast<uint64_t>(INT64_MAX)
            : d <= -9223372036854775808.0 ? static_cast<uint64_t>(INT64_MIN)
                                          : static_cast<uint64_t>(static_cast<int64_t>(d));
        break;
      }
      case RFX_MF_ILT: v = sa < sb; break;
      case RFX_MF_ILE: v = sa <= sb; break;
      case RFX_MF_IEQ: v = a == b; break;
      case RFX_MF_INE: v = a != b; break;
      case RFX_MF_ULT: v = a < b; break;
      case RFX_MF_ULE: v = a <= b; break;
      case RFX_MF_FLT: v = as_f(a) < as_f(b); break;
      case RFX_MF_FLE: v = as_f(a) <= as_f(b); break;
      case RFX_MF_FEQ: v = as_f(a) == as_f(b); break;
      case RFX_MF_FNE: v = as_f(a) != as_f(b); break;
      case RFX_MF_AND: v = (a != 0) && (b != 0); break;
      case RFX_MF_OR: v = (a != 0) || (b != 0); break;
      case RFX_MF_NOT: v = a == 0; break;
      case RFX_MF_ADD_MONTH:
      case RFX_MF_ADD_YEAR: {
        int y, m, d;
        int2date(a, y, m, d);
        if (op.code == RFX_MF_ADD_MONTH) {
          m += static_cast<int>(static_cast<int64_t>(op.imm));
          wrapdate(y, m);
        } else {
          y += static_cast<int>(static_cast<int64_t>(op.imm));
        }
        v = date2int(y, m, d);
        break;
      }
      case RFX_MF_PART_YEAR: {
        int y, m, d;
        int2date(a, y, m, d);
        v = static_cast<uint64_t>(y);
        break;
      }
      case RFX_MF_OUT:
        store_out(out, op.b, pos, a);
        continue;
      default:
        continue;
    }
    R[op.d] = v;
  }
  return R[prog.keep_reg];
}

__global__ __launch_bounds__(kMNT) void k_mapfilter(MfIn in, uint64_t n, rfx_mf_program prog,
                                                    MfOut out, uint32_t ntiles,
                                                    uint64_t* __restrict__ n_kept,
                                                    uint64_t* __restrict__ status,
                                                    uint32_t* __restrict__ ticket) {
  __shared__ uint32_t s_cnt[kMVT * kMNW + 1];
  __shared__ uint64_t s_ball[kMVT * kMNW];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t base = static_cast<uint64_t>(tile) * kMTile;
  // Step j of wave w covers rows base + j*256 + w*64 .. +63.  The row
  // loops stay rolled: each holds one copy of the interpreter.
  uint32_t keep = 0;
#pragma unroll 1
  for (int j = 0; j < kMVT; ++j) {
    const uint64_t r = base + static_cast<uint64_t>(j) * kMNT + tid;
    bool k = false;
    if (r < n) k = prog.n_filter == 0 || run_ops(prog, 0, prog.n_filter, in, out, r, 0) != 0;
    keep |= static_cast<uint32_t>(k) << j;
    const uint64_t m = ballot(k);
    if (lane == 0) {
      s_cnt[j * kMNW + w] = static_cast<uint32_t>(__popcll(m));
      s_ball[j * kMNW + w] = m;
    }
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t c = lane < kMVT * kMNW ? s_cnt[lane] : 0u;
    const uint32_t incl = wave_incl_scan(c);
    if (lane < kMVT * kMNW) s_cnt[lane] = incl - c;
    const uint32_t tile_kept = __shfl(incl, kWave - 1, kWave);
    if (lane == 0) s_cnt[kMVT * kMNW] = tile_kept;
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
#pragma unroll 1
  for (int j = 0; j < kMVT; ++j) {
    if ((keep >> j) & 1u) {
      const uint64_t r = base + static_cast<uint64_t>(j) * kMNT + tid;
      run_ops(prog, prog.n_filter, prog.n_ops, in, out, r,
              prefix + s_cnt[j * kMNW + w] + mbcnt(s_ball[j * kMNW + w]));
    }
  }
  if (tile == ntiles - 1 && tid == 0) *n_kept = prefix + s_cnt[kMVT * kMNW];
}

bool valid_col_type(int t) { return t >= RFX_MF_U8 && t <= RFX_MF_TUPLE16; }

bool valid_code(uint32_t c) {
  return c <= RFX_MF_MOV || (c >= RFX_MF_IADD && c <= RFX_MF_IMOD) ||
         (c >= RFX_MF_FADD && c <= RFX_MF_FDIV) || (c >= RFX_MF_I2F && c <= RFX_MF_F2I) ||
         (c >= RFX_MF_ILT && c <= RFX_MF_FNE) || (c >= RFX_MF_AND && c <= RFX_MF_NOT) ||
         (c >= RFX_MF_ADD_MONTH && c <= RFX_MF_PART_YEAR) || c == RFX_MF_OUT;
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_mapfilter(rfx_ctx* ctx, const rfx_mf_col* cols, int n_cols, uint64_t n_rows,
                             const rfx_mf_program* prog, const rfx_mf_out* outs, int n_outs,
                             uint64_t* n_kept) {
  if (!ctx || !prog || !n_kept || (n_cols && !cols) || (n_outs && !outs))
    return fail(RFX_EINVAL, "rfx_mapfilter: null argument");
  if (n_cols < 0 || n_cols > RFX_MF_MAX_COLS || n_outs < 0 || n_outs > RFX_MF_MAX_OUTS)
    return fail(RFX_EINVAL, "rfx_mapfilter: %d columns, %d outputs", n_cols, n_outs);
  if (prog->n_ops > RFX_MF_MAX_OPS || prog->n_filter > prog->n_ops || prog->keep_reg >= RFX_MF_REGS)
    return fail(RFX_EINVAL, "rfx_mapfilter: program of %u ops (%u filter)", prog->n_ops, prog->n_filter);
  if (n_rows >= (1ull << 40)) return fail(RFX_ERANGE, "rfx_mapfilter: too many rows");
  MfIn in{};
  MfOut out{};
  for (int c = 0; c < n_cols; ++c) {
    if (!valid_col_type(cols[c].type) || (n_rows && !cols[c].data))
      return fail(RFX_EINVAL, "rfx_mapfilter: column %d (type %d)", c, cols[c].type);
    in.p[c] = cols[c].data;
    in.t[c] = cols[c].type;
  }
  for (int k = 0; k < n_outs; ++k) {
    if (outs[k].type < RFX_MF_U8 || outs[k].type > RFX_MF_F64 || (n_rows && !outs[k].data))
      return fail(RFX_EINVAL, "rfx_mapfilter: output %d (type %d)", k, outs[k].type);
    out.p[k] = outs[k].data;
    out.t[k] = outs[k].type;
  }
  // Every operand the program names must exist.
  for (uint32_t pc = 0; pc < prog->n_ops; ++pc) {
    const rfx_mf_op& op = prog->ops[pc];
    const bool filter = pc < prog->n_filter;
    if (op.code == RFX_MF_LDC) {
      if (op.a >= n_cols) return fail(RFX_EINVAL, "rfx_mapfilter: op %u reads column %u", pc, op.a);
      if (cols[op.a].type >= RFX_MF_TUPLE1) {
        const uint32_t shift = op.imm & 0xFF, bits = (op.imm >> 8) & 0xFF;
        const int wb = 8 << (cols[op.a].type - RFX_MF_TUPLE1);
        if (bits == 0 || bits > 64 || shift + bits > static_cast<uint32_t>(wb) ||
            (((op.imm >> 16) & 3) == 2 && bits != 64))
          return fail(RFX_EINVAL, "rfx_mapfilter: op %u field [%u, +%u) of a %d-bit tuple", pc,
                      shift, bits, wb);
      }
    } else if (op.code == RFX_MF_OUT) {
      if (filter || op.b >= n_outs) return fail(RFX_EINVAL, "rfx_mapfilter: op %u writes output %u", pc, op.b);
    } else if (!valid_code(op.code)) {
      return fail(RFX_EINVAL, "rfx_mapfilter: op %u code %u", pc, op.code);
    }
    if (op.d >= RFX_MF_REGS || (op.code != RFX_MF_LDC && (op.a >= RFX_MF_REGS || op.b >= RFX_MF_REGS)))
      return fail(RFX_EINVAL, "rfx_mapfilter: op %u register out of range", pc);
  }
  if (n_rows == 0) return fill_u32(ctx, reinterpret_cast<uint32_t*>(n_kept), 2, 0);
  const uint64_t ntiles = (n_rows + kMTile - 1) / kMTile;
  Carve c(nullptr);
  c.take<uint64_t>(ntiles);
  c.take<uint32_t>(1);
  RFX_TRY(ensure_workspace(ctx, c.off));
  Carve w(ctx->ws);
  uint64_t* status = w.take<uint64_t>(ntiles);
  uint32_t* ticket = w.take<uint32_t>(1);
  RFX_TRY(fill_u32(ctx, reinterpret_cast<uint32_t*>(status),
                   (reinterpret_cast<char*>(ticket + 1) - reinterpret_cast<char*>(status)) / 4, 0));
  RFX_PROF(ctx, "mapfilter");
  k_mapfilter<<<static_cast<uint32_t>(ntiles), kMNT, 0, ctx->stream>>>(
      in, n_rows, *prog, out, static_cast<uint32_t>(ntiles), n_kept, status, ticket);
  RFX_LAUNCHED();
  return RFX_OK;
}
