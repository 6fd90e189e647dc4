// mapfilter.hip - a fused filter / project / arithmetic pass over columns
// (the RA MAPFILTER operator) for gfx950.
//
// Replaces the reference's per-instance generated kernels for arithmetic
// and selection (set_arith / set_convert, RelationalAlgebraKernel.cpp:
// 2001-2215: one new field computed from a C expression over the source
// tuple's fields; set_select, :2301-2593: a comparison filter with
// per-CTA compaction).  A program is a short list of register ops
// (include/rfx_ra.h); it runs in one of two ways, with the same tile
// (mapfilter_dev.h: filter, ballot compaction with look-back, outputs for
// the kept rows in input order) and the same per-op semantics (mf_apply):
//
//  * compiled: the program, with its column and output types, is turned
//    into the source of one kernel, compiled with hiprtc for gfx950 on its
//    first use and cached for the process.  Every op becomes one statement
//    on named registers, so a program costs what the same expression
//    written by hand costs;
//  * interpreted: one kernel runs any program, every lane stepping through
//    the same ops (uniform branches, registers indexed with uniform
//    indices).  It serves while a graph is being captured, when the
//    compiler is unavailable, or when the "mapfilter.mode" knob asks for it.
#include <hip/hiprtc.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "mapfilter_dev.h"
#include "rfx_internal.h"

// The layouts the Python binding (redfox_amd/mapfilter.py) mirrors.
static_assert(sizeof(rfx_mf_op) == 16 && sizeof(rfx_mf_col) == 24 && sizeof(rfx_mf_out) == 16 &&
                  sizeof(rfx_mf_program) == 16 + 16 * RFX_MF_MAX_OPS,
              "MAPFILTER ABI structs");

namespace rfx {
namespace {

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
    if (op.code == RFX_MF_LDC) {
      R[op.d] = load_elem(in.p[op.a], in.t[op.a], r, op.imm);
    } else if (op.code == RFX_MF_LDX) {
      R[op.d] = load_at(in.p[op.a], in.t[op.a], in.rows[op.a], R[op.b], op.imm);
    } else if (op.code == RFX_MF_OUT) {
      store_out(out.p[op.b], out.t[op.b], pos, R[op.a]);
    } else if (op.code != RFX_MF_NOP) {
      R[op.d] = mf_apply(op.code, R[op.a], R[op.b], op.imm);
    }
  }
  return R[prog.keep_reg];
}

// The row loops stay rolled: each holds one copy of the interpreter.  The
// output phase runs the output ops (which store) in place of compute.
__global__ __launch_bounds__(kMNT) void k_mapfilter(MfIn in, uint64_t n, rfx_mf_program prog,
                                                    MfOut out, uint32_t ntiles,
                                                    uint64_t* __restrict__ n_kept,
                                                    uint64_t* __restrict__ status,
                                                    uint32_t* __restrict__ ticket) {
  mapfilter_tile<1, 1>(
      [&](uint64_t r) {
        return prog.n_filter == 0 || run_ops(prog, 0, prog.n_filter, in, out, r, 0) != 0;
      },
      [](uint64_t) { return Empty{}; },
      [&](uint64_t r, uint64_t pos, Empty) {
        run_ops(prog, prog.n_filter, prog.n_ops, in, out, r, pos);
      },
      n, ntiles, n_kept, status, ticket);
}

bool valid_col_type(int t) { return t >= RFX_MF_U8 && t <= RFX_MF_TUPLE16; }

bool valid_code(uint32_t c) {
  return c <= RFX_MF_LDX || (c >= RFX_MF_IADD && c <= RFX_MF_IMOD) ||
         (c >= RFX_MF_FADD && c <= RFX_MF_FDIV) || (c >= RFX_MF_I2F && c <= RFX_MF_F2I) ||
         (c >= RFX_MF_ILT && c <= RFX_MF_FNE) || (c >= RFX_MF_AND && c <= RFX_MF_NOT) ||
         (c >= RFX_MF_ADD_MONTH && c <= RFX_MF_LIKE) || c == RFX_MF_OUT;
}

// ---------------------------------------------------------------------------
// Compiled programs.
// ---------------------------------------------------------------------------

// What the run-time compiler sees: the device headers this library is
// built from (embedded at build time) and stand-ins for the system headers
// they name; hiprtc supplies the HIP device API and the fixed-width
// integer types (in __hip_internal) itself.
constexpr const char kStdintStub[] =
    "#pragma once\n"
    "using __hip_internal::int8_t; using __hip_internal::int16_t;\n"
    "using __hip_internal::int32_t; using __hip_internal::int64_t;\n"
    "using __hip_internal::uint8_t; using __hip_internal::uint16_t;\n"
    "using __hip_internal::uint32_t; using __hip_internal::uint64_t;\n"
    "#ifndef INT64_MAX\n#define INT64_MAX 9223372036854775807LL\n#endif\n"
    "#ifndef INT64_MIN\n#define INT64_MIN (-INT64_MAX - 1LL)\n#endif\n";

struct JitHeader {
  const char* name;
  const char* text;
};
const JitHeader kJitHeaders[] = {
#include "mf_jit_headers.inc"
    {"hip/hip_runtime.h", "#pragma once\n"},
    {"cstdint", kStdintStub},
    {"stdint.h", kStdintStub},
};

// Kernels compiled so far, by device and source.  Modules stay loaded for
// the life of the process (a launch may still be running one), so the cache
// is bounded: past kJitMax programs, new ones are interpreted.
constexpr size_t kJitMax = 1024;
std::mutex g_jit_mu;
std::unordered_map<std::string, hipFunction_t> g_jit_fns;

constexpr const char kJitName[] = "rfx_mf_jit";

// One statement per op.  In the output phase OUT ops copy their register
// into the row's value slot (store_out happens in the store step).
void jit_ops(std::string& s, const rfx_mf_program& p, uint32_t begin, uint32_t end,
             const rfx_mf_col* cols) {
  char line[160];
  uint32_t slot = 0;
  for (uint32_t pc = begin; pc < end; ++pc) {
    const rfx_mf_op& op = p.ops[pc];
    if (op.code == RFX_MF_NOP) continue;
    if (op.code == RFX_MF_LDC)
      std::snprintf(line, sizeof line, "    R[%u] = load_elem(in.p[%u], %d, row, 0x%llxull);\n",
                    op.d, op.a, cols[op.a].type, static_cast<unsigned long long>(op.imm));
    else if (op.code == RFX_MF_LDX)
      std::snprintf(line, sizeof line,
                    "    R[%u] = load_at(in.p[%u], %d, in.rows[%u], R[%u], 0x%llxull);\n", op.d,
                    op.a, cols[op.a].type, op.a, op.b, static_cast<unsigned long long>(op.imm));
    else if (op.code == RFX_MF_OUT)
      std::snprintf(line, sizeof line, "    o.v[%u] = R[%u];\n", slot++, op.a);
    else
      std::snprintf(line, sizeof line, "    R[%u] = mf_apply(%u, R[%u], R[%u], 0x%llxull);\n",
                    op.d, op.code, op.a, op.b, static_cast<unsigned long long>(op.imm));
    s += line;
  }
}

// One kernel for a program with its column and output types.  Short
// phases run their row batches unrolled (more loads in flight per wave);
// long ones, or outputs with many values per row, keep fewer copies.
std::string jit_source(const rfx_mf_program& p, const rfx_mf_col* cols, const rfx_mf_out* outs) {
  uint32_t n_out_ops = 0;
  for (uint32_t pc = p.n_filter; pc < p.n_ops; ++pc) n_out_ops += p.ops[pc].code == RFX_MF_OUT;
  const uint32_t n_emit = p.n_ops - p.n_filter;
  const int uf = p.n_filter <= 24 ? 8 : p.n_filter <= 48 ? 2 : 1;
  const int uo = (n_emit <= 24 && n_out_ops <= 4) ? 8 : (n_emit <= 48 && n_out_ops <= 8) ? 2 : 1;
  std::string s =
      "#include \"mapfilter_dev.h\"\n"
      "using namespace rfx;\n"
      "struct Vals { uint64_t v[" + std::to_string(n_out_ops ? n_out_ops : 1) + "]; };\n"
      "extern \"C\" __global__ __launch_bounds__(kMNT) void rfx_mf_jit(\n"
      "    MfIn in, MfOut out, uint64_t n, uint32_t ntiles, uint64_t* __restrict__ n_kept,\n"
      "    uint64_t* __restrict__ status, uint32_t* __restrict__ ticket) {\n"
      "  auto filter = [&](uint64_t row) -> bool {\n";
  if (p.n_filter == 0) {
    s += "    return true;\n";
  } else {
    s += "    uint64_t R[16] = {};\n";
    jit_ops(s, p, 0, p.n_filter, cols);
    s += "    return R[" + std::to_string(p.keep_reg) + "] != 0;\n";
  }
  s += "  };\n  auto compute = [&](uint64_t row) -> Vals {\n    Vals o{};\n    uint64_t R[16] = {};\n";
  jit_ops(s, p, p.n_filter, p.n_ops, cols);
  s += "    return o;\n  };\n  auto store = [&](uint64_t, uint64_t pos, const Vals& o) {\n";
  char line[128];
  uint32_t slot = 0;
  for (uint32_t pc = p.n_filter; pc < p.n_ops; ++pc) {
    const rfx_mf_op& op = p.ops[pc];
    if (op.code != RFX_MF_OUT) continue;
    std::snprintf(line, sizeof line, "    store_out(out.p[%u], %d, pos, o.v[%u]);\n", op.b,
                  outs[op.b].type, slot++);
    s += line;
  }
  s += "  };\n  mapfilter_tile<" + std::to_string(uf) + ", " + std::to_string(uo) +
       ">(filter, compute, store, n, ntiles, n_kept, status, ticket);\n}\n";
  return s;
}

#define RFX_RTC(call)                                                                  \
  do {                                                                                 \
    hiprtcResult rfx_r_ = (call);                                                      \
    if (rfx_r_ != HIPRTC_SUCCESS) {                                                    \
      hiprtcDestroyProgram(&prog);                                                     \
      return fail(RFX_EHIP, "rfx_mapfilter: %s: %s", #call, hiprtcGetErrorString(rfx_r_)); \
    }                                                                                  \
  } while (0)

// The compiled kernel for `src` on the current device, compiling it on
// first use.  Holds the cache lock while compiling, so concurrent first
// calls compile once.
int jit_function(const std::string& src, hipFunction_t* fn) {
  int device = 0;
  RFX_HIP(hipGetDevice(&device));
  const std::string key = std::to_string(device) + "\n" + src;
  std::lock_guard<std::mutex> lock(g_jit_mu);
  auto it = g_jit_fns.find(key);
  if (it != g_jit_fns.end()) {
    *fn = it->second;
    return RFX_OK;
  }
  if (g_jit_fns.size() >= kJitMax)
    return fail(RFX_ERANGE, "rfx_mapfilter: %zu compiled programs already (the most kept)",
                g_jit_fns.size());
  constexpr int nh = sizeof(kJitHeaders) / sizeof(kJitHeaders[0]);
  const char* names[nh];
  const char* texts[nh];
  for (int i = 0; i < nh; ++i) {
    names[i] = kJitHeaders[i].name;
    texts[i] = kJitHeaders[i].text;
  }
  hiprtcProgram prog = nullptr;
  {
    const hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "mapfilter_program.hip", nh,
                                               texts, names);
    if (r != HIPRTC_SUCCESS)
      return fail(RFX_EHIP, "rfx_mapfilter: hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, sizeof(opts) / sizeof(opts[0]), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      hiprtcGetProgramLog(prog, &log[0]);
      if (log.size() > 400) log.resize(400);
    }
    hiprtcDestroyProgram(&prog);
    return fail(RFX_EHIP, "rfx_mapfilter: compiling the program failed: %s", log.c_str());
  }
  size_t code_bytes = 0;
  RFX_RTC(hiprtcGetCodeSize(prog, &code_bytes));
  std::vector<char> code(code_bytes);
  RFX_RTC(hiprtcGetCode(prog, code.data()));
  hiprtcDestroyProgram(&prog);
  hipModule_t mod = nullptr;
  RFX_HIP(hipModuleLoadData(&mod, code.data()));
  hipFunction_t f = nullptr;
  const hipError_t e = hipModuleGetFunction(&f, mod, kJitName);
  if (e != hipSuccess) {
    (void)hipModuleUnload(mod);
    return fail(RFX_EHIP, "rfx_mapfilter: hipModuleGetFunction: %s", hipGetErrorString(e));
  }
  g_jit_fns.emplace(key, f);
  *fn = f;
  return RFX_OK;
}

bool jit_cached(const std::string& src) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(g_jit_mu);
  return g_jit_fns.count(std::to_string(device) + "\n" + src) != 0;
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
    in.rows[c] = cols[c].rows ? cols[c].rows : n_rows;
    if (in.rows[c] && !cols[c].data)
      return fail(RFX_EINVAL, "rfx_mapfilter: column %d has %llu rows and no data", c,
                  static_cast<unsigned long long>(in.rows[c]));
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
    if (op.code == RFX_MF_LDC || op.code == RFX_MF_LDX) {
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
    } else if (op.code == RFX_MF_SCHUNK && op.imm > 4095) {
      return fail(RFX_EINVAL, "rfx_mapfilter: op %u string chunk %llu", pc,
                  static_cast<unsigned long long>(op.imm));
    }
    if (op.d >= RFX_MF_REGS || op.b >= RFX_MF_REGS ||
        (op.code != RFX_MF_LDC && op.code != RFX_MF_LDX && op.a >= RFX_MF_REGS))
      return fail(RFX_EINVAL, "rfx_mapfilter: op %u register out of range", pc);
  }
  // Which kernel: mode 2 interprets; mode 1 must compile; the default
  // compiles and interprets only if that fails, or while a graph is being
  // captured and the program is not compiled yet (loading a module is not
  // a stream operation a graph can hold).
  hipFunction_t fn = nullptr;
  if (ctx->tune_mf_mode != 2) {
    const std::string src = jit_source(*prog, cols, outs);
    if (ctx->capturing && !jit_cached(src)) {
      if (ctx->tune_mf_mode == 1)
        return fail(RFX_EINVAL, "rfx_mapfilter: cannot compile a program during graph capture");
    } else {
      const int s = jit_function(src, &fn);
      if (s != RFX_OK && ctx->tune_mf_mode == 1) return s;
    }
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
  uint32_t nt = static_cast<uint32_t>(ntiles);
  if (fn) {
    RFX_PROF(ctx, "mapfilter");
    void* args[] = {&in, &out, &n_rows, &nt, &n_kept, &status, &ticket};
    RFX_HIP(hipModuleLaunchKernel(fn, nt, 1, 1, kMNT, 1, 1, 0, ctx->stream, args, nullptr));
    return RFX_OK;
  }
  RFX_PROF(ctx, "mapfilter");
  k_mapfilter<<<nt, kMNT, 0, ctx->stream>>>(in, n_rows, *prog, out, nt, n_kept, status, ticket);
  RFX_LAUNCHED();
  return RFX_OK;
}
