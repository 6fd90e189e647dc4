// rfx_internal.h - host-side internals shared by the librfx_hip sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/rfx_ra.h"

struct rfx_prof_rec {
  const char* tag;
  hipEvent_t start, stop;
};

struct rfx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* ws = nullptr;         // cached device workspace
  uint64_t ws_bytes = 0;
  int device_cus = 256;       // compute units, for persistent grids
  // Per-launch timing (off unless rfx_ctx_profile enabled it).
  bool profiling = false;
  bool capturing = false;     // inside rfx_graph_begin/end: no allocation
  bool ws_borrowed = false;   // ws belongs to the caller: never grown or freed
  int live_graphs = 0;        // instantiated graphs that hold pointers into ws
  uint64_t* cells = nullptr;  // kCtxCells device u64 cells for scalar operands
  // Tile shapes chosen by rfx_ctx_tune (0: the built-in default).
  int tune_sort_ipt = 0;      // keys per thread of the onesweep tile
  int tune_sort_digit = 0;    // radix digit bits (0: chosen from the key bits)
  int tune_sort_status = 0;   // 64: 64-bit look-back status words at any size
  int tune_fb_tile = 0;       // merged keys per find_bounds tile
  int tune_ex_tile = 0;       // merged items per expansion tile
  int tune_sel_vt = 0;        // rows per thread of the select pass
  int tune_mf_mode = 0;       // MAPFILTER: 0 compiled (else interpreted), 1 compiled, 2 interpreted
  std::vector<rfx_prof_rec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  std::map<std::string, std::pair<uint64_t, double>> prof_stats;  // launches, ms
};

namespace rfx {

// Record an error message for rfx_last_error() and return `code`.
int fail(int code, const char* fmt, ...);

// Check a HIP call; on failure record it and return RFX_EHIP.
#define RFX_HIP(call)                                                   \
  do {                                                                  \
    hipError_t rfx_e_ = (call);                                         \
    if (rfx_e_ != hipSuccess)                                           \
      return ::rfx::fail(RFX_EHIP, "%s failed: %s (%s:%d)", #call,      \
                         hipGetErrorString(rfx_e_), __FILE__, __LINE__); \
  } while (0)

#define RFX_TRY(expr)            \
  do {                           \
    int rfx_s_ = (expr);         \
    if (rfx_s_ != RFX_OK) return rfx_s_; \
  } while (0)

// Check the launch that was just issued.
#define RFX_LAUNCHED() RFX_HIP(hipGetLastError())

inline bool valid_width(int b) {
  return b == 1 || b == 2 || b == 4 || b == 8 || b == 16;
}

constexpr uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Make sure ctx->ws holds `bytes` (synchronises the stream before growing).
// Refused while graphs captured on the context are alive.
int ensure_workspace(rfx_ctx* ctx, uint64_t bytes);
// Drop the graph-ownership records of a context being destroyed.
void forget_graphs_of(rfx_ctx* ctx);

// A bump allocator over a workspace region.  Constructed with a null base
// it only measures: `off` ends as the bytes the same sequence needs.
struct Carve {
  uintptr_t base;
  uint64_t off = 0;
  explicit Carve(void* b) : base(reinterpret_cast<uintptr_t>(b)) {}
  template <class T>
  T* take(uint64_t count) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

// The context's device u64 cells (kCtxCells of them, allocated on first
// use and freed with the context): scalar operands that the operators read
// on the device, e.g. a row count the Harmony convention passes by value.
constexpr int kCtxCells = 8;
int ctx_cells(rfx_ctx* ctx, uint64_t** cells);

// Zero `words` u32 words at `p` on the context stream with a fill kernel
// (graph-capture safe, unlike hipMemsetAsync).
int fill_u32(rfx_ctx* ctx, uint32_t* p, uint64_t words, uint32_t value);

// Type-code helpers (RelationalAlgebraKernel.h:81-95).
inline int bytes_of_type(uint64_t type) {
  switch (type) {
    case RFX_I8: return 1;
    case RFX_I16: return 2;
    case RFX_I32: return 4;
    case RFX_I64: return 8;
    case RFX_I128: return 16;
    case RFX_F32: return 4;
    case RFX_F64: return 8;
    default: return 0;
  }
}
// Clamp<bits>/8 (Tuple.h:30-37).
inline int bytes_of_bits(uint64_t bits) {
  if (bits <= 8) return 1;
  if (bits <= 16) return 2;
  if (bits <= 32) return 4;
  if (bits <= 64) return 8;
  if (bits <= 128) return 16;
  return 0;
}

// Reduce-by-key with size variables: rows = *n_ptr / n_div (at most
// n_cap); writes runs * out_a_mul to *out_a and runs * out_b_mul to *out_b.
// val_kind: 0 f64 (op = RFX_AGG_*), 1 u32 sum, 2 count (u32 counts to
// out_vals), 3 f64 with the row count of each run also to out_counts.
// Runs past out_cap are counted but not written.
int reduce_by_key_sized(rfx_ctx* ctx, int op, int val_kind, int kb, void* out_keys,
                        void* out_vals, const void* keys, const void* vals,
                        const uint64_t* n_ptr, uint64_t n_div, uint64_t n_cap, uint64_t* out_a,
                        uint64_t out_a_mul, uint64_t* out_b, uint64_t out_b_mul,
                        uint32_t* out_counts, uint64_t out_cap);
// kind 0: f64 sum, 1: u32 sum, 2: max of u64 bit patterns.
int single_reduce(rfx_ctx* ctx, int kind, void* result, const void* v, uint64_t n);

// Bracket a kernel launch with events when the context is profiling.
void prof_begin(rfx_ctx* ctx, const char* tag);
void prof_end(rfx_ctx* ctx);
struct ProfScope {
  rfx_ctx* c;
  ProfScope(rfx_ctx* ctx, const char* tag) : c(ctx) {
    if (c->profiling) prof_begin(c, tag);
  }
  ~ProfScope() {
    if (c->profiling) prof_end(c);
  }
};
#define RFX_PROF(ctx, tag) ::rfx::ProfScope rfx_prof_scope_##__LINE__((ctx), (tag))

// Abort with the message, as the reference's check() does.
[[noreturn]] void die(const char* where, int status);

}  // namespace rfx
