// serve.hip - the served form: a request loop over device-resident
// relations.
//
// Relations stay in HBM, registered once as schemas (one per plan kind).
// A request names a schema and its parameters.  In continuous mode each
// admitted request gets a slot (a stream with its own operator context) and
// one block of the HBM arena holding every intermediate, output and the
// operators' workspace; its whole plan is enqueued on the slot's stream
// with no host synchronisation (sizes stay in device cells), and an event
// marks its end.  The loop admits requests in arrival order whenever a slot
// is free and the arena can hold the head request, and polls the slots'
// events, so requests that arrive while others run join them on the device.
// Plain mode runs the same plans one request at a time the way Harmony runs
// an operator plan: after every operator the host synchronises and reads
// the size variables back before launching the next one (Runtime.cpp:99-118,
// 506-646, 1054-1314).
//
// Each request ends with a result row in device memory: the output row
// count and one digest per output column, sum over rows i of
// fmix64(w_i + fmix64(salt_c + i)) mod 2^64 (fmix64: the splitmix64
// finaliser, w_i the row's bits), so equal digests mean equal columns, order
// included.  An optional sink receives the output columns themselves.
#include <chrono>
#include <map>
#include <set>
#include <thread>
#include <vector>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kResWords = 8;  // rows, digest[3], flags, admit/done (host), spare
constexpr uint64_t kFlagJoinSize = 1;  // a join's size differs from its registered size

__device__ __forceinline__ uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Start of a request: clear its result row.
__global__ void k_result_clear(uint64_t* res) {
  if (threadIdx.x < kResWords) res[threadIdx.x] = 0;
}

// End of a request: res[0] = min(*n, cap).
__global__ void k_result_rows(uint64_t* res, const uint64_t* n, uint64_t cap) {
  if (threadIdx.x == 0) res[0] = *n < cap ? *n : cap;
}

// Raise `flag` in res[4] unless *total == expect.
__global__ void k_check_total(uint64_t* res, const uint64_t* total, uint64_t expect,
                              uint64_t flag) {
  if (threadIdx.x == 0 && *total != expect) res[4] |= flag;
}

template <int B>
__device__ __forceinline__ uint64_t row_bits(const void* col, uint64_t i) {
  if constexpr (B == 4) return static_cast<const uint32_t*>(col)[i];
  else if constexpr (B == 8) return static_cast<const uint64_t*>(col)[i];
  else {
    const uint64_t* p = static_cast<const uint64_t*>(col);
    return p[2 * i] ^ fmix64(p[2 * i + 1]);
  }
}

template <int B>
__global__ __launch_bounds__(256) void k_digest(const void* __restrict__ col,
                                                const uint64_t* __restrict__ rows, uint64_t salt,
                                                unsigned long long* __restrict__ out) {
  __shared__ uint64_t part[256];
  const uint64_t n = *rows;
  uint64_t h = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride)
    h += fmix64(row_bits<B>(col, i) + fmix64(salt + i));
  part[threadIdx.x] = h;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (static_cast<int>(threadIdx.x) < d) part[threadIdx.x] += part[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(out, static_cast<unsigned long long>(part[0]));
}

template <int B>
__global__ __launch_bounds__(256) void k_copy_rows(void* __restrict__ dst,
                                                   const void* __restrict__ src,
                                                   const uint64_t* __restrict__ rows) {
  using T = typename Elem<B>::T;
  const uint64_t n = *rows;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride)
    static_cast<T*>(dst)[i] = static_cast<const T*>(src)[i];
}

uint32_t grid_for(rfx_ctx* ctx, uint64_t cap) {
  uint64_t b = (cap + 255) / 256;
  const uint64_t m = static_cast<uint64_t>(ctx->device_cus) * 8;
  if (b > m) b = m;
  return static_cast<uint32_t>(b ? b : 1);
}

// Best-fit allocator over [0, cap) with coalescing; offsets, not pointers.
class Arena {
 public:
  void reset(uint64_t cap) {
    by_off_.clear();
    by_size_.clear();
    cap_ = cap;
    used_ = peak_ = 0;
    if (cap) insert(0, cap);
  }
  bool alloc(uint64_t size, uint64_t* off) {
    auto it = by_size_.lower_bound({size, 0});
    if (it == by_size_.end()) return false;
    const uint64_t o = it->second, s = it->first;
    erase(o, s);
    if (s > size) insert(o + size, s - size);
    *off = o;
    used_ += size;
    if (used_ > peak_) peak_ = used_;
    return true;
  }
  void release(uint64_t off, uint64_t size) {
    used_ -= size;
    uint64_t o = off, s = size;
    auto next = by_off_.lower_bound(o);
    if (next != by_off_.end() && next->first == o + s) {
      const uint64_t ns = next->second;
      erase(next->first, ns);
      s += ns;
    }
    auto after = by_off_.lower_bound(o);
    if (after != by_off_.begin()) {
      auto prev = std::prev(after);
      if (prev->first + prev->second == o) {
        const uint64_t po = prev->first, ps = prev->second;
        erase(po, ps);
        o = po;
        s += ps;
      }
    }
    insert(o, s);
  }
  uint64_t capacity() const { return cap_; }
  uint64_t peak() const { return peak_; }
  uint64_t used() const { return used_; }

 private:
  void insert(uint64_t o, uint64_t s) {
    by_off_[o] = s;
    by_size_.insert({s, o});
  }
  void erase(uint64_t o, uint64_t s) {
    by_off_.erase(o);
    by_size_.erase({s, o});
  }
  std::map<uint64_t, uint64_t> by_off_;
  std::set<std::pair<uint64_t, uint64_t>> by_size_;
  uint64_t cap_ = 0, used_ = 0, peak_ = 0;
};

struct Schema {
  int plan = 0;
  int nrel = 0;
  rfx_relation rel[3] = {};
  // Rows each join of the plan produces (join_agg: [0]; join3: the first
  // and the second join), measured once at registration: the relations
  // stay resident and unchanged, so every request's join has these sizes,
  // and its buffers are sized from them.
  uint64_t join_rows[2] = {0, 0};
};

struct Slot {
  rfx_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int64_t req = -1;
  uint64_t off = 0, size = 0;
};

}  // namespace
}  // namespace rfx

struct rfx_server {
  int device = 0;
  char* base = nullptr;
  rfx::Arena arena;
  std::vector<rfx::Slot> slots;
  std::vector<rfx::Schema> schemas;
};

namespace rfx {
namespace {

uint64_t now_ns(std::chrono::steady_clock::time_point t0) {
  return static_cast<uint64_t>(
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
          .count());
}

// Workspace bound for the operators other than the sort: tile status and
// partition arrays of at most 24 bytes per 256 rows, plus the fixed
// 1024-chunk reduction spine.
uint64_t small_ws(uint64_t rows) { return 32 * (rows / 256 + 16) + (256u << 10); }

// One request's plan.  With base == nullptr it only measures: *need gets
// the bytes of its arena block.  Otherwise the operators are enqueued on
// ctx's stream, with the block at base; res is the request's result row.
class Plan {
 public:
  Plan(const Schema& sc, const rfx_request& rq, rfx_ctx* ctx, char* base, bool plain,
       uint64_t* res)
      : sc_(sc), rq_(rq), ctx_(ctx), c_(base), base_(base), plain_(plain), res_(res) {}

  int run(uint64_t* need) {
    switch (sc_.plan) {
      case RFX_PLAN_SELECT_PROJECT: return select_project(need);
      case RFX_PLAN_JOIN_AGG: return join_agg(need);
      case RFX_PLAN_JOIN3: return join3(need);
      case RFX_PLAN_SET_OP: return set_op(need);
    }
    return fail(RFX_EINVAL, "serve: unknown plan %d", sc_.plan);
  }

 private:
  // Close the layout: the operators' workspace goes last.  Returns true
  // when measuring (the caller returns at once).
  bool layout_done(uint64_t ws, uint64_t* need) {
    char* w = c_.take<char>(ws);
    if (!base_) {
      *need = align_up(c_.off, 4096);
      return true;
    }
    ctx_->ws = w;
    ctx_->ws_bytes = ws;
    ctx_->ws_borrowed = true;
    return false;
  }
  // First launch of every plan: clear the request's result row.
  int begin() {
    k_result_clear<<<1, 64, 0, ctx_->stream>>>(res_);
    RFX_LAUNCHED();
    return RFX_OK;
  }
  // After an operator: plain mode waits for it (Harmony synchronises after
  // each launch).
  int step(int status) {
    RFX_TRY(status);
    if (plain_) RFX_HIP(hipStreamSynchronize(ctx_->stream));
    return RFX_OK;
  }
  // Plain mode reads a size variable back to the host.
  int read(const uint64_t* cell, uint64_t* v) {
    RFX_HIP(hipMemcpy(v, cell, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return RFX_OK;
  }
  int copy(void* dst, const void* src, uint64_t bytes) {
    return step(rfx_copy(ctx_, dst, src, bytes));
  }
  int finish(const uint64_t* rows_cell, uint64_t cap) {
    k_result_rows<<<1, 64, 0, ctx_->stream>>>(res_, rows_cell, cap);
    RFX_LAUNCHED();
    return RFX_OK;
  }
  // Plain mode: a size read back, bounded by the buffer it sizes.
  int read_capped(const uint64_t* cell, uint64_t cap, uint64_t* v) {
    RFX_TRY(read(cell, v));
    if (*v > cap) *v = cap;
    return RFX_OK;
  }
  template <int B>
  int digest_col(int c, const void* col, uint64_t cap) {
    k_digest<B><<<grid_for(ctx_, cap), 256, 0, ctx_->stream>>>(
        col, res_, 0x9E3779B97F4A7C15ull * (c + 1), reinterpret_cast<unsigned long long*>(res_ + 1 + c));
    RFX_LAUNCHED();
    if (rq_.sink) {
      char* dst = static_cast<char*>(rq_.sink) + sink_off_;
      sink_off_ += cap * B;
      if (sink_off_ > rq_.sink_bytes)
        return fail(RFX_EINVAL, "serve: sink of %llu bytes is too small",
                    static_cast<unsigned long long>(rq_.sink_bytes));
      k_copy_rows<B><<<grid_for(ctx_, cap), 256, 0, ctx_->stream>>>(dst, col, res_);
      RFX_LAUNCHED();
    }
    return step(RFX_OK);
  }
  int check_total(const uint64_t* total, uint64_t expect) {
    k_check_total<<<1, 64, 0, ctx_->stream>>>(res_, total, expect, kFlagJoinSize);
    RFX_LAUNCHED();
    return RFX_OK;
  }

  // SELECT id, val FROM T WHERE qty <op> constant, T packed 16-byte tuples
  // [127:96] id, [95:64] qty, [63:0] val; output 8-byte tuples (id, val
  // low 32 bits).
  int select_project(uint64_t* need) {
    const rfx_relation& t = sc_.rel[0];
    const uint64_t n = t.rows;
    void* sel = c_.take<u128>(n);
    uint64_t* nsel = c_.take<uint64_t>(1);
    uint64_t* out = c_.take<uint64_t>(n);
    if (layout_done(small_ws(n), need)) return RFX_OK;
    RFX_TRY(begin());
    const int op = rq_.param ? static_cast<int>(rq_.param) : RFX_CMP_LT;
    RFX_TRY(step(rfx_select(ctx_, sel, nsel, t.col[0], n, 16, 64, 32, op, rq_.constant)));
    const uint32_t ish[2] = {96, 0}, bits[2] = {32, 32}, osh[2] = {32, 0};
    if (plain_) {
      uint64_t m = 0;
      RFX_TRY(read_capped(nsel, n, &m));
      RFX_TRY(step(rfx_project(ctx_, out, 8, sel, 16, m, 2, ish, bits, osh)));
    } else {
      RFX_TRY(step(rfx_project_sized(ctx_, out, 8, sel, 16, nsel, n, 2, ish, bits, osh)));
    }
    RFX_TRY(finish(nsel, n));
    return digest_col<8>(0, out, n);
  }

  // SELECT grp, SUM(m), COUNT(*) FROM F JOIN D ON fk = dkey GROUP BY grp:
  // F (fk u32, m f64), D (dkey u32, grp u32 < 2^grp_bits); any key may
  // repeat on either side.
  int join_agg(uint64_t* need) {
    const rfx_relation& f = sc_.rel[0];
    const rfx_relation& d = sc_.rel[1];
    const uint64_t nf = f.rows, nd = d.rows;
    const uint32_t gbits = d.val_bits;  // bits of grp
    const uint64_t groups = 1ull << gbits;
    const uint64_t nj = sc_.join_rows[0];  // joined rows
    uint32_t* fk = c_.take<uint32_t>(nf);
    double* m = c_.take<double>(nf);
    uint32_t* dk = c_.take<uint32_t>(nd);
    uint32_t* dg = c_.take<uint32_t>(nd);
    int32_t* lb = c_.take<int32_t>(nf);
    int32_t* lc = c_.take<int32_t>(nf);
    uint64_t* tot = c_.take<uint64_t>(1);
    uint32_t* key2 = c_.take<uint32_t>(nj ? nj : 1);
    double* val = c_.take<double>(nj ? nj : 1);
    uint32_t* ok = c_.take<uint32_t>(groups);
    double* osum = c_.take<double>(groups);
    uint32_t* ok2 = c_.take<uint32_t>(groups);
    uint32_t* ocnt = c_.take<uint32_t>(groups);
    uint64_t* nout = c_.take<uint64_t>(1);
    uint64_t* nout2 = c_.take<uint64_t>(1);
    uint64_t ws = std::max<uint64_t>(
        std::max(rfx_sort_workspace(nf, 4, f.key_bits, 8), rfx_sort_workspace(nd, 4, d.key_bits, 4)),
        small_ws(2 * nf + nd + nj));
    ws = std::max(ws, rfx_sort_workspace(nj, 4, static_cast<int>(gbits), 8));
    if (layout_done(ws, need)) return RFX_OK;
    RFX_TRY(begin());
    RFX_TRY(copy(fk, f.col[0], 4 * nf));
    RFX_TRY(copy(m, f.col[1], 8 * nf));
    RFX_TRY(copy(dk, d.col[0], 4 * nd));
    RFX_TRY(copy(dg, d.col[1], 4 * nd));
    RFX_TRY(step(rfx_sort_pairs(ctx_, fk, m, nf, 4, f.key_bits, 8)));
    RFX_TRY(step(rfx_sort_pairs(ctx_, dk, dg, nd, 4, d.key_bits, 4)));
    RFX_TRY(step(rfx_join_find_bounds(ctx_, lb, lc, tot, fk, nf, dk, nd, 4)));
    uint64_t t = nj;
    if (plain_) RFX_TRY(read_capped(tot, nj, &t));
    RFX_TRY(check_total(tot, nj));
    const uint32_t fields[10] = {0, 2, 0, gbits, 0,   // grp -> key
                                 1, 1, 0, 64, 0};     // m -> value
    RFX_TRY(step(rfx_join_expand_project(ctx_, lb, lc, nf, tot, t, fk, 4, m, 8, dg, 4, key2, 4,
                                         val, 8, 2, fields)));
    RFX_TRY(step(rfx_sort_pairs(ctx_, key2, val, t, 4, static_cast<int>(gbits), 8)));
    if (plain_) {  // Harmony: set_reduce, then set_count, sizes read back
      uint64_t g = 0;
      RFX_TRY(step(rfx_reduce_by_key_f64(ctx_, ok, osum, nout, key2, val, tot, t, 4, RFX_AGG_SUM)));
      RFX_TRY(read(nout, &g));
      RFX_TRY(step(rfx_count_by_key(ctx_, ok2, ocnt, nout2, key2, tot, t, 4)));
      RFX_TRY(read(nout2, &g));
    } else {  // both aggregates from one reduce-by-key (the same bits)
      // The outputs hold `groups` rows; any run past them (possible only
      // when the join's size differs from the registered one, flagged
      // above) is counted but never written.
      RFX_TRY(step(rfx_reduce_count_by_key_f64(ctx_, ok, osum, ocnt, groups, nout, key2, val, tot,
                                               t, 4, RFX_AGG_SUM)));
    }
    RFX_TRY(finish(nout, groups));
    RFX_TRY(digest_col<4>(0, ok, groups));
    RFX_TRY(digest_col<8>(1, osum, groups));
    return digest_col<4>(2, ocnt, groups);
  }

  // Chain F ⋈ S ⋈ T on F.fk1 = S.skey and S.fk2 = T.tkey, output (T.tval,
  // F.m) in join order: F (fk1 u32, m f64), S (skey u32, fk2 u32), T (tkey
  // u32, tval u32), keys repeating on either side.  S.val_bits bounds fk2.
  int join3(uint64_t* need) {
    const rfx_relation& f = sc_.rel[0];
    const rfx_relation& s = sc_.rel[1];
    const rfx_relation& t = sc_.rel[2];
    const uint64_t nf = f.rows, ns = s.rows, nt = t.rows;
    const uint64_t j1 = sc_.join_rows[0], j2 = sc_.join_rows[1];  // rows of each join
    uint32_t* f1 = c_.take<uint32_t>(nf);
    double* fm = c_.take<double>(nf);
    uint32_t* sk = c_.take<uint32_t>(ns);
    uint32_t* sf = c_.take<uint32_t>(ns);
    int32_t* lb = c_.take<int32_t>(nf);
    int32_t* lc = c_.take<int32_t>(nf);
    uint64_t* tot = c_.take<uint64_t>(1);
    uint32_t* k2 = c_.take<uint32_t>(j1 ? j1 : 1);
    double* v2 = c_.take<double>(j1 ? j1 : 1);
    uint32_t* tk = c_.take<uint32_t>(nt);
    uint32_t* tv = c_.take<uint32_t>(nt);
    int32_t* lb2 = c_.take<int32_t>(j1 ? j1 : 1);
    int32_t* lc2 = c_.take<int32_t>(j1 ? j1 : 1);
    uint64_t* tot2 = c_.take<uint64_t>(1);
    uint32_t* ok = c_.take<uint32_t>(j2 ? j2 : 1);
    double* ov = c_.take<double>(j2 ? j2 : 1);
    uint64_t ws = std::max(rfx_sort_workspace(nf, 4, f.key_bits, 8),
                           rfx_sort_workspace(j1, 4, s.val_bits, 8));
    ws = std::max(ws, rfx_sort_workspace(ns, 4, s.key_bits, 4));
    ws = std::max(ws, rfx_sort_workspace(nt, 4, t.key_bits, 4));
    ws = std::max(ws, small_ws(nf + j1 + j2 + ns + nt));
    if (layout_done(ws, need)) return RFX_OK;
    RFX_TRY(begin());
    RFX_TRY(copy(f1, f.col[0], 4 * nf));
    RFX_TRY(copy(fm, f.col[1], 8 * nf));
    RFX_TRY(copy(sk, s.col[0], 4 * ns));
    RFX_TRY(copy(sf, s.col[1], 4 * ns));
    RFX_TRY(step(rfx_sort_pairs(ctx_, f1, fm, nf, 4, f.key_bits, 8)));
    RFX_TRY(step(rfx_sort_pairs(ctx_, sk, sf, ns, 4, s.key_bits, 4)));
    RFX_TRY(step(rfx_join_find_bounds(ctx_, lb, lc, tot, f1, nf, sk, ns, 4)));
    uint64_t n1 = j1;
    if (plain_) RFX_TRY(read_capped(tot, j1, &n1));
    RFX_TRY(check_total(tot, j1));
    const uint32_t fields1[10] = {0, 2, 0, 32, 0,   // S.fk2 -> key
                                  1, 1, 0, 64, 0};  // F.m -> value
    RFX_TRY(step(rfx_join_expand_project(ctx_, lb, lc, nf, tot, n1, f1, 4, fm, 8, sf, 4, k2, 4,
                                         v2, 8, 2, fields1)));
    RFX_TRY(step(rfx_sort_pairs(ctx_, k2, v2, n1, 4, s.val_bits, 8)));
    RFX_TRY(copy(tk, t.col[0], 4 * nt));
    RFX_TRY(copy(tv, t.col[1], 4 * nt));
    RFX_TRY(step(rfx_sort_pairs(ctx_, tk, tv, nt, 4, t.key_bits, 4)));
    RFX_TRY(step(rfx_join_find_bounds(ctx_, lb2, lc2, tot2, k2, n1, tk, nt, 4)));
    uint64_t n2 = j2;
    if (plain_) RFX_TRY(read_capped(tot2, j2, &n2));
    RFX_TRY(check_total(tot2, j2));
    const uint32_t fields2[10] = {0, 2, 0, 32, 0,   // T.tval -> key
                                  1, 1, 0, 64, 0};  // m -> value
    RFX_TRY(step(rfx_join_expand_project(ctx_, lb2, lc2, n1, tot2, n2, k2, 4, v2, 8, tv, 4, ok, 4,
                                         ov, 8, 2, fields2)));
    RFX_TRY(finish(tot2, j2));
    RFX_TRY(digest_col<4>(0, ok, j2));
    return digest_col<8>(1, ov, j2);
  }

  // A op B on sorted unique u64 sets (unsigned order); param = RFX_SET_*.
  int set_op(uint64_t* need) {
    const rfx_relation& a = sc_.rel[0];
    const rfx_relation& b = sc_.rel[1];
    const int op = static_cast<int>(rq_.param);
    const uint64_t cap = op == RFX_SET_UNION ? a.rows + b.rows : a.rows;
    uint64_t* out = c_.take<uint64_t>(cap ? cap : 1);
    uint64_t* nout = c_.take<uint64_t>(1);
    if (layout_done(small_ws(a.rows + b.rows), need)) return RFX_OK;
    RFX_TRY(begin());
    RFX_TRY(step(rfx_set_op(ctx_, op, out, nout, a.col[0], a.rows, b.col[0], b.rows, 8)));
    if (plain_) {
      uint64_t g = 0;
      RFX_TRY(read(nout, &g));
    }
    RFX_TRY(finish(nout, cap));
    return digest_col<8>(0, out, cap);
  }

  const Schema& sc_;
  const rfx_request& rq_;
  rfx_ctx* ctx_;
  Carve c_;
  char* base_;
  bool plain_;
  uint64_t* res_;
  uint64_t sink_off_ = 0;
};

int validate(const rfx_server* s, const rfx_request& rq) {
  if (rq.schema >= s->schemas.size())
    return fail(RFX_EINVAL, "serve: schema %u not registered", rq.schema);
  const Schema& sc = s->schemas[rq.schema];
  if (sc.plan == RFX_PLAN_SET_OP && rq.param > RFX_SET_INTERSECTION)
    return fail(RFX_EINVAL, "serve: set operation %u", rq.param);
  if (sc.plan == RFX_PLAN_SELECT_PROJECT && rq.param > RFX_CMP_SGE)
    return fail(RFX_EINVAL, "serve: comparison %u", rq.param);
  return RFX_OK;
}

// Device memory for one registration, freed on every path.
struct TempMem {
  std::vector<void*> p;
  ~TempMem() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int take(uint64_t n, T** out) {
    void* q = nullptr;
    if (hipMalloc(&q, (n ? n : 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(RFX_ENOMEM, "serve: %llu bytes to measure a join",
                  static_cast<unsigned long long>(n * sizeof(T)));
    }
    p.push_back(q);
    *out = static_cast<T*>(q);
    return RFX_OK;
  }
};

// Rows of each join of a join_agg or join3 plan over the schema's
// relations (the same sorts and find_bounds the plan runs, once, on a
// context of its own; for join3 the second join's left keys are S.fk2 of
// the first join's rows).
int measure_joins(int device, Schema* sc) {
  const rfx_relation& f = sc->rel[0];
  const rfx_relation& d = sc->rel[1];
  rfx_ctx* ctx = nullptr;
  RFX_TRY(rfx_ctx_create(&ctx, device, nullptr));
  struct Guard {
    rfx_ctx* c;
    ~Guard() { (void)rfx_ctx_destroy(c); }
  } guard{ctx};
  TempMem mem;
  uint32_t *fk, *dk, *dv;
  int32_t *lb, *lc;
  uint64_t* tot;
  RFX_TRY(mem.take(f.rows, &fk));
  RFX_TRY(mem.take(d.rows, &dk));
  RFX_TRY(mem.take(d.rows, &dv));
  RFX_TRY(mem.take(f.rows, &lb));
  RFX_TRY(mem.take(f.rows, &lc));
  RFX_TRY(mem.take(2, &tot));
  RFX_TRY(rfx_copy(ctx, fk, f.col[0], 4 * f.rows));
  RFX_TRY(rfx_copy(ctx, dk, d.col[0], 4 * d.rows));
  RFX_TRY(rfx_copy(ctx, dv, d.col[1], 4 * d.rows));
  RFX_TRY(rfx_sort_pairs(ctx, fk, nullptr, f.rows, 4, static_cast<int>(f.key_bits), 0));
  RFX_TRY(rfx_sort_pairs(ctx, dk, dv, d.rows, 4, static_cast<int>(d.key_bits), 4));
  RFX_TRY(rfx_join_find_bounds(ctx, lb, lc, tot, fk, f.rows, dk, d.rows, 4));
  RFX_TRY(rfx_ctx_synchronize(ctx));
  uint64_t t1 = 0;
  RFX_HIP(hipMemcpy(&t1, tot, 8, hipMemcpyDeviceToHost));
  sc->join_rows[0] = t1;
  if (sc->plan != RFX_PLAN_JOIN3) return RFX_OK;
  if (t1 >= (1ull << 30))
    return fail(RFX_ERANGE, "serve: the first join makes %llu rows (max 2^30-1)",
                static_cast<unsigned long long>(t1));
  const rfx_relation& t = sc->rel[2];
  uint32_t *k2, *tk;
  int32_t *lb2, *lc2;
  RFX_TRY(mem.take(t1, &k2));
  RFX_TRY(mem.take(t.rows, &tk));
  RFX_TRY(mem.take(t1, &lb2));
  RFX_TRY(mem.take(t1, &lc2));
  const uint32_t field[5] = {0, 2, 0, 32, 0};  // S.fk2 of each joined row
  RFX_TRY(rfx_join_expand_project(ctx, lb, lc, f.rows, tot, t1, fk, 4, nullptr, 0, dv, 4, k2, 4,
                                  nullptr, 0, 1, field));
  RFX_TRY(rfx_sort_pairs(ctx, k2, nullptr, t1, 4, static_cast<int>(d.val_bits), 0));
  RFX_TRY(rfx_copy(ctx, tk, t.col[0], 4 * t.rows));
  RFX_TRY(rfx_sort_pairs(ctx, tk, nullptr, t.rows, 4, static_cast<int>(t.key_bits), 0));
  RFX_TRY(rfx_join_find_bounds(ctx, lb2, lc2, tot + 1, k2, t1, tk, t.rows, 4));
  RFX_TRY(rfx_ctx_synchronize(ctx));
  uint64_t t2 = 0;
  RFX_HIP(hipMemcpy(&t2, tot + 1, 8, hipMemcpyDeviceToHost));
  if (t2 >= (1ull << 30))
    return fail(RFX_ERANGE, "serve: the second join makes %llu rows (max 2^30-1)",
                static_cast<unsigned long long>(t2));
  sc->join_rows[1] = t2;
  return RFX_OK;
}

int release_slot(rfx_server* s, Slot& sl) {
  s->arena.release(sl.off, sl.size);
  sl.req = -1;
  sl.ctx->ws = nullptr;
  sl.ctx->ws_bytes = 0;
  return RFX_OK;
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" {

int rfx_server_create(rfx_server** out, int device, int slots, uint64_t budget_bytes) {
  if (!out || slots < 1 || slots > 64 || budget_bytes < (1u << 20))
    return fail(RFX_EINVAL, "rfx_server_create: bad arguments (%d slots, %llu bytes)", slots,
                static_cast<unsigned long long>(budget_bytes));
  RFX_HIP(hipSetDevice(device));
  auto* s = new rfx_server;
  s->device = device;
  const uint64_t bytes = align_up(budget_bytes, 1u << 20);
  if (hipMalloc(reinterpret_cast<void**>(&s->base), bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return fail(RFX_ENOMEM, "rfx_server_create: arena of %llu bytes",
                static_cast<unsigned long long>(bytes));
  }
  s->arena.reset(bytes);
  s->slots.resize(slots);
  for (Slot& sl : s->slots) {
    int st = RFX_OK;
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      st = fail(RFX_EHIP, "rfx_server_create: stream or event creation failed");
    if (st == RFX_OK) st = rfx_ctx_create(&sl.ctx, device, sl.stream);
    if (st != RFX_OK) {
      rfx_server_destroy(s);
      return st;
    }
  }
  *out = s;
  return RFX_OK;
}

int rfx_server_destroy(rfx_server* s) {
  if (!s) return RFX_OK;
  (void)hipSetDevice(s->device);
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.ctx) {
      sl.ctx->ws = nullptr;
      sl.ctx->ws_bytes = 0;
      (void)rfx_ctx_destroy(sl.ctx);
    }
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  if (s->base) (void)hipFree(s->base);
  delete s;
  return RFX_OK;
}

int rfx_server_add_schema(rfx_server* s, int plan, const rfx_relation* rels, int nrels,
                          uint32_t* schema_id) {
  if (!s || !rels || !schema_id) return fail(RFX_EINVAL, "rfx_server_add_schema: null argument");
  static const int want[4] = {1, 2, 3, 2};
  if (plan < 0 || plan > RFX_PLAN_SET_OP || nrels != want[plan])
    return fail(RFX_EINVAL, "rfx_server_add_schema: plan %d takes %d relations, got %d", plan,
                plan >= 0 && plan <= 3 ? want[plan] : 0, nrels);
  Schema sc;
  sc.plan = plan;
  sc.nrel = nrels;
  for (int i = 0; i < nrels; ++i) {
    sc.rel[i] = rels[i];
    if (rels[i].rows && !rels[i].col[0])
      return fail(RFX_EINVAL, "rfx_server_add_schema: relation %d has no key column", i);
    if (rels[i].rows > (1ull << 30) - 1)
      return fail(RFX_ERANGE, "rfx_server_add_schema: relation %d has %llu rows (max 2^30-1)",
                  i, static_cast<unsigned long long>(rels[i].rows));
    if (rels[i].key_bits > 32 || rels[i].val_bits > 32)
      return fail(RFX_EINVAL, "rfx_server_add_schema: key_bits %u / val_bits %u",
                  rels[i].key_bits, rels[i].val_bits);
  }
  if ((plan == RFX_PLAN_JOIN_AGG || plan == RFX_PLAN_JOIN3) &&
      (!rels[0].col[1] || !rels[1].col[1] || (plan == RFX_PLAN_JOIN3 && !rels[2].col[1])))
    return fail(RFX_EINVAL, "rfx_server_add_schema: join relations need a value column");
  if (plan == RFX_PLAN_JOIN_AGG || plan == RFX_PLAN_JOIN3) {
    for (int i = 0; i < nrels; ++i)
      if (rels[i].key_bits < 1)
        return fail(RFX_EINVAL, "rfx_server_add_schema: relation %d needs key_bits", i);
    if (plan == RFX_PLAN_JOIN3 && rels[1].val_bits < 1)
      return fail(RFX_EINVAL, "rfx_server_add_schema: S needs val_bits (bits of fk2)");
  }
  if (plan == RFX_PLAN_JOIN_AGG && (rels[1].val_bits < 1 || rels[1].val_bits > 16))
    return fail(RFX_EINVAL, "rfx_server_add_schema: group column of %u bits (1..16)",
                rels[1].val_bits);
  if (plan == RFX_PLAN_JOIN_AGG || plan == RFX_PLAN_JOIN3)
    RFX_TRY(measure_joins(s->device, &sc));
  *schema_id = static_cast<uint32_t>(s->schemas.size());
  s->schemas.push_back(sc);
  return RFX_OK;
}

int rfx_server_request_bytes(rfx_server* s, const rfx_request* rq, uint64_t* bytes) {
  if (!s || !rq || !bytes) return fail(RFX_EINVAL, "rfx_server_request_bytes: null argument");
  RFX_TRY(validate(s, *rq));
  Plan p(s->schemas[rq->schema], *rq, nullptr, nullptr, false, nullptr);
  return p.run(bytes);
}

int rfx_server_run(rfx_server* s, rfx_request* reqs, uint64_t n, int mode, rfx_serve_stats* st) {
  if (!s || (n && !reqs)) return fail(RFX_EINVAL, "rfx_server_run: null argument");
  if (mode != RFX_SERVE_PLAIN && mode != RFX_SERVE_CONTINUOUS)
    return fail(RFX_EINVAL, "rfx_server_run: mode %d", mode);
  RFX_HIP(hipSetDevice(s->device));
  for (uint64_t i = 0; i < n; ++i) {
    RFX_TRY(validate(s, reqs[i]));
    if (i && reqs[i].arrival_ns < reqs[i - 1].arrival_ns)
      return fail(RFX_EINVAL, "rfx_server_run: arrivals must not decrease (request %llu)",
                  static_cast<unsigned long long>(i));
    RFX_TRY(rfx_server_request_bytes(s, &reqs[i], &reqs[i].mem_bytes));
    if (reqs[i].mem_bytes > s->arena.capacity())
      return fail(RFX_EBUDGET, "rfx_server_run: request %llu needs %llu bytes, budget %llu",
                  static_cast<unsigned long long>(i),
                  static_cast<unsigned long long>(reqs[i].mem_bytes),
                  static_cast<unsigned long long>(s->arena.capacity()));
    reqs[i].admit_ns = reqs[i].done_ns = 0;
    reqs[i].status = RFX_OK;
  }
  uint64_t* res = nullptr;
  RFX_HIP(hipMalloc(reinterpret_cast<void**>(&res), (n ? n : 1) * kResWords * sizeof(uint64_t)));
  s->arena.reset(s->arena.capacity());
  const int nslots = mode == RFX_SERVE_PLAIN ? 1 : static_cast<int>(s->slots.size());
  uint64_t next = 0, finished = 0, blocked = 0;
  int inflight = 0, peak_inflight = 0;
  int status = RFX_OK;
  const auto t0 = std::chrono::steady_clock::now();
  while (finished < n && status == RFX_OK) {
    bool progress = false;
    // Completions.
    for (int k = 0; k < nslots && status == RFX_OK; ++k) {
      Slot& sl = s->slots[k];
      if (sl.req < 0) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) {
        status = fail(RFX_EHIP, "rfx_server_run: request %lld failed: %s",
                      static_cast<long long>(sl.req), hipGetErrorString(q));
        break;
      }
      reqs[sl.req].done_ns = now_ns(t0);
      release_slot(s, sl);
      ++finished;
      --inflight;
      progress = true;
    }
    // Admissions, in arrival order.
    while (status == RFX_OK && next < n && reqs[next].arrival_ns <= now_ns(t0)) {
      int k = 0;
      while (k < nslots && s->slots[k].req >= 0) ++k;
      if (k == nslots) break;
      rfx_request& rq = reqs[next];
      uint64_t off = 0;
      if (!s->arena.alloc(rq.mem_bytes, &off)) {
        ++blocked;
        break;
      }
      Slot& sl = s->slots[k];
      sl.req = static_cast<int64_t>(next);
      sl.off = off;
      sl.size = rq.mem_bytes;
      rq.admit_ns = now_ns(t0);
      Plan p(s->schemas[rq.schema], rq, sl.ctx, s->base + off, mode == RFX_SERVE_PLAIN,
             res + next * kResWords);
      uint64_t unused = 0;
      status = p.run(&unused);
      if (status == RFX_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess)
        status = fail(RFX_EHIP, "rfx_server_run: event record failed");
      rq.status = status;
      ++next;
      ++inflight;
      if (inflight > peak_inflight) peak_inflight = inflight;
      progress = true;
    }
    if (!progress) std::this_thread::yield();
  }
  const uint64_t wall = now_ns(t0);
  for (int k = 0; k < static_cast<int>(s->slots.size()); ++k) {
    Slot& sl = s->slots[k];
    (void)hipStreamSynchronize(sl.stream);
    if (sl.req >= 0) release_slot(s, sl);
  }
  std::vector<uint64_t> host(static_cast<size_t>((n ? n : 1) * kResWords));
  const hipError_t ce =
      hipMemcpy(host.data(), res, host.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
  (void)hipFree(res);
  if (status != RFX_OK) return status;
  if (ce != hipSuccess) return fail(RFX_EHIP, "rfx_server_run: result copy failed");
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t* r = &host[i * kResWords];
    reqs[i].rows = r[0];
    for (int c = 0; c < 3; ++c) reqs[i].digest[c] = r[1 + c];
    if (r[4] & kFlagJoinSize) reqs[i].status = RFX_EINVAL;
  }
  if (st) {
    st->wall_ns = wall;
    st->peak_bytes = s->arena.peak();
    st->budget_bytes = s->arena.capacity();
    st->peak_inflight = static_cast<uint32_t>(peak_inflight);
    st->slots = static_cast<uint32_t>(nslots);
    st->blocked_on_memory = blocked;
  }
  return RFX_OK;
}

}  // extern "C"
