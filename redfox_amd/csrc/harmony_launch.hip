// harmony_launch.hip - rfx_harmony_launch(): one Harmony kernel record
// (family name, kernel type, `code` bytes holding the instance descriptor,
// operand slots) run on the native operators.
//
// Operand slots follow the reference runtime exactly:
// * ComputeKernel (PTX) records: slots 0-2 are the values of the launch's
//   ctaCount, threadCount and sharedBytes variables, every later slot a
//   device address, sizes included (PTXKernelLaunch::execute,
//   Runtime.cpp:506-646);
// * BinaryKernel records: the family's pointer slots are device addresses
//   and the others the first 8 bytes of their variables
//   (BinaryKernelLaunch::execute, Runtime.cpp:672-1038).
// What the generated source baked into each instance (tuple widths, key
// fields, the reduction, the comparison) comes from the descriptor
// (harmony_desc.h).  Sizes in size variables are byte counts, in and out,
// as the reference's kernels read and write them.  Every call finishes
// before it returns; failures are returned, not aborted on.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "harmony_desc.h"
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
int set_op_sized(rfx_ctx* ctx, int op, void* out, uint64_t* n_out, uint64_t mul, const void* a,
                 uint64_t na, const void* b, uint64_t nb, int eb);
int unique_sized(rfx_ctx* ctx, void* out, uint64_t* n_out, uint64_t mul, const void* in,
                 uint64_t n, int eb);
namespace {

// ---- size arithmetic on the device ---------------------------------------
// *out = (*in / div) * mul, or (*in * mul) / div when mul_first.
__global__ void k_size_scale(uint64_t* out, const uint64_t* in, uint64_t div, uint64_t mul,
                             int mul_first) {
  const uint64_t v = *in;
  *out = mul_first ? v * mul / div : v / div * mul;
}
__global__ void k_size_sum(uint64_t* out, const uint64_t* a, const uint64_t* b) {
  *out = *a + (b ? *b : 0);
}
// join_get_result_size: *index = total * 4, *result = total * out_bytes.
__global__ void k_join_sizes(uint64_t* index, uint64_t* result, uint64_t out_bytes) {
  const uint64_t t = *result;
  *index = t * 4;
  *result = t * out_bytes;
}
// Exclusive scan of h[0, n) in place (n <= 1024, one block).
__global__ __launch_bounds__(1024) void k_excl_scan_u32(uint32_t* h, uint32_t n) {
  __shared__ uint32_t tmp[1024 / kWave + 1];
  const uint32_t x = threadIdx.x < n ? h[threadIdx.x] : 0u;
  uint32_t total;
  const uint32_t e = block_excl_scan<1024>(x, tmp, &total);
  if (threadIdx.x < n) h[threadIdx.x] = e;
}
// A compaction whose rows all sit at the front of the temp array: one
// partition holding `kept` rows, the rest empty (entries [0, parts] set).
__global__ void k_select_hist(uint32_t* hist, const uint64_t* kept, uint32_t parts) {
  for (uint32_t i = threadIdx.x; i <= parts; i += blockDim.x)
    hist[i] = i == 0 ? static_cast<uint32_t>(*kept) : 0u;
}
// select_gather (Select.h gather): partition p's rows begin at tile
// offset range(p) * nv of temp and go to dest[hist[p], hist[p+1]).
__global__ void k_select_gather(unsigned char* dest, const unsigned char* temp, uint64_t count,
                                const uint32_t* hist, int w, uint32_t parts, uint32_t nv) {
  const uint32_t p = blockIdx.x;
  const uint64_t tiles = (count + nv - 1) / nv;
  const uint64_t tx = tiles / parts, ty = tiles - tx * parts;
  const uint64_t first = (tx * p + (p < ty ? p : ty)) * nv;
  const uint32_t b = hist[p], e = hist[p + 1];
  for (uint32_t i = threadIdx.x; i < e - b; i += blockDim.x)
    for (int k = 0; k < w; ++k) dest[(static_cast<uint64_t>(b) + i) * w + k] = temp[(first + i) * w + k];
}
// out[i] = in[rows[i]] for w-byte elements.
__global__ void k_permute_rows(unsigned char* out, const unsigned char* in, const uint32_t* rows,
                               uint64_t n, int w) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const uint64_t r = rows[i];
    for (int k = 0; k < w; ++k) out[i * w + k] = in[r * w + k];
  }
}
// The key of each run of equal strings is its first row's address
// (reduce_by_key keeps the first key of a run).
__global__ void k_first_of_run(uint64_t* out, const uint64_t* addr, const uint32_t* group,
                               uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    if (i == 0 || group[i] != group[i - 1]) out[group[i]] = addr[i];
}

uint32_t grid_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return static_cast<uint32_t>(b ? b : 1);
}

int read_u64(const uint64_t* dev, uint64_t* v) {
  RFX_HIP(hipMemcpy(v, dev, 8, hipMemcpyDeviceToHost));
  return RFX_OK;
}
int write_u64(uint64_t* dev, uint64_t v) {
  RFX_HIP(hipMemcpy(dev, &v, 8, hipMemcpyHostToDevice));
  return RFX_OK;
}

// Scratch device memory for the duration of one call (the reference's
// kernels allocate and free theirs inside the call as well).
struct Scratch {
  void* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  int alloc(uint64_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(RFX_ENOMEM, "harmony: %llu bytes of scratch", static_cast<unsigned long long>(bytes));
    }
    return RFX_OK;
  }
};

struct Call {
  rfx_ctx* ctx;
  const char* name;
  const KernelDesc& d;
  bool have_desc;
  const uint64_t* a;  // operand slots after the launch dimensions
  uint32_t n;
  uint64_t dims[3];   // ctaCount, threadCount, sharedBytes (PTX records)

  void* P(uint32_t i) const { return reinterpret_cast<void*>(a[i]); }
  uint64_t* S(uint32_t i) const { return reinterpret_cast<uint64_t*>(a[i]); }
  int need(uint32_t lo, uint32_t hi) const {
    return n >= lo && n <= hi ? RFX_OK
                              : fail(RFX_EINVAL, "%s: %u operands, expected %u..%u", name, n, lo, hi);
  }
  int want(bool ok, const char* what) const {
    return ok ? RFX_OK : fail(RFX_EINVAL, "%s: the descriptor lacks %s", name, what);
  }
};

int launched() {
  RFX_LAUNCHED();
  return RFX_OK;
}

int bytes_checked(const Call& c, const TupleDesc& t, const char* what, int* out) {
  *out = t.bytes();
  if (!*out) return fail(RFX_ENOTSUP, "%s: %s tuple of %u bits is wider than 128", c.name, what, t.total());
  return RFX_OK;
}

// ---- PTX families ---------------------------------------------------------
// split_get_result_size(size, key_size, value_size): d key, d1 value, a source
int ptx_split_size(const Call& c) {
  RFX_TRY(c.need(3, 3));
  RFX_TRY(c.want(c.d.has_d && c.d.has_d1 && c.d.has_a, "d, d1 and a"));
  int kb, vb, sb;
  RFX_TRY(bytes_checked(c, c.d.d, "key", &kb));
  RFX_TRY(bytes_checked(c, c.d.d1, "value", &vb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(1), c.S(0), sb, kb, 0);
  RFX_TRY(launched());
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(2), c.S(0), sb, vb, 0);
  return launched();
}
// split_key_get_result_size(key_size, size): d key, a source
int ptx_split_key_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a, "d and a"));
  int kb, sb;
  RFX_TRY(bytes_checked(c, c.d.d, "key", &kb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), sb, kb, 1);
  return launched();
}
// The first `dom` fields of the source are the key: key = t >> (bits of
// the other fields), value = those bits (Tuple.h stripValues/stripKeys).
int split_shift(const Call& c, uint32_t* shift) {
  RFX_TRY(c.want(c.d.has_a && c.d.dom >= 1 && c.d.dom <= c.d.a.n, "a and dom"));
  *shift = c.d.a.bits_after(c.d.dom - 1);
  return RFX_OK;
}
// set_split(key, value, begin, size): d key, d1 value, a source, dom
int ptx_split(const Call& c) {
  RFX_TRY(c.need(4, 4));
  RFX_TRY(c.want(c.d.has_d && c.d.has_d1, "d and d1"));
  int kb, vb, sb;
  uint32_t sh;
  RFX_TRY(split_shift(c, &sh));
  RFX_TRY(bytes_checked(c, c.d.d, "key", &kb));
  RFX_TRY(bytes_checked(c, c.d.d1, "value", &vb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(3), &bytes));
  return rfx_split(c.ctx, c.P(0), kb, c.P(1), vb, c.P(2), sb, bytes / sb, sh);
}
// set_splitkey(key, begin, size): d key, a source, dom
int ptx_split_key(const Call& c) {
  RFX_TRY(c.need(3, 3));
  RFX_TRY(c.want(c.d.has_d, "d"));
  int kb, sb;
  uint32_t sh;
  RFX_TRY(split_shift(c, &sh));
  RFX_TRY(bytes_checked(c, c.d.d, "key", &kb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(2), &bytes));
  return rfx_split(c.ctx, c.P(0), kb, nullptr, 0, c.P(1), sb, bytes / sb, sh);
}
// merge_get_result_size(size, value_size): d result, a value
int ptx_merge_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a, "d and a"));
  int rb, vb;
  RFX_TRY(bytes_checked(c, c.d.d, "result", &rb));
  RFX_TRY(bytes_checked(c, c.d.a, "value", &vb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), vb, rb, 0);
  return launched();
}
// set_merge(result, key, key_size, value, value_size): d result, a key,
// b value, dom; result = (key << bits after the key fields) | value, over
// the value rows (Split_Merge.h merge).
int ptx_merge(const Call& c) {
  RFX_TRY(c.need(5, 5));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a && c.d.has_b && c.d.dom >= 1 && c.d.dom <= c.d.d.n,
                 "d, a, b and dom"));
  int rb, kb, vb;
  RFX_TRY(bytes_checked(c, c.d.d, "result", &rb));
  RFX_TRY(bytes_checked(c, c.d.a, "key", &kb));
  RFX_TRY(bytes_checked(c, c.d.b, "value", &vb));
  uint64_t vbytes;
  RFX_TRY(read_u64(c.S(4), &vbytes));
  return rfx_merge(c.ctx, c.P(0), rb, c.P(1), kb, c.P(3), vb, vbytes / vb,
                   c.d.d.bits_after(c.d.dom - 1));
}
// generate_get_result_size(generate_size, size): a source
int ptx_generate_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_a, "a"));
  int sb;
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), sb, 4, 0);
  return launched();
}
// set_generate(begin, size)
int ptx_generate(const Call& c) {
  RFX_TRY(c.need(2, 2));
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(1), &bytes));
  return rfx_generate_ones(c.ctx, static_cast<uint32_t*>(c.P(0)), bytes / 4);
}
// join_get_temp_size(lower_bound_size, left_size): d left tuple
int ptx_join_temp_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_d, "d"));
  int lb;
  RFX_TRY(bytes_checked(c, c.d.d, "left", &lb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), lb, 4, 0);
  return launched();
}
// join_get_result_size(index_size, result_size): d output tuple
int ptx_join_result_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_d, "d"));
  int ob;
  RFX_TRY(bytes_checked(c, c.d.d, "output", &ob));
  k_join_sizes<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), ob);
  return launched();
}
// mgpu_join_gather(result, left_key, left_index, right_key, right_index,
// size, [left_value], [right_value]): d output, a left key, b right key,
// c left value, e right value, dom key fields (ModernGPUJoin.h:43-200).
int ptx_join_gather(const Call& c) {
  RFX_TRY(c.need(6, 8));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a && c.d.dom >= 1 && c.d.dom <= c.d.d.n, "d, a and dom"));
  const bool lv = c.d.c.n > 0, rv = c.d.e.n > 0;
  const uint32_t want_n = 6 + lv + rv;
  if (c.n != want_n)
    return fail(RFX_EINVAL, "%s: %u operands for %s left and %s right values", c.name, c.n,
                lv ? "with" : "no", rv ? "with" : "no");
  int ob, kb, lvb = 0, rvb = 0;
  RFX_TRY(bytes_checked(c, c.d.d, "output", &ob));
  RFX_TRY(bytes_checked(c, c.d.a, "left key", &kb));
  if (lv) RFX_TRY(bytes_checked(c, c.d.c, "left value", &lvb));
  if (rv) RFX_TRY(bytes_checked(c, c.d.e, "right value", &rvb));
  // restoreValues<Out, keys>(key) | (left value << right value bits) | right value
  const uint32_t key_shift = c.d.d.bits_after(c.d.dom - 1);
  const uint32_t lval_shift = rv ? c.d.e.total() : 0;
  uint64_t ibytes;
  RFX_TRY(read_u64(c.S(5), &ibytes));
  uint64_t* cells = nullptr;
  RFX_TRY(ctx_cells(c.ctx, &cells));
  const uint64_t total = ibytes / 4;
  RFX_TRY(write_u64(cells, total));
  void* lval = lv ? c.P(6) : nullptr;
  void* rval = rv ? c.P(lv ? 7 : 6) : nullptr;
  return rfx_join_gather(c.ctx, c.P(0), ob, c.P(1), kb, lval, lvb, rval, rvb,
                         static_cast<int32_t*>(c.P(2)), static_cast<int32_t*>(c.P(4)), cells, total,
                         key_shift, lval_shift);
}
// union_get_result_size(size, left, right) / difference_get_result_size(size, left)
int ptx_union_size(const Call& c) {
  RFX_TRY(c.need(3, 3));
  k_size_sum<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), c.S(2));
  return launched();
}
int ptx_difference_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  k_size_sum<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(1), nullptr);
  return launched();
}
// project_get_result_size(size): d result, a source;
// size = size * sizeof(result) / sizeof(source) (Project.h getResultSize).
int ptx_project_size(const Call& c) {
  RFX_TRY(c.need(1, 1));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a, "d and a"));
  int rb, sb;
  RFX_TRY(bytes_checked(c, c.d.d, "result", &rb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  k_size_scale<<<1, 1, 0, c.ctx->stream>>>(c.S(0), c.S(0), sb, rb, 1);
  return launched();
}
// set_project(result, begin, size): d result, a source, map; result field
// k = source field map[k] (a 64-bit extract; Project.h permuteElement),
// over size / sizeof(result) rows.
int ptx_project(const Call& c) {
  RFX_TRY(c.need(3, 3));
  RFX_TRY(c.want(c.d.has_d && c.d.has_a && !c.d.map.empty(), "d, a and map"));
  int rb, sb;
  RFX_TRY(bytes_checked(c, c.d.d, "result", &rb));
  RFX_TRY(bytes_checked(c, c.d.a, "source", &sb));
  uint32_t in_shift[16], bits[16], out_shift[16];
  int nf = 0;
  for (int k = 0; k < c.d.d.n && k < static_cast<int>(c.d.map.size()); ++k) {
    const uint32_t src = c.d.map[k];
    if (src >= static_cast<uint32_t>(c.d.a.n)) continue;  // an unused slot extracts 0
    uint32_t w = c.d.d.bits[k] < c.d.a.bits[src] ? c.d.d.bits[k] : c.d.a.bits[src];
    if (w > 64) w = 64;
    in_shift[nf] = c.d.a.bits_after(src);
    bits[nf] = w;
    out_shift[nf] = c.d.d.bits_after(k);
    ++nf;
  }
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(2), &bytes));
  const uint64_t rows = bytes / rb;
  if (nf == 0) return rows ? fill_u32(c.ctx, static_cast<uint32_t*>(c.P(0)), rows * rb / 4, 0) : RFX_OK;
  return rfx_project(c.ctx, c.P(0), rb, c.P(1), sb, rows, nf, in_shift, bits, out_shift);
}
// set_select(output, histogram, input, bytes): d tuple, cmp; the kept
// rows go to the front of `output`, in input order, as one partition.
int ptx_select(const Call& c) {
  RFX_TRY(c.need(4, 4));
  RFX_TRY(c.want(c.d.has_d && c.d.has_cmp && c.d.cmp.field >= 0 && c.d.cmp.field < c.d.d.n,
                 "d and cmp"));
  int tb;
  RFX_TRY(bytes_checked(c, c.d.d, "tuple", &tb));
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(3), &bytes));
  uint64_t* cells = nullptr;
  RFX_TRY(ctx_cells(c.ctx, &cells));
  const int f = c.d.cmp.field;
  RFX_TRY(rfx_select(c.ctx, c.P(0), cells, c.P(2), bytes / tb, tb, c.d.d.bits_after(f),
                     c.d.d.bits[f], c.d.cmp.op, c.d.cmp.constant));
  const uint32_t parts = c.d.parts > 0 && c.d.parts < 1024 ? c.d.parts : 128;
  k_select_hist<<<1, 256, 0, c.ctx->stream>>>(static_cast<uint32_t*>(c.P(1)), cells, parts);
  return launched();
}
// scan(histogram): exclusive scan of threadCount entries in place.
int ptx_scan(const Call& c) {
  RFX_TRY(c.need(1, 1));
  const uint64_t n = c.dims[1];
  if (n == 0 || n > 1024) return fail(RFX_EINVAL, "%s: %llu entries", c.name, (unsigned long long)n);
  k_excl_scan_u32<<<1, 1024, 0, c.ctx->stream>>>(static_cast<uint32_t*>(c.P(0)), static_cast<uint32_t>(n));
  return launched();
}
// select_get_result_size(size, histogram): d tuple, parts;
// size = histogram[parts] * sizeof(tuple).
int ptx_select_size(const Call& c) {
  RFX_TRY(c.need(2, 2));
  RFX_TRY(c.want(c.d.has_d, "d"));
  int tb;
  RFX_TRY(bytes_checked(c, c.d.d, "tuple", &tb));
  uint32_t h = 0;
  const uint32_t parts = c.d.parts > 0 && c.d.parts < 1024 ? c.d.parts : 128;
  RFX_HIP(hipStreamSynchronize(c.ctx->stream));
  RFX_HIP(hipMemcpy(&h, static_cast<uint32_t*>(c.P(1)) + parts, 4, hipMemcpyDeviceToHost));
  return write_u64(c.S(0), static_cast<uint64_t>(h) * tb);
}
// select_gather(dest, temp, temp_size, histogram): d tuple, parts, nt.
int ptx_select_gather(const Call& c) {
  RFX_TRY(c.need(4, 4));
  RFX_TRY(c.want(c.d.has_d, "d"));
  int tb;
  RFX_TRY(bytes_checked(c, c.d.d, "tuple", &tb));
  uint64_t tbytes;
  RFX_TRY(read_u64(c.S(2), &tbytes));
  const uint32_t parts = c.d.parts > 0 && c.d.parts < 1024 ? c.d.parts : 128;
  const uint32_t nv = c.d.nt > 0 ? static_cast<uint32_t>(c.d.nt) : 128;
  k_select_gather<<<parts, 256, 0, c.ctx->stream>>>(
      static_cast<unsigned char*>(c.P(0)), static_cast<const unsigned char*>(c.P(1)), tbytes / tb,
      static_cast<const uint32_t*>(c.P(3)), tb, parts, nv);
  return launched();
}

// ---- binary families with a descriptor -------------------------------------
// Sort of (string address, value) pairs into compare_string2's order,
// stably (sort_string_pair, ModernGPUSort.cu:148-213).
int sort_string_pairs(rfx_ctx* ctx, uint64_t* keys, void* vals, int vb, uint64_t n) {
  if (n < 2) return RFX_OK;
  const uint64_t vbytes = vals ? n * vb : 0;
  Scratch s;
  RFX_TRY(s.alloc(8 * (n + 1) + 4 * n + 8 * n + vbytes + 64));
  char* base = static_cast<char*>(s.p);
  uint64_t* scratch = reinterpret_cast<uint64_t*>(base);
  uint32_t* rows = reinterpret_cast<uint32_t*>(base + 8 * (n + 1));
  unsigned char* kcopy = reinterpret_cast<unsigned char*>(base + align_up(8 * (n + 1) + 4 * n, 16));
  unsigned char* vcopy = kcopy + 8 * n;
  RFX_TRY(rfx_gen_iota_u32(ctx, rows, n));
  RFX_TRY(rfx_sort_strings_rows(ctx, keys, rows, scratch, n));
  k_permute_rows<<<grid_for(n), 256, 0, ctx->stream>>>(kcopy, reinterpret_cast<unsigned char*>(keys), rows, n, 8);
  RFX_TRY(launched());
  RFX_HIP(hipMemcpyAsync(keys, kcopy, 8 * n, hipMemcpyDeviceToDevice, ctx->stream));
  if (vals) {
    k_permute_rows<<<grid_for(n), 256, 0, ctx->stream>>>(vcopy, static_cast<unsigned char*>(vals), rows, n, vb);
    RFX_TRY(launched());
    RFX_HIP(hipMemcpyAsync(vals, vcopy, vbytes, hipMemcpyDeviceToDevice, ctx->stream));
  }
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  return RFX_OK;
}

// Reduce-by-key over runs of equal strings (count_string, total_string_32,
// total_string_double; Reduce.cu:424-469,665-785): the runs are numbered
// (rfx_string_groups), reduced on those numbers, and each run's key is the
// address of its first row.  val_kind as reduce_by_key_sized.
int reduce_strings(rfx_ctx* ctx, int op, int val_kind, void* out_keys, void* out_vals,
                   const uint64_t* keys, uint64_t* key_size, const void* vals, uint64_t* value_size,
                   uint64_t vb) {
  uint64_t bytes;
  RFX_TRY(read_u64(key_size, &bytes));
  const uint64_t n = bytes / 8;
  if (n == 0) {
    RFX_TRY(write_u64(key_size, 0));
    return write_u64(value_size, 0);
  }
  Scratch s;
  RFX_TRY(s.alloc(4 * n + 4 * n + 4 * n + 64));
  uint32_t* group = static_cast<uint32_t*>(s.p);
  int32_t* scr = reinterpret_cast<int32_t*>(group + n);
  uint32_t* gkeys = reinterpret_cast<uint32_t*>(scr + n);
  uint64_t* cells = nullptr;
  RFX_TRY(ctx_cells(ctx, &cells));
  RFX_TRY(rfx_string_groups(ctx, keys, n, group, scr, cells + 2));
  RFX_TRY(reduce_by_key_sized(ctx, op, val_kind, 4, gkeys, out_vals, group, vals, key_size, 8, n,
                              key_size, 8, value_size, vb, nullptr, ~0ull));
  k_first_of_run<<<grid_for(n), 256, 0, ctx->stream>>>(static_cast<uint64_t*>(out_keys), keys,
                                                       group, n);
  RFX_TRY(launched());
  RFX_HIP(hipStreamSynchronize(ctx->stream));
  return RFX_OK;
}

// set_reduce / set_count(result_key, result_value, key, key_size, value,
// value_size): a key tuple, b value tuple, agg (set_reduce: total|min|max).
// The key width follows the reference's choice among total_8/16/32/64/128
// (RelationalAlgebraKernel.cpp:2800-2908): the packed byte count rounded
// up to 1, 2, 4, 8 or 16.
int bin_reduce(const Call& c, bool count) {
  RFX_TRY(c.need(6, 6));
  RFX_TRY(c.want(c.d.has_a, "a"));
  int op = RFX_AGG_SUM;
  if (!count) {
    if (c.d.agg == "min") op = RFX_AGG_MIN;
    else if (c.d.agg == "max") op = RFX_AGG_MAX;
    else if (c.d.agg != "total" && !c.d.agg.empty())
      return fail(RFX_EINVAL, "%s: aggregate '%s'", c.name, c.d.agg.c_str());
  }
  const bool f64 = !count && c.d.has_b && c.d.b.is_f64();
  const int val_kind = f64 ? 0 : 1;
  const uint64_t vb = f64 ? 8 : 4;
  if (!count && !f64 && op != RFX_AGG_SUM)
    return fail(RFX_ENOTSUP, "%s: %s of 32-bit values", c.name, c.d.agg.c_str());
  if (c.d.a.is_string())
    return reduce_strings(c.ctx, op, val_kind, c.P(0), c.P(1), static_cast<const uint64_t*>(c.P(2)),
                          c.S(3), c.P(4), c.S(5), vb);
  int kb = bytes_of_bits(8ull * c.d.a.packed_bytes());
  if (!kb) return fail(RFX_ENOTSUP, "%s: %d-byte keys", c.name, c.d.a.packed_bytes());
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(3), &bytes));
  RFX_TRY(reduce_by_key_sized(c.ctx, op, val_kind, kb, c.P(0), c.P(1), c.P(2), c.P(4), c.S(3), kb,
                              bytes / kb, c.S(3), kb, c.S(5), vb, nullptr, ~0ull));
  RFX_HIP(hipStreamSynchronize(c.ctx->stream));
  return RFX_OK;
}

// mgpu_join_find_bounds(lower_bound, left_count, size, left, left_bytes,
// right, right_bytes): a left key tuple (a string address selects
// find_bounds_string).
int bin_find_bounds(const Call& c) {
  RFX_TRY(c.need(7, 7));
  RFX_TRY(c.want(c.d.has_a, "a"));
  if (c.d.a.is_string())
    return rfx_join_find_bounds_string(c.ctx, static_cast<int32_t*>(c.P(0)),
                                       static_cast<int32_t*>(c.P(1)), c.S(2),
                                       static_cast<const uint64_t*>(c.P(3)), c.a[4] / 8,
                                       static_cast<const uint64_t*>(c.P(5)), c.a[6] / 8);
  int kb;
  RFX_TRY(bytes_checked(c, c.d.a, "key", &kb));
  return rfx_join_find_bounds(c.ctx, static_cast<int32_t*>(c.P(0)), static_cast<int32_t*>(c.P(1)),
                              c.S(2), c.P(3), c.a[4] / kb, c.P(5), c.a[6] / kb, kb);
}

// set_mgpusortpair(keys, values, key_bytes, key_type, value_type) and
// set_mgpusortkey(keys, bytes, type): d key tuple.  The type operands come
// from the tuples' bit counts (getTupleDataType,
// RelationalAlgebraCompiler.cpp:494-512), so a string key arrives as I64
// and only the key tuple's Pointer field says to sort strings
// (RelationalAlgebraKernel.cpp:2670,2715).
int bin_mgpu_sort(const Call& c, bool pairs) {
  RFX_TRY(c.need(pairs ? 5 : 3, pairs ? 5 : 3));
  const uint64_t kt = c.a[pairs ? 3 : 2];
  if (c.have_desc && c.d.has_d && c.d.d.is_string()) {
    const int vb = pairs ? bytes_of_type(c.a[4]) : 0;
    if (pairs && !vb) return fail(RFX_EINVAL, "%s: value type %llu", c.name, (unsigned long long)c.a[4]);
    return sort_string_pairs(c.ctx, static_cast<uint64_t*>(c.P(0)), pairs ? c.P(1) : nullptr, vb,
                             c.a[pairs ? 2 : 1] / 8);
  }
  const int kb = bytes_of_type(kt), vb = pairs ? bytes_of_type(c.a[4]) : 0;
  if (!kb || (pairs && !vb)) return fail(RFX_EINVAL, "%s: type codes", c.name);
  return rfx_sort_pairs(c.ctx, c.P(0), pairs ? c.P(1) : nullptr, c.a[pairs ? 2 : 1] / kb, kb, 8 * kb, vb);
}

// set_single_reduce(result, begin, bytes): agg total|max|count.
int bin_single_reduce(const Call& c) {
  RFX_TRY(c.need(3, 3));
  const int kind = c.d.agg == "count" ? 1 : c.d.agg == "max" ? 2 : 0;
  if (!c.d.agg.empty() && c.d.agg != "total" && c.d.agg != "count" && c.d.agg != "max")
    return fail(RFX_EINVAL, "%s: aggregate '%s'", c.name, c.d.agg.c_str());
  const uint64_t eb = kind == 1 ? 4 : 8;
  return single_reduce(c.ctx, kind, c.P(0), c.P(1), c.a[2] / eb);
}

int bin_set_family(const Call& c, int op) {
  RFX_TRY(c.need(7, 7));
  const int eb = bytes_of_type(c.a[6]);
  if (!eb || c.a[6] > RFX_I128) return fail(RFX_EINVAL, "%s: type %llu", c.name, (unsigned long long)c.a[6]);
  return set_op_sized(c.ctx, op, c.P(0), c.S(1), eb, c.P(2), c.a[3] / eb, c.P(4), c.a[5] / eb, eb);
}

int bin_unique(const Call& c) {
  RFX_TRY(c.need(3, 3));
  const int eb = bytes_of_type(c.a[2]);
  if (!eb || c.a[2] > RFX_I128) return fail(RFX_EINVAL, "%s: type %llu", c.name, (unsigned long long)c.a[2]);
  uint64_t bytes;
  RFX_TRY(read_u64(c.S(1), &bytes));
  return unique_sized(c.ctx, c.P(0), c.S(1), eb, c.P(0), bytes / eb, eb);
}

int bin_b40c(const Call& c, bool pairs) {
  RFX_TRY(c.need(pairs ? 5 : 3, pairs ? 5 : 3));
  const uint64_t bits = c.a[pairs ? 3 : 2];
  const int kb = bytes_of_bits(bits), vb = pairs ? bytes_of_type(c.a[4]) : 0;
  if (!kb || bits == 0 || (pairs && !vb)) return fail(RFX_EINVAL, "%s: key bits / value type", c.name);
  return rfx_sort_pairs(c.ctx, c.P(0), pairs ? c.P(1) : nullptr, c.a[pairs ? 2 : 1] / kb, kb,
                        static_cast<int>(bits), vb);
}

int bin_join_main(const Call& c) {
  RFX_TRY(c.need(6, 7));
  uint64_t* cells = nullptr;
  RFX_TRY(ctx_cells(c.ctx, &cells));
  const uint64_t total = c.a[2] / 4;
  RFX_TRY(write_u64(cells + 1, total));
  return rfx_join_main(c.ctx, static_cast<int32_t*>(c.P(0)), static_cast<int32_t*>(c.P(1)), cells + 1,
                       total, static_cast<const int32_t*>(c.P(3)), static_cast<const int32_t*>(c.P(4)),
                       c.a[5] / 4);
}

// The family of a kernel name: the name without its trailing instance id.
std::string family_of_name(const char* name) {
  size_t n = std::strlen(name);
  while (n && name[n - 1] >= '0' && name[n - 1] <= '9') --n;
  return std::string(name, n);
}

using Handler = int (*)(const Call&);
int bin_b40c_pair(const Call& c) { return bin_b40c(c, true); }
int bin_b40c_key(const Call& c) { return bin_b40c(c, false); }
int bin_mgpu_pair(const Call& c) { return bin_mgpu_sort(c, true); }
int bin_mgpu_key(const Call& c) { return bin_mgpu_sort(c, false); }
int bin_reduce_total(const Call& c) { return bin_reduce(c, false); }
int bin_count(const Call& c) { return bin_reduce(c, true); }
int bin_union(const Call& c) { return bin_set_family(c, RFX_SET_UNION); }
int bin_difference(const Call& c) { return bin_set_family(c, RFX_SET_DIFFERENCE); }
int bin_intersection(const Call& c) { return bin_set_family(c, RFX_SET_INTERSECTION); }

struct FamilyEntry {
  const char* name;
  int type;        // 1 compute (PTX), 2 binary
  bool needs_desc; // the instance's types are not in the operands
  Handler run;
};
const FamilyEntry kLaunchFamilies[] = {
    {"split_get_result_size", 1, true, ptx_split_size},
    {"split_key_get_result_size", 1, true, ptx_split_key_size},
    {"set_split", 1, true, ptx_split},
    {"set_splitkey", 1, true, ptx_split_key},
    {"merge_get_result_size", 1, true, ptx_merge_size},
    {"set_merge", 1, true, ptx_merge},
    {"generate_get_result_size", 1, true, ptx_generate_size},
    {"set_generate", 1, false, ptx_generate},
    {"join_get_temp_size", 1, true, ptx_join_temp_size},
    {"join_get_result_size", 1, true, ptx_join_result_size},
    {"mgpu_join_gather", 1, true, ptx_join_gather},
    {"union_get_result_size", 1, false, ptx_union_size},
    {"difference_get_result_size", 1, false, ptx_difference_size},
    {"project_get_result_size", 1, true, ptx_project_size},
    {"set_project", 1, true, ptx_project},
    {"set_select", 1, true, ptx_select},
    {"scan", 1, false, ptx_scan},
    {"select_get_result_size", 1, true, ptx_select_size},
    {"select_gather", 1, true, ptx_select_gather},
    {"set_b40csortpair", 2, false, bin_b40c_pair},
    {"set_b40csortkey", 2, false, bin_b40c_key},
    {"set_mgpusortpair", 2, false, bin_mgpu_pair},
    {"set_mgpusortkey", 2, false, bin_mgpu_key},
    {"mgpu_join_find_bounds", 2, true, bin_find_bounds},
    {"mgpu_join_main", 2, false, bin_join_main},
    {"set_reduce", 2, true, bin_reduce_total},
    {"set_count", 2, true, bin_count},
    {"set_single_reduce", 2, true, bin_single_reduce},
    {"set_union", 2, false, bin_union},
    {"set_difference", 2, false, bin_difference},
    {"set_intersection", 2, false, bin_intersection},
    {"set_unique", 2, false, bin_unique},
};

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_harmony_launch(const char* kernel_name, int kernel_type, const void* code,
                                  uint64_t code_bytes, const uint64_t* args, uint32_t nargs) {
  if (!kernel_name || (nargs && !args) || (code_bytes && !code))
    return fail(RFX_EINVAL, "rfx_harmony_launch: null argument");
  if (kernel_type != 1 && kernel_type != 2)
    return fail(RFX_EINVAL, "rfx_harmony_launch: kernel type %d (1 compute, 2 binary)", kernel_type);
  const std::string fam = family_of_name(kernel_name);
  const FamilyEntry* e = nullptr;
  for (const FamilyEntry& f : kLaunchFamilies)
    if (fam == f.name && kernel_type == f.type) e = &f;
  if (!e)
    return fail(RFX_ENOTSUP, "%s: %s kernel family '%s' is not provided", kernel_name,
                kernel_type == 1 ? "compute" : "binary", fam.c_str());
  KernelDesc d;
  const char* text = static_cast<const char*>(code);
  const bool have = is_desc(text, code_bytes);
  if (have) {
    std::string err;
    if (!parse_desc(text, code_bytes, &d, &err))
      return fail(RFX_EINVAL, "%s: descriptor: bad '%s'", kernel_name, err.c_str());
    if (d.family != fam)
      return fail(RFX_EINVAL, "%s: descriptor is for '%s'", kernel_name, d.family.c_str());
  } else if (e->needs_desc) {
    return fail(RFX_EINVAL, "%s: no rfx1 descriptor in the kernel's code", kernel_name);
  }
  Call c{nullptr, kernel_name, d, have, args, nargs, {0, 0, 0}};
  if (kernel_type == 1) {
    if (nargs < 3)
      return fail(RFX_EINVAL, "%s: %u operands, a compute kernel has 3 or more", kernel_name, nargs);
    c.dims[0] = args[0];
    c.dims[1] = args[1];
    c.dims[2] = args[2];
    c.a = args + 3;
    c.n = nargs - 3;
  }
  c.ctx = rfx_default_ctx();
  if (!c.ctx) return fail(RFX_EHIP, "rfx_harmony_launch: no default context");
  RFX_TRY(e->run(c));
  return rfx_ctx_synchronize(c.ctx);
}
