// harmony.hip - the Harmony binary-kernel calling convention over the
// native operators.
//
// Red Fox compiles each RA node into extern "C" wrappers with per-instance
// tuple types (RelationalAlgebraKernel.cpp:849-3100) that Harmony calls
// through JIT'd thunks unpacking an array of 8-byte slots
// (Runtime.cpp:672-1038, 1408-1817).  The functions here take the same
// operands in the same order - byte sizes, u64 type codes, u64* size
// variables in device memory - with one trailing operand where the
// reference baked the tuple types into the generated instance.  They run on
// the default context, finish before returning, and abort on failure as the
// reference's check() does.  rfx_harmony_call() is the packed dispatcher.
#include <cstring>

#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
int set_op_sized(rfx_ctx* ctx, int op, void* out, uint64_t* n_out, uint64_t mul, const void* a,
                 uint64_t na, const void* b, uint64_t nb, int eb);
int unique_sized(rfx_ctx* ctx, void* out, uint64_t* n_out, uint64_t mul, const void* in,
                 uint64_t n, int eb);
namespace {

rfx_ctx* hctx(const char* where) {
  rfx_ctx* c = rfx_default_ctx();
  if (!c) die(where, RFX_EHIP);
  return c;
}

void finish(const char* where, rfx_ctx* c, int status) {
  if (status == RFX_OK) status = rfx_ctx_synchronize(c);
  if (status != RFX_OK) die(where, status);
}

// The default context's device u64 cells, for scalar operands the native
// API reads on the device.
uint64_t* cells(const char* where, rfx_ctx* c) {
  uint64_t* p = nullptr;
  const int st = ctx_cells(c, &p);
  if (st != RFX_OK) die(where, st);
  return p;
}

uint64_t read_cell(const char* where, const uint64_t* dev) {
  uint64_t v = 0;
  if (hipMemcpy(&v, dev, 8, hipMemcpyDeviceToHost) != hipSuccess) die(where, RFX_EHIP);
  return v;
}

void write_cell(const char* where, uint64_t* dev, uint64_t v) {
  if (hipMemcpy(dev, &v, 8, hipMemcpyHostToDevice) != hipSuccess) die(where, RFX_EHIP);
}

__global__ void k_join_result_size(uint64_t* index_size, uint64_t* result_size,
                                   uint64_t out_bytes) {
  const uint64_t total = *result_size;
  *index_size = total * 4;
  *result_size = total * out_bytes;
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" {

void set_b40csortpair(void* keys, void* vals, uint64_t key_bytes, uint64_t key_bits,
                      uint64_t value_type) {
  rfx_ctx* c = hctx("set_b40csortpair");
  const int kb = bytes_of_bits(key_bits), vb = bytes_of_type(value_type);
  if (!kb || !vb) die("set_b40csortpair", fail(RFX_EINVAL, "key_bits %llu / value type %llu",
                                               (unsigned long long)key_bits,
                                               (unsigned long long)value_type));
  finish("set_b40csortpair", c,
         rfx_sort_pairs(c, keys, vals, key_bytes / kb, kb, static_cast<int>(key_bits), vb));
}

void set_b40csortkey(void* keys, uint64_t bytes, uint64_t key_bits) {
  rfx_ctx* c = hctx("set_b40csortkey");
  const int kb = bytes_of_bits(key_bits);
  if (!kb) die("set_b40csortkey", fail(RFX_EINVAL, "key_bits %llu", (unsigned long long)key_bits));
  finish("set_b40csortkey", c,
         rfx_sort_pairs(c, keys, nullptr, bytes / kb, kb, static_cast<int>(key_bits), 0));
}

void set_mgpusortpair(void* keys, void* vals, uint64_t key_bytes, uint64_t key_type,
                      uint64_t value_type) {
  rfx_ctx* c = hctx("set_mgpusortpair");
  const int kb = bytes_of_type(key_type), vb = bytes_of_type(value_type);
  if (!kb || !vb) die("set_mgpusortpair", fail(RFX_EINVAL, "type codes"));
  finish("set_mgpusortpair", c, rfx_sort_pairs(c, keys, vals, key_bytes / kb, kb, 8 * kb, vb));
}

void set_mgpusortkey(void* keys, uint64_t bytes, uint64_t type) {
  rfx_ctx* c = hctx("set_mgpusortkey");
  const int kb = bytes_of_type(type);
  if (!kb) die("set_mgpusortkey", fail(RFX_EINVAL, "type code %llu", (unsigned long long)type));
  finish("set_mgpusortkey", c, rfx_sort_pairs(c, keys, nullptr, bytes / kb, kb, 8 * kb, 0));
}

void mgpu_join_find_bounds(int32_t* lower_bound, int32_t* left_count, uint64_t* result_size,
                           void* left_key, uint64_t left_bytes, void* right_key,
                           uint64_t right_bytes, uint64_t key_type) {
  rfx_ctx* c = hctx("mgpu_join_find_bounds");
  const int kb = bytes_of_type(key_type);
  if (!kb) die("mgpu_join_find_bounds", fail(RFX_EINVAL, "key type %llu", (unsigned long long)key_type));
  finish("mgpu_join_find_bounds", c,
         rfx_join_find_bounds(c, lower_bound, left_count, result_size, left_key, left_bytes / kb,
                              right_key, right_bytes / kb, kb));
}

void mgpu_join_main(int32_t* left_idx, int32_t* right_idx, uint64_t idx_bytes,
                    int32_t* lower_bound, int32_t* left_count, uint64_t lb_bytes) {
  rfx_ctx* c = hctx("mgpu_join_main");
  uint64_t* cell = cells("mgpu_join_main", c);
  const uint64_t total = idx_bytes / 4;
  write_cell("mgpu_join_main", cell, total);
  finish("mgpu_join_main", c,
         rfx_join_main(c, left_idx, right_idx, cell, total, lower_bound, left_count, lb_bytes / 4));
}

void join_get_result_size(uint64_t* index_size, uint64_t* result_size, uint64_t out_bytes) {
  rfx_ctx* c = hctx("join_get_result_size");
  k_join_result_size<<<1, 1, 0, c->stream>>>(index_size, result_size, out_bytes);
  finish("join_get_result_size", c, hipGetLastError() == hipSuccess ? RFX_OK : RFX_EHIP);
}

void mgpu_join_gather(void* result, void* left_key, int32_t* left_index, void* right_key,
                      int32_t* right_index, uint64_t* size, void* left_value, void* right_value,
                      uint64_t desc) {
  (void)right_key;  // equal to the left key on every output row
  rfx_ctx* c = hctx("mgpu_join_gather");
  const int ob = desc & 0xFF, kb = (desc >> 8) & 0xFF, lvb = (desc >> 16) & 0xFF,
            rvb = (desc >> 24) & 0xFF;
  const uint32_t kshift = (desc >> 32) & 0xFF, lvshift = (desc >> 40) & 0xFF;
  uint64_t* cell = cells("mgpu_join_gather", c);
  const uint64_t total = read_cell("mgpu_join_gather", size) / 4;
  write_cell("mgpu_join_gather", cell + 1, total);
  finish("mgpu_join_gather", c,
         rfx_join_gather(c, result, ob, left_key, kb, lvb ? left_value : nullptr, lvb,
                         rvb ? right_value : nullptr, rvb, left_index, right_index, cell + 1, total,
                         kshift, lvshift));
}

void set_reduce(void* result_key, void* result_value, void* key, uint64_t* key_size, void* value,
                uint64_t* value_size, uint64_t desc) {
  rfx_ctx* c = hctx("set_reduce");
  const int kb = desc & 0xFF, op = (desc >> 8) & 0xFF, vtype = (desc >> 16) & 0xFF;
  const int vkind = vtype == RFX_F64 ? 0 : 1;
  const uint64_t vb = vtype == RFX_F64 ? 8 : 4;
  if (!valid_width(kb) || (vtype != RFX_F64 && vtype != RFX_I32))
    die("set_reduce", fail(RFX_EINVAL, "descriptor %llx", (unsigned long long)desc));
  const uint64_t n = read_cell("set_reduce", key_size) / kb;
  finish("set_reduce", c,
         reduce_by_key_sized(c, op, vkind, kb, result_key, result_value, key, value, key_size, kb,
                             n, key_size, kb, value_size, vb, nullptr, ~0ull));
}

void set_count(void* result_key, void* result_value, void* key, uint64_t* key_size, void* value,
               uint64_t* value_size, uint64_t key_bytes) {
  rfx_ctx* c = hctx("set_count");
  const int kb = static_cast<int>(key_bytes);
  if (!valid_width(kb)) die("set_count", fail(RFX_EINVAL, "key width %d", kb));
  const uint64_t n = read_cell("set_count", key_size) / kb;
  finish("set_count", c,
         reduce_by_key_sized(c, RFX_AGG_SUM, 1, kb, result_key, result_value, key, value, key_size,
                             kb, n, key_size, kb, value_size, 4, nullptr, ~0ull));
}

void set_single_reduce(void* result, void* begin, uint64_t bytes, uint64_t kind) {
  rfx_ctx* c = hctx("set_single_reduce");
  const uint64_t eb = kind == 1 ? 4 : 8;
  finish("set_single_reduce", c, single_reduce(c, static_cast<int>(kind), result, begin, bytes / eb));
}

static void set_family(const char* where, int op, void* result, uint64_t* size, void* left,
                       uint64_t left_bytes, void* right, uint64_t right_bytes, uint64_t type) {
  rfx_ctx* c = hctx(where);
  const int eb = bytes_of_type(type);
  if (!eb || type > RFX_I128) die(where, fail(RFX_EINVAL, "type %llu", (unsigned long long)type));
  finish(where, c,
         set_op_sized(c, op, result, size, eb, left, left_bytes / eb, right, right_bytes / eb, eb));
}

void set_union(void* result, uint64_t* size, void* left, uint64_t left_bytes, void* right,
               uint64_t right_bytes, uint64_t type) {
  set_family("set_union", RFX_SET_UNION, result, size, left, left_bytes, right, right_bytes, type);
}
void set_difference(void* result, uint64_t* size, void* left, uint64_t left_bytes, void* right,
                    uint64_t right_bytes, uint64_t type) {
  set_family("set_difference", RFX_SET_DIFFERENCE, result, size, left, left_bytes, right,
             right_bytes, type);
}
void set_intersection(void* result, uint64_t* size, void* left, uint64_t left_bytes, void* right,
                      uint64_t right_bytes, uint64_t type) {
  set_family("set_intersection", RFX_SET_INTERSECTION, result, size, left, left_bytes, right,
             right_bytes, type);
}

void set_unique(void* begin, uint64_t* bytes, uint64_t type) {
  rfx_ctx* c = hctx("set_unique");
  const int eb = bytes_of_type(type);
  if (!eb || type > RFX_I128) die("set_unique", fail(RFX_EINVAL, "type %llu", (unsigned long long)type));
  const uint64_t n = read_cell("set_unique", bytes) / eb;
  finish("set_unique", c, unique_sized(c, begin, bytes, eb, begin, n, eb));
}

int rfx_harmony_call(const char* name, const uint64_t* a, uint32_t nargs) {
  if (!name || (nargs && !a)) return fail(RFX_EINVAL, "rfx_harmony_call: null argument");
  auto is = [&](const char* prefix) { return std::strncmp(name, prefix, std::strlen(prefix)) == 0; };
  auto P = [&](uint32_t i) { return reinterpret_cast<void*>(a[i]); };
  auto need = [&](uint32_t lo, uint32_t hi) {
    return nargs >= lo && nargs <= hi
               ? RFX_OK
               : fail(RFX_EINVAL, "%s: %u operands, expected %u..%u", name, nargs, lo, hi);
  };
  if (is("set_b40csortpair")) {
    RFX_TRY(need(5, 5));
    set_b40csortpair(P(0), P(1), a[2], a[3], a[4]);
  } else if (is("set_b40csortkey")) {
    RFX_TRY(need(3, 3));
    set_b40csortkey(P(0), a[1], a[2]);
  } else if (is("set_mgpusortpair")) {
    RFX_TRY(need(5, 5));
    set_mgpusortpair(P(0), P(1), a[2], a[3], a[4]);
  } else if (is("set_mgpusortkey")) {
    RFX_TRY(need(3, 3));
    set_mgpusortkey(P(0), a[1], a[2]);
  } else if (is("mgpu_join_find_bounds")) {
    RFX_TRY(need(8, 8));
    mgpu_join_find_bounds(static_cast<int32_t*>(P(0)), static_cast<int32_t*>(P(1)),
                          static_cast<uint64_t*>(P(2)), P(3), a[4], P(5), a[6], a[7]);
  } else if (is("mgpu_join_main")) {
    // A 7th operand (the right size) is passed and ignored
    // (RelationalAlgebraCompiler.cpp:5089-5091).
    RFX_TRY(need(6, 7));
    mgpu_join_main(static_cast<int32_t*>(P(0)), static_cast<int32_t*>(P(1)), a[2],
                   static_cast<int32_t*>(P(3)), static_cast<int32_t*>(P(4)), a[5]);
  } else if (is("mgpu_join_gather")) {
    RFX_TRY(need(9, 9));
    mgpu_join_gather(P(0), P(1), static_cast<int32_t*>(P(2)), P(3), static_cast<int32_t*>(P(4)),
                     static_cast<uint64_t*>(P(5)), P(6), P(7), a[8]);
  } else if (is("join_get_result_size")) {
    RFX_TRY(need(3, 3));
    join_get_result_size(static_cast<uint64_t*>(P(0)), static_cast<uint64_t*>(P(1)), a[2]);
  } else if (is("set_union") || is("set_difference") || is("set_intersection")) {
    RFX_TRY(need(7, 7));
    const int op = is("set_union") ? RFX_SET_UNION
                   : is("set_difference") ? RFX_SET_DIFFERENCE : RFX_SET_INTERSECTION;
    set_family(name, op, P(0), static_cast<uint64_t*>(P(1)), P(2), a[3], P(4), a[5], a[6]);
  } else if (is("set_single_reduce")) {
    RFX_TRY(need(4, 4));
    set_single_reduce(P(0), P(1), a[2], a[3]);
  } else if (is("set_unique")) {
    RFX_TRY(need(3, 3));
    set_unique(P(0), static_cast<uint64_t*>(P(1)), a[2]);
  } else if (is("set_reduce")) {
    RFX_TRY(need(7, 7));
    set_reduce(P(0), P(1), P(2), static_cast<uint64_t*>(P(3)), P(4), static_cast<uint64_t*>(P(5)),
               a[6]);
  } else if (is("set_count")) {
    RFX_TRY(need(7, 7));
    set_count(P(0), P(1), P(2), static_cast<uint64_t*>(P(3)), P(4), static_cast<uint64_t*>(P(5)),
              a[6]);
  } else {
    return fail(RFX_EINVAL, "rfx_harmony_call: unknown kernel family '%s'", name);
  }
  return RFX_OK;
}

}  // extern "C"
