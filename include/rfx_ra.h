/*
 * rfx_ra.h - C-ABI of librfx_hip.so, the MI355X (gfx950) relational-algebra
 * operator library that replaces Red Fox's precompiled binary kernels.
 *
 * Two layers are exported:
 *
 *  1. Native entry points rfx_*(): status-returning (0 = ok, negative =
 *     RFX_E*), explicit context (HIP stream + cached workspace), element
 *     counts instead of byte sizes, output sizes written to device memory
 *     without a host round trip.  They are safe to issue back to back on
 *     one stream and to capture into a HIP graph once the context's
 *     workspace has been reserved.
 *
 *  2. Harmony-convention entry points named after the reference's binary
 *     kernel families (set_b40csortpair, mgpu_join_find_bounds, ...).  They
 *     take the same operands in the same order as the extern "C" wrappers
 *     that RelationalAlgebraKernel.cpp generates (byte sizes, u64 type
 *     codes, u64* size variables in device memory), run on the default
 *     context, and are synchronous.  Where the reference baked the tuple
 *     types into each generated instance, one trailing operand carries
 *     them.  On failure they print the error and call abort(), as the
 *     reference's check() does (ModernGPUJoin.cu:99-107).
 *
 * rfx_harmony_call() is the packed-u64 dispatcher that stands in for
 * Harmony's JIT'd _Z_thunk_<name> wrappers (Runtime.cpp:1408-1817).
 *
 * Element widths are 1, 2, 4, 8 or 16 bytes.  16-byte elements are two
 * little-endian u64 words, word 1 most significant; all key comparisons
 * are unsigned, high word first.
 */
#ifndef RFX_RA_H
#define RFX_RA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes                                                           */
/* ---------------------------------------------------------------------- */
#define RFX_OK 0
#define RFX_EINVAL -1   /* bad argument (width, null pointer, op code)   */
#define RFX_ENOMEM -2   /* device allocation failed                      */
#define RFX_EHIP -3     /* a HIP runtime call failed                     */
#define RFX_ERANGE -4   /* size beyond the 32-bit index range            */
#define RFX_ENOTSUP -5  /* combination not supported                     */
#define RFX_EBUDGET -6  /* admission refused: over the HBM budget        */

/* Type codes of nvcc::RelationalAlgebraKernel::DataType
 * (redfox/redfox/nvcc/interface/RelationalAlgebraKernel.h:81-95). */
#define RFX_I8 1
#define RFX_I16 2
#define RFX_I32 3
#define RFX_I64 4
#define RFX_I128 5
#define RFX_F32 9
#define RFX_F64 10
#define RFX_POINTER 11  /* a string: the address of NUL-terminated bytes */

typedef struct rfx_ctx rfx_ctx;

/* ---------------------------------------------------------------------- */
/* Context, memory, timing (plumbing)                                     */
/* ---------------------------------------------------------------------- */
const char* rfx_strerror(int status);
/* Text of the last error on the calling thread ("" if none). */
const char* rfx_last_error(void);
/* Library version string. */
const char* rfx_version(void);

/* Create a context on `device` issuing work to `stream` (a hipStream_t;
 * NULL = the null stream). */
int rfx_ctx_create(rfx_ctx** ctx, int device, void* stream);
int rfx_ctx_destroy(rfx_ctx* ctx);
int rfx_ctx_set_stream(rfx_ctx* ctx, void* stream);
void* rfx_ctx_stream(rfx_ctx* ctx);
/* Make sure the cached workspace holds at least `bytes`. */
int rfx_ctx_reserve(rfx_ctx* ctx, uint64_t bytes);
uint64_t rfx_ctx_workspace_bytes(rfx_ctx* ctx);
/* Tile shapes of the context's operators, for an autotuner: "sort.ipt"
 * (keys per thread of the onesweep tile: 8, 12 or 16, for 4/8-byte keys
 * with 0/4/8-byte values), "sort.digit" (radix digit bits, 8, 9 or 10, for
 * the same shapes; 0 chooses from key_bits: 9 for 17-18 bits, 10 for
 * 19-20, else 8), "join.fb_tile" and "join.ex_tile" (2048 or 4096 merged
 * items per find_bounds / expansion tile), "select.vt" (rows per thread of
 * the select pass: 8, 16, 24 or 32), "mapfilter.mode" (1: run programs
 * compiled for their types only, 2: interpret them only; the default 0
 * compiles and interprets where compiling is not possible).  0 restores
 * the default.  Results do not depend on the choice.  rfx_sort_workspace
 * covers every sort.ipt value at the default digit width; a context tuned
 * to another width grows its own workspace on first use. */
int rfx_ctx_tune(rfx_ctx* ctx, const char* knob, int value);
int rfx_ctx_tuning(rfx_ctx* ctx, const char* knob, int* value);
int rfx_ctx_synchronize(rfx_ctx* ctx);
/* The process-wide context used by the Harmony-convention entry points. */
rfx_ctx* rfx_default_ctx(void);

int rfx_device_count(int* count);
int rfx_set_device(int device);
int rfx_malloc(void** ptr, uint64_t bytes);
int rfx_free(void* ptr);
int rfx_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int rfx_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
int rfx_memcpy_dtod(void* dst, const void* src, uint64_t bytes);
int rfx_memset(void* dst, int value, uint64_t bytes);
int rfx_device_synchronize(void);
int rfx_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* Per-launch timing: while enabled, every kernel launch of the operators
 * is bracketed by HIP events on the context's stream, grouped by tag
 * ("sort.onesweep", "sort.hist", "join.find_bounds", "join.expand",
 * "rbk.up", "rbk.down", "setop", "select", "q1.partial", ...).  Enabling
 * (or disabling) resets the totals.  Reading waits for pending launches. */
int rfx_ctx_profile(rfx_ctx* ctx, int enable);
int rfx_ctx_profile_read(rfx_ctx* ctx, const char* tag, uint64_t* launches,
                         double* total_ms);

/* Events on the context's stream; elapsed time in milliseconds. */
int rfx_event_create(void** ev);
int rfx_event_destroy(void* ev);
int rfx_event_record(void* ev, rfx_ctx* ctx);
int rfx_event_elapsed(float* ms, void* start, void* stop);

/* Device-side synthetic data: splitmix64 stream value i is
 * mix(seed + (i + 1) * 0x9E3779B97F4A7C15).  out[i] = (u32)(v >> 32). */
int rfx_gen_u32(rfx_ctx* ctx, uint32_t* out, uint64_t n, uint64_t seed);
/* out[i] = full 64-bit splitmix64 value. */
int rfx_gen_u64(rfx_ctx* ctx, uint64_t* out, uint64_t n, uint64_t seed);
/* out[i] = i (u32 row index). */
int rfx_gen_iota_u32(rfx_ctx* ctx, uint32_t* out, uint64_t n);
/* out[i] = mix64(rank), rank ~ Zipf(s) over [1, n_max] by rejection-
 * inversion (Hormann & Derflinger 1996), element i drawing from its own
 * splitmix64 sub-stream; mix64 is the splitmix64 finaliser (a bijection),
 * so all 64 bits vary and equal ranks give equal keys. */
int rfx_gen_zipf_u64(rfx_ctx* ctx, uint64_t* out, uint64_t n, uint64_t seed,
                     double s, uint64_t n_max);
/* TPC-H lineitem columns for Q1 from seed (the same values as the host
 * generator in redfox_amd/datagen.py): shipdate U[0,2526), returnflag
 * U{0,1,2}, linestatus U{0,1}, qty U[1,50], price = qty * U{90000..200000}
 * / 100, disc U{0..10} / 100, tax U{0..8} / 100. */
int rfx_gen_lineitem(rfx_ctx* ctx, int32_t* shipdate, uint8_t* returnflag,
                     uint8_t* linestatus, int32_t* qty, double* price,
                     double* disc, double* tax, uint64_t n, uint64_t seed);

/* Streams and graphs: a context bound to its own stream can record a whole
 * plan (any sequence of native operators) into a HIP graph and replay it
 * with one launch.  Reserve the workspace before capturing: growing it
 * during capture fails with RFX_EINVAL.  A graph keeps the workspace's
 * addresses, so while any graph from rfx_graph_end is alive (until
 * rfx_graph_destroy) an operator, tuning change or rfx_ctx_reserve that
 * would grow the workspace fails with RFX_EINVAL instead of freeing it.
 * Destroy a context's graphs before the context.  Profiling must be off. */
int rfx_stream_create(void** stream);
int rfx_stream_destroy(void* stream);
int rfx_stream_synchronize(void* stream);
int rfx_graph_begin(rfx_ctx* ctx);
int rfx_graph_end(rfx_ctx* ctx, void** graph_exec);
int rfx_graph_launch(rfx_ctx* ctx, void* graph_exec);
int rfx_graph_destroy(void* graph_exec);

/* Device-to-device copy ordered on the context's stream (capturable). */
int rfx_copy(rfx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Instance-prefixed keys for running many query instances as one:
 * out[i] = (s << key_bits) | key[i] where s is the segment holding row i
 * (seg_offsets: nseg + 1 ascending device u64, seg_offsets[0] = 0). */
int rfx_segment_keys(rfx_ctx* ctx, uint32_t* out, const uint32_t* key,
                     const uint64_t* seg_offsets, uint32_t nseg, uint64_t n,
                     uint32_t key_bits);
/* Fact rows of the batched join-aggregate workload: row i of instance b
 * (row_off) gets fk = perm[dim_off[b] + rank - 1], rank ~ Zipf(s) over
 * [1, d_b], d_b = dim_off[b+1] - dim_off[b], and measure =
 * U{0..100000} / 100, all from splitmix64 stream `seed`. */
int rfx_gen_fk_batch(rfx_ctx* ctx, uint32_t* fk, double* measure,
                     const uint64_t* row_off, const uint64_t* dim_off,
                     const uint32_t* perm, uint32_t nseg, uint64_t n,
                     uint64_t seed, double s);

/* ---------------------------------------------------------------------- */
/* Native operators                                                       */
/* ---------------------------------------------------------------------- */

/* Stable LSD radix sort (8-bit digits, onesweep with decoupled look-back)
 * of n keys and, if val_bytes > 0, values, ordered by the low key_bits
 * bits of each key.  The result lands in keys/vals.  Replaces
 * redfox::sort_pair/sort_key (b40cSort.cu:58-1021) and the mergesort
 * variants (ModernGPUSort.cu:238-574, with key_bits = 8*key_bytes). */
int rfx_sort_pairs(rfx_ctx* ctx, void* keys, void* vals, uint64_t n,
                   int key_bytes, int key_bits, int val_bytes);
/* Workspace bytes rfx_sort_pairs needs for this shape. */
uint64_t rfx_sort_workspace(uint64_t n, int key_bytes, int key_bits,
                            int val_bytes);

/* Sorted search of a sorted left key column in a sorted right key column
 * (find_bounds_{8,16,32,64,128}, ModernGPUJoin.cu:109-663):
 *   lower_bound[i] = #{r < left[i]},
 *   left_count[i]  = exclusive prefix sum of #{r == left[i]},
 *   *total (device u64) = number of matching pairs. */
int rfx_join_find_bounds(rfx_ctx* ctx, int32_t* lower_bound,
                         int32_t* left_count, uint64_t* total,
                         const void* left, uint64_t nl, const void* right,
                         uint64_t nr, int key_bytes);
/* rfx_join_find_bounds without allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as rfx_join_find_bounds; the launch's scratch
 * (merge-path splits, look-back status words, tile ticket) lives in a
 * static device array of about 1 MiB per GPU, held until the process
 * ends, instead of the context's workspace.  At most 2^28 merged keys
 * (nl + nr) of up to 8 bytes, 2^27 of 16 bytes; more gives RFX_ERANGE
 * before anything runs.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (concurrent streams or threads) race. */
int rfx_join_find_bounds_noalloc(rfx_ctx* ctx, int32_t* lower_bound,
                                 int32_t* left_count, uint64_t* total,
                                 const void* left, uint64_t nl,
                                 const void* right, uint64_t nr, int key_bytes);
/* find_bounds over string keys (find_bounds_string, ModernGPUJoin.cu:
 * 213-312): left/right hold device addresses of NUL-terminated strings,
 * each side sorted in compare_string2's order (:72-97: bytes compared as
 * signed char, a proper prefix first; a null address reads as the empty
 * string).  Outputs as rfx_join_find_bounds; rfx_join_main expands them. */
int rfx_join_find_bounds_string(rfx_ctx* ctx, int32_t* lower_bound, int32_t* left_count,
                                uint64_t* total, const uint64_t* left, uint64_t nl,
                                const uint64_t* right, uint64_t nr);
/* Sort n string addresses in place into compare_string2's order, stably
 * (equal strings keep their order); scratch holds n + 1 u64.  No reference
 * counterpart takes unsorted string columns: the reference's string keys
 * arrive sorted.  Synchronises the stream once (the longest string sets
 * the number of passes). */
int rfx_sort_strings(rfx_ctx* ctx, uint64_t* addresses, uint64_t* scratch, uint64_t n);
/* The same order as a permutation: rows (n u32 row numbers below n, e.g.
 * 0..n-1) is sorted stably by the string of each row, addresses[row]
 * (left unchanged); scratch holds n + 1 u64.  For string-keyed joins of
 * relations whose other columns follow the key. */
int rfx_sort_strings_rows(rfx_ctx* ctx, const uint64_t* addresses, uint32_t* rows,
                          uint64_t* scratch, uint64_t n);
/* Number the runs of equal strings in a sorted array of string addresses:
 * group[i] = number of distinct strings before row i's string (0, 0, 1,
 * ...), *n_groups (device u64) the count; scratch holds n int32.  No
 * reference counterpart: it turns strings into sorted integer group keys
 * for GROUP BY on a string column. */
int rfx_string_groups(rfx_ctx* ctx, const uint64_t* sorted, uint64_t n, uint32_t* group,
                      int32_t* scratch, uint64_t* n_groups);
/* Load-balanced expansion (redfox::join, ModernGPUJoin.cu:665-733):
 * for g < total, left_idx[g] = i with left_count[i] <= g <
 * left_count[i] + count[i], right_idx[g] = lower_bound[i] + g -
 * left_count[i].  `total` is a device pointer (read on the device). */
int rfx_join_main(rfx_ctx* ctx, int32_t* left_idx, int32_t* right_idx,
                  const uint64_t* total, uint64_t total_cap,
                  const int32_t* lower_bound, const int32_t* left_count,
                  uint64_t nl);
/* Gather of joined rows (gather_*, ModernGPUJoin.h:43-200):
 *   out[g] = (lkey[li] << key_shift) | (lval[li] << lval_shift) | rval[ri]
 * lval / rval may be NULL.  `total` is a device pointer. */
int rfx_join_gather(rfx_ctx* ctx, void* out, int out_bytes, const void* lkey,
                    int lkey_bytes, const void* lval, int lval_bytes,
                    const void* rval, int rval_bytes, const int32_t* left_idx,
                    const int32_t* right_idx, const uint64_t* total,
                    uint64_t total_cap, uint32_t key_shift,
                    uint32_t lval_shift);
/* find_bounds + expansion + gather in one pass over the output, without
 * materialising the index columns.  *total must already be known (from
 * rfx_join_find_bounds). */
int rfx_join_expand_gather(rfx_ctx* ctx, void* out, int out_bytes,
                           const void* lkey, int lkey_bytes, const void* lval,
                           int lval_bytes, const void* rval, int rval_bytes,
                           const int32_t* lower_bound,
                           const int32_t* left_count, uint64_t nl,
                           const uint64_t* total, uint64_t total_cap,
                           uint32_t key_shift, uint32_t lval_shift);

/* Expansion fused with a projection of the joined row into one or two
 * columns (the gather + project + split steps of a join feeding an
 * aggregate, in one pass).  fields holds nfields (<= 8) records of five
 * u32: output column (0 or 1), source (0 left key, 1 left value, 2 right
 * value), in_shift, bits, out_shift; out[g] |= ((src >> in_shift) & mask)
 * << out_shift.  out1 may be NULL. */
int rfx_join_expand_project(rfx_ctx* ctx, const int32_t* lower_bound,
                            const int32_t* left_count, uint64_t nl,
                            const uint64_t* total, uint64_t total_cap,
                            const void* lkey, int lkey_bytes, const void* lval,
                            int lval_bytes, const void* rval, int rval_bytes,
                            void* out0, int out0_bytes, void* out1,
                            int out1_bytes, int nfields, const uint32_t* fields);

/* Reduce-by-key over runs of equal adjacent keys (reduce_by_key in
 * Reduce.cu:175-1532).  keys/vals hold *n rows (device u64 *n);
 * out_keys/out_vals receive one row per run and *n_out (device u64) the
 * run count.  op: RFX_AGG_SUM, RFX_AGG_MIN, RFX_AGG_MAX.  Outputs must
 * not overlap keys or vals (as for thrust::reduce_by_key).  The inputs are
 * taken to hold n_cap rows; an output that overlaps them gives RFX_EINVAL,
 * where an output spans min(n_cap, out_cap) rows when out_cap is given
 * and its first row otherwise.  Any other overlap is undefined. */
#define RFX_AGG_SUM 0
#define RFX_AGG_MIN 1
#define RFX_AGG_MAX 2
int rfx_reduce_by_key_f64(rfx_ctx* ctx, void* out_keys, double* out_vals,
                          uint64_t* n_out, const void* keys,
                          const double* vals, const uint64_t* n,
                          uint64_t n_cap, int key_bytes, int op);
int rfx_reduce_by_key_u32(rfx_ctx* ctx, void* out_keys, uint32_t* out_vals,
                          uint64_t* n_out, const void* keys,
                          const uint32_t* vals, const uint64_t* n,
                          uint64_t n_cap, int key_bytes);
/* reduce_by_key_f64 and count_by_key over the same keys at once: run
 * r gets out_keys[r], out_vals[r] and its row count out_counts[r] (the
 * aggregate and COUNT(*) of one GROUP BY, as a Red Fox plan computes them
 * with set_reduce then set_count).  The outputs hold out_cap rows: *n_out
 * counts every run, runs past out_cap are not written. */
int rfx_reduce_count_by_key_f64(rfx_ctx* ctx, void* out_keys, double* out_vals,
                                uint32_t* out_counts, uint64_t out_cap,
                                uint64_t* n_out, const void* keys,
                                const double* vals, const uint64_t* n,
                                uint64_t n_cap, int key_bytes, int op);
/* Count of each run (reduce_by_key of the generate() ones). */
int rfx_count_by_key(rfx_ctx* ctx, void* out_keys, uint32_t* out_counts,
                     uint64_t* n_out, const void* keys, const uint64_t* n,
                     uint64_t n_cap, int key_bytes);
/* Single reductions (Reduce.cu:976-1136): sum of f64, sum of u32, and the
 * maximum of u64 bit patterns.  Results go to device memory. */
int rfx_total_f64(rfx_ctx* ctx, double* result, const double* v, uint64_t n);
int rfx_total_u32(rfx_ctx* ctx, uint32_t* result, const uint32_t* v,
                  uint64_t n);
int rfx_max_u64(rfx_ctx* ctx, uint64_t* result, const uint64_t* v,
                uint64_t n);

/* Set operations on sorted inputs with multiset semantics (Union.cu,
 * Difference.cu, Unique.cu): union keeps max(m,n) copies of a value,
 * difference max(m-n,0), intersection min(m,n).  *n_out (device u64)
 * receives the output element count; `out` must hold na+nb (union) or
 * na elements. */
#define RFX_SET_UNION 0
#define RFX_SET_DIFFERENCE 1
#define RFX_SET_INTERSECTION 2
int rfx_set_op(rfx_ctx* ctx, int op, void* out, uint64_t* n_out,
               const void* a, uint64_t na, const void* b, uint64_t nb,
               int elem_bytes);
/* Drop adjacent duplicates (thrust::unique); in and out may alias only if
 * equal. */
int rfx_unique(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in,
               uint64_t n, int elem_bytes);

/* Stable filter of packed tuples (select_field_constant, Select.h:30-114,
 * plus its scan and gather): keep rows whose field at [field_shift,
 * field_shift + field_bits) compares true against `constant`.
 * Ops: RFX_CMP_*.  *n_out (device u64) receives the kept count. */
#define RFX_CMP_EQ 0
#define RFX_CMP_NE 1
#define RFX_CMP_LT 2
#define RFX_CMP_LE 3
#define RFX_CMP_GT 4
#define RFX_CMP_GE 5
#define RFX_CMP_SLT 6
#define RFX_CMP_SLE 7
#define RFX_CMP_SGT 8
#define RFX_CMP_SGE 9
#define RFX_CMP_DLT 10
#define RFX_CMP_DLE 11
#define RFX_CMP_DGT 12
#define RFX_CMP_DGE 13
int rfx_select(rfx_ctx* ctx, void* out, uint64_t* n_out, const void* in,
               uint64_t n, int tuple_bytes, uint32_t field_shift,
               uint32_t field_bits, int op, uint64_t constant);
/* Field permutation (permute, Project.h:58-130). */
int rfx_project(rfx_ctx* ctx, void* out, int out_bytes, const void* in,
                int in_bytes, uint64_t n, int nfields,
                const uint32_t* in_shift, const uint32_t* bits,
                const uint32_t* out_shift);
/* rfx_project over a device row count: min(*n, n_cap) rows (the size
 * variable a preceding select wrote, as Harmony passes it). */
int rfx_project_sized(rfx_ctx* ctx, void* out, int out_bytes, const void* in,
                      int in_bytes, const uint64_t* n, uint64_t n_cap,
                      int nfields, const uint32_t* in_shift,
                      const uint32_t* bits, const uint32_t* out_shift);
/* split (Split_Merge.h:303-330): key = in >> key_shift, value = low
 * key_shift bits.  val may be NULL (split_key). */
int rfx_split(rfx_ctx* ctx, void* key, int key_bytes, void* val,
              int val_bytes, const void* in, int in_bytes, uint64_t n,
              uint32_t key_shift);
/* merge (Split_Merge.h:441-519): out = (key << key_shift) | value. */
int rfx_merge(rfx_ctx* ctx, void* out, int out_bytes, const void* key,
              int key_bytes, const void* val, int val_bytes, uint64_t n,
              uint32_t key_shift);
/* generate (Split_Merge.h:541-555): out[i] = 1. */
int rfx_generate_ones(rfx_ctx* ctx, uint32_t* out, uint64_t n);

/* Fused TPC-H Q1 filter-group-aggregate over columns (the plan of
 * q1.man.cpp:959-2329 in one pass): rows with shipdate <= cutoff, grouped
 * by g = returnflag*2 + linestatus (6 groups).  out (device, 6*8 f64):
 * [g*8+0] sum_qty, +1 sum_base_price, +2 sum_disc_price, +3 sum_charge,
 * +4 sum_disc, +5 count, +6 avg_qty, +7 avg_price.  avg_disc =
 * sum_disc / count.  Summation order is fixed by n alone. */
int rfx_q1(rfx_ctx* ctx, double* out, const int32_t* shipdate,
           const uint8_t* returnflag, const uint8_t* linestatus,
           const int32_t* qty, const double* price, const double* disc,
           const double* tax, uint64_t n, int32_t cutoff);

/* ---------------------------------------------------------------------- */
/* Served form: a request loop over device-resident relations.            */
/* ---------------------------------------------------------------------- */
/* No single counterpart in the reference: it replaces Harmony's loop over
 * one plan's nodes (Runtime.cpp:303-445, 506-646, 1054-1314) with a loop
 * over many plans' requests.  Plain mode runs one request at a time and,
 * like Harmony, synchronises after every operator and reads each size
 * variable back; continuous mode admits requests in arrival order whenever
 * one of `slots` streams and enough of the HBM budget are free, enqueues
 * each request's whole plan on its slot with sizes kept on the device, and
 * polls one event per slot, so newly arrived requests run beside those
 * already on the GPU.  Every request's intermediates, outputs and operator
 * workspace come from one block of the budget, freed when it completes. */
typedef struct rfx_server rfx_server;

/* Plans (the relations each takes, in order):
 * SELECT_PROJECT: T, 16-byte tuples [127:96] id, [95:64] qty, [63:0] val;
 *   SELECT id, val FROM T WHERE qty <param> constant (param = RFX_CMP_*,
 *   0 meaning RFX_CMP_LT); output 8-byte tuples id << 32 | (u32)val.
 * JOIN_AGG: F (fk u32, m f64), D (dkey u32 unique, grp u32 of val_bits
 *   bits); SELECT grp, SUM(m), COUNT(*) FROM F JOIN D ON fk = dkey GROUP BY
 *   grp; outputs keys u32, sums f64, counts u32.
 * JOIN3: F (fk1 u32, m f64), S (skey u32 unique, fk2 u32 of val_bits
 *   bits), T (tkey u32 unique, tval u32); F JOIN S ON fk1 = skey JOIN T ON
 *   fk2 = tkey; outputs (tval u32, m f64) in join order.
 * SET_OP: A, B sorted unique u64 (unsigned order); A <param> B with param
 *   = RFX_SET_*; output u64.
 * The join plans expect every left row to match exactly one right row
 * (foreign keys into a primary key); a request that breaks this ends with
 * status RFX_EINVAL. */
#define RFX_PLAN_SELECT_PROJECT 0
#define RFX_PLAN_JOIN_AGG 1
#define RFX_PLAN_JOIN3 2
#define RFX_PLAN_SET_OP 3

typedef struct rfx_relation {
  const void* col[2];   /* key column, value column (device) */
  uint64_t rows;
  uint32_t key_bits;    /* significant bits of the key column (radix passes) */
  uint32_t val_bits;    /* significant bits of the value column, where a plan
                           sorts or groups by it */
} rfx_relation;

typedef struct rfx_request {
  uint32_t schema;      /* from rfx_server_add_schema */
  uint32_t param;       /* RFX_CMP_* (select) or RFX_SET_* (set op) */
  uint64_t constant;    /* select constant */
  uint64_t arrival_ns;  /* arrival time after the start of the run */
  void* sink;           /* optional device buffer receiving the output
                           columns, each at capacity rows, in order */
  uint64_t sink_bytes;
  /* filled in by rfx_server_run */
  uint64_t admit_ns, done_ns;  /* since the start of the run */
  uint64_t mem_bytes;   /* bytes of the budget the request held */
  uint64_t rows;        /* output rows */
  uint64_t digest[3];   /* per output column: sum_i fmix64(w_i +
                           fmix64(0x9E3779B97F4A7C15 * (c + 1) + i)) */
  int32_t status;       /* RFX_OK, or the request's error */
  uint32_t reserved;
} rfx_request;

typedef struct rfx_serve_stats {
  uint64_t wall_ns;          /* first admission window to last completion */
  uint64_t peak_bytes;       /* most budget bytes held at once */
  uint64_t budget_bytes;
  uint64_t blocked_on_memory;  /* admissions deferred for lack of budget */
  uint32_t peak_inflight;    /* most requests on the GPU at once */
  uint32_t slots;
} rfx_serve_stats;

#define RFX_SERVE_PLAIN 0
#define RFX_SERVE_CONTINUOUS 1

/* A server on `device` with `slots` streams (1..64) and an HBM budget of
 * budget_bytes for requests (allocated once, up front). */
int rfx_server_create(rfx_server** srv, int device, int slots,
                      uint64_t budget_bytes);
int rfx_server_destroy(rfx_server* srv);
/* Register resident relations for one plan; *schema_id names them.  The
 * relations must stay resident and unchanged while the server runs.  For
 * the join plans the joins are run once here: every join may be
 * many-to-many, and each request's buffers are sized from the measured
 * join sizes (a request whose join makes another number of rows ends
 * with RFX_EINVAL). */
int rfx_server_add_schema(rfx_server* srv, int plan, const rfx_relation* rels,
                          int nrels, uint32_t* schema_id);
/* Budget bytes a request needs (its whole plan, workspace included). */
int rfx_server_request_bytes(rfx_server* srv, const rfx_request* req,
                             uint64_t* bytes);
/* Serve n requests (arrival_ns non-decreasing) and return when all are
 * done.  stats may be NULL. */
int rfx_server_run(rfx_server* srv, rfx_request* reqs, uint64_t n, int mode,
                   rfx_serve_stats* stats);

/* ---------------------------------------------------------------------- */
/* Harmony-convention entry points (synchronous, default context).        */
/* Operand order follows RelationalAlgebraKernel.cpp; "P" = device        */
/* pointer slot, "S" = scalar u64 slot in Harmony's packed arguments.     */
/* ---------------------------------------------------------------------- */

/* set_b40csortpair<id>(K* keys P, V* vals P, u64 key_bytes S,
 *   u64 key_bits S, u64 value_type S) - RelationalAlgebraKernel.cpp:2745-2772.
 * Key width = Clamp(key_bits)/8 as in the generated Tuple_key. */
void set_b40csortpair(void* keys, void* vals, uint64_t key_bytes,
                      uint64_t key_bits, uint64_t value_type);
/* set_b40csortkey<id>(K* keys P, u64 bytes S, u64 bits S) - :2774-2798. */
void set_b40csortkey(void* keys, uint64_t bytes, uint64_t key_bits);
/* set_mgpusortpair<id>(K*, V*, u64 key_bytes, u64 key_type,
 *   u64 value_type) - :2658-2712. */
void set_mgpusortpair(void* keys, void* vals, uint64_t key_bytes,
                      uint64_t key_type, uint64_t value_type);
/* set_mgpusortkey<id>(K*, u64 bytes, u64 type) - :2714-2743. */
void set_mgpusortkey(void* keys, uint64_t bytes, uint64_t type);

/* mgpu_join_find_bounds<id>(int* lower_bound P, int* left_count P,
 *   u64* result_size P, K* left P, u64 left_bytes S, K* right P,
 *   u64 right_bytes S) - :1210-1301; trailing key_type (the generated
 *   instance's key tuple width: RFX_I8..RFX_I128).  *result_size
 *   receives the match count, as ModernGPUJoin.cu:566-567 writes it. */
void mgpu_join_find_bounds(int32_t* lower_bound, int32_t* left_count,
                           uint64_t* result_size, void* left_key,
                           uint64_t left_bytes, void* right_key,
                           uint64_t right_bytes, uint64_t key_type);
/* mgpu_join_main<id>(int* left_idx P, int* right_idx P, u64 idx_bytes S,
 *   int* lower_bound P, int* left_count P, u64 lb_bytes S) - :1377-1397. */
void mgpu_join_main(int32_t* left_idx, int32_t* right_idx,
                    uint64_t idx_bytes, int32_t* lower_bound,
                    int32_t* left_count, uint64_t lb_bytes);
/* join_get_result_size<id>(u64* index_size, u64* result_size)
 * - :1188-1208: index_size = total*4, result_size = total*out_bytes.
 * Trailing out_bytes (sizeof the generated output tuple). */
void join_get_result_size(uint64_t* index_size, uint64_t* result_size,
                          uint64_t out_bytes);
/* mgpu_join_gather<id>(Out* result, LK* left_key, int* left_index,
 *   RK* right_key, int* right_index, u64* size, [LV* left_value],
 *   [RV* right_value]) - :1101-1186, all device pointers; *size is the
 * index byte size.  The generated instance's tuple shapes come in
 * `desc`: bits 0-7 out_bytes, 8-15 lkey_bytes, 16-23 lval_bytes (0 =
 * none), 24-31 rval_bytes (0 = none), 32-39 key_shift, 40-47 lval_shift. */
void mgpu_join_gather(void* result, void* left_key, int32_t* left_index,
                      void* right_key, int32_t* right_index, uint64_t* size,
                      void* left_value, void* right_value, uint64_t desc);

/* set_reduce<id> / set_count<id>(RK* result_key P, RV* result_value P,
 *   K* key P, u64* key_size P, V* value P, u64* value_size P) -
 *   :2800-3045.  Sizes are byte counts, in and out.  Trailing desc:
 *   bits 0-7 key_bytes, 8-15 op (RFX_AGG_*), 16-23 value type code
 *   (RFX_F64 or RFX_I32). */
void set_reduce(void* result_key, void* result_value, void* key,
                uint64_t* key_size, void* value, uint64_t* value_size,
                uint64_t desc);
void set_count(void* result_key, void* result_value, void* key,
               uint64_t* key_size, void* value, uint64_t* value_size,
               uint64_t key_bytes);
/* set_single_reduce<id>(R* result P, T* begin P, u64 bytes S) -
 *   :2910-2946.  Trailing kind: 0 total (f64), 1 count (u32 sum),
 *   2 max (u64 bit patterns). */
void set_single_reduce(void* result, void* begin, uint64_t bytes,
                       uint64_t kind);
/* set_union<id> / set_difference<id>(T* out P, u64* out_bytes P, T* a P,
 *   u64 a_bytes S, T* b P, u64 b_bytes S, u64 type S) - :3077-3100,
 *   :983-1007.  set_intersection follows the same convention (declared
 *   but unimplemented in the reference, :919-922). */
void set_union(void* result, uint64_t* size, void* left, uint64_t left_bytes,
               void* right, uint64_t right_bytes, uint64_t type);
void set_difference(void* result, uint64_t* size, void* left,
                    uint64_t left_bytes, void* right, uint64_t right_bytes,
                    uint64_t type);
void set_intersection(void* result, uint64_t* size, void* left,
                      uint64_t left_bytes, void* right, uint64_t right_bytes,
                      uint64_t type);
/* set_unique<id>(T* begin P, u64* bytes P, u64 type S) - :2948-2968. */
void set_unique(void* begin, uint64_t* bytes, uint64_t type);

/* Packed-argument dispatcher: classify `kernel_name` by family prefix the
 * way BinaryKernelLaunch::execute does (Runtime.cpp:672-1038) and call the
 * entry point with args[i] read as a pointer or a scalar.  Returns
 * RFX_OK, or RFX_EINVAL for an unknown family / wrong argument count. */
int rfx_harmony_call(const char* kernel_name, const uint64_t* args,
                     uint32_t nargs);

/* One Harmony kernel record run on the native operators: what
 * PTXKernelLaunch::execute (Runtime.cpp:506-646, kernel_type 1) and
 * BinaryKernelLaunch::execute (:672-1038, kernel_type 2) do, with the
 * operand slots packed as they pack them: for kernel_type 1, args[0..2]
 * are the values of the ctaCount, threadCount and sharedBytes variables
 * and every later slot a device address; for kernel_type 2, the family's
 * pointer slots are addresses and the rest the first 8 bytes of their
 * variables.  `code` is the HIR kernel's code bytes: an rfx1 descriptor
 * holding what RelationalAlgebraKernel::cudaSourceRepresentation()
 * (RelationalAlgebraKernel.cpp:849-909) bakes into the instance:
 *   rfx1 <family> [d=T] [d1=T] [a=T] [b=T] [c=T] [e=T] [dom=N]
 *        [map=i,j,...] [agg=total|min|max|count] [cmp=FIELD:OP:CONST]
 *        [parts=N] [nt=N]
 * T lists a tuple's field widths in bits, each optionally followed by f
 * (float) or p (a string address), or "-" for none.  Compute families:
 * split_get_result_size, split_key_get_result_size, set_split,
 * set_splitkey, merge_get_result_size, set_merge,
 * generate_get_result_size, set_generate, join_get_temp_size,
 * join_get_result_size, mgpu_join_gather, union_get_result_size,
 * difference_get_result_size, project_get_result_size, set_project,
 * set_select (field-vs-constant), scan, select_get_result_size,
 * select_gather.  Binary families: as rfx_harmony_call but with the
 * reference's operand count, plus string (Pointer) keys in
 * set_mgpusortpair/key, mgpu_join_find_bounds, set_reduce and set_count.
 * Runs on the default context and finishes before returning; failures
 * return RFX_E* (RFX_ENOTSUP for a family that is not provided). */
int rfx_harmony_launch(const char* kernel_name, int kernel_type,
                       const void* code, uint64_t code_bytes,
                       const uint64_t* args, uint32_t nargs);

/* ---------------------------------------------------------------------- */
/* MAPFILTER: fused filter / project / arithmetic over columns            */
/* ---------------------------------------------------------------------- */
/* Replaces the per-instance generated arithmetic and selection kernels
 * (set_arith / set_convert, RelationalAlgebraKernel.cpp:2001-2215;
 * set_select, :2301-2593) with one register program: ops
 * [0, n_filter) run on every row and register keep_reg decides whether it
 * is kept (n_filter == 0 keeps every row); ops [n_filter, n_ops) run on the
 * kept rows, starting from cleared registers, and their OUT ops store to
 * the output columns at the row's rank among kept rows (input order).
 * *n_kept (device) receives the kept count.  Registers hold 64 bits: an
 * int64 (two's complement; C division and remainder, a zero divisor
 * gives 0) or the bits of a double (IEEE, no fused multiply-add).  The
 * date ops follow Date.h (seconds since 1970, TPC-H years 1992-1998).
 * A program runs as a kernel compiled for it and its column and output
 * types (hiprtc, on its first call; kept for the process), or through the
 * interpreter kernel while a graph is being captured and the program is
 * not compiled yet, when compiling fails, or once 1024 programs are
 * compiled (modules stay loaded for the process); "mapfilter.mode" in
 * rfx_ctx_tune forces one.  Both give the same results. */
#define RFX_MF_MAX_COLS 16
#define RFX_MF_MAX_OUTS 8
#define RFX_MF_MAX_OPS 96
#define RFX_MF_REGS 16

/* Column element types; TUPLEn columns are packed tuples of n bytes whose
 * fields LDC extracts. */
#define RFX_MF_U8 1
#define RFX_MF_U16 2
#define RFX_MF_U32 3
#define RFX_MF_U64 4
#define RFX_MF_I8 5
#define RFX_MF_I16 6
#define RFX_MF_I32 7
#define RFX_MF_I64 8
#define RFX_MF_F32 9
#define RFX_MF_F64 10
#define RFX_MF_TUPLE1 11
#define RFX_MF_TUPLE2 12
#define RFX_MF_TUPLE4 13
#define RFX_MF_TUPLE8 14
#define RFX_MF_TUPLE16 15

/* Opcodes: R[d] = f(R[a], R[b], imm). */
#define RFX_MF_NOP 0
#define RFX_MF_LDC 1       /* R[d] = column a at the row; integers widened, floats as
                              double bits; TUPLEn: imm bits 0-7 shift, 8-15 bits,
                              16-17 kind (0 unsigned, 1 signed, 2 double) */
#define RFX_MF_LDI 2       /* R[d] = imm */
#define RFX_MF_MOV 3       /* R[d] = R[a] */
#define RFX_MF_LDX 4       /* R[d] = column a at row R[b] (a gather; imm as for LDC);
                              a row at or past the column's length reads 0 */
#define RFX_MF_IADD 10
#define RFX_MF_ISUB 11
#define RFX_MF_IMUL 12
#define RFX_MF_IDIV 13     /* truncating; R[b] == 0 gives 0 */
#define RFX_MF_IMOD 14     /* sign of the dividend; R[b] == 0 gives 0 */
#define RFX_MF_FADD 20
#define RFX_MF_FSUB 21
#define RFX_MF_FMUL 22
#define RFX_MF_FDIV 23
#define RFX_MF_I2F 30      /* int64 -> double */
#define RFX_MF_U2F 31      /* uint64 -> double */
#define RFX_MF_F2I 32      /* double -> int64, truncating, saturating, NaN -> 0 */
#define RFX_MF_ILT 40      /* comparisons give 1 or 0; I signed, U unsigned, F double */
#define RFX_MF_ILE 41
#define RFX_MF_IEQ 42
#define RFX_MF_INE 43
#define RFX_MF_ULT 44
#define RFX_MF_ULE 45
#define RFX_MF_FLT 46
#define RFX_MF_FLE 47
#define RFX_MF_FEQ 48
#define RFX_MF_FNE 49
#define RFX_MF_AND 50      /* logical */
#define RFX_MF_OR 51
#define RFX_MF_NOT 52      /* R[d] = R[a] == 0 */
#define RFX_MF_ADD_MONTH 60 /* date R[a] plus imm months (AddMonth, :2053-2090) */
#define RFX_MF_ADD_YEAR 61  /* date R[a] plus imm years (AddYear) */
#define RFX_MF_PART_YEAR 62 /* year of date R[a] (PartYear, :2091-2123) */
#define RFX_MF_SCHUNK 63   /* chunk imm of the string at address R[a] (0: the empty
                              string): 8 bytes as big-endian codes, the NUL and
                              what follows it 0, other bytes as signed char
                              -128..127 -> 1..255, so chunks compare (unsigned)
                              in compare_string2's order.  R[a] must be 0 or the
                              address of a NUL-terminated string in device memory */
#define RFX_MF_LIKE 64     /* string at R[a] LIKE pattern at R[b] (addresses as for
                              SCHUNK): '%' matches any run; the pattern's head
                              before the first '%' and tail after the last must
                              match the string's ends, its middle words are found
                              leftmost in turn (Comparisons.h:12-372); 1 or 0 */
#define RFX_MF_OUT 70      /* output column b at the row's rank = R[a], converted to
                              the column type (integers truncated, F32 rounded) */

typedef struct rfx_mf_op {
  uint8_t code, d, a, b;
  uint32_t reserved;
  uint64_t imm;
} rfx_mf_op;

typedef struct rfx_mf_program {
  uint32_t n_filter, n_ops, keep_reg, reserved;
  rfx_mf_op ops[RFX_MF_MAX_OPS];
} rfx_mf_program;

typedef struct rfx_mf_col {
  const void* data;
  int32_t type, reserved;
  uint64_t rows;  /* rows LDX may read (0: the pass's n_rows) */
} rfx_mf_col;

typedef struct rfx_mf_out {
  void* data;
  int32_t type, reserved;
} rfx_mf_out;

int rfx_mapfilter(rfx_ctx* ctx, const rfx_mf_col* cols, int n_cols, uint64_t n_rows,
                  const rfx_mf_program* program, const rfx_mf_out* outs, int n_outs,
                  uint64_t* n_kept);

/* ---------------------------------------------------------------------- */
/* Loading delimited text                                                 */
/* ---------------------------------------------------------------------- */
/* Rows of delimited text (TPC-H .tbl: fields ended by '|'; the reference's
 * TC.tbl: tab-separated) already in device memory, to typed device
 * columns (TC_data_process.cu:20-100: getline, strtok, atoi).  Rows end at
 * '\n' (a '\r' before it is dropped); a final row needs no newline.  A row
 * is split at every delimiter first and each field read from its own
 * bytes, whatever the delimiter (a space or a sign included).  Two
 * delimiters in a row make an empty field (it reads 0), where strtok
 * would skip it; TPC-H and TC.tbl text has no empty fields, so the two
 * agree there.  Field f of every row goes to columns[f] as types[f]:
 *   RFX_TBL_SKIP  not stored            RFX_TBL_I32/I64/U32/U8  atoi-style
 *   RFX_TBL_F64   decimal digits[.digits] (mantissa / 10^k, correctly
 *                 rounded below 2^53)   RFX_TBL_DATE  YYYY-MM-DD as i64
 *                 seconds (Date.h)      RFX_TBL_CHAR  the field's first byte
 *   RFX_TBL_STR   the field's device address (u64), a pointer key for
 *                 rfx_join_find_bounds_string: the byte ending the field
 *                 (delimiter, '\r' or newline) is overwritten with NUL, so
 *                 the text is consumed; a field that runs to the end of the
 *                 text reads as the null pointer (the empty string)
 * *rows (device u64) gets the row count; rows past `capacity` are counted,
 * not stored.  *short_rows (device u64) counts rows with fewer fields than
 * n_fields (their missing fields read 0).  Synchronises the stream. */
#define RFX_TBL_MAX_COLS 32
#define RFX_TBL_SKIP 0
#define RFX_TBL_I32 1
#define RFX_TBL_I64 2
#define RFX_TBL_U32 3
#define RFX_TBL_U8 4
#define RFX_TBL_F64 5
#define RFX_TBL_DATE 6
#define RFX_TBL_CHAR 7
#define RFX_TBL_STR 8
int rfx_tbl_parse(rfx_ctx* ctx, void* text, uint64_t bytes, char delim,
                  const int* types, int n_fields, void* const* columns,
                  uint64_t capacity, uint64_t* rows, uint64_t* short_rows);

/* ---------------------------------------------------------------------- */
/* Harmony IR programs                                                    */
/* ---------------------------------------------------------------------- */
/* A loaded HIR file: a u64 byte count, a KernelControlFlowGraph message and
 * `testCount` (u64 count, Test message) records (HarmonyIR.proto;
 * RelationalAlgebraCompiler.cpp:9502-9520 writes it, Module::read,
 * Module.cpp:193-333, reads it).  Replaces harmony::runtime::Runtime
 * (Runtime.cpp:65-445,1040-1314): blocks run in file order, variables are
 * allocated before their first Out operand and freed after their last In
 * operand, GetSize/Resize/UpdateSize act on the size words, and binary
 * kernels run through rfx_harmony_call with operands packed as device
 * addresses or 8-byte values per family.  Declared sizes are byte counts,
 * as Runtime treats them.  PTX compute kernels return RFX_ENOTSUP. */
typedef struct rfx_hir rfx_hir;

typedef struct rfx_hir_summary {
  char name[128];
  uint32_t blocks, variables, kernels, tests;
  uint32_t entry, exit;
} rfx_hir_summary;

/* Parse a HIR file image (no device work).  RFX_EINVAL on a malformed or
 * inconsistent file. */
int rfx_hir_load(rfx_hir** out, const void* bytes, uint64_t nbytes);
void rfx_hir_destroy(rfx_hir* program);
int rfx_hir_info(const rfx_hir* program, rfx_hir_summary* summary);
/* The operation list of block `block` (file order) as text:
 * "alloc <var>", "run <kernel>", "free <var>" separated by "; ". */
int rfx_hir_schedule(const rfx_hir* program, uint32_t block, char* text, uint64_t cap);
/* Replace an input variable's initial contents (and size) for later runs. */
int rfx_hir_set_input(rfx_hir* program, uint32_t variable, const void* data, uint64_t bytes);
/* Run the program on the default device; variables still allocated at the
 * end (inputs, outputs) stay readable until the next run or destroy. */
int rfx_hir_run(rfx_hir* program);
int rfx_hir_variable(const rfx_hir* program, uint32_t variable, void** device_addr,
                     uint64_t* bytes);
int rfx_hir_read(const rfx_hir* program, uint32_t variable, void* dst, uint64_t cap,
                 uint64_t* bytes);
/* Run every self-test record: set its inputs, run, and compare each listed
 * output byte for byte (size and contents).  The loader of the reference
 * reads these records but never runs them (HirManager.cpp:76-163).
 * `report` receives one line per failure. */
int rfx_hir_run_tests(rfx_hir* program, uint32_t* passed, uint32_t* failed, char* report,
                      uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
