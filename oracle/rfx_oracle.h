/*
 * rfx_oracle.h - CPU restatement of the Red Fox relational-algebra hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library's results are compared against (tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py).  Nothing in the product links it.
 *
 * Every function states the reference code it restates (paths relative to
 * the reference checkout).  Element widths are byte counts 1, 2, 4, 8 or 16;
 * 16-byte elements are two little-endian u64 words with word 1 the most
 * significant (ra::tuple::PackedNBytes<2>, redfox/redfox/ra/interface/Tuple.h:232-625),
 * compared high word first (ModernGPUJoin.cu:27-39, Union.cu:22-32).
 * All comparisons are unsigned.
 */
#ifndef RFX_ORACLE_H
#define RFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tuple codec (Tuple.h:776-959) ---------------------------------- */
/* A tuple word of up to 128 bits: w[0] low, w[1] high. */
uint64_t rfxo_tuple_extract(const uint64_t w[2], const uint32_t* bits,
                            int nfields, int index);
void rfxo_tuple_insert(uint64_t w[2], const uint32_t* bits, int nfields,
                       int index, uint64_t value);
/* BitsAfter<index, Tuple> (Tuple.h:718-725) */
uint32_t rfxo_bits_after(const uint32_t* bits, int nfields, int index);
/* Clamp<bits>/8 (Tuple.h:30-37): bytes of the word holding a tuple */
uint32_t rfxo_tuple_bytes(const uint32_t* bits, int nfields);

/* ---- sort (b40cSort.cu:58-1021, ModernGPUSort.cu:238-574) ----------- */
/* Stable sort of n (key, value) pairs by the low key_bits bits of the key.
 * val_bytes may be 0 (keys only).  Returns 0, or -1 on bad widths. */
int rfxo_sort_pairs(void* keys, void* vals, uint64_t n, int key_bytes,
                    int key_bits, int val_bytes);

/* ---- join (ModernGPUJoin.cu:504-590,665-733; join.cuh:45-99) -------- */
/* lower_bound[i] = #{r < l[i]}; left_count = exclusive scan of
 * #{r == l[i]}.  Returns the total number of matches. */
uint64_t rfxo_join_find_bounds(int32_t* lower_bound, int32_t* left_count,
                               const void* left, uint64_t nl,
                               const void* right, uint64_t nr, int key_bytes);
/* Left-major, right-minor expansion of the scanned counts. */
void rfxo_join_main(int32_t* left_idx, int32_t* right_idx, uint64_t total,
                    const int32_t* lower_bound, const int32_t* left_count,
                    uint64_t nl);
/* Gather (ModernGPUJoin.h:43-200):
 *   out[g] = (lkey[li] << key_shift) | (lval[li] << lval_shift) | rval[ri]
 * lval/rval may be NULL (width 0).  All shifts are in the out word. */
void rfxo_join_gather(void* out, int out_bytes, const void* lkey, int lkey_bytes,
                      const void* lval, int lval_bytes, const void* rval,
                      int rval_bytes, const int32_t* left_idx,
                      const int32_t* right_idx, uint64_t total,
                      uint32_t key_shift, uint32_t lval_shift);
/* Serial CPUJoin restated from moderngpu tests/benchmarkjoin.cu:40-140
 * (inner join).  Independent of the three functions above; used to pin
 * them.  Returns total. */
uint64_t rfxo_cpujoin(int32_t* a_idx, int32_t* b_idx, uint64_t cap,
                      const void* a, uint64_t na, const void* b, uint64_t nb,
                      int key_bytes);

/* ---- reduce-by-key (Reduce.cu:175-1532) ----------------------------- */
enum { RFXO_SUM = 0, RFXO_MIN = 1, RFXO_MAX = 2 };
/* FP64 reduce over runs of equal adjacent keys.  Sums use Neumaier
 * compensation (compensated=1) or a plain serial loop (0). */
uint64_t rfxo_reduce_by_key_f64(void* out_keys, double* out_vals,
                                const void* keys, const double* vals,
                                uint64_t n, int key_bytes, int op,
                                int compensated);
/* u32 sum over runs (thrust::plus<unsigned>, wraps mod 2^32). */
uint64_t rfxo_reduce_by_key_u32(void* out_keys, uint32_t* out_vals,
                                const void* keys, const uint32_t* vals,
                                uint64_t n, int key_bytes);
/* Single reductions (Reduce.cu:976-1136). */
double rfxo_total_double(const double* v, uint64_t n, int compensated);
uint32_t rfxo_count_u32(const uint32_t* v, uint64_t n);
uint64_t rfxo_max_double_bits(const uint64_t* v, uint64_t n);

/* ---- set operations (Union.cu, Difference.cu, Unique.cu) ------------ */
/* Multiset semantics on sorted inputs: union max(m,n), difference
 * max(m-n,0), intersection min(m,n).  Return the output element count. */
uint64_t rfxo_set_union(void* out, const void* a, uint64_t na, const void* b,
                        uint64_t nb, int elem_bytes);
uint64_t rfxo_set_difference(void* out, const void* a, uint64_t na,
                             const void* b, uint64_t nb, int elem_bytes);
uint64_t rfxo_set_intersection(void* out, const void* a, uint64_t na,
                               const void* b, uint64_t nb, int elem_bytes);
uint64_t rfxo_unique(void* data, uint64_t n, int elem_bytes);

/* ---- select / project / split / merge (Select.h, Project.h,
 *      Split_Merge.h) ------------------------------------------------ */
enum {
  RFXO_EQ = 0, RFXO_NE, RFXO_LT, RFXO_LE, RFXO_GT, RFXO_GE, /* unsigned */
  RFXO_SLT, RFXO_SLE, RFXO_SGT, RFXO_SGE,                   /* signed  */
  RFXO_DLT, RFXO_DLE, RFXO_DGT, RFXO_DGE                    /* double  */
};
/* Stable filter: keep rows whose field (shift, bits) compares true
 * against the constant.  Returns kept row count. */
uint64_t rfxo_select(void* out, const void* in, uint64_t n, int tuple_bytes,
                     uint32_t field_shift, uint32_t field_bits, int op,
                     uint64_t constant);
/* Field permutation: out field f = in bits [in_shift[f], +bits[f])
 * placed at out_shift[f]. */
void rfxo_project(void* out, int out_bytes, const void* in, int in_bytes,
                  uint64_t n, int nfields, const uint32_t* in_shift,
                  const uint32_t* bits, const uint32_t* out_shift);
/* key = in >> key_shift; value = in & ((1 << key_shift) - 1). */
void rfxo_split(void* key, int key_bytes, void* val, int val_bytes,
                const void* in, int in_bytes, uint64_t n, uint32_t key_shift);
/* out = (key << key_shift) | value. */
void rfxo_merge(void* out, int out_bytes, const void* key, int key_bytes,
                const void* val, int val_bytes, uint64_t n, uint32_t key_shift);

/* ---- TPC-H Q1 (q1.man.cpp:959-1815) --------------------------------- */
/* Per group g = returnflag*2 + linestatus (6 groups):
 *   out[g*8 + 0] sum_qty, 1 sum_base_price, 2 sum_disc_price,
 *   3 sum_charge, 4 sum_disc, 5 count (as double), 6,7 unused.
 * Rows with shipdate > cutoff are dropped. */
void rfxo_q1(double* out, const int32_t* shipdate, const uint8_t* flag,
             const uint8_t* status, const int32_t* qty, const double* price,
             const double* disc, const double* tax, uint64_t n,
             int32_t cutoff, int compensated);

/* ---- dates (redfox/redfox/ra/interface/Date.h:17-118) --------------- */
/* Seconds since 1970 <-> (year, month, day) as Date.h computes them for the
 * TPC-H years; the three date operations of set_arith
 * (RelationalAlgebraKernel.cpp:2053-2123) over n values:
 * op 0 AddMonth by k, 1 AddYear by k, 2 PartYear. */
uint64_t rfxo_date2int(int year, int month, int day);
void rfxo_int2date(uint64_t seconds, int* year, int* month, int* day);
void rfxo_date_op(int op, const uint64_t* in, uint64_t* out, uint64_t n, int64_t k);


#ifdef __cplusplus
}
#endif
#endif
