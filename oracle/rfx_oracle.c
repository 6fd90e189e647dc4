/*
 * rfx_oracle.c - CPU restatement of the Red Fox RA hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  It is
 * deliberately serial and simple.  Citations are paths in the reference
 * checkout.
 */
#include "rfx_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- element helpers ----------------------------------------------- */

typedef struct { uint64_t lo, hi; } u128;

static u128 ld(const void* base, uint64_t i, int eb) {
  const uint8_t* p = (const uint8_t*)base + i * (uint64_t)eb;
  u128 r = {0, 0};
  switch (eb) {
    case 1: r.lo = p[0]; break;
    case 2: { uint16_t v; memcpy(&v, p, 2); r.lo = v; } break;
    case 4: { uint32_t v; memcpy(&v, p, 4); r.lo = v; } break;
    case 8: memcpy(&r.lo, p, 8); break;
    case 16: memcpy(&r.lo, p, 8); memcpy(&r.hi, p + 8, 8); break;
    default: break;
  }
  return r;
}

static void st(void* base, uint64_t i, int eb, u128 v) {
  uint8_t* p = (uint8_t*)base + i * (uint64_t)eb;
  switch (eb) {
    case 1: p[0] = (uint8_t)v.lo; break;
    case 2: { uint16_t w = (uint16_t)v.lo; memcpy(p, &w, 2); } break;
    case 4: { uint32_t w = (uint32_t)v.lo; memcpy(p, &w, 4); } break;
    case 8: memcpy(p, &v.lo, 8); break;
    case 16: memcpy(p, &v.lo, 8); memcpy(p + 8, &v.hi, 8); break;
    default: break;
  }
}

/* Unsigned compare, high word first (compare_sort_gpu128,
 * ModernGPUJoin.cu:27-39). */
static int cmp(u128 a, u128 b) {
  if (a.hi != b.hi) return a.hi < b.hi ? -1 : 1;
  if (a.lo != b.lo) return a.lo < b.lo ? -1 : 1;
  return 0;
}

static u128 shl(u128 v, uint32_t s) {
  u128 r = {0, 0};
  if (s == 0) return v;
  if (s >= 128) return r;
  if (s >= 64) { r.hi = v.lo << (s - 64); return r; }
  r.hi = (v.hi << s) | (v.lo >> (64 - s));
  r.lo = v.lo << s;
  return r;
}

static u128 shr(u128 v, uint32_t s) {
  u128 r = {0, 0};
  if (s == 0) return v;
  if (s >= 128) return r;
  if (s >= 64) { r.lo = v.hi >> (s - 64); return r; }
  r.lo = (v.lo >> s) | (v.hi << (64 - s));
  r.hi = v.hi >> s;
  return r;
}

static u128 mask_bits(u128 v, uint32_t bits) {
  if (bits >= 128) return v;
  if (bits >= 64) {
    if (bits > 64) v.hi &= (~0ull) >> (128 - bits); else v.hi = 0;
    return v;
  }
  v.hi = 0;
  v.lo &= bits ? ((~0ull) >> (64 - bits)) : 0;
  return v;
}

static u128 or128(u128 a, u128 b) { u128 r = {a.lo | b.lo, a.hi | b.hi}; return r; }

static int valid_width(int eb) {
  return eb == 1 || eb == 2 || eb == 4 || eb == 8 || eb == 16;
}

/* ---- tuple codec (Tuple.h:776-959) --------------------------------- */

/* BitsAfter<index> = total bits - bits up to index - bits at index
 * (Tuple.h:718-725): field 0 sits in the most significant bits. */
uint32_t rfxo_bits_after(const uint32_t* bits, int nfields, int index) {
  uint32_t total = 0, upto = 0;
  for (int f = 0; f < nfields; ++f) total += bits[f];
  for (int f = 0; f < index; ++f) upto += bits[f];
  return total - upto - bits[index];
}

/* Clamp<bits> (Tuple.h:30-37) in bytes. */
uint32_t rfxo_tuple_bytes(const uint32_t* bits, int nfields) {
  uint32_t total = 0;
  for (int f = 0; f < nfields; ++f) total += bits[f];
  if (total <= 8) return 1;
  if (total <= 16) return 2;
  if (total <= 32) return 4;
  if (total <= 64) return 8;
  if (total <= 128) return 16;
  return 0;
}

/* extract<T, Index, Tuple> (Tuple.h:776-790): (tuple >> shift) & mask. */
uint64_t rfxo_tuple_extract(const uint64_t w[2], const uint32_t* bits,
                            int nfields, int index) {
  u128 t = {w[0], w[1]};
  uint32_t s = rfxo_bits_after(bits, nfields, index);
  return mask_bits(shr(t, s), bits[index]).lo;
}

/* insert<T, Index, Tuple> (Tuple.h:846-857): clear the field, or in the
 * shifted value masked to the field. */
void rfxo_tuple_insert(uint64_t w[2], const uint32_t* bits, int nfields,
                       int index, uint64_t value) {
  uint32_t s = rfxo_bits_after(bits, nfields, index);
  u128 ones = {~0ull, ~0ull};
  u128 m = shl(mask_bits(ones, bits[index]), s);
  u128 v = {value, 0};
  v = shl(v, s);
  w[0] = (w[0] & ~m.lo) | (v.lo & m.lo);
  w[1] = (w[1] & ~m.hi) | (v.hi & m.hi);
}

/* ---- sort ---------------------------------------------------------- */

/* Stable LSD radix sort on 8-bit digits of the low key_bits bits.  This is
 * the semantics of Enactor::Sort<LARGE_PROBLEM, key_bits, 0>
 * (b40cSort.cu:58-826, b40c radix_sort/enactor.cuh:272-306): a stable sort
 * by the key's low key_bits bits, values following their keys.  With
 * key_bits = 8*key_bytes it is also the stable MergesortPairs/Keys order
 * (ModernGPUSort.cu:238-574). */
int rfxo_sort_pairs(void* keys, void* vals, uint64_t n, int key_bytes,
                    int key_bits, int val_bytes) {
  if (!valid_width(key_bytes)) return -1;
  if (val_bytes != 0 && !valid_width(val_bytes)) return -1;
  if (key_bits <= 0 || key_bits > 8 * key_bytes) return -1;
  if (n < 2) return 0;
  uint8_t* kt = (uint8_t*)malloc(n * key_bytes);
  uint8_t* vt = val_bytes ? (uint8_t*)malloc(n * val_bytes) : NULL;
  uint64_t* cnt = (uint64_t*)malloc(256 * sizeof(uint64_t));
  uint8_t *ksrc = (uint8_t*)keys, *kdst = kt;
  uint8_t *vsrc = (uint8_t*)vals, *vdst = vt;
  int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    uint32_t shift = 8u * (uint32_t)p;
    uint32_t width = (uint32_t)key_bits - shift;
    if (width > 8) width = 8;
    uint32_t dmask = (1u << width) - 1u;
    memset(cnt, 0, 256 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i)
      cnt[shr(ld(ksrc, i, key_bytes), shift).lo & dmask]++;
    uint64_t run = 0;
    for (int d = 0; d < 256; ++d) { uint64_t c = cnt[d]; cnt[d] = run; run += c; }
    for (uint64_t i = 0; i < n; ++i) {
      u128 k = ld(ksrc, i, key_bytes);
      uint64_t dst = cnt[shr(k, shift).lo & dmask]++;
      memcpy(kdst + dst * key_bytes, ksrc + i * key_bytes, key_bytes);
      if (val_bytes)
        memcpy(vdst + dst * val_bytes, vsrc + i * val_bytes, val_bytes);
    }
    uint8_t* t = ksrc; ksrc = kdst; kdst = t;
    t = vsrc; vsrc = vdst; vdst = t;
  }
  if (ksrc != (uint8_t*)keys) {
    memcpy(keys, ksrc, n * key_bytes);
    if (val_bytes) memcpy(vals, vsrc, n * val_bytes);
  }
  free(kt);
  free(vt);
  free(cnt);
  return 0;
}

/* ---- join ---------------------------------------------------------- */

/* find_bounds_{8..128} (ModernGPUJoin.cu:504-590): SortedSearch<Lower>
 * gives lower_bound[i]; SortedEqualityCount gives upper - lower; Scan
 * turns the counts into an exclusive prefix sum (int) and returns the
 * total, which is written to *result_size. */
uint64_t rfxo_join_find_bounds(int32_t* lower_bound, int32_t* left_count,
                               const void* left, uint64_t nl,
                               const void* right, uint64_t nr, int key_bytes) {
  uint64_t lo = 0, hi = 0, run = 0;
  for (uint64_t i = 0; i < nl; ++i) {
    u128 k = ld(left, i, key_bytes);
    while (lo < nr && cmp(ld(right, lo, key_bytes), k) < 0) ++lo;
    if (hi < lo) hi = lo;
    while (hi < nr && cmp(ld(right, hi, key_bytes), k) <= 0) ++hi;
    lower_bound[i] = (int32_t)lo;
    uint64_t c = hi - lo;
    left_count[i] = (int32_t)run;
    run += c;
  }
  return run;
}

/* redfox::join (ModernGPUJoin.cu:665-733) -> KernelLeftJoin<.., false>
 * (join.cuh:45-99): output g belongs to the last left row i whose scanned
 * count is <= g; its right index is lower_bound[i] + (g - scan[i]). */
void rfxo_join_main(int32_t* left_idx, int32_t* right_idx, uint64_t total,
                    const int32_t* lower_bound, const int32_t* left_count,
                    uint64_t nl) {
  uint64_t i = 0;
  for (uint64_t g = 0; g < total; ++g) {
    while (i + 1 < nl && (uint64_t)(uint32_t)left_count[i + 1] <= g) ++i;
    left_idx[g] = (int32_t)i;
    right_idx[g] = lower_bound[i] + (int32_t)(g - (uint32_t)left_count[i]);
  }
}

/* gather_{key_key,key_value,value_key,value_value} (ModernGPUJoin.h:43-200):
 * restoreValues<Out, keys>(lkey) = lkey << BitsAfter<keys-1, Out>, or'ed
 * with the left value shifted by RightValue::bits and the right value. */
void rfxo_join_gather(void* out, int out_bytes, const void* lkey, int lkey_bytes,
                      const void* lval, int lval_bytes, const void* rval,
                      int rval_bytes, const int32_t* left_idx,
                      const int32_t* right_idx, uint64_t total,
                      uint32_t key_shift, uint32_t lval_shift) {
  for (uint64_t g = 0; g < total; ++g) {
    uint64_t li = (uint32_t)left_idx[g], ri = (uint32_t)right_idx[g];
    u128 o = shl(ld(lkey, li, lkey_bytes), key_shift);
    if (lval) o = or128(o, shl(ld(lval, li, lval_bytes), lval_shift));
    if (rval) o = or128(o, ld(rval, ri, rval_bytes));
    o = mask_bits(o, 8u * (uint32_t)out_bytes);
    st(out, g, out_bytes, o);
  }
}

/* Inner-join CPUJoin from moderngpu tests/benchmarkjoin.cu:40-140: one
 * merge walk records each a's lower bound in b, a second walk its upper
 * bound, then the counts are scanned and expanded.  Written here as a
 * second, independent path to the join indices. */
uint64_t rfxo_cpujoin(int32_t* a_idx, int32_t* b_idx, uint64_t cap,
                      const void* a, uint64_t na, const void* b, uint64_t nb,
                      int key_bytes) {
  uint64_t* lb = (uint64_t*)malloc((na + 1) * sizeof(uint64_t));
  uint64_t* ub = (uint64_t*)malloc((na + 1) * sizeof(uint64_t));
  /* walk 1: advance b while b < a (a wins ties) */
  uint64_t ai = 0, bi = 0;
  while (ai < na) {
    if (bi < nb && cmp(ld(b, bi, key_bytes), ld(a, ai, key_bytes)) < 0) ++bi;
    else lb[ai++] = bi;
  }
  /* walk 2: advance b while !(a < b) (b wins ties) */
  ai = bi = 0;
  while (ai < na) {
    if (bi < nb && cmp(ld(a, ai, key_bytes), ld(b, bi, key_bytes)) >= 0) ++bi;
    else ub[ai++] = bi;
  }
  uint64_t out = 0;
  for (ai = 0; ai < na; ++ai)
    for (uint64_t j = lb[ai]; j < ub[ai]; ++j) {
      if (out < cap) { a_idx[out] = (int32_t)ai; b_idx[out] = (int32_t)j; }
      ++out;
    }
  free(lb);
  free(ub);
  return out;
}

/* ---- reduce-by-key ------------------------------------------------- */

typedef struct { double s, c; } nsum;

/* Neumaier's improvement of Kahan summation (Neumaier 1974). */
static void nsum_add(nsum* a, double x) {
  double t = a->s + x;
  if ((a->s >= 0 ? a->s : -a->s) >= (x >= 0 ? x : -x)) a->c += (a->s - t) + x;
  else a->c += (x - t) + a->s;
  a->s = t;
}

/* total_*_double / min_32_double (Reduce.cu:665-975,1415-1532):
 * thrust::reduce_by_key over runs of equal adjacent keys with plus<double>
 * or minimum<double>.  Returns the number of runs. */
uint64_t rfxo_reduce_by_key_f64(void* out_keys, double* out_vals,
                                const void* keys, const double* vals,
                                uint64_t n, int key_bytes, int op,
                                int compensated) {
  uint64_t groups = 0, i = 0;
  while (i < n) {
    u128 k = ld(keys, i, key_bytes);
    nsum acc = {0.0, 0.0};
    double plain = 0.0, ext = vals[i];
    uint64_t j = i;
    for (; j < n && cmp(ld(keys, j, key_bytes), k) == 0; ++j) {
      double v = vals[j];
      nsum_add(&acc, v);
      plain += v;
      if (op == RFXO_MIN && v < ext) ext = v;
      if (op == RFXO_MAX && v > ext) ext = v;
    }
    st(out_keys, groups, key_bytes, k);
    if (op == RFXO_SUM) out_vals[groups] = compensated ? acc.s + acc.c : plain;
    else out_vals[groups] = ext;
    ++groups;
    i = j;
  }
  return groups;
}

/* count_* and total_64_32 (Reduce.cu:175-664,665-818): reduce_by_key
 * with plus<unsigned int>; count is this over the generate() ones
 * (Split_Merge.h:541-555). */
uint64_t rfxo_reduce_by_key_u32(void* out_keys, uint32_t* out_vals,
                                const void* keys, const uint32_t* vals,
                                uint64_t n, int key_bytes) {
  uint64_t groups = 0, i = 0;
  while (i < n) {
    u128 k = ld(keys, i, key_bytes);
    uint32_t acc = 0;
    uint64_t j = i;
    for (; j < n && cmp(ld(keys, j, key_bytes), k) == 0; ++j) acc += vals[j];
    st(out_keys, groups, key_bytes, k);
    out_vals[groups++] = acc;
    i = j;
  }
  return groups;
}

/* total_double (Reduce.cu:976-1033): thrust::reduce, plus<double>. */
double rfxo_total_double(const double* v, uint64_t n, int compensated) {
  nsum acc = {0.0, 0.0};
  double plain = 0.0;
  for (uint64_t i = 0; i < n; ++i) { nsum_add(&acc, v[i]); plain += v[i]; }
  return compensated ? acc.s + acc.c : plain;
}

/* count (Reduce.cu:1034-1090): thrust::reduce of u32 values. */
uint32_t rfxo_count_u32(const uint32_t* v, uint64_t n) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) acc += v[i];
  return acc;
}

/* max_double (Reduce.cu:1091-1136): maximum over the u64 bit patterns. */
uint64_t rfxo_max_double_bits(const uint64_t* v, uint64_t n) {
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; ++i) if (v[i] > m) m = v[i];
  return m;
}

/* ---- set operations ------------------------------------------------ */

/* thrust::set_union (Union.cu:47-148): for each value, max(m, n) copies. */
uint64_t rfxo_set_union(void* out, const void* a, uint64_t na, const void* b,
                        uint64_t nb, int eb) {
  uint64_t i = 0, j = 0, o = 0;
  while (i < na && j < nb) {
    u128 x = ld(a, i, eb), y = ld(b, j, eb);
    int c = cmp(x, y);
    if (c < 0) { st(out, o++, eb, x); ++i; }
    else if (c > 0) { st(out, o++, eb, y); ++j; }
    else { st(out, o++, eb, x); ++i; ++j; }
  }
  for (; i < na; ++i) st(out, o++, eb, ld(a, i, eb));
  for (; j < nb; ++j) st(out, o++, eb, ld(b, j, eb));
  return o;
}

/* thrust::set_difference (Difference.cu:35-143): max(m - n, 0) copies. */
uint64_t rfxo_set_difference(void* out, const void* a, uint64_t na,
                             const void* b, uint64_t nb, int eb) {
  uint64_t i = 0, j = 0, o = 0;
  while (i < na && j < nb) {
    u128 x = ld(a, i, eb), y = ld(b, j, eb);
    int c = cmp(x, y);
    if (c < 0) { st(out, o++, eb, x); ++i; }
    else if (c > 0) ++j;
    else { ++i; ++j; }
  }
  for (; i < na; ++i) st(out, o++, eb, ld(a, i, eb));
  return o;
}

/* Intersection is declared but not implemented in the reference
 * (RelationalAlgebraKernel.cpp:919-922); thrust::set_intersection's
 * min(m, n) multiset semantics are used. */
uint64_t rfxo_set_intersection(void* out, const void* a, uint64_t na,
                               const void* b, uint64_t nb, int eb) {
  uint64_t i = 0, j = 0, o = 0;
  while (i < na && j < nb) {
    u128 x = ld(a, i, eb), y = ld(b, j, eb);
    int c = cmp(x, y);
    if (c < 0) ++i;
    else if (c > 0) ++j;
    else { st(out, o++, eb, x); ++i; ++j; }
  }
  return o;
}

/* thrust::unique (Unique.cu:36-114): drop adjacent duplicates in place. */
uint64_t rfxo_unique(void* data, uint64_t n, int eb) {
  if (n == 0) return 0;
  uint64_t o = 1;
  u128 last = ld(data, 0, eb);
  for (uint64_t i = 1; i < n; ++i) {
    u128 x = ld(data, i, eb);
    if (cmp(x, last) != 0) { st(data, o++, eb, x); last = x; }
  }
  return o;
}

/* ---- select / project / split / merge ------------------------------ */

static int64_t sext(uint64_t v, uint32_t bits) {
  if (bits == 0 || bits >= 64) return (int64_t)v;
  uint64_t m = 1ull << (bits - 1);
  return (int64_t)((v ^ m) - m);
}

/* Comparison functors of Comparisons.h:115-300.  The double variants
 * compare with a 1e-6 margin exactly as ltdouble/ledouble/gtdouble/
 * gedouble do. */
static int compare_op(uint64_t field, uint32_t bits, int op, uint64_t k) {
  double x, y;
  switch (op) {
    case RFXO_EQ: return field == k;
    case RFXO_NE: return field != k;
    case RFXO_LT: return field < k;
    case RFXO_LE: return field <= k;
    case RFXO_GT: return field > k;
    case RFXO_GE: return field >= k;
    case RFXO_SLT: return sext(field, bits) < sext(k, bits);
    case RFXO_SLE: return sext(field, bits) <= sext(k, bits);
    case RFXO_SGT: return sext(field, bits) > sext(k, bits);
    case RFXO_SGE: return sext(field, bits) >= sext(k, bits);
    default: break;
  }
  memcpy(&x, &field, 8);
  memcpy(&y, &k, 8);
  switch (op) {
    case RFXO_DLT: return (y - x) > 1e-6;
    case RFXO_DLE: return (y - x) > -1e-6;
    case RFXO_DGT: return (x - y) > 1e-6;
    case RFXO_DGE: return (x - y) > -1e-6;
    default: return 0;
  }
}

/* select_field_constant (Select.h:30-114) followed by the histogram scan
 * and gather (Scan.h:18, Select.h:1055-1130): a stable filter. */
uint64_t rfxo_select(void* out, const void* in, uint64_t n, int tuple_bytes,
                     uint32_t field_shift, uint32_t field_bits, int op,
                     uint64_t constant) {
  uint64_t o = 0;
  for (uint64_t i = 0; i < n; ++i) {
    u128 t = ld(in, i, tuple_bytes);
    uint64_t f = mask_bits(shr(t, field_shift), field_bits).lo;
    if (compare_op(f, field_bits, op, constant)) st(out, o++, tuple_bytes, t);
  }
  return o;
}

/* permute (Project.h:58-130): each output field is an input field moved
 * to its new position. */
void rfxo_project(void* out, int out_bytes, const void* in, int in_bytes,
                  uint64_t n, int nfields, const uint32_t* in_shift,
                  const uint32_t* bits, const uint32_t* out_shift) {
  for (uint64_t i = 0; i < n; ++i) {
    u128 t = ld(in, i, in_bytes), o = {0, 0};
    for (int f = 0; f < nfields; ++f)
      o = or128(o, shl(mask_bits(shr(t, in_shift[f]), bits[f]), out_shift[f]));
    st(out, i, out_bytes, mask_bits(o, 8u * (uint32_t)out_bytes));
  }
}

/* split (Split_Merge.h:303-330): key = stripValues = in >> BitsAfter,
 * value = stripKeys = in & ((1 << BitsAfter) - 1). */
void rfxo_split(void* key, int key_bytes, void* val, int val_bytes,
                const void* in, int in_bytes, uint64_t n, uint32_t key_shift) {
  for (uint64_t i = 0; i < n; ++i) {
    u128 t = ld(in, i, in_bytes);
    st(key, i, key_bytes, mask_bits(shr(t, key_shift), 8u * (uint32_t)key_bytes));
    if (val) st(val, i, val_bytes, mask_bits(mask_bits(t, key_shift), 8u * (uint32_t)val_bytes));
  }
}

/* merge (Split_Merge.h:441-519): restoreValues(key) | value. */
void rfxo_merge(void* out, int out_bytes, const void* key, int key_bytes,
                const void* val, int val_bytes, uint64_t n, uint32_t key_shift) {
  for (uint64_t i = 0; i < n; ++i) {
    u128 o = shl(ld(key, i, key_bytes), key_shift);
    o = or128(o, ld(val, i, val_bytes));
    st(out, i, out_bytes, mask_bits(o, 8u * (uint32_t)out_bytes));
  }
}

/* ---- TPC-H Q1 ------------------------------------------------------ */

/* Q1 as planned in q1.man.cpp: filter shipdate <= 1998-12-01 - delta
 * (:959-1012), disc_price = price * (1 - disc) (:1735-1815), charge =
 * disc_price * (1 + tax) (:2110-2180), COUNT and TOTAL grouped by
 * (RETURNFLAG, LINESTATUS). */
void rfxo_q1(double* out, const int32_t* shipdate, const uint8_t* flag,
             const uint8_t* status, const int32_t* qty, const double* price,
             const double* disc, const double* tax, uint64_t n,
             int32_t cutoff, int compensated) {
  nsum acc[6][5];
  double plain[6][5];
  uint64_t count[6];
  memset(acc, 0, sizeof acc);
  memset(plain, 0, sizeof plain);
  memset(count, 0, sizeof count);
  for (uint64_t i = 0; i < n; ++i) {
    if (shipdate[i] > cutoff) continue;
    int g = flag[i] * 2 + status[i];
    double dp = price[i] * (1.0 - disc[i]);
    double ch = dp * (1.0 + tax[i]);
    double v[5] = {(double)qty[i], price[i], dp, ch, disc[i]};
    for (int f = 0; f < 5; ++f) { nsum_add(&acc[g][f], v[f]); plain[g][f] += v[f]; }
    count[g]++;
  }
  for (int g = 0; g < 6; ++g) {
    for (int f = 0; f < 5; ++f)
      out[g * 8 + f] = compensated ? acc[g][f].s + acc[g][f].c : plain[g][f];
    out[g * 8 + 5] = (double)count[g];
    out[g * 8 + 6] = 0.0;
    out[g * 8 + 7] = 0.0;
  }
}

/* ---- dates ------------------------------------------------------------ */
/* wrapdate (Date.h:17-25): months past 12 carry into the year, and the
 * month becomes month % 12 (so a multiple of 12 gives month 0). */
static void wrapdate(int* year, int* month) {
  if (*month > 12) {
    int inc_year = (*month - 1) / 12;
    *year += inc_year;
    *month = *month % 12;
  }
}

/* int2date (Date.h:27-90): days since 1992-01-01, then at most six year
 * steps (366, 365, 365, 365, 366, 365), then the month from cumulative
 * day counts (one more from March on in 1992 and 1996). */
void rfxo_int2date(uint64_t seconds, int* year, int* month, int* day) {
  uint64_t days = seconds / 86400 - 8035;
  unsigned year_int = 1992, month_int = 0;
  static const unsigned steps[6] = {366, 365, 365, 365, 366, 365};
  for (int s = 0; s < 6; ++s) {
    if (days >= steps[s]) {
      days -= steps[s];
      year_int++;
    } else {
      break;
    }
  }
  int month_days[13] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334, 365};
  if (year_int == 1992 || year_int == 1996)
    for (int i = 2; i < 13; ++i) month_days[i]++;
  for (int i = 1; i < 13; ++i) {
    if (days < (uint64_t)month_days[i] && days >= (uint64_t)month_days[i - 1]) {
      month_int = (unsigned)i;
      days -= (uint64_t)month_days[i - 1];
      break;
    }
  }
  *year = (int)year_int;
  *month = (int)month_int;
  *day = (int)(days + 1);
}

/* date2int (Date.h:92-118): 365-day years from 1970, months by fixed
 * lengths (February 28), plus the leap days of 1972-1988 and 1992/1996
 * by the cases Date.h lists. */
uint64_t rfxo_date2int(int year_int, int month_int, int date_int) {
  uint64_t days = (uint64_t)(int64_t)((year_int - 1970) * 365);
  static const int len[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  for (int m = 1; m <= 11; ++m)
    if (month_int > m) days += (uint64_t)len[m - 1];
  days += (uint64_t)(int64_t)(date_int - 1);
  if (year_int == 1992 && month_int <= 2) days += 5;
  if (year_int == 1992 && month_int > 2) days += 6;
  if (year_int > 1992 && year_int < 1996) days += 6;
  if (year_int == 1996 && month_int <= 2) days += 6;
  if (year_int == 1996 && month_int > 2) days += 7;
  if (year_int > 1996) days += 7;
  return days * 86400;
}

/* The date expressions of _arithSource (RelationalAlgebraKernel.cpp:
 * 2053-2123): AddMonth converts, adds to the month, wraps and converts
 * back; AddYear adds to the year; PartYear returns the year. */
void rfxo_date_op(int op, const uint64_t* in, uint64_t* out, uint64_t n, int64_t k) {
  for (uint64_t i = 0; i < n; ++i) {
    int year, month, day;
    rfxo_int2date(in[i], &year, &month, &day);
    if (op == 0) {
      month += (int)k;
      wrapdate(&year, &month);
      out[i] = rfxo_date2int(year, month, day);
    } else if (op == 1) {
      year += (int)k;
      out[i] = rfxo_date2int(year, month, day);
    } else {
      out[i] = (uint64_t)(int64_t)year;
    }
  }
}
