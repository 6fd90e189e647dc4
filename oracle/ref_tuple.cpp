// ref_tuple.cpp - C entry points over the reference's own tuple codec
// (redfox/redfox/ra/interface/Tuple.h, compiled from the reference
// checkout by oracle/Makefile into oracle/_ref/libref_tuple.so).
//
// TEST INFRASTRUCTURE ONLY: tests/test_oracle.py compares the oracle's
// codec (rfx_oracle.c) with these on every field of a fixed set of tuple
// shapes.  Words are passed as two little-endian u64 (w[0] low).
#include <stdint.h>

#include <redfox/ra/interface/Tuple.h>

using namespace ra::tuple;

typedef Tuple<32, 32> S0;          // config-1 rows, TPC-H/data/A
typedef Tuple<32, 32, 32> S1;      // join output (key, r_val, s_val)
typedef Tuple<2, 1, 64> S2;        // Q1 aggregate output
typedef Tuple<16, 8, 8> S3;        // 32-bit word
typedef Tuple<3, 23, 38> S4;       // 64-bit word, odd widths
typedef Tuple<64, 64> S5;          // 128-bit word, word-aligned fields
typedef Tuple<5, 3> S6;            // 8-bit word

static const int kFields[] = {2, 3, 3, 3, 3, 2, 2};

template <class B> struct W;
template <> struct W<unsigned char> {
  static unsigned char from(const uint64_t* w) { return (unsigned char)w[0]; }
  static void to(unsigned char v, uint64_t* w) { w[0] = v; w[1] = 0; }
};
template <> struct W<unsigned short> {
  static unsigned short from(const uint64_t* w) { return (unsigned short)w[0]; }
  static void to(unsigned short v, uint64_t* w) { w[0] = v; w[1] = 0; }
};
template <> struct W<unsigned int> {
  static unsigned int from(const uint64_t* w) { return (unsigned int)w[0]; }
  static void to(unsigned int v, uint64_t* w) { w[0] = v; w[1] = 0; }
};
template <> struct W<unsigned long long> {
  static unsigned long long from(const uint64_t* w) { return w[0]; }
  static void to(unsigned long long v, uint64_t* w) { w[0] = v; w[1] = 0; }
};
template <> struct W<PackedNBytes<2> > {
  static PackedNBytes<2> from(const uint64_t* w) {
    PackedNBytes<2> p;
    p.a[0] = w[0];
    p.a[1] = w[1];
    return p;
  }
  static void to(const PackedNBytes<2>& v, uint64_t* w) { w[0] = v.a[0]; w[1] = v.a[1]; }
};

template <class S, unsigned I>
static uint64_t ex(const uint64_t* w) {
  typedef typename S::BasicType B;
  B t = W<B>::from(w);
  return (uint64_t)extract<unsigned long long, I, S>(t);
}
template <class S, unsigned I>
static void in(uint64_t* w, uint64_t v) {
  typedef typename S::BasicType B;
  B t = W<B>::from(w);
  t = insert<unsigned long long, I, S>(t, (unsigned long long)v);
  W<B>::to(t, w);
}
template <class S, unsigned K>
static void sv(const uint64_t* w, uint64_t* o) {
  typedef typename S::BasicType B;
  W<B>::to(stripValues<S, K>(W<B>::from(w)), o);
}
template <class S, unsigned K>
static void sk(const uint64_t* w, uint64_t* o) {
  typedef typename S::BasicType B;
  W<B>::to(stripKeys<S, K>(W<B>::from(w)), o);
}
template <class S, unsigned K>
static void rv(const uint64_t* w, uint64_t* o) {
  typedef typename S::BasicType B;
  W<B>::to(restoreValues<S, K>(W<B>::from(w)), o);
}

#define FOR_SHAPES(X) X(0, S0) X(1, S1) X(2, S2) X(3, S3) X(4, S4) X(5, S5) X(6, S6)

extern "C" {

int ref_shape_count(void) { return 7; }
int ref_shape_fields(int s) { return (s >= 0 && s < 7) ? kFields[s] : 0; }
int ref_shape_bytes(int s) {
#define B(i, S) if (s == i) return (int)sizeof(typename S::BasicType);
  FOR_SHAPES(B)
#undef B
  return 0;
}
int ref_shape_bits(int s, int f) {
#define B(i, S)                          \
  if (s == i) {                          \
    if (f == 0) return S::bits0;         \
    if (f == 1) return S::bits1;         \
    if (f == 2) return S::bits2;         \
    return 0;                            \
  }
  FOR_SHAPES(B)
#undef B
  return 0;
}

uint64_t ref_extract(int s, const uint64_t* w, int f) {
#define B(i, S)                                 \
  if (s == i) {                                 \
    if (f == 0) return ex<S, 0>(w);             \
    if (f == 1) return ex<S, 1>(w);             \
    if (f == 2) return ex<S, 2>(w);             \
  }
  FOR_SHAPES(B)
#undef B
  return 0;
}

void ref_insert(int s, uint64_t* w, int f, uint64_t v) {
#define B(i, S)                                 \
  if (s == i) {                                 \
    if (f == 0) in<S, 0>(w, v);                 \
    if (f == 1) in<S, 1>(w, v);                 \
    if (f == 2) in<S, 2>(w, v);                 \
  }
  FOR_SHAPES(B)
#undef B
}

// keys = number of leading key fields (1 or 2).
void ref_strip_values(int s, int keys, const uint64_t* w, uint64_t* o) {
#define B(i, S)                                 \
  if (s == i) {                                 \
    if (keys == 1) sv<S, 1>(w, o);              \
    if (keys == 2) sv<S, 2>(w, o);              \
  }
  FOR_SHAPES(B)
#undef B
}
void ref_strip_keys(int s, int keys, const uint64_t* w, uint64_t* o) {
#define B(i, S)                                 \
  if (s == i) {                                 \
    if (keys == 1) sk<S, 1>(w, o);              \
    if (keys == 2) sk<S, 2>(w, o);              \
  }
  FOR_SHAPES(B)
#undef B
}
void ref_restore_values(int s, int keys, const uint64_t* w, uint64_t* o) {
#define B(i, S)                                 \
  if (s == i) {                                 \
    if (keys == 1) rv<S, 1>(w, o);              \
    if (keys == 2) rv<S, 2>(w, o);              \
  }
  FOR_SHAPES(B)
#undef B
}

}  // extern "C"
