// rfx_cpu.cpp - the reported CPU baseline: the reference's operator
// algorithms restated in C++ and run with OpenMP on the host's cores.
//
// TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg (and its
// CPU test), never by the product.  The reference's own CPU path (Harmony
// kernels through GPU-Ocelot's LLVM x86 backend) cannot be built here, so
// this is a baseline of the same algorithms, not of that path:
//
//   sort     stable LSD radix sort, 8-bit digits (b40cSort.cu:58-1021): per
//            pass each thread counts its contiguous chunk, the counts are
//            scanned digit-major then thread-minor, each thread scatters its
//            chunk in order (stable);
//   join     sorted search of each left key (lower and upper bound,
//            ModernGPUJoin.cu:504-590), exclusive scan of the counts,
//            left-major expansion with the gather of (key, r_val, s_val)
//            (ModernGPUJoin.cu:665-733, join.cuh:45-99, ModernGPUJoin.h:
//            43-200);
//   reduce   reduce-by-key over runs of equal adjacent keys with Neumaier
//            FP64 sums and counts (Reduce.cu:175-1532): chunks start at run
//            heads, a count pass, a scan, a write pass;
//   q1       filter and 6-group FP64 sums (q1.man.cpp:959-1815), per-thread
//            Neumaier partials combined in thread order;
//   select   stable compaction (Select.h:30-1130): count pass, scan, write;
//   mapfilter the same compaction computing price*(1-disc) and
//            price*(1-disc)*(1+tax) for the kept rows;
//   set ops  std::set_union / set_intersection / set_difference (Union.cu:
//            47-148, Difference.cu:35-143) on key ranges split at A's first
//            occurrences, so equal keys stay in one piece.
//
// Every rfxc_* timing entry runs the operator `reps` times after one
// warm-up, copying its inputs back before each run outside the timed
// region, and returns the median wall time in seconds (steady_clock).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

template <class Prep, class Run>
double median_seconds(int reps, Prep prep, Run run) {
  std::vector<double> t;
  for (int r = 0; r <= reps; ++r) {
    prep();
    const auto t0 = clk::now();
    run();
    const auto t1 = clk::now();
    if (r) t.push_back(std::chrono::duration<double>(t1 - t0).count());
  }
  std::sort(t.begin(), t.end());
  return t.empty() ? 0.0 : t[t.size() / 2];
}

// [lo, hi) of chunk `c` of `parts` over n items.
inline void chunk(uint64_t n, int parts, int c, uint64_t* lo, uint64_t* hi) {
  *lo = n * c / parts;
  *hi = n * (c + 1) / parts;
}

template <class K, class V>
void lsd_sort(K* keys, V* vals, K* tk, V* tv, uint64_t n, int key_bits) {
  const int passes = (key_bits + 7) / 8;
  const int T = omp_get_max_threads();
  std::vector<uint64_t> cnt(static_cast<size_t>(T) * 256);
  K* src_k = keys;
  V* src_v = vals;
  K* dst_k = tk;
  V* dst_v = tv;
  for (int p = 0; p < passes; ++p) {
    const int sh = 8 * p;
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(n, T, t, &lo, &hi);
      uint64_t* c = &cnt[static_cast<size_t>(t) * 256];
      std::fill(c, c + 256, 0);
      for (uint64_t i = lo; i < hi; ++i) ++c[(src_k[i] >> sh) & 0xFF];
#pragma omp barrier
#pragma omp single
      {
        uint64_t run = 0;
        for (int d = 0; d < 256; ++d)
          for (int u = 0; u < T; ++u) {
            const uint64_t x = cnt[static_cast<size_t>(u) * 256 + d];
            cnt[static_cast<size_t>(u) * 256 + d] = run;
            run += x;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) {
        const uint64_t o = c[(src_k[i] >> sh) & 0xFF]++;
        dst_k[o] = src_k[i];
        dst_v[o] = src_v[i];
      }
    }
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (src_k != keys) {
    std::memcpy(keys, src_k, n * sizeof(K));
    std::memcpy(vals, src_v, n * sizeof(V));
  }
}

// Exclusive scan of per-chunk totals in place; returns the total.
uint64_t scan_chunks(std::vector<uint64_t>& v) {
  uint64_t run = 0;
  for (auto& x : v) {
    const uint64_t y = x;
    x = run;
    run += y;
  }
  return run;
}

struct Neumaier {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    if ((s >= 0 ? s : -s) >= (x >= 0 ? x : -x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

}  // namespace

extern "C" {

int rfxc_threads(void) { return omp_get_max_threads(); }
void rfxc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// Sort of (u32 key, u32 value) or (u64 key, u32 value) pairs in place.
double rfxc_sort_pairs(void* keys, uint32_t* vals, const void* src_keys, const uint32_t* src_vals,
                       uint64_t n, int key_bytes, int reps) {
  std::vector<uint8_t> tk(n * key_bytes);
  std::vector<uint32_t> tv(n);
  auto prep = [&] {
    std::memcpy(keys, src_keys, n * key_bytes);
    std::memcpy(vals, src_vals, n * 4);
  };
  if (key_bytes == 4)
    return median_seconds(reps, prep, [&] {
      lsd_sort(static_cast<uint32_t*>(keys), vals, reinterpret_cast<uint32_t*>(tk.data()),
               tv.data(), n, 32);
    });
  return median_seconds(reps, prep, [&] {
    lsd_sort(static_cast<uint64_t*>(keys), vals, reinterpret_cast<uint64_t*>(tk.data()),
             tv.data(), n, 64);
  });
}

// Join core over sorted u32 key columns: out (2 u64 per row) receives
// (key << 64) | (r_val << 32) | s_val in (left row, right row) order, at
// most cap rows; *total the match count.
double rfxc_join(uint64_t* out, uint64_t cap, uint64_t* total, const uint32_t* rk,
                 const uint32_t* rv, uint64_t nr, const uint32_t* sk, const uint32_t* sv,
                 uint64_t ns, int reps) {
  std::vector<uint32_t> lb(nr), cnt(nr);
  std::vector<uint64_t> start(nr);
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<uint64_t> part(T);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(nr, T, t, &lo, &hi);
      uint64_t sum = 0;
      for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t* l = std::lower_bound(sk, sk + ns, rk[i]);
        const uint32_t* u = std::upper_bound(l, sk + ns, rk[i]);
        lb[i] = static_cast<uint32_t>(l - sk);
        cnt[i] = static_cast<uint32_t>(u - l);
        sum += cnt[i];
      }
      part[t] = sum;
#pragma omp barrier
#pragma omp single
      *total = scan_chunks(part);
      uint64_t run = part[t];
      for (uint64_t i = lo; i < hi; ++i) {
        start[i] = run;
        run += cnt[i];
      }
      for (uint64_t i = lo; i < hi; ++i)
        for (uint32_t j = 0; j < cnt[i]; ++j) {
          const uint64_t g = start[i] + j;
          if (g >= cap) break;
          out[2 * g] = (static_cast<uint64_t>(rv[i]) << 32) | sv[lb[i] + j];
          out[2 * g + 1] = rk[i];
        }
    }
  });
}

// GROUP BY over sorted u32 keys: SUM(v) (Neumaier) and COUNT(*) per run.
double rfxc_reduce_sum_count(uint32_t* out_keys, double* out_sums, uint32_t* out_counts,
                             uint64_t* runs, const uint32_t* keys, const double* v, uint64_t n,
                             int reps) {
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<uint64_t> first(T + 1), part(T);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(n, T, t, &lo, &hi);
      // A chunk starts at the first run head at or after its nominal start.
      while (lo > 0 && lo < n && keys[lo] == keys[lo - 1]) ++lo;
      first[t] = lo;
      if (t == T - 1) first[T] = n;
#pragma omp barrier
      const uint64_t a = first[t], b = first[t + 1] > a ? first[t + 1] : a;
      uint64_t heads = 0;
      for (uint64_t i = a; i < b; ++i) heads += (i == a || keys[i] != keys[i - 1]);
      part[t] = heads;
#pragma omp barrier
#pragma omp single
      *runs = scan_chunks(part);
      uint64_t r = part[t];
      uint64_t i = a;
      while (i < b) {
        Neumaier s;
        uint64_t j = i;
        for (; j < b && keys[j] == keys[i]; ++j) s.add(v[j]);
        out_keys[r] = keys[i];
        out_sums[r] = s.value();
        out_counts[r] = static_cast<uint32_t>(j - i);
        ++r;
        i = j;
      }
    }
  });
}

// TPC-H Q1 over columns, the oracle's output layout (6 groups x 8 doubles).
double rfxc_q1(double* out, const int32_t* shipdate, const uint8_t* flag, const uint8_t* status,
               const int32_t* qty, const double* price, const double* disc, const double* tax,
               uint64_t n, int32_t cutoff, int reps) {
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<Neumaier> acc(static_cast<size_t>(T) * 30);
    std::vector<uint64_t> count(static_cast<size_t>(T) * 6, 0);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(n, T, t, &lo, &hi);
      Neumaier a[30];
      uint64_t c[6] = {0, 0, 0, 0, 0, 0};
      for (uint64_t i = lo; i < hi; ++i) {
        if (shipdate[i] > cutoff) continue;
        const int g = flag[i] * 2 + status[i];
        const double dp = price[i] * (1.0 - disc[i]);
        const double ch = dp * (1.0 + tax[i]);
        a[g * 5 + 0].add(static_cast<double>(qty[i]));
        a[g * 5 + 1].add(price[i]);
        a[g * 5 + 2].add(dp);
        a[g * 5 + 3].add(ch);
        a[g * 5 + 4].add(disc[i]);
        ++c[g];
      }
      for (int k = 0; k < 30; ++k) acc[static_cast<size_t>(t) * 30 + k] = a[k];
      for (int g = 0; g < 6; ++g) count[static_cast<size_t>(t) * 6 + g] = c[g];
    }
    for (int g = 0; g < 6; ++g) {
      uint64_t c = 0;
      for (int f = 0; f < 5; ++f) {
        Neumaier s;
        for (int t = 0; t < T; ++t) {
          s.add(acc[static_cast<size_t>(t) * 30 + g * 5 + f].s);
          s.add(acc[static_cast<size_t>(t) * 30 + g * 5 + f].c);
        }
        out[g * 8 + f] = s.value();
      }
      for (int t = 0; t < T; ++t) c += count[static_cast<size_t>(t) * 6 + g];
      out[g * 8 + 5] = static_cast<double>(c);
      out[g * 8 + 6] = out[g * 8 + 7] = 0.0;
    }
  });
}

// Stable selection of 16-byte tuples (lo, hi u64 words) whose 32-bit field
// at bit 64 (the high word's low half) is below k.
double rfxc_select16(uint64_t* out, uint64_t* kept, const uint64_t* in, uint64_t n, uint32_t k,
                     int reps) {
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<uint64_t> part(T);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(n, T, t, &lo, &hi);
      uint64_t c = 0;
      for (uint64_t i = lo; i < hi; ++i) c += static_cast<uint32_t>(in[2 * i + 1]) < k;
      part[t] = c;
#pragma omp barrier
#pragma omp single
      *kept = scan_chunks(part);
      uint64_t o = part[t];
      for (uint64_t i = lo; i < hi; ++i)
        if (static_cast<uint32_t>(in[2 * i + 1]) < k) {
          out[2 * o] = in[2 * i];
          out[2 * o + 1] = in[2 * i + 1];
          ++o;
        }
    }
  });
}

// WHERE shipdate <= cutoff: disc_price = price*(1-disc), charge =
// disc_price*(1+tax) for the kept rows, in row order.
double rfxc_mapfilter(double* out_dp, double* out_ch, uint64_t* kept, const int32_t* shipdate,
                      const double* price, const double* disc, const double* tax, uint64_t n,
                      int32_t cutoff, int reps) {
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<uint64_t> part(T);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      uint64_t lo, hi;
      chunk(n, T, t, &lo, &hi);
      uint64_t c = 0;
      for (uint64_t i = lo; i < hi; ++i) c += shipdate[i] <= cutoff;
      part[t] = c;
#pragma omp barrier
#pragma omp single
      *kept = scan_chunks(part);
      uint64_t o = part[t];
      for (uint64_t i = lo; i < hi; ++i)
        if (shipdate[i] <= cutoff) {
          const double dp = price[i] * (1.0 - disc[i]);
          out_dp[o] = dp;
          out_ch[o] = dp * (1.0 + tax[i]);
          ++o;
        }
    }
  });
}

// Multiset union (op 0), difference (1) or intersection (2) of sorted u64
// columns.  A is split at the first occurrences of evenly spaced keys and B
// at the same keys, so equal keys stay in one piece.
double rfxc_set_op(uint64_t* out, uint64_t* nout, const uint64_t* a, uint64_t na,
                   const uint64_t* b, uint64_t nb, int op, int reps) {
  return median_seconds(reps, [] {}, [&] {
    const int T = omp_get_max_threads();
    std::vector<uint64_t> as(T + 1), bs(T + 1), part(T);
    as[0] = bs[0] = 0;
    as[T] = na;
    bs[T] = nb;
    for (int t = 1; t < T; ++t) {
      const uint64_t i = na * t / T;
      if (i >= na) {
        as[t] = na;
        bs[t] = nb;
        continue;
      }
      as[t] = std::lower_bound(a, a + na, a[i]) - a;
      bs[t] = std::lower_bound(b, b + nb, a[i]) - b;
    }
    auto piece = [&](int t, uint64_t* dst) -> uint64_t {
      const uint64_t *a0 = a + as[t], *a1 = a + std::max(as[t], as[t + 1]);
      const uint64_t *b0 = b + bs[t], *b1 = b + std::max(bs[t], bs[t + 1]);
      if (!dst) {  // count only
        uint64_t c = 0;
        struct Counter {
          uint64_t* c;
          Counter& operator*() { return *this; }
          Counter& operator=(uint64_t) { ++*c; return *this; }
          Counter& operator++() { return *this; }
          Counter operator++(int) { return *this; }
        } it{&c};
        if (op == 0) std::set_union(a0, a1, b0, b1, it);
        else if (op == 1) std::set_difference(a0, a1, b0, b1, it);
        else std::set_intersection(a0, a1, b0, b1, it);
        return c;
      }
      uint64_t* e;
      if (op == 0) e = std::set_union(a0, a1, b0, b1, dst);
      else if (op == 1) e = std::set_difference(a0, a1, b0, b1, dst);
      else e = std::set_intersection(a0, a1, b0, b1, dst);
      return static_cast<uint64_t>(e - dst);
    };
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      part[t] = piece(t, nullptr);
#pragma omp barrier
#pragma omp single
      *nout = scan_chunks(part);
      piece(t, out + part[t]);
    }
  });
}

}  // extern "C"
