// q1.hip - TPC-H Q1 filter-group-aggregate in one pass over the lineitem
// columns, for gfx950.
//
// The reference plan (q1.man.cpp:959-2329) lowers Q1 to seven key joins,
// 24 sorts, 5 reduce_by_key and a count over packed tuples
// (RelationalAlgebraCompiler.cpp:7852-8641).  Here the same query - filter
// shipdate <= cutoff, disc_price = price * (1 - disc), charge = disc_price *
// (1 + tax), COUNT and TOTAL grouped by (returnflag, linestatus) - reads each
// of the seven columns once (34 bytes per row) and keeps the six groups'
// partial sums in registers.  Rows are dealt to a fixed grid (kQBlocks x
// 256 threads, four consecutive rows per step), and the partials are
// combined by fixed trees, so the FP64 sums depend on n alone.
#include "rfx_device.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

constexpr int kQNT = 256;
constexpr uint32_t kQBlocks = 2048;
constexpr int kQGroups = 6;
constexpr int kQSums = 5;  // qty, base price, disc price, charge, disc

struct Acc {
  double s[kQGroups][kQSums];
  uint32_t c[kQGroups];
};

__device__ __forceinline__ void add_row(Acc& a, int32_t sd, uint32_t g, int32_t qty, double price,
                                        double disc, double tax, int32_t cutoff) {
  if (sd > cutoff) return;
  const double dp = price * (1.0 - disc);
  const double ch = dp * (1.0 + tax);
  const double q = static_cast<double>(qty);
#pragma unroll
  for (int k = 0; k < kQGroups; ++k) {
    const bool m = g == static_cast<uint32_t>(k);
    a.s[k][0] += m ? q : 0.0;
    a.s[k][1] += m ? price : 0.0;
    a.s[k][2] += m ? dp : 0.0;
    a.s[k][3] += m ? ch : 0.0;
    a.s[k][4] += m ? disc : 0.0;
    a.c[k] += m;
  }
}

__global__ __launch_bounds__(kQNT) void k_q1_partial(
    const int32_t* __restrict__ shipdate, const uint8_t* __restrict__ flag,
    const uint8_t* __restrict__ status, const int32_t* __restrict__ qty,
    const double* __restrict__ price, const double* __restrict__ disc,
    const double* __restrict__ tax, uint64_t n, int32_t cutoff, double* __restrict__ part) {
  __shared__ double red[kQNT / kWave][kQGroups * (kQSums + 1)];
  Acc a;
#pragma unroll
  for (int k = 0; k < kQGroups; ++k) {
#pragma unroll
    for (int f = 0; f < kQSums; ++f) a.s[k][f] = 0.0;
    a.c[k] = 0;
  }
  const uint64_t gtid = blockIdx.x * static_cast<uint64_t>(kQNT) + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(kQBlocks) * kQNT;
  // Four consecutive rows per step: 16-byte loads of the 4-byte columns,
  // 4-byte loads of the byte columns, 2 x 16 bytes of each double column.
  const uint64_t n4 = n / 4;
  for (uint64_t v = gtid; v < n4; v += stride) {
    const int4 sd = reinterpret_cast<const int4*>(shipdate)[v];
    const uint32_t fl = reinterpret_cast<const uint32_t*>(flag)[v];
    const uint32_t st = reinterpret_cast<const uint32_t*>(status)[v];
    const int4 q = reinterpret_cast<const int4*>(qty)[v];
    const double2 p0 = reinterpret_cast<const double2*>(price)[2 * v];
    const double2 p1 = reinterpret_cast<const double2*>(price)[2 * v + 1];
    const double2 d0 = reinterpret_cast<const double2*>(disc)[2 * v];
    const double2 d1 = reinterpret_cast<const double2*>(disc)[2 * v + 1];
    const double2 t0 = reinterpret_cast<const double2*>(tax)[2 * v];
    const double2 t1 = reinterpret_cast<const double2*>(tax)[2 * v + 1];
    add_row(a, sd.x, (fl & 0xFF) * 2 + (st & 0xFF), q.x, p0.x, d0.x, t0.x, cutoff);
    add_row(a, sd.y, ((fl >> 8) & 0xFF) * 2 + ((st >> 8) & 0xFF), q.y, p0.y, d0.y, t0.y, cutoff);
    add_row(a, sd.z, ((fl >> 16) & 0xFF) * 2 + ((st >> 16) & 0xFF), q.z, p1.x, d1.x, t1.x, cutoff);
    add_row(a, sd.w, (fl >> 24) * 2 + (st >> 24), q.w, p1.y, d1.y, t1.y, cutoff);
  }
  for (uint64_t i = 4 * n4 + gtid; i < n; i += stride)
    add_row(a, shipdate[i], flag[i] * 2u + status[i], qty[i], price[i], disc[i], tax[i], cutoff);

  // Block sum of each of the 36 values: a wave tree, then the waves in order.
  const uint32_t lane = lane_id(), w = wave_id();
#pragma unroll
  for (int k = 0; k < kQGroups; ++k) {
#pragma unroll
    for (int f = 0; f < kQSums; ++f) {
      const double s = wave_sum(a.s[k][f]);
      if (lane == 0) red[w][k * (kQSums + 1) + f] = s;
    }
    const double c = static_cast<double>(wave_sum(a.c[k]));
    if (lane == 0) red[w][k * (kQSums + 1) + kQSums] = c;
  }
  __syncthreads();
  if (threadIdx.x < kQGroups * (kQSums + 1)) {
    double s = 0.0;
    for (int ww = 0; ww < kQNT / kWave; ++ww) s += red[ww][threadIdx.x];
    part[static_cast<uint64_t>(threadIdx.x) * kQBlocks + blockIdx.x] = s;
  }
}

// One block per aggregate value: a fixed tree over the kQBlocks partials.
__global__ __launch_bounds__(1024) void k_q1_final(const double* __restrict__ part,
                                                   double* __restrict__ out) {
  __shared__ double sm[1024];
  const uint32_t v = blockIdx.x;  // k * 6 + f
  const double* p = part + static_cast<uint64_t>(v) * kQBlocks;
  sm[threadIdx.x] = p[threadIdx.x] + p[threadIdx.x + 1024];
  __syncthreads();
  for (int d = 512; d > 0; d >>= 1) {
    if (threadIdx.x < static_cast<uint32_t>(d)) sm[threadIdx.x] += sm[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int k = v / (kQSums + 1), f = v % (kQSums + 1);
    out[k * 8 + f] = sm[0];
  }
}

__global__ void k_q1_avgs(double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k < kQGroups) {
    const double c = out[k * 8 + 5];
    out[k * 8 + 6] = c > 0 ? out[k * 8 + 0] / c : 0.0;
    out[k * 8 + 7] = c > 0 ? out[k * 8 + 1] / c : 0.0;
  }
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" int rfx_q1(rfx_ctx* ctx, double* out, const int32_t* shipdate,
                      const uint8_t* returnflag, const uint8_t* linestatus, const int32_t* qty,
                      const double* price, const double* disc, const double* tax, uint64_t n,
                      int32_t cutoff) {
  if (!ctx || !out) return fail(RFX_EINVAL, "rfx_q1: null argument");
  if (n && (!shipdate || !returnflag || !linestatus || !qty || !price || !disc || !tax))
    return fail(RFX_EINVAL, "rfx_q1: null column");
  const uintptr_t al = reinterpret_cast<uintptr_t>(shipdate) | reinterpret_cast<uintptr_t>(qty) |
                       reinterpret_cast<uintptr_t>(price) | reinterpret_cast<uintptr_t>(disc) |
                       reinterpret_cast<uintptr_t>(tax);
  const uintptr_t al4 = reinterpret_cast<uintptr_t>(returnflag) |
                        reinterpret_cast<uintptr_t>(linestatus);
  if ((al % 16) || (al4 % 4))
    return fail(RFX_EINVAL, "rfx_q1: columns must be 16-byte aligned (byte columns 4-byte)");
  if (n >= (1ull << 32)) return fail(RFX_ERANGE, "rfx_q1: too many rows");
  Carve c0(nullptr);
  c0.take<double>(static_cast<uint64_t>(kQGroups) * (kQSums + 1) * kQBlocks);
  RFX_TRY(ensure_workspace(ctx, c0.off));
  Carve w(ctx->ws);
  double* part = w.take<double>(static_cast<uint64_t>(kQGroups) * (kQSums + 1) * kQBlocks);
  {
    RFX_PROF(ctx, "q1.partial");
    k_q1_partial<<<kQBlocks, kQNT, 0, ctx->stream>>>(shipdate, returnflag, linestatus, qty, price,
                                                       disc, tax, n, cutoff, part);
    RFX_LAUNCHED();
  }
  k_q1_final<<<kQGroups * (kQSums + 1), 1024, 0, ctx->stream>>>(part, out);
  RFX_LAUNCHED();
  k_q1_avgs<<<1, 64, 0, ctx->stream>>>(out);
  RFX_LAUNCHED();
  return RFX_OK;
}
