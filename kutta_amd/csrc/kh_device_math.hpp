// Device-side arithmetic shared by the kernels. Everything is compiled with
// -ffp-contract=off and without fast-math: a*b+c is a rounded product then a
// rounded sum, division and sqrt are correctly rounded, exactly as the serial
// reference evaluates them.
#pragma once

#include <hip/hip_runtime.h>

#include "kh_internal.hpp"

namespace kh {

// ----------------------------------------------------------- min and max
// kmax/kmin of kernels/prelude.h as the reference's serial CPU build evaluates
// them (x86-64 std::fmax/fmin): a NaN operand yields the other operand, and
// otherwise the first operand is returned only when strictly larger (smaller),
// so equal operands, +0 and -0 included, yield the second. The hardware
// v_max_f32/v_min_f32 order -0 below +0 and would flip the sign of zero
// results (limiters, clamps) against the reference.
__device__ __forceinline__ float kmax(float a, float b) {
  return a != a ? b : (b != b ? a : (a > b ? a : b));
}
__device__ __forceinline__ float kmin(float a, float b) {
  return a != a ? b : (b != b ? a : (a < b ? a : b));
}

// ------------------------------------------------------------ upward rounding
// Upper bound of a float sum (kernels/prelude.h:82-100): TwoSum, then step up
// one ulp only when round-to-nearest went down.
__device__ __forceinline__ float add_up(float a, float b) {
  float s = a + b;
  float part = s - a;
  float err = (a - (s - part)) + (b - part);
  return err > 0.0f ? nextafterf(s, __builtin_inff()) : s;
}
struct UpperSum {
  float sum, error;
};
__device__ __forceinline__ UpperSum upper_add(UpperSum a, float b) {
  float s = a.sum + b;
  float part = s - a.sum;
  float err = (a.sum - (s - part)) + (b - part);
  return {s, add_up(a.error, err)};
}
__device__ __forceinline__ float upper_value(UpperSum a) { return add_up(a.sum, a.error); }

// ------------------------------------------------- compensated row operations
// Dekker split product error (kernels/linalg.kern.h:21-30). When |a|, |b| are
// within the reference's 1e30 guard and 2^-100 <= |a*b| <= 2^100, no split
// piece, partial product or sum overflows or underflows, so Dekker's result
// is the exact rounding error of a*b, which is also what one FMA returns; that
// path costs one instruction instead of twelve. Outside the window the split
// runs as the reference writes it.
__device__ __forceinline__ float product_error(float a, float b, float product) {
  const float m = fabsf(product);
  if (m >= 0x1p-100f && m <= 0x1p100f && fabsf(a) <= 1.0e30f && fabsf(b) <= 1.0e30f)
    return __builtin_fmaf(a, b, -product);
  if (fabsf(a) > 1.0e30f || fabsf(b) > 1.0e30f) return 0.0f;
  float ca = 4097.0f * a, cb = 4097.0f * b;
  float ah = ca - (ca - a), bh = cb - (cb - b);
  float al = a - ah, bl = b - bh;
  return ((ah * bh - product) + ah * bl + al * bh) + al * bl;
}

// Neumaier accumulator for b - sum a*x with product errors
// (kernels/linalg.kern.h:32-45). Callers feed entries in row order.
struct RowResidual {
  float s, e;
  __device__ __forceinline__ explicit RowResidual(float b) : s(b), e(0.0f) {}
  __device__ __forceinline__ void add(float a, float x) {
    float prod = a * x;
    float t = -prod;
    float nx = s + t;
    e += fabsf(s) >= fabsf(t) ? (s - nx) + t : (t - nx) + s;
    e -= product_error(a, x, prod);
    s = nx;
  }
  // The same step with the entry's term t = -(a x) and its product error
  // formed beforehand (by another thread): the serial part of add().
  __device__ __forceinline__ void add_split(float t, float error) {
    float nx = s + t;
    e += fabsf(s) >= fabsf(t) ? (s - nx) + t : (t - nx) + s;
    e -= error;
    s = nx;
  }
  __device__ __forceinline__ float value() const { return s + e; }
};

// Kahan accumulator (kernels/linalg.kern.h:5-16, kernels/amg.kern.h:5-16).
struct Kahan {
  float s = 0.0f, e = 0.0f;
  __device__ __forceinline__ void add(float term_in) {
    float t = term_in - e;
    float nx = s + t;
    e = (nx - s) - t;
    s = nx;
  }
};

// Neumaier accumulator (kernels/amg.kern.h:27-38).
struct Neumaier {
  float s = 0.0f, e = 0.0f;
  __device__ __forceinline__ void add(float t) {
    float nx = s + t;
    e += fabsf(s) >= fabsf(t) ? (s - nx) + t : (t - nx) + s;
    s = nx;
  }
  __device__ __forceinline__ float value() const { return s + e; }
};

// ----------------------------------------------------------- scalar pairs
// Normalized (hi, lo) float pairs (kernels/scalars.kern.h:3-42).
struct Pair {
  float hi, lo;
};
__device__ __forceinline__ Pair pload(const float* p, int off) { return {p[off], p[off + 1]}; }
__device__ __forceinline__ void pstore(float* p, int off, Pair v) {
  p[off] = v.hi;
  p[off + 1] = v.lo;
}
__device__ __forceinline__ Pair padd(Pair a, Pair b) {
  float s = a.hi + b.hi;
  if (s - s != 0.0f) return {s, 0.0f};
  float v = s - a.hi;
  float e = (a.hi - (s - v)) + (b.hi - v);
  e = e + a.lo + b.lo;
  float h = s + e;
  return {h, e - (h - s)};
}
__device__ __forceinline__ Pair pneg(Pair a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ Pair pmul(Pair a, Pair b) {
  float h = a.hi * b.hi;
  float l = __builtin_fmaf(a.hi, b.hi, -h) + a.hi * b.lo + a.lo * b.hi;
  return padd({h, 0.0f}, {l, 0.0f});
}
__device__ __forceinline__ Pair pdiv(Pair a, Pair b) {
  float q = a.hi / b.hi;
  Pair r = padd(a, pneg(pmul({q, 0.0f}, b)));
  return padd({q, 0.0f}, {(r.hi + r.lo) / b.hi, 0.0f});
}
__device__ __forceinline__ float pfloat(Pair a) { return a.hi + a.lo; }
__device__ __forceinline__ bool ppositive(Pair a) {
  return a.hi > 0.0f || (a.hi == 0.0f && a.lo > 0.0f);
}
__device__ __forceinline__ bool pfinite(Pair a) {
  return a.hi - a.hi == 0.0f && a.lo - a.lo == 0.0f;
}
__device__ __forceinline__ bool pzero(Pair a) { return a.hi == 0.0f && a.lo == 0.0f; }
__device__ __forceinline__ bool ple(Pair a, Pair b) {
  return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo);
}
__device__ __forceinline__ bool pcoefficient(Pair a) {
  float v = pfloat(a);
  return pfinite(a) && v - v == 0.0f && (v != 0.0f || pzero(a));
}
// FP64 value to a normalized pair (runtime/cpu/cpu_device.cpp:83-86).
__device__ __forceinline__ Pair to_pair(double value) {
  float hi = static_cast<float>(value);
  float lo = (hi - hi == 0.0f) ? static_cast<float>(value - double(hi)) : 0.0f;
  return {hi, lo};
}

// ------------------------------------------------------- scalar state machine
// Single-lane bodies of kernels/scalars.kern.h:63-121, shared by the plain
// scalar kernels and by the fused kernels' last workgroup.
__device__ __forceinline__ void pcg_alpha(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rho = pload(s, 0), pq = pload(s, 2);
  Pair alpha = pdiv(rho, pq);
  if (!ppositive(rho) || !ppositive(pq) || !pcoefficient(alpha) || pzero(alpha)) status[0] = 2;
  else pstore(s, 8, alpha);
}
__device__ __forceinline__ void pcg_beta(float* s, int* status, int flexible) {
  if (status[0] != 0) return;
  Pair next = pload(s, 4);
  Pair numer = flexible != 0 ? pneg(pmul(pload(s, 8), pload(s, 6))) : next;
  Pair beta = pdiv(numer, pload(s, 0));
  if (!ppositive(next) || !pcoefficient(beta) || (pzero(beta) && !pzero(numer))) status[0] = 2;
  else {
    pstore(s, 10, beta);
    pstore(s, 0, next);
  }
}
__device__ __forceinline__ void solver_check_body(float* s, int* status, int pcg) {
  if (status[0] != 0) return;
  if (pcg != 0) ++status[1];
  Pair rr = pload(s, 18);
  if (!pfinite(rr) || ple(rr, pload(s, 20)) || !ple(rr, pload(s, 22)) || status[1] % 50 == 0)
    status[0] = 1;
  if (pcg == 0) pstore(s, 14, pload(s, 0));
}

// ---------------------------------------------------------- block reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Deterministic sum over the workgroup; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds /* kBlock/64 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double total = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) total += lds[w];
  __syncthreads();
  return total;
}

// Grid reductions are two launches: each workgroup stores its partial sums at
// a fixed slot (no fences, no atomics), and a one-workgroup finisher sums the
// slots in a fixed order after the kernel boundary has made them visible.
// A device-scope fence per workgroup would write back and invalidate the L2 of
// every XCD and costs far more than the second launch.
template <int N>
__device__ __forceinline__ void store_partials(const double (&mine)[N], double* partials) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) partials[blockIdx.x * N + k] = mine[k];
  }
}
// Finisher body: the totals of `count` partial slots, valid on thread 0. Each
// thread adds slots threadIdx.x, +blockDim.x, ... in that order; the loads of
// kSumAhead consecutive slots of a thread are issued before their adds, so
// the finisher pays one memory round trip per kSumAhead slots instead of one
// per slot (a few thousand partials at a million rows), with the same sums.
constexpr int kSumAhead = 8;
template <int N>
__device__ __forceinline__ void sum_partials(const double* partials, unsigned count,
                                             double (&total)[N]) {
  __shared__ double lds[N][kBlock / 64];
  double v[N];
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = 0;
  const unsigned stride = blockDim.x;
  unsigned b = threadIdx.x;
  for (; b + (kSumAhead - 1) * stride < count; b += kSumAhead * stride) {
    double p[kSumAhead][N];
#pragma unroll
    for (int u = 0; u < kSumAhead; ++u)
#pragma unroll
      for (int k = 0; k < N; ++k) p[u][k] = partials[(b + u * stride) * N + k];
#pragma unroll
    for (int u = 0; u < kSumAhead; ++u)
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] += p[u][k];
  }
  for (; b < count; b += stride)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += partials[b * N + k];
#pragma unroll
  for (int k = 0; k < N; ++k) total[k] = block_sum(v[k], lds[k]);
}

// --------------------------------------------------- prefetched row products
// Entries of one CSR row are processed in chunks of kChunk: all column ids and
// values of a chunk are loaded first, then all gathered x values, then the
// compensated sum runs over the registers in entry order. The arithmetic is
// the sequential loop's, but memory latency is paid per chunk instead of per
// entry.
constexpr int kChunk = 8;

// A matrix value stored as bfloat16 (the mixed path's coarse AMG levels);
// value_of widens either storage to the FP32 the row arithmetic uses.
struct bf16v {
  unsigned short bits;
};
__device__ __forceinline__ float value_of(float v) { return v; }
__device__ __forceinline__ float value_of(bf16v v) { return __uint_as_float(unsigned(v.bits) << 16); }

// x[j] for a gather; Prolonged adds the coarse correction of j's aggregate,
// the value the reference's in-place prolongation would have stored.
struct Plain {
  const float* __restrict__ x;
  __device__ __forceinline__ float operator()(int j) const { return x[j]; }
};
struct Prolonged {
  const float* __restrict__ x;
  const int* __restrict__ aggregate;
  const float* __restrict__ coarse;
  __device__ __forceinline__ float operator()(int j) const {
    const int g = aggregate[j];
    return g >= 0 ? x[j] + coarse[g] : x[j];
  }
};

// Kahan-compensated sum of a[j] * x[col[j]] over [begin, end), C entries per
// chunk (kernels/linalg.kern.h:5-16).
template <int C, class X, class V>
__device__ __forceinline__ float row_product_chunked(int begin, int end,
                                                     const int* __restrict__ cols,
                                                     const V* __restrict__ values, X gather) {
  Kahan k;
  for (int j = begin; j < end; j += C) {
    int c[C];
    float v[C], xv[C];
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) {
        c[u] = cols[j + u];
        v[u] = value_of(values[j + u]);
      }
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) xv[u] = gather(c[u]);
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) k.add(v[u] * xv[u]);
  }
  return k.s;
}

// Compensated residual b - sum a[j] * x[col[j]] over [begin, end), C entries
// per chunk (kernels/linalg.kern.h:32-45).
template <int C, class X, class V>
__device__ __forceinline__ float row_residual_chunked(float b, int begin, int end,
                                                      const int* __restrict__ cols,
                                                      const V* __restrict__ values,
                                                      X gather) {
  RowResidual acc(b);
  for (int j = begin; j < end; j += C) {
    int c[C];
    float v[C], xv[C];
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) {
        c[u] = cols[j + u];
        v[u] = value_of(values[j + u]);
      }
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) xv[u] = gather(c[u]);
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (j + u < end) acc.add(v[u], xv[u]);
  }
  return acc.value();
}

// Coarse AMG levels have 10-30 entries per row. When any row of the wave is
// longer than one chunk, the wave takes chunks of twice the width, so such a
// row costs one gather round trip instead of two. The choice is uniform per
// wave and the entry order is unchanged.
template <class X, class V>
__device__ __forceinline__ float row_product(int begin, int end, const int* __restrict__ cols,
                                             const V* __restrict__ values, X gather) {
  if (__any(end - begin > kChunk))
    return row_product_chunked<2 * kChunk>(begin, end, cols, values, gather);
  return row_product_chunked<kChunk>(begin, end, cols, values, gather);
}
template <class X, class V>
__device__ __forceinline__ float row_residual(float b, int begin, int end,
                                              const int* __restrict__ cols,
                                              const V* __restrict__ values, X gather) {
  if (__any(end - begin > kChunk))
    return row_residual_chunked<2 * kChunk>(b, begin, end, cols, values, gather);
  return row_residual_chunked<kChunk>(b, begin, end, cols, values, gather);
}
// The same with a third width, four chunks, for levels whose rows run past two
// chunks (the last levels where pairwise aggregation stalls have rows of up
// to 70 entries): a separate instantiation, so the registers the wide chunk
// needs do not lower the occupancy of the kernels that never take it.
template <class X, class V>
__device__ __forceinline__ float row_residual_wide(float b, int begin, int end,
                                                   const int* __restrict__ cols,
                                                   const V* __restrict__ values, X gather) {
  if (__any(end - begin > 2 * kChunk))
    return row_residual_chunked<4 * kChunk>(b, begin, end, cols, values, gather);
  return row_residual(b, begin, end, cols, values, gather);
}

// Maps workgroups so that each of the 8 XCDs (which receive workgroups
// round-robin) processes one contiguous range of the index space, keeping
// neighbour gathers inside one XCD's L2.
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

}  // namespace kh
