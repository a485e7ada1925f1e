// FP64 dense coarsest-level solve for the mixed-precision AMG path (no
// reference counterpart): the coarsest operator, stored FP32 in CSR, is
// expanded into a dense FP64 matrix and inverted in place by Gauss-Jordan
// elimination without pivoting (the operator is symmetric positive definite:
// a pinned M-matrix and its Galerkin products), one elimination step per
// launch over a ping-pong pair of buffers. Each V-cycle then applies the
// inverse as an FP64 matrix-vector product, one workgroup per row.
#include "kh_device_math.hpp"

namespace kh {
namespace {

__global__ void __launch_bounds__(kBlock) k_dense_clear(long count, double* __restrict__ dense) {
  const long i = long(blockIdx.x) * kBlock + threadIdx.x;
  if (i < count) dense[i] = 0.0;
}
__global__ void __launch_bounds__(kBlock) k_dense_scatter(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, double* __restrict__ dense) {
  const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
  if (i >= n) return;
  for (int j = rows[i]; j < rows[i + 1]; ++j) dense[long(i) * n + cols[j]] += double(values[j]);
}
// One Gauss-Jordan step on pivot k: src holds the matrix after k steps, dst
// receives it after k + 1 (after all n steps, the inverse).
__global__ void __launch_bounds__(kBlock) k_gauss_jordan(int n, int k, const double* __restrict__ src,
                                                         double* __restrict__ dst) {
  const long e = long(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= long(n) * n) return;
  const int i = int(e / n), j = int(e - long(i) * n);
  const double p = src[long(k) * n + k];
  double v;
  if (i == k && j == k) v = 1.0 / p;
  else if (i == k) v = src[long(k) * n + j] / p;
  else if (j == k) v = -src[long(i) * n + k] / p;
  else v = src[e] - src[long(i) * n + k] * src[long(k) * n + j] / p;
  dst[e] = v;
}
// x = inverse b in FP64, rounded to FP32; one workgroup per row.
__global__ void __launch_bounds__(kBlock) k_dense_apply(int n, const double* __restrict__ inverse,
    const float* __restrict__ b, float* __restrict__ x, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = int(blockIdx.x);
  double v = 0.0;
  for (int j = threadIdx.x; j < n; j += kBlock) v += inverse[long(i) * n + j] * double(b[j]);
  __shared__ double lds[kBlock / 64];
  v = block_sum(v, lds);
  if (threadIdx.x == 0) x[i] = float(v);
}

#define KB(k) static_cast<const int*>(buf(a, k))
#define KCF(k) static_cast<const float*>(buf(a, k))
#define KD(k) static_cast<double*>(buf(a, k))

// rows, cols, values, inverse, scratch (n = rows of the dense operator; both
// FP64 buffers n*n): the dense inverse of the CSR operator, left in `inverse`.
void l_dense_invert(kh_device* d, int n, const kh_arg* a) {
  if (n > 8192) throw std::runtime_error("dense_invert: at most 8192 rows");
  const long count = long(n) * n;
  const dim3 grid(unsigned((count + kBlock - 1) / kBlock));
  double* first = KD(3);
  double* second = KD(4);
  double* src = (n % 2 == 0) ? first : second;  // the last step writes `first`
  k_dense_clear<<<grid, dim3(kBlock), 0, d->stream>>>(count, src);
  k_dense_scatter<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, KB(0), KB(1), KCF(2), src);
  for (int k = 0; k < n; ++k) {
    double* dst = src == first ? second : first;
    k_gauss_jordan<<<grid, dim3(kBlock), 0, d->stream>>>(n, k, src, dst);
    src = dst;
  }
}
// inverse, b, x, status
void l_dense_apply(kh_device* d, int n, const kh_arg* a) {
  k_dense_apply<<<dim3(unsigned(n)), dim3(kBlock), 0, d->stream>>>(
      n, KD(0), KCF(1), static_cast<float*>(buf(a, 2)), KB(3));
}

}  // namespace

const std::vector<KernelEntry>& dense_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_DENSE_INVERT, "dense_invert", "bbbbb", l_dense_invert},
      {KH_DENSE_APPLY, "dense_apply", "bbbb", l_dense_apply},
  };
  return table;
}

}  // namespace kh
