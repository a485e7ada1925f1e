// FP64 dense coarsest-level solve for the mixed-precision AMG path (no
// reference counterpart): the coarse operator, stored FP32 in CSR, is
// expanded into a dense FP64 matrix and inverted in place by blocked
// Gauss-Jordan elimination without pivoting (the operator is symmetric
// positive definite: a pinned M-matrix and its Galerkin products). Per block
// of kGjBlock pivots K:
//   P = A_KK^-1 (one workgroup, element-wise Gauss-Jordan in LDS), A_KK = P
//   A_KJ = P A_KJ                (row panel, J outside K)
//   A_IJ -= A_IK A_KJ            (trailing update, I and J outside K)
//   A_IK = -A_IK P               (column panel)
// The three products are 64 x 64 tiles on the FP64 MFMA
// (v_mfma_f64_16x16x4_f64). Each V-cycle then applies the inverse as an FP64
// matrix-vector product, one workgroup per row.
#include "kh_device_math.hpp"

namespace kh {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kGjBlock = 64;          // pivots per block, and the tile edge
constexpr int kGjStride = kGjBlock + 1;  // LDS row stride (doubles), padded

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

// P = A_KK^-1 for K = [k0, k0 + b): element-wise Gauss-Jordan on the block in
// LDS, pivot by pivot with the pivot's reciprocal r = 1/p (a_pp' = r,
// a_pj' = a_pj r, a_ip' = -a_ip r, a_ij' = a_ij - a_ip (a_pj r)); written to
// P (kGjBlock x kGjBlock) and A_KK. One 1024-thread workgroup, four entries
// per thread.
constexpr int kGjPivotThreads = 1024;
__global__ void __launch_bounds__(kGjPivotThreads) k_gj_pivot(int n, int k0, int b, double* __restrict__ a,
                                                              double* __restrict__ p) {
  __shared__ double t[kGjBlock * kGjStride];
  __shared__ double row[kGjBlock], col[kGjBlock];
  for (int e = threadIdx.x; e < b * b; e += kGjPivotThreads) {
    const int i = e / b, j = e % b;
    t[i * kGjStride + j] = a[long(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  for (int q = 0; q < b; ++q) {
    const double r = 1.0 / t[q * kGjStride + q];
    if (threadIdx.x < b) {
      row[threadIdx.x] = t[q * kGjStride + threadIdx.x] * r;  // the new pivot row (a_qj r)
      col[threadIdx.x] = t[threadIdx.x * kGjStride + q];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < b * b; e += kGjPivotThreads) {
      const int i = e / b, j = e % b;
      double v;
      if (i == q && j == q) v = r;
      else if (i == q) v = row[j];
      else if (j == q) v = -col[i] * r;
      else v = t[i * kGjStride + j] - col[i] * row[j];
      t[i * kGjStride + j] = v;
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kGjBlock * kGjBlock; e += kGjPivotThreads) {
    const int i = e / kGjBlock, j = e % kGjBlock;
    const double v = (i < b && j < b) ? t[i * kGjStride + j] : 0.0;
    p[e] = v;
    if (i < b && j < b) a[long(k0 + i) * n + k0 + j] = v;
  }
}

// The three panel and update products of one block step, 64 x 64 output
// tiles of C = L R with L, R staged in LDS (zero outside the matrix and the
// block). Wave w computes the 16 x 16 sub-blocks (w, 0..3); lane l of a
// v_mfma_f64_16x16x4_f64 supplies L[16 br + l%16][4 s + l/16] and
// R[4 s + l/16][16 bc + l%16] and its accumulator register q holds row
// 4 q + l/16, column l % 16 (the FP64 accumulator interleaves rows; checked
// against a CPU product on the device).
enum class GjTile { Rows, Update, Cols };
template <GjTile M>
__global__ void __launch_bounds__(kBlock) k_gj_tile(int n, int k0, int b, double* __restrict__ a,
                                                    const double* __restrict__ p) {
  __shared__ double L[kGjBlock * kGjStride];
  __shared__ double R[kGjBlock * kGjStride];
  // Tile origin: Rows tiles columns, Cols tiles rows, Update tiles both.
  int i0 = 0, j0 = 0;
  if constexpr (M == GjTile::Rows) j0 = int(blockIdx.x) * kGjBlock;
  if constexpr (M == GjTile::Cols) i0 = int(blockIdx.x) * kGjBlock;
  if constexpr (M == GjTile::Update) {
    i0 = int(blockIdx.y) * kGjBlock;
    j0 = int(blockIdx.x) * kGjBlock;
    if (i0 == k0 || j0 == k0) return;  // whole tile inside the block's rows or columns
  }
  if constexpr (M == GjTile::Rows) {
    if (j0 == k0) return;
  }
  if constexpr (M == GjTile::Cols) {
    if (i0 == k0) return;
  }
  for (int e = threadIdx.x; e < kGjBlock * kGjBlock; e += kBlock) {
    const int r = e / kGjBlock, c = e % kGjBlock;
    double lv = 0.0, rv = 0.0;
    if constexpr (M == GjTile::Rows) {
      lv = p[e];  // zero outside b x b
      if (r < b && j0 + c < n) rv = a[long(k0 + r) * n + j0 + c];
    } else if constexpr (M == GjTile::Update) {
      if (i0 + r < n && c < b) lv = a[long(i0 + r) * n + k0 + c];
      if (r < b && j0 + c < n) rv = a[long(k0 + r) * n + j0 + c];
    } else {
      if (i0 + r < n && c < b) lv = a[long(i0 + r) * n + k0 + c];
      rv = p[e];
    }
    L[r * kGjStride + c] = lv;
    R[r * kGjStride + c] = rv;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  f64x4 acc[4];
#pragma unroll
  for (int bc = 0; bc < 4; ++bc) acc[bc] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < kGjBlock / 4; ++s) {
    const double lv = L[(16 * wave + (lane & 15)) * kGjStride + 4 * s + (lane >> 4)];
#pragma unroll
    for (int bc = 0; bc < 4; ++bc) {
      const double rv = R[(4 * s + (lane >> 4)) * kGjStride + 16 * bc + (lane & 15)];
      acc[bc] = __builtin_amdgcn_mfma_f64_16x16x4f64(lv, rv, acc[bc], 0, 0, 0);
    }
  }
#pragma unroll
  for (int bc = 0; bc < 4; ++bc) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * wave + 4 * q + (lane >> 4), c = 16 * bc + (lane & 15);
      if constexpr (M == GjTile::Rows) {
        if (r < b && j0 + c < n) a[long(k0 + r) * n + j0 + c] = acc[bc][q];
      } else if constexpr (M == GjTile::Update) {
        if (i0 + r < n && j0 + c < n) a[long(i0 + r) * n + j0 + c] -= acc[bc][q];
      } else {
        if (i0 + r < n && c < b) a[long(i0 + r) * n + k0 + c] = -acc[bc][q];
      }
    }
  }
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

// rows, cols, values, inverse, scratch (n = rows of the dense operator;
// inverse n*n doubles, scratch at least kGjBlock^2): the dense inverse of the
// CSR operator, left in `inverse`.
void l_dense_invert(kh_device* d, int n, const kh_arg* a) {
  if (n > 8192) throw std::runtime_error("dense_invert: at most 8192 rows");
  if (n < 1) return;
  const long count = long(n) * n;
  double* m = KD(3);
  double* p = KD(4);
  k_dense_clear<<<dim3(unsigned((count + kBlock - 1) / kBlock)), dim3(kBlock), 0, d->stream>>>(count, m);
  k_dense_scatter<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, KB(0), KB(1), KCF(2), m);
  const unsigned tiles = unsigned((n + kGjBlock - 1) / kGjBlock);
  for (int k0 = 0; k0 < n; k0 += kGjBlock) {
    const int b = std::min(kGjBlock, n - k0);
    k_gj_pivot<<<dim3(1), dim3(kGjPivotThreads), 0, d->stream>>>(n, k0, b, m, p);
    k_gj_tile<GjTile::Rows><<<dim3(tiles), dim3(kBlock), 0, d->stream>>>(n, k0, b, m, p);
    k_gj_tile<GjTile::Update><<<dim3(tiles, tiles), dim3(kBlock), 0, d->stream>>>(n, k0, b, m, p);
    k_gj_tile<GjTile::Cols><<<dim3(tiles), dim3(kBlock), 0, d->stream>>>(n, k0, b, m, p);
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
