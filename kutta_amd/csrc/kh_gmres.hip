// Flexible GMRES(m) building blocks (Saad, "A flexible inner-outer
// preconditioned GMRES algorithm", SIAM J. Sci. Comput. 14, 1993): an additive
// solver with no reference counterpart. Basis vectors are FP32 in HBM; the
// Hessenberg matrix, the Givens rotations and the residual recurrence are FP64
// in a small device block, so a cycle runs without host round trips:
//
//   gmres_start     beta = ||r||, g = beta e1, V0 = r / beta
//   gmres_arnoldi   classical Gram-Schmidt twice on w = A M^-1 v_j against
//                   V_0..V_j (one fused multi-dot per pass), the new column's
//                   Givens rotation, the residual estimate |g_{j+1}| against
//                   the target, V_{j+1} = w / h_{j+1,j}
//   gmres_update    back substitution H y = g, x += Z y
//
// Device block layout (doubles): H (m+1) x m column-major at 0, g (m+1),
// cs (m), sn (m), y (m), h (m+1) scratch, target, m. status[0] stops the
// cycle (1 converged or broke down), status[1] counts iterations, status[2]
// is the number of columns built.
#include "kh_device_math.hpp"

namespace kh {
namespace {

constexpr int kMaxBasis = 64;  // restart length limit (m + 1 vectors)

// Row of this thread, XCD-contiguous as in the linear-algebra kernels.
__device__ __forceinline__ int row_index() { return int(xcd_block() * kBlock + threadIdx.x); }

struct GmresLayout {
  int m;
  __device__ explicit GmresLayout(int restart) : m(restart) {}
  __device__ int h(int row, int col) const { return col * (m + 1) + row; }
  __device__ int g() const { return (m + 1) * m; }
  __device__ int cs() const { return g() + m + 1; }
  __device__ int sn() const { return cs() + m; }
  __device__ int y() const { return sn() + m; }
  __device__ int scratch() const { return y() + m; }
  __device__ int target() const { return scratch() + m + 1; }
};

// Partials of r.r over workgroups.
__global__ void __launch_bounds__(kBlock) k_gmres_norm_partials(int n, const float* __restrict__ r,
    double* __restrict__ partials) {
  const int i = row_index();
  const double v = i < n ? double(r[i]) * double(r[i]) : 0.0;
  __shared__ double lds[kBlock / 64];
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
// beta = sqrt(sum), g = beta e1, H cleared; a zero residual stops the cycle.
__global__ void __launch_bounds__(kBlock) k_gmres_start(const double* __restrict__ partials,
    unsigned count, double* __restrict__ block, int m, int* __restrict__ status) {
  double total[1];
  sum_partials<1>(partials, count, total);
  const GmresLayout L(m);
  for (int k = threadIdx.x; k < L.target(); k += blockDim.x) block[k] = 0.0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double beta = sqrt(total[0]);
  block[L.g()] = beta;
  block[L.scratch()] = beta;
  if (!(beta > 0.0) || !isfinite(beta)) status[0] = 1;
}
// V0 = r / beta (beta kept in the scratch slot by k_gmres_start).
__global__ void __launch_bounds__(kBlock) k_gmres_scale(int n, const float* __restrict__ w,
    float* __restrict__ v, const double* __restrict__ block, int slot, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i < n) v[i] = float(double(w[i]) / block[slot]);
}
// Partials of V_c . w for c = 0..k-1, one slot per vector per workgroup.
__global__ void __launch_bounds__(kBlock) k_gmres_dots(int n, const float* __restrict__ basis, int k,
    const float* __restrict__ w, double* __restrict__ partials, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  const double wi = i < n ? double(w[i]) : 0.0;
  __shared__ double lds[kMaxBasis][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < k; ++c) {
    const double v = i < n ? double(basis[std::size_t(c) * n + i]) * wi : 0.0;
    const double s = wave_sum(v);
    if (lane == 0) lds[c][wave] = s;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += lds[c][q];
    partials[std::size_t(blockIdx.x) * kMaxBasis + c] = t;
  }
}
// h[c] = sum of the partials of vector c, in workgroup order.
__global__ void __launch_bounds__(kBlock) k_gmres_dots_finish(const double* __restrict__ partials,
    unsigned count, int k, double* __restrict__ block, int m, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const GmresLayout L(m);
  __shared__ double lds[kBlock / 64];
  for (int c = 0; c < k; ++c) {
    double v = 0.0;
    for (unsigned b = threadIdx.x; b < count; b += blockDim.x) v += partials[std::size_t(b) * kMaxBasis + c];
    const double t = block_sum(v, lds);
    if (threadIdx.x == 0) block[L.scratch() + c] = t;
    __syncthreads();
  }
}
// w -= sum_c h[c] V_c; adds h into column j of H; with `norm`, partials of w.w.
__global__ void __launch_bounds__(kBlock) k_gmres_subtract(int n, const float* __restrict__ basis, int k,
    float* __restrict__ w, const double* __restrict__ block, int m, int norm, double* __restrict__ partials,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const GmresLayout L(m);
  __shared__ double h[kMaxBasis];
  for (int c = threadIdx.x; c < k; c += blockDim.x) h[c] = block[L.scratch() + c];
  __syncthreads();
  const int i = row_index();
  double wi = 0.0;
  if (i < n) {
    double projection = 0.0;
    for (int c = 0; c < k; ++c) projection += h[c] * double(basis[std::size_t(c) * n + i]);
    const float updated = float(double(w[i]) - projection);
    w[i] = updated;
    wi = double(updated);
  }
  if (norm) {
    __shared__ double lds[kBlock / 64];
    const double t = block_sum(wi * wi, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}
// Column j's first pass into H (the second pass adds to it in k_gmres_arnoldi).
__global__ void k_gmres_keep(double* __restrict__ block, int m, int j, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const GmresLayout L(m);
  for (int c = threadIdx.x; c <= j; c += blockDim.x) block[L.h(c, j)] = block[L.scratch() + c];
}
// h_{j+1,j} = ||w||, the second pass added to column j, the previous rotations
// applied to it and a new rotation zeroing h_{j+1,j}; the residual estimate
// |g_{j+1}| stops the cycle at the target, or on breakdown.
__global__ void __launch_bounds__(kBlock) k_gmres_arnoldi(const double* __restrict__ partials,
    unsigned count, double* __restrict__ block, int m, int j, int* __restrict__ status) {
  if (status[0] != 0) return;
  double total[1];
  sum_partials<1>(partials, count, total);
  if (threadIdx.x != 0) return;
  const GmresLayout L(m);
  double* H = block;
  for (int c = 0; c <= j; ++c) H[L.h(c, j)] += block[L.scratch() + c];
  const double norm = sqrt(total[0]);
  H[L.h(j + 1, j)] = norm;
  block[L.scratch()] = norm;  // for V_{j+1} = w / norm
  for (int c = 0; c < j; ++c) {
    const double cs = block[L.cs() + c], sn = block[L.sn() + c];
    const double a = H[L.h(c, j)], b = H[L.h(c + 1, j)];
    H[L.h(c, j)] = cs * a + sn * b;
    H[L.h(c + 1, j)] = -sn * a + cs * b;
  }
  const double a = H[L.h(j, j)], b = H[L.h(j + 1, j)];
  const double d = hypot(a, b);
  const double cs = d > 0.0 ? a / d : 1.0, sn = d > 0.0 ? b / d : 0.0;
  block[L.cs() + j] = cs;
  block[L.sn() + j] = sn;
  H[L.h(j, j)] = d;
  H[L.h(j + 1, j)] = 0.0;
  const double gj = block[L.g() + j];
  block[L.g() + j] = cs * gj;
  block[L.g() + j + 1] = -sn * gj;
  ++status[1];
  status[2] = j + 1;
  const double estimate = fabs(block[L.g() + j + 1]);
  if (!(norm > 0.0) || !isfinite(estimate) || estimate <= block[L.target()] || !(d > 0.0)) status[0] = 1;
}
// y = H^-1 g over the built columns (back substitution), then x += Z y.
__global__ void k_gmres_solve(double* __restrict__ block, int m, const int* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const GmresLayout L(m);
  const int k = status[2];
  for (int r = k - 1; r >= 0; --r) {
    double v = block[L.g() + r];
    for (int c = r + 1; c < k; ++c) v -= block[L.h(r, c)] * block[L.y() + c];
    const double d = block[L.h(r, r)];
    block[L.y() + r] = d != 0.0 ? v / d : 0.0;
  }
}
__global__ void __launch_bounds__(kBlock) k_gmres_correct(int n, const float* __restrict__ basis,
    float* __restrict__ x, const double* __restrict__ block, int m, const int* __restrict__ status) {
  const GmresLayout L(m);
  __shared__ double y[kMaxBasis];
  const int k = status[2];
  for (int c = threadIdx.x; c < k; c += blockDim.x) y[c] = block[L.y() + c];
  __syncthreads();
  const int i = row_index();
  if (i >= n) return;
  double v = 0.0;
  for (int c = 0; c < k; ++c) v += y[c] * double(basis[std::size_t(c) * n + i]);
  x[i] = float(double(x[i]) + v);
}

#define KB(k) static_cast<const int*>(buf(a, k))
#define KF(k) static_cast<float*>(buf(a, k))
#define KCF(k) static_cast<const float*>(buf(a, k))
#define KD(k) static_cast<double*>(buf(a, k))
#define GRID(n) dim3(blocks_for(n)), dim3(kBlock), 0, d->stream
#define ONE dim3(1), dim3(kBlock), 0, d->stream

void require_restart(int m) {
  if (m < 1 || m + 1 > kMaxBasis) throw std::runtime_error("gmres: restart length must be in [1, 63]");
}
// r, V0, block, m, status, partials
void l_gmres_start(kh_device* d, int n, const kh_arg* a) {
  const int m = i32(a, 3);
  require_restart(m);
  const unsigned g = blocks_for(std::size_t(n));
  int* st = static_cast<int*>(buf(a, 4));
  k_gmres_norm_partials<<<GRID(n)>>>(n, KCF(0), KD(5));
  k_gmres_start<<<ONE>>>(KD(5), g, KD(2), m, st);
  k_gmres_scale<<<GRID(n)>>>(n, KCF(0), KF(1), KD(2), (m + 1) * m + 4 * m + 1, st);
}
// basis, w, block, m, j, status, partials: CGS2 of w against V_0..V_j, the
// rotation of column j and V_{j+1}.
void l_gmres_arnoldi(kh_device* d, int n, const kh_arg* a) {
  const int m = i32(a, 3), j = i32(a, 4);
  require_restart(m);
  if (j < 0 || j >= m) throw std::runtime_error("gmres_arnoldi: column out of range");
  const unsigned g = blocks_for(std::size_t(n));
  const float* basis = KCF(0);
  float* w = KF(1);
  double* block = KD(2);
  int* st = static_cast<int*>(buf(a, 5));
  double* partials = KD(6);
  const int k = j + 1;
  k_gmres_dots<<<GRID(n)>>>(n, basis, k, w, partials, st);
  k_gmres_dots_finish<<<ONE>>>(partials, g, k, block, m, st);
  k_gmres_subtract<<<GRID(n)>>>(n, basis, k, w, block, m, 0, partials, st);
  k_gmres_keep<<<ONE>>>(block, m, j, st);
  k_gmres_dots<<<GRID(n)>>>(n, basis, k, w, partials, st);
  k_gmres_dots_finish<<<ONE>>>(partials, g, k, block, m, st);
  k_gmres_subtract<<<GRID(n)>>>(n, basis, k, w, block, m, 1, partials, st);
  k_gmres_arnoldi<<<ONE>>>(partials, g, block, m, j, st);
  const int scratch = (m + 1) * m + 4 * m + 1;
  k_gmres_scale<<<GRID(n)>>>(n, w, const_cast<float*>(basis) + std::size_t(k) * n, block, scratch, st);
}
// basis Z, x, block, m, status: y = H^-1 g, x += Z y (runs after a stop too).
void l_gmres_update(kh_device* d, int n, const kh_arg* a) {
  const int m = i32(a, 3);
  require_restart(m);
  const int* st = KB(4);
  k_gmres_solve<<<dim3(1), dim3(64), 0, d->stream>>>(KD(2), m, st);
  k_gmres_correct<<<GRID(n)>>>(n, KCF(0), KF(1), KD(2), m, st);
}

}  // namespace

const std::vector<KernelEntry>& gmres_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_GMRES_START, "gmres_start", "bbbibb", l_gmres_start},
      {KH_GMRES_ARNOLDI, "gmres_arnoldi", "bbbiibb", l_gmres_arnoldi},
      {KH_GMRES_UPDATE, "gmres_update", "bbbib", l_gmres_update},
  };
  return table;
}

}  // namespace kh
