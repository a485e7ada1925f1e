// Finite-volume kernels of the SIMPLEC pressure path for gfx950. One lane
// owns one cell and gathers over its faces; each lane writes only its own CSR
// row, so no atomics are needed and results are deterministic.
#include "kh_device_math.hpp"

namespace kh {
namespace {

__device__ __forceinline__ int cell_index() { return int(xcd_block() * kBlock + threadIdx.x); }

// kernels/fv_flow.kern.h:164-183 — pressure-correction M-matrix row: clear the
// row's neighbour entries, subtract each face coefficient from its (possibly
// shared) entry, eliminate the reference cell's row and column, and build the
// diagonal as the upward-rounded sum of the stored off-diagonals plus boundary
// anchors; b = -sum sign * flux.
__global__ void __launch_bounds__(kBlock) k_fv_pressure_assemble(int n,
    const int* __restrict__ offsets, const int* __restrict__ faces,
    const float* __restrict__ sign, const int* __restrict__ adjacent,
    const int* __restrict__ entry, const int* __restrict__ rows, const int* __restrict__ diag,
    const float* __restrict__ coefficient, const float* __restrict__ flux,
    float* __restrict__ values, float* __restrict__ rhs, int reference) {
  const int i = cell_index();
  if (i >= n) return;
  const int begin = offsets[i], end = offsets[i + 1];
  for (int j = begin; j < end; ++j)
    if (adjacent[j] >= 0) values[entry[j]] = 0.0f;
  UpperSum ap{0.0f, 0.0f};
  float b = 0.0f;
  for (int j = begin; j < end; ++j) {
    const int f = faces[j], nb = adjacent[j];
    const float d = coefficient[f];
    b -= sign[j] * flux[f];
    if (nb < 0 || nb == reference) ap = upper_add(ap, d);
    if (nb >= 0 && nb != reference && i != reference) values[entry[j]] -= d;
  }
  const int di = diag[i];
  for (int j = rows[i]; j < rows[i + 1]; ++j)
    if (j != di) ap = upper_add(ap, -values[j]);
  values[di] = i == reference ? 1.0f : upper_value(ap);
  rhs[i] = i == reference ? 0.0f : b;
}

#define KB(k) static_cast<const int*>(buf(a, k))
#define KF(k) static_cast<float*>(buf(a, k))
#define KCF(k) static_cast<const float*>(buf(a, k))
#define GRID(n) dim3(blocks_for(n)), dim3(kBlock), 0, d->stream

void l_fv_pressure_assemble(kh_device* d, int n, const kh_arg* a) {
  k_fv_pressure_assemble<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KB(3), KB(4), KB(5), KB(6), KCF(7),
                                      KCF(8), KF(9), KF(10), i32(a, 11));
}

}  // namespace

const std::vector<KernelEntry>& fv_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_FV_PRESSURE_ASSEMBLE, "fv_pressure_assemble", "bbbbbbbbbbbi", l_fv_pressure_assemble},
  };
  return table;
}

}  // namespace kh
