// Linear-algebra kernels of the pressure solve for gfx950: CSR products and
// compensated residuals, l1-Jacobi smoothing, aggregation transfers, Galerkin
// refresh, Krylov vector updates and the device-scalar state machine, plus
// fused MI355X variants that remove launches and HBM round trips.
//
// Per-element arithmetic follows the reference kernel bodies operation for
// operation (cited per kernel), so results are bitwise equal to the serial
// reference wherever no FP64 reduction order is involved.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "kh_device_math.hpp"

namespace kh {
namespace {

// Global index of this thread for an n-index launch, XCD-contiguous.
__device__ __forceinline__ int row_index() { return int(xcd_block() * kBlock + threadIdx.x); }

// ------------------------------------------------------------ CSR products
// kernels/linalg.kern.h:5-16
__device__ __forceinline__ float spmv_row(int i, const int* __restrict__ rows,
                                          const int* __restrict__ cols,
                                          const float* __restrict__ values,
                                          const float* __restrict__ x) {
  return row_product(rows[i], rows[i + 1], cols, values, Plain{x});
}
// kernels/linalg.kern.h:32-45
template <class V>
__device__ __forceinline__ float residual_row(int i, const int* __restrict__ rows,
                                              const int* __restrict__ cols,
                                              const V* __restrict__ values,
                                              const float* __restrict__ x,
                                              const float* __restrict__ b) {
  return row_residual(b[i], rows[i], rows[i + 1], cols, values, Plain{x});
}

__global__ void __launch_bounds__(kBlock) k_csr_spmv(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values,
    const float* __restrict__ x, float* __restrict__ y, const int* __restrict__ status) {
  if (status && status[0] != 0) return;
  const int i = row_index();
  if (i < n) y[i] = spmv_row(i, rows, cols, values, x);
}
__global__ void __launch_bounds__(kBlock) k_csr_residual(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values,
    const float* __restrict__ x, const float* __restrict__ b, float* __restrict__ r,
    const int* __restrict__ status) {
  if (status && status[0] != 0) return;
  const int i = row_index();
  if (i < n) r[i] = residual_row(i, rows, cols, values, x, b);
}
// kernels/linalg.kern.h:52-70
__global__ void __launch_bounds__(kBlock) k_krylov_diagonal(int n, const int* __restrict__ rows,
    const int* __restrict__ diag, const float* __restrict__ values, float* __restrict__ inverse,
    float* __restrict__ invalid, int use_l1) {
  const int i = row_index();
  if (i >= n) return;
  const int di = diag[i];
  float d = values[di];
  float off = 0.0f, bad = 0.0f;
  for (int j = rows[i]; j < rows[i + 1]; ++j) {
    float v = values[j];
    if (v - v != 0.0f) bad = 1.0f;
    if (j != di) off += fabsf(v);
  }
  float inv = d != 0.0f ? 1.0f / d : 0.0f;
  float l1 = kmax(fabsf(d), 0.5f * (fabsf(d) + off));
  inverse[i] = use_l1 != 0 ? (l1 != 0.0f ? 1.0f / l1 : 0.0f) : inv;
  invalid[i] = (bad != 0.0f || inv - inv != 0.0f || l1 - l1 != 0.0f)
                   ? 2.0f
                   : (use_l1 != 0 && d < 0.0f ? 3.0f : (d == 0.0f ? 1.0f : 0.0f));
}
// kernels/linalg.kern.h:72-107 — elementwise Krylov updates.
__global__ void k_krylov_multiply(int n, const float* __restrict__ r,
    const float* __restrict__ inverse, float* __restrict__ z, const int* __restrict__ status) {
  if (status && status[0] != 0) return;
  const int i = row_index();
  if (i < n) z[i] = r[i] * inverse[i];
}
__global__ void k_krylov_smooth(int n, const float* __restrict__ r, const float* __restrict__ inv,
                                float* __restrict__ x, float omega) {
  const int i = row_index();
  if (i < n) x[i] += omega * inv[i] * r[i];
}
__global__ void k_krylov_shift(int n, float* __restrict__ x, float mean) {
  const int i = row_index();
  if (i < n) x[i] -= mean;
}
__global__ void k_krylov_direction(int n, const float* __restrict__ z, float* __restrict__ p,
                                   float beta) {
  const int i = row_index();
  if (i < n) p[i] = z[i] + beta * p[i];
}
__global__ void k_krylov_step(int n, float* __restrict__ x, float* __restrict__ r,
    const float* __restrict__ p, const float* __restrict__ ap, float alpha) {
  const int i = row_index();
  if (i < n) {
    x[i] += alpha * p[i];
    r[i] -= alpha * ap[i];
  }
}
__global__ void k_krylov_bicg_direction(int n, const float* __restrict__ r,
    const float* __restrict__ v, float* __restrict__ p, float beta, float omega) {
  const int i = row_index();
  if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
__global__ void k_krylov_difference(int n, const float* __restrict__ a, const float* __restrict__ b,
                                    float* __restrict__ out, float scale) {
  const int i = row_index();
  if (i < n) out[i] = a[i] - scale * b[i];
}
__global__ void k_krylov_bicg_step(int n, float* __restrict__ x, const float* __restrict__ p,
    const float* __restrict__ s, float alpha, float omega) {
  const int i = row_index();
  if (i < n) x[i] += alpha * p[i] + omega * s[i];
}

// ------------------------------------------------------------- AMG kernels
// kernels/linalg.kern.h:126-147. The reference runs all sweeps in one thread;
// here one lane owns each row and the lanes of one wave exchange the sweep's
// iterate through LDS. Every row's arithmetic is the reference's, and all
// rows read the previous sweep's iterate, so the result is identical.
constexpr int kCoarseMax = 24;
__global__ void __launch_bounds__(64) k_amg_coarse_solve(const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, float* __restrict__ x,
    const float* __restrict__ b, const float* __restrict__ inverse,
    const int* __restrict__ status, int count, int sweeps) {
  if (status[0] != 0) return;
  __shared__ float cur[kCoarseMax];
  const int row = threadIdx.x;
  if (row < count) cur[row] = 0.0f;
  __syncthreads();
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    float r = 0.0f;
    if (row < count) {
      RowResidual acc(b[row]);
      for (int j = rows[row]; j < rows[row + 1]; ++j) acc.add(values[j], cur[cols[j]]);
      r = acc.value();
    }
    __syncthreads();
    if (row < count) cur[row] += (2.0f / 3.0f) * inverse[row] * r;
    __syncthreads();
  }
  if (row < count) x[row] = cur[row];
}
// kernels/linalg.kern.h:150-156. Wide: the instantiation for long rows
// (row_residual_wide), same arithmetic.
template <class V, bool Wide = false>
__global__ void __launch_bounds__(kBlock) k_amg_smooth(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const V* __restrict__ values, const float* __restrict__ x,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ next,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float r = Wide ? row_residual_wide(b[i], rows[i], rows[i + 1], cols, values, Plain{x})
                       : residual_row(i, rows, cols, values, x, b);
  next[i] = x[i] + (2.0f / 3.0f) * inverse[i] * r;
}
// The same sweep with one wavefront per row, for levels whose rows are long
// (the last levels where pairwise aggregation stalls have rows of up to ~100
// entries): the 64 lanes load the row's entries and gathered x values and
// form each product and its error (independent of the running sum) in
// parallel, up to kRowWaveEntries per row at once; lane 0 then runs the
// compensated sum over them in entry order, the arithmetic of RowResidual.
constexpr int kRowWaveSlots = 4;                       // entries per lane per pass
constexpr int kRowWaveEntries = 64 * kRowWaveSlots;    // entries per pass
// What a wavefront-per-row launch writes for row i: the sweep's new value
// gather(i) + (2/3) l1_i r_i (amg_smooth, or the prolongation folded into the
// sweep when the gather adds the coarse correction), or the residual r_i
// alone (csr_residual).
enum class RowWaveOut { Smooth, Residual };
template <RowWaveOut Out, class X>
__global__ void __launch_bounds__(kBlock) k_rowwave(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, X gather,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ out,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  __shared__ float term[kBlock / 64][kRowWaveEntries];
  __shared__ float error[kBlock / 64][kRowWaveEntries];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = int(xcd_block()) * (kBlock / 64) + wave;
  if (i >= n) return;  // uniform per wavefront; no block-wide barrier follows
  const int begin = rows[i], end = rows[i + 1];
  float s = b[i], e = 0.0f;
  for (int base = begin; base < end; base += kRowWaveEntries) {
    float a[kRowWaveSlots], xv[kRowWaveSlots];
    int c[kRowWaveSlots];
#pragma unroll
    for (int k = 0; k < kRowWaveSlots; ++k) {
      const int j = base + k * 64 + lane;
      if (j < end) {
        c[k] = cols[j];
        a[k] = values[j];
      }
    }
#pragma unroll
    for (int k = 0; k < kRowWaveSlots; ++k)
      if (base + k * 64 + lane < end) xv[k] = gather(c[k]);
#pragma unroll
    for (int k = 0; k < kRowWaveSlots; ++k) {
      if (base + k * 64 + lane < end) {
        const float prod = a[k] * xv[k];
        term[wave][k * 64 + lane] = -prod;
        error[wave][k * 64 + lane] = product_error(a[k], xv[k], prod);
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane == 0) {
      const int count = min(end - base, kRowWaveEntries);
      for (int k = 0; k < count; ++k) {
        const float t = term[wave][k];
        const float nx = s + t;
        e += fabsf(s) >= fabsf(t) ? (s - nx) + t : (t - nx) + s;
        e -= error[wave][k];
        s = nx;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (lane == 0) {
    if constexpr (Out == RowWaveOut::Smooth) out[i] = gather(i) + (2.0f / 3.0f) * inverse[i] * (s + e);
    else out[i] = s + e;
  }
}

// kernels/linalg.kern.h:162-170
__global__ void __launch_bounds__(kBlock) k_amg_smooth_zero(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ x,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ next,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float bi = b[i];
  if (bi - bi != 0.0f) {
    const float r = residual_row(i, rows, cols, values, x, b);
    next[i] = x[i] + (2.0f / 3.0f) * inverse[i] * r;
    return;
  }
  next[i] = 0.0f + (2.0f / 3.0f) * inverse[i] * (bi + 0.0f);
}
// kernels/amg.kern.h:5-16
__global__ void k_amg_restrict(int n, const int* __restrict__ offsets,
    const int* __restrict__ members, const float* __restrict__ fine, float* __restrict__ coarse,
    const int* __restrict__ status) {
  if (status && status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  Kahan k;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) k.add(fine[members[j]]);
  coarse[i] = k.s;
}
// kernels/amg.kern.h:18-21
__global__ void k_amg_prolong(int n, const int* __restrict__ aggregate,
    const float* __restrict__ coarse, float* __restrict__ fine, const int* __restrict__ status) {
  if (status && status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const int g = aggregate[i];
  if (g >= 0) fine[i] += coarse[g];
}
// kernels/amg.kern.h:27-38
__global__ void k_amg_galerkin(int n, const int* __restrict__ offsets,
    const int* __restrict__ entries, const float* __restrict__ fine, float* __restrict__ coarse) {
  const int i = row_index();
  if (i >= n) return;
  Neumaier acc;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) acc.add(fine[entries[j]]);
  coarse[i] = acc.value();
}
// kernels/amg.kern.h:44-58
__global__ void k_amg_m_matrix_diagonal(int n, const int* __restrict__ offsets,
    const int* __restrict__ members, const int* __restrict__ fine_rows,
    const int* __restrict__ fine_diag, const float* __restrict__ fine,
    const int* __restrict__ rows, const int* __restrict__ diag, float* coarse) {
  const int i = row_index();
  if (i >= n) return;
  UpperSum d{0.0f, 0.0f};
  for (int m = offsets[i]; m < offsets[i + 1]; ++m) {
    const int row = members[m];
    const int fd = fine_diag[row];
    UpperSum surplus{fine[fd], 0.0f};
    for (int j = fine_rows[row]; j < fine_rows[row + 1]; ++j)
      if (j != fd) surplus = upper_add(surplus, fine[j]);
    d = upper_add(d, upper_value(surplus));
  }
  const int di = diag[i];
  for (int j = rows[i]; j < rows[i + 1]; ++j)
    if (j != di) d = upper_add(d, -coarse[j]);
  coarse[di] = upper_value(d);
}

// ------------------------------------------------------------------- BLAS-1
// kernels/blas.kern.h:5-19
__global__ void k_axpy(int n, float* __restrict__ y, const float* __restrict__ x, float a) {
  const int i = row_index();
  if (i < n) y[i] = y[i] + a * x[i];
}
__global__ void k_xpby(int n, float* __restrict__ y, const float* __restrict__ x, float b) {
  const int i = row_index();
  if (i < n) y[i] = x[i] + b * y[i];
}
__global__ void k_axpbyz(int n, float* __restrict__ z, const float* __restrict__ x,
                         const float* __restrict__ y, float a, float b) {
  const int i = row_index();
  if (i < n) z[i] = a * x[i] + b * y[i];
}
__global__ void k_scale_copy(int n, float* __restrict__ y, const float* __restrict__ x, float a) {
  const int i = row_index();
  if (i < n) y[i] = a * x[i];
}
__global__ void k_mul(int n, float* __restrict__ z, const float* __restrict__ x,
                      const float* __restrict__ y) {
  const int i = row_index();
  if (i < n) z[i] = x[i] * y[i];
}

// ------------------------------------------------------- scalar kernels
// kernels/scalars.kern.h:51-55
__global__ void k_solver_limits(int n, float* s, float th, float tl, float dh, float dl) {
  const int i = row_index();
  if (i < n) s[i] = i == 20 ? th : (i == 21 ? tl : (i == 22 ? dh : (i == 23 ? dl : 0.0f)));
}
// kernels/scalars.kern.h:56-62
__global__ void k_solver_reset(float* s, int* status, int iterations) {
  if (threadIdx.x != 0) return;
  status[0] = 0; status[1] = iterations; status[2] = 0;
  s[8] = 1.0f; s[9] = 0.0f; s[12] = 1.0f; s[13] = 0.0f; s[14] = 1.0f; s[15] = 0.0f;
}
__global__ void k_solver_pcg_alpha(float* s, int* status) {
  if (threadIdx.x == 0) pcg_alpha(s, status);
}
__global__ void k_solver_pcg_beta(float* s, int* status, int flexible) {
  if (threadIdx.x == 0) pcg_beta(s, status, flexible);
}
// kernels/scalars.kern.h:81-89
__device__ void bicg_beta(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rho = pload(s, 0), omega = pload(s, 12);
  Pair beta = pmul(pdiv(rho, pload(s, 14)), pdiv(pload(s, 8), omega));
  if (pzero(rho) || !pfinite(rho) || pzero(omega) || !pcoefficient(omega) ||
      !pcoefficient(beta) || pzero(beta))
    status[0] = 2;
  else {
    pstore(s, 10, beta);
    status[2] = 0;
  }
}
// kernels/scalars.kern.h:90-97
__device__ void bicg_alpha(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rv = pload(s, 2);
  Pair alpha = pdiv(pload(s, 0), rv);
  if (pzero(rv) || !pfinite(rv) || !pcoefficient(alpha) || pzero(alpha)) status[0] = 2;
  else pstore(s, 8, alpha);
}
// kernels/scalars.kern.h:98-103
__device__ void bicg_s(const float* s, int* status) {
  if (status[0] != 0) return;
  ++status[1];
  status[2] = ple(pload(s, 16), pload(s, 20)) ? 1 : 0;
}
// kernels/scalars.kern.h:104-112
__device__ void bicg_omega(float* s, int* status) {
  if (status[0] != 0) return;
  if (status[2] != 0) {
    pstore(s, 12, {0.0f, 0.0f});
    return;
  }
  Pair tt = pload(s, 2), ts = pload(s, 4);
  Pair omega = pdiv(ts, tt);
  if (!ppositive(tt) || !pfinite(tt) || pzero(omega) || !pcoefficient(omega)) status[0] = 2;
  else pstore(s, 12, omega);
}
__global__ void k_solver_bicg_beta(float* s, int* status) {
  if (threadIdx.x == 0) bicg_beta(s, status);
}
__global__ void k_solver_bicg_alpha(float* s, int* status) {
  if (threadIdx.x == 0) bicg_alpha(s, status);
}
__global__ void k_solver_bicg_s(const float* s, int* status) {
  if (threadIdx.x == 0) bicg_s(s, status);
}
__global__ void k_solver_bicg_omega(float* s, int* status) {
  if (threadIdx.x == 0) bicg_omega(s, status);
}
__global__ void k_solver_check(float* s, int* status, int pcg) {
  if (threadIdx.x == 0) solver_check_body(s, status, pcg);
}
// kernels/scalars.kern.h:122-127
__global__ void k_solver_step(int n, float* __restrict__ x, float* __restrict__ r,
    const float* __restrict__ p, const float* __restrict__ q, const float* __restrict__ s,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float alpha = pfloat(pload(s, 8));
  x[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}
// kernels/scalars.kern.h:128-131
__global__ void k_solver_direction(int n, const float* __restrict__ z, float* __restrict__ p,
    const float* __restrict__ s, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i < n) p[i] = z[i] + pfloat(pload(s, 10)) * p[i];
}
// kernels/scalars.kern.h:157-186: the optional K-cycle's inner corrections.
// Slots of a level's 12 scalars: 0 the step's denominator (p.Ap for SPD,
// |Ap|^2 otherwise), 2 its numerator, 4 t.t then the conjugation product,
// 6 b.b, 8 alpha, 10 beta. `first` makes the level's status follow the
// parent's; a failed first step is phase 2 (keep the first correction
// unscaled), a failed second one phase 1 (keep the first step).
__device__ __forceinline__ void inner_alpha_body(float* s, int* status, const int* parent, int first) {
  if (first != 0) status[0] = parent[0] == 0 ? 0 : 2;
  if (status[0] != 0) return;
  const Pair denom = pload(s, 0), alpha = pdiv(pload(s, 2), denom);
  if (!ppositive(denom) || !pcoefficient(alpha) || (pzero(alpha) && !pzero(pload(s, 2))))
    status[0] = first != 0 ? 2 : 1;
  else pstore(s, 8, alpha);
}
__device__ __forceinline__ void inner_check_body(const float* s, int* status) {
  if (status[0] != 0) return;
  if (ple(pload(s, 4), pmul({0.0625f, 0.0f}, pload(s, 6)))) status[0] = 1;
}
__device__ __forceinline__ void inner_beta_body(float* s, int* status) {
  if (status[0] != 0) return;
  const Pair beta = pdiv(pneg(pload(s, 4)), pload(s, 0));
  if (!pcoefficient(beta) || (pzero(beta) && !pzero(pload(s, 4)))) status[0] = 1;
  else pstore(s, 10, beta);
}
__global__ void k_amg_inner_alpha(float* s, int* status, const int* parent, int first) {
  if (threadIdx.x == 0) inner_alpha_body(s, status, parent, first);
}
// Unguarded: x = alpha p and t = b - alpha q while running, else x = p, t = 0.
__global__ void k_amg_inner_first(int n, const float* __restrict__ b, const float* __restrict__ p,
    const float* __restrict__ q, float* __restrict__ x, float* __restrict__ t,
    const float* __restrict__ s, const int* __restrict__ status) {
  const int i = row_index();
  if (i >= n) return;
  const bool running = status[0] == 0;
  const float alpha = pfloat(pload(s, 8));
  x[i] = running ? alpha * p[i] : p[i];
  t[i] = running ? b[i] - alpha * q[i] : 0.0f;
}
// Adaptive exit: the second correction is skipped once t.t <= b.b / 16.
__global__ void k_amg_inner_check(const float* s, int* status) {
  if (threadIdx.x == 0) inner_check_body(s, status);
}
__global__ void k_amg_inner_beta(float* s, int* status) {
  if (threadIdx.x == 0) inner_beta_body(s, status);
}
__global__ void k_amg_inner_step(int n, float* __restrict__ x, const float* __restrict__ p,
    const float* __restrict__ s, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i < n) x[i] += pfloat(pload(s, 8)) * p[i];
}

// kernels/scalars.kern.h:132-137
__global__ void k_solver_bicg_direction(int n, const float* __restrict__ r,
    const float* __restrict__ v, float* __restrict__ p, const float* __restrict__ s,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float beta = pfloat(pload(s, 10)), omega = pfloat(pload(s, 12));
  p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
// kernels/scalars.kern.h:138-141
__global__ void k_solver_bicg_difference(int n, const float* __restrict__ r,
    const float* __restrict__ q, float* __restrict__ sv, const float* __restrict__ s,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i < n) sv[i] = r[i] - pfloat(pload(s, 8)) * q[i];
}
// kernels/scalars.kern.h:142-149
__global__ void k_solver_bicg_step(int n, float* __restrict__ x, float* __restrict__ r,
    const float* __restrict__ z, const float* __restrict__ zs, const float* __restrict__ sv,
    const float* __restrict__ t, const float* __restrict__ s, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float alpha = pfloat(pload(s, 8)), omega = pfloat(pload(s, 12));
  if (status[2] != 0) {
    x[i] += alpha * z[i];
    r[i] = sv[i];
  } else {
    x[i] += alpha * z[i] + omega * zs[i];
    r[i] = sv[i] - omega * t[i];
  }
}
// kernels/scalars.kern.h:150-152
__global__ void k_solver_project(int n, float* __restrict__ x, const float* __restrict__ sum,
                                 float count_hi, float count_lo) {
  const int i = row_index();
  if (i < n) x[i] -= pfloat(pdiv(pload(sum, 0), {count_hi, count_lo}));
}

// ------------------------------------------------------------ fused kernels
// A fused reduction is a row kernel that stores one FP64 partial per
// workgroup (store_partials) and a one-workgroup finisher that sums the slots
// in order and applies the scalar update of the reference sequence it replaces.

// q = A p, partials of p.q: csr_spmv_guarded + dot_into(p,q,2).
__global__ void __launch_bounds__(kBlock) k_fused_spmv_dot(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ p,
    float* __restrict__ q, const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float qi = spmv_row(i, rows, cols, values, p);
    q[i] = qi;
    mine[0] = double(p[i]) * double(qi);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// x += alpha p, r -= alpha q, partials of r.r: solver_step + dot_into(r,r,18).
__global__ void __launch_bounds__(kBlock) k_fused_step_dot(int n, float* __restrict__ x,
    float* __restrict__ r, const float* __restrict__ p, const float* __restrict__ q,
    const float* __restrict__ scalars, const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  const float alpha = pfloat(pload(scalars, 8));
  double mine[1] = {0.0};
  if (i < n) {
    x[i] += alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    mine[0] = double(ri) * double(ri);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// Final smoothing sweep of the preconditioner, partials of r.z:
// amg_smooth + dot_into(r,z,4). Here b is the Krylov residual r.
__global__ void __launch_bounds__(kBlock) k_fused_smooth_dot(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ x,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ next,
    const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float bi = b[i];
    const float r = row_residual(bi, rows[i], rows[i + 1], cols, values, Plain{x});
    const float z = x[i] + (2.0f / 3.0f) * inverse[i] * r;
    next[i] = z;
    mine[0] = double(bi) * double(z);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// Smoothing sweep whose output is summed for the constant-nullspace
// projection that follows: amg_smooth + sum_into(next).
__global__ void __launch_bounds__(kBlock) k_fused_smooth_partial_sum(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const float* __restrict__ values,
    const float* __restrict__ x, const float* __restrict__ b, const float* __restrict__ inverse,
    float* __restrict__ next, const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float r = residual_row(i, rows, cols, values, x, b);
    const float z = x[i] + (2.0f / 3.0f) * inverse[i] * r;
    next[i] = z;
    mine[0] = double(z);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// Mean removal, partials of r.z: solver_project + dot_into(r,z,4).
__global__ void __launch_bounds__(kBlock) k_fused_project_dot(int n, float* __restrict__ z,
    const float* __restrict__ sum, float count_hi, float count_lo, const float* __restrict__ r,
    const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float zi = z[i] - pfloat(pdiv(pload(sum, 0), {count_hi, count_lo}));
    z[i] = zi;
    mine[0] = double(r[i]) * double(zi);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}

// BiCGStab with the Jacobi preconditioner (linalg/solver.cpp:314-361). Each
// vector kernel of an iteration also applies the diagonal inverse to its
// output or produces the partials of the reduction that follows it, so an
// iteration is direction, shadow-dot SpMV, difference, two-dot SpMV and
// step, each reduction with its one-workgroup finisher.

// Partials of x.y: dot_into(x, y) ahead of a scalar update.
__device__ __forceinline__ void fused_dot_body(int n, const float* __restrict__ x,
    const float* __restrict__ y, const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {i < n ? double(x[i]) * double(y[i]) : 0.0};
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// p = r + beta (p - omega q), z = p * inverse: solver_bicg_direction +
// krylov_multiply_guarded.
__device__ __forceinline__ void fused_bicg_direction_body(int n, const float* __restrict__ r,
    const float* __restrict__ q, float* __restrict__ p, const float* __restrict__ inverse,
    float* __restrict__ z, const float* __restrict__ scalars, const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float beta = pfloat(pload(scalars, 10)), omega = pfloat(pload(scalars, 12));
  const float pi = r[i] + beta * (p[i] - omega * q[i]);
  p[i] = pi;
  z[i] = pi * inverse[i];
}
// q = A z, partials of shadow.q: csr_spmv_guarded + dot_into(shadow, q, 2).
__device__ __forceinline__ void fused_spmv_shadow_dot_body(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ z,
    float* __restrict__ q, const float* __restrict__ shadow, const int* __restrict__ status,
    double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float qi = spmv_row(i, rows, cols, values, z);
    q[i] = qi;
    mine[0] = double(shadow[i]) * double(qi);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// s = r - alpha q, zs = s * inverse, partials of s.s: solver_bicg_difference +
// dot_into(s, s, 16) + krylov_multiply_guarded.
__device__ __forceinline__ void fused_bicg_difference_body(int n, const float* __restrict__ r,
    const float* __restrict__ q, float* __restrict__ sv, const float* __restrict__ inverse,
    float* __restrict__ zs, const float* __restrict__ scalars, const int* __restrict__ status,
    double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    const float si = r[i] - pfloat(pload(scalars, 8)) * q[i];
    sv[i] = si;
    zs[i] = si * inverse[i];
    mine[0] = double(si) * double(si);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
// t = A zs, partials of t.t and t.s: csr_spmv_guarded + dot2_into(t,t,t,s,2).
__device__ __forceinline__ void fused_spmv_dot2_body(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ zs,
    float* __restrict__ t, const float* __restrict__ sv, const int* __restrict__ status,
    double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[2] = {0.0, 0.0};
  if (i < n) {
    const float ti = spmv_row(i, rows, cols, values, zs);
    t[i] = ti;
    mine[0] = double(ti) * double(ti);
    mine[1] = double(ti) * double(sv[i]);
  }
  __shared__ double lds[2][kBlock / 64];
  mine[0] = block_sum(mine[0], lds[0]);
  mine[1] = block_sum(mine[1], lds[1]);
  store_partials<2>(mine, partials);
}
// x += alpha z + omega zs, r = s - omega t (or the s exit), partials of r.r
// and shadow.r: solver_bicg_step + dot_into(r, r, 18) + the next
// iteration's dot_into(shadow, r, 0).
__device__ __forceinline__ void fused_bicg_step_body(int n, float* __restrict__ x,
    float* __restrict__ r, const float* __restrict__ z, const float* __restrict__ zs,
    const float* __restrict__ sv, const float* __restrict__ t, const float* __restrict__ shadow,
    const float* __restrict__ scalars, const int* __restrict__ status, double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[2] = {0.0, 0.0};
  if (i < n) {
    const float alpha = pfloat(pload(scalars, 8)), omega = pfloat(pload(scalars, 12));
    float ri;
    if (status[2] != 0) {
      x[i] += alpha * z[i];
      ri = sv[i];
    } else {
      x[i] += alpha * z[i] + omega * zs[i];
      ri = sv[i] - omega * t[i];
    }
    r[i] = ri;
    mine[0] = double(ri) * double(ri);
    mine[1] = double(shadow[i]) * double(ri);
  }
  __shared__ double lds[2][kBlock / 64];
  mine[0] = block_sum(mine[0], lds[0]);
  mine[1] = block_sum(mine[1], lds[1]);
  store_partials<2>(mine, partials);
}

// Single-system kernels (one BiCGStab solve per launch).
__global__ void __launch_bounds__(kBlock) k_fused_dot(int n, const float* __restrict__ x,
    const float* __restrict__ y, const int* __restrict__ status, double* __restrict__ partials) {
  fused_dot_body(n, x, y, status, partials);
}
__global__ void __launch_bounds__(kBlock) k_fused_bicg_direction(int n, const float* __restrict__ r,
    const float* __restrict__ q, float* __restrict__ p, const float* __restrict__ inverse,
    float* __restrict__ z, const float* __restrict__ scalars, const int* __restrict__ status) {
  fused_bicg_direction_body(n, r, q, p, inverse, z, scalars, status);
}
__global__ void __launch_bounds__(kBlock) k_fused_spmv_shadow_dot(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ z,
    float* __restrict__ q, const float* __restrict__ shadow, const int* __restrict__ status,
    double* __restrict__ partials) {
  fused_spmv_shadow_dot_body(n, rows, cols, values, z, q, shadow, status, partials);
}
__global__ void __launch_bounds__(kBlock) k_fused_bicg_difference(int n, const float* __restrict__ r,
    const float* __restrict__ q, float* __restrict__ sv, const float* __restrict__ inverse,
    float* __restrict__ zs, const float* __restrict__ scalars, const int* __restrict__ status,
    double* __restrict__ partials) {
  fused_bicg_difference_body(n, r, q, sv, inverse, zs, scalars, status, partials);
}
__global__ void __launch_bounds__(kBlock) k_fused_spmv_dot2(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ zs,
    float* __restrict__ t, const float* __restrict__ sv, const int* __restrict__ status,
    double* __restrict__ partials) {
  fused_spmv_dot2_body(n, rows, cols, values, zs, t, sv, status, partials);
}
__global__ void __launch_bounds__(kBlock) k_fused_bicg_step(int n, float* __restrict__ x,
    float* __restrict__ r, const float* __restrict__ z, const float* __restrict__ zs,
    const float* __restrict__ sv, const float* __restrict__ t, const float* __restrict__ shadow,
    const float* __restrict__ scalars, const int* __restrict__ status, double* __restrict__ partials) {
  fused_bicg_step_body(n, x, r, z, zs, sv, t, shadow, scalars, status, partials);
}

// Batched kernels (no reference counterpart): lane blockIdx.y of a
// kh_bicg_lane table runs the single-system body on its own vectors, values,
// scalars, status and partials region (partials + lane * stride), for the
// lanes set in `mask`. Each lane's arithmetic, grid and partial order are the
// single launch's, so its iterates are bitwise those of a solve of its own.
#define LANE_OR_RETURN                                  \
  const int lane = int(blockIdx.y);                     \
  if (((mask >> lane) & 1) == 0) return;                \
  const kh_bicg_lane& L = lanes[lane];                  \
  double* part = partials + std::size_t(lane) * stride;
__global__ void __launch_bounds__(kBlock) k_batch_bicg_direction(int n, const kh_bicg_lane* __restrict__ lanes,
                                                                 int mask, double* partials, unsigned stride) {
  LANE_OR_RETURN
  (void)part;
  fused_bicg_direction_body(n, L.r, L.q, L.p, L.inverse, L.z, L.scalars, L.status);
}
__global__ void __launch_bounds__(kBlock) k_batch_spmv_shadow_dot(int n, const kh_bicg_lane* __restrict__ lanes,
                                                                  int mask, double* partials, unsigned stride) {
  LANE_OR_RETURN
  fused_spmv_shadow_dot_body(n, L.row_ptr, L.col, L.val, L.z, L.q, L.shadow, L.status, part);
}
__global__ void __launch_bounds__(kBlock) k_batch_bicg_difference(int n, const kh_bicg_lane* __restrict__ lanes,
                                                                  int mask, double* partials, unsigned stride) {
  LANE_OR_RETURN
  fused_bicg_difference_body(n, L.r, L.q, L.s, L.inverse, L.zs, L.scalars, L.status, part);
}
__global__ void __launch_bounds__(kBlock) k_batch_spmv_dot2(int n, const kh_bicg_lane* __restrict__ lanes,
                                                            int mask, double* partials, unsigned stride) {
  LANE_OR_RETURN
  fused_spmv_dot2_body(n, L.row_ptr, L.col, L.val, L.zs, L.t, L.s, L.status, part);
}
__global__ void __launch_bounds__(kBlock) k_batch_bicg_step(int n, const kh_bicg_lane* __restrict__ lanes,
                                                            int mask, double* partials, unsigned stride) {
  LANE_OR_RETURN
  fused_bicg_step_body(n, L.x, L.r, L.z, L.zs, L.s, L.t, L.shadow, L.scalars, L.status, part);
}

// The K-cycle's inner corrections in four fused steps (no reference
// counterpart: linalg/amg.cpp:320-336 with each dot product folded into the
// kernel that produces its operand and each scalar kernel into the
// reduction's finisher; the scalars and status go through the reference's
// transitions, and the reductions are the same FP64 partial sums as the
// fused PCG kernels').
// q = A src and the partials of (d.q, rhs.d), d = src (SPD) or q (GCR),
// guarded by `guard`; with `copy` set, copy = src first, unguarded (the
// reference's dev.copy(p, z)).
__global__ void __launch_bounds__(kBlock) k_inner_spmv_dot2(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ src,
    float* __restrict__ copy, float* __restrict__ q, const float* __restrict__ rhs,
    const int* __restrict__ guard, int spd, double* __restrict__ partials) {
  const int i = row_index();
  if (copy != nullptr && i < n) copy[i] = src[i];
  if (guard[0] != 0) return;
  double mine[2] = {0.0, 0.0};
  if (i < n) {
    const float qi = spmv_row(i, rows, cols, values, src);
    q[i] = qi;
    const float d = spd != 0 ? src[i] : qi;
    mine[0] = double(d) * double(qi);
    mine[1] = double(rhs[i]) * double(d);
  }
  __shared__ double lds[2][kBlock / 64];
  mine[0] = block_sum(mine[0], lds[0]);
  mine[1] = block_sum(mine[1], lds[1]);
  store_partials<2>(mine, partials);
}
// The first correction (amg_inner_first, unguarded) and the partials of
// (t.t, b.b).
__global__ void __launch_bounds__(kBlock) k_inner_first_dot2(int n, const float* __restrict__ b,
    const float* __restrict__ p, const float* __restrict__ q, float* __restrict__ x, float* __restrict__ t,
    const float* __restrict__ s, const int* __restrict__ status, double* __restrict__ partials) {
  const int i = row_index();
  double mine[2] = {0.0, 0.0};
  if (i < n) {
    const bool running = status[0] == 0;
    const float alpha = pfloat(pload(s, 8));
    const float ti = running ? b[i] - alpha * q[i] : 0.0f;
    x[i] = running ? alpha * p[i] : p[i];
    t[i] = ti;
    mine[0] = double(ti) * double(ti);
    mine[1] = double(b[i]) * double(b[i]);
  }
  __shared__ double lds[2][kBlock / 64];
  mine[0] = block_sum(mine[0], lds[0]);
  mine[1] = block_sum(mine[1], lds[1]);
  store_partials<2>(mine, partials);
}
// The partials of w.q, w = z (SPD) or w = A z stored in `w` (GCR), guarded.
__global__ void __launch_bounds__(kBlock) k_inner_beta_dot(int n, const int* __restrict__ rows,
    const int* __restrict__ cols, const float* __restrict__ values, const float* __restrict__ z,
    float* __restrict__ w, const float* __restrict__ q, const int* __restrict__ status, int spd,
    double* __restrict__ partials) {
  if (status[0] != 0) return;
  const int i = row_index();
  double mine[1] = {0.0};
  if (i < n) {
    float v = z[i];
    if (spd == 0) {
      v = spmv_row(i, rows, cols, values, z);
      w[i] = v;
    }
    mine[0] = double(v) * double(q[i]);
  }
  __shared__ double lds[kBlock / 64];
  mine[0] = block_sum(mine[0], lds);
  store_partials<1>(mine, partials);
}
enum class FinishInner { First, Check, Beta, Second };
template <FinishInner F>
__global__ void __launch_bounds__(kBlock) k_inner_finish(const double* __restrict__ partials, unsigned count,
                                                         float* s, int* status, const int* parent) {
  if constexpr (F == FinishInner::First) {
    if (parent[0] == 0) {
      double total[2];
      sum_partials<2>(partials, count, total);
      if (threadIdx.x == 0) {
        pstore(s, 0, to_pair(total[0]));
        pstore(s, 2, to_pair(total[1]));
      }
    }
    if (threadIdx.x == 0) inner_alpha_body(s, status, parent, 1);
  } else if constexpr (F == FinishInner::Check) {
    double total[2];
    sum_partials<2>(partials, count, total);
    if (threadIdx.x != 0) return;
    pstore(s, 4, to_pair(total[0]));
    pstore(s, 6, to_pair(total[1]));
    inner_check_body(s, status);
  } else if constexpr (F == FinishInner::Beta) {
    if (status[0] != 0) return;
    double total[1];
    sum_partials<1>(partials, count, total);
    if (threadIdx.x != 0) return;
    pstore(s, 4, to_pair(total[0]));
    inner_beta_body(s, status);
  } else {
    if (status[0] != 0) return;
    double total[2];
    sum_partials<2>(partials, count, total);
    if (threadIdx.x != 0) return;
    pstore(s, 0, to_pair(total[0]));
    pstore(s, 2, to_pair(total[1]));
    inner_alpha_body(s, status, parent, 0);
  }
}

// Finishers: the FP64 total of `count` partials as a normalized pair, then the
// reference's scalar kernel on it.
enum class Finish { Alpha, Check, Beta, Pair, BicgBeta, BicgAlpha, BicgS };
template <Finish F>
__device__ __forceinline__ void finish_body(const double* __restrict__ partials, unsigned count,
                                            float* scalars, int* status, float* dst) {
  if (status[0] != 0) return;
  double total[1];
  sum_partials<1>(partials, count, total);
  if (threadIdx.x != 0) return;
  const Pair value = to_pair(total[0]);
  if constexpr (F == Finish::Alpha) {
    pstore(scalars, 2, value);
    pcg_alpha(scalars, status);
  } else if constexpr (F == Finish::Check) {
    pstore(scalars, 18, value);
    solver_check_body(scalars, status, 1);
  } else if constexpr (F == Finish::Beta) {
    pstore(scalars, 4, value);
    pcg_beta(scalars, status, 0);
  } else if constexpr (F == Finish::BicgBeta) {
    pstore(scalars, 0, value);
    bicg_beta(scalars, status);
  } else if constexpr (F == Finish::BicgAlpha) {
    pstore(scalars, 2, value);
    bicg_alpha(scalars, status);
  } else if constexpr (F == Finish::BicgS) {
    pstore(scalars, 16, value);
    bicg_s(scalars, status);
  } else {
    pstore(dst, 0, value);
  }
}
// Two-slot finishers. Omega: t.t, t.s into slots 2 and 4, then bicg_omega.
// CheckBeta: r.r into slot 18 and the check (which keeps slot 0 as the
// previous rho in slot 14), then, if the iteration goes on, shadow.r into
// slot 0 and the next iteration's bicg_beta.
enum class Finish2 { Omega, CheckBeta };
template <Finish2 F>
__device__ __forceinline__ void finish2_body(const double* __restrict__ partials, unsigned count,
                                             float* scalars, int* status) {
  if (status[0] != 0) return;
  double total[2];
  sum_partials<2>(partials, count, total);
  if (threadIdx.x != 0) return;
  if constexpr (F == Finish2::Omega) {
    pstore(scalars, 2, to_pair(total[0]));
    pstore(scalars, 4, to_pair(total[1]));
    bicg_omega(scalars, status);
  } else {
    pstore(scalars, 18, to_pair(total[0]));
    solver_check_body(scalars, status, 0);
    if (status[0] != 0) return;
    pstore(scalars, 0, to_pair(total[1]));
    bicg_beta(scalars, status);
  }
}
template <Finish F>
__global__ void __launch_bounds__(kBlock) k_finish(const double* __restrict__ partials,
                                                   unsigned count, float* scalars, int* status,
                                                   float* dst) {
  finish_body<F>(partials, count, scalars, status, dst);
}
template <Finish2 F>
__global__ void __launch_bounds__(kBlock) k_finish2(const double* __restrict__ partials,
                                                    unsigned count, float* scalars, int* status) {
  finish2_body<F>(partials, count, scalars, status);
}
// Batched finishers: one workgroup per lane (blockIdx.y).
template <Finish F>
__global__ void __launch_bounds__(kBlock) k_batch_finish(const kh_bicg_lane* __restrict__ lanes, int mask,
                                                         double* partials, unsigned stride, unsigned count) {
  LANE_OR_RETURN
  finish_body<F>(part, count, L.scalars, L.status, nullptr);
}
template <Finish2 F>
__global__ void __launch_bounds__(kBlock) k_batch_finish2(const kh_bicg_lane* __restrict__ lanes, int mask,
                                                          double* partials, unsigned stride, unsigned count) {
  LANE_OR_RETURN
  finish2_body<F>(part, count, L.scalars, L.status);
}

// Restricted residual without storing the fine residual:
// csr_residual_guarded + amg_restrict_guarded. A workgroup owns kRestrictRows
// consecutive coarse rows, whose members are consecutive in `members`. The
// members are processed in tiles of one per lane: each lane computes one
// member row's compensated residual into LDS, then the lane of each coarse row
// adds its members' residuals in member order (Kahan), carrying the sum from
// tile to tile. Every member row gets a lane of its own, as in a sweep.
constexpr int kRestrictRows = 64;
template <class V, int kGroupRows = kRestrictRows>
__global__ void __launch_bounds__(kBlock) k_fused_residual_restrict(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const V* __restrict__ values,
    const float* __restrict__ x, const float* __restrict__ b, const int* __restrict__ offsets,
    const int* __restrict__ members, float* __restrict__ coarse, const int* __restrict__ status) {
  if (status[0] != 0) return;
  __shared__ float residual[kBlock];
  static_assert(kGroupRows <= kBlock, "one owner lane per coarse row");
  const int g0 = int(xcd_block()) * kGroupRows;
  if (g0 >= n) return;
  const int g1 = min(g0 + kGroupRows, n);
  const int first = offsets[g0], last = offsets[g1];
  const int g = g0 + int(threadIdx.x);
  const bool owner = threadIdx.x < kGroupRows && g < g1;
  const int mine = owner ? offsets[g] : 0, mine_end = owner ? offsets[g + 1] : 0;
  Kahan k;
  for (int tile = first; tile < last; tile += kBlock) {
    const int m = tile + int(threadIdx.x);
    if (m < last) {
      const int row = members[m];
      residual[threadIdx.x] = row_residual(b[row], rows[row], rows[row + 1], cols, values, Plain{x});
    }
    __syncthreads();
    if (owner) {
      const int lo = max(mine, tile), hi = min(mine_end, tile + kBlock);
      for (int j = lo; j < hi; ++j) k.add(residual[j - tile]);
    }
    __syncthreads();
  }
  if (owner) coarse[g] = k.s;
}
// Prolongation followed by one smoothing sweep without storing x + P c:
// amg_prolong_guarded + amg_smooth.
template <class V>
__global__ void __launch_bounds__(kBlock) k_fused_prolong_smooth(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const V* __restrict__ values,
    const float* __restrict__ x, const float* __restrict__ b, const float* __restrict__ inverse,
    float* __restrict__ next, const int* __restrict__ aggregate, const float* __restrict__ coarse,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const Prolonged gather{x, aggregate, coarse};
  const float r = row_residual(b[i], rows[i], rows[i + 1], cols, values, gather);
  next[i] = gather(i) + (2.0f / 3.0f) * inverse[i] * r;
}

// fill(x, 0) followed by amg_smooth_zero, reading the zero iterate as the
// constant +0.0f instead of from memory (kernels/linalg.kern.h:162-170).
template <class V>
__global__ void __launch_bounds__(kBlock) k_fused_smooth_from_zero(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const V* __restrict__ values,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ next,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  const int i = row_index();
  if (i >= n) return;
  const float bi = b[i];
  if (bi - bi != 0.0f) {
    RowResidual acc(bi);
    for (int j = rows[i]; j < rows[i + 1]; ++j) acc.add(value_of(values[j]), 0.0f);
    next[i] = 0.0f + (2.0f / 3.0f) * inverse[i] * acc.value();
    return;
  }
  next[i] = 0.0f + (2.0f / 3.0f) * inverse[i] * (bi + 0.0f);
}

// The coarsest level's coarse_sweeps l1-Jacobi sweeps (linalg/amg.cpp:287-293)
// in one workgroup: the first sweep is k_fused_smooth_from_zero's row update,
// the others k_amg_smooth's, each row's arithmetic unchanged; the iterate
// ping-pongs between two LDS arrays (a workgroup barrier between sweeps) and
// only the last sweep's iterate is stored, into x. The operator stays in
// global memory, where it is L2-resident at these sizes. One launch replaces
// coarse_sweeps launches of a few microseconds each on a level with a few
// thousand rows.
constexpr int kCoarseSweepThreads = 1024;
__global__ void __launch_bounds__(kCoarseSweepThreads) k_fused_coarse_sweeps(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const float* __restrict__ values,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ x,
    const int* __restrict__ status, int sweeps) {
  if (status[0] != 0) return;
  __shared__ float iterate[2][KH_COARSE_SWEEPS_MAX_ROWS];
  for (int i = threadIdx.x; i < n; i += kCoarseSweepThreads) {
    const float bi = b[i];
    float next;
    if (bi - bi != 0.0f) {
      RowResidual acc(bi);
      for (int j = rows[i]; j < rows[i + 1]; ++j) acc.add(values[j], 0.0f);
      next = 0.0f + (2.0f / 3.0f) * inverse[i] * acc.value();
    } else {
      next = 0.0f + (2.0f / 3.0f) * inverse[i] * (bi + 0.0f);
    }
    iterate[0][i] = next;
  }
  int cur = 0;
  for (int t = 1; t < sweeps; ++t) {
    __syncthreads();
    const float* in = iterate[cur];
    float* out = iterate[cur ^ 1];
    for (int i = threadIdx.x; i < n; i += kCoarseSweepThreads) {
      const float r = row_residual(b[i], rows[i], rows[i + 1], cols, values, Plain{in});
      out[i] = in[i] + (2.0f / 3.0f) * inverse[i] * r;
    }
    cur ^= 1;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kCoarseSweepThreads) x[i] = iterate[cur][i];
}

// The same sweeps with the operator read once per launch and each sweep in
// two phases. Phase 1 is entry-parallel: thread t holds entries t,
// t + threads, ... (Cap of them) in registers and forms each entry's term
// -(a x) and product error into LDS. Phase 2 is row-serial: each thread owns consecutive rows
// (table: first row, end row; rows longer than 16 entries alone, in the
// first wavefronts) and runs the compensated sum of each row over the
// precomputed terms, in entry order, then updates the row in place (no other
// thread reads x in phase 2). The arithmetic is RowResidual::add split in
// two, so the iterates are k_fused_coarse_sweeps' bits; a long row's serial
// part is three additions per entry instead of the whole product.
constexpr int kResidentThreads = KH_RESIDENT_THREADS;
template <int Cap>
__global__ void __launch_bounds__(kResidentThreads) k_coarse_sweeps_resident(int n,
    const int* __restrict__ rows, const int* __restrict__ cols, const float* __restrict__ values,
    const float* __restrict__ b, const float* __restrict__ inverse, float* __restrict__ x,
    const int* __restrict__ status, int sweeps, const int* __restrict__ table) {
  static_assert(Cap == 8 || Cap == 16, "Cap is 8 or 16");
  if (status[0] != 0) return;
  __shared__ float iterate[KH_RESIDENT_MAX_ROWS];
  __shared__ float2 bw[KH_RESIDENT_MAX_ROWS];  // b and (2/3) l1 inverse, as the sweeps form it
  __shared__ int row_start[KH_RESIDENT_MAX_ROWS + 1];
  __shared__ float term[KH_RESIDENT_THREADS * 16];
  __shared__ float error[KH_RESIDENT_THREADS * 16];
  const int t = threadIdx.x;
  const int nnz = rows[n];
  if (nnz > kResidentThreads * Cap) return;  // not this partition's operator (uniform, before any barrier)
  for (int i = t; i < n; i += kResidentThreads) {
    bw[i] = make_float2(b[i], (2.0f / 3.0f) * inverse[i]);
    iterate[i] = 0.0f;
  }
  for (int i = t; i <= n; i += kResidentThreads) row_start[i] = rows[i];
  // Entry t + k*threads is this thread's k-th: neighbouring lanes touch
  // neighbouring words of the term and error arrays (no LDS bank conflicts).
  float val[Cap];
  int col[Cap];
#pragma unroll
  for (int k = 0; k < Cap; ++k) {
    const int j = t + k * kResidentThreads;
    val[k] = j < nnz ? values[j] : 0.0f;
    col[k] = j < nnz ? cols[j] : 0;
  }
  const int r0 = table[2 * t], r1 = table[2 * t + 1];
  for (int s = 0; s < sweeps; ++s) {
    __syncthreads();
    // Phase 1: terms and product errors of this thread's entries.
#pragma unroll
    for (int c = 0; c < Cap; c += 8) {
      float xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = iterate[col[c + u]];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = t + (c + u) * kResidentThreads;
        if (j < nnz) {
          const float prod = val[c + u] * xv[u];
          term[j] = -prod;
          error[j] = product_error(val[c + u], xv[u], prod);
        }
      }
    }
    __syncthreads();
    // Phase 2: each row's compensated sum in entry order, then its update.
    for (int r = r0; r < r1; ++r) {
      const float2 q = bw[r];
      if (s == 0 && q.x - q.x == 0.0f) {
        iterate[r] = 0.0f + q.y * (q.x + 0.0f);
        continue;
      }
      RowResidual acc(q.x);
      const int e1 = row_start[r + 1];
      for (int j = row_start[r]; j < e1; j += 8) {
        float tv[8], ev[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (j + u < e1) {
            tv[u] = term[j + u];
            ev[u] = error[j + u];
          }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (j + u < e1) acc.add_split(tv[u], ev[u]);
      }
      iterate[r] = iterate[r] + q.y * acc.value();
    }
  }
  __syncthreads();
  for (int i = t; i < n; i += kResidentThreads) x[i] = iterate[i];
}

// FP32 -> bfloat16, round to nearest even (NaN stays NaN): the mixed path's
// coarse-level value copies.
__global__ void __launch_bounds__(kBlock) k_to_bf16(int n, const float* __restrict__ src, bf16v* __restrict__ dst) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  unsigned u = __float_as_uint(src[i]);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    dst[i].bits = (unsigned short)((u >> 16) | 0x40u);
    return;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  dst[i].bits = (unsigned short)(u >> 16);
}

// -------------------------------------------------------- coarse-grid tail
// Below a few thousand rows a level's launches are pure latency: each sweep
// is a handful of dependent memory round trips plus a kernel boundary. The
// tail runs every remaining level of the V-cycle (linalg/amg.cpp:283-310) in
// one workgroup, with sweeps separated by workgroup barriers and everything
// in LDS: each level's iterate and right-hand side for the whole cycle, and,
// on each visit to a level, a copy of its operator (row starts, columns,
// values, l1 inverse, aggregate map). The code is kept small on purpose: one
// workgroup re-executes it for every sweep from one compute unit's
// instruction cache. Every row is computed by the same expression as the
// per-level kernels (l1-Jacobi, kernels/linalg.kern.h:150-170; restriction,
// amg.kern.h:5-16; coarse solve, linalg.kern.h:126-147), and Jacobi rows only
// read the previous sweep, so the iterates are those of the launch sequence
// it replaces.
constexpr int kTailThreads = 1024;
constexpr int kTailPoolWords = KH_TAIL_POOL_WORDS;

// Words an operator copy of a level takes: row starts, columns, values, the
// l1 inverse and the aggregate map.
__device__ __forceinline__ int tail_stage_words(int rows, int nnz) { return 3 * rows + 1 + 2 * nnz; }

// A level's operator, copied into LDS.
struct TailOperator {
  const int* row_ptr;
  const int* col;
  const float* val;
  const float* inverse;
  const int* aggregate;
  int rows;
};

// Compensated residual b - sum a[j] x[col[j]] of row i, entries in order,
// columns, values and gathers loaded a chunk at a time.
template <class X>
__device__ __forceinline__ float tail_residual(const TailOperator& L, float b, int i, X gather) {
  return row_residual_chunked<kChunk>(b, L.row_ptr[i], L.row_ptr[i + 1], L.col, L.val, gather);
}

enum class TailSweep { Plain, FromZero, Prolonged };

// One sweep over a level: every row's new value goes to `out`, computed from
// `in`. Jacobi rows read only the previous iterate, so `in` and `out` differ
// (FromZero reads no iterate); one barrier ends the sweep.
__device__ void tail_sweep(TailSweep kind, const TailOperator& L, const float* in, const float* b,
                           const float* coarse, float* out) {
  for (int i = threadIdx.x; i < L.rows; i += kTailThreads) {
    const float bi = b[i];
    const float w = (2.0f / 3.0f) * L.inverse[i];
    float value;
    if (kind == TailSweep::FromZero) {
      value = bi - bi != 0.0f ? 0.0f + w * tail_residual(L, bi, i, [](int) { return 0.0f; })
                              : 0.0f + w * (bi + 0.0f);
    } else if (kind == TailSweep::Prolonged) {
      const Prolonged gather{in, L.aggregate, coarse};
      value = gather(i) + w * tail_residual(L, bi, i, gather);
    } else {
      value = in[i] + w * tail_residual(L, bi, i, Plain{in});
    }
    out[i] = value;
  }
  __syncthreads();
}
// `count` sweeps from the zero iterate, alternating between x and `next` so
// that the last writes x.
__device__ void tail_smooth_from_zero(const TailOperator& L, float* x, const float* b, float* next, int count) {
  const float* cur = nullptr;
  for (int t = 1; t <= count; ++t) {
    float* out = (count - t) % 2 == 0 ? x : next;
    tail_sweep(t == 1 ? TailSweep::FromZero : TailSweep::Plain, L, cur, b, nullptr, out);
    cur = out;
  }
}

// Restricted residual of level L into the next level's right-hand side.
__device__ void tail_restrict(const TailOperator& L, const int* __restrict__ offsets,
                              const int* __restrict__ members, int coarse_rows, const float* x,
                              const float* b, float* coarse_b) {
  for (int g = threadIdx.x; g < coarse_rows; g += kTailThreads) {
    Kahan k;
    for (int m = offsets[g]; m < offsets[g + 1]; ++m) {
      const int row = members[m];
      k.add(tail_residual(L, b[row], row, Plain{x}));
    }
    coarse_b[g] = k.s;
  }
  __syncthreads();
}

// Constant-nullspace projection: sum_into + solver_project
// (linalg/detail.hpp:13-19; kernels/scalars.kern.h:150-152).
__device__ void tail_project(float* x, int rows, double* red, float* pair) {
  double v = 0.0;
  for (int i = threadIdx.x; i < rows; i += kTailThreads) v += double(x[i]);
  v = block_sum(v, red);
  if (threadIdx.x == 0) {
    const Pair s = to_pair(v);
    pair[0] = s.hi;
    pair[1] = s.lo;
  }
  __syncthreads();
  const float c = static_cast<float>(rows);
  const float mean =
      pfloat(pdiv({pair[0], pair[1]}, {c, static_cast<float>(double(rows) - double(c))}));
  for (int i = threadIdx.x; i < rows; i += kTailThreads) x[i] -= mean;
  __syncthreads();
}

// Copies `count` words from global memory into LDS, eight loads in flight
// per lane.
__device__ void tail_copy(uint32_t* dst, const uint32_t* __restrict__ src, int count) {
  constexpr int kUnroll = 8;
  int k = threadIdx.x;
  for (; k + (kUnroll - 1) * kTailThreads < count; k += kUnroll * kTailThreads) {
    uint32_t v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = src[k + u * kTailThreads];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) dst[k + u * kTailThreads] = v[u];
  }
  for (; k < count; k += kTailThreads) dst[k] = src[k];
}

// Copies level T's operator into LDS at `stage` and describes it.
__device__ TailOperator tail_stage(const kh_tail_level& T, bool has_aggregate, uint32_t* stage) {
  const int rows = T.rows, nnz = T.nnz;
  uint32_t* row_ptr = stage;
  uint32_t* col = row_ptr + rows + 1;
  uint32_t* val = col + nnz;
  uint32_t* inverse = val + nnz;
  uint32_t* aggregate = inverse + rows;
  tail_copy(row_ptr, reinterpret_cast<const uint32_t*>(T.row_ptr), rows + 1);
  tail_copy(col, reinterpret_cast<const uint32_t*>(T.col), nnz);
  tail_copy(val, reinterpret_cast<const uint32_t*>(T.val), nnz);
  tail_copy(inverse, reinterpret_cast<const uint32_t*>(T.inverse), rows);
  if (has_aggregate) tail_copy(aggregate, reinterpret_cast<const uint32_t*>(T.aggregate), rows);
  __syncthreads();
  return {reinterpret_cast<const int*>(row_ptr), reinterpret_cast<const int*>(col),
          reinterpret_cast<const float*>(val), reinterpret_cast<const float*>(inverse),
          reinterpret_cast<const int*>(aggregate), rows};
}

__global__ void __launch_bounds__(kTailThreads) k_fused_coarse_tail(
    const kh_tail_level* __restrict__ levels, int count, int pre, int post, int coarse_sweeps,
    int nullspace, const float* __restrict__ rhs, float* __restrict__ out,
    const int* __restrict__ status) {
  if (status[0] != 0) return;
  __shared__ uint32_t pool[kTailPoolWords];
  __shared__ double red[kTailThreads / 64];
  __shared__ float pair[2];
  // Level k keeps its iterate at vectors[base[k]] and right-hand side right
  // after; `next` holds a sweep's new values; `stage` the visited operator.
  __shared__ int base[KH_TAIL_MAX_LEVELS];
  // The table is validated by every thread alike, so an invalid one returns
  // uniformly before any barrier.
  if (count < 1 || count > KH_TAIL_MAX_LEVELS) return;
  int used = 0, widest = 0, largest = 0;
  for (int k = 0; k < count; ++k) {
    const int rows = levels[k].rows, nnz = levels[k].nnz;
    if (rows < 1 || rows > KH_TAIL_MAX_ROWS || nnz < 0 || nnz > kTailPoolWords) return;
    if (k + 1 < count && levels[k].coarse_rows != levels[k + 1].rows) return;
    if (threadIdx.x == 0) base[k] = used;
    used += 2 * rows;
    widest = max(widest, rows);
    largest = max(largest, tail_stage_words(rows, nnz));
  }
  if (used + widest + largest > kTailPoolWords) return;
  __syncthreads();
  float* vectors = reinterpret_cast<float*>(pool);
  float* next = vectors + used;
  uint32_t* stage = pool + used + widest;
  {
    const int rows = levels[0].rows;
    for (int i = threadIdx.x; i < rows; i += kTailThreads) vectors[base[0] + rows + i] = rhs[i];
  }
  __syncthreads();
  for (int k = 0; k + 1 < count; ++k) {
    const kh_tail_level T = levels[k];
    float* x = vectors + base[k];
    const float* b = x + T.rows;
    float* coarse_b = vectors + base[k + 1] + T.coarse_rows;
    const TailOperator L = tail_stage(T, true, stage);
    tail_smooth_from_zero(L, x, b, next, pre);
    tail_restrict(L, T.member_offsets, T.members, T.coarse_rows, x, b, coarse_b);
    if (nullspace) tail_project(coarse_b, T.coarse_rows, red, pair);
  }
  {
    const kh_tail_level T = levels[count - 1];
    float* x = vectors + base[count - 1];
    const float* b = x + T.rows;
    const TailOperator L = tail_stage(T, false, stage);
    if (T.rows <= kCoarseMax) {
      // amg_coarse_solve: sweeps from the zero iterate, general residual form
      // (both buffers start at zero, so the first sweep may read either).
      for (int i = threadIdx.x; i < T.rows; i += kTailThreads) x[i] = next[i] = 0.0f;
      __syncthreads();
      for (int t = 0; t < coarse_sweeps; ++t) {
        float* out = (coarse_sweeps - 1 - t) % 2 == 0 ? x : next;
        tail_sweep(TailSweep::Plain, L, out == x ? next : x, b, nullptr, out);
      }
    } else {
      tail_smooth_from_zero(L, x, b, next, coarse_sweeps);
    }
    if (nullspace) tail_project(x, T.rows, red, pair);
  }
  for (int k = count - 2; k >= 0; --k) {
    const kh_tail_level T = levels[k];
    float* x = vectors + base[k];
    const float* b = x + T.rows;
    const TailOperator L = tail_stage(T, true, stage);
    // The prolongation sweep reads x, so it writes `next`; with an odd sweep
    // count the result returns to x by a copy, then the remaining sweeps
    // alternate so that the last writes x.
    tail_sweep(TailSweep::Prolonged, L, x, b, vectors + base[k + 1], next);
    const float* cur = next;
    if (post % 2 == 1) {
      for (int i = threadIdx.x; i < T.rows; i += kTailThreads) x[i] = next[i];
      __syncthreads();
      cur = x;
    }
    for (int t = 2; t <= post; ++t) {
      float* out = (post - t) % 2 == 0 ? x : next;
      tail_sweep(TailSweep::Plain, L, cur, b, nullptr, out);
      cur = out;
    }
    if (nullspace) tail_project(x, T.rows, red, pair);
  }
  for (int i = threadIdx.x; i < levels[0].rows; i += kTailThreads) out[i] = vectors[base[0] + i];
}

// ----------------------------------------------------------------- launchers
#define KB(k) static_cast<const int*>(buf(a, k))
#define KF(k) static_cast<float*>(buf(a, k))
#define KCF(k) static_cast<const float*>(buf(a, k))
#define GRID(n) dim3(blocks_for(n)), dim3(kBlock), 0, d->stream

void l_csr_spmv(kh_device* d, int n, const kh_arg* a) {
  k_csr_spmv<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), nullptr);
}
void l_csr_spmv_guarded(kh_device* d, int n, const kh_arg* a) {
  k_csr_spmv<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), KB(5));
}
void l_csr_residual(kh_device* d, int n, const kh_arg* a) {
  k_csr_residual<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5), nullptr);
}
void l_csr_residual_guarded(kh_device* d, int n, const kh_arg* a) {
  k_csr_residual<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5), KB(6));
}
void l_krylov_diagonal(kh_device* d, int n, const kh_arg* a) {
  k_krylov_diagonal<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KF(3), KF(4), i32(a, 5));
}
void l_krylov_multiply(kh_device* d, int n, const kh_arg* a) {
  k_krylov_multiply<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), nullptr);
}
void l_krylov_multiply_guarded(kh_device* d, int n, const kh_arg* a) {
  k_krylov_multiply<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), KB(3));
}
void l_krylov_smooth(kh_device* d, int n, const kh_arg* a) {
  k_krylov_smooth<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), f32(a, 3));
}
void l_krylov_shift(kh_device* d, int n, const kh_arg* a) {
  k_krylov_shift<<<GRID(n)>>>(n, KF(0), f32(a, 1));
}
void l_krylov_direction(kh_device* d, int n, const kh_arg* a) {
  k_krylov_direction<<<GRID(n)>>>(n, KCF(0), KF(1), f32(a, 2));
}
void l_krylov_step(kh_device* d, int n, const kh_arg* a) {
  k_krylov_step<<<GRID(n)>>>(n, KF(0), KF(1), KCF(2), KCF(3), f32(a, 4));
}
void l_krylov_bicg_direction(kh_device* d, int n, const kh_arg* a) {
  k_krylov_bicg_direction<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), f32(a, 3), f32(a, 4));
}
void l_krylov_difference(kh_device* d, int n, const kh_arg* a) {
  k_krylov_difference<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), f32(a, 3));
}
void l_krylov_bicg_step(kh_device* d, int n, const kh_arg* a) {
  k_krylov_bicg_step<<<GRID(n)>>>(n, KF(0), KCF(1), KCF(2), f32(a, 3), f32(a, 4));
}
void l_amg_coarse_solve(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 7);
  if (count < 0 || count > kCoarseMax) throw std::runtime_error("amg_coarse_solve: count must be in [0, 24]");
  if (n < 1) return;
  k_amg_coarse_solve<<<dim3(1), dim3(64), 0, d->stream>>>(KB(0), KB(1), KCF(2), KF(3), KCF(4),
                                                           KCF(5), KB(6), count, i32(a, 8));
}
void l_amg_smooth(kh_device* d, int n, const kh_arg* a) {
  if (d->packet_start) {
    hipExtLaunchKernelGGL(k_amg_smooth<float>, dim3(blocks_for(n)), dim3(kBlock), 0, d->stream,
                          d->packet_start, d->packet_stop, 0, n, KB(0), KB(1), KCF(2), KCF(3),
                          KCF(4), KCF(5), KF(6), KB(7));
  } else {
    k_amg_smooth<float><<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6), KB(7));
  }
}
void l_amg_smooth_zero(kh_device* d, int n, const kh_arg* a) {
  k_amg_smooth_zero<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6), KB(7));
}
void l_amg_restrict(kh_device* d, int n, const kh_arg* a) {
  k_amg_restrict<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KF(3), nullptr);
}
void l_amg_restrict_guarded(kh_device* d, int n, const kh_arg* a) {
  k_amg_restrict<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KF(3), KB(4));
}
void l_amg_prolong(kh_device* d, int n, const kh_arg* a) {
  k_amg_prolong<<<GRID(n)>>>(n, KB(0), KCF(1), KF(2), nullptr);
}
void l_amg_prolong_guarded(kh_device* d, int n, const kh_arg* a) {
  k_amg_prolong<<<GRID(n)>>>(n, KB(0), KCF(1), KF(2), KB(3));
}
void l_amg_galerkin(kh_device* d, int n, const kh_arg* a) {
  k_amg_galerkin<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KF(3));
}
void l_amg_m_matrix_diagonal(kh_device* d, int n, const kh_arg* a) {
  k_amg_m_matrix_diagonal<<<GRID(n)>>>(n, KB(0), KB(1), KB(2), KB(3), KCF(4), KB(5), KB(6), KF(7));
}
void l_axpy(kh_device* d, int n, const kh_arg* a) { k_axpy<<<GRID(n)>>>(n, KF(0), KCF(1), f32(a, 2)); }
void l_xpby(kh_device* d, int n, const kh_arg* a) { k_xpby<<<GRID(n)>>>(n, KF(0), KCF(1), f32(a, 2)); }
void l_axpbyz(kh_device* d, int n, const kh_arg* a) {
  k_axpbyz<<<GRID(n)>>>(n, KF(0), KCF(1), KCF(2), f32(a, 3), f32(a, 4));
}
void l_scale_copy(kh_device* d, int n, const kh_arg* a) {
  k_scale_copy<<<GRID(n)>>>(n, KF(0), KCF(1), f32(a, 2));
}
void l_mul(kh_device* d, int n, const kh_arg* a) { k_mul<<<GRID(n)>>>(n, KF(0), KCF(1), KCF(2)); }
#define SINGLE dim3(1), dim3(64), 0, d->stream
#define ROWWAVE_GRID dim3(unsigned((n + kBlock / 64 - 1) / (kBlock / 64))), dim3(kBlock), 0, d->stream
void l_amg_smooth_rowwave(kh_device* d, int n, const kh_arg* a) {
  k_rowwave<RowWaveOut::Smooth><<<ROWWAVE_GRID>>>(n, KB(0), KB(1), KCF(2), Plain{KCF(3)}, KCF(4), KCF(5), KF(6),
                                                  KB(7));
}
// rows, cols, values, x, b, r, status
void l_csr_residual_rowwave(kh_device* d, int n, const kh_arg* a) {
  k_rowwave<RowWaveOut::Residual><<<ROWWAVE_GRID>>>(n, KB(0), KB(1), KCF(2), Plain{KCF(3)}, KCF(4), nullptr,
                                                    KF(5), KB(6));
}
// rows, cols, values, x, b, inverse, next, aggregate, coarse, status
void l_prolong_smooth_rowwave(kh_device* d, int n, const kh_arg* a) {
  k_rowwave<RowWaveOut::Smooth><<<ROWWAVE_GRID>>>(n, KB(0), KB(1), KCF(2), Prolonged{KCF(3), KB(7), KCF(8)},
                                                  KCF(4), KCF(5), KF(6), KB(9));
}
void l_amg_smooth_wide(kh_device* d, int n, const kh_arg* a) {
  k_amg_smooth<float, true><<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6), KB(7));
}
void l_solver_limits(kh_device* d, int n, const kh_arg* a) {
  k_solver_limits<<<GRID(n)>>>(n, KF(0), f32(a, 1), f32(a, 2), f32(a, 3), f32(a, 4));
}
void l_solver_reset(kh_device* d, int, const kh_arg* a) {
  k_solver_reset<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)), i32(a, 2));
}
void l_solver_pcg_alpha(kh_device* d, int, const kh_arg* a) {
  k_solver_pcg_alpha<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)));
}
void l_solver_pcg_beta(kh_device* d, int, const kh_arg* a) {
  k_solver_pcg_beta<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)), i32(a, 2));
}
void l_solver_bicg_beta(kh_device* d, int, const kh_arg* a) {
  k_solver_bicg_beta<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)));
}
void l_solver_bicg_alpha(kh_device* d, int, const kh_arg* a) {
  k_solver_bicg_alpha<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)));
}
void l_solver_bicg_s(kh_device* d, int, const kh_arg* a) {
  k_solver_bicg_s<<<SINGLE>>>(KCF(0), static_cast<int*>(buf(a, 1)));
}
void l_solver_bicg_omega(kh_device* d, int, const kh_arg* a) {
  k_solver_bicg_omega<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)));
}
void l_solver_check(kh_device* d, int, const kh_arg* a) {
  k_solver_check<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)), i32(a, 2));
}
void l_solver_step(kh_device* d, int n, const kh_arg* a) {
  k_solver_step<<<GRID(n)>>>(n, KF(0), KF(1), KCF(2), KCF(3), KCF(4), KB(5));
}
void l_solver_direction(kh_device* d, int n, const kh_arg* a) {
  k_solver_direction<<<GRID(n)>>>(n, KCF(0), KF(1), KCF(2), KB(3));
}
void l_amg_inner_alpha(kh_device* d, int, const kh_arg* a) {
  k_amg_inner_alpha<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)), KB(2), i32(a, 3));
}
void l_amg_inner_first(kh_device* d, int n, const kh_arg* a) {
  k_amg_inner_first<<<GRID(n)>>>(n, KCF(0), KCF(1), KCF(2), KF(3), KF(4), KCF(5), KB(6));
}
void l_amg_inner_check(kh_device* d, int, const kh_arg* a) {
  k_amg_inner_check<<<SINGLE>>>(KCF(0), static_cast<int*>(buf(a, 1)));
}
void l_amg_inner_beta(kh_device* d, int, const kh_arg* a) {
  k_amg_inner_beta<<<SINGLE>>>(KF(0), static_cast<int*>(buf(a, 1)));
}
void l_amg_inner_step(kh_device* d, int n, const kh_arg* a) {
  k_amg_inner_step<<<GRID(n)>>>(n, KF(0), KCF(1), KCF(2), KB(3));
}
void l_solver_bicg_direction(kh_device* d, int n, const kh_arg* a) {
  k_solver_bicg_direction<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), KCF(3), KB(4));
}
void l_solver_bicg_difference(kh_device* d, int n, const kh_arg* a) {
  k_solver_bicg_difference<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), KCF(3), KB(4));
}
void l_solver_bicg_step(kh_device* d, int n, const kh_arg* a) {
  k_solver_bicg_step<<<GRID(n)>>>(n, KF(0), KF(1), KCF(2), KCF(3), KCF(4), KCF(5), KCF(6), KB(7));
}
void l_solver_project(kh_device* d, int n, const kh_arg* a) {
  k_solver_project<<<GRID(n)>>>(n, KF(0), KCF(1), f32(a, 2), f32(a, 3));
}
// Fused reductions use one partial slot per workgroup of 256 rows.
unsigned reduce_grid(int n) {
  const unsigned blocks = blocks_for(std::size_t(n));
  if (blocks > unsigned(kMaxReduceBlocks))
    throw std::runtime_error("fused reduction kernel: n exceeds 16,777,216 rows");
  return blocks;
}
#define FINISH dim3(1), dim3(kBlock), 0, d->stream
void l_fused_spmv_dot_alpha(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 6));
  k_fused_spmv_dot<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), st, d->partials);
  k_finish<Finish::Alpha><<<FINISH>>>(d->partials, g, KF(5), st, nullptr);
}
void l_inner_first_step(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 8));
  k_inner_spmv_dot2<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), KF(5), KCF(6), KB(9), i32(a, 10),
                                 d->partials);
  k_inner_finish<FinishInner::First><<<FINISH>>>(d->partials, g, KF(7), st, KB(9));
}
void l_inner_first_update(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 6));
  k_inner_first_dot2<<<GRID(n)>>>(n, KCF(0), KCF(1), KCF(2), KF(3), KF(4), KCF(5), st, d->partials);
  k_inner_finish<FinishInner::Check><<<FINISH>>>(d->partials, g, KF(5), st, nullptr);
}
void l_inner_beta_step(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 7));
  k_inner_beta_dot<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), KCF(5), st, i32(a, 8), d->partials);
  k_inner_finish<FinishInner::Beta><<<FINISH>>>(d->partials, g, KF(6), st, nullptr);
}
void l_inner_second_step(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 7));
  k_inner_spmv_dot2<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), nullptr, KF(4), KCF(5), st, i32(a, 9),
                                 d->partials);
  k_inner_finish<FinishInner::Second><<<FINISH>>>(d->partials, g, KF(6), st, KB(8));
}
void l_fused_step_dot_check(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 5));
  k_fused_step_dot<<<GRID(n)>>>(n, KF(0), KF(1), KCF(2), KCF(3), KCF(4), st, d->partials);
  k_finish<Finish::Check><<<FINISH>>>(d->partials, g, KF(4), st, nullptr);
}
void l_fused_residual_restrict(kh_device* d, int n, const kh_arg* a) {
  const unsigned groups = unsigned((n + kRestrictRows - 1) / kRestrictRows);
  k_fused_residual_restrict<float><<<dim3(groups), dim3(kBlock), 0, d->stream>>>(
      n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KB(5), KB(6), KF(7), KB(8));
}
void l_fused_residual_restrict_wide(kh_device* d, int n, const kh_arg* a) {
  const unsigned groups = unsigned((n + 127) / 128);
  k_fused_residual_restrict<float, 128><<<dim3(groups), dim3(kBlock), 0, d->stream>>>(
      n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KB(5), KB(6), KF(7), KB(8));
}
void l_fused_coarse_sweeps(kh_device* d, int n, const kh_arg* a) {
  if (n > KH_COARSE_SWEEPS_MAX_ROWS)
    throw std::runtime_error("fused_coarse_sweeps: more rows than KH_COARSE_SWEEPS_MAX_ROWS");
  if (i32(a, 7) < 1) throw std::runtime_error("fused_coarse_sweeps: at least one sweep");
  k_fused_coarse_sweeps<<<dim3(1), dim3(kCoarseSweepThreads), 0, d->stream>>>(
      n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5), KB(6), i32(a, 7));
}
void l_coarse_sweeps_resident(kh_device* d, int n, const kh_arg* a) {
  if (n < 1 || n > KH_RESIDENT_MAX_ROWS)
    throw std::runtime_error("coarse_sweeps_resident: rows outside 1..KH_RESIDENT_MAX_ROWS");
  if (i32(a, 7) < 1) throw std::runtime_error("coarse_sweeps_resident: at least one sweep");
  const dim3 grid(1), block(kResidentThreads);
  switch (i32(a, 9)) {
    case 8:
      k_coarse_sweeps_resident<8><<<grid, block, 0, d->stream>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5),
                                                                  KB(6), i32(a, 7), KB(8));
      break;
    case 16:
      k_coarse_sweeps_resident<16><<<grid, block, 0, d->stream>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5),
                                                                   KB(6), i32(a, 7), KB(8));
      break;
    default:
      throw std::runtime_error("coarse_sweeps_resident: cap must be 8 or 16");
  }
}
void l_fused_prolong_smooth(kh_device* d, int n, const kh_arg* a) {
  k_fused_prolong_smooth<float><<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6),
                                             KB(7), KCF(8), KB(9));
}
void l_fused_smooth_dot_beta(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 7));
  k_fused_smooth_dot<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6), st,
                                  d->partials);
  k_finish<Finish::Beta><<<FINISH>>>(d->partials, g, KF(8), st, nullptr);
}
void l_fused_smooth_sum(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = const_cast<int*>(KB(7));
  k_fused_smooth_partial_sum<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KCF(5), KF(6),
                                          st, d->partials);
  k_finish<Finish::Pair><<<FINISH>>>(d->partials, g, nullptr, st, KF(8));
}
void l_fused_smooth_from_zero(kh_device* d, int n, const kh_arg* a) {
  k_fused_smooth_from_zero<float><<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KCF(4), KF(5), KB(6));
}
// The bfloat16-valued variants (values argument: KH_TO_BF16's output).
#define KBF(k) static_cast<const bf16v*>(buf(a, k))
void l_bf16_amg_smooth(kh_device* d, int n, const kh_arg* a) {
  k_amg_smooth<bf16v><<<GRID(n)>>>(n, KB(0), KB(1), KBF(2), KCF(3), KCF(4), KCF(5), KF(6), KB(7));
}
void l_bf16_residual_restrict(kh_device* d, int n, const kh_arg* a) {
  const unsigned groups = unsigned((n + kRestrictRows - 1) / kRestrictRows);
  k_fused_residual_restrict<bf16v><<<dim3(groups), dim3(kBlock), 0, d->stream>>>(
      n, KB(0), KB(1), KBF(2), KCF(3), KCF(4), KB(5), KB(6), KF(7), KB(8));
}
void l_bf16_prolong_smooth(kh_device* d, int n, const kh_arg* a) {
  k_fused_prolong_smooth<bf16v><<<GRID(n)>>>(n, KB(0), KB(1), KBF(2), KCF(3), KCF(4), KCF(5), KF(6),
                                             KB(7), KCF(8), KB(9));
}
void l_bf16_smooth_from_zero(kh_device* d, int n, const kh_arg* a) {
  k_fused_smooth_from_zero<bf16v><<<GRID(n)>>>(n, KB(0), KB(1), KBF(2), KCF(3), KCF(4), KF(5), KB(6));
}
// n = entries; src (FP32), dst (bf16)
void l_to_bf16(kh_device* d, int n, const kh_arg* a) {
  k_to_bf16<<<GRID(n)>>>(n, KCF(0), static_cast<bf16v*>(buf(a, 1)));
}
void l_fused_coarse_tail(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1);
  if (count < 1 || count > KH_TAIL_MAX_LEVELS)
    throw std::runtime_error("fused_coarse_tail: level count must be in [1, 16]");
  if (i32(a, 2) < 1 || i32(a, 3) < 1 || i32(a, 4) < 1)
    throw std::runtime_error("fused_coarse_tail: sweep counts must be positive");
  if (n < 1) return;
  k_fused_coarse_tail<<<dim3(1), dim3(kTailThreads), 0, d->stream>>>(
      static_cast<const kh_tail_level*>(buf(a, 0)), count, i32(a, 2), i32(a, 3), i32(a, 4),
      i32(a, 5), KCF(6), KF(7), KB(8));
}
void l_fused_dot_bicg_beta(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 3));
  k_fused_dot<<<GRID(n)>>>(n, KCF(0), KCF(1), st, d->partials);
  k_finish<Finish::BicgBeta><<<FINISH>>>(d->partials, g, KF(2), st, nullptr);
}
void l_fused_bicg_direction(kh_device* d, int n, const kh_arg* a) {
  k_fused_bicg_direction<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), KCF(3), KF(4), KCF(5), KB(6));
}
void l_fused_spmv_shadow_alpha(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 7));
  k_fused_spmv_shadow_dot<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), KCF(5), st, d->partials);
  k_finish<Finish::BicgAlpha><<<FINISH>>>(d->partials, g, KF(6), st, nullptr);
}
void l_fused_bicg_difference_s(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 6));
  k_fused_bicg_difference<<<GRID(n)>>>(n, KCF(0), KCF(1), KF(2), KCF(3), KF(4), KCF(5), st, d->partials);
  k_finish<Finish::BicgS><<<FINISH>>>(d->partials, g, KF(5), st, nullptr);
}
void l_fused_spmv_dot2_omega(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 7));
  k_fused_spmv_dot2<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KCF(3), KF(4), KCF(5), st, d->partials);
  k_finish2<Finish2::Omega><<<FINISH>>>(d->partials, g, KF(6), st);
}
void l_fused_bicg_step_check_beta(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 8));
  k_fused_bicg_step<<<GRID(n)>>>(n, KF(0), KF(1), KCF(2), KCF(3), KCF(4), KCF(5), KCF(6), KCF(7), st,
                                 d->partials);
  k_finish2<Finish2::CheckBeta><<<FINISH>>>(d->partials, g, KF(7), st);
}
// Batched BiCGStab-Jacobi launches: n = rows of every lane; lanes (device
// array of kh_bicg_lane), lane count (1..KH_BATCH_MAX_LANES), mask of the
// lanes that take the step. Lane k's partials occupy reduce_grid(n) x 2
// doubles from d->partials + k * that.
unsigned batch_stride(int n, int count) {
  const unsigned g = reduce_grid(n);
  if (count < 1 || count > KH_BATCH_MAX_LANES || std::size_t(g) * 2 * count > std::size_t(kMaxReduceBlocks) * 2)
    throw std::runtime_error("batched BiCGStab: lane count out of range for this row count");
  return 2 * g;
}
#define BATCH_GRID(n, lanes) dim3(blocks_for(n), unsigned(lanes)), dim3(kBlock), 0, d->stream
#define BATCH_FINISH(lanes) dim3(1, unsigned(lanes)), dim3(kBlock), 0, d->stream
void l_batch_bicg_direction(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1), mask = i32(a, 2);
  const unsigned stride = batch_stride(n, count);
  const auto* lanes = static_cast<const kh_bicg_lane*>(buf(a, 0));
  k_batch_bicg_direction<<<BATCH_GRID(n, count)>>>(n, lanes, mask, d->partials, stride);
}
void l_batch_spmv_shadow_alpha(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1), mask = i32(a, 2);
  const unsigned stride = batch_stride(n, count), g = reduce_grid(n);
  const auto* lanes = static_cast<const kh_bicg_lane*>(buf(a, 0));
  k_batch_spmv_shadow_dot<<<BATCH_GRID(n, count)>>>(n, lanes, mask, d->partials, stride);
  k_batch_finish<Finish::BicgAlpha><<<BATCH_FINISH(count)>>>(lanes, mask, d->partials, stride, g);
}
void l_batch_bicg_difference_s(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1), mask = i32(a, 2);
  const unsigned stride = batch_stride(n, count), g = reduce_grid(n);
  const auto* lanes = static_cast<const kh_bicg_lane*>(buf(a, 0));
  k_batch_bicg_difference<<<BATCH_GRID(n, count)>>>(n, lanes, mask, d->partials, stride);
  k_batch_finish<Finish::BicgS><<<BATCH_FINISH(count)>>>(lanes, mask, d->partials, stride, g);
}
void l_batch_spmv_dot2_omega(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1), mask = i32(a, 2);
  const unsigned stride = batch_stride(n, count), g = reduce_grid(n);
  const auto* lanes = static_cast<const kh_bicg_lane*>(buf(a, 0));
  k_batch_spmv_dot2<<<BATCH_GRID(n, count)>>>(n, lanes, mask, d->partials, stride);
  k_batch_finish2<Finish2::Omega><<<BATCH_FINISH(count)>>>(lanes, mask, d->partials, stride, g);
}
void l_batch_bicg_step_check_beta(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 1), mask = i32(a, 2);
  const unsigned stride = batch_stride(n, count), g = reduce_grid(n);
  const auto* lanes = static_cast<const kh_bicg_lane*>(buf(a, 0));
  k_batch_bicg_step<<<BATCH_GRID(n, count)>>>(n, lanes, mask, d->partials, stride);
  k_batch_finish2<Finish2::CheckBeta><<<BATCH_FINISH(count)>>>(lanes, mask, d->partials, stride, g);
}
void l_fused_project_dot_beta(kh_device* d, int n, const kh_arg* a) {
  const unsigned g = reduce_grid(n);
  int* st = static_cast<int*>(buf(a, 6));
  k_fused_project_dot<<<GRID(n)>>>(n, KF(0), KCF(1), f32(a, 2), f32(a, 3), KCF(4), st, d->partials);
  k_finish<Finish::Beta><<<FINISH>>>(d->partials, g, KF(5), st, nullptr);
}

}  // namespace

const std::vector<KernelEntry>& linalg_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_CSR_SPMV, "csr_spmv", "bbbbb", l_csr_spmv},
      {KH_CSR_SPMV_GUARDED, "csr_spmv_guarded", "bbbbbb", l_csr_spmv_guarded},
      {KH_CSR_RESIDUAL, "csr_residual", "bbbbbb", l_csr_residual},
      {KH_CSR_RESIDUAL_GUARDED, "csr_residual_guarded", "bbbbbbb", l_csr_residual_guarded},
      {KH_KRYLOV_DIAGONAL, "krylov_diagonal", "bbbbbi", l_krylov_diagonal},
      {KH_KRYLOV_MULTIPLY, "krylov_multiply", "bbb", l_krylov_multiply},
      {KH_KRYLOV_MULTIPLY_GUARDED, "krylov_multiply_guarded", "bbbb", l_krylov_multiply_guarded},
      {KH_KRYLOV_SMOOTH, "krylov_smooth", "bbbf", l_krylov_smooth},
      {KH_KRYLOV_SHIFT, "krylov_shift", "bf", l_krylov_shift},
      {KH_KRYLOV_DIRECTION, "krylov_direction", "bbf", l_krylov_direction},
      {KH_KRYLOV_STEP, "krylov_step", "bbbbf", l_krylov_step},
      {KH_KRYLOV_BICG_DIRECTION, "krylov_bicg_direction", "bbbff", l_krylov_bicg_direction},
      {KH_KRYLOV_DIFFERENCE, "krylov_difference", "bbbf", l_krylov_difference},
      {KH_KRYLOV_BICG_STEP, "krylov_bicg_step", "bbbff", l_krylov_bicg_step},
      {KH_AMG_COARSE_SOLVE, "amg_coarse_solve", "bbbbbbbii", l_amg_coarse_solve},
      {KH_AMG_SMOOTH, "amg_smooth", "bbbbbbbb", l_amg_smooth, true},
      {KH_AMG_SMOOTH_ZERO, "amg_smooth_zero", "bbbbbbbb", l_amg_smooth_zero},
      {KH_AMG_RESTRICT, "amg_restrict", "bbbb", l_amg_restrict},
      {KH_AMG_PROLONG, "amg_prolong", "bbb", l_amg_prolong},
      {KH_AMG_GALERKIN, "amg_galerkin", "bbbb", l_amg_galerkin},
      {KH_AMG_M_MATRIX_DIAGONAL, "amg_m_matrix_diagonal", "bbbbbbbb", l_amg_m_matrix_diagonal},
      {KH_AMG_RESTRICT_GUARDED, "amg_restrict_guarded", "bbbbb", l_amg_restrict_guarded},
      {KH_AMG_PROLONG_GUARDED, "amg_prolong_guarded", "bbbb", l_amg_prolong_guarded},
      {KH_AXPY, "axpy", "bbf", l_axpy},
      {KH_XPBY, "xpby", "bbf", l_xpby},
      {KH_AXPBYZ, "axpbyz", "bbbff", l_axpbyz},
      {KH_SCALE_COPY, "scale_copy", "bbf", l_scale_copy},
      {KH_MUL, "mul", "bbb", l_mul},
      {KH_SOLVER_LIMITS, "solver_limits", "bffff", l_solver_limits},
      {KH_SOLVER_RESET, "solver_reset", "bbi", l_solver_reset},
      {KH_SOLVER_PCG_ALPHA, "solver_pcg_alpha", "bb", l_solver_pcg_alpha},
      {KH_SOLVER_PCG_BETA, "solver_pcg_beta", "bbi", l_solver_pcg_beta},
      {KH_SOLVER_BICG_BETA, "solver_bicg_beta", "bb", l_solver_bicg_beta},
      {KH_SOLVER_BICG_ALPHA, "solver_bicg_alpha", "bb", l_solver_bicg_alpha},
      {KH_SOLVER_BICG_S, "solver_bicg_s", "bb", l_solver_bicg_s},
      {KH_SOLVER_BICG_OMEGA, "solver_bicg_omega", "bb", l_solver_bicg_omega},
      {KH_SOLVER_CHECK, "solver_check", "bbi", l_solver_check},
      {KH_SOLVER_STEP, "solver_step", "bbbbbb", l_solver_step},
      {KH_SOLVER_DIRECTION, "solver_direction", "bbbb", l_solver_direction},
      {KH_SOLVER_BICG_DIRECTION, "solver_bicg_direction", "bbbbb", l_solver_bicg_direction},
      {KH_SOLVER_BICG_DIFFERENCE, "solver_bicg_difference", "bbbbb", l_solver_bicg_difference},
      {KH_SOLVER_BICG_STEP, "solver_bicg_step", "bbbbbbbb", l_solver_bicg_step},
      {KH_SOLVER_PROJECT, "solver_project", "bbff", l_solver_project},
      {KH_FUSED_SPMV_DOT_ALPHA, "fused_spmv_dot_alpha", "bbbbbbb", l_fused_spmv_dot_alpha},
      {KH_FUSED_DOT_BICG_BETA, "fused_dot_bicg_beta", "bbbb", l_fused_dot_bicg_beta},
      {KH_FUSED_BICG_DIRECTION, "fused_bicg_direction", "bbbbbbb", l_fused_bicg_direction},
      {KH_FUSED_SPMV_SHADOW_ALPHA, "fused_spmv_shadow_alpha", "bbbbbbbb", l_fused_spmv_shadow_alpha},
      {KH_FUSED_BICG_DIFFERENCE_S, "fused_bicg_difference_s", "bbbbbbb", l_fused_bicg_difference_s},
      {KH_FUSED_SPMV_DOT2_OMEGA, "fused_spmv_dot2_omega", "bbbbbbbb", l_fused_spmv_dot2_omega},
      {KH_FUSED_BICG_STEP_CHECK_BETA, "fused_bicg_step_check_beta", "bbbbbbbbb", l_fused_bicg_step_check_beta},
      {KH_BATCH_BICG_DIRECTION, "batch_bicg_direction", "bii", l_batch_bicg_direction},
      {KH_TO_BF16, "to_bf16", "bb", l_to_bf16},
      {KH_BF16_AMG_SMOOTH, "bf16_amg_smooth", "bbbbbbbb", l_bf16_amg_smooth},
      {KH_BF16_SMOOTH_FROM_ZERO, "bf16_smooth_from_zero", "bbbbbbb", l_bf16_smooth_from_zero},
      {KH_BF16_RESIDUAL_RESTRICT, "bf16_residual_restrict", "bbbbbbbbb", l_bf16_residual_restrict},
      {KH_BF16_PROLONG_SMOOTH, "bf16_prolong_smooth", "bbbbbbbbbb", l_bf16_prolong_smooth},
      {KH_BATCH_SPMV_SHADOW_ALPHA, "batch_spmv_shadow_alpha", "bii", l_batch_spmv_shadow_alpha},
      {KH_BATCH_BICG_DIFFERENCE_S, "batch_bicg_difference_s", "bii", l_batch_bicg_difference_s},
      {KH_BATCH_SPMV_DOT2_OMEGA, "batch_spmv_dot2_omega", "bii", l_batch_spmv_dot2_omega},
      {KH_BATCH_BICG_STEP_CHECK_BETA, "batch_bicg_step_check_beta", "bii", l_batch_bicg_step_check_beta},
      {KH_FUSED_STEP_DOT_CHECK, "fused_step_dot_check", "bbbbbb", l_fused_step_dot_check},
      {KH_FUSED_RESIDUAL_RESTRICT, "fused_residual_restrict", "bbbbbbbbb", l_fused_residual_restrict},
      {KH_FUSED_PROLONG_SMOOTH, "fused_prolong_smooth", "bbbbbbbbbb", l_fused_prolong_smooth},
      {KH_FUSED_SMOOTH_DOT_BETA, "fused_smooth_dot_beta", "bbbbbbbbb", l_fused_smooth_dot_beta},
      {KH_FUSED_SMOOTH_SUM, "fused_smooth_sum", "bbbbbbbbb", l_fused_smooth_sum},
      {KH_FUSED_PROJECT_DOT_BETA, "fused_project_dot_beta", "bbffbbb", l_fused_project_dot_beta},
      {KH_FUSED_SMOOTH_FROM_ZERO, "fused_smooth_from_zero", "bbbbbbb", l_fused_smooth_from_zero},
      {KH_FUSED_COARSE_TAIL, "fused_coarse_tail", "biiiiibbb", l_fused_coarse_tail},
      {KH_FUSED_COARSE_SWEEPS, "fused_coarse_sweeps", "bbbbbbbi", l_fused_coarse_sweeps},
      {KH_FUSED_RESIDUAL_RESTRICT_WIDE, "fused_residual_restrict_wide", "bbbbbbbbb",
       l_fused_residual_restrict_wide},
      {KH_AMG_INNER_ALPHA, "amg_inner_alpha", "bbbi", l_amg_inner_alpha},
      {KH_AMG_INNER_FIRST, "amg_inner_first", "bbbbbbb", l_amg_inner_first},
      {KH_AMG_INNER_CHECK, "amg_inner_check", "bb", l_amg_inner_check},
      {KH_AMG_INNER_BETA, "amg_inner_beta", "bb", l_amg_inner_beta},
      {KH_AMG_INNER_STEP, "amg_inner_step", "bbbb", l_amg_inner_step},
      {KH_AMG_SMOOTH_WIDE, "amg_smooth_wide", "bbbbbbbb", l_amg_smooth_wide},
      {KH_AMG_SMOOTH_ROWWAVE, "amg_smooth_rowwave", "bbbbbbbb", l_amg_smooth_rowwave},
      {KH_COARSE_SWEEPS_RESIDENT, "coarse_sweeps_resident", "bbbbbbbibi", l_coarse_sweeps_resident},
      {KH_CSR_RESIDUAL_ROWWAVE, "csr_residual_rowwave", "bbbbbbb", l_csr_residual_rowwave},
      {KH_PROLONG_SMOOTH_ROWWAVE, "prolong_smooth_rowwave", "bbbbbbbbbb", l_prolong_smooth_rowwave},
      {KH_INNER_FIRST_STEP, "inner_first_step", "bbbbbbbbbbi", l_inner_first_step},
      {KH_INNER_FIRST_UPDATE, "inner_first_update", "bbbbbbb", l_inner_first_update},
      {KH_INNER_BETA_STEP, "inner_beta_step", "bbbbbbbbi", l_inner_beta_step},
      {KH_INNER_SECOND_STEP, "inner_second_step", "bbbbbbbbbi", l_inner_second_step},
  };
  return table;
}

}  // namespace kh

// Row partition of KH_COARSE_SWEEPS_RESIDENT: the entries per thread of
// phase 1 (8, or 16 above 8 x KH_RESIDENT_THREADS entries) and the rows of
// each thread in phase 2. Rows longer than 16 entries take a thread each,
// the first threads, in row order; the other rows are cut greedily, in order,
// into chunks of at most `target` entries (a longer row alone; the smallest
// target from their entries per remaining thread up that needs no more
// threads), starting at the next wavefront.
extern "C" int kh_resident_partition(const int32_t* row_ptr, const int32_t* cols, int32_t rows, int32_t* table,
                                     int32_t* cap_out) {
  constexpr int T = KH_RESIDENT_THREADS;
  constexpr int kLong = 16;
  (void)cols;
  if (!row_ptr || !table || !cap_out || rows < 1 || rows > KH_RESIDENT_MAX_ROWS || row_ptr[0] != 0) return 1;
  for (int32_t r = 0; r < rows; ++r)
    if (row_ptr[r + 1] <= row_ptr[r]) return 1;
  const long nnz = row_ptr[rows];
  if (nnz > 16L * T) return 1;
  struct Chunk {
    int32_t first, last;
  };
  std::vector<Chunk> chunks;
  long short_entries = 0;
  for (int32_t r = 0; r < rows; ++r) {
    const int32_t len = row_ptr[r + 1] - row_ptr[r];
    if (len > kLong) chunks.push_back({r, r + 1});
    else short_entries += len;
  }
  const std::size_t lead = (chunks.size() + 63) / 64 * 64;
  if (lead > std::size_t(T) || (lead == std::size_t(T) && short_entries > 0)) return 1;
  const long room = T - long(lead);
  chunks.resize(lead, Chunk{rows, rows});
  std::vector<Chunk> shorts;
  bool fits = short_entries == 0;
  for (long target = std::max(1L, (short_entries + room - 1) / std::max(1L, room)); !fits; ++target) {
    shorts.clear();
    long open = 0;
    for (int32_t r = 0; r < rows; ++r) {
      const int32_t len = row_ptr[r + 1] - row_ptr[r];
      if (len > kLong) {
        open = 0;
      } else if (open == 0 || open + len > target) {
        shorts.push_back({r, r + 1});
        open = len;
      } else {
        shorts.back().last = r + 1;
        open += len;
      }
    }
    fits = long(shorts.size()) <= room;
  }
  chunks.insert(chunks.end(), shorts.begin(), shorts.end());
  for (int t = 0; t < T; ++t) {
    const Chunk c = t < int(chunks.size()) ? chunks[t] : Chunk{rows, rows};
    table[2 * t] = c.first;
    table[2 * t + 1] = c.last;
  }
  *cap_out = nnz <= 8L * T ? 8 : 16;
  return 0;
}
