// ORACLE — test infrastructure only (see oracle.hpp). Serial restatement of the
// reference's linear-algebra kernel bodies and its serial reductions.
#include <cmath>

#include "oracle.hpp"

namespace oracle {

// kernels/prelude.h:82-87: advance to the next float only when round-to-nearest
// rounded the exact sum down.
float kadd_up(float a, float b) {
  float s = a + b;
  float part = s - a;
  float err = (a - (s - part)) + (b - part);
  return err > 0.0f ? std::nextafter(s, INFINITY) : s;
}
// kernels/prelude.h:92-100
UpperSum kupper_add(UpperSum a, float b) {
  float s = a.sum + b;
  float part = s - a.sum;
  float err = (a.sum - (s - part)) + (b - part);
  return {s, kadd_up(a.error, err)};
}
float kupper_value(UpperSum a) { return kadd_up(a.sum, a.error); }

// runtime/cpu/cpu_device.cpp:37-43
double sum(const float* x, std::size_t n) {
  double s = 0;
  for (std::size_t i = 0; i < n; ++i) s += x[i];
  return s;
}
// runtime/cpu/cpu_device.cpp:44-51
double dot(const float* x, const float* y, std::size_t n) {
  double s = 0;
  for (std::size_t i = 0; i < n; ++i) s += double(x[i]) * double(y[i]);
  return s;
}
// runtime/cpu/cpu_device.cpp:52-62 (NaN propagates)
double max_abs(const float* x, std::size_t n) {
  double m = 0;
  for (std::size_t i = 0; i < n; ++i) {
    const double v = std::fabs(x[i]);
    if (v > m || std::isnan(v)) m = v;
  }
  return m;
}
// runtime/cpu/cpu_device.cpp:83-86
void store_pair(float* p, double value) {
  p[0] = static_cast<float>(value);
  p[1] = std::isfinite(p[0]) ? static_cast<float>(value - double(p[0])) : 0.0f;
}
// runtime/cpu/cpu_device.cpp:72-89
void reduce_into(const float* x, const float* y, const float* u, const float* v,
                 std::size_t n, float* dst, Reduction mode) {
  double a = 0, b = 0;
  for (std::size_t i = 0; i < n; ++i) {
    a += mode == Reduction::Sum ? double(x[i]) : double(x[i]) * double(y[i]);
    if (mode == Reduction::Dot2) b += double(u[i]) * double(v[i]);
  }
  store_pair(dst, a);
  if (mode == Reduction::Dot2) store_pair(dst + 2, b);
}

namespace kern {

// kernels/linalg.kern.h:5-16 — Kahan-compensated row sum.
void csr_spmv(int n, const int* rows, const int* cols, const float* values,
              const float* x, float* y) {
  for (int i = 0; i < n; ++i) {
    float s = 0.0f, e = 0.0f;
    for (int j = rows[i]; j < rows[i + 1]; ++j) {
      float t = values[j] * x[cols[j]] - e;
      float nx = s + t;
      e = (nx - s) - t;
      s = nx;
    }
    y[i] = s;
  }
}
// kernels/linalg.kern.h:111-114
void csr_spmv_guarded(int n, const int* rows, const int* cols, const float* values,
                      const float* x, float* y, const int* status) {
  if (status[0] == 0) csr_spmv(n, rows, cols, values, x, y);
}
// kernels/linalg.kern.h:21-30 — Dekker split product error.
float csr_product_error(float a, float b, float product) {
  if (std::fabs(a) > 1.0e30f || std::fabs(b) > 1.0e30f) return 0.0f;
  float ca = 4097.0f * a, cb = 4097.0f * b;
  float ah = ca - (ca - a), bh = cb - (cb - b);
  float al = a - ah, bl = b - bh;
  return ((ah * bh - product) + ah * bl + al * bh) + al * bl;
}
// kernels/linalg.kern.h:32-45 — Neumaier sum of b - sum a*x plus product errors.
float csr_row_residual(int i, const int* rows, const int* cols, const float* values,
                       const float* x, const float* b) {
  float s = b[i], e = 0.0f;
  for (int j = rows[i]; j < rows[i + 1]; ++j) {
    float prod = values[j] * x[cols[j]];
    float t = -prod;
    float nx = s + t;
    e += std::fabs(s) >= std::fabs(t) ? (s - nx) + t : (t - nx) + s;
    e -= csr_product_error(values[j], x[cols[j]], prod);
    s = nx;
  }
  return s + e;
}
// kernels/linalg.kern.h:47-50
void csr_residual(int n, const int* rows, const int* cols, const float* values,
                  const float* x, const float* b, float* r) {
  for (int i = 0; i < n; ++i) r[i] = csr_row_residual(i, rows, cols, values, x, b);
}
// kernels/linalg.kern.h:115-118
void csr_residual_guarded(int n, const int* rows, const int* cols, const float* values,
                          const float* x, const float* b, float* r, const int* status) {
  if (status[0] == 0) csr_residual(n, rows, cols, values, x, b, r);
}
// kernels/linalg.kern.h:52-70 — (l1) inverse diagonal and a validity code.
void krylov_diagonal(int n, const int* rows, const int* diag, const float* values,
                     float* inverse, float* invalid, int use_l1) {
  for (int i = 0; i < n; ++i) {
    float d = values[diag[i]];
    float off = 0.0f, bad = 0.0f;
    for (int j = rows[i]; j < rows[i + 1]; ++j) {
      float v = values[j];
      if (v - v != 0.0f) bad = 1.0f;
      if (j != diag[i]) off += std::fabs(v);
    }
    float inv = d != 0.0f ? 1.0f / d : 0.0f;
    float l1 = std::fmax(std::fabs(d), 0.5f * (std::fabs(d) + off));
    inverse[i] = use_l1 != 0 ? (l1 != 0.0f ? 1.0f / l1 : 0.0f) : inv;
    invalid[i] = (bad != 0.0f || inv - inv != 0.0f || l1 - l1 != 0.0f)
                     ? 2.0f
                     : (use_l1 != 0 && d < 0.0f ? 3.0f : (d == 0.0f ? 1.0f : 0.0f));
  }
}
// kernels/linalg.kern.h:72-75
void krylov_multiply(int n, const float* r, const float* inverse, float* z) {
  for (int i = 0; i < n; ++i) z[i] = r[i] * inverse[i];
}
// kernels/linalg.kern.h:119-122
void krylov_multiply_guarded(int n, const float* r, const float* inverse, float* z,
                             const int* status) {
  if (status[0] == 0) krylov_multiply(n, r, inverse, z);
}
// kernels/linalg.kern.h:126-147 — one owner solves the whole <=24-row system.
void amg_coarse_solve(int n, const int* rows, const int* cols, const float* values,
                      float* x, const float* b, const float* inverse, const int* status,
                      int count, int sweeps) {
  if (n < 1 || status[0] != 0) return;
  float cur[24], r[24];
  for (int row = 0; row < count; ++row) cur[row] = 0.0f;
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    for (int row = 0; row < count; ++row) r[row] = csr_row_residual(row, rows, cols, values, cur, b);
    for (int row = 0; row < count; ++row) cur[row] += (2.0f / 3.0f) * inverse[row] * r[row];
  }
  for (int row = 0; row < count; ++row) x[row] = cur[row];
}
// kernels/linalg.kern.h:150-156 — l1-Jacobi sweep with the compensated residual.
void amg_smooth(int n, const int* rows, const int* cols, const float* values,
                const float* x, const float* b, const float* inverse, float* next,
                const int* status) {
  if (status[0] != 0) return;
  for (int i = 0; i < n; ++i) {
    float r = csr_row_residual(i, rows, cols, values, x, b);
    next[i] = x[i] + (2.0f / 3.0f) * inverse[i] * r;
  }
}
// kernels/linalg.kern.h:162-170 — first sweep from x = 0.
void amg_smooth_zero(int n, const int* rows, const int* cols, const float* values,
                     const float* x, const float* b, const float* inverse, float* next,
                     const int* status) {
  if (status[0] != 0) return;
  for (int i = 0; i < n; ++i) {
    if (b[i] - b[i] != 0.0f) {
      float r = csr_row_residual(i, rows, cols, values, x, b);
      next[i] = x[i] + (2.0f / 3.0f) * inverse[i] * r;
    } else {
      next[i] = 0.0f + (2.0f / 3.0f) * inverse[i] * (b[i] + 0.0f);
    }
  }
}
// kernels/amg.kern.h:5-16 — R = P^T, Kahan sum over aggregate members.
void amg_restrict(int n, const int* offsets, const int* members, const float* fine,
                  float* coarse) {
  for (int i = 0; i < n; ++i) {
    float s = 0.0f, e = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      float t = fine[members[j]] - e;
      float nx = s + t;
      e = (nx - s) - t;
      s = nx;
    }
    coarse[i] = s;
  }
}
// kernels/amg.kern.h:84-87
void amg_restrict_guarded(int n, const int* offsets, const int* members,
                          const float* fine, float* coarse, const int* status) {
  if (status[0] == 0) amg_restrict(n, offsets, members, fine, coarse);
}
// kernels/amg.kern.h:18-21 — piecewise-constant interpolation.
void amg_prolong(int n, const int* aggregate, const float* coarse, float* fine) {
  for (int i = 0; i < n; ++i)
    if (aggregate[i] >= 0) fine[i] += coarse[aggregate[i]];
}
// kernels/amg.kern.h:88-91
void amg_prolong_guarded(int n, const int* aggregate, const float* coarse, float* fine,
                         const int* status) {
  if (status[0] == 0) amg_prolong(n, aggregate, coarse, fine);
}
// kernels/amg.kern.h:27-38 — Neumaier sum over each coarse entry's fine entries.
void amg_galerkin(int n, const int* offsets, const int* entries, const float* fine,
                  float* coarse) {
  for (int i = 0; i < n; ++i) {
    float s = 0.0f, e = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      float t = fine[entries[j]];
      float nx = s + t;
      e += std::fabs(s) >= std::fabs(t) ? (s - nx) + t : (t - nx) + s;
      s = nx;
    }
    coarse[i] = s + e;
  }
}
// kernels/amg.kern.h:44-58 — M-matrix diagonal from surpluses and coarse edges.
void amg_m_matrix_diagonal(int n, const int* offsets, const int* members,
                           const int* fine_rows, const int* fine_diag, const float* fine,
                           const int* rows, const int* diag, float* coarse) {
  for (int i = 0; i < n; ++i) {
    UpperSum d{0.0f, 0.0f};
    for (int m = offsets[i]; m < offsets[i + 1]; ++m) {
      int row = members[m];
      UpperSum surplus{fine[fine_diag[row]], 0.0f};
      for (int j = fine_rows[row]; j < fine_rows[row + 1]; ++j)
        if (j != fine_diag[row]) surplus = kupper_add(surplus, fine[j]);
      d = kupper_add(d, kupper_value(surplus));
    }
    for (int j = rows[i]; j < rows[i + 1]; ++j)
      if (j != diag[i]) d = kupper_add(d, -coarse[j]);
    coarse[diag[i]] = kupper_value(d);
  }
}
// kernels/blas.kern.h:5 and :14
void axpy(int n, float* y, const float* x, float a) {
  for (int i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}
void scale_copy(int n, float* y, const float* x, float a) {
  for (int i = 0; i < n; ++i) y[i] = a * x[i];
}

// ------------------------------------------------------- scalar pair arithmetic
// kernels/scalars.kern.h:3-42: normalized (hi, lo) float pairs.
namespace {
struct Pair { float hi, lo; };
Pair kp(float hi, float lo) { return {hi, lo}; }
Pair load(const float* p, int off) { return {p[off], p[off + 1]}; }
void store(float* p, int off, Pair v) { p[off] = v.hi; p[off + 1] = v.lo; }
Pair add(Pair a, Pair b) {
  float s = a.hi + b.hi;
  if (s - s != 0.0f) return kp(s, 0.0f);
  float v = s - a.hi;
  float e = (a.hi - (s - v)) + (b.hi - v);
  e = e + a.lo + b.lo;
  float h = s + e;
  return kp(h, e - (h - s));
}
Pair neg(Pair a) { return kp(-a.hi, -a.lo); }
Pair mul(Pair a, Pair b) {
  float h = a.hi * b.hi;
  float l = std::fma(a.hi, b.hi, -h) + a.hi * b.lo + a.lo * b.hi;
  return add(kp(h, 0.0f), kp(l, 0.0f));
}
Pair div(Pair a, Pair b) {
  float q = a.hi / b.hi;
  Pair r = add(a, neg(mul(kp(q, 0.0f), b)));
  return add(kp(q, 0.0f), kp((r.hi + r.lo) / b.hi, 0.0f));
}
float to_float(Pair a) { return a.hi + a.lo; }
bool positive(Pair a) { return a.hi > 0.0f || (a.hi == 0.0f && a.lo > 0.0f); }
bool finite(Pair a) { return a.hi - a.hi == 0.0f && a.lo - a.lo == 0.0f; }
bool zero(Pair a) { return a.hi == 0.0f && a.lo == 0.0f; }
bool le(Pair a, Pair b) { return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo); }
bool coefficient(Pair a) {
  float v = to_float(a);
  return finite(a) && v - v == 0.0f && (v != 0.0f || zero(a));
}
}  // namespace

// kernels/scalars.kern.h:51-55
void solver_limits(int n, float* s, float th, float tl, float dh, float dl) {
  for (int i = 0; i < n; ++i)
    s[i] = i == 20 ? th : (i == 21 ? tl : (i == 22 ? dh : (i == 23 ? dl : 0.0f)));
}
// kernels/scalars.kern.h:56-62
void solver_reset(float* s, int* status, int iterations) {
  status[0] = 0; status[1] = iterations; status[2] = 0;
  s[8] = 1.0f; s[9] = 0.0f; s[12] = 1.0f; s[13] = 0.0f; s[14] = 1.0f; s[15] = 0.0f;
}
// kernels/scalars.kern.h:63-71
void solver_pcg_alpha(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rho = load(s, 0), pq = load(s, 2);
  Pair alpha = div(rho, pq);
  if (!positive(rho) || !positive(pq) || !coefficient(alpha) || zero(alpha)) status[0] = 2;
  else store(s, 8, alpha);
}
// kernels/scalars.kern.h:72-80
void solver_pcg_beta(float* s, int* status, int flexible) {
  if (status[0] != 0) return;
  Pair next = load(s, 4);
  Pair numer = flexible != 0 ? neg(mul(load(s, 8), load(s, 6))) : next;
  Pair beta = div(numer, load(s, 0));
  if (!positive(next) || !coefficient(beta) || (zero(beta) && !zero(numer))) status[0] = 2;
  else { store(s, 10, beta); store(s, 0, next); }
}
// kernels/scalars.kern.h:81-89
void solver_bicg_beta(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rho = load(s, 0), omega = load(s, 12);
  Pair beta = mul(div(rho, load(s, 14)), div(load(s, 8), omega));
  if (zero(rho) || !finite(rho) || zero(omega) || !coefficient(omega) || !coefficient(beta) ||
      zero(beta))
    status[0] = 2;
  else { store(s, 10, beta); status[2] = 0; }
}
// kernels/scalars.kern.h:90-97
void solver_bicg_alpha(float* s, int* status) {
  if (status[0] != 0) return;
  Pair rv = load(s, 2);
  Pair alpha = div(load(s, 0), rv);
  if (zero(rv) || !finite(rv) || !coefficient(alpha) || zero(alpha)) status[0] = 2;
  else store(s, 8, alpha);
}
// kernels/scalars.kern.h:98-103
void solver_bicg_s(const float* s, int* status) {
  if (status[0] != 0) return;
  ++status[1];
  status[2] = le(load(s, 16), load(s, 20)) ? 1 : 0;
}
// kernels/scalars.kern.h:104-112
void solver_bicg_omega(float* s, int* status) {
  if (status[0] != 0) return;
  if (status[2] != 0) { store(s, 12, kp(0.0f, 0.0f)); return; }
  Pair tt = load(s, 2), ts = load(s, 4);
  Pair omega = div(ts, tt);
  if (!positive(tt) || !finite(tt) || zero(omega) || !coefficient(omega)) status[0] = 2;
  else store(s, 12, omega);
}
// kernels/scalars.kern.h:113-121
void solver_check(float* s, int* status, int pcg) {
  if (status[0] != 0) return;
  if (pcg != 0) ++status[1];
  Pair rr = load(s, 18);
  if (!finite(rr) || le(rr, load(s, 20)) || !le(rr, load(s, 22)) || status[1] % 50 == 0)
    status[0] = 1;
  if (pcg == 0) store(s, 14, load(s, 0));
}
// kernels/scalars.kern.h:122-127
void solver_step(int n, float* x, float* r, const float* p, const float* q,
                 const float* s, const int* status) {
  if (status[0] != 0) return;
  float alpha = to_float(load(s, 8));
  for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
}
// kernels/scalars.kern.h:128-131
void solver_direction(int n, const float* z, float* p, const float* s, const int* status) {
  if (status[0] != 0) return;
  float beta = to_float(load(s, 10));
  for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
}
// kernels/scalars.kern.h:132-137
void solver_bicg_direction(int n, const float* r, const float* v, float* p, const float* s,
                           const int* status) {
  if (status[0] != 0) return;
  float beta = to_float(load(s, 10)), omega = to_float(load(s, 12));
  for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
// kernels/scalars.kern.h:138-141
void solver_bicg_difference(int n, const float* r, const float* q, float* sv,
                            const float* s, const int* status) {
  if (status[0] != 0) return;
  float alpha = to_float(load(s, 8));
  for (int i = 0; i < n; ++i) sv[i] = r[i] - alpha * q[i];
}
// kernels/scalars.kern.h:142-149
void solver_bicg_step(int n, float* x, float* r, const float* z, const float* zs,
                      const float* sv, const float* t, const float* s, const int* status) {
  if (status[0] != 0) return;
  float alpha = to_float(load(s, 8)), omega = to_float(load(s, 12));
  for (int i = 0; i < n; ++i) {
    if (status[2] != 0) { x[i] += alpha * z[i]; r[i] = sv[i]; }
    else { x[i] += alpha * z[i] + omega * zs[i]; r[i] = sv[i] - omega * t[i]; }
  }
}
// kernels/scalars.kern.h:157-166 — the K-cycle's inner step length. `first`
// makes the level's status follow the parent's (a frozen parent freezes the
// level with phase 2); a failed first step is phase 2 (keep the unscaled
// first correction), a failed second one phase 1 (keep the first step).
void amg_inner_alpha(float* s, int* status, const int* parent, int first) {
  if (first != 0) status[0] = parent[0] == 0 ? 0 : 2;
  if (status[0] != 0) return;
  Pair denom = load(s, 0), alpha = div(load(s, 2), denom);
  if (!positive(denom) || !coefficient(alpha) || (zero(alpha) && !zero(load(s, 2))))
    status[0] = first != 0 ? 2 : 1;
  else store(s, 8, alpha);
}
// kernels/scalars.kern.h:167-172 — unguarded: x = alpha p and t = b - alpha q
// while running, else x = p and t = 0.
void amg_inner_first(int n, const float* b, const float* p, const float* q, float* x, float* t,
                     const float* s, const int* status) {
  const float alpha = to_float(load(s, 8));
  for (int i = 0; i < n; ++i) {
    x[i] = status[0] == 0 ? alpha * p[i] : p[i];
    t[i] = status[0] == 0 ? b[i] - alpha * q[i] : 0.0f;
  }
}
// kernels/scalars.kern.h:173-176 — adaptive exit: t.t <= b.b / 16.
void amg_inner_check(const float* s, int* status) {
  if (status[0] != 0) return;
  if (le(load(s, 4), mul(kp(0.0625f, 0.0f), load(s, 6)))) status[0] = 1;
}
// kernels/scalars.kern.h:177-181
void amg_inner_beta(float* s, int* status) {
  if (status[0] != 0) return;
  Pair beta = div(neg(load(s, 4)), load(s, 0));
  if (!coefficient(beta) || (zero(beta) && !zero(load(s, 4)))) status[0] = 1;
  else store(s, 10, beta);
}
// kernels/scalars.kern.h:182-186
void amg_inner_step(int n, float* x, const float* p, const float* s, const int* status) {
  if (status[0] != 0) return;
  const float alpha = to_float(load(s, 8));
  for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
}
// kernels/scalars.kern.h:150-152
void solver_project(int n, float* x, const float* sum, float count_hi, float count_lo) {
  for (int i = 0; i < n; ++i) x[i] -= to_float(div(load(sum, 0), kp(count_hi, count_lo)));
}

}  // namespace kern
}  // namespace oracle
