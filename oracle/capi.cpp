// ORACLE — test infrastructure only (see oracle.hpp). A plain C interface to
// the oracle for the Python tests (ctypes). Every entry point returns 0 on
// success; on failure it returns nonzero and oracle_last_error() names why.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_error;
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const NonfiniteError& e) {
    g_error = e.what();
    return 2;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}
struct SolverHandle {
  Solver solver;
  CsrDevice matrix;
  SolverHandle(SolverKind k, const CsrHost& a, const AmgOptions& o) : solver(k, a, o), matrix(upload(a)) {}
};
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

// Checks the window in which the device's one-FMA product error replaces the
// reference's Dekker split (kernels/linalg.kern.h:21-30): over `count` seeded
// pairs (half with uniformly random bits, half with exponents in [-60, 60)),
// returns how many inside the window 2^-100 <= |a*b| <= 2^100, |a|,|b| <= 1e30
// give a Dekker result whose bits differ from fma(a, b, -a*b); *inside receives
// how many pairs fell in the window.
long oracle_product_error_window(long count, unsigned long long seed, long* inside) {
  std::uint64_t state = seed;
  auto next = [&] {  // splitmix64
    std::uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  long differ = 0, in = 0;
  for (long k = 0; k < count; ++k) {
    std::uint64_t r = next();
    std::uint32_t ua = std::uint32_t(r), ub = std::uint32_t(r >> 32);
    float a, b;
    std::memcpy(&a, &ua, 4);
    std::memcpy(&b, &ub, 4);
    if (k & 1) {
      int ea = 0, eb = 0;
      const std::uint64_t e = next();
      a = std::ldexp(std::frexp(a, &ea), int(e % 120) - 60);
      b = std::ldexp(std::frexp(b, &eb), int((e >> 32) % 120) - 60);
    }
    if (!std::isfinite(a) || !std::isfinite(b)) continue;
    const float p = a * b;
    const float m = std::fabs(p);
    if (!(m >= 0x1p-100f && m <= 0x1p100f && std::fabs(a) <= 1.0e30f && std::fabs(b) <= 1.0e30f)) continue;
    ++in;
    const float d = kern::csr_product_error(a, b, p), f = std::fma(a, b, -p);
    std::uint32_t bd, bf;
    std::memcpy(&bd, &d, 4);
    std::memcpy(&bf, &f, 4);
    if (bd != bf) ++differ;
  }
  *inside = in;
  return differ;
}

// ---- reductions (runtime/cpu/cpu_device.cpp:37-89)
double oracle_sum(const float* x, long n) { return sum(x, std::size_t(n)); }
double oracle_dot(const float* x, const float* y, long n) { return dot(x, y, std::size_t(n)); }
double oracle_max_abs(const float* x, long n) { return max_abs(x, std::size_t(n)); }
void oracle_reduce_into(int mode, const float* x, const float* y, const float* u, const float* v,
                        long n, float* dst) {
  reduce_into(x, y, u, v, std::size_t(n), dst, mode == 0 ? Reduction::Sum
                                                   : mode == 1 ? Reduction::Dot : Reduction::Dot2);
}

// ---- kernels, one call per launch over [0, n)
void oracle_csr_spmv(int n, const int* rows, const int* cols, const float* values,
                     const float* x, float* y) { kern::csr_spmv(n, rows, cols, values, x, y); }
void oracle_csr_residual(int n, const int* rows, const int* cols, const float* values,
                         const float* x, const float* b, float* r) {
  kern::csr_residual(n, rows, cols, values, x, b, r);
}
void oracle_krylov_diagonal(int n, const int* rows, const int* diag, const float* values,
                            float* inverse, float* invalid, int use_l1) {
  kern::krylov_diagonal(n, rows, diag, values, inverse, invalid, use_l1);
}
void oracle_amg_smooth(int n, const int* rows, const int* cols, const float* values,
                       const float* x, const float* b, const float* inverse, float* next,
                       const int* status) {
  kern::amg_smooth(n, rows, cols, values, x, b, inverse, next, status);
}
void oracle_amg_smooth_zero(int n, const int* rows, const int* cols, const float* values,
                            const float* x, const float* b, const float* inverse, float* next,
                            const int* status) {
  kern::amg_smooth_zero(n, rows, cols, values, x, b, inverse, next, status);
}
void oracle_amg_coarse_solve(const int* rows, const int* cols, const float* values, float* x,
                             const float* b, const float* inverse, const int* status, int count,
                             int sweeps) {
  kern::amg_coarse_solve(1, rows, cols, values, x, b, inverse, status, count, sweeps);
}
void oracle_amg_restrict(int n, const int* offsets, const int* members, const float* fine,
                         float* coarse) { kern::amg_restrict(n, offsets, members, fine, coarse); }
void oracle_amg_prolong(int n, const int* aggregate, const float* coarse, float* fine) {
  kern::amg_prolong(n, aggregate, coarse, fine);
}
void oracle_amg_galerkin(int n, const int* offsets, const int* entries, const float* fine,
                         float* coarse) { kern::amg_galerkin(n, offsets, entries, fine, coarse); }
void oracle_amg_m_matrix_diagonal(int n, const int* offsets, const int* members,
                                  const int* fine_rows, const int* fine_diag, const float* fine,
                                  const int* rows, const int* diag, float* coarse) {
  kern::amg_m_matrix_diagonal(n, offsets, members, fine_rows, fine_diag, fine, rows, diag, coarse);
}
// The K-cycle's inner kernels (kernels/scalars.kern.h:157-186): which 0 alpha
// (flag = first), 1 first (n entries), 2 check, 3 beta, 4 step (n entries).
void oracle_amg_inner(int which, int n, const float* b, const float* p, const float* q, float* x,
                      float* t, float* scalars, int* status, const int* parent, int flag) {
  switch (which) {
    case 0: kern::amg_inner_alpha(scalars, status, parent, flag); break;
    case 1: kern::amg_inner_first(n, b, p, q, x, t, scalars, status); break;
    case 2: kern::amg_inner_check(scalars, status); break;
    case 3: kern::amg_inner_beta(scalars, status); break;
    case 4: kern::amg_inner_step(n, x, p, scalars, status); break;
  }
}
void oracle_solver_scalar(int which, float* scalars, int* status, int flag) {
  switch (which) {
    case 0: kern::solver_reset(scalars, status, flag); break;
    case 1: kern::solver_pcg_alpha(scalars, status); break;
    case 2: kern::solver_pcg_beta(scalars, status, flag); break;
    case 3: kern::solver_bicg_beta(scalars, status); break;
    case 4: kern::solver_bicg_alpha(scalars, status); break;
    case 5: kern::solver_bicg_s(scalars, status); break;
    case 6: kern::solver_bicg_omega(scalars, status); break;
    case 7: kern::solver_check(scalars, status, flag); break;
  }
}
void oracle_solver_step(int n, float* x, float* r, const float* p, const float* q,
                        const float* scalars, const int* status) {
  kern::solver_step(n, x, r, p, q, scalars, status);
}
void oracle_solver_direction(int n, const float* z, float* p, const float* scalars,
                             const int* status) { kern::solver_direction(n, z, p, scalars, status); }
void oracle_solver_project(int n, float* x, const float* sum_pair, float hi, float lo) {
  kern::solver_project(n, x, sum_pair, hi, lo);
}
float oracle_kadd_up(float a, float b) { return kadd_up(a, b); }

// ---- host CSR objects
void* oracle_poisson(int nx, int ny, int nz, double anisotropy, int random, int neumann,
                     double wind) {
  return new CsrHost(poisson(nx, ny, nz, anisotropy, random != 0, neumann != 0, wind));
}
void* oracle_csr_from_arrays(int n, const int* row_ptr, const int* col, const int* diag,
                             const double* val) {
  auto* a = new CsrHost;
  a->n_rows = n;
  a->row_ptr.assign(row_ptr, row_ptr + n + 1);
  const int nnz = row_ptr[n];
  a->col.assign(col, col + nnz);
  a->diag.assign(diag, diag + n);
  a->val.assign(val, val + nnz);
  return a;
}
void oracle_csr_free(void* h) { delete static_cast<CsrHost*>(h); }
int oracle_csr_rows(void* h) { return static_cast<CsrHost*>(h)->n_rows; }
int oracle_csr_nnz(void* h) { return static_cast<CsrHost*>(h)->nnz(); }
void oracle_csr_copy(void* h, int* row_ptr, int* col, int* diag, double* val) {
  auto* a = static_cast<CsrHost*>(h);
  std::memcpy(row_ptr, a->row_ptr.data(), a->row_ptr.size() * sizeof(int));
  std::memcpy(col, a->col.data(), a->col.size() * sizeof(int));
  std::memcpy(diag, a->diag.data(), a->diag.size() * sizeof(int));
  std::memcpy(val, a->val.data(), a->val.size() * sizeof(double));
}

// ---- one level of host aggregation (linalg/amg.cpp:160-174)
int oracle_aggregate(void* csr, const int* options, double threshold, void** out) {
  return guard([&] {
    AmgOptions o;
    o.aggregation_passes = options[5];
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    o.strength_threshold = threshold;
    *out = new Coarsening(aggregate(*static_cast<CsrHost*>(csr), o));
  });
}
long oracle_coarsening_array(void* c, int which, int* out) {
  const auto* g = static_cast<Coarsening*>(c);
  const std::vector<Index>* v = which == 0 ? &g->aggregate : which == 1 ? &g->member_offsets
      : which == 2 ? &g->members : which == 3 ? &g->entry_offsets : which == 4 ? &g->entries
      : which == 5 ? &g->matrix.row_ptr : which == 6 ? &g->matrix.col : &g->matrix.diag;
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(int));
  return long(v->size());
}
long oracle_coarsening_values(void* c, double* out) {
  const auto* g = static_cast<Coarsening*>(c);
  if (out) std::memcpy(out, g->matrix.val.data(), g->matrix.val.size() * sizeof(double));
  return long(g->matrix.val.size());
}
void oracle_coarsening_free(void* c) { delete static_cast<Coarsening*>(c); }

// ---- solvers (linalg/amg.cpp:345-382, linalg/solver.cpp:75-176)
// options[11]: pre, post, coarse_sweeps, coarse_rows, max_levels, passes, nullspace,
// structure, cycle (0 V, 1 K), K-cycle depth (0: every level), fewest rows of a
// K-cycle level (0: no bound)
int oracle_solver_create(int kind, void* csr, const int* options, double threshold, void** out) {
  return guard([&] {
    AmgOptions o;
    o.pre_sweeps = options[0]; o.post_sweeps = options[1]; o.coarse_sweeps = options[2];
    o.coarse_rows = options[3]; o.max_levels = options[4]; o.aggregation_passes = options[5];
    o.nullspace = options[6] ? NullSpace::Constant : NullSpace::None;
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    ORACLE_REQUIRE((options[8] == 0 || options[8] == 1) && options[9] >= 0 && options[10] >= 0,
                   "AMG: invalid options");
    o.cycle = options[8] ? AmgCycle::K : AmgCycle::V;
    o.k_depth = options[9];
    o.k_min_rows = options[10];
    o.strength_threshold = threshold;
    *out = new SolverHandle(static_cast<SolverKind>(kind), *static_cast<CsrHost*>(csr), o);
  });
}
void oracle_solver_free(void* h) { delete static_cast<SolverHandle*>(h); }
// Replaces the operator's float values (same pattern) and refreshes.
int oracle_solver_refresh(void* h, const float* values) {
  return guard([&] {
    auto* s = static_cast<SolverHandle*>(h);
    std::memcpy(s->matrix.val.data(), values, s->matrix.val.size() * sizeof(float));
    s->solver.refresh(s->matrix);
  });
}
// control: rel_tol, abs_tol; ints: max_iter, interval. stats: iterations, initial, final, converged
int oracle_solver_solve(void* h, float* x, const float* b, double rel_tol, double abs_tol,
                        int max_iter, int interval, double* stats) {
  return guard([&] {
    auto* s = static_cast<SolverHandle*>(h);
    const auto n = std::size_t(s->matrix.n_rows);
    std::vector<float> xv(x, x + n), bv(b, b + n);
    SolveControl c{rel_tol, abs_tol, max_iter, interval};
    auto st = s->solver.solve(s->matrix, xv, bv, c);
    std::memcpy(x, xv.data(), n * sizeof(float));
    stats[0] = st.iterations; stats[1] = st.initial_residual; stats[2] = st.final_residual;
    stats[3] = st.converged ? 1 : 0;
  });
}
int oracle_solver_describe(void* h, char* out, int capacity) {
  const auto text = static_cast<SolverHandle*>(h)->solver.describe();
  std::strncpy(out, text.c_str(), std::size_t(capacity));
  if (capacity > 0) out[capacity - 1] = 0;
  return int(text.size());
}
int oracle_solver_levels(void* h, int* rows, int* nnz, int capacity) {
  const auto levels = static_cast<SolverHandle*>(h)->solver.levels();
  for (std::size_t i = 0; i < levels.size() && int(i) < capacity; ++i) {
    rows[i] = levels[i].rows;
    nnz[i] = levels[i].nnz;
  }
  return int(levels.size());
}
// which: 0 aggregate, 1 member_offsets, 2 members, 3 entry_offsets, 4 entries.
// Returns the length; copies when out is non-null.
long oracle_solver_transfer(void* h, int level, int which, int* out) {
  const auto* t = static_cast<SolverHandle*>(h)->solver.transfer(std::size_t(level));
  if (!t) return -1;
  const std::vector<Index>* v = which == 0 ? &t->aggregate : which == 1 ? &t->member_offsets
      : which == 2 ? &t->members : which == 3 ? &t->entry_offsets : &t->entries;
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(int));
  return long(v->size());
}
// One preconditioner application to r (n floats) into z.
int oracle_solver_apply(void* h, const float* r, float* z) {
  return guard([&] {
    auto* s = static_cast<SolverHandle*>(h);
    const std::size_t n = std::size_t(s->matrix.n_rows);
    std::vector<float> rv(r, r + n), zv(n, 0.0f);
    s->solver.apply(s->matrix, rv, zv);
    std::memcpy(z, zv.data(), n * sizeof(float));
  });
}
// A level's arrays: which 0 row_ptr, 1 col, 2 diag (level >= 1), 3 the
// l1-Jacobi inverse (any level). Returns the length; copies when out is
// non-null; -1 when absent.
long oracle_solver_level_array(void* h, int level, int which, void* out) {
  const auto& solver = static_cast<SolverHandle*>(h)->solver;
  if (which == 3) {
    const auto* l1 = solver.level_l1(std::size_t(level));
    if (!l1) return -1;
    if (out) std::memcpy(out, l1->data(), l1->size() * sizeof(float));
    return long(l1->size());
  }
  const auto* m = solver.level_matrix(std::size_t(level));
  if (!m || which < 0 || which > 2) return -1;
  const std::vector<Index>& v = which == 0 ? m->row_ptr : which == 1 ? m->col : m->diag;
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(Index));
  return long(v.size());
}
// Coarse operator values after the latest refresh (level >= 1).
long oracle_solver_level_values(void* h, int level, float* out) {
  const auto* m = static_cast<SolverHandle*>(h)->solver.level_matrix(std::size_t(level));
  if (!m) return -1;
  if (out) std::memcpy(out, m->val.data(), m->val.size() * sizeof(float));
  return long(m->val.size());
}

}  // extern "C"
