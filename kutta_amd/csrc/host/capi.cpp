// C ABI of the host library (include/kutta_amd.h).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "learned.hpp"
#include "kutta_amd.h"

using namespace kutta_amd;
using capi::guard;
using capi::Matrix;
using capi::MatrixHandle;
using capi::Solver;

std::string& kutta_amd::capi::last_error() {
  thread_local std::string error;
  return error;
}

namespace {
DeviceCsr view_of(void* m) { return static_cast<MatrixHandle*>(m)->view(); }
}  // namespace

extern "C" {

const char* ka_last_error(void) { return capi::last_error().c_str(); }

int ka_device_create(int ordinal, void** out) {
  return guard([&] { *out = new HipDevice(ordinal); });
}
void ka_device_destroy(void* dev) { delete static_cast<HipDevice*>(dev); }
kh_device* ka_device_raw(void* dev) { return static_cast<HipDevice*>(dev)->raw(); }

int ka_csr_create(int n_rows, const int32_t* row_ptr, const int32_t* col, const int32_t* diag,
                  const double* val, void** out) {
  return guard([&] {
    KA_REQUIRE(n_rows >= 0, "CSR: negative row count");
    auto* a = new CsrHost;
    a->n_rows = n_rows;
    a->row_ptr.assign(row_ptr, row_ptr + n_rows + 1);
    const int32_t nnz = row_ptr[n_rows];
    a->col.assign(col, col + nnz);
    a->diag.assign(diag, diag + n_rows);
    a->val.assign(val, val + nnz);
    *out = a;
  });
}
void ka_csr_destroy(void* csr) { delete static_cast<CsrHost*>(csr); }
int ka_csr_rows(void* csr) { return static_cast<CsrHost*>(csr)->n_rows; }
int ka_csr_nnz(void* csr) { return static_cast<CsrHost*>(csr)->nnz(); }
int ka_csr_copy(void* csr, int32_t* row_ptr, int32_t* col, int32_t* diag, double* val) {
  return guard([&] {
    const auto* a = static_cast<CsrHost*>(csr);
    std::memcpy(row_ptr, a->row_ptr.data(), a->row_ptr.size() * sizeof(int32_t));
    std::memcpy(col, a->col.data(), a->col.size() * sizeof(int32_t));
    std::memcpy(diag, a->diag.data(), a->diag.size() * sizeof(int32_t));
    std::memcpy(val, a->val.data(), a->val.size() * sizeof(double));
  });
}

int ka_aggregate(void* csr, const int32_t* options, double threshold, void** out) {
  return guard([&] {
    AmgOptions o;
    o.aggregation_passes = options[5];
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    o.strength_threshold = threshold;
    const auto& a = *static_cast<CsrHost*>(csr);
    validate_host(a);
    *out = new Coarsening(aggregate(a, o));
  });
}
long ka_coarsening_array(void* c, int which, int32_t* out) {
  const auto* g = static_cast<Coarsening*>(c);
  const std::vector<Index>* v = nullptr;
  switch (which) {
    case 0: v = &g->aggregate; break;
    case 1: v = &g->member_offsets; break;
    case 2: v = &g->members; break;
    case 3: v = &g->entry_offsets; break;
    case 4: v = &g->entries; break;
    case 5: v = &g->matrix.row_ptr; break;
    case 6: v = &g->matrix.col; break;
    case 7: v = &g->matrix.diag; break;
    default: return -1;
  }
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
  return long(v->size());
}
long ka_coarsening_values(void* c, double* out) {
  const auto* g = static_cast<Coarsening*>(c);
  if (out) std::memcpy(out, g->matrix.val.data(), g->matrix.val.size() * sizeof(double));
  return long(g->matrix.val.size());
}
void ka_coarsening_destroy(void* c) { delete static_cast<Coarsening*>(c); }

int ka_matrix_upload(void* dev, void* csr, void** out) {
  return guard([&] {
    auto* d = static_cast<HipDevice*>(dev);
    const auto& a = *static_cast<CsrHost*>(csr);
    validate_host(a);
    *out = static_cast<MatrixHandle*>(new Matrix(d, OwnedCsr(*d, a)));
  });
}
void ka_matrix_destroy(void* m) { delete static_cast<MatrixHandle*>(m); }
float* ka_matrix_values(void* m) { return view_of(m).val; }
const int32_t* ka_matrix_row_ptr(void* m) { return view_of(m).row_ptr; }
const int32_t* ka_matrix_col(void* m) { return view_of(m).col; }
const int32_t* ka_matrix_diag(void* m) { return view_of(m).diag; }

int ka_solver_create(void* dev, int kind, void* csr, const int32_t* options, double threshold,
                     void** out) {
  return guard([&] {
    KA_REQUIRE(kind >= 0 && kind <= 6, "solver: unknown SolverKind");
    AmgOptions o;
    o.pre_sweeps = options[0];
    o.post_sweeps = options[1];
    o.coarse_sweeps = options[2];
    o.coarse_rows = options[3];
    o.max_levels = options[4];
    o.aggregation_passes = options[5];
    o.nullspace = options[6] ? NullSpace::Constant : NullSpace::None;
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    KA_REQUIRE((options[8] == 0 || options[8] == 1) && options[9] >= 0 && options[10] >= 0, "AMG: invalid options");
    o.cycle = options[8] ? AmgCycle::K : AmgCycle::V;
    o.k_depth = options[9];
    o.k_min_rows = options[10];
    o.strength_threshold = threshold;
    auto* d = static_cast<HipDevice*>(dev);
    *out = new Solver{d, KrylovSolver(*d, static_cast<SolverKind>(kind), *static_cast<CsrHost*>(csr), o)};
  });
}
void ka_solver_destroy(void* s) { delete static_cast<Solver*>(s); }
int ka_solver_refresh(void* s, void* m) {
  return guard([&] { static_cast<Solver*>(s)->solver.refresh(view_of(m)); });
}
int ka_solver_solve(void* s, void* m, float* x, const float* b, double rel_tol, double abs_tol,
                    int max_iter, int interval, double* stats) {
  return guard([&] {
    SolveControl c{rel_tol, abs_tol, max_iter, interval};
    const auto st = static_cast<Solver*>(s)->solver.solve(view_of(m), x, b, c);
    stats[0] = st.iterations;
    stats[1] = st.initial_residual;
    stats[2] = st.final_residual;
    stats[3] = st.converged ? 1 : 0;
  });
}
int ka_solver_solve_batched(void* const* solvers, void* const* matrices, float* const* x, const float* const* b,
                            int count, double rel_tol, double abs_tol, int max_iter, int interval, double* stats) {
  return guard([&] {
    KA_REQUIRE(count >= 1 && count <= KH_BATCH_MAX_LANES, "batched solve: bad system count");
    SolveControl c{rel_tol, abs_tol, max_iter, interval};
    std::vector<KrylovSolver*> ss;
    std::vector<DeviceCsr> as;
    std::vector<float*> xs;
    std::vector<const float*> bs;
    for (int k = 0; k < count; ++k) {
      ss.push_back(&static_cast<Solver*>(solvers[k])->solver);
      as.push_back(view_of(matrices[k]));
      xs.push_back(x[k]);
      bs.push_back(b[k]);
    }
    const auto st = KrylovSolver::solve_batched(ss, as, xs, bs, c);
    for (int k = 0; k < count; ++k) {
      stats[4 * k] = st[k].iterations;
      stats[4 * k + 1] = st[k].initial_residual;
      stats[4 * k + 2] = st[k].final_residual;
      stats[4 * k + 3] = st[k].converged ? 1 : 0;
    }
  });
}
int ka_solver_describe(void* s, char* out, int capacity) {
  const auto text = static_cast<Solver*>(s)->solver.describe();
  if (capacity > 0) {
    std::strncpy(out, text.c_str(), std::size_t(capacity));
    out[capacity - 1] = 0;
  }
  return int(text.size());
}
int ka_solver_set_plan(void* s, int restrict_mode, int coarse_sweeps, int prolong_mode, int sweep_mode) {
  return guard([&] {
    static_cast<Solver*>(s)->solver.set_plan(restrict_mode, coarse_sweeps, prolong_mode, sweep_mode);
  });
}
int ka_solver_plan(void* s, char* out, int capacity) {
  const auto text = static_cast<Solver*>(s)->solver.plan();
  if (capacity > 0) {
    std::strncpy(out, text.c_str(), std::size_t(capacity));
    out[capacity - 1] = 0;
  }
  return int(text.size());
}
int ka_solver_levels(void* s, int32_t* rows, int32_t* nnz, int capacity) {
  const auto levels = static_cast<Solver*>(s)->solver.levels();
  for (std::size_t i = 0; i < levels.size() && int(i) < capacity; ++i) {
    rows[i] = levels[i].first;
    nnz[i] = levels[i].second;
  }
  return int(levels.size());
}
long ka_solver_transfer(void* s, int level, int which, int32_t* out) {
  try {
    const auto v = static_cast<Solver*>(s)->solver.transfer(std::size_t(level), which);
    if (v.empty()) return -1;
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    return long(v.size());
  } catch (const std::exception& e) {
    capi::last_error() = e.what();
    return -1;
  }
}
long ka_solver_level_values(void* s, int level, float* out) {
  try {
    const auto v = static_cast<Solver*>(s)->solver.level_values(std::size_t(level));
    if (v.empty()) return -1;
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(float));
    return long(v.size());
  } catch (const std::exception& e) {
    capi::last_error() = e.what();
    return -1;
  }
}
int ka_solver_set_fused(void* s, int fused) {
  return guard([&] { static_cast<Solver*>(s)->solver.set_fused(fused != 0); });
}
int ka_solver_host_waits(void* s) { return static_cast<Solver*>(s)->solver.host_waits(); }
int ka_solver_set_direct(void* solver, int rows, int32_t* level) {
  return guard([&] {
    const int l = static_cast<Solver*>(solver)->solver.set_direct(rows);
    if (level) *level = l;
  });
}
int ka_solver_set_dense_tolerance(void* solver, double tolerance) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_dense_tolerance(tolerance); });
}
int ka_solver_dense_refreshes(void* solver, int64_t* inversions, int64_t* skips, double* staleness) {
  return guard([&] {
    long i = 0, k = 0;
    double d = 0;
    static_cast<Solver*>(solver)->solver.dense_refreshes(&i, &k, &d);
    if (inversions) *inversions = i;
    if (skips) *skips = k;
    if (staleness) *staleness = d;
  });
}
int ka_solver_set_learned(void* solver, int level, int layers, uint64_t seed, int capacity) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_learned(level, layers, seed, capacity); });
}
int ka_solver_learned_mode(void* solver, int mode, double reduction, double max_relative) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_learned_mode(mode, reduction, max_relative); });
}
int ka_solver_learned_targets(void* solver, double rel_tol, int32_t* pairs) {
  return guard([&] {
    const int k = static_cast<Solver*>(solver)->solver.learned_targets(rel_tol);
    if (pairs) *pairs = k;
  });
}
int ka_solver_learned_train(void* solver, int steps, float lr, int pairs, double* losses, double* relative) {
  return guard([&] {
    const auto l = static_cast<Solver*>(solver)->solver.learned_train(steps, lr, pairs);
    if (losses) std::copy(l.loss.begin(), l.loss.end(), losses);
    if (relative) std::copy(l.relative.begin(), l.relative.end(), relative);
  });
}
int ka_solver_learned_evaluate(void* solver, int first, int count, double* relative) {
  return guard([&] { *relative = static_cast<Solver*>(solver)->solver.learned_evaluate(first, count); });
}
int ka_solver_learned_report(void* solver, double* out) {
  return guard([&] {
    const auto r = static_cast<Solver*>(solver)->solver.learned_report();
    const double v[11] = {double(r.level), double(r.rows), double(r.mode), double(r.harvested),
                          double(r.targets), double(r.applications), double(r.fallbacks),
                          double(r.solves), r.last_loss, r.relative, double(r.gated)};
    std::copy(v, v + 11, out);
  });
}
int ka_solver_learned_parameters(void* solver, float* theta, int64_t count, int set) {
  return guard([&] {
    auto& s = static_cast<Solver*>(solver)->solver;
    if (set) {
      s.set_learned_parameters(std::vector<float>(theta, theta + count));
    } else {
      const auto v = s.learned_parameters();
      KA_REQUIRE(int64_t(v.size()) == count, "learned: parameter count mismatch");
      std::copy(v.begin(), v.end(), theta);
    }
  });
}
int ka_solver_autotune_tail(void* solver, void* matrix, int repeats, int32_t* level) {
  return guard([&] {
    const int l = static_cast<Solver*>(solver)->solver.autotune_tail(view_of(matrix), repeats);
    if (level) *level = l;
  });
}
int ka_solver_set_graphs(void* solver, int on) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_graphs(on != 0); });
}
int ka_solver_set_coarse_bf16(void* solver, int on) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_coarse_bf16(on != 0); });
}
int ka_solver_set_restart(void* solver, int m) {
  return guard([&] { static_cast<Solver*>(solver)->solver.set_restart(m); });
}
int ka_solver_set_tail(void* s, int first, int* level) {
  return guard([&] {
    const int used = static_cast<Solver*>(s)->solver.set_tail(first);
    if (level) *level = used;
  });
}

}  // extern "C"
