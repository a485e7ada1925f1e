// ORACLE — test infrastructure only (see oracle.hpp). Serial restatement of the
// reference's host-side aggregation AMG (linalg/amg.cpp) and device-scalar
// Krylov driver (linalg/solver.cpp), running the kernels of linalg_kernels.cpp.
#include <algorithm>
#include <array>
#include <cmath>
#include <iomanip>
#include <limits>
#include <memory>
#include <numeric>
#include <optional>
#include <sstream>
#include <tuple>

#include "oracle.hpp"

namespace oracle {

// linalg/csr.cpp:11-31 and :53-59
void validate_host(const CsrHost& p) {
  ORACLE_REQUIRE(p.n_rows >= 0, "CSR: negative row count");
  ORACLE_REQUIRE(p.row_ptr.size() == std::size_t(p.n_rows) + 1 &&
                     p.diag.size() == std::size_t(p.n_rows), "CSR: invalid array sizes");
  ORACLE_REQUIRE(p.row_ptr.front() == 0 && p.row_ptr.back() == p.nnz(), "CSR: invalid row offsets");
  for (Index i = 0; i < p.n_rows; ++i) {
    ORACLE_REQUIRE(p.row_ptr[i] >= 0 && p.row_ptr[i] <= p.row_ptr[i + 1] &&
                       p.row_ptr[i + 1] <= p.nnz(), "CSR: invalid row range");
    Index previous = -1;
    for (Index j = p.row_ptr[i]; j < p.row_ptr[i + 1]; ++j) {
      ORACLE_REQUIRE(p.col[j] > previous && p.col[j] < p.n_rows,
                     "CSR: columns must be in range, unique and sorted");
      previous = p.col[j];
    }
    ORACLE_REQUIRE(p.diag[i] >= p.row_ptr[i] && p.diag[i] < p.row_ptr[i + 1] &&
                       p.col[p.diag[i]] == i, "CSR: missing or incorrect diagonal");
  }
  ORACLE_REQUIRE(p.val.size() == p.col.size(), "CSR: value count mismatch");
  for (double v : p.val)
    ORACLE_REQUIRE(std::isfinite(v) && std::abs(v) <= std::numeric_limits<float>::max(),
                   "CSR: nonfinite or out-of-range value");
}

// linalg/csr.cpp:61-76 (pattern upload zeroes values, then values upload)
CsrDevice upload(const CsrHost& a) {
  CsrDevice d;
  d.n_rows = a.n_rows;
  d.row_ptr = a.row_ptr;
  d.col = a.col;
  d.diag = a.diag;
  d.val.resize(a.val.size());
  for (std::size_t j = 0; j < a.val.size(); ++j) d.val[j] = static_cast<float>(a.val[j]);
  return d;
}

// linalg/solver.cpp:37-53
std::vector<float> diagonal_values(const CsrHost& a, bool l1, bool nullspace) {
  std::vector<float> inverse(a.n_rows);
  for (Index i = 0; i < a.n_rows; ++i) {
    const float d = static_cast<float>(a.val[a.diag[i]]);
    ORACLE_REQUIRE(!l1 || d >= 0, "AMG: requires positive diagonals");
    ORACLE_REQUIRE(d != 0 || (nullspace && a.n_rows == 1), "solver: invalid values or zero diagonal");
    float off = 0;
    if (l1)
      for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j)
        if (j != a.diag[i]) off += std::abs(static_cast<float>(a.val[j]));
    const float scale = l1 ? std::max(std::abs(d), 0.5f * (std::abs(d) + off)) : d;
    inverse[i] = scale != 0 ? 1.0f / scale : 0;
    ORACLE_REQUIRE(std::isfinite(scale) && std::isfinite(inverse[i]),
                   "solver: invalid values or diagonal inverse");
  }
  return inverse;
}

// linalg/amg.cpp:25-75 — greedy strongest-unassigned matching.
std::vector<Index> pairs(const CsrHost& a, double threshold, bool exclude_isolated) {
  const Index n = a.n_rows;
  std::vector<Index> offsets(std::size_t(n) + 1, 0), sources(a.nnz());
  for (auto c : a.col) ++offsets[c + 1];
  std::partial_sum(offsets.begin(), offsets.end(), offsets.begin());
  auto cursor = offsets;
  for (Index i = 0; i < n; ++i)
    for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j) sources[cursor[a.col[j]]++] = i;
  auto neighbors = [&](Index i, auto visit) {
    Index out = a.row_ptr[i], in = offsets[i];
    while (out < a.row_ptr[i + 1] || in < offsets[i + 1]) {
      const Index oc = out < a.row_ptr[i + 1] ? a.col[out] : n;
      const Index ic = in < offsets[i + 1] ? sources[in] : n;
      const Index col = std::min(oc, ic);
      double w = 0;
      if (oc == col) w = std::abs(a.val[out++]);
      if (ic == col) {
        const auto first = a.col.begin() + a.row_ptr[col];
        const auto last = a.col.begin() + a.row_ptr[col + 1];
        const auto entry = std::lower_bound(first, last, i) - a.col.begin();
        w = std::max(w, std::abs(a.val[entry]));
        ++in;
      }
      if (col != i && w != 0) visit(col, w);
    }
  };
  std::vector<Index> map(n, -1);
  Index next = 0;
  for (Index i = 0; i < n; ++i) {
    if (map[i] != -1) continue;
    double strongest = 0;
    neighbors(i, [&](Index, double w) { strongest = std::max(strongest, w); });
    if (exclude_isolated && strongest == 0) continue;
    Index partner = -1;
    double best = 0;
    neighbors(i, [&](Index col, double w) {
      if (map[col] == -1 && w > best && w >= threshold * strongest) { partner = col; best = w; }
    });
    map[i] = next;
    if (partner != -1) map[partner] = next;
    ++next;
  }
  return map;
}

// linalg/amg.cpp:77-158 — coarse operator, member lists and canonical entry maps.
Coarsening coarsen(const CsrHost& fine, std::vector<Index> aggregate, bool entry_map,
                   bool m_matrix) {
  Coarsening c;
  c.aggregate = std::move(aggregate);
  const Index nc = c.aggregate.empty() ? 0
      : *std::max_element(c.aggregate.begin(), c.aggregate.end()) + 1;
  auto& cp = c.matrix;
  cp.n_rows = nc;
  cp.row_ptr.assign(std::size_t(nc) + 1, 0);
  cp.diag.resize(nc);
  c.member_offsets.assign(std::size_t(nc) + 1, 0);
  for (auto g : c.aggregate) if (g >= 0) ++c.member_offsets[g + 1];
  std::partial_sum(c.member_offsets.begin(), c.member_offsets.end(), c.member_offsets.begin());
  c.members.resize(c.member_offsets.back());
  auto cursor = c.member_offsets;
  for (Index i = 0; i < fine.n_rows; ++i)
    if (c.aggregate[i] >= 0) c.members[cursor[c.aggregate[i]]++] = i;
  std::vector<std::tuple<Index, Index, Index>> row;
  if (entry_map) c.entry_offsets.push_back(0);
  for (Index g = 0; g < nc; ++g) {
    row.clear();
    for (Index m = c.member_offsets[g]; m < c.member_offsets[g + 1]; ++m) {
      const Index i = c.members[m];
      for (Index j = fine.row_ptr[i]; j < fine.row_ptr[i + 1]; ++j)
        if (c.aggregate[fine.col[j]] >= 0) {
          Index canonical = j;
          if (m_matrix) {
            const auto col = fine.col[j];
            const auto first = fine.col.begin() + fine.row_ptr[col];
            const auto last = fine.col.begin() + fine.row_ptr[col + 1];
            canonical = std::min(j, static_cast<Index>(std::lower_bound(first, last, i) - fine.col.begin()));
          }
          row.emplace_back(c.aggregate[fine.col[j]], canonical, j);
        }
    }
    std::sort(row.begin(), row.end());
    Index previous = -1;
    float s = 0, e = 0;
    for (auto [col, canonical, source] : row) {
      (void)canonical;
      if (col != previous) {
        if (previous != -1) cp.val.back() = s + e;
        s = e = 0;
        if (entry_map && previous != -1) c.entry_offsets.push_back(Index(c.entries.size()));
        if (col == g) cp.diag[g] = Index(cp.col.size());
        cp.col.push_back(col);
        cp.val.push_back(0);
        previous = col;
      }
      if (entry_map) c.entries.push_back(source);
      const float t = static_cast<float>(fine.val[source]);
      const float nx = s + t;
      e += std::abs(s) >= std::abs(t) ? (s - nx) + t : (t - nx) + s;
      s = nx;
    }
    if (previous != -1) cp.val.back() = s + e;
    if (entry_map) c.entry_offsets.push_back(Index(c.entries.size()));
    cp.row_ptr[g + 1] = Index(cp.col.size());
    if (m_matrix) {
      UpperSum d{0, 0};
      for (Index m = c.member_offsets[g]; m < c.member_offsets[g + 1]; ++m) {
        const auto i = c.members[m];
        UpperSum surplus{float(fine.val[fine.diag[i]]), 0};
        for (Index j = fine.row_ptr[i]; j < fine.row_ptr[i + 1]; ++j)
          if (j != fine.diag[i]) surplus = kupper_add(surplus, float(fine.val[j]));
        d = kupper_add(d, kupper_value(surplus));
      }
      for (Index j = cp.row_ptr[g]; j < cp.row_ptr[g + 1]; ++j)
        if (j != cp.diag[g]) d = kupper_add(d, -float(cp.val[j]));
      cp.val[cp.diag[g]] = kupper_value(d);
    }
  }
  return c;
}

// linalg/amg.cpp:160-174 — repeated pairing on the intermediate operator.
Coarsening aggregate(const CsrHost& a, const AmgOptions& options) {
  const bool m = options.structure == Structure::SymmetricMMatrix;
  auto map = pairs(a, options.strength_threshold, true);
  for (int pass = 1; pass < options.aggregation_passes; ++pass) {
    auto intermediate = coarsen(a, std::move(map), false, m);
    auto next = pairs(intermediate.matrix, options.strength_threshold, false);
    map = std::move(intermediate.aggregate);
    for (auto& g : map) if (g >= 0) g = next[g];
  }
  return coarsen(a, std::move(map), true, m);
}

namespace {
using V = std::vector<float>;

// One hierarchy level (linalg/amg.cpp:179-196). The fine level borrows the
// caller's operator; coarse levels own theirs.
struct Level {
  CsrDevice matrix;
  V l1, x, b, r, projection;
  // K-cycle workspaces of a coarse level (linalg/amg.cpp:185-193).
  V p, q, z, t, scalars;
  std::vector<int> status;
  Coarsening transfer;  // valid when a coarser level exists
  bool has_transfer = false;
};

// Constant-nullspace projection (linalg/detail.hpp:13-19).
void project(V& x, V& scalar) {
  if (x.empty()) return;
  reduce_into(x.data(), x.data(), x.data(), x.data(), x.size(), scalar.data(), Reduction::Sum);
  const float count = static_cast<float>(x.size());
  kern::solver_project(int(x.size()), x.data(), scalar.data(), count,
                       static_cast<float>(double(x.size()) - double(count)));
}
double norm(const V& x) { return std::sqrt(dot(x.data(), x.data(), x.size())); }
}  // namespace

struct Solver::Impl {
  SolverKind kind;
  AmgOptions options;
  bool amg, nullspace;
  Index n = 0, nnz = 0;
  double complexity = 0;
  std::vector<Level> levels;  // AMG only
  V jacobi_inverse, jacobi_invalid;
  V r, z, p, q, shadow, s, t, zs, best;
  V scalars = V(24, 0.0f);
  std::vector<int> status = std::vector<int>(3, 0);
  V projection = V(2, 0.0f);
  double best_residual = 0;
  int previous_iterations = 0;
  bool verified_current = false;

  Impl(SolverKind k, const CsrHost& a, const AmgOptions& o)
      : kind(k), options(o),
        amg(k == SolverKind::AmgPcg || k == SolverKind::AmgBicgstab),
        nullspace(o.nullspace == NullSpace::Constant), n(a.n_rows), nnz(a.nnz()) {
    validate_host(a);
    if (amg) {
      ORACLE_REQUIRE(options.pre_sweeps > 0 && options.post_sweeps > 0 &&
                         options.coarse_sweeps > 0 && options.coarse_rows >= 1 &&
                         options.max_levels >= 1 && options.aggregation_passes >= 1 &&
                         options.aggregation_passes <= 8 && options.strength_threshold >= 0 &&
                         options.strength_threshold <= 1, "AMG: invalid options");
      ORACLE_REQUIRE(k != SolverKind::AmgPcg || options.pre_sweeps == options.post_sweeps,
                     "AMG-PCG requires equal pre/post sweeps");
      if (options.structure == Structure::SymmetricMMatrix) check_m_matrix(a);
      build_hierarchy(a);
    } else {
      jacobi_inverse = diagonal_values(a, false, false);
      jacobi_invalid.assign(n, 0.0f);
    }
    r.assign(n, 0); z.assign(n, 0); p.assign(n, 0); q.assign(n, 0); best.assign(n, 0);
    if (bicg()) { shadow.assign(n, 0); s.assign(n, 0); t.assign(n, 0); zs.assign(n, 0); }
  }
  bool bicg() const { return kind == SolverKind::BicgstabJacobi || kind == SolverKind::AmgBicgstab; }

  // linalg/amg.cpp:364-379
  static void check_m_matrix(const CsrHost& a) {
    for (Index i = 0; i < a.n_rows; ++i) {
      long double total = 0;
      for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j) {
        const float value = float(a.val[j]);
        total += value;
        const auto col = a.col[j];
        const auto first = a.col.begin() + a.row_ptr[col], last = a.col.begin() + a.row_ptr[col + 1];
        const auto opposite = std::lower_bound(first, last, i);
        ORACLE_REQUIRE(opposite != last && *opposite == i &&
                           value == float(a.val[opposite - a.col.begin()]) && (col == i || value <= 0),
                       "AMG M-matrix: expected symmetric nonpositive edges");
      }
      ORACLE_REQUIRE(total >= 0, "AMG M-matrix: expected diagonally dominant rows");
    }
  }

  // linalg/amg.cpp:200-226
  void build_hierarchy(const CsrHost& a) {
    Level fine;
    fine.l1 = diagonal_values(a, true, nullspace);
    fine.r.assign(a.n_rows, 0.0f);
    if (nullspace) fine.projection.assign(2, 0.0f);
    levels.push_back(std::move(fine));
    const CsrHost* current = &a;
    CsrHost owned;
    std::size_t total = a.nnz();
    while (current->n_rows > options.coarse_rows && levels.size() < std::size_t(options.max_levels)) {
      auto c = aggregate(*current, options);
      const auto active = c.members.size();
      if (active == 0 || c.matrix.n_rows >= 0.85 * double(active)) break;
      total += c.matrix.nnz();
      Level l;
      l.l1 = diagonal_values(c.matrix, true, nullspace);
      l.x.assign(c.matrix.n_rows, 0.0f);
      l.b.assign(c.matrix.n_rows, 0.0f);
      l.r.assign(c.matrix.n_rows, 0.0f);
      if (nullspace) l.projection.assign(2, 0.0f);
      if (options.cycle == AmgCycle::K) {
        for (V* v : {&l.p, &l.q, &l.z, &l.t}) v->assign(c.matrix.n_rows, 0.0f);
        l.scalars.assign(12, 0.0f);
        l.status.assign(1, 0);
      }
      l.matrix = upload(c.matrix);
      owned = c.matrix;
      levels.back().transfer = std::move(c);
      levels.back().has_transfer = true;
      levels.push_back(std::move(l));
      current = &owned;
    }
    complexity = a.nnz() ? double(total) / a.nnz() : 0;
  }

  // linalg/amg.cpp:228-256
  void refresh_amg(const CsrDevice& a) {
    const CsrDevice* m = &a;
    std::vector<double> invalid, allowed;
    for (std::size_t i = 0; i < levels.size(); ++i) {
      auto& l = levels[i];
      kern::krylov_diagonal(m->n_rows, m->row_ptr.data(), m->diag.data(), m->val.data(),
                            l.l1.data(), l.r.data(), 1);
      allowed.push_back(nullspace && m->n_rows == 1 ? 1 : 0);
      invalid.push_back(max_abs(l.r.data(), m->n_rows));
      if (i + 1 < levels.size()) {
        auto& coarse = levels[i + 1].matrix;
        auto& t = l.transfer;
        kern::amg_galerkin(coarse.nnz(), t.entry_offsets.data(), t.entries.data(), m->val.data(),
                           coarse.val.data());
        if (options.structure == Structure::SymmetricMMatrix)
          kern::amg_m_matrix_diagonal(coarse.n_rows, t.member_offsets.data(), t.members.data(),
                                      m->row_ptr.data(), m->diag.data(), m->val.data(),
                                      coarse.row_ptr.data(), coarse.diag.data(), coarse.val.data());
        m = &coarse;
      }
    }
    for (std::size_t i = 0; i < invalid.size(); ++i) {
      ORACLE_REQUIRE(invalid[i] != 3, "AMG: requires positive diagonals");
      ORACLE_REQUIRE(invalid[i] <= allowed[i], "AMG: invalid values or zero diagonal");
    }
  }

  // linalg/amg.cpp:272-282 — ping-pong between x and the level's r. `st` is
  // the status that guards the sweeps: the solver's, or an inner correction's.
  void smooth(Level& l, const CsrDevice& a, const V& b, V& x, int sweeps, bool zero_start,
              const int* st) {
    for (int sweep = 0; sweep < sweeps; ++sweep) {
      if (zero_start && sweep == 0)
        kern::amg_smooth_zero(a.n_rows, a.row_ptr.data(), a.col.data(), a.val.data(), x.data(),
                              b.data(), l.l1.data(), l.r.data(), st);
      else
        kern::amg_smooth(a.n_rows, a.row_ptr.data(), a.col.data(), a.val.data(), x.data(),
                         b.data(), l.l1.data(), l.r.data(), st);
      std::swap(x, l.r);
    }
  }
  bool k_inner(std::size_t coarse) const {
    return options.cycle == AmgCycle::K && coarse + 1 < levels.size() &&
           (options.k_depth == 0 || coarse <= std::size_t(options.k_depth)) &&
           levels[coarse].matrix.n_rows >= options.k_min_rows;
  }
  // linalg/amg.cpp:283-310 — V-cycle, or K-cycle where k_inner holds.
  void cycle(std::size_t index, const CsrDevice& a, const V& b, V& x, const int* st) {
    auto& l = levels[index];
    std::fill(x.begin(), x.end(), 0.0f);
    if (index + 1 == levels.size()) {
      if (a.n_rows <= 24)
        kern::amg_coarse_solve(1, a.row_ptr.data(), a.col.data(), a.val.data(), x.data(), b.data(),
                               l.l1.data(), st, a.n_rows, options.coarse_sweeps);
      else smooth(l, a, b, x, options.coarse_sweeps, true, st);
    } else {
      smooth(l, a, b, x, options.pre_sweeps, true, st);
      kern::csr_residual_guarded(a.n_rows, a.row_ptr.data(), a.col.data(), a.val.data(), x.data(),
                                 b.data(), l.r.data(), st);
      auto& c = levels[index + 1];
      kern::amg_restrict_guarded(c.matrix.n_rows, l.transfer.member_offsets.data(),
                                 l.transfer.members.data(), l.r.data(), c.b.data(), st);
      if (nullspace) project(c.b, c.projection);
      if (k_inner(index + 1)) inner(index + 1, st);
      else cycle(index + 1, c.matrix, c.b, c.x, st);
      kern::amg_prolong_guarded(a.n_rows, l.transfer.aggregate.data(), c.x.data(), x.data(), st);
      smooth(l, a, b, x, options.post_sweeps, false, st);
    }
    if (nullspace && index != 0) project(x, l.projection);
  }

  // linalg/amg.cpp:315-336 — two preconditioned Krylov corrections of the
  // coarse equation A x = b at `index` (A-conjugate for SPD, GCR otherwise),
  // the second skipped on the device once the first leaves |t| <= |b| / 4.
  void inner(std::size_t index, const int* parent) {
    auto& l = levels[index];
    const auto& a = l.matrix;
    const int n = a.n_rows;
    const bool spd = kind == SolverKind::AmgPcg;
    float* s = l.scalars.data();
    int* st = l.status.data();
    auto spmv = [&](const V& in, V& out, const int* guard) {
      kern::csr_spmv_guarded(n, a.row_ptr.data(), a.col.data(), a.val.data(), in.data(), out.data(), guard);
    };
    auto dot2 = [&](const V& x, const V& y, const V& u, const V& v, int offset) {
      reduce_into(x.data(), y.data(), u.data(), v.data(), n, s + offset, Reduction::Dot2);
    };
    cycle(index, a, l.b, l.z, parent);
    l.p = l.z;
    spmv(l.p, l.q, parent);
    dot2(spd ? l.p : l.q, l.q, l.b, spd ? l.p : l.q, 0);
    kern::amg_inner_alpha(s, st, parent, 1);
    kern::amg_inner_first(n, l.b.data(), l.p.data(), l.q.data(), l.x.data(), l.t.data(), s, st);
    dot2(l.t, l.t, l.b, l.b, 4);
    kern::amg_inner_check(s, st);
    cycle(index, a, l.t, l.z, st);
    if (!spd) spmv(l.z, l.r, st);
    const V& w = spd ? l.z : l.r;
    reduce_into(w.data(), l.q.data(), w.data(), l.q.data(), n, s + 4, Reduction::Dot);
    kern::amg_inner_beta(s, st);
    kern::solver_direction(n, l.z.data(), l.p.data(), s, st);
    spmv(l.p, l.q, st);
    dot2(spd ? l.p : l.q, l.q, l.t, spd ? l.p : l.q, 0);
    kern::amg_inner_alpha(s, st, parent, 0);
    kern::amg_inner_step(n, l.x.data(), l.p.data(), s, st);
  }
  bool variable() const { return amg && options.cycle == AmgCycle::K; }

  // linalg/solver.cpp:197-200
  void apply(const CsrDevice& a, const V& rin, V& zout) {
    if (amg) cycle(0, a, rin, zout, status.data());
    else kern::krylov_multiply_guarded(int(rin.size()), rin.data(), jacobi_inverse.data(),
                                       zout.data(), status.data());
    if (nullspace) project(zout, projection);
  }
  void residual(const CsrDevice& a, const V& x, const V& b, V& out) {
    kern::csr_residual(a.n_rows, a.row_ptr.data(), a.col.data(), a.val.data(), x.data(), b.data(),
                       out.data());
  }
  // linalg/solver.cpp:201-222 (the speculative residual of :230-244 reproduces
  // exactly these numbers and is not restated).
  bool verified(const CsrDevice& a, V& x, const V& b, double target, SolveStats& st) {
    if (nullspace) project(x, projection);
    residual(a, x, b, r);
    st.final_residual = norm(r);
    verified_current = true;
    st.converged = std::isfinite(st.final_residual) && st.final_residual <= target;
    if (st.final_residual < best_residual) { best_residual = st.final_residual; best = x; }
    if (nullspace) project(r, projection);
    return st.converged;
  }
  static bool diverged(const SolveStats& st) {
    return !std::isfinite(st.final_residual) || st.final_residual > 1e3 * st.initial_residual;
  }
  void reset(int iterations) { kern::solver_reset(scalars.data(), status.data(), iterations); }
  bool replacement(const CsrDevice& a, V& x, const V& b, double target, SolveStats& st,
                   int& restarts, double& last) {
    if (verified(a, x, b, target, st) || diverged(st)) return false;
    restarts = st.final_residual >= 0.9 * last ? restarts + 1 : 0;
    last = st.final_residual;
    return restarts < 12;
  }
  int batch_interval(const SolveControl& c) const {
    return previous_iterations > 0 && previous_iterations < c.residual_check_interval
               ? previous_iterations + 1 : c.residual_check_interval;
  }
  void dot_into(const V& x, const V& y, int offset) {
    reduce_into(x.data(), y.data(), x.data(), y.data(), n, scalars.data() + offset, Reduction::Dot);
  }

  // linalg/solver.cpp:271-313
  void pcg(const CsrDevice& a, V& x, const V& b, const SolveControl& c, double target,
           SolveStats& st) {
    auto restart = [&] {
      reset(st.iterations);
      apply(a, r, z);
      p = z;
      dot_into(r, z, 0);
    };
    restart();
    int restarts = 0;
    double last = st.initial_residual;
    int interval = batch_interval(c);
    const int* rp = a.row_ptr.data();
    while (st.iterations < c.max_iter) {
      const int batch = std::min(interval, c.max_iter - st.iterations);
      verified_current = false;
      for (int j = 0; j < batch; ++j) {
        kern::csr_spmv_guarded(n, rp, a.col.data(), a.val.data(), p.data(), q.data(), status.data());
        dot_into(p, q, 2);
        kern::solver_pcg_alpha(scalars.data(), status.data());
        kern::solver_step(n, x.data(), r.data(), p.data(), q.data(), scalars.data(), status.data());
        dot_into(r, r, 18);
        kern::solver_check(scalars.data(), status.data(), 1);
        if (st.iterations + j + 1 == c.max_iter) continue;
        apply(a, r, z);
        if (variable())
          reduce_into(r.data(), z.data(), z.data(), q.data(), n, scalars.data() + 4, Reduction::Dot2);
        else dot_into(r, z, 4);
        kern::solver_pcg_beta(scalars.data(), status.data(), variable() ? 1 : 0);
        kern::solver_direction(n, z.data(), p.data(), scalars.data(), status.data());
      }
      st.iterations = status[1];
      if (status[0] == 2) break;
      if (status[0] == 1) {
        if (!replacement(a, x, b, target, st, restarts, last) || st.iterations == c.max_iter) break;
        if (st.final_residual <= 10 * target) interval = 1;
        restart();
      }
    }
  }

  // linalg/solver.cpp:314-361
  void bicgstab(const CsrDevice& a, V& x, const V& b, const SolveControl& c, double target,
                SolveStats& st) {
    auto restart = [&] {
      reset(st.iterations);
      shadow = r;
      std::fill(p.begin(), p.end(), 0.0f);
      std::fill(q.begin(), q.end(), 0.0f);
    };
    restart();
    int restarts = 0;
    double last = st.initial_residual;
    int interval = batch_interval(c);
    const int* rp = a.row_ptr.data();
    float* sc = scalars.data();
    int* stt = status.data();
    while (st.iterations < c.max_iter) {
      const int batch = std::min(interval, c.max_iter - st.iterations);
      verified_current = false;
      for (int j = 0; j < batch; ++j) {
        dot_into(shadow, r, 0);
        kern::solver_bicg_beta(sc, stt);
        kern::solver_bicg_direction(n, r.data(), q.data(), p.data(), sc, stt);
        apply(a, p, z);
        kern::csr_spmv_guarded(n, rp, a.col.data(), a.val.data(), z.data(), q.data(), stt);
        dot_into(shadow, q, 2);
        kern::solver_bicg_alpha(sc, stt);
        kern::solver_bicg_difference(n, r.data(), q.data(), s.data(), sc, stt);
        dot_into(s, s, 16);
        kern::solver_bicg_s(sc, stt);
        apply(a, s, zs);
        kern::csr_spmv_guarded(n, rp, a.col.data(), a.val.data(), zs.data(), t.data(), stt);
        reduce_into(t.data(), t.data(), t.data(), s.data(), n, sc + 2, Reduction::Dot2);
        kern::solver_bicg_omega(sc, stt);
        kern::solver_bicg_step(n, x.data(), r.data(), z.data(), zs.data(), s.data(), t.data(), sc, stt);
        dot_into(r, r, 18);
        kern::solver_check(sc, stt, 0);
      }
      st.iterations = status[1];
      if (status[0] == 2) break;
      if (status[0] == 1) {
        if (!replacement(a, x, b, target, st, restarts, last) || st.iterations == c.max_iter) break;
        if (st.final_residual <= 10 * target) interval = 1;
        restart();
      }
    }
  }

  // linalg/solver.cpp:97-176
  SolveStats solve(const CsrDevice& a, V& x, const V& b, const SolveControl& c) {
    ORACLE_REQUIRE(a.n_rows == n && a.nnz() == nnz, "solver: device or matrix mismatch");
    ORACLE_REQUIRE(x.size() == std::size_t(n) && b.size() == x.size(), "solve: vector size mismatch");
    ORACLE_REQUIRE(c.max_iter >= 0 && std::isfinite(c.rel_tol) && c.rel_tol >= 0 &&
                       std::isfinite(c.abs_tol) && c.abs_tol >= 0 && c.residual_check_interval > 0,
                   "solve: invalid control");
    if (!nullspace) residual(a, x, b, r);
    const double e0 = dot(b.data(), b.data(), b.size());
    const double e1 = dot(x.data(), x.data(), x.size());
    const double e2 = nullspace ? sum(b.data(), n) : dot(r.data(), r.data(), r.size());
    const double bnorm = std::sqrt(e0);
    if (!(std::isfinite(bnorm) && std::isfinite(std::sqrt(e1))))
      throw NonfiniteError("solve: nonfinite input vector");
    SolveStats st;
    if (nullspace) {
      if (n != 0) {
        ORACLE_REQUIRE(std::abs(e2) <= 1e-6 * std::sqrt(double(n)) * bnorm,
                       "solve: incompatible RHS for constant Neumann nullspace");
        project(x, projection);
      }
      residual(a, x, b, r);
      st.initial_residual = norm(r);
    } else st.initial_residual = std::sqrt(e2);
    st.final_residual = st.initial_residual;
    verified_current = true;
    const double target = std::max(c.abs_tol, c.rel_tol * st.initial_residual);
    if (!std::isfinite(st.initial_residual)) return st;
    if (st.final_residual <= target) { st.converged = true; return st; }
    if (nullspace) project(r, projection);
    best_residual = st.initial_residual;
    best = x;
    std::array<float, 24> limits{};
    store_pair(&limits[20], target * target);
    store_pair(&limits[22], 1e6 * st.initial_residual * st.initial_residual);
    kern::solver_limits(24, scalars.data(), limits[20], limits[21], limits[22], limits[23]);
    reset(0);
    if (kind == SolverKind::PcgJacobi || kind == SolverKind::AmgPcg) pcg(a, x, b, c, target, st);
    else if (bicg()) bicgstab(a, x, b, c, target, st);
    else ORACLE_REQUIRE(false, "solver kind not restated");
    if (!verified_current || nullspace) verified(a, x, b, target, st);
    if (!std::isfinite(st.final_residual) || st.final_residual > best_residual) {
      x = best;
      verified(a, x, b, target, st);
    }
    previous_iterations = st.converged ? st.iterations : 0;
    return st;
  }

  void refresh(const CsrDevice& a) {
    ORACLE_REQUIRE(a.n_rows == n && a.nnz() == nnz, "solver: device or matrix mismatch");
    if (amg) refresh_amg(a);
    else {
      kern::krylov_diagonal(a.n_rows, a.row_ptr.data(), a.diag.data(), a.val.data(),
                            jacobi_inverse.data(), jacobi_invalid.data(), 0);
      ORACLE_REQUIRE(max_abs(jacobi_invalid.data(), n) == 0, "Jacobi: invalid values or zero diagonal");
    }
  }
};

Solver::Solver(SolverKind kind, const CsrHost& a, const AmgOptions& options)
    : impl_(new Impl(kind, a, options)) {}
Solver::~Solver() { delete impl_; }
void Solver::refresh(const CsrDevice& a) { impl_->refresh(a); }
SolveStats Solver::solve(const CsrDevice& a, std::vector<float>& x, const std::vector<float>& b,
                         const SolveControl& control) {
  return impl_->solve(a, x, b, control);
}
// linalg/amg.cpp:263-269 and linalg/solver.cpp:178-184
std::string Solver::describe() const {
  std::ostringstream out;
  const bool pcg = impl_->kind == SolverKind::PcgJacobi || impl_->kind == SolverKind::AmgPcg;
  out << (pcg ? "PCG" : "BiCGStab") << ", ";
  if (impl_->amg)
    out << "aggregation AMG " << (impl_->options.cycle == AmgCycle::K ? 'K' : 'V') << "-cycle, "
        << impl_->levels.size() << " levels, op complexity "
        << std::fixed << std::setprecision(3) << impl_->complexity;
  else out << "Jacobi";
  if (impl_->nullspace) out << ", constant-nullspace projection";
  return out.str();
}
std::vector<LevelInfo> Solver::levels() const {
  std::vector<LevelInfo> out;
  for (std::size_t i = 0; i < impl_->levels.size(); ++i) {
    const auto& l = impl_->levels[i];
    out.push_back({i == 0 ? impl_->n : l.matrix.n_rows, i == 0 ? impl_->nnz : l.matrix.nnz()});
  }
  return out;
}
const Coarsening* Solver::transfer(std::size_t level) const {
  if (level >= impl_->levels.size() || !impl_->levels[level].has_transfer) return nullptr;
  return &impl_->levels[level].transfer;
}
const CsrDevice* Solver::level_matrix(std::size_t level) const {
  if (level == 0 || level >= impl_->levels.size()) return nullptr;
  return &impl_->levels[level].matrix;
}
const std::vector<float>* Solver::level_l1(std::size_t level) const {
  if (level >= impl_->levels.size()) return nullptr;
  return &impl_->levels[level].l1;
}
void Solver::apply(const CsrDevice& a, const std::vector<float>& r, std::vector<float>& z) {
  ORACLE_REQUIRE(a.n_rows == impl_->n && r.size() == std::size_t(impl_->n) && z.size() == r.size(),
                 "apply: size mismatch");
  impl_->status.assign(3, 0);
  impl_->apply(a, r, z);
}

// tests/linalg/matrices.hpp:13-15
std::uint32_t hash(std::uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; return x ^ (x >> 16);
}
// tests/linalg/matrices.hpp:17-62 — finite-difference Poisson test operators.
CsrHost poisson(Index nx, Index ny, Index nz, double anisotropy, bool random, bool neumann,
                double wind) {
  CsrHost a;
  a.n_rows = nx * ny * nz;
  a.diag.resize(a.n_rows);
  const Index dims[] = {nx, ny, nz};
  const Index strides[] = {1, nx, nx * ny};
  const int dimension = nz == 1 ? 2 : 3;
  std::vector<std::pair<Index, double>> row;
  for (Index i = 0; i < a.n_rows; ++i) {
    double diagonal = 0;
    row.clear();
    for (int d = 0; d < dimension; ++d) {
      const Index coord = (i / strides[d]) % dims[d];
      for (int sign : {-1, 1}) {
        const Index j = i + sign * strides[d];
        const bool inside = coord + sign >= 0 && coord + sign < dims[d];
        double weight = d == 0 ? anisotropy : 1;
        if (random && inside) {
          auto key = hash(std::uint32_t(std::min(i, j)) ^ (hash(std::uint32_t(std::max(i, j))) + 0x9e3779b9U));
          weight *= 0.1 + 9.9 * double(key % 1000) / 999;
        }
        if (inside) row.emplace_back(j, -weight);
        if (inside || !neumann) diagonal += weight;
      }
    }
    if (wind != 0) {
      diagonal += wind;
      if (i % nx != 0) {
        auto left = std::find_if(row.begin(), row.end(), [i](auto e) { return e.first == i - 1; });
        left->second -= wind;
      }
    }
    row.emplace_back(i, diagonal);
    std::sort(row.begin(), row.end());
    for (auto [j, value] : row) {
      if (i == j) a.diag[i] = Index(a.col.size());
      a.col.push_back(j);
      a.val.push_back(static_cast<float>(value));
    }
    a.row_ptr.push_back(Index(a.col.size()));
  }
  return a;
}

}  // namespace oracle
