// Host orchestration of the device Krylov solvers and the aggregation-AMG
// V-cycle (linalg/solver.cpp:75-393, linalg/amg.cpp:176-342), expressed as
// kernel launches through the kutta_hip C ABI.
//
// The MI355X path fuses the reference's launch sequence where the fused kernel
// computes the identical floats: spmv+dot+alpha, step+dot+check, residual+
// restriction, prolongation+first post-sweep, and the last fine sweep with the
// r.z reduction and beta. set_fused(false) runs the reference sequence kernel
// by kernel instead; both give the same iterates.
#include "krylov.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <iomanip>
#include <sstream>

namespace kutta_amd {

OwnedCsr::OwnedCsr(HipDevice& dev, const CsrHost& a)
    : n_rows(a.n_rows), nnz(a.nnz()), row_ptr(to_device(dev, a.row_ptr)),
      col(to_device(dev, a.col)), diag(to_device(dev, a.diag)), val(dev, a.val.size()) {
  std::vector<float> v(a.val.size());
  for (std::size_t j = 0; j < v.size(); ++j) v[j] = static_cast<float>(a.val[j]);
  val.upload(v);
}

namespace {

// Normalized (hi, lo) pair of a double (runtime/cpu/cpu_device.cpp:83-86).
std::array<float, 2> pair_of(double value) {
  const float hi = static_cast<float>(value);
  return {hi, std::isfinite(hi) ? static_cast<float>(value - double(hi)) : 0.0f};
}

struct Level {
  Index rows = 0, nnz = 0;
  OwnedCsr matrix;  // coarse levels only; level 0 uses the caller's operator
  DeviceArray<float> l1, x, b, r, projection;
  bool has_transfer = false;
  DeviceArray<Index> aggregate, member_offsets, members, entry_offsets, entries;
  Index coarse_rows = 0;
};

// What the caller wants fused into the last sweep of the fine level.
enum class Tail { None, DotBeta };

}  // namespace

struct KrylovSolver::Impl {
  HipDevice& dev;
  SolverKind kind;
  AmgOptions options;
  bool amg, nullspace, fused = true;
  Index n = 0, nnz = 0;
  double complexity = 0;
  std::vector<Level> levels;
  DeviceArray<float> jacobi_inverse, jacobi_invalid;
  DeviceArray<float> r, z, p, q, shadow, s, t, zs, best;
  DeviceArray<float> scalars;
  DeviceArray<Index> status;
  DeviceArray<float> projection;
  double best_residual = 0;
  int previous_iterations = 0;
  bool verified_current = false;
  int waits = 0;
  // Coarse-grid tail: levels [tail_first, end) run as one KH_FUSED_COARSE_TAIL
  // launch (0: off). tail_request < 0 picks the level automatically.
  int tail_request = -1;
  int tail_first = 0;
  DeviceArray<kh_tail_level> tail_table;

  Impl(HipDevice& d, SolverKind k, const CsrHost& a, const AmgOptions& o)
      : dev(d), kind(k), options(o),
        amg(k == SolverKind::AmgPcg || k == SolverKind::AmgBicgstab || k == SolverKind::AmgFgmres),
        nullspace(o.nullspace == NullSpace::Constant), n(a.n_rows), nnz(a.nnz()) {
    validate_host(a);
    if (amg) {
      KA_REQUIRE(options.pre_sweeps > 0 && options.post_sweeps > 0 && options.coarse_sweeps > 0 &&
                     options.coarse_rows >= 1 && options.max_levels >= 1 &&
                     options.aggregation_passes >= 1 && options.aggregation_passes <= 8 &&
                     std::isfinite(options.strength_threshold) && options.strength_threshold >= 0 &&
                     options.strength_threshold <= 1,
                 "AMG: invalid options");
      KA_REQUIRE(k != SolverKind::AmgPcg || options.pre_sweeps == options.post_sweeps,
                 "AMG-PCG requires equal pre/post sweeps");
      KA_REQUIRE(k != SolverKind::AmgFgmres || options.nullspace == NullSpace::None,
                 "FGMRES: constant-nullspace operators are not supported");
      if (options.structure == Structure::SymmetricMMatrix) check_m_matrix(a);
      build_hierarchy(a);
      plan_tail();
    } else {
      jacobi_inverse = to_device(dev, diagonal_values(a, false, false));
      jacobi_invalid = DeviceArray<float>(dev, n);
    }
    r = DeviceArray<float>(dev, n);
    z = DeviceArray<float>(dev, n);
    p = DeviceArray<float>(dev, n);
    q = DeviceArray<float>(dev, n);
    best = DeviceArray<float>(dev, n);
    if (bicg()) {
      shadow = DeviceArray<float>(dev, n);
      s = DeviceArray<float>(dev, n);
      t = DeviceArray<float>(dev, n);
      zs = DeviceArray<float>(dev, n);
    }
    scalars = DeviceArray<float>(dev, 24);
    status = DeviceArray<Index>(dev, 3);
    projection = DeviceArray<float>(dev, 2);
    dev.fill(scalars.data(), 0.0f, 24);
    dev.fill_int(status.data(), 0, 3);
  }
  bool bicg() const { return kind == SolverKind::BicgstabJacobi || kind == SolverKind::AmgBicgstab; }
  bool gmres() const { return kind == SolverKind::FgmresJacobi || kind == SolverKind::AmgFgmres; }

  // Flexible GMRES(m) (Saad 1993), right preconditioned by the V-cycle or the
  // Jacobi inverse: the cycle's Arnoldi steps, rotations and convergence test
  // run on the device (kh_gmres.hip); the host polls every
  // residual_check_interval steps and verifies the true residual after each
  // cycle, restarting from it.
  // Mixed-precision AMG (no reference counterpart): level `direct_level` is
  // solved by an FP64 dense inverse (Gauss-Jordan at refresh, one GEMV per
  // V-cycle) instead of coarser levels and sweeps; 0 is off.
  int direct_level = 0;
  bool direct_ready = false;
  DeviceArray<double> dense_inverse, dense_scratch;
  int set_direct(Index rows) {
    direct_level = 0;
    direct_ready = false;
    dense_inverse = DeviceArray<double>();
    dense_scratch = DeviceArray<double>();
    if (rows <= 0) {
      plan_tail();
      return 0;
    }
    KA_REQUIRE(amg && options.structure == Structure::SymmetricMMatrix && !nullspace,
               "AMG: the dense coarse solve needs a symmetric M-matrix operator without nullspace");
    KA_REQUIRE(rows <= 8192, "AMG: the dense coarse solve takes at most 8192 rows");
    for (std::size_t i = 1; i < levels.size(); ++i)
      if (levels[i].rows <= rows) {
        direct_level = int(i);
        break;
      }
    KA_REQUIRE(direct_level > 0, "AMG: no coarse level has at most " + std::to_string(rows) + " rows");
    const std::size_t nc = std::size_t(levels[direct_level].rows);
    dense_inverse = DeviceArray<double>(dev, nc * nc);
    dense_scratch = DeviceArray<double>(dev, nc * nc);
    plan_tail();
    return direct_level;
  }
  int restart_length = 30;
  DeviceArray<float> basis_v, basis_z, gmres_w;
  DeviceArray<double> gmres_block, gmres_partials;
  std::size_t gmres_target_slot() const {
    const std::size_t m = std::size_t(restart_length);
    return (m + 1) * m + 5 * m + 2;
  }
  void gmres_buffers() {
    const std::size_t m = std::size_t(restart_length);
    if (basis_v.size() != (m + 1) * std::size_t(n)) {
      basis_v = DeviceArray<float>(dev, (m + 1) * std::size_t(n));
      basis_z = DeviceArray<float>(dev, m * std::size_t(n));
    }
    if (gmres_w.size() != std::size_t(n)) gmres_w = DeviceArray<float>(dev, n);
    if (gmres_block.size() != gmres_target_slot() + 1) gmres_block = DeviceArray<double>(dev, gmres_target_slot() + 1);
    const std::size_t blocks = (std::size_t(n) + 255) / 256;
    if (gmres_partials.size() != blocks * 64) gmres_partials = DeviceArray<double>(dev, std::max<std::size_t>(blocks * 64, 1));
  }
  void fgmres(const DeviceCsr& a, float* x, const float* b, const SolveControl& c, double target,
              SolveStats& st) {
    const int m = restart_length;
    gmres_buffers();
    {
      std::vector<double> host(gmres_block.size(), 0.0);
      host[gmres_target_slot()] = target;
      gmres_block.upload(host);
    }
    int restarts = 0;
    double last = st.initial_residual;
    while (st.iterations < c.max_iter) {
      reset(st.iterations);
      dev.launch(KH_GMRES_START, n, {r, basis_v, gmres_block, m, status, gmres_partials});
      const int steps = std::min(m, c.max_iter - st.iterations);
      verified_current = false;
      for (int j = 0; j < steps; ++j) {
        float* vj = basis_v.data() + std::size_t(j) * n;
        float* zj = basis_z.data() + std::size_t(j) * n;
        apply(a, vj, zj, Tail::None);
        dev.launch(KH_CSR_SPMV_GUARDED, n, {a.row_ptr, a.col, a.val, zj, gmres_w, status});
        dev.launch(KH_GMRES_ARNOLDI, n, {basis_v, gmres_w, gmres_block, m, j, status, gmres_partials});
        if ((j + 1) % c.residual_check_interval == 0 && j + 1 < steps && poll()[0] != 0) break;
      }
      dev.launch(KH_GMRES_UPDATE, n, {basis_z, x, gmres_block, m, status});
      st.iterations = poll()[1];
      if (verified(a, x, b, target, st) || diverged(st)) break;
      restarts = st.final_residual >= 0.9 * last ? restarts + 1 : 0;
      last = st.final_residual;
      if (restarts >= 12) break;
    }
  }

  // linalg/amg.cpp:200-226
  void build_hierarchy(const CsrHost& a) {
    levels.emplace_back();
    levels.back().rows = a.n_rows;
    levels.back().nnz = a.nnz();
    levels.back().l1 = to_device(dev, diagonal_values(a, true, nullspace));
    levels.back().r = DeviceArray<float>(dev, a.n_rows);
    if (nullspace) levels.back().projection = DeviceArray<float>(dev, 2);
    const CsrHost* current = &a;
    CsrHost owned;
    std::size_t total = a.nnz();
    while (current->n_rows > options.coarse_rows && levels.size() < std::size_t(options.max_levels)) {
      Coarsening c = aggregate(*current, options);
      const auto active = c.members.size();
      if (active == 0 || c.matrix.n_rows >= 0.85 * double(active)) break;
      Level& fine = levels.back();
      fine.has_transfer = true;
      fine.coarse_rows = c.matrix.n_rows;
      fine.aggregate = to_device(dev, c.aggregate);
      fine.member_offsets = to_device(dev, c.member_offsets);
      fine.members = to_device(dev, c.members);
      fine.entry_offsets = to_device(dev, c.entry_offsets);
      fine.entries = to_device(dev, c.entries);
      total += c.matrix.nnz();
      Level coarse;
      coarse.rows = c.matrix.n_rows;
      coarse.nnz = c.matrix.nnz();
      coarse.l1 = to_device(dev, diagonal_values(c.matrix, true, nullspace));
      coarse.x = DeviceArray<float>(dev, c.matrix.n_rows);
      coarse.b = DeviceArray<float>(dev, c.matrix.n_rows);
      coarse.r = DeviceArray<float>(dev, c.matrix.n_rows);
      if (nullspace) coarse.projection = DeviceArray<float>(dev, 2);
      coarse.matrix = OwnedCsr(dev, c.matrix);
      owned = std::move(c.matrix);
      levels.push_back(std::move(coarse));
      current = &owned;
    }
    complexity = a.nnz() ? double(total) / a.nnz() : 0;
  }

  // Whether levels [first, end) fit one KH_FUSED_COARSE_TAIL workgroup: all
  // iterates and right-hand sides, one sweep's new values and the largest
  // level's operator copy in LDS at once.
  bool tail_fits(std::size_t first) const {
    if (first < 1 || first >= levels.size() || levels.size() - first > KH_TAIL_MAX_LEVELS) return false;
    std::size_t vectors = 0, widest = 0, largest = 0;
    for (std::size_t i = first; i < levels.size(); ++i) {
      const std::size_t rows = std::size_t(levels[i].rows), nnz = std::size_t(levels[i].nnz);
      if (rows > KH_TAIL_MAX_ROWS) return false;
      vectors += 2 * rows;
      widest = std::max(widest, rows);
      largest = std::max(largest, 3 * rows + 1 + 2 * nnz);
    }
    return vectors <= KH_TAIL_LDS_FLOATS && vectors + widest + largest <= KH_TAIL_POOL_WORDS;
  }

  // Picks the tail's first level: the requested one, or else the first level
  // of at most kTailAutoRows rows from which the rest of the hierarchy fits
  // one workgroup. Above that size one CU's compensated sweeps cost more than
  // a level's own launches, whose sweeps spread over the whole chip.
  static constexpr Index kTailAutoRows = 2048;
  void plan_tail() {
    tail_first = 0;
    tail_table = DeviceArray<kh_tail_level>();
    if (!amg || tail_request == 0 || direct_level > 0) return;
    if (tail_request > 0) {
      KA_REQUIRE(tail_fits(std::size_t(tail_request)),
                 "AMG: coarse tail level " + std::to_string(tail_request) + " does not fit one workgroup");
      tail_first = tail_request;
    } else {
      for (std::size_t i = 1; i < levels.size(); ++i)
        if (levels[i].rows <= kTailAutoRows && tail_fits(i)) {
          tail_first = int(i);
          break;
        }
    }
    if (tail_first == 0) return;
    std::vector<kh_tail_level> table;
    for (std::size_t i = std::size_t(tail_first); i < levels.size(); ++i) {
      const Level& l = levels[i];
      const DeviceCsr m = l.matrix.view();
      kh_tail_level t{};
      t.row_ptr = m.row_ptr;
      t.col = m.col;
      t.val = m.val;
      t.inverse = l.l1.data();
      if (l.has_transfer) {
        t.member_offsets = l.member_offsets.data();
        t.members = l.members.data();
        t.aggregate = l.aggregate.data();
        t.coarse_rows = l.coarse_rows;
      }
      t.rows = l.rows;
      t.nnz = l.nnz;
      table.push_back(t);
    }
    tail_table = to_device(dev, table);
  }

  DeviceCsr level_operator(std::size_t i, const DeviceCsr& fine) const {
    return i == 0 ? fine : levels[i].matrix.view();
  }

  // linalg/amg.cpp:228-256
  void refresh_amg(const DeviceCsr& a) {
    std::vector<kh_reduction> checks;
    std::vector<double> allowed;
    for (std::size_t i = 0; i < levels.size(); ++i) {
      Level& l = levels[i];
      const DeviceCsr m = level_operator(i, a);
      dev.launch(KH_KRYLOV_DIAGONAL, m.n_rows, {m.row_ptr, m.diag, m.val, l.l1, l.r, 1});
      allowed.push_back(nullspace && m.n_rows == 1 ? 1 : 0);
      checks.push_back(HipDevice::max_abs(l.r.data(), m.n_rows));
      if (i + 1 < levels.size()) {
        const DeviceCsr c = levels[i + 1].matrix.view();
        dev.launch(KH_AMG_GALERKIN, c.nnz, {l.entry_offsets, l.entries, m.val, c.val});
        if (options.structure == Structure::SymmetricMMatrix)
          dev.launch(KH_AMG_M_MATRIX_DIAGONAL, c.n_rows,
                     {l.member_offsets, l.members, m.row_ptr, m.diag, m.val, c.row_ptr, c.diag, c.val});
      }
    }
    const auto invalid = dev.reduce_many(checks);
    ++waits;
    for (std::size_t i = 0; i < invalid.size(); ++i) {
      KA_REQUIRE(invalid[i] != 3, "AMG: requires positive diagonals");
      KA_REQUIRE(invalid[i] <= allowed[i], "AMG: invalid values or zero diagonal");
    }
    if (direct_level > 0) {
      const DeviceCsr c = levels[direct_level].matrix.view();
      dev.launch(KH_DENSE_INVERT, c.n_rows, {c.row_ptr, c.col, c.val, dense_inverse, dense_scratch});
      direct_ready = true;
    }
  }

  // Constant-nullspace projection (linalg/detail.hpp:13-19).
  void project(float* x, std::size_t count, float* sum) {
    if (count == 0) return;
    dev.reduce_into(KH_INTO_SUM, x, x, x, x, count, sum, 0);
    project_with(x, count, sum);
  }
  void project_with(float* x, std::size_t count, float* sum) {
    const float c = static_cast<float>(count);
    dev.launch(KH_SOLVER_PROJECT, count,
               {x, sum, c, static_cast<float>(double(count) - double(c))});
  }

  // One l1-Jacobi sweep (linalg/amg.cpp:272-282). `first` marks the sweep that
  // starts from the zero iterate the cycle's fill would have stored.
  void sweep(Level& l, const DeviceCsr& a, const float* b, const float* in, float* out, bool first,
             float* zero_x) {
    const int* st = status.data();
    if (first) {
      if (fused) {
        dev.launch(KH_FUSED_SMOOTH_FROM_ZERO, a.n_rows, {a.row_ptr, a.col, a.val, b, l.l1, out, st});
      } else {
        dev.launch(KH_AMG_SMOOTH_ZERO, a.n_rows, {a.row_ptr, a.col, a.val, zero_x, b, l.l1, out, st});
      }
    } else {
      dev.launch(KH_AMG_SMOOTH, a.n_rows, {a.row_ptr, a.col, a.val, in, b, l.l1, out, st});
    }
  }

  // What a V-cycle completed beyond z = M^-1 b.
  struct CycleResult {
    bool dot_beta = false;   // dot_into(r,z,4) and solver_pcg_beta already done
    bool sum_ready = false;  // the fine-level projection sum is already in `projection`
  };

  // linalg/amg.cpp:283-310. Sweep outputs alternate between x and the level's
  // scratch so that the last sweep writes x.
  CycleResult cycle(std::size_t index, const DeviceCsr& a, const float* b, float* x, Tail tail) {
    Level& l = levels[index];
    float* w = l.r.data();
    const int* st = status.data();
    CycleResult result;
    if (fused && tail_first > 0 && int(index) == tail_first) {
      dev.launch(KH_FUSED_COARSE_TAIL, 1,
                 {tail_table, int(levels.size() - index), options.pre_sweeps, options.post_sweeps,
                  options.coarse_sweeps, nullspace ? 1 : 0, b, x, st});
      return result;
    }
    if (direct_level > 0 && int(index) == direct_level) {
      KA_REQUIRE(direct_ready, "AMG: refresh before solving with the dense coarse solve");
      dev.launch(KH_DENSE_APPLY, a.n_rows, {dense_inverse, b, x, st});
      return result;
    }
    // The reference clears x first; only the unfused sequence reads that zero.
    if (!fused) dev.fill(x, 0.0f, a.n_rows);
    const bool coarsest = index + 1 == levels.size();
    float* sum = index == 0 ? projection.data() : (nullspace ? l.projection.data() : nullptr);
    if (coarsest && a.n_rows <= 24) {
      dev.launch(KH_AMG_COARSE_SOLVE, 1,
                 {a.row_ptr, a.col, a.val, x, b, l.l1, st, int(a.n_rows), options.coarse_sweeps});
      if (nullspace && index != 0) project(x, a.n_rows, sum);
      return result;
    }
    const int pre = coarsest ? options.coarse_sweeps : options.pre_sweeps;
    const int post = coarsest ? 0 : options.post_sweeps;
    const int total = pre + post;
    // The final sweep is a plain smoothing sweep (not the first sweep, not the
    // prolongation sweep), so a reduction can ride along with it.
    const bool last_plain = total >= 2 && (coarsest || post >= 2);
    const bool fuse_dot = fused && index == 0 && tail == Tail::DotBeta && !nullspace && last_plain;
    const bool fuse_sum = fused && nullspace && last_plain;
    auto target = [&](int t) { return (total - t) % 2 == 0 ? x : w; };
    const float* current = nullptr;
    auto run_sweep = [&](int t, float* out) {
      if (t == total && fuse_dot) {
        dev.launch(KH_FUSED_SMOOTH_DOT_BETA, a.n_rows,
                   {a.row_ptr, a.col, a.val, current, b, l.l1, out, status.data(), scalars.data()});
        result.dot_beta = true;
      } else if (t == total && fuse_sum) {
        dev.launch(KH_FUSED_SMOOTH_SUM, a.n_rows, {a.row_ptr, a.col, a.val, current, b, l.l1, out, st, sum});
      } else {
        sweep(l, a, b, current, out, t == 1, x);
      }
      current = out;
    };
    for (int t = 1; t <= pre; ++t) run_sweep(t, target(t));
    if (!coarsest) {
      Level& c = levels[index + 1];
      const DeviceCsr ca = c.matrix.view();
      if (fused) {
        dev.launch(KH_FUSED_RESIDUAL_RESTRICT, ca.n_rows,
                   {a.row_ptr, a.col, a.val, current, b, l.member_offsets, l.members, c.b, st});
      } else {
        float* rbuf = current == x ? w : x;
        dev.launch(KH_CSR_RESIDUAL_GUARDED, a.n_rows, {a.row_ptr, a.col, a.val, current, b, rbuf, st});
        dev.launch(KH_AMG_RESTRICT_GUARDED, ca.n_rows, {l.member_offsets, l.members, rbuf, c.b, st});
      }
      if (nullspace) project(c.b.data(), ca.n_rows, c.projection.data());
      cycle(index + 1, ca, c.b.data(), c.x.data(), Tail::None);
      for (int t = pre + 1; t <= total; ++t) {
        float* out = target(t);
        if (t == pre + 1) {
          if (fused) {
            dev.launch(KH_FUSED_PROLONG_SMOOTH, a.n_rows,
                       {a.row_ptr, a.col, a.val, current, b, l.l1, out, l.aggregate, c.x, st});
          } else {
            float* xin = const_cast<float*>(current);
            dev.launch(KH_AMG_PROLONG_GUARDED, a.n_rows, {l.aggregate, c.x, xin, st});
            dev.launch(KH_AMG_SMOOTH, a.n_rows, {a.row_ptr, a.col, a.val, xin, b, l.l1, out, st});
          }
          current = out;
        } else {
          run_sweep(t, out);
        }
      }
    }
    if (nullspace) {
      if (index != 0) {
        if (fuse_sum) project_with(x, a.n_rows, sum);
        else project(x, a.n_rows, sum);
      } else {
        result.sum_ready = fuse_sum;
      }
    }
    return result;
  }

  // Preconditioner application (linalg/solver.cpp:197-200). With tail ==
  // DotBeta it also performs dot_into(r,z,4) + solver_pcg_beta(0), as the PCG
  // loop does right after apply (linalg/solver.cpp:295-298).
  void apply(const DeviceCsr& a, const float* rin, float* zout, Tail tail) {
    CycleResult done;
    if (amg) {
      done = cycle(0, a, rin, zout, tail);
    } else {
      dev.launch(KH_KRYLOV_MULTIPLY_GUARDED, n, {rin, jacobi_inverse, zout, status});
    }
    if (done.dot_beta) return;
    if (nullspace) {
      if (!done.sum_ready) dev.reduce_into(KH_INTO_SUM, zout, zout, zout, zout, n, projection.data(), 0);
      if (tail == Tail::DotBeta && fused) {
        const float c = static_cast<float>(n);
        dev.launch(KH_FUSED_PROJECT_DOT_BETA, n,
                   {zout, projection, c, static_cast<float>(double(n) - double(c)), r, scalars, status});
        return;
      }
      project_with(zout, n, projection.data());
    }
    if (tail == Tail::DotBeta) {
      dot_into(r.data(), zout, 4);
      dev.launch(KH_SOLVER_PCG_BETA, 1, {scalars, status, 0});
    }
  }

  void residual(const DeviceCsr& a, const float* x, const float* b, float* out) {
    dev.launch(KH_CSR_RESIDUAL, a.n_rows, {a.row_ptr, a.col, a.val, x, b, out});
  }
  double norm(const float* v) {
    ++waits;
    return std::sqrt(dev.reduce_many({HipDevice::dot(v, v, n)})[0]);
  }
  // linalg/solver.cpp:201-222
  bool verified(const DeviceCsr& a, float* x, const float* b, double target, SolveStats& st) {
    if (nullspace) project(x, n, projection.data());
    residual(a, x, b, r.data());
    st.final_residual = norm(r.data());
    verified_current = true;
    st.converged = std::isfinite(st.final_residual) && st.final_residual <= target;
    if (st.final_residual < best_residual) {
      best_residual = st.final_residual;
      dev.copy(best.data(), x, sizeof(float) * n);
    }
    if (nullspace) project(r.data(), n, projection.data());
    return st.converged;
  }
  static bool diverged(const SolveStats& st) {
    return !std::isfinite(st.final_residual) || st.final_residual > 1e3 * st.initial_residual;
  }
  void reset(int iterations) { dev.launch(KH_SOLVER_RESET, 1, {scalars, status, iterations}); }
  bool replacement(const DeviceCsr& a, float* x, const float* b, double target, SolveStats& st,
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
  std::array<Index, 3> poll() {
    ++waits;
    const auto host = status.download();
    return {host[0], host[1], host[2]};
  }
  void dot_into(const float* x, const float* y, int offset) {
    dev.reduce_into(KH_INTO_DOT, x, y, x, y, n, scalars.data(), offset);
  }

  // linalg/solver.cpp:271-313
  void pcg(const DeviceCsr& a, float* x, const float* b, const SolveControl& c, double target,
           SolveStats& st) {
    auto restart = [&] {
      reset(st.iterations);
      apply(a, r.data(), z.data(), Tail::None);
      dev.copy(p.data(), z.data(), sizeof(float) * n);
      dot_into(r.data(), z.data(), 0);
    };
    restart();
    int restarts = 0;
    double last = st.initial_residual;
    int interval = batch_interval(c);
    while (st.iterations < c.max_iter) {
      const int batch = std::min(interval, c.max_iter - st.iterations);
      verified_current = false;
      for (int j = 0; j < batch; ++j) {
        if (fused) {
          dev.launch(KH_FUSED_SPMV_DOT_ALPHA, n, {a.row_ptr, a.col, a.val, p, q, scalars, status});
          dev.launch(KH_FUSED_STEP_DOT_CHECK, n, {x, r, p, q, scalars, status});
        } else {
          dev.launch(KH_CSR_SPMV_GUARDED, n, {a.row_ptr, a.col, a.val, p, q, status});
          dot_into(p.data(), q.data(), 2);
          dev.launch(KH_SOLVER_PCG_ALPHA, 1, {scalars, status});
          dev.launch(KH_SOLVER_STEP, n, {x, r, p, q, scalars, status});
          dot_into(r.data(), r.data(), 18);
          dev.launch(KH_SOLVER_CHECK, 1, {scalars, status, 1});
        }
        if (st.iterations + j + 1 == c.max_iter) continue;
        apply(a, r.data(), z.data(), Tail::DotBeta);
        dev.launch(KH_SOLVER_DIRECTION, n, {z, p, scalars, status});
      }
      const auto state = poll();
      st.iterations = state[1];
      if (state[0] == 2) break;
      if (state[0] == 1) {
        if (!replacement(a, x, b, target, st, restarts, last) || st.iterations == c.max_iter) break;
        if (st.final_residual <= 10 * target) interval = 1;
        restart();
      }
    }
  }

  // linalg/solver.cpp:314-361. With the Jacobi preconditioner and fused
  // kernels an iteration is five launches (each with its reduction and scalar
  // update); the first rho and beta after a restart are one more.
  void bicgstab(const DeviceCsr& a, float* x, const float* b, const SolveControl& c, double target,
                SolveStats& st) {
    const bool jacobi_fused = fused && !amg;
    auto restart = [&] {
      reset(st.iterations);
      dev.copy(shadow.data(), r.data(), sizeof(float) * n);
      dev.fill(p.data(), 0.0f, n);
      dev.fill(q.data(), 0.0f, n);
      if (jacobi_fused) dev.launch(KH_FUSED_DOT_BICG_BETA, n, {shadow, r, scalars, status});
    };
    restart();
    int restarts = 0;
    double last = st.initial_residual;
    int interval = batch_interval(c);
    while (st.iterations < c.max_iter) {
      const int batch = std::min(interval, c.max_iter - st.iterations);
      verified_current = false;
      for (int j = 0; j < batch; ++j) {
        if (jacobi_fused) {
          dev.launch(KH_FUSED_BICG_DIRECTION, n, {r, q, p, jacobi_inverse, z, scalars, status});
          dev.launch(KH_FUSED_SPMV_SHADOW_ALPHA, n, {a.row_ptr, a.col, a.val, z, q, shadow, scalars, status});
          dev.launch(KH_FUSED_BICG_DIFFERENCE_S, n, {r, q, s, jacobi_inverse, zs, scalars, status});
          dev.launch(KH_FUSED_SPMV_DOT2_OMEGA, n, {a.row_ptr, a.col, a.val, zs, t, s, scalars, status});
          dev.launch(KH_FUSED_BICG_STEP_CHECK_BETA, n, {x, r, z, zs, s, t, shadow, scalars, status});
          continue;
        }
        dot_into(shadow.data(), r.data(), 0);
        dev.launch(KH_SOLVER_BICG_BETA, 1, {scalars, status});
        dev.launch(KH_SOLVER_BICG_DIRECTION, n, {r, q, p, scalars, status});
        apply(a, p.data(), z.data(), Tail::None);
        dev.launch(KH_CSR_SPMV_GUARDED, n, {a.row_ptr, a.col, a.val, z, q, status});
        dot_into(shadow.data(), q.data(), 2);
        dev.launch(KH_SOLVER_BICG_ALPHA, 1, {scalars, status});
        dev.launch(KH_SOLVER_BICG_DIFFERENCE, n, {r, q, s, scalars, status});
        dot_into(s.data(), s.data(), 16);
        dev.launch(KH_SOLVER_BICG_S, 1, {scalars, status});
        apply(a, s.data(), zs.data(), Tail::None);
        dev.launch(KH_CSR_SPMV_GUARDED, n, {a.row_ptr, a.col, a.val, zs, t, status});
        dev.reduce_into(KH_INTO_DOT2, t.data(), t.data(), t.data(), s.data(), n, scalars.data(), 2);
        dev.launch(KH_SOLVER_BICG_OMEGA, 1, {scalars, status});
        dev.launch(KH_SOLVER_BICG_STEP, n, {x, r, z, zs, s, t, scalars, status});
        dot_into(r.data(), r.data(), 18);
        dev.launch(KH_SOLVER_CHECK, 1, {scalars, status, 0});
      }
      const auto state = poll();
      st.iterations = state[1];
      if (state[0] == 2) break;
      if (state[0] == 1) {
        if (!replacement(a, x, b, target, st, restarts, last) || st.iterations == c.max_iter) break;
        if (st.final_residual <= 10 * target) interval = 1;
        restart();
      }
    }
  }

  // linalg/solver.cpp:97-176
  SolveStats solve(const DeviceCsr& a, float* x, const float* b, const SolveControl& c) {
    KA_REQUIRE(a.n_rows == n && a.nnz == nnz, "solver: device or matrix mismatch");
    KA_REQUIRE(x != nullptr && b != nullptr && x != b, "solve: vector size or alias mismatch");
    KA_REQUIRE(c.max_iter >= 0 && std::isfinite(c.rel_tol) && c.rel_tol >= 0 &&
                   std::isfinite(c.abs_tol) && c.abs_tol >= 0 && c.residual_check_interval > 0,
               "solve: invalid control");
    if (!nullspace) residual(a, x, b, r.data());
    ++waits;
    const auto entry = dev.reduce_many({HipDevice::dot(b, b, n), HipDevice::dot(x, x, n),
                                        nullspace ? HipDevice::sum(b, n) : HipDevice::dot(r.data(), r.data(), n)});
    const double bnorm = std::sqrt(entry[0]);
    if (!(std::isfinite(bnorm) && std::isfinite(std::sqrt(entry[1]))))
      throw NonfiniteError("solve: nonfinite input vector");
    SolveStats st;
    if (nullspace) {
      if (n != 0) {
        KA_REQUIRE(std::abs(entry[2]) <= 1e-6 * std::sqrt(double(n)) * bnorm,
                   "solve: incompatible RHS for constant Neumann nullspace");
        project(x, n, projection.data());
      }
      residual(a, x, b, r.data());
      st.initial_residual = norm(r.data());
    } else st.initial_residual = std::sqrt(entry[2]);
    st.final_residual = st.initial_residual;
    verified_current = true;
    const double target = std::max(c.abs_tol, c.rel_tol * st.initial_residual);
    if (!std::isfinite(st.initial_residual)) return st;
    if (st.final_residual <= target) {
      st.converged = true;
      return st;
    }
    if (nullspace) project(r.data(), n, projection.data());
    best_residual = st.initial_residual;
    dev.copy(best.data(), x, sizeof(float) * n);
    if (kind != SolverKind::Jacobi) {
      const auto tol = pair_of(target * target);
      const auto div = pair_of(1e6 * st.initial_residual * st.initial_residual);
      dev.launch(KH_SOLVER_LIMITS, 24, {scalars, tol[0], tol[1], div[0], div[1]});
    }
    reset(0);
    if (kind == SolverKind::PcgJacobi || kind == SolverKind::AmgPcg) pcg(a, x, b, c, target, st);
    else if (bicg()) bicgstab(a, x, b, c, target, st);
    else if (gmres()) fgmres(a, x, b, c, target, st);
    else jacobi(a, x, b, c, target, st);
    if (!verified_current || nullspace) verified(a, x, b, target, st);
    if (!std::isfinite(st.final_residual) || st.final_residual > best_residual) {
      dev.copy(x, best.data(), sizeof(float) * n);
      verified(a, x, b, target, st);
    }
    previous_iterations = st.converged ? st.iterations : 0;
    return st;
  }

  // linalg/solver.cpp:151-160: damped Jacobi sweeps.
  void jacobi(const DeviceCsr& a, float* x, const float* b, const SolveControl& c, double target,
              SolveStats& st) {
    for (int i = 0; i < c.max_iter; ++i) {
      apply(a, r.data(), z.data(), Tail::None);
      dev.launch(KH_AXPY, n, {x, z, 2.0f / 3.0f});
      verified_current = false;
      ++st.iterations;
      if (st.iterations % c.residual_check_interval == 0 || st.iterations == c.max_iter) {
        if (verified(a, x, b, target, st) || diverged(st)) break;
      } else {
        residual(a, x, b, r.data());
      }
    }
  }

  void refresh(const DeviceCsr& a) {
    KA_REQUIRE(a.n_rows == n && a.nnz == nnz, "solver: device or matrix mismatch");
    if (amg) {
      refresh_amg(a);
    } else {
      dev.launch(KH_KRYLOV_DIAGONAL, n, {a.row_ptr, a.diag, a.val, jacobi_inverse, jacobi_invalid, 0});
      ++waits;
      KA_REQUIRE(dev.reduce_many({HipDevice::max_abs(jacobi_invalid.data(), n)})[0] == 0,
                 "Jacobi: invalid values or zero diagonal");
    }
  }
};

KrylovSolver::KrylovSolver(HipDevice& dev, SolverKind kind, const CsrHost& a, const AmgOptions& options)
    : impl_(std::make_unique<Impl>(dev, kind, a, options)) {}
KrylovSolver::~KrylovSolver() = default;
void KrylovSolver::refresh(const DeviceCsr& a) { impl_->refresh(a); }
SolveStats KrylovSolver::solve(const DeviceCsr& a, float* x, const float* b, const SolveControl& c) {
  return impl_->solve(a, x, b, c);
}
void KrylovSolver::set_fused(bool fused) { impl_->fused = fused; }
int KrylovSolver::set_tail(int first) {
  const int previous = impl_->tail_request;
  impl_->tail_request = first;
  try {
    impl_->plan_tail();
  } catch (...) {
    impl_->tail_request = previous;
    impl_->plan_tail();
    throw;
  }
  return impl_->tail_first;
}
int KrylovSolver::tail_level() const { return impl_->tail_first; }
int KrylovSolver::host_waits() const { return impl_->waits; }
int KrylovSolver::set_direct(Index rows) { return impl_->set_direct(rows); }
void KrylovSolver::set_restart(int m) {
  KA_REQUIRE(m >= 1 && m <= 63, "FGMRES: restart length must be in [1, 63]");
  impl_->restart_length = m;
}

// linalg/amg.cpp:263-269, linalg/solver.cpp:178-184
std::string KrylovSolver::describe() const {
  std::ostringstream out;
  const bool pcg = impl_->kind == SolverKind::PcgJacobi || impl_->kind == SolverKind::AmgPcg;
  out << (impl_->kind == SolverKind::Jacobi ? "damped Jacobi"
          : pcg                             ? "PCG"
          : impl_->gmres()                  ? "FGMRES(" + std::to_string(impl_->restart_length) + ")"
                                            : "BiCGStab")
      << ", ";
  if (impl_->amg)
    out << "aggregation AMG V-cycle, " << impl_->levels.size() << " levels, op complexity "
        << std::fixed << std::setprecision(3) << impl_->complexity;
  else out << "Jacobi";
  if (impl_->nullspace) out << ", constant-nullspace projection";
  if (impl_->direct_level > 0)
    out << ", FP64 dense solve of level " << impl_->direct_level << " ("
        << impl_->levels[impl_->direct_level].rows << " rows)";
  return out.str();
}
std::vector<std::pair<Index, Index>> KrylovSolver::levels() const {
  std::vector<std::pair<Index, Index>> out;
  for (const auto& l : impl_->levels) out.emplace_back(l.rows, l.nnz);
  return out;
}
std::vector<Index> KrylovSolver::transfer(std::size_t level, int which) const {
  if (level >= impl_->levels.size() || !impl_->levels[level].has_transfer) return {};
  const auto& l = impl_->levels[level];
  const DeviceArray<Index>& a = which == 0 ? l.aggregate : which == 1 ? l.member_offsets
      : which == 2 ? l.members : which == 3 ? l.entry_offsets : l.entries;
  return a.download();
}
std::vector<float> KrylovSolver::level_values(std::size_t level) const {
  if (level == 0 || level >= impl_->levels.size()) return {};
  return impl_->levels[level].matrix.val.download();
}

}  // namespace kutta_amd
