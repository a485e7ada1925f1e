// Device-resident Krylov solvers with the aggregation-AMG preconditioner for
// the MI355X backend: the HIP-native counterpart of make_amg_solver /
// make_solver and LinearSolver (linalg/solver.hpp:223-293, linalg/amg.hpp:8-27).
// Iteration state (scalars and status) never leaves the device between polls;
// the host only reads a 3-int status once per batch.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "amg_setup.hpp"

namespace kutta_amd {

// linalg/solver.hpp:223-235
struct SolveControl {
  double rel_tol = 1e-2;
  double abs_tol = 1e-12;
  int max_iter = 200;
  int residual_check_interval = 8;
};
struct SolveStats {
  int iterations = 0;
  double initial_residual = 0;
  double final_residual = 0;
  bool converged = false;
};
// linalg/solver.hpp:237-243
enum class SolverKind : std::uint8_t { Jacobi, PcgJacobi, BicgstabJacobi, AmgPcg, AmgBicgstab,
                                       // Flexible GMRES(m) (no reference counterpart).
                                       FgmresJacobi, AmgFgmres };

// Non-owning device CSR view (linalg/csr.hpp:31-38).
struct DeviceCsr {
  Index n_rows = 0, nnz = 0;
  const Index* row_ptr = nullptr;
  const Index* col = nullptr;
  const Index* diag = nullptr;
  float* val = nullptr;
};
// Owning device CSR.
struct OwnedCsr {
  Index n_rows = 0, nnz = 0;
  DeviceArray<Index> row_ptr, col, diag;
  DeviceArray<float> val;
  OwnedCsr() = default;
  OwnedCsr(HipDevice& dev, const CsrHost& a);
  DeviceCsr view() const { return {n_rows, nnz, row_ptr.data(), col.data(), diag.data(), val.data()}; }
};

// State of the learned coarse correction stage (no reference counterpart).
struct LearnedReport {
  int level = 0, rows = 0, mode = 0, harvested = 0, targets = 0;
  long applications = 0, fallbacks = 0, solves = 0, gated = 0;
  double last_loss = 0;
  double relative = -1;  // unexplained fraction of the last epoch (< 0: not trained)
};
struct TrainResult;

class KrylovSolver {
 public:
  KrylovSolver(HipDevice& dev, SolverKind kind, const CsrHost& a, const AmgOptions& options);
  ~KrylovSolver();
  // Recomputes value-dependent state (l1 diagonals, coarse operators) for new
  // values in the same pattern (linalg/amg.cpp:228-256).
  void refresh(const DeviceCsr& a);
  // x and b are device vectors of n_rows floats (linalg/solver.cpp:97-176).
  SolveStats solve(const DeviceCsr& a, float* x, const float* b, const SolveControl& control);
  std::string describe() const;
  // The V-cycle forms the autotuner picked (restriction per level, coarsest
  // sweeps, tail level); no reference counterpart.
  std::string plan() const;
  // Sets every level's restriction form (0 fused, 1 wide, 2 split, 3 split
  // with a wavefront per residual row; -1 keeps), the coarsest level's sweeps
  // (0 a launch each, 1 one workgroup, 2 one workgroup reading the operator
  // once; -1 keeps), the prolongation form (0 folded, 1 split, 2 folded with
  // a wavefront per row) and the long-row sweep form (0 thread, 1 wavefront).
  void set_plan(int restrict_mode, int coarse_sweeps, int prolong_mode = -1, int sweep_mode = -1);
  std::vector<std::pair<Index, Index>> levels() const;  // (rows, nnz) per level
  // Host copies for parity checks: transfer arrays and coarse values.
  std::vector<Index> transfer(std::size_t level, int which) const;
  std::vector<float> level_values(std::size_t level) const;
  // Fused-kernel use can be disabled to run the reference kernel sequence.
  void set_fused(bool fused);
  // Replay each fused PCG iteration from a captured HIP graph (default on;
  // identical iterates). No reference counterpart.
  void set_graphs(bool on);
  bool graphs_enabled() const;
  // Solves up to KH_BATCH_MAX_LANES independent systems with fused
  // BiCGStab-Jacobi solvers at once (one launch per step for all of them):
  // each system's iterates, statistics and x are bitwise those of
  // solvers[k]->solve(a[k], x[k], b[k], control). No reference counterpart.
  static std::vector<SolveStats> solve_batched(const std::vector<KrylovSolver*>& solvers,
                                               const std::vector<DeviceCsr>& a, const std::vector<float*>& x,
                                               const std::vector<const float*>& b, const SolveControl& control);
  // First level of the one-workgroup coarse-grid tail: < 0 automatic, 0 off.
  // Returns the level in use (0: none).
  int set_tail(int first);
  int tail_level() const;
  // Times `repeats` V-cycles per admissible tail level (and none) on the
  // refreshed operator and keeps the fastest; returns it. Identical iterates
  // either way. No reference counterpart.
  int autotune_tail(const DeviceCsr& a, int repeats);
  // Restart length of the FGMRES kinds (default 30, at most 63).
  void set_restart(int m);
  // Mixed-precision AMG: the first coarse level of at most `rows` rows is
  // solved by an FP64 dense inverse, computed at the next refresh(); 0 turns
  // it off. Returns that level (0: none). No reference counterpart.
  int set_direct(Index rows);
  // Lazy dense refresh (no reference counterpart): a refresh recomputes the
  // dense inverse only when the level's values have moved more than
  // `tolerance` (relative l2) from those it was computed from; 0 (default)
  // recomputes at every refresh. The counts of inversions and kept inverses
  // and the last relative distance are reported by dense_refreshes().
  void set_dense_tolerance(double tolerance);
  void dense_refreshes(long* inversions, long* skips, double* staleness) const;
  // Mixed-precision AMG: coarse levels smoothed, restricted and prolonged
  // from bfloat16 copies of their values (FP32 arithmetic), made at the next
  // refresh(). No reference counterpart.
  void set_coarse_bf16(bool on);
  // Learned coarse correction stage of the AMG FGMRES kind (no reference
  // counterpart): a message-passing network of `layers` layers on coarse
  // level `level`, with room for `capacity` harvested pairs.
  void set_learned(int level, int layers, std::uint64_t seed, int capacity);
  // 0 off, 1 harvest (plain V-cycle steps record restricted residuals),
  // 2 apply; `reduction` is the least average fall of the residual estimate
  // per step that keeps the stage on within a solve, and a network whose
  // last training epoch left more than `max_relative` of its targets' mean
  // square unexplained is not applied.
  void set_learned_mode(int mode, double reduction = 0.5, double max_relative = 0.5);
  // Exact level corrections of the harvested residuals; returns the pairs.
  int learned_targets(double rel_tol);
  // Adam steps over the pairs: the loss and unexplained fraction per step.
  // Adam steps over the first `pairs` harvested pairs (0: all of them).
  TrainResult learned_train(int steps, float lr, int pairs = 0);
  // Mean unexplained fraction on pairs [first, first + count), not trained.
  double learned_evaluate(int first, int count);
  LearnedReport learned_report() const;
  std::vector<float> learned_parameters() const;
  void set_learned_parameters(const std::vector<float>& theta);
  int host_waits() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace kutta_amd
