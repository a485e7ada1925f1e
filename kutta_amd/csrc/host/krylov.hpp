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
  std::vector<std::pair<Index, Index>> levels() const;  // (rows, nnz) per level
  // Host copies for parity checks: transfer arrays and coarse values.
  std::vector<Index> transfer(std::size_t level, int which) const;
  std::vector<float> level_values(std::size_t level) const;
  // Fused-kernel use can be disabled to run the reference kernel sequence.
  void set_fused(bool fused);
  // First level of the one-workgroup coarse-grid tail: < 0 automatic, 0 off.
  // Returns the level in use (0: none).
  int set_tail(int first);
  int tail_level() const;
  // Restart length of the FGMRES kinds (default 30, at most 63).
  void set_restart(int m);
  // Mixed-precision AMG: the first coarse level of at most `rows` rows is
  // solved by an FP64 dense inverse, computed at the next refresh(); 0 turns
  // it off. Returns that level (0: none). No reference counterpart.
  int set_direct(Index rows);
  int host_waits() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace kutta_amd
