// ORACLE — test infrastructure only. A serial CPU restatement of the Kutta
// reference's SIMPLEC pressure-correction path (ezzy1630/Kutta). Only tests,
// the smoke check and the CPU-baseline leg of the benchmark may load it, and
// only to check or time against; the product never links it.
//
// Arithmetic contract: compiled with -O2 -ffp-contract=off and no fast-math,
// float expressions are evaluated exactly in the reference's order, and every
// reduction is a sequential FP64 loop (runtime/cpu/cpu_device.cpp:37-89).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

using Index = std::int32_t;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NonfiniteError : Error {
  using Error::Error;
};

#define ORACLE_REQUIRE(cond, msg)                                              \
  do {                                                                         \
    if (!(cond)) throw ::oracle::Error(std::string(__FILE__) + ":" +           \
                                       std::to_string(__LINE__) + ": " + (msg)); \
  } while (0)

// ---------------------------------------------------------------- float helpers
// kernels/prelude.h:82-100
float kadd_up(float a, float b);
struct UpperSum { float sum, error; };
UpperSum kupper_add(UpperSum a, float b);
float kupper_value(UpperSum a);

// ------------------------------------------------------------- serial reductions
// runtime/cpu/cpu_device.cpp:37-62 and :72-89
double sum(const float* x, std::size_t n);
double dot(const float* x, const float* y, std::size_t n);
double max_abs(const float* x, std::size_t n);
enum class Reduction { Sum, Dot, Dot2 };
void reduce_into(const float* x, const float* y, const float* u, const float* v,
                 std::size_t n, float* dst, Reduction mode);
void store_pair(float* p, double value);

// -------------------------------------------------------------- linalg kernels
// One call runs the kernel body for every index in [0, n), as CpuDevice::launch
// does (runtime/cpu/cpu_device.cpp:64-68).
namespace kern {
void csr_spmv(int n, const int* rows, const int* cols, const float* values,
              const float* x, float* y);
void csr_spmv_guarded(int n, const int* rows, const int* cols, const float* values,
                      const float* x, float* y, const int* status);
float csr_product_error(float a, float b, float product);
float csr_row_residual(int i, const int* rows, const int* cols, const float* values,
                       const float* x, const float* b);
void csr_residual(int n, const int* rows, const int* cols, const float* values,
                  const float* x, const float* b, float* r);
void csr_residual_guarded(int n, const int* rows, const int* cols, const float* values,
                          const float* x, const float* b, float* r, const int* status);
void krylov_diagonal(int n, const int* rows, const int* diag, const float* values,
                     float* inverse, float* invalid, int use_l1);
void krylov_multiply(int n, const float* r, const float* inverse, float* z);
void krylov_multiply_guarded(int n, const float* r, const float* inverse, float* z,
                             const int* status);
void amg_coarse_solve(int n, const int* rows, const int* cols, const float* values,
                      float* x, const float* b, const float* inverse, const int* status,
                      int count, int sweeps);
void amg_smooth(int n, const int* rows, const int* cols, const float* values,
                const float* x, const float* b, const float* inverse, float* next,
                const int* status);
void amg_smooth_zero(int n, const int* rows, const int* cols, const float* values,
                     const float* x, const float* b, const float* inverse, float* next,
                     const int* status);
void amg_restrict(int n, const int* offsets, const int* members, const float* fine,
                  float* coarse);
void amg_restrict_guarded(int n, const int* offsets, const int* members,
                          const float* fine, float* coarse, const int* status);
void amg_prolong(int n, const int* aggregate, const float* coarse, float* fine);
void amg_prolong_guarded(int n, const int* aggregate, const float* coarse, float* fine,
                         const int* status);
void amg_galerkin(int n, const int* offsets, const int* entries, const float* fine,
                  float* coarse);
void amg_m_matrix_diagonal(int n, const int* offsets, const int* members,
                           const int* fine_rows, const int* fine_diag, const float* fine,
                           const int* rows, const int* diag, float* coarse);
void axpy(int n, float* y, const float* x, float a);
void scale_copy(int n, float* y, const float* x, float a);

// Device-scalar state machine (kernels/scalars.kern.h:51-152).
void solver_limits(int n, float* scalars, float tol_hi, float tol_lo, float div_hi,
                   float div_lo);
void solver_reset(float* scalars, int* status, int iterations);
void solver_pcg_alpha(float* scalars, int* status);
void solver_pcg_beta(float* scalars, int* status, int flexible);
void solver_bicg_beta(float* scalars, int* status);
void solver_bicg_alpha(float* scalars, int* status);
void solver_bicg_s(const float* scalars, int* status);
void solver_bicg_omega(float* scalars, int* status);
void solver_check(float* scalars, int* status, int pcg);
void solver_step(int n, float* x, float* r, const float* p, const float* q,
                 const float* scalars, const int* status);
void solver_direction(int n, const float* z, float* p, const float* scalars,
                      const int* status);
void solver_bicg_direction(int n, const float* r, const float* v, float* p,
                           const float* scalars, const int* status);
void solver_bicg_difference(int n, const float* r, const float* q, float* s,
                            const float* scalars, const int* status);
void solver_bicg_step(int n, float* x, float* r, const float* z, const float* zs,
                      const float* s, const float* t, const float* scalars,
                      const int* status);
void solver_project(int n, float* x, const float* sum, float count_hi, float count_lo);
// The optional K-cycle's inner corrections (kernels/scalars.kern.h:157-186).
void amg_inner_alpha(float* scalars, int* status, const int* parent, int first);
void amg_inner_first(int n, const float* b, const float* p, const float* q, float* x, float* t,
                     const float* scalars, const int* status);
void amg_inner_check(const float* scalars, int* status);
void amg_inner_beta(float* scalars, int* status);
void amg_inner_step(int n, float* x, const float* p, const float* scalars, const int* status);
}  // namespace kern

// ------------------------------------------------------------------------- CSR
// linalg/csr.hpp:16-38. Host values are double (CsrMatrixHost); the device copy
// of the values is float.
struct CsrHost {
  Index n_rows = 0;
  std::vector<Index> row_ptr{0}, col, diag;
  std::vector<double> val;
  Index nnz() const { return static_cast<Index>(col.size()); }
};
void validate_host(const CsrHost& a);

// linalg/solver.hpp:223-246, linalg/amg.hpp:183-197
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
enum class SolverKind : std::uint8_t { Jacobi, PcgJacobi, BicgstabJacobi, AmgPcg, AmgBicgstab };
enum class NullSpace : std::uint8_t { None, Constant };
enum class Structure : std::uint8_t { General, SymmetricMMatrix };
// linalg/amg.hpp:6 — V: one recursive coarse correction; K: two inner
// preconditioned Krylov corrections per coarse level.
enum class AmgCycle : std::uint8_t { V, K };
struct AmgOptions {
  int pre_sweeps = 5, post_sweeps = 5, coarse_sweeps = 32;
  Index coarse_rows = 24;
  int max_levels = 24, aggregation_passes = 3;
  double strength_threshold = 0.25;
  NullSpace nullspace = NullSpace::None;
  Structure structure = Structure::General;
  AmgCycle cycle = AmgCycle::V;
  // K-cycle only: the inner corrections stop at this coarse level (plain
  // recursion below it); 0 uses them wherever the reference does (every
  // level with two levels below it).
  int k_depth = 0;
  // K-cycle only: inner corrections only on coarse levels of at least this
  // many rows (0: no bound).
  Index k_min_rows = 0;
};

// The aggregation of one level (linalg/amg.cpp:17-20).
struct Coarsening {
  CsrHost matrix;
  std::vector<Index> aggregate, member_offsets, members, entry_offsets, entries;
};
std::vector<Index> pairs(const CsrHost& a, double threshold, bool exclude_isolated);
Coarsening coarsen(const CsrHost& fine, std::vector<Index> aggregate, bool entry_map,
                   bool m_matrix);
Coarsening aggregate(const CsrHost& a, const AmgOptions& options);
std::vector<float> diagonal_values(const CsrHost& a, bool l1, bool nullspace);

// A CSR operator as the device holds it: pattern plus float values.
struct CsrDevice {
  Index n_rows = 0;
  std::vector<Index> row_ptr, col, diag;
  std::vector<float> val;
  Index nnz() const { return static_cast<Index>(col.size()); }
};
CsrDevice upload(const CsrHost& a);

struct LevelInfo {
  Index rows = 0, nnz = 0;
};

// Krylov solver with optional aggregation-AMG preconditioner
// (linalg/solver.cpp:75-393, linalg/amg.cpp:198-342).
class Solver {
 public:
  Solver(SolverKind kind, const CsrHost& a, const AmgOptions& options);
  ~Solver();
  void refresh(const CsrDevice& a);
  SolveStats solve(const CsrDevice& a, std::vector<float>& x, const std::vector<float>& b,
                   const SolveControl& control);
  std::string describe() const;
  std::vector<LevelInfo> levels() const;
  // The hierarchy's transfer arrays and coarse values, for parity checks.
  const Coarsening* transfer(std::size_t level) const;
  const CsrDevice* level_matrix(std::size_t level) const;
  const std::vector<float>* level_l1(std::size_t level) const;
  // One preconditioner application z = M^-1 r with a running status (the
  // V- or K-cycle, or Jacobi), for checks of the cycle itself.
  void apply(const CsrDevice& a, const std::vector<float>& r, std::vector<float>& z);

 private:
  struct Impl;
  Impl* impl_;
};

// tests/linalg/matrices.hpp:13-62
std::uint32_t hash(std::uint32_t x);
CsrHost poisson(Index nx, Index ny, Index nz, double anisotropy, bool random, bool neumann,
                double wind);

}  // namespace oracle
