// Host setup of the aggregation hierarchy: pairwise matching and the coarse
// operators with their canonical Galerkin entry maps. The aggregates and every
// coarse value must equal the reference's bit for bit (linalg/amg.cpp:25-174),
// so the order of each float sum is the reference's.
#pragma once

#include <vector>

#include "hip_device.hpp"

namespace kutta_amd {

// linalg/csr.hpp:16-28: host CSR, sorted columns, diagonal always present.
struct CsrHost {
  Index n_rows = 0;
  std::vector<Index> row_ptr{0}, col, diag;
  std::vector<double> val;
  Index nnz() const { return static_cast<Index>(col.size()); }
};
void validate_host(const CsrHost& a);

// Upper-bound float sums (kernels/prelude.h:82-100): add_up rounds a + b up
// when the rounding dropped a positive error; UpperSum carries the error term.
float add_up(float a, float b);
struct UpperSum {
  float sum = 0, error = 0;
  void add(float b);
  float value() const;
};

enum class NullSpace : std::uint8_t { None, Constant };
enum class Structure : std::uint8_t { General, SymmetricMMatrix };

// linalg/amg.hpp:8-20 (V-cycle options).
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
  // K-cycle only (no reference counterpart): the inner corrections stop at
  // this coarse level, plain recursion below; 0 uses them wherever the
  // reference does (every coarse level with two levels below it).
  int k_depth = 0;
  // K-cycle only (no reference counterpart): inner corrections only on coarse
  // levels of at least this many rows (0: no bound), where a level's
  // launches move enough data to be worth a second visit.
  Index k_min_rows = 0;
};

struct Coarsening {
  CsrHost matrix;
  std::vector<Index> aggregate, member_offsets, members, entry_offsets, entries;
};

// Greedy strongest-unassigned pairing (linalg/amg.cpp:25-75).
std::vector<Index> pair_rows(const CsrHost& a, double threshold, bool exclude_isolated);
// Coarse operator for a fine->aggregate map (linalg/amg.cpp:77-158).
Coarsening coarsen(const CsrHost& fine, std::vector<Index> aggregate, bool entry_map, bool m_matrix);
// `aggregation_passes` rounds of pairing (linalg/amg.cpp:160-174).
Coarsening aggregate(const CsrHost& a, const AmgOptions& options);
// Host (l1) inverse diagonal (linalg/solver.cpp:37-53).
std::vector<float> diagonal_values(const CsrHost& a, bool l1, bool nullspace);
// Symmetric nonpositive edges with diagonally dominant rows (linalg/amg.cpp:364-379).
void check_m_matrix(const CsrHost& a);

}  // namespace kutta_amd
