// ORACLE — test infrastructure only. Builds the reference's own kernel bodies
// (src/kutta/kernels/*.kern.h, read in place under the reference tree) into
// oracle/_ref/libref_kernels.so, with one extern "C" loop per kernel running
// the body for every index in [0, n) as the reference's serial CPU backend
// does (runtime/cpu/cpu_device.cpp:64-68). Used only to cross-check the
// oracle's restatement.
#include <cstddef>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

#include "kutta/kernels/prelude.h"

// Kernel bodies live in their own namespace, as the reference's generated CPU
// table arranges them (tools/kernelgen.py:176-186).
namespace refk {
#include "kutta/kernels/linalg.kern.h"
#include "kutta/kernels/amg.kern.h"
#include "kutta/kernels/blas.kern.h"
#include "kutta/kernels/scalars.kern.h"
}  // namespace refk
// The finite-volume bodies use the prelude's pair type, which scalars.kern.h
// shadows inside refk, so they get a namespace of their own.
namespace reff {
#include "kutta/kernels/fv_flow.kern.h"
#include "kutta/kernels/fv_gradient.kern.h"
#include "kutta/kernels/fv_transport.kern.h"
#include "kutta/kernels/fv_paired.kern.h"
#include "kutta/kernels/fv_rans.kern.h"
#include "kutta/kernels/sst.kern.h"
#include "kutta/kernels/fv_fluent_schemes.kern.h"
}  // namespace reff

#define LOOP(name, ...) for (int i = 0; i < n; ++i) refk::name(i, __VA_ARGS__)

// Name-keyed runner over the reference's finite-volume bodies: args[k] points
// at the k-th parameter (a buffer, or a 4-byte scalar), as for oracle_fv_each.
namespace {
template <class T>
T from_arg(void* const* a, std::size_t k) {
  if constexpr (std::is_pointer_v<T>) return static_cast<T>(a[k]);
  else {
    T v;
    std::memcpy(&v, a[k], sizeof v);
    return v;
  }
}
template <class... P, std::size_t... K>
void run_each(void (*body)(int, P...), int n, void* const* a, std::index_sequence<K...>) {
  const std::tuple<P...> args{from_arg<P>(a, K)...};
  for (int i = 0; i < n; ++i) std::apply([&](P... p) { body(i, p...); }, args);
}
template <class... P>
void run_each(void (*body)(int, P...), int n, void* const* a) {
  run_each(body, n, a, std::index_sequence_for<P...>{});
}
using Runner = void (*)(int, void* const*);
#define REF_FV(name) {#name, [](int n, void* const* a) { run_each(&reff::name, n, a); }}
const std::map<std::string, Runner>& fv_table() {
  static const std::map<std::string, Runner> t = {
      REF_FV(fv_velocity_boundary), REF_FV(fv_pressure_boundary), REF_FV(fv_pressure_faces),
      REF_FV(fv_momentum_source), REF_FV(fv_momentum_faces), REF_FV(fv_convective_outflow),
      REF_FV(fv_peclet_side_velocity), REF_FV(fv_rhie_chow), REF_FV(fv_correct_flux),
      REF_FV(fv_correct_velocity), REF_FV(fv_monitor_cell), REF_FV(fv_speed),
      REF_FV(fv_continuity), REF_FV(fv_wall_pressure_force), REF_FV(fv_wall_viscous_force),
      REF_FV(fv_fluent_momentum_scale), REF_FV(fv_physical_time_terms),
      REF_FV(fv_face_time_defect), REF_FV(fv_face_time_flux), REF_FV(fv_cell_mass_error),
      REF_FV(fv_gradient), REF_FV(fv_paired_gradient), REF_FV(fv_range_shift),
      REF_FV(fv_limiter), REF_FV(fv_paired_limiter), REF_FV(fv_paired_gradient_limiter),
      REF_FV(fv_vector_gradient), REF_FV(fv_scale_gradient), REF_FV(fv_limiter_scale),
      REF_FV(fv_paired_limiter_scale), REF_FV(fv_paired_gradient_limiter_scale),
      REF_FV(fv_jump_gradient), REF_FV(fv_face_jump), REF_FV(fv_gather_jump_gradient),
      REF_FV(fv_paired_jump_gradient), REF_FV(fv_lsq_restore_inverse),
      REF_FV(fv_diffusion_correction_bound), REF_FV(fv_transport_faces),
      REF_FV(fv_transport_assemble), REF_FV(fv_fixed_rows), REF_FV(fv_floor),
      REF_FV(fv_scalar_steady_measure), REF_FV(fv_scalar_set_fixed), REF_FV(fv_scalar_monitor),
      REF_FV(fv_compensated_pressure_jump), REF_FV(fv_paired_pressure_jump),
      REF_FV(fv_compensated_rhie_chow), REF_FV(fv_paired_pressure_boundary_low),
      REF_FV(fv_paired_pressure_faces), REF_FV(fv_paired_momentum_residual),
      REF_FV(fv_compensated_momentum_residual), REF_FV(fv_paired_accumulate),
      REF_FV(fv_compensated_pressure_force), REF_FV(fv_paired_velocity_boundary),
      REF_FV(fv_paired_advection), REF_FV(fv_paired_face_residual), REF_FV(fv_paired_flux),
      REF_FV(fv_paired_flux_offset), REF_FV(fv_paired_continuity), REF_FV(fv_paired_pressure_rhs),
      REF_FV(fv_paired_correct_flux), REF_FV(fv_paired_correct_velocity), REF_FV(fv_paired_value),
      REF_FV(fv_paired_vector_values), REF_FV(fv_compensated_correct_flux),
      REF_FV(fv_compensated_correct_velocity),
      REF_FV(fv_backflow_pressure), REF_FV(fv_effective_viscosity), REF_FV(fv_transpose_stress),
      REF_FV(fv_stress_source), REF_FV(fv_boundary_momentum), REF_FV(fv_momentum_diagonal),
      REF_FV(fv_sst_force), REF_FV(fv_turbulence_boundary),
      REF_FV(sst_rates), REF_FV(sst_blending), REF_FV(sst_viscosity), REF_FV(sst_blending_viscosity),
      REF_FV(sst_wall), REF_FV(sst_wall_omega), REF_FV(sst_wall_cells), REF_FV(sst_sources),
      REF_FV(sst_face_diffusion), REF_FV(sst_sustain),
      REF_FV(fv_simple_inverse), REF_FV(fv_standard_pressure_faces),
      REF_FV(fv_cell_pressure_boundary), REF_FV(fv_force_to_shear),
      REF_FV(fv_poor_cell_pressure_faces), REF_FV(fv_poor_pressure_interfaces),
  };
  return t;
}
}  // namespace

extern "C" {
int ref_fv_each(const char* name, int n, void* const* args) {
  const auto& t = fv_table();
  const auto it = t.find(name);
  if (it == t.end()) return 1;
  it->second(n, args);
  return 0;
}
void ref_csr_spmv(int n, const int* r, const int* c, const float* v, const float* x, float* y) {
  LOOP(csr_spmv, r, c, v, x, y);
}
void ref_csr_residual(int n, const int* r, const int* c, const float* v, const float* x,
                      const float* b, float* out) {
  LOOP(csr_residual, r, c, v, x, b, out);
}
void ref_krylov_diagonal(int n, const int* r, const int* d, const float* v, float* inv,
                         float* invalid, int l1) {
  LOOP(krylov_diagonal, r, d, v, inv, invalid, l1);
}
void ref_amg_smooth(int n, const int* r, const int* c, const float* v, const float* x,
                    const float* b, const float* inv, float* next, const int* st) {
  LOOP(amg_smooth, r, c, v, x, b, inv, next, st);
}
void ref_amg_smooth_zero(int n, const int* r, const int* c, const float* v, const float* x,
                         const float* b, const float* inv, float* next, const int* st) {
  LOOP(amg_smooth_zero, r, c, v, x, b, inv, next, st);
}
void ref_amg_coarse_solve(const int* r, const int* c, const float* v, float* x, const float* b,
                          const float* inv, const int* st, int count, int sweeps) {
  refk::amg_coarse_solve(0, r, c, v, x, b, inv, st, count, sweeps);
}
void ref_amg_restrict(int n, const int* off, const int* mem, const float* f, float* c) {
  LOOP(amg_restrict, off, mem, f, c);
}
void ref_amg_prolong(int n, const int* agg, const float* c, float* f) { LOOP(amg_prolong, agg, c, f); }
void ref_amg_galerkin(int n, const int* off, const int* ent, const float* f, float* c) {
  LOOP(amg_galerkin, off, ent, f, c);
}
void ref_amg_m_matrix_diagonal(int n, const int* off, const int* mem, const int* fr, const int* fd,
                               const float* f, const int* r, const int* d, float* c) {
  LOOP(amg_m_matrix_diagonal, off, mem, fr, fd, f, r, d, c);
}
void ref_solver_scalar(int which, float* s, int* st, int flag) {
  switch (which) {
    case 0: refk::solver_reset(0, s, st, flag); break;
    case 1: refk::solver_pcg_alpha(0, s, st); break;
    case 2: refk::solver_pcg_beta(0, s, st, flag); break;
    case 3: refk::solver_bicg_beta(0, s, st); break;
    case 4: refk::solver_bicg_alpha(0, s, st); break;
    case 5: refk::solver_bicg_s(0, s, st); break;
    case 6: refk::solver_bicg_omega(0, s, st); break;
    case 7: refk::solver_check(0, s, st, flag); break;
  }
}
void ref_amg_inner(int which, int n, const float* b, const float* p, const float* q, float* x,
                   float* t, float* s, int* st, const int* parent, int flag) {
  switch (which) {
    case 0: refk::amg_inner_alpha(0, s, st, parent, flag); break;
    case 1: LOOP(amg_inner_first, b, p, q, x, t, s, st); break;
    case 2: refk::amg_inner_check(0, s, st); break;
    case 3: refk::amg_inner_beta(0, s, st); break;
    case 4: LOOP(amg_inner_step, x, p, s, st); break;
  }
}
void ref_solver_project(int n, float* x, const float* sum, float hi, float lo) {
  LOOP(solver_project, x, sum, hi, lo);
}
float ref_kadd_up(float a, float b) { return kadd_up(a, b); }
void ref_fv_pressure_assemble(int n, const int* offsets, const int* faces, const float* sign,
                              const int* adjacent, const int* entry, const int* rows,
                              const int* diag, const float* coefficient, const float* flux,
                              float* values, float* rhs, int reference) {
  for (int i = 0; i < n; ++i) reff::fv_pressure_assemble(i, offsets, faces, sign, adjacent, entry, rows, diag, coefficient, flux,
       values, rhs, reference);
}
}
