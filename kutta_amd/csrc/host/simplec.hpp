// Steady SIMPLEC solver on the MI355X backend: the HIP-native counterpart of
// kutta::fv::SteadySolver (fv/steady_solver.hpp:286-360) for laminar flow in a
// closed domain (wall and symmetry patches, the reference cell pinned), with
// the reference's default numerics: paired (hi+lo) or float state,
// second-order upwind momentum limited by Barth-Jespersen, least-squares
// gradients, momentum-weighted Rhie-Chow faces, the automatic pseudo time
// step and the AMG-PCG pressure correction. Every field lives in HBM; the host
// reads one reduction per iteration (the time step's speed scale) plus the
// linear solvers' convergence polls, as the reference does.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "krylov.hpp"
#include "mesh.hpp"

namespace kutta_amd {

// fv/steady_solver.hpp:23-51, 65 (the supported conditions), coded by the
// boundary_kind of the face kernels: a pressure outlet with a backflow
// velocity is kind 2, with normal backflow (the reference's default) kind 4.
struct PatchCondition {
  // MassFlowInlet (fv/steady_solver.hpp:60-65) is a velocity inlet (kind 1)
  // on the faces; its code here only selects the set-up.
  enum Kind : int {
    Wall = 0, VelocityInlet = 1, PressureOutlet = 2, Symmetry = 3, NormalBackflowOutlet = 4, MassFlowInlet = 5
  };
  std::string patch;
  int kind = Wall;
  Vec3 velocity{};        // wall, inlet or backflow velocity (uniform)
  double pressure = 0;    // outlet static pressure
  bool total_backflow = false;  // BackflowPressure::Total at an outlet
  // A velocity profile of the face centroid for an inlet or an outlet's
  // backflow (VelocityProfile / BackflowProfile as a function); replaces
  // `velocity` when set.
  std::function<Vec3(Vec3)> profile;
  // MassFlowInlet: kg/s into the domain over the whole patch, and an
  // optional inward direction (zero: along each face's inward normal).
  double mass_flow = 0;
  Vec3 direction{};
  bool has_direction = false;  // the optional direction is given
  // SST inflow turbulence at inlets and outlets (fv/steady_solver.hpp:28-35):
  // -1 none, 0 TurbulenceValues{k, omega}, 1 IntensityLength, 2
  // IntensityViscosityRatio; a and b are the pair's two values.
  int turbulence = -1;
  double turbulence_a = 0, turbulence_b = 0;
};

// SstSettings (fv/steady_solver.hpp:102-118) with SstOptions
// (turbulence/sst.hpp:46-78, no intermittency) as model[19]: has_ambient,
// ambient k, ambient omega, production_limiter, production_clip,
// limit_omega_production, omega_production_molecular_ratio, low_re,
// gamma_set Menter2003, viscosity limiter Vorticity, kato_launder,
// vorticity_production, wall_production, cdkw_floor, limit_cross_diffusion,
// cross_diffusion_clip, omega_floor, k_floor, max_viscosity_ratio.
struct SstSettings {
  double model[19] = {0, 0, 0, 1, 10, 1, 0, 0, 0, 0, 0, 0, 1, 1e-10, 1, 1e6, 1e-20, 1e-14, 1e5};
  bool k_second_order = true, omega_second_order = true;
  double k_relaxation = .75, omega_relaxation = .75;
  SolveControl linear{1e-3, 1e-20, 100, 8};
  double steady_tolerance = 1e-6;
  bool compensated_transport = true, pseudo_transient = false;
};

// fv/steady_solver.hpp:193-271, the members this solver honours; the others
// keep their reference defaults, which are the only ones accepted.
struct SolverSettings {
  double density = 1, viscosity = .01;
  double momentum_relaxation = .7, pressure_relaxation = 1;
  double pseudo_time_scale = 1;
  int monitor_interval = 20;
  bool paired_state = true;
  double paired_state_slack_ulps = 1;
  Vec3 reference_point{};
  SolveControl momentum_linear{1e-2, 1e-20, 100, 8};
  SolveControl pressure_linear{1e-3, 1e-20, 300, 8};
  // Solve the three momentum components in one batched launch sequence when
  // that is bitwise the reference's order (no reference counterpart).
  bool batched_momentum = true;
  double pseudo_time_step = 0;  // a fixed pseudo time step; 0 is automatic
  bool sst = false;             // k-omega SST with sst_settings
  SstSettings sst_settings;
};

// fv/steady_solver.hpp:273-284
struct ResidualValue {
  double raw = 0, scaled = 0, fluent_scaled = 0, fluent_normalization = 0;
};
struct IterationReport {
  int iteration = 0;
  ResidualValue u, v, w, continuity;
  double corrected_mass_imbalance = 0, maximum_velocity = 0, pseudo_time_step = 0;
  SolveStats momentum[3], pressure;
  ResidualValue k, omega;
  SolveStats turbulence[2];
};

// fv/steady_solver.hpp:286-292 (FP64 sums of the per-face wall forces).
struct ForceReport {
  Vec3 pressure, viscous, area_vector;
  double reference_pressure = 0;
  Vec3 force() const { return {pressure.x + viscous.x, pressure.y + viscous.y, pressure.z + viscous.z}; }
  // fv/steady_solver.cpp:1780-1784
  double coefficient(Vec3 direction, double density, double velocity, double area, double pressure) const;
};

class SteadySolver {
 public:
  SteadySolver(HipDevice& dev, PolyMesh mesh, std::vector<PatchCondition> conditions,
               SolverSettings settings = {});
  ~SteadySolver();
  void iterate(int count);
  // SteadySolver::initialize (fv/steady_solver.cpp:1480-1491): the velocity
  // and pressure of every cell in the solver's cell order (that of
  // geometry().cell_centroid); a closed domain's pressure is shifted so the
  // reference cell holds zero.
  void initialize(const std::vector<Vec3>& velocity, const std::vector<double>& pressure);
  // SteadySolver::initialize_fields (fv/steady_solver.cpp:1493-1524) in the
  // solver's cell order: also the SST k and omega (empty without SST), then
  // three monitor passes that establish the momentum coefficients.
  void initialize_fields(const std::vector<Vec3>& velocity, const std::vector<double>& pressure,
                         const std::vector<double>& k, const std::vector<double>& omega);
  IterationReport monitor();
  const IterationReport& report() const;
  // Pressure and viscous force of the fluid on the named wall patches
  // (fv/steady_solver.cpp:1711-1736).
  ForceReport forces(const std::vector<std::string>& patches, double reference_pressure = 0);
  // Device array by name (u, v, w, p, mass, ...; the names of the oracle's
  // restatement) and its length; nullptr when unknown or not allocated.
  const float* field(const std::string& name, std::size_t* size) const;
  Index reference_cell() const;
  const MeshGeometry& geometry() const;
  Index cells() const;
  // Host waits of the linear solvers so far (pressure, momentum).
  int host_waits() const;
  // SstModel::update (turbulence/sst.cpp:101-158, no intermittency) on this
  // solver's mesh, walls as its wall patches. options[19]: has_ambient,
  // ambient k, ambient omega, production_limiter, production_clip,
  // limit_omega_production, omega_production_molecular_ratio, low_re,
  // gamma_set Menter2003, viscosity limiter Vorticity, kato_launder,
  // vorticity_production, wall_production, cdkw_floor,
  // limit_cross_diffusion, cross_diffusion_clip, omega_floor, k_floor,
  // max_viscosity_ratio (SstOptions, turbulence/sst.hpp:46-78). inputs[25]:
  // device arrays velocity x y z, grad u (x y z), grad v, grad w, density,
  // viscosity, k, omega, grad k (x y z), grad omega (x y z) per cell, then
  // wall velocity x y z per boundary face (SstInputs). The fields
  // (SstFields) are read by field(name); the integer ones hold int32 bits.
  void sst_update(const double* options, const float* const* inputs);
  // Batched momentum solve on or off; returns whether it is in use (it
  // stays off where it would change the iterates: symmetry faces that are not
  // axis-aligned).
  bool set_batched_momentum(bool on);
  bool batched_momentum() const;
  // HIP graph replay of the pressure PCG iterations (default on; identical
  // iterates).
  void set_graphs(bool on);
  // Rebuilds the pressure AMG with other options (no reference counterpart;
  // e.g. the K-cycle). Symmetric M-matrix structure, no nullspace.
  void set_pressure_options(const AmgOptions& options);
  // The pressure solver's hierarchy (rows, nonzeros per level) and plan.
  std::vector<std::pair<Index, Index>> pressure_levels() const;
  std::string pressure_plan() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace kutta_amd
