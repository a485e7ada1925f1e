/* kutta_amd.h — C ABI of the MI355X host library (libkutta_amd.so): the
 * HIP-native pressure solver and the mesh/operator setup of Kutta's SIMPLEC
 * pressure-correction path, built on the kutta_hip device ABI.
 *
 * Every int-returning function returns 0 on success, 2 for a non-finite
 * (diverged) input, 1 for any other error; ka_last_error() describes it.
 */
#ifndef KUTTA_AMD_H
#define KUTTA_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "kutta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* ka_last_error(void);

/* A device (one HIP stream). ka_device_raw exposes the kutta_hip handle for
 * buffer management through kh_alloc/kh_upload/... on the same stream. */
int ka_device_create(int ordinal, void** out);
void ka_device_destroy(void* dev);
kh_device* ka_device_raw(void* dev);

/* Host CSR operator (linalg/csr.hpp:16-28): columns sorted, diagonal present,
 * double values. */
int ka_csr_create(int n_rows, const int32_t* row_ptr, const int32_t* col, const int32_t* diag,
                  const double* val, void** out);
void ka_csr_destroy(void* csr);
int ka_csr_rows(void* csr);
int ka_csr_nnz(void* csr);
int ka_csr_copy(void* csr, int32_t* row_ptr, int32_t* col, int32_t* diag, double* val);

/* Host aggregation of one level (linalg/amg.cpp:160-174), no device needed.
 * options as for ka_solver_create. Arrays (which): 0 aggregate,
 * 1 member_offsets, 2 members, 3 entry_offsets, 4 entries, 5 coarse row_ptr,
 * 6 coarse col, 7 coarse diag. ka_coarsening_values gives the coarse values. */
int ka_aggregate(void* csr, const int32_t* options, double threshold, void** out);
long ka_coarsening_array(void* coarsening, int which, int32_t* out);
long ka_coarsening_values(void* coarsening, double* out);
void ka_coarsening_destroy(void* coarsening);

/* Device CSR (linalg/csr.hpp:31-48): upload(Device&, CsrMatrixHost&). The value
 * array is the part that changes between solves. */
int ka_matrix_upload(void* dev, void* csr, void** out);
void ka_matrix_destroy(void* matrix);
float* ka_matrix_values(void* matrix);
const int32_t* ka_matrix_row_ptr(void* matrix);
const int32_t* ka_matrix_col(void* matrix);
const int32_t* ka_matrix_diag(void* matrix);

/* Pressure-correction operator of a box mesh on [0,1]^3 as the SIMPLEC solver
 * sets it up (fv/steady_solver.cpp:240-371): make_box_tet/hex
 * (io/generators.cpp:98-103), an optional seeded point perturbation of
 * relative amplitude `amplitude` (no reference counterpart), RCM renumbering,
 * FP64 geometry, the FV stencil and fv_pressure_assemble on the device with the
 * face diffusion coefficients and zero flux. The handle is also a matrix handle
 * (its assembled device operator) for ka_solver_refresh / ka_solver_solve. */
int ka_pressure_problem_create(void* dev, int nx, int ny, int nz, int tet, double amplitude,
                               unsigned long long seed, int reference, void** out);
void ka_pressure_problem_destroy(void* problem);
/* Host CSR of the assembled operator, owned by the problem (for ka_solver_create). */
void* ka_pressure_problem_csr(void* problem);
/* Re-runs fv_pressure_assemble on the device. */
int ka_pressure_problem_assemble(void* problem, int reference);
/* reference: cell to pin (row/column eliminated), -1 none (pure Neumann), -2 the
 * cell holding the origin, the SIMPLEC default (fv/steady_solver.cpp:247).
 * info[7]: cells, faces, internal faces, nnz, cell-face entries, redraw rounds,
 * pinned cell;
 * seconds[6]: box, perturb, renumber, geometry, stencil, upload+assembly. */
int ka_pressure_problem_info(void* problem, int32_t* info, double* seconds);
/* Named host arrays (kind 0 int32, 1 float32, 2 float64); -1 if unknown. */
long ka_pressure_problem_array(void* problem, const char* name, int kind, void* out);

/* The host half of the same setup (mesh, renumbering, geometry, stencil),
 * without a device; arrays by the same names as ka_pressure_problem_array. */
int ka_mesh_setup_create(int nx, int ny, int nz, int tet, double amplitude, unsigned long long seed,
                         void** out);
void ka_mesh_setup_destroy(void* setup);
long ka_mesh_setup_array(void* setup, const char* name, int kind, void* out);

/* Solvers (linalg/solver.hpp:237-293, linalg/amg.hpp:8-27).
 * kind: 0 Jacobi, 1 PcgJacobi, 2 BicgstabJacobi, 3 AmgPcg, 4 AmgBicgstab;
 * additive, no reference counterpart: 5 FgmresJacobi, 6 AmgFgmres (flexible
 * GMRES(m), right preconditioned, restart length 30 unless set).
 * options[11]: pre_sweeps, post_sweeps, coarse_sweeps, coarse_rows, max_levels,
 *             aggregation_passes, nullspace (0 none, 1 constant),
 *             structure (0 general, 1 symmetric M-matrix),
 *             cycle (0 V, 1 K: linalg/amg.hpp:6), and two bounds on where
 *             the K-cycle's inner corrections run (no reference counterpart):
 *             the deepest coarse level (0 every level the reference gives
 *             them) and the fewest rows of such a level (0 no bound). */
int ka_solver_create(void* dev, int kind, void* csr, const int32_t* options, double threshold,
                     void** out);
void ka_solver_destroy(void* solver);
int ka_solver_refresh(void* solver, void* matrix);
/* x and b are device vectors of n_rows floats. stats[4]: iterations,
 * initial residual, final residual, converged. */
int ka_solver_solve(void* solver, void* matrix, float* x, const float* b, double rel_tol,
                    double abs_tol, int max_iter, int interval, double* stats);
/* Up to KH_BATCH_MAX_LANES independent systems with fused BiCGStab-Jacobi
 * solvers (kind 2) at once, one launch per step for all (no reference
 * counterpart): each system's x and stats[4k..4k+3] are bitwise those of its
 * own ka_solver_solve. */
int ka_solver_solve_batched(void* const* solvers, void* const* matrices, float* const* x, const float* const* b,
                            int count, double rel_tol, double abs_tol, int max_iter, int interval, double* stats);
int ka_solver_describe(void* solver, char* out, int capacity);
/* The V-cycle forms the autotuner (ka_solver_autotune_tail) picked: the
 * restriction form per level, the coarsest level's sweeps, the tail level.
 * No reference counterpart; all forms give the same iterates. */
int ka_solver_plan(void* solver, char* out, int capacity);
/* Forces the plan: every level's restriction form (0 one kernel, 1 the same
 * with 128 coarse rows per workgroup, 2 residual then member sums, 3 the same
 * with a wavefront per residual row; -1 keeps), the coarsest level's sweeps
 * (0 a launch per sweep, 1 one workgroup, 2 one workgroup reading the
 * operator once, where the level has a resident partition, else as 1; -1
 * keeps), every level's prolongation form (0 folded into the next sweep, 1
 * in place then a plain sweep, 2 folded, a wavefront per row; -1 keeps) and
 * the sweep form of levels with rows longer than KH_WIDE_ROW (0 a thread per
 * row, 1 a wavefront per row; -1 keeps). */
int ka_solver_set_plan(void* solver, int restrict_mode, int coarse_sweeps, int prolong_mode,
                       int sweep_mode);
int ka_solver_levels(void* solver, int32_t* rows, int32_t* nnz, int capacity);
/* which: 0 aggregate, 1 member_offsets, 2 members, 3 entry_offsets, 4 entries.
 * Returns the length (copies when out is non-null), or -1. */
long ka_solver_transfer(void* solver, int level, int which, int32_t* out);
long ka_solver_level_values(void* solver, int level, float* out);
/* 1: fused MI355X kernels (default); 0: the reference kernel sequence. */
int ka_solver_set_fused(void* solver, int fused);
/* 1 (default): each fused PCG iteration is replayed from a HIP graph captured
 * on first use for the operator and x; 0: launched kernel by kernel. The
 * iterates are identical. No reference counterpart. */
int ka_solver_set_graphs(void* solver, int on);
int ka_solver_host_waits(void* solver);
/* First AMG level of the one-workgroup coarse-grid tail (KH_FUSED_COARSE_TAIL)
 * on the fused path: first < 0 picks it automatically (the default), 0 turns
 * the tail off, k >= 1 requests level k (an error if levels k.. do not fit).
 * *level receives the level in use, 0 for none. No reference counterpart: the
 * iterates are those of the per-level launches either way. */
int ka_solver_set_tail(void* solver, int first, int32_t* level);
/* The tail autotuner: times `repeats` V-cycles on the refreshed operator for
 * each admissible first level (and none) and keeps the fastest in *level.
 * The iterates are identical whichever is chosen. No reference counterpart. */
int ka_solver_autotune_tail(void* solver, void* matrix, int repeats, int32_t* level);
/* Restart length m (1..63) of the FGMRES kinds. No reference counterpart. */
int ka_solver_set_restart(void* solver, int m);
/* Learned coarse correction stage of the AMG FGMRES kind (no reference
 * counterpart). set_learned: a network of `layers` (1..4) message-passing
 * layers on coarse level `level`, the composite aggregation from the fine
 * level, and room for `capacity` harvested pairs. mode: 0 off; 1 harvest
 * (each plain V-cycle step of FGMRES records the restricted residual
 * R (v - A z)); 2 apply (z += P net(R (v - A z)) after each V-cycle; not applied
 * when the last training epoch left more than `max_relative` of the targets'
 * mean square unexplained, and dropped for the rest of a solve when the
 * residual estimate falls by less than `reduction` per step on average; 0.5
 * and 0.5 by default). targets: exact level corrections of the harvested
 * residuals by AMG-PCG to rel_tol. train: Adam steps cycling over the first
 * `pairs` pairs (0: all), losses[steps] and the unexplained fractions
 * relative[steps] (null skipped). evaluate: the mean unexplained fraction on
 * pairs [first, first + count) without training on them (held-out pairs).
 * report out[11]: level, rows, mode, harvested, targets, applications,
 * fallbacks, solves in mode 2, last loss, last epoch's unexplained fraction
 * (-1 before training), solves not applied by that gate.
 * parameters: get (set = 0) or set (set = 1) the network's parameters. */
int ka_solver_set_learned(void* solver, int level, int layers, uint64_t seed, int capacity);
int ka_solver_learned_mode(void* solver, int mode, double reduction, double max_relative);
int ka_solver_learned_targets(void* solver, double rel_tol, int32_t* pairs);
int ka_solver_learned_train(void* solver, int steps, float lr, int pairs, double* losses, double* relative);
int ka_solver_learned_evaluate(void* solver, int first, int count, double* relative);
int ka_solver_learned_report(void* solver, double* out);
int ka_solver_learned_parameters(void* solver, float* theta, int64_t count, int set);
/* Mixed-precision AMG (no reference counterpart): coarse levels (1 and
 * below) smoothed, restricted and prolonged from bfloat16 copies of their
 * values, FP32 arithmetic; the copies are made at the next ka_solver_refresh.
 * The coarse tail and the dense level keep FP32. Not with the nullspace. */
int ka_solver_set_coarse_bf16(void* solver, int on);
/* Mixed-precision AMG (no reference counterpart): the first coarse level of
 * at most `rows` rows (<= 8192) is solved by an FP64 dense inverse computed at
 * the next ka_solver_refresh; coarser levels and the coarse tail go unused.
 * rows = 0 turns it off. *level receives the level (0: none). Symmetric
 * M-matrix operators without nullspace only. */
int ka_solver_set_direct(void* solver, int rows, int32_t* level);
/* Lazy refresh of that dense inverse (no reference counterpart): with
 * tolerance > 0, ka_solver_refresh recomputes it only when the level's values
 * have moved more than `tolerance` in relative l2 distance from the values it
 * was computed from (an FP64 |a - s|/|s| in the refresh's existing host wait);
 * 0 (default) recomputes at every refresh. */
int ka_solver_set_dense_tolerance(void* solver, double tolerance);
/* Inversions done, inverses kept, and the last relative distance. */
int ka_solver_dense_refreshes(void* solver, int64_t* inversions, int64_t* skips, double* staleness);

/* Steady laminar SIMPLEC (fv/steady_solver.hpp:286-360: SteadySolver
 * constructor, iterate, monitor, report, field access) on a box mesh
 * make_box_hex(nx, ny, nz) on [0,1]^3 (io/generators.cpp:98-103), closed
 * domain. kinds[6]/wall[18]: one condition per patch in the generator's order
 * xmin, xmax, ymin, ymax, zmin, zmax; kind 0 Wall{velocity wall[3p..3p+2]},
 * 3 Symmetry. settings[11]: density, viscosity, momentum_relaxation,
 * pressure_relaxation, pseudo_time_scale, monitor_interval, paired_state,
 * paired_state_slack_ulps, reference_point x, y, z (SolverSettings,
 * fv/steady_solver.hpp:193-271; the members not listed keep their defaults). */
int ka_simplec_create(void* dev, int nx, int ny, int nz, const int32_t* kinds, const double* wall,
                      const double* settings, void** out);
/* As ka_simplec_create on the box [0, extent[0]] x [0, extent[1]] x
 * [0, extent[2]] (the unit cube when extent is NULL) with inlets and outlets
 * (fv/steady_solver.hpp:38-51, set up as fv/steady_solver.cpp:314-345):
 * kinds[p] 0 wall, 1 velocity inlet,
 * 2 pressure outlet with a backflow velocity, 3 symmetry, 4 pressure outlet
 * with normal backflow; velocity[3p..3p+2] the wall, inlet or backflow
 * velocity; pressure[p] the outlet static pressure and total[p] != 0 for
 * BackflowPressure::Total (pressure, total may be NULL for closed domains).
 * A domain with an outlet pins no reference cell (:247). */
int ka_simplec_create_open(void* dev, int nx, int ny, int nz, const double* extent, const int32_t* kinds,
                           const double* velocity, const double* pressure, const int32_t* total,
                           const double* settings, void** out);
/* In ka_simplec_create_open and ka_simplec_create_profiled, kinds[p] == 5 is
 * a MassFlowInlet (fv/steady_solver.hpp:60-65; velocities as
 * fv/boundary_physics.cpp:14-26): pressure[p] holds its mass flow (kg/s into
 * the domain over the whole patch) and velocity[3p..3p+2] its inward
 * direction, zero for each face's inward normal. */
/* As ka_simplec_create_open, with the velocity of each patch whose
 * profiled[p] != 0 (a velocity inlet, or an outlet's backflow velocity) given
 * by profile(p, face centroid[3], velocity[3], user) at every face of the
 * patch (VelocityProfile and BackflowProfile as functions,
 * fv/steady_solver.hpp:22, 44; evaluated as fv/steady_solver.cpp:26-29,
 * 317, 337). */
typedef void (*ka_boundary_profile)(int32_t patch, const double* point, double* velocity, void* user);
int ka_simplec_create_profiled(void* dev, int nx, int ny, int nz, const double* extent, const int32_t* kinds,
                               const double* velocity, const double* pressure, const int32_t* total,
                               const int32_t* profiled, ka_boundary_profile profile, void* user,
                               const double* settings, void** out);
/* SteadySolver construction on a general mesh (fv/steady_solver.hpp:313-315,
 * fv/steady_solver.cpp:203-597), laminar or with k-omega SST. The mesh is
 * the PolyMesh arrays (mesh/mesh.hpp:36-68): points[3*n_points],
 * face_offsets[n_faces+1] into face_nodes, owner[n_faces],
 * neighbour[n_internal] (internal faces first), and n_patches contiguous
 * patches of boundary faces named in patch_names ('\n' separated) with
 * patch_start/patch_size. Per patch: kinds[p] (0 wall, 1 velocity inlet, 2
 * pressure outlet with a backflow velocity, 3 symmetry, 4 pressure outlet
 * with normal backflow, 5 mass-flow inlet), velocity[3p..3p+2] (a mass-flow
 * inlet's direction), pressure[p] (a mass-flow inlet's kg/s), total[p]
 * (BackflowPressure::Total), flags[p] bit 0 (a mass-flow direction is
 * given), turbulence[p] (-1 none, 0 TurbulenceValues{k, omega}, 1
 * IntensityLength, 2 IntensityViscosityRatio; fv/steady_solver.hpp:28-35)
 * with its pair turbulence_ab[2p..2p+1], required at inlets and outlets
 * under SST. settings[12]: those of ka_simplec_create, then the fixed pseudo
 * time step (0: automatic). sst: NULL for laminar flow, else SstSettings
 * (fv/steady_solver.hpp:102-118) as 29 values: the 19 SstOptions of
 * ka_simplec_sst_update, k second order (0/1: FirstOrderUpwind/SOU), omega
 * second order, k relaxation, omega relaxation, linear rel_tol, abs_tol,
 * max_iter, steady_tolerance, compensated_transport (0/1),
 * pseudo_transient (0/1). */
int ka_simplec_create_mesh(void* dev, int n_points, const double* points, int n_faces, const int32_t* face_offsets,
                           const int32_t* face_nodes, const int32_t* owner, int n_internal, const int32_t* neighbour,
                           int n_cells, int n_patches, const char* patch_names, const int32_t* patch_start,
                           const int32_t* patch_size, const int32_t* kinds, const double* velocity,
                           const double* pressure, const int32_t* total, const int32_t* flags,
                           const int32_t* turbulence, const double* turbulence_ab, const double* settings,
                           const double* sst, void** out);
/* SteadySolver::initialize_fields (fv/steady_solver.cpp:1493-1524): u, v, w,
 * p and, with SST, k and omega (NULL without) per cell in the solver's cell
 * order; then the reference's three monitor passes. */
int ka_simplec_initialize_fields(void* solver, const double* u, const double* v, const double* w, const double* p,
                                 const double* k, const double* omega);
/* SteadySolver::initialize (fv/steady_solver.cpp:1480-1491): u, v, w, p hold
 * one value per cell in the solver's cell order (that of
 * ka_simplec_centroids); a closed domain's pressure is shifted so that the
 * reference cell holds zero. */
int ka_simplec_initialize(void* solver, const double* u, const double* v, const double* w, const double* p);
/* SstModel::update (turbulence/sst.cpp:101-158; SstOptions and SstInputs,
 * turbulence/sst.hpp:46-124; no intermittency coupling) on the solver's mesh
 * with its wall patches as walls. options[19]: has_ambient, ambient k,
 * ambient omega, production_limiter, production_clip,
 * limit_omega_production, omega_production_molecular_ratio, low_re,
 * gamma_set == Menter2003, viscosity_limiter == Vorticity, kato_launder,
 * vorticity_production, wall_production, cdkw_floor, limit_cross_diffusion,
 * cross_diffusion_clip, omega_floor, k_floor, max_viscosity_ratio.
 * inputs[25]: device pointers to velocity x, y, z; grad u (d/dx, d/dy,
 * d/dz), grad v, grad w; density; viscosity; k; omega; grad k (x, y, z);
 * grad omega (x, y, z) per cell; wall velocity x, y, z per boundary face.
 * The SstFields are then read with ka_simplec_field by name (mu_t, k_source,
 * ..., wall_shear_x, fixed_omega_mask and selected_wall_face as int32 bits). */
int ka_simplec_sst_update(void* solver, const double* options, const float* const* inputs);
void ka_simplec_destroy(void* solver);
/* SteadySolver::iterate(count): count SIMPLEC outer iterations. */
int ka_simplec_iterate(void* solver, int count);
/* The IterationReport (fv/steady_solver.hpp:276-284), after running monitor()
 * first when run_monitor != 0. out[36]: iteration; u, v, w, continuity as
 * (raw, scaled, fluent_scaled, fluent_normalization); corrected mass
 * imbalance; maximum velocity; pseudo time step; momentum iterations x3;
 * pressure iterations; pressure final residual; pressure converged; then
 * SST's k and omega residuals (4 each, as u) and k, omega iterations (zero
 * without SST). */
int ka_simplec_report(void* solver, int run_monitor, double* out);
/* Downloads device field `name` (u v w p mass u_lo v_lo w_lo p_lo mass_lo ...)
 * into out when non-null; returns its length, or -1 if unknown. */
long ka_simplec_field(void* solver, const char* name, float* out);
/* SteadySolver::forces (fv/steady_solver.cpp:1711-1736): the fluid's pressure
 * and viscous force on the wall patches named in `patches` (comma separated),
 * FP64 sums of the per-face forces. out[10]: pressure xyz, viscous xyz, area
 * vector xyz, reference pressure (the ForceReport). */
int ka_simplec_forces(void* solver, const char* patches, double reference_pressure, double* out);
/* ForceReport::coefficient (fv/steady_solver.cpp:1780-1784) of a report as
 * laid out above, along `direction`, with reference density, velocity, area
 * and pressure. */
int ka_simplec_force_coefficient(const double* report, const double* direction, double density,
                                 double velocity, double area, double pressure, double* out);
/* info[4]: cells, pinned reference cell, linear-solver host waits so far,
 * iterations. centroids: 3 doubles per cell of the renumbered mesh. */
int ka_simplec_info(void* solver, int32_t* info);
/* Batched momentum solve (no reference counterpart; on by default): the
 * three components are assembled first and solved by one batched
 * BiCGStab-Jacobi, which gives bitwise the reference's sequential fields when
 * every symmetry face is axis-aligned; elsewhere it stays off. *used
 * receives whether it is in use. */
int ka_simplec_set_batched_momentum(void* solver, int on, int32_t* used);
/* HIP graph replay of the pressure solve's PCG iterations (default on;
 * identical iterates; ka_solver_set_graphs). */
int ka_simplec_set_graphs(void* solver, int on);
/* Rebuilds the pressure solve's AMG with other options (options[11] and the
 * strength threshold as for ka_solver_create; symmetric M-matrix structure,
 * no nullspace), e.g. the K-cycle. No reference counterpart: the reference
 * fixes them (fv/steady_solver.cpp:573-578). The next step refreshes the new
 * hierarchy with the current operator. */
int ka_simplec_set_pressure_options(void* solver, const int32_t* options, double threshold);
/* The pressure solver's hierarchy (rows and nonzeros per level, up to
 * capacity; returns the level count) and its autotuned V-cycle plan (as
 * ka_solver_plan). No reference counterpart. */
int ka_simplec_pressure_levels(void* solver, int32_t* rows, int32_t* nnz, int capacity);
int ka_simplec_pressure_plan(void* solver, char* out, int capacity);
int ka_simplec_centroids(void* solver, double* out);

/* ---- Learned coarse correction (no reference counterpart) ----------------
 * A message-passing network (hidden width 64, `layers` layers, 4 input
 * features) over a coarse operator's graph; bf16 MFMA layers with FP32
 * accumulation, FP32 master weights and Adam. Parameter layout: W0 (4x64),
 * b0 (64), per layer Ws (64x64), Wn (64x64), b (64), then wo (64), bo (1).
 * Vector arguments r, e, target are device pointers of the graph's node
 * count; grad, outputs and losses are host arrays. */
int64_t ka_learned_parameter_count(int layers); /* -1 outside 1..4 */
int ka_learned_create(void* dev, int layers, uint64_t seed, void** out);
void ka_learned_destroy(void* net);
/* Node features and neighbour weights of a symmetric operator with positive
 * diagonal (a ka_matrix handle); again after its values change. */
int ka_learned_set_graph(void* net, void* matrix);
int ka_learned_set_parameters(void* net, const float* theta, int64_t count);
int ka_learned_parameters(void* net, float* theta, int64_t count);
/* e = (network output) * rms(r) / mean diagonal. */
int ka_learned_infer(void* net, const float* r, float* e);
/* loss[2]: the mean squared error of the normalised output against
 * target / scale and the mean squared normalised target; grad (count floats)
 * and the normalised outputs (nodes floats); null outputs are skipped. */
int ka_learned_gradient(void* net, const float* r, const float* target, float* grad, float* outputs,
                        double* loss);
/* `steps` Adam steps cycling over `pairs` (r, target) pairs stored
 * contiguously; losses[steps] receives the loss before each step and
 * relative[steps] that loss over the mean squared normalised target (null
 * arrays are skipped). */
int ka_learned_train(void* net, const float* r, const float* target, int pairs, int steps, float lr,
                     double* losses, double* relative);

#ifdef __cplusplus
}
#endif

#endif /* KUTTA_AMD_H */
