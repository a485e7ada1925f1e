// Steady SIMPLEC on the MI355X backend (fv/steady_solver.cpp:203-597 setup,
// :1303-1369 step, :1417-1477 monitor), laminar closed-domain subset. The
// order of launches and host reductions is the reference's, so every field
// equals the serial CPU backend's bit for bit; the layout is the device's.
#include "simplec.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <numeric>

namespace kutta_amd {
namespace {

using F = DeviceArray<float>;
using I = DeviceArray<Index>;

struct Vec3F {
  F x, y, z;
  Vec3F() = default;
  Vec3F(HipDevice& dev, std::size_t n) : x(dev, n), y(dev, n), z(dev, n) {}
  F& operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
  const F& operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};

double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double length_of(Vec3 a) { return std::sqrt(dot(a, a)); }

// fv/operators.hpp:52 BoundaryObservation.
enum class Observation { Omitted, FaceValue, NormalFootValue };

// One cell's least-squares fit (fv/operators.cpp:97-137): weighted rows
// d/|d|^2 appended in cell-face order (zero for omitted boundary faces) and
// the inverse of the normal matrix, eigen-decomposed by cyclic Jacobi so that
// rank-deficient directions get unit eigenvalues.
void cell_fit(const PolyMesh& m, const MeshGeometry& g, const CellFaces& cf,
              const std::vector<Observation>& obs, Index c, std::vector<Vec3>& delta, double inv[3][3]) {
  double a[3][3]{}, q[3][3]{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const Index ni = m.n_internal_faces();
  for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
    const Index f = cf.faces[j];
    const bool internal = f < ni;
    if (!internal && obs[f - ni] == Observation::Omitted) {
      delta.push_back({});
      continue;
    }
    const Index other = internal ? (m.owner[f] == c ? m.neighbour[f] : m.owner[f]) : -1;
    Vec3 d = sub(internal ? g.cell_centroid[other] : g.face_centroid[f], g.cell_centroid[c]);
    if (!internal && obs[f - ni] == Observation::NormalFootValue) {
      const Vec3 s = g.face_area[f];
      d = scale(s, dot(d, s) / dot(s, s));
    }
    const double w = 1 / dot(d, d), x[]{d.x, d.y, d.z};
    delta.push_back(scale(d, w));
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) a[r][k] += w * x[r] * x[k];
  }
  for (int sweep = 0; sweep < 24; ++sweep) {
    int p = 0, r = 1;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j)
        if (std::abs(a[i][j]) > std::abs(a[p][r])) p = i, r = j;
    if (std::abs(a[p][r]) < 1e-14) break;
    const double angle = .5 * std::atan2(2 * a[p][r], a[r][r] - a[p][p]);
    const double cs = std::cos(angle), sn = std::sin(angle);
    const double app = a[p][p], arr = a[r][r], apr = a[p][r];
    a[p][p] = cs * cs * app - 2 * cs * sn * apr + sn * sn * arr;
    a[r][r] = sn * sn * app + 2 * cs * sn * apr + cs * cs * arr;
    a[p][r] = a[r][p] = 0;
    for (int i = 0; i < 3; ++i) {
      if (i != p && i != r) {
        const double x = a[i][p], y = a[i][r];
        a[i][p] = a[p][i] = cs * x - sn * y;
        a[i][r] = a[r][i] = sn * x + cs * y;
      }
      const double x = q[i][p], y = q[i][r];
      q[i][p] = cs * x - sn * y;
      q[i][r] = sn * x + cs * y;
    }
  }
  for (int r = 0; r < 3; ++r)
    for (int j = 0; j < 3; ++j) inv[r][j] = 0;
  for (int k = 0; k < 3; ++k) {
    const double e = a[k][k] > 1e-10 ? 1 / a[k][k] : 1;
    for (int r = 0; r < 3; ++r)
      for (int j = 0; j < 3; ++j) inv[r][j] += q[r][k] * e * q[j][k];
  }
}

// The LSQ gradient's boundary-face rows (one per boundary face, FP32) and the
// six-plane FP32 inverse; interior rows are the stencil's (fv/operators.cpp:166-211).
struct LsqFit {
  F bx, by, bz, inverse;
};
LsqFit least_squares(HipDevice& dev, const PolyMesh& m, const MeshGeometry& g, const CellFaces& cf,
                     const std::vector<Observation>& obs) {
  const std::size_t n = m.n_cells, ni = m.n_internal_faces(), nb = m.n_faces() - ni;
  std::vector<Vec3> delta;
  delta.reserve(cf.faces.size());
  std::vector<float> inverse(6 * n);
  for (Index c = 0; c < m.n_cells; ++c) {
    double inv[3][3];
    cell_fit(m, g, cf, obs, c, delta, inv);
    const double v[]{inv[0][0], inv[0][1], inv[0][2], inv[1][1], inv[1][2], inv[2][2]};
    for (int k = 0; k < 6; ++k) inverse[c + k * n] = float(v[k]);
  }
  std::vector<float> bx(nb, 0.0f), by(nb, 0.0f), bz(nb, 0.0f);
  for (std::size_t j = 0; j < cf.faces.size(); ++j)
    if (std::size_t(cf.faces[j]) >= ni) {
      const auto b = cf.faces[j] - ni;
      bx[b] = float(delta[j].x);
      by[b] = float(delta[j].y);
      bz[b] = float(delta[j].z);
    }
  return {to_device(dev, bx), to_device(dev, by), to_device(dev, bz), to_device(dev, inverse)};
}

// fv/operators.cpp:139-164: how far a fit's extrapolation to each boundary
// face centroid can amplify its cell's neighbour differences.
std::vector<double> extrapolation_gain(const PolyMesh& m, const MeshGeometry& g, const CellFaces& cf,
                                       const std::vector<Observation>& obs) {
  const Index ni = m.n_internal_faces();
  std::vector<double> gain(m.n_faces() - ni, 0.0);
  std::vector<Vec3> delta;
  for (Index c = 0; c < m.n_cells; ++c) {
    double inv[3][3];
    delta.clear();
    cell_fit(m, g, cf, obs, c, delta, inv);
    for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
      const Index f = cf.faces[j];
      if (f < ni) continue;
      const Vec3 r = sub(g.face_centroid[f], g.cell_centroid[c]);
      const Vec3 w{r.x * inv[0][0] + r.y * inv[1][0] + r.z * inv[2][0],
                   r.x * inv[0][1] + r.y * inv[1][1] + r.z * inv[2][1],
                   r.x * inv[0][2] + r.y * inv[1][2] + r.z * inv[2][2]};
      for (const auto& d : delta) gain[f - ni] += std::abs(dot(w, d));
    }
  }
  return gain;
}

// mesh/mesh.cpp:172-187
std::vector<double> orthogonal_quality(const PolyMesh& m, const MeshGeometry& g) {
  std::vector<double> q(m.n_cells, 1);
  const auto cosine = [](Vec3 a, Vec3 b) { return dot(a, b) / (length_of(a) * length_of(b)); };
  for (Index f = 0; f < m.n_faces(); ++f) {
    const Index o = m.owner[f];
    q[o] = std::min(q[o], cosine(g.face_area[f], sub(g.face_centroid[f], g.cell_centroid[o])));
    if (f >= m.n_internal_faces()) continue;
    const Index nb = m.neighbour[f];
    const Vec3 link = sub(g.cell_centroid[nb], g.cell_centroid[o]);
    q[o] = std::min(q[o], cosine(g.face_area[f], link));
    q[nb] = std::min({q[nb], cosine(g.face_area[f], sub(g.cell_centroid[nb], g.face_centroid[f])),
                      cosine(g.face_area[f], link)});
  }
  for (auto& v : q) v = std::max(v, 0.);
  return q;
}

void upload3(HipDevice& dev, Vec3F& out, std::size_t n, const std::function<Vec3(std::size_t)>& at) {
  std::vector<float> x(n), y(n), z(n);
  for (std::size_t i = 0; i < n; ++i) {
    const Vec3 v = at(i);
    x[i] = float(v.x);
    y[i] = float(v.y);
    z[i] = float(v.z);
  }
  out.x = to_device(dev, x);
  out.y = to_device(dev, y);
  out.z = to_device(dev, z);
}

}  // namespace

// fv/transport.hpp:50-106: one momentum component's equation. The operator's
// values and the face-correction scratch are shared by all three (and by the
// pressure equation), as in the reference.
struct Equation {
  F rhs, diagonal, consistent_inverse, limiter;
  Vec3F grad;
  std::unique_ptr<KrylovSolver> solver;
};

// A turbulence transport equation (fv/transport.hpp:50-106) with the
// compensated solve (fv/transport.cpp:51-55: low words of the assembled RHS
// and diagonal, the delta and its residual) and the steady-stopping work
// arrays (:64-69). It shares the operator values and scratch of momentum.
struct ScalarEquation {
  F rhs, diagonal, consistent_inverse, limiter, rhs_low, diagonal_low, delta, residual;
  F work_residual, work_absolute, work_scale;
  Vec3F grad;
  std::unique_ptr<KrylovSolver> solver;
};

struct SteadySolver::Impl {
  HipDevice& dev;
  SolverSettings settings;
  std::vector<PatchCondition> patch_condition;
  PolyMesh host;
  MeshGeometry geom;
  CellFaces cf;
  Stencil stencil;
  Index n = 0, nf = 0, ni = 0, nb = 0;
  Index reference = -1;
  bool paired = false;

  // fv/device_mesh.hpp:36-67 (wall distance is a turbulence-model input).
  I owner, neighbour, offsets, faces;
  F sign, weight, volume, wall_normal_distance, wall_distance;
  Vec3F area, owner_to_face, neighbour_to_face, owner_to_neighbour;
  // fv/operators.hpp:16-25
  I adjacent, matrix_entry;
  F lsq_x, lsq_y, lsq_z, diffusion, nonorth_x, nonorth_y, nonorth_z;
  LsqFit velocity_lsq, pressure_lsq;
  OwnedCsr matrix;
  std::unique_ptr<KrylovSolver> pressure_solver;
  CsrHost pressure_initial;  // the operator the pressure hierarchy was built from
  Equation momentum[3];
  F face_correction, zero;

  Vec3F velocity, face_velocity, body, source, boundary_velocity, prescribed_velocity;
  Vec3F limited_gradient[3], grad_p, grad_correction, absolute_residual, fluent_scale;
  F pressure, mass, viscosity, boundary_pressure, prescribed_pressure;
  I boundary_kind, scalar_kind, total_pressure;
  std::map<std::string, F> sst_fields;
  I sst_wall;
  // SST k-omega in the step (fv/steady_solver.cpp:111-127, 484-542): the
  // strain, k and omega fits, the face-jump fit, the transported fields and
  // the boundary values the model reads.
  bool sst = false;
  LsqFit strain_lsq, k_lsq, omega_lsq;
  Vec3F jump_delta, strain_base[3], strain_gradient[3], transpose_stress;
  F jump_inverse, jump_forward, jump_reverse;
  F kinetic, omega, rho_cell, mu_cell, boundary_k, boundary_omega, turbulence_a, turbulence_b;
  I turbulence_spec, turbulence_kind, zero_mask;
  ScalarEquation k_eq, omega_eq;
  F pressure_rhs, correction, correction_boundary, coefficient, face_pressure, pressure_jump;
  F residual_u, residual_v, residual_w, continuity, provisional_continuity, scale_cells, speed;
  F wall_diagonal, wall_rhs, physical_diagonal, consistent_inverse;
  Vec3F velocity_lo, pressure_gradient_lo;
  F pressure_lo, mass_lo, previous_mass, pressure_boundary_lo, imbalance, momentum_delta;

  float limiter_epsilon2[3]{-1, -1, -1};
  double paired_speed = 0, paired_pressure_magnitude = 0;
  double length = 0, bc_speed = 0, pressure_speed = 0, continuity_scale = 0;
  double first_five_continuity[5]{};
  float time_step = 0;
  double velocity_scale = 1;
  IterationReport report;

  Impl(HipDevice& d, PolyMesh mesh, std::vector<PatchCondition> conditions, SolverSettings s);

  F zeros(std::size_t count) {
    F a(dev, count);
    dev.fill(a.data(), 0.0f, count);
    return a;
  }
  I zeros_int(std::size_t count) {
    I a(dev, count);
    dev.fill_int(a.data(), 0, count);
    return a;
  }
  Vec3F zeros3(std::size_t count) {
    Vec3F v(dev, count);
    for (int k = 0; k < 3; ++k) dev.fill(v[k].data(), 0.0f, count);
    return v;
  }
  void zero3(Vec3F& v, std::size_t count) {
    for (int k = 0; k < 3; ++k) dev.fill(v[k].data(), 0.0f, count);
  }
  double max_abs(const F& x, std::size_t count) { return dev.reduce_many({HipDevice::max_abs(x.data(), count)})[0]; }

  // :843-851 (segregated solvers: the maximum cell speed).
  void update_time(double maximum_speed) {
    velocity_scale = std::max({.1 * bc_speed, maximum_speed, pressure_speed, 1e-10});
    time_step = settings.pseudo_time_step > 0 ? float(settings.pseudo_time_step)
                                              : float(.3 * length * settings.pseudo_time_scale / velocity_scale);
  }
  // :857-873 (a fixed step is not refreshed)
  void refresh_time_step() {
    if (settings.pseudo_time_step > 0) return;
    dev.launch(KH_FV_SPEED, n, {velocity.x, velocity.y, velocity.z, speed});
    update_time(max_abs(speed, n));
  }
  // :874-885
  void update_boundaries() {
    dev.launch(KH_FV_VELOCITY_BOUNDARY, nb,
               {owner, boundary_kind, area.x, area.y, area.z, velocity.x, velocity.y, velocity.z,
                prescribed_velocity.x, prescribed_velocity.y, prescribed_velocity.z, mass, boundary_velocity.x,
                boundary_velocity.y, boundary_velocity.z, scalar_kind, ni, float(settings.density)});
  }
  // :599-606
  void apply_backflow_pressure() {
    dev.launch(KH_FV_BACKFLOW_PRESSURE, nb,
               {owner, total_pressure, mass, velocity.x, velocity.y, velocity.z, boundary_pressure, ni,
                float(settings.density)});
  }
  // fv/operators.cpp:296-302
  void gradient(const LsqFit& fit, const F& cell, const F& bc, Vec3F& out) {
    dev.launch(KH_FV_GRADIENT, n,
               {offsets, adjacent, lsq_x, lsq_y, lsq_z, fit.bx, fit.by, fit.bz, fit.inverse, cell, bc, out.x,
                out.y, out.z, n});
  }
  void pressure_boundary(const F& p, const Vec3F& g, const F& prescribed, F& out) {
    dev.launch(KH_FV_PRESSURE_BOUNDARY, nb,
               {owner, boundary_kind, owner_to_face.x, owner_to_face.y, owner_to_face.z, p, g.x, g.y, g.z,
                prescribed, out, ni});
  }
  // :886-889 with fv/operators.cpp:311-318
  void reconstruct_velocity() {
    const auto& fit = velocity_lsq;
    auto &g0 = momentum[0].grad, &g1 = momentum[1].grad, &g2 = momentum[2].grad;
    dev.launch(KH_FV_VECTOR_GRADIENT, n,
               {offsets, adjacent, lsq_x, lsq_y, lsq_z, fit.bx, fit.by, fit.bz, fit.inverse, velocity.x,
                velocity.y, velocity.z, boundary_velocity.x, boundary_velocity.y, boundary_velocity.z, g0.x,
                g0.y, g0.z, g1.x, g1.y, g1.z, g2.x, g2.y, g2.z, n});
  }
  // :896-925 (second-order pressure faces, no poor cells).
  void update_pressure() {
    pressure_boundary(pressure, grad_p, prescribed_pressure, boundary_pressure);
    dev.launch(KH_FV_CELL_PRESSURE_BOUNDARY, nb, {owner, boundary_kind, pressure, boundary_pressure, ni});
    apply_backflow_pressure();
    gradient(pressure_lsq, pressure, boundary_pressure, grad_p);
    pressure_boundary(pressure, grad_p, prescribed_pressure, boundary_pressure);
    apply_backflow_pressure();
    dev.launch(KH_FV_PRESSURE_FACES, nf,
               {owner, neighbour, owner_to_face.x, owner_to_face.y, owner_to_face.z, neighbour_to_face.x,
                neighbour_to_face.y, neighbour_to_face.z, pressure, grad_p.x, grad_p.y, grad_p.z,
                boundary_pressure, face_pressure, pressure_jump, ni});
  }
  // :955-961
  struct Slack {
    float velocity, pressure, flux_per_area;
  };
  Slack paired_slack() const {
    const double ulps = settings.paired_state_slack_ulps * double(std::numeric_limits<float>::epsilon());
    const double s = std::max(bc_speed, paired_speed);
    const double p = std::max(settings.density * s * s, paired_pressure_magnitude);
    return {float(ulps * s), float(ulps * p), float(ulps * settings.density * s)};
  }
  // :1139-1248 (Barth-Jespersen limiter, momentum-weighted faces).
  void update_flux() {
    if (paired) dev.copy(previous_mass.data(), mass.data(), std::size_t(nf) * 4);
    reconstruct_velocity();
    for (int k = 0; k < 3; ++k) {
      auto& eq = momentum[k];
      auto& g = limited_gradient[k];
      if (paired)
        dev.launch(KH_FV_PAIRED_LIMITER_SCALE, n,
                   {offsets, adjacent, faces, sign, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                    neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, velocity[k], velocity_lo[k],
                    boundary_velocity[k], scalar_kind, eq.grad.x, eq.grad.y, eq.grad.z, eq.limiter, g.x, g.y,
                    g.z, limiter_epsilon2[k]});
      else
        dev.launch(KH_FV_LIMITER_SCALE, n,
                   {offsets, adjacent, faces, sign, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                    neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, velocity[k],
                    boundary_velocity[k], eq.grad.x, eq.grad.y, eq.grad.z, eq.limiter, g.x, g.y, g.z,
                    limiter_epsilon2[k]});
    }
    const auto &gu = limited_gradient[0], &gv = limited_gradient[1], &gw = limited_gradient[2];
    dev.launch(KH_FV_MOMENTUM_FACES, nf,
               {owner, neighbour, boundary_kind, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, velocity.x, velocity.y,
                velocity.z, gu.x, gu.y, gu.z, gv.x, gv.y, gv.z, gw.x, gw.y, gw.z, physical_diagonal,
                boundary_velocity.x, boundary_velocity.y, boundary_velocity.z, face_velocity.x, face_velocity.y,
                face_velocity.z, ni});
    dev.launch(KH_FV_RHIE_CHOW, nf,
               {owner, neighbour, boundary_kind, area.x, area.y, area.z, diffusion, volume, physical_diagonal,
                consistent_inverse, face_velocity.x, face_velocity.y, face_velocity.z, pressure_jump, mass,
                coefficient, ni, float(settings.density)});
    if (paired)
      dev.launch(KH_FV_PAIRED_FLUX, nf,
                 {owner, neighbour, boundary_kind, area.x, area.y, area.z, owner_to_face.x, owner_to_face.y,
                  owner_to_face.z, neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z,
                  physical_diagonal, physical_diagonal, volume, diffusion, velocity_lo.x, velocity_lo.y,
                  velocity_lo.z, pressure_lo, pressure_gradient_lo.x, pressure_gradient_lo.y,
                  pressure_gradient_lo.z, previous_mass, mass, mass_lo, ni, float(settings.density),
                  paired_slack().flux_per_area});
  }
  // :765-776: the SST wall-function shear coefficient at walls (laminar: the
  // coefficient argument is an unused placeholder).
  void wall_terms(int k) {
    const F& coefficient_of_wall = sst ? sst_fields.at("wall_shear_coefficient") : prescribed_pressure;
    dev.launch(KH_FV_BOUNDARY_MOMENTUM, n,
               {offsets, faces, boundary_kind, area.x, area.y, area.z, coefficient_of_wall, viscosity, diffusion,
                velocity.x, velocity.y, velocity.z, boundary_velocity.x, boundary_velocity.y,
                boundary_velocity.z, wall_diagonal, wall_rhs, ni, k, sst ? 1 : 0});
  }
  // fv/transport.cpp:83-121 with the paired limiter of with_low (:963-976).
  void assemble(int k, float* values = nullptr) {
    auto& eq = momentum[k];
    if (!values) values = matrix.val.data();
    const F& field = velocity[k];
    const F& bc = boundary_velocity[k];
    if (paired)
      dev.launch(KH_FV_PAIRED_LIMITER, n,
                 {offsets, adjacent, faces, sign, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                  neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, field, velocity_lo[k], bc,
                  scalar_kind, eq.grad.x, eq.grad.y, eq.grad.z, eq.limiter, limiter_epsilon2[k]});
    else
      dev.launch(KH_FV_LIMITER, n,
                 {offsets, adjacent, faces, sign, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                  neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, field, bc, eq.grad.x,
                  eq.grad.y, eq.grad.z, eq.limiter, limiter_epsilon2[k]});
    dev.launch(KH_FV_TRANSPORT_FACES, nf,
               {owner, neighbour, weight, owner_to_face.x, owner_to_face.y, owner_to_face.z, neighbour_to_face.x,
                neighbour_to_face.y, neighbour_to_face.z, nonorth_x, nonorth_y, nonorth_z, eq.grad.x, eq.grad.y,
                eq.grad.z, eq.limiter, mass, viscosity, scalar_kind, face_correction, ni, 1, 0});
    dev.launch(KH_FV_TRANSPORT_ASSEMBLE, n,
               {offsets, faces, sign, adjacent, matrix_entry, matrix.row_ptr, matrix.diag, diffusion, mass,
                viscosity, scalar_kind, bc, face_correction, source[k], volume, field, zero, wall_diagonal,
                wall_rhs, values, eq.rhs, zero, eq.diagonal, zero, eq.consistent_inverse, ni,
                float(settings.density), time_step, float(settings.momentum_relaxation), 0});
  }
  // :1045-1060 (paired momentum without compensated assembly).
  void paired_momentum_residual(int k, F& r, float* values = nullptr) {
    auto& eq = momentum[k];
    if (!values) values = matrix.val.data();
    dev.launch(KH_FV_PAIRED_MOMENTUM_RESIDUAL, n,
               {matrix.row_ptr, matrix.col, values, matrix.diag, eq.diagonal, zero, velocity[k],
                velocity_lo[k], pressure_gradient_lo[k], volume, eq.rhs, zero, r});
  }
  // The three momentum components in one batched BiCGStab solve (no
  // reference counterpart). The reference assembles and solves u, v and w in
  // turn (:1320-1328); component k's assembly reads only its own velocity,
  // and the boundary momentum terms (fv_rans.kern.h:46-74) read the others
  // only through products with the face normal's other components, which are
  // exactly zero on axis-aligned symmetry faces (no wall functions in the
  // laminar solver). On such meshes assembling all three first and solving
  // them together gives bitwise the reference's fields; elsewhere the
  // components stay sequential.
  bool batch_momentum = false;
  bool pressure_tuned = false;
  F momentum_values[3], momentum_deltas[3];
  // With SST wall functions the wall faces' terms read the other components
  // the same way (the tangential projection), so walls must be axis-aligned
  // too.
  bool batch_eligible() {
    const auto ax = area.x.download(), ay = area.y.download(), az = area.z.download();
    const auto kinds = boundary_kind.download();
    for (Index b = 0; b < nb; ++b) {
      if (kinds[b] != 3 && !(sst && kinds[b] == 0)) continue;
      const Index f = ni + b;
      const int nonzero = (ax[f] != 0.0f) + (ay[f] != 0.0f) + (az[f] != 0.0f);
      if (nonzero > 1) return false;
    }
    return true;
  }
  bool set_batched(bool on) {
    batch_momentum = on && batch_eligible();
    for (int k = 0; k < 3; ++k) {
      momentum_values[k] = batch_momentum ? F(dev, matrix.val.size()) : F();
      momentum_deltas[k] = batch_momentum && paired ? zeros(n) : F();
    }
    return batch_momentum;
  }
  void solve_momentum_batched() {
    std::vector<KrylovSolver*> solvers;
    std::vector<DeviceCsr> ops;
    std::vector<float*> xs;
    std::vector<const float*> bs;
    for (int k = 0; k < 3; ++k) {
      wall_terms(k);
      assemble(k, momentum_values[k].data());
      DeviceCsr op = matrix.view();
      op.val = momentum_values[k].data();
      if (paired) {
        paired_momentum_residual(k, residual_u, op.val);
        dev.copy(momentum[k].rhs.data(), residual_u.data(), std::size_t(n) * 4);
        dev.fill(momentum_deltas[k].data(), 0.0f, n);
      }
      momentum[k].solver->refresh(op);
      solvers.push_back(momentum[k].solver.get());
      ops.push_back(op);
      xs.push_back(paired ? momentum_deltas[k].data() : velocity[k].data());
      bs.push_back(momentum[k].rhs.data());
    }
    const auto stats = KrylovSolver::solve_batched(solvers, ops, xs, bs, settings.momentum_linear);
    for (int k = 0; k < 3; ++k) {
      report.momentum[k] = stats[k];
      if (paired)
        dev.launch(KH_FV_PAIRED_ACCUMULATE, n,
                   {momentum_deltas[k], velocity[k], velocity_lo[k], paired_slack().velocity});
    }
    // The shared operator values end as the last component's, as in turn.
    dev.copy(matrix.val.data(), momentum_values[2].data(), matrix.val.size() * sizeof(float));
  }
  SolveStats solve_momentum(int k, F& field) {
    auto& eq = momentum[k];
    eq.solver->refresh(matrix.view());
    return eq.solver->solve(matrix.view(), field.data(), eq.rhs.data(), settings.momentum_linear);
  }
  // :607-614
  void update_momentum_diagonal() {
    dev.launch(KH_FV_MOMENTUM_DIAGONAL, n,
               {momentum[0].diagonal, momentum[1].diagonal, momentum[2].diagonal, momentum[0].consistent_inverse,
                momentum[1].consistent_inverse, momentum[2].consistent_inverse, physical_diagonal,
                consistent_inverse});
  }
  // :1026-1034
  void update_pressure_gradient_lo() {
    zero3(pressure_gradient_lo, n);
    pressure_boundary(pressure_lo, pressure_gradient_lo, correction_boundary, pressure_boundary_lo);
    gradient(pressure_lsq, pressure_lo, pressure_boundary_lo, pressure_gradient_lo);
  }
  // :943-949
  void refresh_paired_scales() {
    if (!paired) return;
    dev.launch(KH_FV_SPEED, n, {velocity.x, velocity.y, velocity.z, speed});
    const auto m = dev.reduce_many({HipDevice::max_abs(speed.data(), n), HipDevice::max_abs(pressure.data(), n)});
    paired_speed = m[0];
    paired_pressure_magnitude = m[1];
  }
  void momentum_source() {
    dev.launch(KH_FV_MOMENTUM_SOURCE, n,
               {offsets, faces, sign, area.x, area.y, area.z, face_pressure, volume, body.x, body.y, body.z,
                source.x, source.y, source.z});
  }
  void step();
  IterationReport monitor();
  ForceReport forces(const std::vector<std::string>& patches, double reference_pressure);
  void initialize_fields(const std::vector<Vec3>& u, const std::vector<double>& p, const std::vector<double>& k,
                         const std::vector<double>& w);

  // SstModel::update (turbulence/sst.cpp:101-158, no intermittency): in[25]
  // as ka_simplec_sst_update's inputs; `gradients` runs where the reference
  // calls reconstruct_scalar_gradients (:126).
  void sst_kernels(const double* o, const float* const* in, const std::function<void()>& gradients);
  // :654-680 (the fused vector fit): strain-rate derivatives of each velocity
  // component, corrected by the face jumps (fv/operators.cpp:254-267).
  void reconstruct_strain() {
    const auto& fit = strain_lsq;
    dev.launch(KH_FV_VECTOR_GRADIENT, n,
               {offsets, adjacent, lsq_x, lsq_y, lsq_z, fit.bx, fit.by, fit.bz, fit.inverse, velocity.x,
                velocity.y, velocity.z, boundary_velocity.x, boundary_velocity.y, boundary_velocity.z,
                strain_base[0].x, strain_base[0].y, strain_base[0].z, strain_base[1].x, strain_base[1].y,
                strain_base[1].z, strain_base[2].x, strain_base[2].y, strain_base[2].z, n});
    for (int k = 0; k < 3; ++k) {
      const auto& base = strain_base[k];
      auto& out = strain_gradient[k];
      dev.launch(KH_FV_FACE_JUMP, nf,
                 {owner, neighbour, owner_to_face.x, owner_to_face.y, owner_to_face.z, neighbour_to_face.x,
                  neighbour_to_face.y, neighbour_to_face.z, velocity[k], boundary_velocity[k], base.x, base.y,
                  base.z, jump_forward, jump_reverse, ni});
      dev.launch(KH_FV_GATHER_JUMP_GRADIENT, n,
                 {offsets, faces, sign, jump_delta.x, jump_delta.y, jump_delta.z, jump_inverse, jump_forward,
                  jump_reverse, base.x, base.y, base.z, out.x, out.y, out.z, n});
    }
  }
  // :646-651
  void turbulence_boundaries() {
    dev.launch(KH_FV_TURBULENCE_BOUNDARY, nb,
               {owner, boundary_kind, turbulence_spec, turbulence_a, turbulence_b, mass, boundary_velocity.x,
                boundary_velocity.y, boundary_velocity.z, kinetic, omega, sst_fields.at("wall_face_omega"),
                boundary_k, boundary_omega, turbulence_kind, ni, float(settings.density),
                float(settings.viscosity)});
  }
  // :681-712 (SST): the strain fit when the velocity changed, the model with
  // the turbulence boundary values and the k and omega fits between its wall
  // and interior passes, then the face effective viscosity.
  void update_turbulence(bool velocity_changed = true) {
    if (!sst) return;
    if (velocity_changed) reconstruct_strain();
    const float* in[25] = {velocity.x.data(), velocity.y.data(), velocity.z.data(),
                           strain_gradient[0].x.data(), strain_gradient[0].y.data(), strain_gradient[0].z.data(),
                           strain_gradient[1].x.data(), strain_gradient[1].y.data(), strain_gradient[1].z.data(),
                           strain_gradient[2].x.data(), strain_gradient[2].y.data(), strain_gradient[2].z.data(),
                           rho_cell.data(), mu_cell.data(), kinetic.data(), omega.data(),
                           k_eq.grad.x.data(), k_eq.grad.y.data(), k_eq.grad.z.data(),
                           omega_eq.grad.x.data(), omega_eq.grad.y.data(), omega_eq.grad.z.data(),
                           boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data()};
    sst_kernels(settings.sst_settings.model, in, [&] {
      turbulence_boundaries();
      gradient(k_lsq, kinetic, boundary_k, k_eq.grad);
      gradient(omega_lsq, omega, boundary_omega, omega_eq.grad);
    });
    dev.launch(KH_FV_EFFECTIVE_VISCOSITY, nf,
               {owner, neighbour, boundary_kind, weight, sst_fields.at("mu_t"), viscosity, ni,
                float(settings.viscosity)});
  }
  // :713-727
  void add_transpose_stress() {
    if (!sst) return;
    const auto &u = momentum[0].grad, &v = momentum[1].grad, &w = momentum[2].grad;
    dev.launch(KH_FV_TRANSPOSE_STRESS, nf,
               {owner, neighbour, boundary_kind, weight, viscosity, area.x, area.y, area.z, u.x, u.y, u.z, v.x,
                v.y, v.z, w.x, w.y, w.z, transpose_stress.x, transpose_stress.y, transpose_stress.z, ni});
    dev.launch(KH_FV_STRESS_SOURCE, n,
               {offsets, faces, sign, volume, transpose_stress.x, transpose_stress.y, transpose_stress.z, source.x,
                source.y, source.z});
  }
  // :778-795 with TransportEquation::assemble_reconstructed
  // (fv/transport.cpp:83-121): Barth-Jespersen limited k or omega with the
  // model's face diffusivity, sources and Patankar sinks, compensated low
  // words; omega's wall cells become fixed rows.
  void assemble_turbulence(bool is_omega) {
    const auto& opt = settings.sst_settings;
    auto& eq = is_omega ? omega_eq : k_eq;
    const F& field = is_omega ? omega : kinetic;
    const F& bc = is_omega ? boundary_omega : boundary_k;
    const F& mu = sst_fields.at(is_omega ? "face_diffusivity_omega" : "face_diffusivity_k");
    const F& src = sst_fields.at(is_omega ? "omega_source" : "k_source");
    const F& sink = sst_fields.at(is_omega ? "omega_sink" : "k_sink");
    const bool second_order = is_omega ? opt.omega_second_order : opt.k_second_order;
    const float dt = opt.pseudo_transient ? time_step : 0.0f;
    const float relaxation = float(is_omega ? opt.omega_relaxation : opt.k_relaxation);
    const bool compensated = opt.compensated_transport;
    dev.launch(KH_FV_LIMITER, n,
               {offsets, adjacent, faces, sign, owner_to_face.x, owner_to_face.y, owner_to_face.z,
                neighbour_to_face.x, neighbour_to_face.y, neighbour_to_face.z, field, bc, eq.grad.x, eq.grad.y,
                eq.grad.z, eq.limiter, -1.0f});
    dev.launch(KH_FV_TRANSPORT_FACES, nf,
               {owner, neighbour, weight, owner_to_face.x, owner_to_face.y, owner_to_face.z, neighbour_to_face.x,
                neighbour_to_face.y, neighbour_to_face.z, nonorth_x, nonorth_y, nonorth_z, eq.grad.x, eq.grad.y,
                eq.grad.z, eq.limiter, mass, mu, turbulence_kind, face_correction, ni, second_order ? 1 : 0, 0});
    dev.launch(KH_FV_TRANSPORT_ASSEMBLE, n,
               {offsets, faces, sign, adjacent, matrix_entry, matrix.row_ptr, matrix.diag, diffusion, mass, mu,
                turbulence_kind, bc, face_correction, src, volume, field, sink, zero, zero, matrix.val, eq.rhs,
                compensated ? eq.rhs_low : zero, eq.diagonal, compensated ? eq.diagonal_low : zero,
                eq.consistent_inverse, ni, float(settings.density), dt, relaxation, compensated ? 1 : 0});
    if (is_omega)
      dev.launch(KH_FV_FIXED_ROWS, n,
                 {matrix.row_ptr, matrix.diag, sst_fields.at("fixed_omega_mask"), sst_fields.at("fixed_omega"),
                  matrix.val, eq.rhs});
  }
  // TransportEquation::physical_residual (fv/transport.cpp:57-62).
  void physical_residual(ScalarEquation& eq, const F& field, F& result, const F* excluded) {
    if (!settings.sst_settings.compensated_transport) {
      dev.launch(KH_CSR_RESIDUAL, n, {matrix.row_ptr, matrix.col, matrix.val, field, eq.rhs, result});
      return;
    }
    dev.launch(KH_FV_COMPENSATED_MOMENTUM_RESIDUAL, n,
               {matrix.row_ptr, matrix.col, matrix.val, matrix.diag, eq.diagonal, eq.diagonal_low, field, zero, zero,
                volume, eq.rhs, eq.rhs_low, result});
    if (excluded) dev.launch(KH_FV_SCALAR_SET_FIXED, n, {*excluded, zero, result});
  }
  // TransportEquation::solve (fv/transport.cpp:123-159), Jacobi BiCGStab; the
  // steady-stopping sums share one host wait. `excluded` holds int32 bits.
  SolveStats solve_scalar(ScalarEquation& eq, F& field, const F* excluded) {
    const auto& opt = settings.sst_settings;
    SolveControl control = opt.linear;
    const bool compensated = opt.compensated_transport;
    if (compensated) physical_residual(eq, field, eq.residual, excluded);
    if (opt.steady_tolerance > 0) {
      if (!compensated)
        dev.launch(KH_CSR_RESIDUAL, n, {matrix.row_ptr, matrix.col, matrix.val, field, eq.rhs, eq.work_residual});
      else
        dev.copy(eq.work_residual.data(), eq.residual.data(), std::size_t(n) * 4);
      if (excluded)
        dev.launch(KH_FV_SCALAR_STEADY_MEASURE, n,
                   {field, eq.diagonal, eq.work_residual, *excluded, eq.work_absolute, eq.work_scale, 1});
      else
        dev.launch(KH_FV_SCALAR_STEADY_MEASURE, n,
                   {field, eq.diagonal, eq.work_residual, matrix.diag, eq.work_absolute, eq.work_scale, 0});
      const auto r = dev.reduce_many({HipDevice::sum(eq.work_absolute.data(), n), HipDevice::sum(eq.work_scale.data(), n),
                                      HipDevice::dot(eq.work_residual.data(), eq.work_residual.data(), n)});
      if (r[0] <= opt.steady_tolerance * r[1] && std::isfinite(r[2]))
        control.abs_tol = std::max(control.abs_tol, std::nextafter(std::sqrt(r[2]), std::numeric_limits<double>::infinity()));
    }
    eq.solver->refresh(matrix.view());
    if (!compensated) return eq.solver->solve(matrix.view(), field.data(), eq.rhs.data(), control);
    dev.fill(eq.delta.data(), 0.0f, n);
    const auto stats = eq.solver->solve(matrix.view(), eq.delta.data(), eq.residual.data(), control);
    dev.launch(KH_AXPBYZ, n, {field, field, eq.delta, 1.0f, 1.0f});
    return stats;
  }
  // :796-818 (no transition model).
  void solve_turbulence() {
    if (!sst) return;
    const auto& opt = settings.sst_settings;
    update_boundaries();
    update_turbulence();
    assemble_turbulence(false);
    report.turbulence[0] = solve_scalar(k_eq, kinetic, nullptr);
    dev.launch(KH_FV_FLOOR, n, {kinetic, float(opt.model[17])});
    update_turbulence(false);
    assemble_turbulence(true);
    const F& mask = sst_fields.at("fixed_omega_mask");
    const F& fixed = sst_fields.at("fixed_omega");
    const bool set_fixed = opt.steady_tolerance > 0 || opt.compensated_transport;
    if (set_fixed) dev.launch(KH_FV_SCALAR_SET_FIXED, n, {mask, fixed, omega});
    report.turbulence[1] = solve_scalar(omega_eq, omega, &mask);
    if (set_fixed) dev.launch(KH_FV_SCALAR_SET_FIXED, n, {mask, fixed, omega});
    dev.launch(KH_FV_FLOOR, n, {omega, float(opt.model[16])});
  }
  // :819-837
  void monitor_turbulence() {
    if (!sst) return;
    const auto measure = [&](bool is_omega) {
      assemble_turbulence(is_omega);
      auto& eq = is_omega ? omega_eq : k_eq;
      const F& q = is_omega ? omega : kinetic;
      const F& mask = sst_fields.at("fixed_omega_mask");
      physical_residual(eq, q, residual_u, is_omega ? &mask : nullptr);
      dev.launch(KH_FV_SCALAR_MONITOR, n,
                 {q, eq.diagonal, residual_u,
                  is_omega ? static_cast<const void*>(mask.data()) : static_cast<const void*>(zero_mask.data()),
                  volume, absolute_residual.x, scale_cells,
                  fluent_scale.x, float(settings.density),
                  settings.sst_settings.pseudo_transient ? time_step : 0.0f});
      const auto r = dev.reduce_many({HipDevice::sum(absolute_residual.x.data(), n), HipDevice::sum(scale_cells.data(), n),
                                      HipDevice::sum(fluent_scale.x.data(), n)});
      const double raw = r[0];
      const auto scaled = [&](double d) {
        return d > 0 ? raw / d : (raw == 0 ? 0 : std::numeric_limits<double>::infinity());
      };
      return ResidualValue{raw, scaled(r[1]), scaled(r[2]), r[2]};
    };
    report.k = measure(false);
    report.omega = measure(true);
  }
};

SteadySolver::Impl::Impl(HipDevice& d, PolyMesh mesh, std::vector<PatchCondition> conditions, SolverSettings s)
    : dev(d), settings(s), host(std::move(mesh)) {
  KA_REQUIRE(s.density > 0 && std::isfinite(s.density) && s.viscosity > 0 && std::isfinite(s.viscosity) &&
                 s.momentum_relaxation > 0 && s.momentum_relaxation <= 1 && s.pressure_relaxation > 0 &&
                 s.pressure_relaxation <= 1 && s.pseudo_time_scale > 0 && s.monitor_interval > 0 &&
                 std::isfinite(s.paired_state_slack_ulps) && s.paired_state_slack_ulps > 0 &&
                 std::isfinite(s.pseudo_time_step) && s.pseudo_time_step >= 0,
             "FV invalid steady solver settings");
  std::map<std::string, PatchCondition> by_name;
  bool outlet = false;
  for (auto& c : conditions) {
    KA_REQUIRE(c.kind >= PatchCondition::Wall && c.kind <= PatchCondition::MassFlowInlet,
               "SIMPLEC: walls, velocity inlets, pressure outlets and symmetry are supported");
    KA_REQUIRE(std::isfinite(c.velocity.x) && std::isfinite(c.velocity.y) && std::isfinite(c.velocity.z),
               "FV nonfinite boundary velocity");
    KA_REQUIRE(std::isfinite(c.pressure), "FV nonfinite outlet pressure");
    outlet = outlet || c.kind == PatchCondition::PressureOutlet || c.kind == PatchCondition::NormalBackflowOutlet;
    KA_REQUIRE(by_name.emplace(c.patch, c).second, "FV duplicate patch condition");
  }
  KA_REQUIRE(by_name.size() == host.patches.size(), "FV supply exactly one condition per patch");
  for (const auto& p : host.patches) {
    const auto it = by_name.find(p.name);
    KA_REQUIRE(it != by_name.end(), "FV missing boundary: " + p.name);
    patch_condition.push_back(it->second);
  }
  renumber_cells_rcm(host);
  geom = compute_geometry(host);
  KA_REQUIRE(host.n_cells > 0, "FV steady solver needs a nonempty mesh");
  // :229-247: a pressure boundary sets the pressure level; a closed domain
  // pins the cell holding the reference point.
  reference = outlet ? -1 : locate_cell(host, geom, s.reference_point);
  {
    const auto cn = build_cell_neighbours(host);
    std::vector<char> seen(host.n_cells, 0);
    std::vector<Index> stack{0};
    seen[0] = 1;
    for (std::size_t k = 0; k < stack.size(); ++k)
      for (Index j = cn.offsets[stack[k]]; j < cn.offsets[stack[k] + 1]; ++j)
        if (!seen[cn.cells[j]]) seen[cn.cells[j]] = 1, stack.push_back(cn.cells[j]);
    KA_REQUIRE(stack.size() == std::size_t(host.n_cells), "FV disconnected domains require separate solvers");
  }
  validate(host, geom);
  n = host.n_cells;
  nf = host.n_faces();
  ni = host.n_internal_faces();
  nb = nf - ni;
  cf = build_cell_faces(host);
  stencil = make_stencil(host, geom);

  // upload_mesh (fv/device_mesh.cpp:208-256).
  owner = to_device(dev, host.owner);
  neighbour = to_device(dev, host.neighbour);
  upload3(dev, area, nf, [&](std::size_t f) { return geom.face_area[f]; });
  const auto rp = [&](std::size_t f) { return sub(geom.face_centroid[f], geom.cell_centroid[host.owner[f]]); };
  const auto rn = [&](std::size_t f) { return sub(geom.face_centroid[f], geom.cell_centroid[host.neighbour[f]]); };
  {
    std::vector<float> w(ni);
    for (Index f = 0; f < ni; ++f) {
      const double on = dot(rp(f), geom.face_area[f]), nn = -dot(rn(f), geom.face_area[f]);
      w[f] = float(nn / (on + nn));
    }
    weight = to_device(dev, w);
  }
  upload3(dev, owner_to_face, nf, rp);
  upload3(dev, neighbour_to_face, ni, rn);
  upload3(dev, owner_to_neighbour, nf, [&](std::size_t f) {
    return f < std::size_t(ni) ? sub(geom.cell_centroid[host.neighbour[f]], geom.cell_centroid[host.owner[f]])
                               : rp(f);
  });
  offsets = to_device(dev, cf.offsets);
  faces = to_device(dev, cf.faces);
  sign = to_device(dev, std::vector<float>(cf.signs.begin(), cf.signs.end()));
  {
    std::vector<float> v(n);
    for (Index c = 0; c < n; ++c) {
      KA_REQUIRE(std::isfinite(geom.cell_volume[c]) && geom.cell_volume[c] > 0,
                 "FV requires finite positive cell volumes");
      v[c] = float(geom.cell_volume[c]);
    }
    volume = to_device(dev, v);
    std::vector<float> wd(nb, 0.0f);
    for (std::size_t p = 0; p < host.patches.size(); ++p)
      if (patch_condition[p].kind == PatchCondition::Wall)
        for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f)
          wd[f - ni] = float(std::abs(dot(rp(f), geom.face_area[f])) / length_of(geom.face_area[f]));
    wall_normal_distance = to_device(dev, wd);
    // fv/device_mesh.cpp:232-245: every cell's distance to the nearest wall
    // triangle (each wall face fanned about its node average), on the device.
    std::vector<double> tri, cent(3 * std::size_t(n));
    for (std::size_t p = 0; p < host.patches.size(); ++p)
      if (patch_condition[p].kind == PatchCondition::Wall)
        for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
          const Index b0 = host.face_offsets[f], s = host.face_offsets[f + 1] - b0;
          const Index* ns = host.face_nodes.data() + b0;
          const Vec3 origin = host.points[ns[0]];
          Vec3 d{};
          for (Index k = 0; k < s; ++k) d = add(d, sub(host.points[ns[k]], origin));
          const Vec3 center = add(origin, scale(d, 1 / double(s)));
          for (Index k = 0; k < s; ++k)
            for (const Vec3& v : {center, host.points[ns[k]], host.points[ns[(k + 1) % s]]}) {
              tri.push_back(v.x);
              tri.push_back(v.y);
              tri.push_back(v.z);
            }
        }
    for (Index c = 0; c < n; ++c) {
      cent[3 * std::size_t(c)] = geom.cell_centroid[c].x;
      cent[3 * std::size_t(c) + 1] = geom.cell_centroid[c].y;
      cent[3 * std::size_t(c) + 2] = geom.cell_centroid[c].z;
    }
    const auto count = tri.size() / 9;
    KA_REQUIRE(count < std::size_t(1) << 31, "FV: too many wall triangles");
    // Triangles in Morton order of their centroids, so that each group of 32
    // is spatially compact and its box prunes well (the patches' face order
    // makes long strips). The nearest distance is a minimum over the same set
    // of exact FP64 distances, so the order does not change it.
    if (count > 1) {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      std::vector<double> mid(3 * count);
      for (std::size_t t = 0; t < count; ++t)
        for (int c = 0; c < 3; ++c) {
          mid[3 * t + c] = (tri[9 * t + c] + tri[9 * t + 3 + c] + tri[9 * t + 6 + c]) / 3;
          lo[c] = std::min(lo[c], mid[3 * t + c]);
          hi[c] = std::max(hi[c], mid[3 * t + c]);
        }
      const auto spread = [](std::uint64_t v) {  // 21 bits to every third bit
        v &= 0x1fffffULL;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
      };
      std::vector<std::pair<std::uint64_t, std::uint32_t>> order(count);
      for (std::size_t t = 0; t < count; ++t) {
        std::uint64_t key = 0;
        for (int c = 0; c < 3; ++c) {
          const double extent = hi[c] - lo[c];
          const double u = extent > 0 ? (mid[3 * t + c] - lo[c]) / extent : 0.0;
          key |= spread(std::uint64_t(std::min(u, 1.0) * 2097151.0)) << c;
        }
        order[t] = {key, std::uint32_t(t)};
      }
      std::sort(order.begin(), order.end());
      std::vector<double> sorted(tri.size());
      for (std::size_t t = 0; t < count; ++t)
        std::copy_n(tri.begin() + 9 * std::size_t(order[t].second), 9, sorted.begin() + 9 * t);
      tri.swap(sorted);
    }
    // Bounding box of each group of 32 consecutive triangles (the kernel's
    // pruning unit; the reference's Bounds::extend over the vertices).
    constexpr std::size_t kGroup = 32;
    std::vector<double> box;
    for (std::size_t g = 0; g < count; g += kGroup) {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (std::size_t t = g; t < std::min(count, g + kGroup); ++t)
        for (int v = 0; v < 3; ++v)
          for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], tri[9 * t + 3 * v + c]);
            hi[c] = std::max(hi[c], tri[9 * t + 3 * v + c]);
          }
      box.insert(box.end(), {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
    }
    if (tri.empty()) tri.assign(9, 0.0), box.assign(6, 0.0);
    DeviceArray<double> tri_dev(dev, tri.size()), box_dev(dev, box.size()), cent_dev(dev, cent.size());
    tri_dev.upload(tri);
    box_dev.upload(box);
    cent_dev.upload(cent);
    wall_distance = F(dev, n);
    dev.launch(KH_WALL_DISTANCE, n, {cent_dev, tri_dev, box_dev, int(count), wall_distance});
    dev.synchronize();
  }
  adjacent = to_device(dev, stencil.adjacent);
  matrix_entry = to_device(dev, stencil.matrix_entry);
  lsq_x = to_device(dev, stencil.lsq_x);
  lsq_y = to_device(dev, stencil.lsq_y);
  lsq_z = to_device(dev, stencil.lsq_z);
  diffusion = to_device(dev, stencil.diffusion);
  nonorth_x = to_device(dev, stencil.nonorth_x);
  nonorth_y = to_device(dev, stencil.nonorth_y);
  nonorth_z = to_device(dev, stencil.nonorth_z);

  // Fields (:262-296).
  velocity = zeros3(n);
  face_velocity = zeros3(nf);
  body = zeros3(n);
  source = zeros3(n);
  boundary_velocity = zeros3(nb);
  for (auto& g : limited_gradient) g = zeros3(n);
  grad_p = zeros3(n);
  grad_correction = zeros3(n);
  absolute_residual = zeros3(n);
  fluent_scale = zeros3(n);
  pressure = zeros(n);
  mass = zeros(nf);
  viscosity = F(dev, nf);
  dev.fill(viscosity.data(), float(s.viscosity), nf);
  boundary_pressure = zeros(nb);
  scalar_kind = zeros_int(nb);
  pressure_rhs = zeros(n);
  correction = zeros(n);
  correction_boundary = zeros(nb);
  coefficient = zeros(nf);
  face_pressure = zeros(nf);
  pressure_jump = zeros(nf);
  for (F* v : {&residual_u, &residual_v, &residual_w, &continuity, &provisional_continuity, &scale_cells, &speed,
               &wall_diagonal, &wall_rhs})
    *v = zeros(n);
  paired = s.paired_state;
  if (paired) {
    velocity_lo = zeros3(n);
    pressure_gradient_lo = zeros3(n);
    pressure_lo = zeros(n);
    mass_lo = zeros(nf);
    previous_mass = zeros(nf);
    pressure_boundary_lo = zeros(nb);
    imbalance = zeros(n);
    momentum_delta = zeros(n);
  }
  total_pressure = zeros_int(nb);
  physical_diagonal = F(dev, n);
  dev.fill(physical_diagonal.data(), 1.0f, n);
  consistent_inverse = zeros(n);

  // Boundary kinds, prescribed values and LSQ observations (:314-400).
  std::vector<Index> kinds(nb, 0), totals(nb, 0);
  std::vector<float> pu(nb, 0.0f), pv(nb, 0.0f), pw(nb, 0.0f), pp(nb, 0.0f);
  // :305-311: a mass-flow inlet's total and projected areas.
  std::vector<double> inlet_area(host.patches.size(), 0.0), inlet_projected(host.patches.size(), 0.0);
  for (std::size_t p = 0; p < host.patches.size(); ++p)
    if (patch_condition[p].kind == PatchCondition::MassFlowInlet) {
      const Vec3 dir = patch_condition[p].direction;
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
        inlet_area[p] += length_of(geom.face_area[f]);
        if (patch_condition[p].has_direction && length_of(dir) > 0)
          inlet_projected[p] -= dot(scale(dir, 1 / length_of(dir)), geom.face_area[f]);
      }
    }
  double pmin = std::numeric_limits<double>::infinity(), pmax = -pmin;
  for (std::size_t p = 0; p < host.patches.size(); ++p)
    for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
      const Index b = f - ni;
      const auto& c = patch_condition[p];
      kinds[b] = c.kind == PatchCondition::MassFlowInlet ? PatchCondition::VelocityInlet : c.kind;
      Vec3 v{};
      if (c.kind == PatchCondition::Wall || c.kind == PatchCondition::VelocityInlet ||
          c.kind == PatchCondition::PressureOutlet)
        v = c.profile && c.kind != PatchCondition::Wall ? c.profile(geom.face_centroid[f]) : c.velocity;
      if (c.kind == PatchCondition::MassFlowInlet) {  // fv/boundary_physics.cpp:14-26
        const Vec3 a = geom.face_area[f];
        KA_REQUIRE(std::isfinite(c.mass_flow) && c.mass_flow >= 0 && inlet_area[p] > 0 && s.density > 0,
                   "mass-flow-inlet: nonnegative mass flow and positive area/density required");
        if (c.has_direction) {
          KA_REQUIRE(std::isfinite(c.direction.x) && std::isfinite(c.direction.y) && std::isfinite(c.direction.z) &&
                         length_of(c.direction) > 0 && inlet_projected[p] > 0,
                     "mass-flow-inlet: nonzero inward direction required");
          const Vec3 d = scale(c.direction, 1 / length_of(c.direction));
          KA_REQUIRE(dot(d, a) < 0, "mass-flow-inlet: direction must enter every face");
          v = scale(d, c.mass_flow / (s.density * inlet_projected[p]));
        } else {
          v = scale(a, -c.mass_flow / (s.density * inlet_area[p] * length_of(a)));
        }
      }
      KA_REQUIRE(std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z), "FV nonfinite field value");
      if (c.kind == PatchCondition::PressureOutlet || c.kind == PatchCondition::NormalBackflowOutlet) {
        totals[b] = c.total_backflow ? 1 : 0;
        pp[b] = float(c.pressure);
        pmin = std::min(pmin, c.pressure);
        pmax = std::max(pmax, c.pressure);
      }
      pu[b] = float(v.x);
      pv[b] = float(v.y);
      pw[b] = float(v.z);
      bc_speed = std::max(bc_speed, length_of(v));
    }
  pressure_speed = outlet ? .1 * std::sqrt((pmax - pmin) / s.density) : 0;
  total_pressure = to_device(dev, totals);
  // :351-358: velocity observed at walls and inlets, its normal foot at
  // symmetry planes; pressure observed at outlets.
  std::vector<Observation> observe_velocity(nb), observe_pressure(nb);
  for (Index b = 0; b < nb; ++b) {
    observe_velocity[b] = kinds[b] == 3                  ? Observation::NormalFootValue
                          : (kinds[b] == 0 || kinds[b] == 1) ? Observation::FaceValue
                                                             : Observation::Omitted;
    observe_pressure[b] = (kinds[b] == 2 || kinds[b] == 4) ? Observation::FaceValue : Observation::Omitted;
  }
  boundary_kind = to_device(dev, kinds);
  prescribed_pressure = to_device(dev, pp);
  prescribed_velocity.x = to_device(dev, pu);
  prescribed_velocity.y = to_device(dev, pv);
  prescribed_velocity.z = to_device(dev, pw);

  CsrHost pattern;
  pattern.n_rows = n;
  pattern.row_ptr = stencil.row_ptr;
  pattern.col = stencil.col;
  pattern.diag = stencil.diag;
  pattern.val.assign(stencil.col.size(), 0.0);
  matrix = OwnedCsr(dev, pattern);

  // The projected face reconstruction (fv/face_projection.cpp:136-150) corrects
  // faces whose cell-to-face vectors leave the face normal; none may here.
  for (Index f = 0; f < nf; ++f) {
    if (f >= ni && observe_velocity[f - ni] != Observation::FaceValue) continue;
    const Index o = host.owner[f];
    const Vec3 unit = scale(geom.face_area[f], 1 / length_of(geom.face_area[f]));
    const Vec3 a = sub(geom.face_centroid[f], geom.cell_centroid[o]);
    const Vec3 b = f < ni ? sub(geom.cell_centroid[host.neighbour[f]], geom.face_centroid[f]) : Vec3{};
    const double dp = dot(a, unit), dn = dot(b, unit);
    const double roundoff =
        256 * std::numeric_limits<double>::epsilon() *
        std::max({1., length_of(geom.face_centroid[f]), length_of(geom.cell_centroid[o]),
                  f < ni ? length_of(geom.cell_centroid[host.neighbour[f]]) : 0.});
    KA_REQUIRE(length_of(sub(a, scale(unit, dp))) <= roundoff && length_of(sub(b, scale(unit, dn))) <= roundoff,
               "SIMPLEC: meshes that need the projected face reconstruction are not supported");
  }
  face_correction = zeros(nf);
  zero = zeros(n);
  CsrHost identity = pattern;
  for (auto j : identity.diag) identity.val[j] = 1;
  for (auto& eq : momentum) {
    eq.rhs = zeros(n);
    eq.diagonal = zeros(n);
    eq.consistent_inverse = zeros(n);
    eq.limiter = zeros(n);
    eq.grad = zeros3(n);
    eq.solver = std::make_unique<KrylovSolver>(dev, SolverKind::BicgstabJacobi, identity, AmgOptions{});
  }
  velocity_lsq = least_squares(dev, host, geom, cf, observe_velocity);
  {
    const auto gain = extrapolation_gain(host, geom, cf, observe_pressure);
    for (Index b = 0; b < nb; ++b)
      if (observe_pressure[b] == Observation::Omitted && gain[b] > 1) observe_pressure[b] = Observation::FaceValue;
  }
  pressure_lsq = least_squares(dev, host, geom, cf, observe_pressure);
  {
    const auto q = orthogonal_quality(host, geom);
    KA_REQUIRE(std::all_of(q.begin(), q.end(), [](double v) { return v > .05; }),
               "SIMPLEC: cells at or below the poor-quality threshold are not supported");
  }
  sst = s.sst;
  if (sst) {  // :471-542
    const auto& t = s.sst_settings;
    const double* o = t.model;
    KA_REQUIRE(t.k_relaxation > 0 && t.k_relaxation <= 1 && t.omega_relaxation > 0 && t.omega_relaxation <= 1,
               "FV invalid SST relaxation");
    KA_REQUIRE(std::isfinite(o[4]) && o[4] > 0 && o[13] > 0 && o[15] > 0 && o[16] > 0 && o[17] > 0 && o[18] > 0 &&
                   std::isfinite(o[6]) && o[6] >= 0 && !(o[10] != 0 && o[11] != 0),
               "SST: invalid options");
    KA_REQUIRE(std::isfinite(t.steady_tolerance) && t.steady_tolerance >= 0, "FV invalid scalar steady tolerance");
    std::vector<Observation> strain_rows(nb), observe_k(nb), observe_omega(nb);
    for (Index b = 0; b < nb; ++b) {
      strain_rows[b] = kinds[b] != 0 ? Observation::FaceValue : Observation::Omitted;
      observe_k[b] = kinds[b] == 1 ? Observation::FaceValue : Observation::Omitted;
      observe_omega[b] = (kinds[b] == 0 || kinds[b] == 1) ? Observation::FaceValue : Observation::Omitted;
      if (kinds[b] == 3) observe_k[b] = observe_omega[b] = Observation::NormalFootValue;
    }
    strain_lsq = least_squares(dev, host, geom, cf, strain_rows);
    {  // face_jump_least_squares (fv/operators.cpp:234-253)
      std::vector<Vec3> rows;
      rows.reserve(cf.faces.size());
      std::vector<float> inverse(6 * std::size_t(n));
      for (Index c = 0; c < n; ++c) {
        double a[6]{};
        for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
          const Vec3 r = sub(geom.face_centroid[cf.faces[j]], geom.cell_centroid[c]);
          const double w = 1 / dot(r, r);
          rows.push_back(scale(r, w));
          a[0] += w * r.x * r.x;
          a[1] += w * r.x * r.y;
          a[2] += w * r.x * r.z;
          a[3] += w * r.y * r.y;
          a[4] += w * r.y * r.z;
          a[5] += w * r.z * r.z;
        }
        const double det = a[0] * (a[3] * a[5] - a[4] * a[4]) - a[1] * (a[1] * a[5] - a[4] * a[2]) +
                           a[2] * (a[1] * a[4] - a[3] * a[2]);
        KA_REQUIRE(det > 1e-14, "FV face-jump fit is rank deficient");
        const double inv[6]{(a[3] * a[5] - a[4] * a[4]) / det, (a[2] * a[4] - a[1] * a[5]) / det,
                            (a[1] * a[4] - a[2] * a[3]) / det, (a[0] * a[5] - a[2] * a[2]) / det,
                            (a[1] * a[2] - a[0] * a[4]) / det, (a[0] * a[3] - a[1] * a[1]) / det};
        for (int k = 0; k < 6; ++k) inverse[c + k * std::size_t(n)] = float(inv[k]);
      }
      upload3(dev, jump_delta, rows.size(), [&](std::size_t j) { return rows[j]; });
      jump_inverse = to_device(dev, inverse);
    }
    jump_forward = zeros(nf);
    jump_reverse = zeros(std::max<Index>(ni, 1));
    for (int k = 0; k < 3; ++k) {
      strain_base[k] = zeros3(n);
      strain_gradient[k] = zeros3(n);
    }
    k_lsq = least_squares(dev, host, geom, cf, observe_k);
    omega_lsq = least_squares(dev, host, geom, cf, observe_omega);
    kinetic = F(dev, n);
    dev.fill(kinetic.data(), float(o[17]), n);
    omega = F(dev, n);
    dev.fill(omega.data(), float(o[16]), n);
    rho_cell = F(dev, n);
    dev.fill(rho_cell.data(), float(s.density), n);
    mu_cell = F(dev, n);
    dev.fill(mu_cell.data(), float(s.viscosity), n);
    boundary_k = zeros(nb);
    boundary_omega = zeros(nb);
    turbulence_kind = zeros_int(nb);
    zero_mask = zeros_int(n);
    transpose_stress = zeros3(nf);
    std::vector<Index> specs(nb, 0), wall(nb, 0);
    std::vector<float> ta(nb, 0.0f), tb(nb, 0.0f);
    for (std::size_t p = 0; p < host.patches.size(); ++p) {
      const auto& c = patch_condition[p];
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f)
        wall[f - ni] = c.kind == PatchCondition::Wall ? 1 : 0;
      if (c.kind == PatchCondition::Wall || c.kind == PatchCondition::Symmetry) continue;
      KA_REQUIRE(c.turbulence >= 0 && c.turbulence <= 2, "FV missing turbulence specification: " + host.patches[p].name);
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
        const Index b = f - ni;
        specs[b] = c.turbulence;
        ta[b] = float(c.turbulence_a);
        tb[b] = float(c.turbulence_b);
        KA_REQUIRE(ta[b] >= 0 && tb[b] > 0 && std::isfinite(ta[b]) && std::isfinite(tb[b]),
                   "FV invalid turbulence specification: " + host.patches[p].name);
      }
    }
    turbulence_spec = to_device(dev, specs);
    turbulence_a = to_device(dev, ta);
    turbulence_b = to_device(dev, tb);
    sst_wall = to_device(dev, wall);
    CsrHost unit = pattern;
    for (auto j : unit.diag) unit.val[j] = 1;
    for (auto* eq : {&k_eq, &omega_eq}) {
      for (F* v : {&eq->rhs, &eq->diagonal, &eq->consistent_inverse, &eq->limiter, &eq->rhs_low, &eq->diagonal_low,
                   &eq->delta, &eq->residual, &eq->work_residual, &eq->work_absolute, &eq->work_scale})
        *v = zeros(n);
      eq->grad = zeros3(n);
      eq->solver = std::make_unique<KrylovSolver>(dev, SolverKind::BicgstabJacobi, unit, AmgOptions{});
    }
    for (const char* name : {"mu_t", "k_source", "k_sink", "omega_source", "omega_sink", "strain", "vorticity", "f1",
                             "f2", "cdkw", "diffusivity_k", "diffusivity_omega", "production_k", "fixed_omega",
                             "wall_production", "fixed_omega_mask", "selected_wall_face"})
      sst_fields[name] = zeros(n);
    for (const char* name : {"wall_tau", "wall_u_star", "wall_u_tau", "wall_y_plus", "wall_y_plus_utau",
                             "wall_shear_x", "wall_shear_y", "wall_shear_z", "wall_shear_coefficient",
                             "wall_face_omega", "wall_cell_omega", "face_production"})
      sst_fields[name] = zeros(nb);
    sst_fields["face_diffusivity_k"] = zeros(nf);
    sst_fields["face_diffusivity_omega"] = zeros(nf);
  }
  // The initial pressure operator (:553-572): stored-float M-matrix values.
  CsrHost initial = pattern;
  for (Index c = 0; c < n; ++c) {
    UpperSum ap{0, 0};
    for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
      const Index f = cf.faces[j];
      const Index other = f < ni ? (host.owner[f] == c ? host.neighbour[f] : host.owner[f]) : -1;
      if (other < 0 && kinds[f - ni] != 2 && kinds[f - ni] != 4) continue;
      const Vec3 dvec = sub(other >= 0 ? geom.cell_centroid[other] : geom.face_centroid[f], geom.cell_centroid[c]);
      const float a = float(dot(geom.face_area[f], geom.face_area[f]) / std::abs(dot(geom.face_area[f], dvec)));
      if (other < 0 || other == reference) ap.add(a);
      if (other >= 0 && other != reference && c != reference) {
        const auto first = initial.col.begin() + initial.row_ptr[c], last = initial.col.begin() + initial.row_ptr[c + 1];
        const auto at = std::lower_bound(first, last, other) - initial.col.begin();
        initial.val[at] = float(float(initial.val[at]) - a);
      }
    }
    for (Index j = initial.row_ptr[c]; j < initial.row_ptr[c + 1]; ++j)
      if (j != initial.diag[c]) ap.add(-float(initial.val[j]));
    initial.val[initial.diag[c]] = c == reference ? 1 : ap.value();
  }
  AmgOptions amg;
  amg.structure = Structure::SymmetricMMatrix;
  if (sst) amg.pre_sweeps = amg.post_sweeps = 2;  // :573-578
  pressure_solver = std::make_unique<KrylovSolver>(dev, SolverKind::AmgPcg, initial, amg);
  pressure_initial = std::move(initial);
  const double total = std::accumulate(geom.cell_volume.begin(), geom.cell_volume.end(), 0.);
  Vec3 lo = host.points.front(), hi = lo;
  for (auto p : host.points) {
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  length = std::min(std::cbrt(total), std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z}));
  update_time(bc_speed);
  dev.fill(momentum[0].diagonal.data(), 1.0f, n);
  dev.fill(momentum[0].consistent_inverse.data(), 0.0f, n);
  update_boundaries();
  update_pressure();
  update_flux();
  set_batched(s.batched_momentum);
}

// :1303-1369
void SteadySolver::Impl::step() {
  refresh_time_step();
  update_boundaries();
  update_pressure();
  update_turbulence();
  momentum_source();
  reconstruct_velocity();
  add_transpose_stress();
  if (batch_momentum) solve_momentum_batched();
  for (int k = 0; k < 3 && !batch_momentum; ++k) {
    wall_terms(k);
    assemble(k);
    if (!paired) {
      report.momentum[k] = solve_momentum(k, velocity[k]);
      continue;
    }
    // Delta form: the same Krylov iterates, with the pair's residual.
    paired_momentum_residual(k, residual_u);
    dev.copy(momentum[k].rhs.data(), residual_u.data(), std::size_t(n) * 4);
    dev.fill(momentum_delta.data(), 0.0f, n);
    report.momentum[k] = solve_momentum(k, momentum_delta);
    dev.launch(KH_FV_PAIRED_ACCUMULATE, n, {momentum_delta, velocity[k], velocity_lo[k], paired_slack().velocity});
  }
  update_momentum_diagonal();
  update_boundaries();
  update_flux();
  if (paired)
    dev.launch(KH_FV_PAIRED_CONTINUITY, n, {offsets, faces, sign, mass, mass_lo, imbalance, provisional_continuity});
  else
    dev.launch(KH_FV_CONTINUITY, n, {offsets, faces, sign, mass, provisional_continuity});
  dev.launch(KH_FV_PRESSURE_ASSEMBLE, n,
             {offsets, faces, sign, adjacent, matrix_entry, matrix.row_ptr, matrix.diag, coefficient, mass,
              matrix.val, pressure_rhs, reference});
  if (paired) dev.launch(KH_FV_PAIRED_PRESSURE_RHS, n, {imbalance, pressure_rhs, reference});
  pressure_solver->refresh(matrix.view());
  if (!pressure_tuned) {  // the first assembled operator: pick the coarse tail by timing
    pressure_solver->autotune_tail(matrix.view(), 3);
    pressure_tuned = true;
  }
  dev.fill(correction.data(), 0.0f, n);
  report.pressure = pressure_solver->solve(matrix.view(), correction.data(), pressure_rhs.data(), settings.pressure_linear);
  if (paired)
    dev.launch(KH_FV_PAIRED_CORRECT_FLUX, nf,
               {owner, neighbour, area.x, area.y, area.z, correction, coefficient, mass, mass_lo, ni,
                paired_slack().flux_per_area});
  else
    dev.launch(KH_FV_CORRECT_FLUX, nf, {owner, neighbour, correction, coefficient, mass, ni});
  // The correction has homogeneous Neumann data on fixed-velocity boundaries.
  zero3(grad_correction, n);
  pressure_boundary(correction, grad_correction, correction_boundary, boundary_pressure);
  gradient(pressure_lsq, correction, boundary_pressure, grad_correction);
  if (paired) {
    const auto slack = paired_slack();
    dev.launch(KH_FV_PAIRED_CORRECT_VELOCITY, n,
               {volume, consistent_inverse, physical_diagonal, grad_correction.x, grad_correction.y,
                grad_correction.z, correction, velocity.x, velocity.y, velocity.z, pressure, velocity_lo.x,
                velocity_lo.y, velocity_lo.z, pressure_lo, float(settings.pressure_relaxation), slack.velocity,
                slack.pressure});
    update_pressure_gradient_lo();
  } else {
    dev.launch(KH_FV_CORRECT_VELOCITY, n,
               {volume, consistent_inverse, physical_diagonal, grad_correction.x, grad_correction.y,
                grad_correction.z, correction, velocity.x, velocity.y, velocity.z, pressure,
                float(settings.pressure_relaxation)});
  }
  solve_turbulence();
  // finish_step (:950-954)
  ++report.iteration;
  if (report.iteration <= 5 || report.iteration % 20 == 0) refresh_paired_scales();
  if (report.iteration <= 5 || report.iteration % settings.monitor_interval == 0) monitor();
}

// :1417-1477
IterationReport SteadySolver::Impl::monitor() {
  refresh_time_step();
  update_boundaries();
  update_pressure();
  update_turbulence();
  momentum_source();
  reconstruct_velocity();
  add_transpose_stress();
  F* residuals[]{&residual_u, &residual_v, &residual_w};
  for (int k = 0; k < 3; ++k) {
    wall_terms(k);
    assemble(k);
    if (paired)
      paired_momentum_residual(k, *residuals[k]);
    else
      dev.launch(KH_CSR_RESIDUAL, n, {matrix.row_ptr, matrix.col, matrix.val, velocity[k], momentum[k].rhs, *residuals[k]});
  }
  update_momentum_diagonal();
  dev.launch(KH_FV_MONITOR_CELL, n,
             {offsets, faces, sign, mass, velocity.x, velocity.y, velocity.z, physical_diagonal, residual_u,
              residual_v, residual_w, continuity, scale_cells, speed, absolute_residual.x, absolute_residual.y,
              absolute_residual.z});
  if (paired) dev.launch(KH_FV_PAIRED_CONTINUITY, n, {offsets, faces, sign, mass, mass_lo, imbalance, continuity});
  dev.launch(KH_FV_FLUENT_MOMENTUM_SCALE, n,
             {momentum[0].diagonal, momentum[1].diagonal, momentum[2].diagonal, volume, speed, fluent_scale.x,
              fluent_scale.y, fluent_scale.z, float(settings.density), time_step});
  // All of the monitor's sums behind one wait (each equals its single call).
  const auto r = dev.reduce_many(
      {HipDevice::sum(scale_cells.data(), n), HipDevice::sum(provisional_continuity.data(), n),
       HipDevice::sum(continuity.data(), n), HipDevice::sum(absolute_residual.x.data(), n),
       HipDevice::sum(absolute_residual.y.data(), n), HipDevice::sum(absolute_residual.z.data(), n),
       HipDevice::sum(fluent_scale.x.data(), n), HipDevice::sum(fluent_scale.y.data(), n),
       HipDevice::sum(fluent_scale.z.data(), n), HipDevice::max_abs(speed.data(), n)});
  const double denominator = r[0], raw = r[1];
  report.corrected_mass_imbalance = r[2];
  if (report.iteration > 0 && report.iteration <= 5) first_five_continuity[report.iteration - 1] = raw;
  continuity_scale = *std::max_element(std::begin(first_five_continuity), std::end(first_five_continuity));
  const auto scaled = [](double v, double d) {
    return d > 0 ? v / d : (v == 0 ? 0 : std::numeric_limits<double>::infinity());
  };
  ResidualValue* values[]{&report.u, &report.v, &report.w};
  for (int k = 0; k < 3; ++k) {
    values[k]->raw = r[3 + k];
    values[k]->scaled = scaled(values[k]->raw, denominator);
    values[k]->fluent_normalization = r[6 + k];
    values[k]->fluent_scaled = scaled(values[k]->raw, values[k]->fluent_normalization);
  }
  report.continuity = {raw, scaled(raw, continuity_scale), scaled(raw, continuity_scale), continuity_scale};
  report.maximum_velocity = r[9];
  report.pseudo_time_step = time_step;
  monitor_turbulence();
  return report;
}

// :1711-1736 with viscous_face_forces (:1373-1416) for laminar walls.
ForceReport SteadySolver::Impl::forces(const std::vector<std::string>& patches, double reference_pressure) {
  KA_REQUIRE(std::isfinite(reference_pressure), "FV nonfinite force reference pressure");
  std::vector<std::string> names(patches);
  std::sort(names.begin(), names.end());
  KA_REQUIRE(std::adjacent_find(names.begin(), names.end()) == names.end(), "FV duplicate force patch");
  std::vector<float> selected(nb, 0.0f);
  ForceReport out;
  for (const auto& name : names) {
    std::size_t p = 0;
    while (p < host.patches.size() && host.patches[p].name != name) ++p;
    KA_REQUIRE(p < host.patches.size() && patch_condition[p].kind == PatchCondition::Wall,
               "FV force patch must name a wall: " + name);
    for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
      selected[f - ni] = 1.0f;
      out.area_vector = {out.area_vector.x + geom.face_area[f].x, out.area_vector.y + geom.face_area[f].y,
                         out.area_vector.z + geom.face_area[f].z};
    }
  }
  update_boundaries();
  update_pressure();
  reconstruct_velocity();
  const F mask = to_device(dev, selected);
  Vec3F pf = zeros3(nb), vf = zeros3(nb);
  dev.launch(KH_FV_WALL_PRESSURE_FORCE, nb,
             {mask, area.x, area.y, area.z, face_pressure, pf.x, pf.y, pf.z, ni, float(reference_pressure)});
  const auto &gu = momentum[0].grad, &gv = momentum[1].grad, &gw = momentum[2].grad;
  if (sst) {  // viscous_face_forces (:1376-1380): the wall-function shear
    update_turbulence();
    dev.launch(KH_FV_SST_FORCE, nb,
               {mask, area.x, area.y, area.z, sst_fields.at("wall_shear_x"), sst_fields.at("wall_shear_y"),
                sst_fields.at("wall_shear_z"), vf.x, vf.y, vf.z, ni});
  } else
  dev.launch(KH_FV_WALL_VISCOUS_FORCE, nb,
             {owner, mask, area.x, area.y, area.z, owner_to_face.x, owner_to_face.y, owner_to_face.z, velocity.x,
              velocity.y, velocity.z, gu.x, gu.y, gu.z, gv.x, gv.y, gv.z, gw.x, gw.y, gw.z, boundary_velocity.x,
              boundary_velocity.y, boundary_velocity.z, vf.x, vf.y, vf.z, ni, float(settings.viscosity)});
  const auto r = dev.reduce_many({HipDevice::sum(pf.x.data(), nb), HipDevice::sum(pf.y.data(), nb),
                                  HipDevice::sum(pf.z.data(), nb), HipDevice::sum(vf.x.data(), nb),
                                  HipDevice::sum(vf.y.data(), nb), HipDevice::sum(vf.z.data(), nb)});
  out.pressure = {r[0], r[1], r[2]};
  out.viscous = {r[3], r[4], r[5]};
  out.reference_pressure = reference_pressure;
  return out;
}

double ForceReport::coefficient(Vec3 direction, double density, double velocity, double area,
                                double pressure) const {
  KA_REQUIRE(density > 0 && velocity > 0 && area > 0, "FV positive force references required");
  const Vec3 f = force();
  const double shift = reference_pressure - pressure;
  const Vec3 adjusted{f.x + area_vector.x * shift, f.y + area_vector.y * shift, f.z + area_vector.z * shift};
  return dot(adjusted, direction) / (.5 * density * velocity * velocity * area);
}

SteadySolver::SteadySolver(HipDevice& dev, PolyMesh mesh, std::vector<PatchCondition> conditions,
                           SolverSettings settings)
    : impl_(std::make_unique<Impl>(dev, std::move(mesh), std::move(conditions), settings)) {}
SteadySolver::~SteadySolver() = default;
void SteadySolver::iterate(int count) {
  KA_REQUIRE(count >= 0, "FV negative iteration count");
  while (count-- > 0) impl_->step();
}
void SteadySolver::initialize(const std::vector<Vec3>& u, const std::vector<double>& p) {
  auto& s = *impl_;
  KA_REQUIRE(!s.sst, "FV SST initialization requires initialize_fields with k and omega");
  const auto n = std::size_t(s.n);
  KA_REQUIRE(u.size() == n && p.size() == n, "FV initialization needs one velocity and pressure per cell");
  const double ref = s.reference >= 0 ? p[std::size_t(s.reference)] : 0;
  std::vector<float> ux(n), uy(n), uz(n), pf(n);
  for (std::size_t c = 0; c < n; ++c) {
    const double value = p[c] - ref;
    KA_REQUIRE(std::isfinite(value), "FV nonfinite pressure initialization");
    ux[c] = float(u[c].x);
    uy[c] = float(u[c].y);
    uz[c] = float(u[c].z);
    pf[c] = float(value);
  }
  if (s.paired)  // reset_paired (:977-988)
    for (auto* v : {&s.velocity_lo.x, &s.velocity_lo.y, &s.velocity_lo.z, &s.pressure_gradient_lo.x,
                    &s.pressure_gradient_lo.y, &s.pressure_gradient_lo.z, &s.pressure_lo, &s.mass_lo})
      s.dev.fill(v->data(), 0.0f, v->size());
  s.velocity.x.upload(ux);
  s.velocity.y.upload(uy);
  s.velocity.z.upload(uz);
  s.pressure.upload(pf);
  s.update_boundaries();
  s.update_pressure();
  s.dev.fill(s.physical_diagonal.data(), 1.0f, n);
  s.dev.fill(s.consistent_inverse.data(), 0.0f, n);
  s.update_flux();
}
void SteadySolver::Impl::sst_kernels(const double* o, const float* const* in,
                                     const std::function<void()>& gradients) {
  KA_REQUIRE(std::isfinite(o[4]) && o[4] > 0 && o[13] > 0 && o[15] > 0 && o[16] > 0 && o[17] > 0 && o[18] > 0 &&
                 std::isfinite(o[6]) && o[6] >= 0 && !(o[10] != 0 && o[11] != 0),
             "SST: invalid options");
  for (int k = 0; k < 25; ++k) KA_REQUIRE(in[k] != nullptr, "SST: missing input array");
  if (sst_fields.empty()) {  // SstModel constructor (:60-99) on a laminar solver's walls
    std::vector<Index> wall(nb, 0);
    for (std::size_t p = 0; p < host.patches.size(); ++p)
      if (patch_condition[p].kind == PatchCondition::Wall)
        for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f)
          wall[f - ni] = 1;
    sst_wall = to_device(dev, wall);
    for (const char* name : {"mu_t", "k_source", "k_sink", "omega_source", "omega_sink", "strain", "vorticity",
                             "f1", "f2", "cdkw", "diffusivity_k", "diffusivity_omega", "production_k",
                             "fixed_omega", "wall_production", "fixed_omega_mask", "selected_wall_face"})
      sst_fields[name] = F(dev, n);
    for (const char* name : {"wall_tau", "wall_u_star", "wall_u_tau", "wall_y_plus", "wall_y_plus_utau",
                             "wall_shear_x", "wall_shear_y", "wall_shear_z", "wall_shear_coefficient",
                             "wall_face_omega", "wall_cell_omega", "face_production"})
      sst_fields[name] = F(dev, nb);
    sst_fields["face_diffusivity_k"] = F(dev, nf);
    sst_fields["face_diffusivity_omega"] = F(dev, nf);
  }
  auto d = [&](const char* name) { return sst_fields.at(name).data(); };
  const float* const* g = in + 3;
  const float *rho = in[12], *mu = in[13], *k = in[14], *om = in[15];
  const float* const* gk = in + 16;
  const float* const* go = in + 19;
  const float* const* wu = in + 22;
  const float k_floor = float(o[17]), omega_floor = float(o[16]);
  dev.launch(KH_SST_WALL, nb,
             {owner, sst_wall, ni, area.x, area.y, area.z, in[0], in[1], in[2], wu[0], wu[1], wu[2],
              wall_normal_distance, rho, mu, k, om, d("wall_tau"), d("wall_shear_x"), d("wall_shear_y"),
              d("wall_shear_z"), d("wall_u_star"), d("wall_u_tau"), d("wall_y_plus"), d("wall_y_plus_utau"),
              d("face_production"), d("wall_shear_coefficient"), k_floor, omega_floor});
  dev.launch(KH_SST_WALL_OMEGA, nb,
             {owner, sst_wall, ni, wall_normal_distance, rho, mu, d("wall_u_star"), d("wall_cell_omega"),
              d("wall_face_omega")});
  dev.launch(KH_SST_WALL_CELLS, n,
             {offsets, faces, sst_wall, wall_normal_distance, d("wall_cell_omega"), d("face_production"), ni,
              d("fixed_omega_mask"), d("selected_wall_face"), d("fixed_omega"), d("wall_production")});
  if (gradients) gradients();
  dev.launch(KH_SST_RATES, n, {g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], d("strain"), d("vorticity")});
  dev.launch(KH_SST_BLENDING_VISCOSITY, n,
             {rho, mu, k, om, wall_distance, gk[0], gk[1], gk[2], go[0], go[1], go[2], d("strain"), d("vorticity"),
              d("f1"), d("f2"), d("cdkw"), d("mu_t"), d("diffusivity_k"), d("diffusivity_omega"), float(o[13]),
              k_floor, omega_floor, int(o[7] != 0), int(o[9] != 0), float(o[18])});
  const int production_mode = o[10] != 0 ? 1 : (o[11] != 0 ? 2 : 0);
  dev.launch(KH_SST_SOURCES, n,
             {rho, mu, k, om, d("strain"), d("vorticity"), d("f1"), d("cdkw"), d("mu_t"), d("fixed_omega_mask"),
              d("wall_production"), d("k_source"), d("k_sink"), d("omega_source"), d("omega_sink"),
              d("production_k"), int(o[7] != 0), int(o[8] != 0), production_mode, int(o[3] != 0), int(o[5] != 0),
              int(o[12] != 0), float(o[4]), int(o[14] != 0), float(o[15]), k_floor, omega_floor, float(o[6])});
  if (o[0] != 0)
    dev.launch(KH_SST_SUSTAIN, n, {rho, d("f1"), d("k_source"), d("omega_source"), float(o[1]), float(o[2])});
  dev.launch(KH_SST_FACE_DIFFUSION, nf,
             {owner, neighbour, weight, sst_wall, mu, d("diffusivity_k"), d("diffusivity_omega"), ni,
              d("face_diffusivity_k"), d("face_diffusivity_omega")});
}
void SteadySolver::sst_update(const double* o, const float* const* in) { impl_->sst_kernels(o, in, {}); }

// SteadySolver::initialize_fields (:1493-1524).
void SteadySolver::Impl::initialize_fields(const std::vector<Vec3>& u, const std::vector<double>& p,
                                           const std::vector<double>& k, const std::vector<double>& w) {
  const auto count = std::size_t(n);
  if (paired)  // reset_paired (:977-988)
    for (auto* v : {&velocity_lo.x, &velocity_lo.y, &velocity_lo.z, &pressure_gradient_lo.x,
                    &pressure_gradient_lo.y, &pressure_gradient_lo.z, &pressure_lo, &mass_lo})
      dev.fill(v->data(), 0.0f, v->size());
  KA_REQUIRE(u.size() == count && p.size() == count, "FV initialization field size mismatch");
  KA_REQUIRE(sst ? (k.size() == count && w.size() == count) : (k.empty() && w.empty()),
             "FV turbulence initialization size mismatch");
  const double ref = reference >= 0 ? p[std::size_t(reference)] : 0;
  std::vector<float> ux(count), uy(count), uz(count), pf(count), kf(sst ? count : 0), wf(sst ? count : 0);
  for (std::size_t c = 0; c < count; ++c) {
    ux[c] = float(u[c].x);
    uy[c] = float(u[c].y);
    uz[c] = float(u[c].z);
    pf[c] = float(p[c] - ref);
    KA_REQUIRE(std::isfinite(pf[c]), "FV nonfinite pressure initialization");
    if (sst) {
      KA_REQUIRE(std::isfinite(k[c]) && k[c] >= 0 && std::isfinite(w[c]) && w[c] > 0,
                 "FV invalid turbulence initialization");
      kf[c] = float(k[c]);
      wf[c] = float(w[c]);
      KA_REQUIRE(std::isfinite(kf[c]) && std::isfinite(wf[c]), "FV turbulence initialization exceeds float storage");
    }
  }
  velocity.x.upload(ux);
  velocity.y.upload(uy);
  velocity.z.upload(uz);
  pressure.upload(pf);
  if (sst) {
    kinetic.upload(kf);
    omega.upload(wf);
  }
  dev.fill(physical_diagonal.data(), 1.0f, count);
  dev.fill(consistent_inverse.data(), 0.0f, count);
  update_boundaries();
  update_pressure();
  update_flux();
  update_turbulence();
  for (int pass = 0; pass < 3; ++pass) {
    monitor();
    refresh_paired_scales();
    update_flux();
    dev.launch(KH_FV_CONTINUITY, n, {offsets, faces, sign, mass, provisional_continuity});
    update_boundaries();
    update_pressure();
  }
}
void SteadySolver::initialize_fields(const std::vector<Vec3>& u, const std::vector<double>& p,
                                     const std::vector<double>& k, const std::vector<double>& w) {
  impl_->initialize_fields(u, p, k, w);
}
IterationReport SteadySolver::monitor() { return impl_->monitor(); }
ForceReport SteadySolver::forces(const std::vector<std::string>& patches, double reference_pressure) {
  return impl_->forces(patches, reference_pressure);
}
const IterationReport& SteadySolver::report() const { return impl_->report; }
Index SteadySolver::reference_cell() const { return impl_->reference; }
const MeshGeometry& SteadySolver::geometry() const { return impl_->geom; }
Index SteadySolver::cells() const { return impl_->n; }
bool SteadySolver::set_batched_momentum(bool on) { return impl_->set_batched(on); }
void SteadySolver::set_graphs(bool on) { impl_->pressure_solver->set_graphs(on); }
// No reference counterpart: the pressure AMG options are fixed there
// (steady_solver.cpp:573-578). The hierarchy is rebuilt from the operator it
// was first built from; the next step refreshes it with the current values.
void SteadySolver::set_pressure_options(const AmgOptions& options) {
  Impl& I = *impl_;
  KA_REQUIRE(options.structure == Structure::SymmetricMMatrix && options.nullspace == NullSpace::None,
             "SIMPLEC: the pressure operator is a pinned symmetric M-matrix");
  auto next = std::make_unique<KrylovSolver>(I.dev, SolverKind::AmgPcg, I.pressure_initial, options);
  next->set_graphs(I.pressure_solver->graphs_enabled());
  I.pressure_solver = std::move(next);
  I.pressure_tuned = false;
}
std::vector<std::pair<Index, Index>> SteadySolver::pressure_levels() const {
  return impl_->pressure_solver->levels();
}
std::string SteadySolver::pressure_plan() const { return impl_->pressure_solver->plan(); }
bool SteadySolver::batched_momentum() const { return impl_->batch_momentum; }
int SteadySolver::host_waits() const {
  int waits = impl_->pressure_solver->host_waits();
  for (const auto& eq : impl_->momentum) waits += eq.solver->host_waits();
  return waits;
}
const float* SteadySolver::field(const std::string& name, std::size_t* size) const {
  auto& s = *impl_;
  const std::map<std::string, const F*> table = {
      {"u", &s.velocity.x}, {"v", &s.velocity.y}, {"w", &s.velocity.z}, {"p", &s.pressure},
      {"mass", &s.mass}, {"u_lo", &s.velocity_lo.x}, {"v_lo", &s.velocity_lo.y}, {"w_lo", &s.velocity_lo.z},
      {"p_lo", &s.pressure_lo}, {"mass_lo", &s.mass_lo}, {"correction", &s.correction},
      {"coefficient", &s.coefficient}, {"face_pressure", &s.face_pressure},
      {"boundary_u", &s.boundary_velocity.x}, {"boundary_pressure", &s.boundary_pressure},
      {"physical_diagonal", &s.physical_diagonal}, {"consistent_inverse", &s.consistent_inverse},
      {"grad_px", &s.grad_p.x}, {"face_u", &s.face_velocity.x}, {"matrix_values", &s.matrix.val},
      {"pressure_rhs", &s.pressure_rhs}, {"volume", &s.volume}, {"weight", &s.weight},
      {"owner_to_face_x", &s.owner_to_face.x}, {"neighbour_to_face_y", &s.neighbour_to_face.y},
      {"owner_to_neighbour_z", &s.owner_to_neighbour.z}, {"area_x", &s.area.x},
      {"wall_normal_distance", &s.wall_normal_distance}, {"wall_distance", &s.wall_distance},
      {"velocity_lsq_inverse", &s.velocity_lsq.inverse},
      {"pressure_lsq_inverse", &s.pressure_lsq.inverse}, {"velocity_lsq_boundary_x", &s.velocity_lsq.bx},
      {"pressure_lsq_boundary_y", &s.pressure_lsq.by}, {"k", &s.kinetic}, {"omega", &s.omega},
      {"viscosity", &s.viscosity}, {"boundary_k", &s.boundary_k}, {"boundary_omega", &s.boundary_omega},
      {"source_x", &s.source.x}, {"strain_dudy", &s.strain_gradient[0].y}, {"k_grad_y", &s.k_eq.grad.y},
      {"omega_rhs", &s.omega_eq.rhs}};
  const auto it = table.find(name);
  const F* array = it != table.end() ? it->second : nullptr;
  if (!array) {
    const auto sst = s.sst_fields.find(name);
    if (sst != s.sst_fields.end()) array = &sst->second;
  }
  if (!array || array->empty()) return nullptr;
  if (size) *size = array->size();
  return array->data();
}

}  // namespace kutta_amd
