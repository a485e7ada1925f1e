// ORACLE — test infrastructure only. Serial restatement of the reference's
// steady SIMPLEC solver (src/kutta/fv/steady_solver.cpp) for the laminar
// closed-domain configuration: wall and symmetry patches, paired (hi+lo) or
// float state, second-order upwind momentum with the Barth-Jespersen limiter,
// least-squares gradients, momentum-weighted Rhie-Chow faces, the automatic
// pseudo time step and the AMG-PCG pressure correction. It also restates the
// host setup the step consumes: upload_mesh's FP32 face and cell arrays
// (fv/device_mesh.cpp:208-256), boundary_least_squares and its extrapolation
// gain (fv/operators.cpp:92-211), orthogonal_quality (mesh/mesh.cpp:172-187)
// and the initial pressure operator (steady_solver.cpp:553-572).
//
// Kernels run through oracle_fv_each (the serial restatements in
// fv_kernels.cpp), reductions are the serial FP64 loops of cpu_device.cpp,
// linear solves are oracle::Solver. Every configuration outside the subset
// above raises an error instead of approximating it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "oracle.hpp"
#include "oracle_mesh.hpp"

extern "C" int oracle_fv_each(const char* name, int n, void* const* args);

namespace oracle::simplec {
namespace {

using V = std::vector<float>;
using VI = std::vector<int>;

struct Vec3f {
  V x, y, z;
  void assign(std::size_t n, float value = 0.0f) {
    x.assign(n, value);
    y.assign(n, value);
    z.assign(n, value);
  }
  V& operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
};

double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 operator*(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(Vec3 a) { return std::sqrt(dot(a, a)); }

// One serial kernel launch in oracle_fv_each's convention: args[k] is the
// buffer pointer itself, or points at a 4-byte scalar held in slot[k].
union Slot {
  float f;
  int i;
};
inline void* put(Slot&, float* p) { return p; }
inline void* put(Slot&, const float* p) { return const_cast<float*>(p); }
inline void* put(Slot&, int* p) { return p; }
inline void* put(Slot&, const int* p) { return const_cast<int*>(p); }
inline void* put(Slot& s, float v) { s.f = v; return &s; }
inline void* put(Slot& s, int v) { s.i = v; return &s; }
template <class... T>
void fv(const char* name, int n, T... a) {
  Slot slot[sizeof...(T)];
  void* args[sizeof...(T)];
  int k = 0;
  ((args[k] = put(slot[k], a), ++k), ...);
  if (n > 0 && oracle_fv_each(name, n, args) != 0)
    throw Error(std::string("oracle SIMPLEC: unknown kernel ") + name);
}

enum class Observation { Omitted, FaceValue, NormalFootValue };

// fv/device_mesh.cpp:208-256.
struct DeviceMesh {
  int n_cells = 0, n_faces = 0, n_internal = 0;
  int n_boundary() const { return n_faces - n_internal; }
  VI owner, neighbour, offsets, faces;
  Vec3f area, owner_to_face, neighbour_to_face, owner_to_neighbour;
  V weight, volume, sign, wall_normal_distance, wall_distance;
};

// fv/operators.hpp:33-50: rows of the fit (interior rows are the stencil's),
// boundary-face rows and the six-plane inverse.
struct LeastSquares {
  Vec3f boundary_delta;
  V inverse;
};

}  // namespace

struct Condition {
  // The boundary_kind codes of fv_flow.kern.h: 0 wall (velocity), 1 velocity
  // inlet (velocity), 2 pressure outlet with a backflow velocity (velocity),
  // 3 symmetry, 4 pressure outlet with normal backflow; outlets hold `pressure`.
  int kind = 0;
  Vec3 velocity{};
  double pressure = 0;
  int total = 0;  // BackflowPressure::Total at an outlet
  std::function<Vec3(Vec3)> profile;  // inlet or backflow velocity of the face centroid
  // kind 5, MassFlowInlet (fv/steady_solver.hpp:60-65; a velocity inlet on
  // the faces): kg/s into the domain and an optional inward direction.
  double mass_flow = 0;
  Vec3 direction{};
  bool has_direction = false;  // the optional direction is given
  // SST inflow turbulence (fv/steady_solver.hpp:28-35) at inlets and
  // outlets: -1 none, 0 k and omega, 1 intensity and length, 2 intensity and
  // viscosity ratio; a and b are the pair's values.
  int turbulence = -1;
  double turbulence_a = 0, turbulence_b = 0;
};

// SstSettings (fv/steady_solver.hpp:102-118) with SstOptions
// (turbulence/sst.hpp:46-78) as model[19]: has_ambient, ambient k, ambient
// omega, production_limiter, production_clip, limit_omega_production,
// omega_production_molecular_ratio, low_re, gamma_set Menter2003, viscosity
// limiter Vorticity, kato_launder, vorticity_production, wall_production,
// cdkw_floor, limit_cross_diffusion, cross_diffusion_clip, omega_floor,
// k_floor, max_viscosity_ratio.
struct SstSettings {
  double model[19] = {0, 0, 0, 1, 10, 1, 0, 0, 0, 0, 0, 0, 1, 1e-10, 1, 1e6, 1e-20, 1e-14, 1e5};
  bool k_second_order = true, omega_second_order = true;
  double k_relaxation = .75, omega_relaxation = .75;
  SolveControl linear{1e-3, 1e-20, 100, 8};
  double steady_tolerance = 1e-6;
  bool compensated_transport = true, pseudo_transient = false;
};

struct Settings {
  double density = 1, viscosity = .01;
  double momentum_relaxation = .7, pressure_relaxation = 1;
  double pseudo_time_scale = 1;
  int monitor_interval = 20;
  bool paired_state = true;
  double paired_state_slack_ulps = 1;
  Vec3 reference_point{};
  SolveControl momentum_linear{1e-2, 1e-20, 100, 8};
  SolveControl pressure_linear{1e-3, 1e-20, 300, 8};
  double pseudo_time_step = 0;  // fixed pseudo time step; 0 is automatic
  bool sst = false;
  SstSettings sst_settings;
};

struct ResidualValue {
  double raw = 0, scaled = 0, fluent_scaled = 0, fluent_normalization = 0;
};
// fv/steady_solver.hpp:286-292
struct ForceReport {
  Vec3 pressure, viscous, area_vector;
  double reference_pressure = 0;
};
// fv/steady_solver.hpp:276-284
struct Report {
  int iteration = 0;
  ResidualValue u, v, w, continuity;
  double corrected_mass_imbalance = 0, maximum_velocity = 0, pseudo_time_step = 0;
  SolveStats momentum[3], pressure;
  ResidualValue k, omega;
  SolveStats turbulence[2];
};

namespace {

// fv/operators.cpp:97-137
void cell_fit(const PolyMesh& m, const Geometry& g, const CellFaces& cf,
              const std::vector<Observation>& obs, Index c, std::vector<Vec3>& delta,
              double inv[3][3]) {
  double a[3][3]{}, q[3][3]{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
    const auto f = cf.faces[j];
    const bool internal = f < m.n_internal_faces();
    if (!internal && obs[f - m.n_internal_faces()] == Observation::Omitted) {
      delta.push_back({});
      continue;
    }
    const auto n = internal ? (m.owner[f] == c ? m.neighbour[f] : m.owner[f]) : -1;
    auto d = (internal ? g.cell_centroid[n] : g.face_centroid[f]) - g.cell_centroid[c];
    if (!internal && obs[f - m.n_internal_faces()] == Observation::NormalFootValue) {
      const auto area = g.face_area[f];
      d = area * (dot(d, area) / dot(area, area));
    }
    const double weight = 1 / dot(d, d), x[]{d.x, d.y, d.z};
    delta.push_back(d * weight);
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) a[r][k] += weight * x[r] * x[k];
  }
  for (int sweep = 0; sweep < 24; ++sweep) {
    int p = 0, r = 1;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j)
        if (std::abs(a[i][j]) > std::abs(a[p][r])) {
          p = i;
          r = j;
        }
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
    const double eigen = a[k][k] > 1e-10 ? 1 / a[k][k] : 1;
    for (int r = 0; r < 3; ++r)
      for (int j = 0; j < 3; ++j) inv[r][j] += q[r][k] * eigen * q[j][k];
  }
}

void to_float(Vec3f& out, const std::vector<Vec3>& v) {
  out.assign(v.size());
  for (std::size_t i = 0; i < v.size(); ++i) {
    out.x[i] = float(v[i].x);
    out.y[i] = float(v[i].y);
    out.z[i] = float(v[i].z);
  }
}

// fv/operators.cpp:29-34
std::vector<Vec3> boundary_rows(const PolyMesh& m, const CellFaces& cf, const std::vector<Vec3>& rows) {
  std::vector<Vec3> result(m.n_faces() - m.n_internal_faces());
  for (Index j = 0; j < Index(cf.faces.size()); ++j)
    if (cf.faces[j] >= m.n_internal_faces()) result[cf.faces[j] - m.n_internal_faces()] = rows[j];
  return result;
}

}  // namespace

// fv/operators.cpp:139-164
std::vector<double> boundary_extrapolation_gain(const PolyMesh& m, const Geometry& g,
                                                const std::vector<Observation>& obs) {
  const auto cf = build_cell_faces(m);
  const auto ni = m.n_internal_faces();
  std::vector<double> gain(m.n_faces() - ni);
  std::vector<Vec3> delta;
  for (Index c = 0; c < m.n_cells; ++c) {
    double inv[3][3];
    delta.clear();
    cell_fit(m, g, cf, obs, c, delta, inv);
    for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
      const auto f = cf.faces[j];
      if (f < ni) continue;
      const auto r = g.face_centroid[f] - g.cell_centroid[c];
      const Vec3 w{r.x * inv[0][0] + r.y * inv[1][0] + r.z * inv[2][0],
                   r.x * inv[0][1] + r.y * inv[1][1] + r.z * inv[2][1],
                   r.x * inv[0][2] + r.y * inv[1][2] + r.z * inv[2][2]};
      for (const auto& d : delta) gain[f - ni] += std::abs(dot(w, d));
    }
  }
  return gain;
}

// fv/operators.cpp:166-211 with a stencil: interior rows are the stencil's
// rows, so only the boundary-face rows and the inverse are stored.
LeastSquares boundary_least_squares(const PolyMesh& m, const Geometry& g,
                                    const std::vector<Observation>& obs) {
  const auto cf = build_cell_faces(m);
  std::vector<Vec3> delta;
  LeastSquares fit;
  fit.inverse.assign(6 * std::size_t(m.n_cells), 0.0f);
  for (Index c = 0; c < m.n_cells; ++c) {
    double inv[3][3];
    cell_fit(m, g, cf, obs, c, delta, inv);
    const double v[]{inv[0][0], inv[0][1], inv[0][2], inv[1][1], inv[1][2], inv[2][2]};
    for (int k = 0; k < 6; ++k) fit.inverse[c + std::size_t(k) * m.n_cells] = float(v[k]);
  }
  to_float(fit.boundary_delta, boundary_rows(m, cf, delta));
  return fit;
}

// mesh/mesh.cpp:172-187
std::vector<double> orthogonal_quality(const PolyMesh& m, const Geometry& g) {
  std::vector<double> quality(m.n_cells, 1);
  const auto cosine = [](Vec3 a, Vec3 b) { return dot(a, b) / (norm(a) * norm(b)); };
  for (Index f = 0; f < m.n_faces(); ++f) {
    const auto o = m.owner[f];
    quality[o] = std::min(quality[o], cosine(g.face_area[f], g.face_centroid[f] - g.cell_centroid[o]));
    if (f >= m.n_internal_faces()) continue;
    const auto n = m.neighbour[f];
    const auto link = g.cell_centroid[n] - g.cell_centroid[o];
    quality[o] = std::min({quality[o], cosine(g.face_area[f], link)});
    quality[n] = std::min({quality[n], cosine(g.face_area[f], g.cell_centroid[n] - g.face_centroid[f]),
                           cosine(g.face_area[f], link)});
  }
  for (auto& q : quality) q = std::max(q, 0.);
  return quality;
}

// fv/device_mesh.cpp:18-64: squared distance from p to a triangle (the
// plane projection when it falls inside, else the nearest edge).
double segment_distance2(Vec3 p, Vec3 a, Vec3 b) {
  const Vec3 ab = b - a;
  const double length2 = dot(ab, ab);
  const double t = length2 == 0 ? 0 : std::clamp(dot(p - a, ab) / length2, 0., 1.);
  const Vec3 r = p - (a + ab * t);
  return dot(r, r);
}
struct WallTriangle {
  Vec3 a, b, c;
  double distance2(Vec3 p) const {
    const Vec3 ab = b - a, ac = c - a, ap = p - a;
    const Vec3 n = cross(ab, ac);
    const double nn = dot(n, n);
    if (nn > 0) {
      const double h = dot(ap, n);
      const Vec3 q = p - n * (h / nn);
      if (dot(cross(ab, q - a), n) >= 0 && dot(cross(c - b, q - b), n) >= 0 && dot(cross(a - c, q - c), n) >= 0)
        return h * h / nn;
    }
    return std::min({segment_distance2(p, a, b), segment_distance2(p, b, c), segment_distance2(p, c, a)});
  }
};

// fv/device_mesh.cpp:208-256 (patch_is_wall marks the wall patches).
DeviceMesh upload_mesh(const PolyMesh& mesh, const Geometry& g, const std::vector<bool>& patch_is_wall) {
  validate(mesh);
  DeviceMesh m;
  m.n_cells = mesh.n_cells;
  m.n_faces = mesh.n_faces();
  m.n_internal = mesh.n_internal_faces();
  m.owner = mesh.owner;
  m.neighbour = mesh.neighbour;
  const auto faces = std::size_t(m.n_faces), internal = std::size_t(m.n_internal);
  to_float(m.area, g.face_area);
  const auto rp = [&](std::size_t f) { return g.face_centroid[f] - g.cell_centroid[mesh.owner[f]]; };
  const auto rn = [&](std::size_t f) { return g.face_centroid[f] - g.cell_centroid[mesh.neighbour[f]]; };
  m.weight.resize(internal);
  for (std::size_t f = 0; f < internal; ++f) {
    const double owner_normal = dot(rp(f), g.face_area[f]);
    const double neighbour_normal = -dot(rn(f), g.face_area[f]);
    m.weight[f] = float(neighbour_normal / (owner_normal + neighbour_normal));
  }
  std::vector<Vec3> a(faces), b(internal), c(faces);
  for (std::size_t f = 0; f < faces; ++f) {
    a[f] = rp(f);
    c[f] = f < internal ? g.cell_centroid[mesh.neighbour[f]] - g.cell_centroid[mesh.owner[f]] : rp(f);
  }
  for (std::size_t f = 0; f < internal; ++f) b[f] = rn(f);
  to_float(m.owner_to_face, a);
  to_float(m.neighbour_to_face, b);
  to_float(m.owner_to_neighbour, c);
  const auto cf = build_cell_faces(mesh);
  m.offsets = cf.offsets;
  m.faces = cf.faces;
  m.sign.assign(cf.signs.begin(), cf.signs.end());
  m.volume.resize(m.n_cells);
  for (Index i = 0; i < m.n_cells; ++i) {
    ORACLE_REQUIRE(std::isfinite(g.cell_volume[i]) && g.cell_volume[i] > 0,
                   "FV requires finite positive cell volumes");
    m.volume[i] = float(g.cell_volume[i]);
  }
  m.wall_normal_distance.assign(m.n_boundary(), 0.0f);
  for (std::size_t p = 0; p < mesh.patches.size(); ++p)
    if (patch_is_wall[p])
      for (Index f = mesh.patches[p].start; f < mesh.patches[p].start + mesh.patches[p].size; ++f)
        m.wall_normal_distance[f - m.n_internal] =
            float(std::abs(dot(rp(std::size_t(f)), g.face_area[f])) / norm(g.face_area[f]));
  // :232-245 with WallBvh (:82-157): the minimum over every wall triangle of
  // the squared distance, which the BVH's pruning does not change.
  std::vector<WallTriangle> triangles;
  for (std::size_t p = 0; p < mesh.patches.size(); ++p)
    if (patch_is_wall[p])
      for (Index f = mesh.patches[p].start; f < mesh.patches[p].start + mesh.patches[p].size; ++f) {
        const Index b = mesh.face_offsets[f], e = mesh.face_offsets[f + 1];
        const Vec3 origin = mesh.points[mesh.face_nodes[b]];
        Vec3 d{};
        for (Index k = b; k < e; ++k) d = d + (mesh.points[mesh.face_nodes[k]] - origin);
        const Vec3 center = origin + d * (1 / double(e - b));
        for (Index k = b; k < e; ++k)
          triangles.push_back({center, mesh.points[mesh.face_nodes[k]],
                               mesh.points[mesh.face_nodes[k + 1 < e ? k + 1 : b]]});
      }
  m.wall_distance.resize(m.n_cells);
  for (Index c = 0; c < m.n_cells; ++c) {
    double best = std::numeric_limits<double>::infinity();
    for (const auto& t : triangles) best = std::min(best, t.distance2(g.cell_centroid[c]));
    m.wall_distance[c] = float(std::sqrt(best));
  }
  return m;
}

namespace {

// fv/transport.hpp:9-25
struct TransportSettings {
  bool second_order = true, limit = true;
  float density = 1, pseudo_time_step = 0, relaxation = 1;
  const V* limiter_low = nullptr;
  float limiter_epsilon2 = -1;
};

// fv/transport.cpp:12-160: one equation sharing the pressure operator's
// values, the face-correction scratch and the zero array; BiCGStab-Jacobi.
struct TransportEquation {
  CsrDevice* matrix = nullptr;
  V rhs, diagonal, consistent_inverse, limiter;
  Vec3f grad;
  V* face_correction = nullptr;
  V* zero = nullptr;
  std::unique_ptr<Solver> solver;
};

// A turbulence transport equation (fv/transport.hpp:50-106) with the
// compensated solve (enable_compensated_solve, transport.cpp:51-55) and the
// steady-stopping work arrays (set_steady_tolerance, :64-69).
struct ScalarEquation {
  V rhs, diagonal, consistent_inverse, limiter, rhs_low, diagonal_low, delta, residual;
  V work_residual, work_absolute, work_scale;
  Vec3f grad;
  std::unique_ptr<Solver> solver;
};

// face_jump_least_squares (fv/operators.cpp:234-253): per cell-face entry the
// face-centroid row r/|r|^2 and the six-plane inverse of sum r r^T/|r|^2.
void face_jump_fit(const PolyMesh& m, const Geometry& g, Vec3f& delta, V& inverse) {
  const auto cf = build_cell_faces(m);
  std::vector<Vec3> rows;
  const auto n = std::size_t(m.n_cells);
  inverse.assign(6 * n, 0.0f);
  for (Index c = 0; c < m.n_cells; ++c) {
    double a[6]{};
    for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
      const Vec3 r = g.face_centroid[cf.faces[j]] - g.cell_centroid[c];
      const double w = 1 / dot(r, r);
      rows.push_back(r * w);
      a[0] += w * r.x * r.x;
      a[1] += w * r.x * r.y;
      a[2] += w * r.x * r.z;
      a[3] += w * r.y * r.y;
      a[4] += w * r.y * r.z;
      a[5] += w * r.z * r.z;
    }
    const double det = a[0] * (a[3] * a[5] - a[4] * a[4]) - a[1] * (a[1] * a[5] - a[4] * a[2]) +
                       a[2] * (a[1] * a[4] - a[3] * a[2]);
    ORACLE_REQUIRE(det > 1e-14, "FV face-jump fit is rank deficient");
    const double inv[6]{(a[3] * a[5] - a[4] * a[4]) / det, (a[2] * a[4] - a[1] * a[5]) / det,
                        (a[1] * a[4] - a[2] * a[3]) / det, (a[0] * a[5] - a[2] * a[2]) / det,
                        (a[1] * a[2] - a[0] * a[4]) / det, (a[0] * a[3] - a[1] * a[1]) / det};
    for (int k = 0; k < 6; ++k) inverse[c + k * n] = float(inv[k]);
  }
  to_float(delta, rows);
}

}  // namespace

class Steady {
 public:
  Steady(PolyMesh input, const std::vector<Condition>& conditions, const Settings& s);
  void step();
  Report monitor();
  void iterate(int count) {
    ORACLE_REQUIRE(count >= 0, "FV negative iteration count");
    while (count-- > 0) step();
  }
  // SteadySolver::initialize (:1480-1491) from per-cell values in the
  // solver's cell order.
  void initialize(const std::vector<Vec3>& u, const std::vector<double>& p);
  // SteadySolver::initialize_fields (:1493-1524) in the solver's cell order;
  // k and omega are required with SST and ignored without.
  void initialize_fields(const std::vector<Vec3>& u, const std::vector<double>& p, const std::vector<double>& k,
                         const std::vector<double>& omega);
  const Report& report() const { return report_; }
  ForceReport forces(const std::vector<std::string>& patches, double reference_pressure);
  // Named state arrays, for parity checks.
  const V* array(const std::string& name) const;
  // SstModel::update (turbulence/sst.cpp:101-158, no intermittency) on this
  // solver's mesh: options[19] as ka_simplec_sst_update, in[25] the input
  // arrays. The fields are kept by name (integer fields bit-copied into
  // float storage) for array().
  void sst_update(const double* options, const float* const* in);
  Index reference() const { return reference_; }
  int cells() const { return mesh.n_cells; }
  // Rebuilds the pressure AMG with other options from the operator it was
  // first built from (no reference counterpart; e.g. the K-cycle).
  void set_pressure_options(const AmgOptions& options) {
    ORACLE_REQUIRE(options.structure == Structure::SymmetricMMatrix && options.nullspace == NullSpace::None,
                   "SIMPLEC: the pressure operator is a pinned symmetric M-matrix");
    pressure_solver = std::make_unique<Solver>(SolverKind::AmgPcg, pressure_initial, options);
  }
  const Geometry& geometry() const { return geom; }

 private:
  void update_time(double maximum_speed);
  void refresh_time_step();
  void update_boundaries();
  void apply_backflow_pressure();
  void gradient(const LeastSquares& fit, const V& cell, const V& bc, Vec3f& out);
  void reconstruct_velocity();
  void update_pressure();
  void update_flux();
  void assemble(int k, const TransportSettings& t);
  void wall_terms(int k);
  void paired_momentum_residual(int k, V& r);
  SolveStats solve_transport(int k, V& field, const SolveControl& control);
  void update_momentum_diagonal();
  void update_pressure_gradient_lo();
  void refresh_paired_scales();
  void finish_step();
  TransportSettings transport() const;
  // SST k-omega in the step (steady_solver.cpp:638-818).
  void sst_kernels(const double* options, const float* const* in, const std::function<void()>& gradients);
  void reconstruct_strain();
  void turbulence_boundaries();
  void update_turbulence(bool velocity_changed = true);
  void add_transpose_stress();
  void assemble_turbulence(bool is_omega);
  void physical_residual(ScalarEquation& eq, const V& field, V& result, const VI* excluded);
  SolveStats solve_scalar(ScalarEquation& eq, V& field, const VI* excluded);
  void solve_turbulence();
  void monitor_turbulence();
  struct Slack {
    float velocity, pressure, flux_per_area;
  };
  Slack paired_slack() const;

  Settings settings;
  std::vector<Condition> conditions_;
  PolyMesh host;
  Geometry geom;
  DeviceMesh mesh;
  Stencil stencil;
  LeastSquares velocity_lsq, pressure_lsq;
  CsrDevice matrix;
  std::unique_ptr<Solver> pressure_solver;
  CsrHost pressure_initial;
  TransportEquation momentum[3];
  V face_correction_scratch, zero;
  Index reference_ = -1;
  std::map<std::string, V> sst_fields_;
  bool paired = false;
  // SST state (steady_solver.cpp:111-127, 484-542).
  bool sst = false;
  LeastSquares strain_lsq, k_lsq, omega_lsq;
  Vec3f jump_delta, strain_base[3], strain_gradient[3], transpose_stress;
  V jump_inverse, jump_forward, jump_reverse;
  V kinetic, omega, rho_cell, mu_cell, boundary_k, boundary_omega, turbulence_a, turbulence_b;
  VI turbulence_spec, turbulence_kind, zero_mask, sst_wall, fixed_omega_mask, selected_wall_face;
  ScalarEquation k_eq, omega_eq;

  Vec3f velocity, face_velocity, body, source, boundary_velocity, prescribed_velocity;
  Vec3f limited_gradient[3], grad_p, grad_correction, absolute_residual, fluent_scale;
  V pressure, mass, viscosity, boundary_pressure, prescribed_pressure;
  VI boundary_kind, scalar_kind, total_pressure;
  V pressure_rhs, correction, correction_boundary, coefficient, face_pressure, pressure_jump;
  V residual_u, residual_v, residual_w, continuity, provisional_continuity, scale, speed;
  V wall_diagonal, wall_rhs, physical_diagonal, consistent_inverse;
  Vec3f velocity_lo, pressure_gradient_lo;
  V pressure_lo, mass_lo, previous_mass, pressure_boundary_lo, imbalance, momentum_delta;

  float limiter_epsilon2[3]{-1, -1, -1};
  double paired_speed = 0, paired_pressure_magnitude = 0;
  double length = 0, bc_speed = 0, pressure_speed = 0, continuity_scale = 0;
  double first_five_continuity[5]{};
  float time_step = 0;
  double velocity_scale = 1;
  Report report_;
};

// fv/steady_solver.cpp:203-597, the laminar closed-domain subset.
Steady::Steady(PolyMesh input, const std::vector<Condition>& conditions, const Settings& s)
    : settings(s), conditions_(conditions), host(std::move(input)) {
  ORACLE_REQUIRE(s.density > 0 && std::isfinite(s.density) && s.viscosity > 0 &&
                     std::isfinite(s.viscosity) && s.momentum_relaxation > 0 &&
                     s.momentum_relaxation <= 1 && s.pressure_relaxation > 0 &&
                     s.pressure_relaxation <= 1 && s.pseudo_time_scale > 0 &&
                     s.monitor_interval > 0 && std::isfinite(s.paired_state_slack_ulps) &&
                     s.paired_state_slack_ulps > 0,
                 "FV invalid steady solver settings");
  ORACLE_REQUIRE(std::isfinite(s.pseudo_time_step) && s.pseudo_time_step >= 0, "FV invalid steady solver settings");
  ORACLE_REQUIRE(conditions.size() == host.patches.size(), "FV supply exactly one condition per patch");
  std::vector<bool> patch_is_wall;
  bool outlet = false;
  for (const auto& c : conditions) {
    ORACLE_REQUIRE(c.kind >= 0 && c.kind <= 5, "SIMPLEC oracle: walls, velocity and mass-flow inlets, pressure outlets, symmetry");
    ORACLE_REQUIRE(std::isfinite(c.velocity.x) && std::isfinite(c.velocity.y) && std::isfinite(c.velocity.z),
                   "FV nonfinite boundary velocity");
    ORACLE_REQUIRE(std::isfinite(c.pressure), "FV nonfinite outlet pressure");
    patch_is_wall.push_back(c.kind == 0);
    outlet = outlet || c.kind == 2 || c.kind == 4;
  }
  renumber_cells_rcm(host);
  geom = compute_geometry(host);
  ORACLE_REQUIRE(host.n_cells > 0, "FV steady solver needs a nonempty mesh");
  // :229-247: a pressure boundary sets the pressure level; a closed domain
  // pins the cell holding the reference point.
  reference_ = outlet ? -1 : locate_cell(host, geom, s.reference_point);
  {
    const auto nb = build_cell_neighbours(host);
    std::vector<bool> seen(host.n_cells);
    std::vector<Index> stack{0};
    seen[0] = true;
    for (std::size_t k = 0; k < stack.size(); ++k)
      for (Index j = nb.offsets[stack[k]]; j < nb.offsets[stack[k] + 1]; ++j)
        if (!seen[nb.cells[j]]) {
          seen[nb.cells[j]] = true;
          stack.push_back(nb.cells[j]);
        }
    ORACLE_REQUIRE(stack.size() == std::size_t(host.n_cells), "FV disconnected domains require separate solvers");
  }
  mesh = upload_mesh(host, geom, patch_is_wall);
  stencil = make_stencil(host, geom);
  const auto n = std::size_t(mesh.n_cells), nf = std::size_t(mesh.n_faces), nb = std::size_t(mesh.n_boundary());
  velocity.assign(n);
  face_velocity.assign(nf);
  body.assign(n);
  source.assign(n);
  boundary_velocity.assign(nb);
  prescribed_velocity.assign(nb);
  for (auto& g : limited_gradient) g.assign(n);
  grad_p.assign(n);
  grad_correction.assign(n);
  absolute_residual.assign(n);
  fluent_scale.assign(n);
  pressure.assign(n, 0.0f);
  mass.assign(nf, 0.0f);
  viscosity.assign(nf, float(s.viscosity));
  boundary_pressure.assign(nb, 0.0f);
  prescribed_pressure.assign(nb, 0.0f);
  boundary_kind.assign(nb, 0);
  scalar_kind.assign(nb, 0);
  pressure_rhs.assign(n, 0.0f);
  correction.assign(n, 0.0f);
  correction_boundary.assign(nb, 0.0f);
  coefficient.assign(nf, 0.0f);
  face_pressure.assign(nf, 0.0f);
  pressure_jump.assign(nf, 0.0f);
  for (V* v : {&residual_u, &residual_v, &residual_w, &continuity, &provisional_continuity, &scale, &speed,
               &wall_diagonal, &wall_rhs})
    v->assign(n, 0.0f);
  paired = s.paired_state;
  if (paired) {
    velocity_lo.assign(n);
    pressure_gradient_lo.assign(n);
    pressure_lo.assign(n, 0.0f);
    mass_lo.assign(nf, 0.0f);
    previous_mass.assign(nf, 0.0f);
    pressure_boundary_lo.assign(nb, 0.0f);
    imbalance.assign(n, 0.0f);
    momentum_delta.assign(n, 0.0f);
  }
  total_pressure.assign(nb, 0);
  physical_diagonal.assign(n, 1.0f);
  consistent_inverse.assign(n, 0.0f);

  // Boundary kinds and prescribed values (:314-371).
  std::vector<Index> kinds(nb);
  std::vector<Vec3> velocities(nb);
  // :305-311: a mass-flow inlet's total and projected areas.
  std::vector<double> inlet_area(host.patches.size(), 0.0), inlet_projected(host.patches.size(), 0.0);
  for (std::size_t p = 0; p < host.patches.size(); ++p)
    if (conditions[p].kind == 5)
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
        inlet_area[p] += norm(geom.face_area[f]);
        const Vec3 dir = conditions[p].direction;
        if (conditions[p].has_direction && norm(dir) > 0)
          inlet_projected[p] -= dot(dir * (1 / norm(dir)), geom.face_area[f]);
      }
  double pmin = std::numeric_limits<double>::infinity(), pmax = -pmin;
  for (std::size_t p = 0; p < host.patches.size(); ++p)
    for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
      const auto b = f - mesh.n_internal;
      const auto& c = conditions[p];
      kinds[b] = c.kind == 5 ? 1 : c.kind;
      if (c.kind == 5) {  // fv/boundary_physics.cpp:14-26
        const Vec3 a = geom.face_area[f];
        ORACLE_REQUIRE(std::isfinite(c.mass_flow) && c.mass_flow >= 0 && inlet_area[p] > 0 && settings.density > 0,
                       "mass-flow-inlet: nonnegative mass flow and positive area/density required");
        if (c.has_direction) {
          ORACLE_REQUIRE(std::isfinite(c.direction.x) && std::isfinite(c.direction.y) &&
                             std::isfinite(c.direction.z) && norm(c.direction) > 0 && inlet_projected[p] > 0,
                         "mass-flow-inlet: nonzero inward direction required");
          const Vec3 d = c.direction * (1 / norm(c.direction));
          ORACLE_REQUIRE(dot(d, a) < 0, "mass-flow-inlet: direction must enter every face");
          velocities[b] = d * (c.mass_flow / (settings.density * inlet_projected[p]));
        } else {
          velocities[b] = a * (-c.mass_flow / (settings.density * inlet_area[p] * norm(a)));
        }
      }
      if (c.kind == 0 || c.kind == 1 || c.kind == 2)
        velocities[b] = c.profile && c.kind != 0 ? c.profile(geom.face_centroid[f]) : c.velocity;
      ORACLE_REQUIRE(std::isfinite(velocities[b].x) && std::isfinite(velocities[b].y) &&
                         std::isfinite(velocities[b].z),
                     "FV nonfinite field value");
      if (c.kind == 2 || c.kind == 4) {
        prescribed_pressure[b] = float(c.pressure);
        total_pressure[b] = c.total;
        pmin = std::min(pmin, c.pressure);
        pmax = std::max(pmax, c.pressure);
      }
      bc_speed = std::max(bc_speed, norm(velocities[b]));
    }
  pressure_speed = outlet ? .1 * std::sqrt((pmax - pmin) / settings.density) : 0;
  std::vector<Observation> independent_velocity(nb), independent_pressure(nb);
  for (std::size_t b = 0; b < nb; ++b) {
    independent_velocity[b] = kinds[b] == 3 ? Observation::NormalFootValue
                              : (kinds[b] == 0 || kinds[b] == 1) ? Observation::FaceValue
                                                                 : Observation::Omitted;
    independent_pressure[b] = (kinds[b] == 2 || kinds[b] == 4) ? Observation::FaceValue : Observation::Omitted;
  }
  boundary_kind.assign(kinds.begin(), kinds.end());
  prescribed_velocity.assign(nb);
  for (std::size_t b = 0; b < nb; ++b) {
    prescribed_velocity.x[b] = float(velocities[b].x);
    prescribed_velocity.y[b] = float(velocities[b].y);
    prescribed_velocity.z[b] = float(velocities[b].z);
  }
  matrix = upload(stencil.pattern);
  matrix.val.assign(matrix.col.size(), 0.0f);

  // The projected face reconstruction corrects only faces whose cell-to-face
  // vectors leave the face normal (face_projection.cpp:136-150); this
  // restatement covers meshes where it corrects none.
  {
    const auto internal = mesh.n_internal;
    for (Index face = 0; face < mesh.n_faces; ++face) {
      if (face >= internal && independent_velocity[face - internal] != Observation::FaceValue) continue;
      const auto owner = host.owner[face];
      const auto area = geom.face_area[face], unit = area * (1 / norm(area));
      const auto rp = geom.face_centroid[face] - geom.cell_centroid[owner];
      const auto rn = face < internal ? geom.cell_centroid[host.neighbour[face]] - geom.face_centroid[face] : Vec3{};
      const double dp = dot(rp, unit), dn = dot(rn, unit);
      const double roundoff = 256 * std::numeric_limits<double>::epsilon() *
                              std::max({1., norm(geom.face_centroid[face]), norm(geom.cell_centroid[owner]),
                                        face < internal ? norm(geom.cell_centroid[host.neighbour[face]]) : 0.});
      ORACLE_REQUIRE(norm(rp - unit * dp) <= roundoff && norm(rn - unit * dn) <= roundoff,
                     "SIMPLEC oracle: meshes needing the projected face reconstruction are not covered");
    }
  }
  face_correction_scratch.assign(nf, 0.0f);
  zero.assign(n, 0.0f);
  CsrHost identity = stencil.pattern;
  identity.val.assign(identity.col.size(), 0.0);
  for (auto j : identity.diag) identity.val[j] = 1;
  for (auto& eq : momentum) {
    eq.matrix = &matrix;
    eq.rhs.assign(n, 0.0f);
    eq.diagonal.assign(n, 0.0f);
    eq.consistent_inverse.assign(n, 0.0f);
    eq.limiter.assign(n, 0.0f);
    eq.grad.assign(n);
    eq.face_correction = &face_correction_scratch;
    eq.zero = &zero;
    eq.solver = std::make_unique<Solver>(SolverKind::BicgstabJacobi, identity, AmgOptions{});
  }
  velocity_lsq = boundary_least_squares(host, geom, independent_velocity);
  {
    const auto gain = boundary_extrapolation_gain(host, geom, independent_pressure);
    for (std::size_t b = 0; b < nb; ++b)
      if (independent_pressure[b] == Observation::Omitted && gain[b] > 1)
        independent_pressure[b] = Observation::FaceValue;
  }
  pressure_lsq = boundary_least_squares(host, geom, independent_pressure);
  {
    const auto quality = orthogonal_quality(host, geom);
    for (Index c = 0; c < mesh.n_cells; ++c)
      ORACLE_REQUIRE(quality[c] > .05, "SIMPLEC oracle: poor-quality cells are not covered");
  }
  sst = s.sst;
  if (sst) {  // :471-542
    const auto& t = s.sst_settings;
    const double* o = t.model;
    ORACLE_REQUIRE(t.k_relaxation > 0 && t.k_relaxation <= 1 && t.omega_relaxation > 0 && t.omega_relaxation <= 1,
                   "FV invalid SST relaxation");
    ORACLE_REQUIRE(std::isfinite(o[4]) && o[4] > 0 && o[13] > 0 && o[15] > 0 && o[16] > 0 && o[17] > 0 &&
                       o[18] > 0 && std::isfinite(o[6]) && o[6] >= 0 && !(o[10] != 0 && o[11] != 0),
                   "SST: invalid options");
    ORACLE_REQUIRE(std::isfinite(t.steady_tolerance) && t.steady_tolerance >= 0, "FV invalid scalar steady tolerance");
    std::vector<Observation> strain_rows(nb), independent_k(nb), independent_omega(nb);
    for (std::size_t b = 0; b < nb; ++b) {
      strain_rows[b] = kinds[b] != 0 ? Observation::FaceValue : Observation::Omitted;
      independent_k[b] = kinds[b] == 1 ? Observation::FaceValue : Observation::Omitted;
      independent_omega[b] = (kinds[b] == 0 || kinds[b] == 1) ? Observation::FaceValue : Observation::Omitted;
      if (kinds[b] == 3) independent_k[b] = independent_omega[b] = Observation::NormalFootValue;
    }
    strain_lsq = boundary_least_squares(host, geom, strain_rows);
    face_jump_fit(host, geom, jump_delta, jump_inverse);
    jump_forward.assign(nf, 0.0f);
    jump_reverse.assign(mesh.n_internal, 0.0f);
    for (int k = 0; k < 3; ++k) {
      strain_base[k].assign(n);
      strain_gradient[k].assign(n);
    }
    k_lsq = boundary_least_squares(host, geom, independent_k);
    omega_lsq = boundary_least_squares(host, geom, independent_omega);
    kinetic.assign(n, float(o[17]));
    omega.assign(n, float(o[16]));
    rho_cell.assign(n, float(s.density));
    mu_cell.assign(n, float(s.viscosity));
    boundary_k.assign(nb, 0.0f);
    boundary_omega.assign(nb, 0.0f);
    turbulence_kind.assign(nb, 0);
    zero_mask.assign(n, 0);
    transpose_stress.assign(nf);
    turbulence_spec.assign(nb, 0);
    turbulence_a.assign(nb, 0.0f);
    turbulence_b.assign(nb, 0.0f);
    sst_wall.assign(nb, 0);
    for (std::size_t p = 0; p < host.patches.size(); ++p) {
      const auto& c = conditions[p];
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f)
        sst_wall[f - mesh.n_internal] = c.kind == 0 ? 1 : 0;
      if (c.kind == 0 || c.kind == 3) continue;
      ORACLE_REQUIRE(c.turbulence >= 0 && c.turbulence <= 2,
                     "FV missing turbulence specification: " + host.patches[p].name);
      for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
        const auto b = f - mesh.n_internal;
        turbulence_spec[b] = c.turbulence;
        turbulence_a[b] = float(c.turbulence_a);
        turbulence_b[b] = float(c.turbulence_b);
        ORACLE_REQUIRE(turbulence_a[b] >= 0 && turbulence_b[b] > 0 && std::isfinite(turbulence_a[b]) &&
                           std::isfinite(turbulence_b[b]),
                       "FV invalid turbulence specification: " + host.patches[p].name);
      }
    }
    CsrHost unit = stencil.pattern;
    unit.val.assign(unit.col.size(), 0.0);
    for (auto j : unit.diag) unit.val[j] = 1;
    for (auto* eq : {&k_eq, &omega_eq}) {
      for (V* v : {&eq->rhs, &eq->diagonal, &eq->consistent_inverse, &eq->limiter, &eq->rhs_low, &eq->diagonal_low,
                   &eq->delta, &eq->residual, &eq->work_residual, &eq->work_absolute, &eq->work_scale})
        v->assign(n, 0.0f);
      eq->grad.assign(n);
      eq->solver = std::make_unique<Solver>(SolverKind::BicgstabJacobi, unit, AmgOptions{});
    }
    for (const char* name : {"mu_t", "k_source", "k_sink", "omega_source", "omega_sink", "strain", "vorticity",
                             "f1", "f2", "cdkw", "diffusivity_k", "diffusivity_omega", "production_k",
                             "fixed_omega", "wall_production"})
      sst_fields_[name].assign(n, 0.0f);
    for (const char* name : {"wall_tau", "wall_u_star", "wall_u_tau", "wall_y_plus", "wall_y_plus_utau",
                             "wall_shear_x", "wall_shear_y", "wall_shear_z", "wall_shear_coefficient",
                             "wall_face_omega", "wall_cell_omega", "face_production"})
      sst_fields_[name].assign(nb, 0.0f);
    sst_fields_["face_diffusivity_k"].assign(nf, 0.0f);
    sst_fields_["face_diffusivity_omega"].assign(nf, 0.0f);
    fixed_omega_mask.assign(n, 0);
    selected_wall_face.assign(n, 0);
  }
  // The initial pressure operator (:553-572).
  CsrHost initial = stencil.pattern;
  initial.val.assign(initial.col.size(), 0.0);
  {
    const auto cf = build_cell_faces(host);
    for (Index c = 0; c < mesh.n_cells; ++c) {
      UpperSum ap{0, 0};
      for (Index j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
        const auto f = cf.faces[j];
        const auto other = f < mesh.n_internal ? (host.owner[f] == c ? host.neighbour[f] : host.owner[f]) : -1;
        if (other < 0 && kinds[f - mesh.n_internal] != 2 && kinds[f - mesh.n_internal] != 4) continue;
        const auto d = (other >= 0 ? geom.cell_centroid[other] : geom.face_centroid[f]) - geom.cell_centroid[c];
        const float a = float(dot(geom.face_area[f], geom.face_area[f]) / std::abs(dot(geom.face_area[f], d)));
        if (other < 0 || other == reference_) ap = kupper_add(ap, a);
        if (other >= 0 && other != reference_ && c != reference_) {
          const auto begin = initial.col.begin() + initial.row_ptr[c], end = initial.col.begin() + initial.row_ptr[c + 1];
          const auto index = std::lower_bound(begin, end, other) - initial.col.begin();
          initial.val[index] = float(float(initial.val[index]) - a);
        }
      }
      for (Index j = initial.row_ptr[c]; j < initial.row_ptr[c + 1]; ++j)
        if (j != initial.diag[c]) ap = kupper_add(ap, -float(initial.val[j]));
      initial.val[initial.diag[c]] = c == reference_ ? 1 : kupper_value(ap);
    }
  }
  AmgOptions amg;
  amg.structure = Structure::SymmetricMMatrix;
  if (sst) amg.pre_sweeps = amg.post_sweeps = 2;  // :573-578
  pressure_solver = std::make_unique<Solver>(SolverKind::AmgPcg, initial, amg);
  pressure_initial = std::move(initial);
  const double volume = std::accumulate(geom.cell_volume.begin(), geom.cell_volume.end(), 0.);
  Vec3 lo = host.points.front(), hi = lo;
  for (auto p : host.points) {
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  length = std::min(std::cbrt(volume), std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z}));
  update_time(bc_speed);
  std::fill(momentum[0].diagonal.begin(), momentum[0].diagonal.end(), 1.0f);
  std::fill(momentum[0].consistent_inverse.begin(), momentum[0].consistent_inverse.end(), 0.0f);
  update_boundaries();
  update_pressure();
  update_flux();
}

void Steady::initialize(const std::vector<Vec3>& u, const std::vector<double>& p) {
  ORACLE_REQUIRE(!sst, "FV SST initialization requires initialize_fields with k and omega");
  const auto n = std::size_t(mesh.n_cells);
  ORACLE_REQUIRE(u.size() == n && p.size() == n, "FV initialization needs one value per cell");
  if (paired) {  // reset_paired (:977-988)
    for (auto* v : {&velocity_lo.x, &velocity_lo.y, &velocity_lo.z, &pressure_gradient_lo.x,
                    &pressure_gradient_lo.y, &pressure_gradient_lo.z, &pressure_lo, &mass_lo})
      std::fill(v->begin(), v->end(), 0.0f);
  }
  const double ref = reference_ >= 0 ? p[std::size_t(reference_)] : 0;
  for (std::size_t c = 0; c < n; ++c) {
    const double value = p[c] - ref;
    ORACLE_REQUIRE(std::isfinite(value), "FV nonfinite pressure initialization");
    velocity.x[c] = float(u[c].x);
    velocity.y[c] = float(u[c].y);
    velocity.z[c] = float(u[c].z);
    pressure[c] = float(value);
  }
  update_boundaries();
  update_pressure();
  std::fill(physical_diagonal.begin(), physical_diagonal.end(), 1.0f);
  std::fill(consistent_inverse.begin(), consistent_inverse.end(), 0.0f);
  update_flux();
}

// :843-851 (segregated: the maximum cell speed is the velocity scale).
void Steady::update_time(double maximum_speed) {
  velocity_scale = std::max({.1 * bc_speed, maximum_speed, pressure_speed, 1e-10});
  time_step = settings.pseudo_time_step > 0 ? float(settings.pseudo_time_step)
                                            : float(.3 * length * settings.pseudo_time_scale / velocity_scale);
}

// :857-873, no ramp, Barth-Jespersen limiter: a fixed step is not refreshed.
void Steady::refresh_time_step() {
  if (settings.pseudo_time_step > 0) return;
  fv("fv_speed", mesh.n_cells, velocity.x.data(), velocity.y.data(), velocity.z.data(), speed.data());
  const auto maximum = max_abs(speed.data(), speed.size());
  update_time(maximum);
}

// :874-885
void Steady::update_boundaries() {
  fv("fv_velocity_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(), mesh.area.x.data(),
     mesh.area.y.data(), mesh.area.z.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(),
     prescribed_velocity.x.data(), prescribed_velocity.y.data(), prescribed_velocity.z.data(), mass.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(), scalar_kind.data(),
     mesh.n_internal, float(settings.density));
}

// :599-606
void Steady::apply_backflow_pressure() {
  fv("fv_backflow_pressure", mesh.n_boundary(), mesh.owner.data(), total_pressure.data(), mass.data(),
     velocity.x.data(), velocity.y.data(), velocity.z.data(), boundary_pressure.data(), mesh.n_internal,
     float(settings.density));
}

// fv/operators.cpp:296-302
void Steady::gradient(const LeastSquares& fit, const V& cell, const V& bc, Vec3f& out) {
  fv("fv_gradient", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), stencil.lsq_x.data(),
     stencil.lsq_y.data(), stencil.lsq_z.data(), fit.boundary_delta.x.data(), fit.boundary_delta.y.data(),
     fit.boundary_delta.z.data(), fit.inverse.data(), cell.data(), bc.data(), out.x.data(), out.y.data(),
     out.z.data(), mesh.n_cells);
}

// :886-889 with fv/operators.cpp:311-318
void Steady::reconstruct_velocity() {
  auto& g0 = momentum[0].grad;
  auto& g1 = momentum[1].grad;
  auto& g2 = momentum[2].grad;
  const auto& fit = velocity_lsq;
  fv("fv_vector_gradient", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), stencil.lsq_x.data(),
     stencil.lsq_y.data(), stencil.lsq_z.data(), fit.boundary_delta.x.data(), fit.boundary_delta.y.data(),
     fit.boundary_delta.z.data(), fit.inverse.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(), g0.x.data(),
     g0.y.data(), g0.z.data(), g1.x.data(), g1.y.data(), g1.z.data(), g2.x.data(), g2.y.data(), g2.z.data(),
     mesh.n_cells);
}

// :896-925, second-order pressure interpolation with no poor cells.
void Steady::update_pressure() {
  const auto boundary = [&](const V& p, Vec3f& g, const V& prescribed, V& out) {
    fv("fv_pressure_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(),
       mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), p.data(),
       g.x.data(), g.y.data(), g.z.data(), prescribed.data(), out.data(), mesh.n_internal);
  };
  boundary(pressure, grad_p, prescribed_pressure, boundary_pressure);
  fv("fv_cell_pressure_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(), pressure.data(),
     boundary_pressure.data(), mesh.n_internal);
  apply_backflow_pressure();
  gradient(pressure_lsq, pressure, boundary_pressure, grad_p);
  boundary(pressure, grad_p, prescribed_pressure, boundary_pressure);
  apply_backflow_pressure();
  fv("fv_pressure_faces", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), mesh.owner_to_face.x.data(),
     mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), mesh.neighbour_to_face.x.data(),
     mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(), pressure.data(), grad_p.x.data(),
     grad_p.y.data(), grad_p.z.data(), boundary_pressure.data(), face_pressure.data(), pressure_jump.data(),
     mesh.n_internal);
}

// :943-961
Steady::Slack Steady::paired_slack() const {
  const double ulps = settings.paired_state_slack_ulps * double(std::numeric_limits<float>::epsilon());
  const double speed_scale = std::max(bc_speed, paired_speed);
  const double pressure_scale = std::max(settings.density * speed_scale * speed_scale, paired_pressure_magnitude);
  return {float(ulps * speed_scale), float(ulps * pressure_scale), float(ulps * settings.density * speed_scale)};
}

// :1139-1248, limited, paired or float, momentum-weighted faces.
void Steady::update_flux() {
  if (paired) previous_mass = mass;
  reconstruct_velocity();
  for (int k = 0; k < 3; ++k) {
    auto& eq = momentum[k];
    auto& g = limited_gradient[k];
    if (paired)
      fv("fv_paired_limiter_scale", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), mesh.faces.data(),
         mesh.sign.data(), mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
         mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
         velocity[k].data(), velocity_lo[k].data(), boundary_velocity[k].data(), scalar_kind.data(),
         eq.grad.x.data(), eq.grad.y.data(), eq.grad.z.data(), eq.limiter.data(), g.x.data(), g.y.data(),
         g.z.data(), limiter_epsilon2[k]);
    else
      fv("fv_limiter_scale", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), mesh.faces.data(),
         mesh.sign.data(), mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
         mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
         velocity[k].data(), boundary_velocity[k].data(), eq.grad.x.data(), eq.grad.y.data(), eq.grad.z.data(),
         eq.limiter.data(), g.x.data(), g.y.data(), g.z.data(), limiter_epsilon2[k]);
  }
  const auto& gu = limited_gradient[0];
  const auto& gv = limited_gradient[1];
  const auto& gw = limited_gradient[2];
  fv("fv_momentum_faces", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), boundary_kind.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
     mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
     velocity.x.data(), velocity.y.data(), velocity.z.data(), gu.x.data(), gu.y.data(), gu.z.data(),
     gv.x.data(), gv.y.data(), gv.z.data(), gw.x.data(), gw.y.data(), gw.z.data(), physical_diagonal.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(),
     face_velocity.x.data(), face_velocity.y.data(), face_velocity.z.data(), mesh.n_internal);
  fv("fv_rhie_chow", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), boundary_kind.data(),
     mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), stencil.diffusion.data(), mesh.volume.data(),
     physical_diagonal.data(), consistent_inverse.data(), face_velocity.x.data(), face_velocity.y.data(),
     face_velocity.z.data(), pressure_jump.data(), mass.data(), coefficient.data(), mesh.n_internal,
     float(settings.density));
  if (paired)
    fv("fv_paired_flux", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), boundary_kind.data(),
       mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), mesh.owner_to_face.x.data(),
       mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), mesh.neighbour_to_face.x.data(),
       mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(), physical_diagonal.data(),
       physical_diagonal.data(), mesh.volume.data(), stencil.diffusion.data(), velocity_lo.x.data(),
       velocity_lo.y.data(), velocity_lo.z.data(), pressure_lo.data(), pressure_gradient_lo.x.data(),
       pressure_gradient_lo.y.data(), pressure_gradient_lo.z.data(), previous_mass.data(), mass.data(),
       mass_lo.data(), mesh.n_internal, float(settings.density), paired_slack().flux_per_area);
}

// :1250-1256 (SIMPLEC: relaxation is the momentum relaxation).
TransportSettings Steady::transport() const {
  TransportSettings t;
  t.density = float(settings.density);
  t.pseudo_time_step = time_step;
  t.relaxation = float(settings.momentum_relaxation);
  return t;
}

// :765-776 (laminar: the coefficient argument is an unused placeholder).
// SST: the wall-function shear coefficient acts at walls.
void Steady::wall_terms(int k) {
  const V& coefficient_of_wall = sst ? sst_fields_.at("wall_shear_coefficient") : prescribed_pressure;
  fv("fv_boundary_momentum", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), boundary_kind.data(),
     mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), coefficient_of_wall.data(), viscosity.data(),
     stencil.diffusion.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(), wall_diagonal.data(),
     wall_rhs.data(), mesh.n_internal, k, sst ? 1 : 0);
}

// fv/transport.cpp:83-121 with limit_gradient (fv/operators.cpp:345-362) and
// with_low (steady_solver.cpp:963-976): component k of momentum.
void Steady::assemble(int k, const TransportSettings& t) {
  auto& eq = momentum[k];
  const V& field = velocity[k];
  const V& bc = boundary_velocity[k];
  const V& src = source[k];
  if (t.limiter_low)
    fv("fv_paired_limiter", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), mesh.faces.data(),
       mesh.sign.data(), mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
       mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
       field.data(), t.limiter_low->data(), bc.data(), scalar_kind.data(), eq.grad.x.data(), eq.grad.y.data(),
       eq.grad.z.data(), eq.limiter.data(), t.limiter_epsilon2);
  else
    fv("fv_limiter", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), mesh.faces.data(),
       mesh.sign.data(), mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
       mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
       field.data(), bc.data(), eq.grad.x.data(), eq.grad.y.data(), eq.grad.z.data(), eq.limiter.data(),
       t.limiter_epsilon2);
  auto& correction_faces = *eq.face_correction;
  fv("fv_transport_faces", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), mesh.weight.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
     mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
     stencil.nonorth_x.data(), stencil.nonorth_y.data(), stencil.nonorth_z.data(), eq.grad.x.data(),
     eq.grad.y.data(), eq.grad.z.data(), eq.limiter.data(), mass.data(), viscosity.data(), scalar_kind.data(),
     correction_faces.data(), mesh.n_internal, t.second_order ? 1 : 0, 0);
  fv("fv_transport_assemble", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
     stencil.adjacent.data(), stencil.matrix_entry.data(), matrix.row_ptr.data(), matrix.diag.data(),
     stencil.diffusion.data(), mass.data(), viscosity.data(), scalar_kind.data(), bc.data(),
     correction_faces.data(), src.data(), mesh.volume.data(), field.data(), eq.zero->data(),
     wall_diagonal.data(), wall_rhs.data(), matrix.val.data(), eq.rhs.data(), eq.zero->data(),
     eq.diagonal.data(), eq.zero->data(), eq.consistent_inverse.data(), mesh.n_internal, t.density,
     t.pseudo_time_step, t.relaxation, 0);
}

// :1045-1060 (legacy paired mode: no compensated assembly).
void Steady::paired_momentum_residual(int k, V& r) {
  auto& eq = momentum[k];
  fv("fv_paired_momentum_residual", mesh.n_cells, matrix.row_ptr.data(), matrix.col.data(), matrix.val.data(),
     matrix.diag.data(), eq.diagonal.data(), eq.zero->data(), velocity[k].data(), velocity_lo[k].data(),
     pressure_gradient_lo[k].data(), mesh.volume.data(), eq.rhs.data(), eq.zero->data(), r.data());
}

// fv/transport.cpp:123-159 (Jacobi preconditioner, no steady stopping).
SolveStats Steady::solve_transport(int k, V& field, const SolveControl& control) {
  auto& eq = momentum[k];
  eq.solver->refresh(matrix);
  return eq.solver->solve(matrix, field, eq.rhs, control);
}

// :607-614
void Steady::update_momentum_diagonal() {
  fv("fv_momentum_diagonal", mesh.n_cells, momentum[0].diagonal.data(), momentum[1].diagonal.data(),
     momentum[2].diagonal.data(), momentum[0].consistent_inverse.data(), momentum[1].consistent_inverse.data(),
     momentum[2].consistent_inverse.data(), physical_diagonal.data(), consistent_inverse.data());
}

// :1026-1034
void Steady::update_pressure_gradient_lo() {
  auto& g = pressure_gradient_lo;
  g.assign(mesh.n_cells);
  fv("fv_pressure_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), pressure_lo.data(),
     g.x.data(), g.y.data(), g.z.data(), correction_boundary.data(), pressure_boundary_lo.data(),
     mesh.n_internal);
  gradient(pressure_lsq, pressure_lo, pressure_boundary_lo, g);
}

// :943-949
void Steady::refresh_paired_scales() {
  if (!paired) return;
  fv("fv_speed", mesh.n_cells, velocity.x.data(), velocity.y.data(), velocity.z.data(), speed.data());
  paired_speed = max_abs(speed.data(), speed.size());
  paired_pressure_magnitude = max_abs(pressure.data(), pressure.size());
}

// :950-954
void Steady::finish_step() {
  ++report_.iteration;
  if (report_.iteration <= 5 || report_.iteration % 20 == 0) refresh_paired_scales();
  if (report_.iteration <= 5 || report_.iteration % settings.monitor_interval == 0) monitor();
}

// :1303-1369
void Steady::step() {
  refresh_time_step();
  update_boundaries();
  update_pressure();
  update_turbulence();
  fv("fv_momentum_source", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
     mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), face_pressure.data(), mesh.volume.data(),
     body.x.data(), body.y.data(), body.z.data(), source.x.data(), source.y.data(), source.z.data());
  reconstruct_velocity();
  add_transpose_stress();
  auto t = transport();
  for (int k = 0; k < 3; ++k) {
    t.limiter_epsilon2 = limiter_epsilon2[k];
    t.limiter_low = paired ? &velocity_lo[k] : nullptr;
    wall_terms(k);
    assemble(k, t);
    if (!paired) {
      report_.momentum[k] = solve_transport(k, velocity[k], settings.momentum_linear);
      continue;
    }
    paired_momentum_residual(k, residual_u);
    momentum[k].rhs = residual_u;
    std::fill(momentum_delta.begin(), momentum_delta.end(), 0.0f);
    report_.momentum[k] = solve_transport(k, momentum_delta, settings.momentum_linear);
    fv("fv_paired_accumulate", mesh.n_cells, momentum_delta.data(), velocity[k].data(), velocity_lo[k].data(),
       paired_slack().velocity);
  }
  update_momentum_diagonal();
  update_boundaries();
  update_flux();
  if (paired)
    fv("fv_paired_continuity", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
       mass.data(), mass_lo.data(), imbalance.data(), provisional_continuity.data());
  else
    fv("fv_continuity", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(), mass.data(),
       provisional_continuity.data());
  fv_pressure_assemble(mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
                       stencil.adjacent.data(), stencil.matrix_entry.data(), matrix.row_ptr.data(),
                       matrix.diag.data(), coefficient.data(), mass.data(), matrix.val.data(),
                       pressure_rhs.data(), reference_);
  if (paired) fv("fv_paired_pressure_rhs", mesh.n_cells, imbalance.data(), pressure_rhs.data(), reference_);
  pressure_solver->refresh(matrix);
  std::fill(correction.begin(), correction.end(), 0.0f);
  report_.pressure = pressure_solver->solve(matrix, correction, pressure_rhs, settings.pressure_linear);
  if (paired)
    fv("fv_paired_correct_flux", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), mesh.area.x.data(),
       mesh.area.y.data(), mesh.area.z.data(), correction.data(), coefficient.data(), mass.data(),
       mass_lo.data(), mesh.n_internal, paired_slack().flux_per_area);
  else
    fv("fv_correct_flux", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), correction.data(),
       coefficient.data(), mass.data(), mesh.n_internal);
  grad_correction.assign(mesh.n_cells);
  fv("fv_pressure_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), correction.data(),
     grad_correction.x.data(), grad_correction.y.data(), grad_correction.z.data(), correction_boundary.data(),
     boundary_pressure.data(), mesh.n_internal);
  gradient(pressure_lsq, correction, boundary_pressure, grad_correction);
  if (paired) {
    const auto slack = paired_slack();
    fv("fv_paired_correct_velocity", mesh.n_cells, mesh.volume.data(), consistent_inverse.data(),
       physical_diagonal.data(), grad_correction.x.data(), grad_correction.y.data(), grad_correction.z.data(),
       correction.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(), pressure.data(),
       velocity_lo.x.data(), velocity_lo.y.data(), velocity_lo.z.data(), pressure_lo.data(),
       float(settings.pressure_relaxation), slack.velocity, slack.pressure);
    update_pressure_gradient_lo();
  } else {
    fv("fv_correct_velocity", mesh.n_cells, mesh.volume.data(), consistent_inverse.data(),
       physical_diagonal.data(), grad_correction.x.data(), grad_correction.y.data(), grad_correction.z.data(),
       correction.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(), pressure.data(),
       float(settings.pressure_relaxation));
  }
  solve_turbulence();
  finish_step();
}

// :1417-1477
Report Steady::monitor() {
  refresh_time_step();
  update_boundaries();
  update_pressure();
  update_turbulence();
  fv("fv_momentum_source", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
     mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), face_pressure.data(), mesh.volume.data(),
     body.x.data(), body.y.data(), body.z.data(), source.x.data(), source.y.data(), source.z.data());
  reconstruct_velocity();
  add_transpose_stress();
  auto t = transport();
  V* residuals[]{&residual_u, &residual_v, &residual_w};
  for (int k = 0; k < 3; ++k) {
    t.limiter_epsilon2 = limiter_epsilon2[k];
    t.limiter_low = paired ? &velocity_lo[k] : nullptr;
    wall_terms(k);
    assemble(k, t);
    if (paired)
      paired_momentum_residual(k, *residuals[k]);
    else
      kern::csr_residual(mesh.n_cells, matrix.row_ptr.data(), matrix.col.data(), matrix.val.data(),
                         velocity[k].data(), momentum[k].rhs.data(), residuals[k]->data());
  }
  update_momentum_diagonal();
  fv("fv_monitor_cell", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(), mass.data(),
     velocity.x.data(), velocity.y.data(), velocity.z.data(), physical_diagonal.data(), residual_u.data(),
     residual_v.data(), residual_w.data(), continuity.data(), scale.data(), speed.data(),
     absolute_residual.x.data(), absolute_residual.y.data(), absolute_residual.z.data());
  if (paired)
    fv("fv_paired_continuity", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
       mass.data(), mass_lo.data(), imbalance.data(), continuity.data());
  fv("fv_fluent_momentum_scale", mesh.n_cells, momentum[0].diagonal.data(), momentum[1].diagonal.data(),
     momentum[2].diagonal.data(), mesh.volume.data(), speed.data(), fluent_scale.x.data(), fluent_scale.y.data(),
     fluent_scale.z.data(), float(settings.density), time_step);
  const std::size_t n = mesh.n_cells;
  const double denominator = sum(scale.data(), n);
  const double raw = sum(provisional_continuity.data(), n);
  report_.corrected_mass_imbalance = sum(continuity.data(), n);
  if (report_.iteration > 0 && report_.iteration <= 5) first_five_continuity[report_.iteration - 1] = raw;
  continuity_scale = *std::max_element(std::begin(first_five_continuity), std::end(first_five_continuity));
  const auto scaled = [](double r, double d) {
    return d > 0 ? r / d : (r == 0 ? 0 : std::numeric_limits<double>::infinity());
  };
  ResidualValue* values[]{&report_.u, &report_.v, &report_.w};
  for (int k = 0; k < 3; ++k) {
    values[k]->raw = sum(absolute_residual[k].data(), n);
    values[k]->scaled = scaled(values[k]->raw, denominator);
    values[k]->fluent_normalization = sum(fluent_scale[k].data(), n);
    values[k]->fluent_scaled = scaled(values[k]->raw, values[k]->fluent_normalization);
  }
  report_.continuity = {raw, scaled(raw, continuity_scale), scaled(raw, continuity_scale), continuity_scale};
  report_.maximum_velocity = max_abs(speed.data(), n);
  report_.pseudo_time_step = time_step;
  monitor_turbulence();
  return report_;
}

// :1711-1736 with viscous_face_forces (:1373-1416), laminar walls without
// the projected face reconstruction: FP64 sums of the per-face forces.
ForceReport Steady::forces(const std::vector<std::string>& patches, double reference_pressure) {
  ORACLE_REQUIRE(std::isfinite(reference_pressure), "FV nonfinite force reference pressure");
  const std::size_t nb = mesh.n_boundary();
  V selected(nb, 0.0f);
  ForceReport out;
  std::vector<std::string> seen;
  for (const auto& name : patches) {
    ORACLE_REQUIRE(std::find(seen.begin(), seen.end(), name) == seen.end(), "FV duplicate force patch");
    seen.push_back(name);
  }
  std::sort(seen.begin(), seen.end());
  for (const auto& name : seen) {
    std::size_t p = 0;
    while (p < host.patches.size() && host.patches[p].name != name) ++p;
    ORACLE_REQUIRE(p < host.patches.size() && conditions_[p].kind == 0, "FV force patch must name a wall: " + name);
    for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f) {
      selected[f - mesh.n_internal] = 1.0f;
      out.area_vector = {out.area_vector.x + geom.face_area[f].x, out.area_vector.y + geom.face_area[f].y,
                         out.area_vector.z + geom.face_area[f].z};
    }
  }
  update_boundaries();
  update_pressure();
  reconstruct_velocity();
  Vec3f pf, vf;
  pf.assign(nb);
  vf.assign(nb);
  fv("fv_wall_pressure_force", int(nb), selected.data(), mesh.area.x.data(), mesh.area.y.data(),
     mesh.area.z.data(), face_pressure.data(), pf.x.data(), pf.y.data(), pf.z.data(), mesh.n_internal,
     float(reference_pressure));
  const auto &gu = momentum[0].grad, &gv = momentum[1].grad, &gw = momentum[2].grad;
  if (sst) {  // :1376-1380
    update_turbulence();
    fv("fv_sst_force", int(nb), selected.data(), mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(),
       sst_fields_.at("wall_shear_x").data(), sst_fields_.at("wall_shear_y").data(),
       sst_fields_.at("wall_shear_z").data(), vf.x.data(), vf.y.data(), vf.z.data(), mesh.n_internal);
  } else
  fv("fv_wall_viscous_force", int(nb), mesh.owner.data(), selected.data(), mesh.area.x.data(),
     mesh.area.y.data(), mesh.area.z.data(), mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(),
     mesh.owner_to_face.z.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(), gu.x.data(),
     gu.y.data(), gu.z.data(), gv.x.data(), gv.y.data(), gv.z.data(), gw.x.data(), gw.y.data(), gw.z.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(), vf.x.data(),
     vf.y.data(), vf.z.data(), mesh.n_internal, float(settings.viscosity));
  out.pressure = {sum(pf.x.data(), nb), sum(pf.y.data(), nb), sum(pf.z.data(), nb)};
  out.viscous = {sum(vf.x.data(), nb), sum(vf.y.data(), nb), sum(vf.z.data(), nb)};
  out.reference_pressure = reference_pressure;
  return out;
}

const V* Steady::array(const std::string& name) const {
  const std::map<std::string, const V*> table = {
      {"u", &velocity.x},         {"v", &velocity.y},          {"w", &velocity.z},
      {"p", &pressure},           {"mass", &mass},             {"u_lo", &velocity_lo.x},
      {"v_lo", &velocity_lo.y},   {"w_lo", &velocity_lo.z},    {"p_lo", &pressure_lo},
      {"mass_lo", &mass_lo},      {"correction", &correction}, {"coefficient", &coefficient},
      {"face_pressure", &face_pressure}, {"boundary_u", &boundary_velocity.x},
      {"boundary_pressure", &boundary_pressure}, {"physical_diagonal", &physical_diagonal},
      {"consistent_inverse", &consistent_inverse}, {"grad_px", &grad_p.x},
      {"face_u", &face_velocity.x}, {"matrix_values", &matrix.val}, {"pressure_rhs", &pressure_rhs},
      {"volume", &mesh.volume}, {"weight", &mesh.weight}, {"owner_to_face_x", &mesh.owner_to_face.x},
      {"neighbour_to_face_y", &mesh.neighbour_to_face.y}, {"owner_to_neighbour_z", &mesh.owner_to_neighbour.z},
      {"area_x", &mesh.area.x}, {"wall_normal_distance", &mesh.wall_normal_distance},
      {"wall_distance", &mesh.wall_distance},
      {"velocity_lsq_inverse", &velocity_lsq.inverse}, {"pressure_lsq_inverse", &pressure_lsq.inverse},
      {"velocity_lsq_boundary_x", &velocity_lsq.boundary_delta.x},
      {"pressure_lsq_boundary_y", &pressure_lsq.boundary_delta.y}, {"k", &kinetic}, {"omega", &omega},
      {"viscosity", &viscosity}, {"boundary_k", &boundary_k}, {"boundary_omega", &boundary_omega},
      {"source_x", &source.x}, {"strain_dudy", &strain_gradient[0].y}, {"k_grad_y", &k_eq.grad.y},
      {"omega_rhs", &omega_eq.rhs}};
  const auto it = table.find(name);
  if (it != table.end()) return it->second;
  const auto sst = sst_fields_.find(name);
  return sst == sst_fields_.end() ? nullptr : &sst->second;
}

// SstModel::update (turbulence/sst.cpp:101-158, no intermittency): in[25] as
// ka_simplec_sst_update's inputs; `gradients` runs where the reference calls
// reconstruct_scalar_gradients (:126), between the wall cells and the rates.
void Steady::sst_kernels(const double* o, const float* const* in, const std::function<void()>& gradients) {
  const int n = mesh.n_cells, nf = mesh.n_faces, nb = mesh.n_boundary(), ni = mesh.n_internal;
  ORACLE_REQUIRE(o[4] > 0 && o[13] > 0 && o[15] > 0 && o[16] > 0 && o[17] > 0 && o[18] > 0 && o[6] >= 0 &&
                     !(o[10] != 0 && o[11] != 0),
                 "SST: invalid options");
  auto& F = sst_fields_;
  if (sst_wall.size() != std::size_t(nb)) {  // the standalone model on a laminar solver's walls
    sst_wall.assign(nb, 0);
    for (std::size_t p = 0; p < host.patches.size(); ++p)
      if (conditions_[p].kind == 0)
        for (Index f = host.patches[p].start; f < host.patches[p].start + host.patches[p].size; ++f)
          sst_wall[f - ni] = 1;
  }
  for (const char* name : {"mu_t", "k_source", "k_sink", "omega_source", "omega_sink", "strain", "vorticity", "f1",
                           "f2", "cdkw", "diffusivity_k", "diffusivity_omega", "production_k", "fixed_omega",
                           "wall_production"})
    F[name].resize(n, 0.0f);
  for (const char* name : {"wall_tau", "wall_u_star", "wall_u_tau", "wall_y_plus", "wall_y_plus_utau",
                           "wall_shear_x", "wall_shear_y", "wall_shear_z", "wall_shear_coefficient",
                           "wall_face_omega", "wall_cell_omega", "face_production"})
    F[name].resize(nb, 0.0f);
  F["face_diffusivity_k"].resize(nf, 0.0f);
  F["face_diffusivity_omega"].resize(nf, 0.0f);
  fixed_omega_mask.resize(n, 0);
  selected_wall_face.resize(n, 0);
  const auto ux = in[0], uy = in[1], uz = in[2];
  const float* const* g = in + 3;  // du/dx, du/dy, du/dz, dv/dx, ...
  const float *rho = in[12], *mu = in[13], *k = in[14], *om = in[15];
  const float* const* gk = in + 16;
  const float* const* go = in + 19;
  const float* const* wu = in + 22;
  const float k_floor = float(o[17]), omega_floor = float(o[16]);
  auto d = [&](const char* name) { return F.at(name).data(); };
  fv("sst_wall", nb, mesh.owner.data(), sst_wall.data(), ni, mesh.area.x.data(), mesh.area.y.data(),
     mesh.area.z.data(), ux, uy, uz, wu[0], wu[1], wu[2], mesh.wall_normal_distance.data(), rho, mu, k, om,
     d("wall_tau"), d("wall_shear_x"), d("wall_shear_y"), d("wall_shear_z"), d("wall_u_star"), d("wall_u_tau"),
     d("wall_y_plus"), d("wall_y_plus_utau"), d("face_production"), d("wall_shear_coefficient"), k_floor,
     omega_floor);
  fv("sst_wall_omega", nb, mesh.owner.data(), sst_wall.data(), ni, mesh.wall_normal_distance.data(), rho, mu,
     d("wall_u_star"), d("wall_cell_omega"), d("wall_face_omega"));
  fv("sst_wall_cells", n, mesh.offsets.data(), mesh.faces.data(), sst_wall.data(), mesh.wall_normal_distance.data(),
     d("wall_cell_omega"), d("face_production"), ni, fixed_omega_mask.data(), selected_wall_face.data(),
     d("fixed_omega"), d("wall_production"));
  if (gradients) gradients();
  fv("sst_rates", n, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], d("strain"), d("vorticity"));
  fv("sst_blending_viscosity", n, rho, mu, k, om, mesh.wall_distance.data(), gk[0], gk[1], gk[2], go[0], go[1],
     go[2], d("strain"), d("vorticity"), d("f1"), d("f2"), d("cdkw"), d("mu_t"), d("diffusivity_k"),
     d("diffusivity_omega"), float(o[13]), k_floor, omega_floor, int(o[7] != 0), int(o[9] != 0), float(o[18]));
  const int production_mode = o[10] != 0 ? 1 : (o[11] != 0 ? 2 : 0);
  fv("sst_sources", n, rho, mu, k, om, d("strain"), d("vorticity"), d("f1"), d("cdkw"), d("mu_t"),
     fixed_omega_mask.data(), d("wall_production"), d("k_source"), d("k_sink"), d("omega_source"), d("omega_sink"),
     d("production_k"), int(o[7] != 0), int(o[8] != 0), production_mode, int(o[3] != 0), int(o[5] != 0),
     int(o[12] != 0), float(o[4]), int(o[14] != 0), float(o[15]), k_floor, omega_floor, float(o[6]));
  if (o[0] != 0)
    fv("sst_sustain", n, rho, d("f1"), d("k_source"), d("omega_source"), float(o[1]), float(o[2]));
  fv("sst_face_diffusion", nf, mesh.owner.data(), mesh.neighbour.data(), mesh.weight.data(), sst_wall.data(), mu,
     d("diffusivity_k"), d("diffusivity_omega"), ni, d("face_diffusivity_k"), d("face_diffusivity_omega"));
  for (const auto& [name, v] : {std::pair<const char*, const VI*>{"fixed_omega_mask", &fixed_omega_mask},
                                {"selected_wall_face", &selected_wall_face}}) {
    auto& out = F[name];
    out.resize(v->size());
    std::memcpy(out.data(), v->data(), v->size() * sizeof(int));
  }
}

void Steady::sst_update(const double* o, const float* const* in) { sst_kernels(o, in, {}); }

// :654-680, the default (fused vector fit) path: the strain-rate derivative
// fit of each velocity component plus its face-jump correction.
void Steady::reconstruct_strain() {
  const auto& fit = strain_lsq;
  fv("fv_vector_gradient", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), stencil.lsq_x.data(),
     stencil.lsq_y.data(), stencil.lsq_z.data(), fit.boundary_delta.x.data(), fit.boundary_delta.y.data(),
     fit.boundary_delta.z.data(), fit.inverse.data(), velocity.x.data(), velocity.y.data(), velocity.z.data(),
     boundary_velocity.x.data(), boundary_velocity.y.data(), boundary_velocity.z.data(), strain_base[0].x.data(),
     strain_base[0].y.data(), strain_base[0].z.data(), strain_base[1].x.data(), strain_base[1].y.data(),
     strain_base[1].z.data(), strain_base[2].x.data(), strain_base[2].y.data(), strain_base[2].z.data(),
     mesh.n_cells);
  for (int k = 0; k < 3; ++k) {  // jump_corrected_gradient (fv/operators.cpp:254-267)
    const auto& base = strain_base[k];
    auto& out = strain_gradient[k];
    fv("fv_face_jump", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), mesh.owner_to_face.x.data(),
       mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(), mesh.neighbour_to_face.x.data(),
       mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(), velocity[k].data(),
       boundary_velocity[k].data(), base.x.data(), base.y.data(), base.z.data(), jump_forward.data(),
       jump_reverse.data(), mesh.n_internal);
    fv("fv_gather_jump_gradient", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
       jump_delta.x.data(), jump_delta.y.data(), jump_delta.z.data(), jump_inverse.data(), jump_forward.data(),
       jump_reverse.data(), base.x.data(), base.y.data(), base.z.data(), out.x.data(), out.y.data(), out.z.data(),
       mesh.n_cells);
  }
}

// :646-651
void Steady::turbulence_boundaries() {
  fv("fv_turbulence_boundary", mesh.n_boundary(), mesh.owner.data(), boundary_kind.data(), turbulence_spec.data(),
     turbulence_a.data(), turbulence_b.data(), mass.data(), boundary_velocity.x.data(), boundary_velocity.y.data(),
     boundary_velocity.z.data(), kinetic.data(), omega.data(), sst_fields_.at("wall_face_omega").data(),
     boundary_k.data(), boundary_omega.data(), turbulence_kind.data(), mesh.n_internal, float(settings.density),
     float(settings.viscosity));
}

// :681-712 (SST): strain fit when the velocity changed, SstModel::update with
// the turbulence boundary values and k, omega fits between its wall and
// interior passes, then the face effective viscosity.
void Steady::update_turbulence(bool velocity_changed) {
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
  fv("fv_effective_viscosity", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), boundary_kind.data(),
     mesh.weight.data(), sst_fields_.at("mu_t").data(), viscosity.data(), mesh.n_internal, float(settings.viscosity));
}

// :713-727
void Steady::add_transpose_stress() {
  if (!sst) return;
  const auto &u = momentum[0].grad, &v = momentum[1].grad, &w = momentum[2].grad;
  fv("fv_transpose_stress", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), boundary_kind.data(),
     mesh.weight.data(), viscosity.data(), mesh.area.x.data(), mesh.area.y.data(), mesh.area.z.data(), u.x.data(),
     u.y.data(), u.z.data(), v.x.data(), v.y.data(), v.z.data(), w.x.data(), w.y.data(), w.z.data(),
     transpose_stress.x.data(), transpose_stress.y.data(), transpose_stress.z.data(), mesh.n_internal);
  fv("fv_stress_source", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(), mesh.volume.data(),
     transpose_stress.x.data(), transpose_stress.y.data(), transpose_stress.z.data(), source.x.data(),
     source.y.data(), source.z.data());
}

// :778-795 with TransportEquation::assemble_reconstructed (fv/transport.cpp:
// 83-121): Barth-Jespersen limited k or omega, the model's face diffusivity,
// sources and Patankar sinks, compensated low words; then omega's fixed
// wall-cell rows.
void Steady::assemble_turbulence(bool is_omega) {
  const auto& opt = settings.sst_settings;
  auto& eq = is_omega ? omega_eq : k_eq;
  const V& field = is_omega ? omega : kinetic;
  const V& bc = is_omega ? boundary_omega : boundary_k;
  const V& mu = sst_fields_.at(is_omega ? "face_diffusivity_omega" : "face_diffusivity_k");
  const V& src = sst_fields_.at(is_omega ? "omega_source" : "k_source");
  const V& sink = sst_fields_.at(is_omega ? "omega_sink" : "k_sink");
  const bool second_order = is_omega ? opt.omega_second_order : opt.k_second_order;
  const float dt = opt.pseudo_transient ? time_step : 0.0f;
  const float relaxation = float(is_omega ? opt.omega_relaxation : opt.k_relaxation);
  const bool compensated = opt.compensated_transport;
  fv("fv_limiter", mesh.n_cells, mesh.offsets.data(), stencil.adjacent.data(), mesh.faces.data(), mesh.sign.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
     mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
     field.data(), bc.data(), eq.grad.x.data(), eq.grad.y.data(), eq.grad.z.data(), eq.limiter.data(), -1.0f);
  fv("fv_transport_faces", mesh.n_faces, mesh.owner.data(), mesh.neighbour.data(), mesh.weight.data(),
     mesh.owner_to_face.x.data(), mesh.owner_to_face.y.data(), mesh.owner_to_face.z.data(),
     mesh.neighbour_to_face.x.data(), mesh.neighbour_to_face.y.data(), mesh.neighbour_to_face.z.data(),
     stencil.nonorth_x.data(), stencil.nonorth_y.data(), stencil.nonorth_z.data(), eq.grad.x.data(),
     eq.grad.y.data(), eq.grad.z.data(), eq.limiter.data(), mass.data(), mu.data(), turbulence_kind.data(),
     face_correction_scratch.data(), mesh.n_internal, second_order ? 1 : 0, 0);
  fv("fv_transport_assemble", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(),
     stencil.adjacent.data(), stencil.matrix_entry.data(), matrix.row_ptr.data(), matrix.diag.data(),
     stencil.diffusion.data(), mass.data(), mu.data(), turbulence_kind.data(), bc.data(),
     face_correction_scratch.data(), src.data(), mesh.volume.data(), field.data(), sink.data(), zero.data(),
     zero.data(), matrix.val.data(), eq.rhs.data(), compensated ? eq.rhs_low.data() : zero.data(),
     eq.diagonal.data(), compensated ? eq.diagonal_low.data() : zero.data(), eq.consistent_inverse.data(),
     mesh.n_internal, float(settings.density), dt, relaxation, compensated ? 1 : 0);
  if (is_omega)
    fv("fv_fixed_rows", mesh.n_cells, matrix.row_ptr.data(), matrix.diag.data(), fixed_omega_mask.data(),
       sst_fields_.at("fixed_omega").data(), matrix.val.data(), eq.rhs.data());
}

// TransportEquation::physical_residual (fv/transport.cpp:57-62).
void Steady::physical_residual(ScalarEquation& eq, const V& field, V& result, const VI* excluded) {
  if (!settings.sst_settings.compensated_transport) {
    kern::csr_residual(mesh.n_cells, matrix.row_ptr.data(), matrix.col.data(), matrix.val.data(), field.data(),
                       eq.rhs.data(), result.data());
    return;
  }
  fv("fv_compensated_momentum_residual", mesh.n_cells, matrix.row_ptr.data(), matrix.col.data(),
     matrix.val.data(), matrix.diag.data(), eq.diagonal.data(), eq.diagonal_low.data(), field.data(), zero.data(),
     zero.data(), mesh.volume.data(), eq.rhs.data(), eq.rhs_low.data(), result.data());
  if (excluded) fv("fv_scalar_set_fixed", mesh.n_cells, excluded->data(), zero.data(), result.data());
}

// TransportEquation::solve (fv/transport.cpp:123-159), Jacobi BiCGStab.
SolveStats Steady::solve_scalar(ScalarEquation& eq, V& field, const VI* excluded) {
  const auto& opt = settings.sst_settings;
  const std::size_t n = mesh.n_cells;
  SolveControl control = opt.linear;
  const bool compensated = opt.compensated_transport;
  if (compensated) physical_residual(eq, field, eq.residual, excluded);
  if (opt.steady_tolerance > 0) {
    if (!compensated)
      kern::csr_residual(mesh.n_cells, matrix.row_ptr.data(), matrix.col.data(), matrix.val.data(), field.data(),
                         eq.rhs.data(), eq.work_residual.data());
    else
      eq.work_residual = eq.residual;
    fv("fv_scalar_steady_measure", mesh.n_cells, field.data(), eq.diagonal.data(), eq.work_residual.data(),
       excluded ? excluded->data() : matrix.diag.data(), eq.work_absolute.data(), eq.work_scale.data(),
       excluded ? 1 : 0);
    const double defect = sum(eq.work_absolute.data(), n), scale_sum = sum(eq.work_scale.data(), n);
    const double r2 = ::oracle::dot(eq.work_residual.data(), eq.work_residual.data(), n);
    if (defect <= opt.steady_tolerance * scale_sum && std::isfinite(r2))
      control.abs_tol = std::max(control.abs_tol, std::nextafter(std::sqrt(r2), std::numeric_limits<double>::infinity()));
  }
  eq.solver->refresh(matrix);
  if (!compensated) return eq.solver->solve(matrix, field, eq.rhs, control);
  std::fill(eq.delta.begin(), eq.delta.end(), 0.0f);
  const auto stats = eq.solver->solve(matrix, eq.delta, eq.residual, control);
  for (std::size_t i = 0; i < n; ++i) field[i] = 1.0f * field[i] + 1.0f * eq.delta[i];  // axpbyz(1, 1)
  return stats;
}

// :796-818 (no transition model).
void Steady::solve_turbulence() {
  if (!sst) return;
  const auto& opt = settings.sst_settings;
  update_boundaries();
  update_turbulence();
  assemble_turbulence(false);
  report_.turbulence[0] = solve_scalar(k_eq, kinetic, nullptr);
  fv("fv_floor", mesh.n_cells, kinetic.data(), float(opt.model[17]));
  update_turbulence(false);
  assemble_turbulence(true);
  const auto& fixed = sst_fields_.at("fixed_omega");
  const bool set_fixed = opt.steady_tolerance > 0 || opt.compensated_transport;
  if (set_fixed) fv("fv_scalar_set_fixed", mesh.n_cells, fixed_omega_mask.data(), fixed.data(), omega.data());
  report_.turbulence[1] = solve_scalar(omega_eq, omega, &fixed_omega_mask);
  if (set_fixed) fv("fv_scalar_set_fixed", mesh.n_cells, fixed_omega_mask.data(), fixed.data(), omega.data());
  fv("fv_floor", mesh.n_cells, omega.data(), float(opt.model[16]));
}

// :819-837
void Steady::monitor_turbulence() {
  if (!sst) return;
  const std::size_t n = mesh.n_cells;
  const auto measure = [&](bool is_omega) {
    assemble_turbulence(is_omega);
    auto& eq = is_omega ? omega_eq : k_eq;
    const V& q = is_omega ? omega : kinetic;
    physical_residual(eq, q, residual_u, is_omega ? &fixed_omega_mask : nullptr);
    fv("fv_scalar_monitor", mesh.n_cells, q.data(), eq.diagonal.data(), residual_u.data(),
       is_omega ? fixed_omega_mask.data() : zero_mask.data(), mesh.volume.data(), absolute_residual.x.data(),
       scale.data(), fluent_scale.x.data(), float(settings.density),
       settings.sst_settings.pseudo_transient ? time_step : 0.0f);
    const double raw = sum(absolute_residual.x.data(), n), den = sum(scale.data(), n);
    const double fluent_den = sum(fluent_scale.x.data(), n);
    const auto scaled = [&](double d) {
      return d > 0 ? raw / d : (raw == 0 ? 0 : std::numeric_limits<double>::infinity());
    };
    return ResidualValue{raw, scaled(den), scaled(fluent_den), fluent_den};
  };
  report_.k = measure(false);
  report_.omega = measure(true);
}

// SteadySolver::initialize_fields (:1493-1524).
void Steady::initialize_fields(const std::vector<Vec3>& u, const std::vector<double>& p,
                               const std::vector<double>& k, const std::vector<double>& w) {
  const auto n = std::size_t(mesh.n_cells);
  if (paired) {  // reset_paired (:977-988)
    for (auto* v : {&velocity_lo.x, &velocity_lo.y, &velocity_lo.z, &pressure_gradient_lo.x,
                    &pressure_gradient_lo.y, &pressure_gradient_lo.z, &pressure_lo, &mass_lo})
      std::fill(v->begin(), v->end(), 0.0f);
  }
  ORACLE_REQUIRE(u.size() == n && p.size() == n, "FV initialization field size mismatch");
  ORACLE_REQUIRE(sst ? (k.size() == n && w.size() == n) : (k.empty() && w.empty()),
                 "FV turbulence initialization size mismatch");
  const double ref = reference_ >= 0 ? p[std::size_t(reference_)] : 0;
  for (std::size_t c = 0; c < n; ++c) {
    const float value = float(p[c] - ref);
    ORACLE_REQUIRE(std::isfinite(value), "FV nonfinite pressure initialization");
    velocity.x[c] = float(u[c].x);
    velocity.y[c] = float(u[c].y);
    velocity.z[c] = float(u[c].z);
    pressure[c] = value;
    if (sst) {
      ORACLE_REQUIRE(std::isfinite(k[c]) && k[c] >= 0 && std::isfinite(w[c]) && w[c] > 0,
                     "FV invalid turbulence initialization");
      kinetic[c] = float(k[c]);
      omega[c] = float(w[c]);
      ORACLE_REQUIRE(std::isfinite(kinetic[c]) && std::isfinite(omega[c]),
                     "FV turbulence initialization exceeds float storage");
    }
  }
  std::fill(physical_diagonal.begin(), physical_diagonal.end(), 1.0f);
  std::fill(consistent_inverse.begin(), consistent_inverse.end(), 0.0f);
  update_boundaries();
  update_pressure();
  update_flux();
  update_turbulence();
  for (int pass = 0; pass < 3; ++pass) {
    monitor();
    refresh_paired_scales();
    update_flux();
    fv("fv_continuity", mesh.n_cells, mesh.offsets.data(), mesh.faces.data(), mesh.sign.data(), mass.data(),
       provisional_continuity.data());
    update_boundaries();
    update_pressure();
  }
}

}  // namespace oracle::simplec

namespace {
thread_local std::string g_simplec_error;
}

extern "C" {
const char* oracle_simplec_last_error() { return g_simplec_error.c_str(); }

// A box (nx, ny, nz; hex) on [0,1]^3 with one condition per patch in the
// generator's order xmin, xmax, ymin, ymax, zmin, zmax: kind[p] 0 wall with
// velocity wall[3p..3p+2], 3 symmetry. settings: density, viscosity,
// momentum relaxation, pressure relaxation, pseudo-time scale, monitor
// interval, paired state (0/1), slack ulps, reference point x, y, z.
void* oracle_simplec_create(int nx, int ny, int nz, const int* kind, const double* wall,
                            const double* settings) {
  try {
    using namespace oracle;
    using namespace oracle::simplec;
    auto mesh = make_box(nx, ny, nz, {0, 0, 0}, {1, 1, 1}, false);
    std::vector<Condition> conditions(mesh.patches.size());
    for (std::size_t p = 0; p < conditions.size(); ++p)
      conditions[p] = {kind[p], {wall[3 * p], wall[3 * p + 1], wall[3 * p + 2]}};
    Settings s;
    s.density = settings[0];
    s.viscosity = settings[1];
    s.momentum_relaxation = settings[2];
    s.pressure_relaxation = settings[3];
    s.pseudo_time_scale = settings[4];
    s.monitor_interval = int(settings[5]);
    s.paired_state = settings[6] != 0;
    s.paired_state_slack_ulps = settings[7];
    s.reference_point = {settings[8], settings[9], settings[10]};
    return new Steady(std::move(mesh), conditions, s);
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return nullptr;
  }
}
// As oracle_simplec_create on [0, extent] with inlets and outlets: kind[p] 0 wall, 1
// velocity inlet, 2 pressure outlet with backflow velocity, 3 symmetry,
// 4 pressure outlet with normal backflow; velocity[3p..3p+2] for kinds 0-2,
// pressure[p] and total[p] (backflow total pressure) for outlets.
typedef void (*oracle_boundary_profile)(int patch, const double* point, double* velocity, void* user);
void* oracle_simplec_create_profiled(int nx, int ny, int nz, const double* extent, const int* kind,
                                     const double* velocity, const double* pressure, const int* total,
                                     const int* profiled, oracle_boundary_profile fn, void* user,
                                     const double* settings);
void* oracle_simplec_create_open(int nx, int ny, int nz, const double* extent, const int* kind,
                                 const double* velocity, const double* pressure, const int* total,
                                 const double* settings) {
  return oracle_simplec_create_profiled(nx, ny, nz, extent, kind, velocity, pressure, total, nullptr, nullptr,
                                        nullptr, settings);
}
// As oracle_simplec_create_open with a profile function for the patches
// whose profiled[p] != 0.
void* oracle_simplec_create_profiled(int nx, int ny, int nz, const double* extent, const int* kind,
                                     const double* velocity, const double* pressure, const int* total,
                                     const int* profiled, oracle_boundary_profile fn, void* user,
                                     const double* settings) {
  try {
    using namespace oracle;
    using namespace oracle::simplec;
    auto mesh = make_box(nx, ny, nz, {0, 0, 0}, {extent[0], extent[1], extent[2]}, false);
    std::vector<Condition> conditions(mesh.patches.size());
    for (std::size_t p = 0; p < conditions.size(); ++p)
      conditions[p] = {kind[p], {velocity[3 * p], velocity[3 * p + 1], velocity[3 * p + 2]}, pressure[p], total[p], {}};
    for (auto& c : conditions)
      if (c.kind == 5) {  // mass flow in pressure[p], inward direction in velocity (zero: none)
        c.mass_flow = c.pressure;
        c.direction = c.velocity;
        c.has_direction = !(c.direction.x == 0 && c.direction.y == 0 && c.direction.z == 0);
        c.velocity = {};
        c.pressure = 0;
      }
    for (std::size_t p = 0; p < conditions.size(); ++p)
      if (profiled && profiled[p]) {
        ORACLE_REQUIRE(fn && (kind[p] == 1 || kind[p] == 2), "FV: profiles apply to inlets and backflow");
        conditions[p].profile = [fn, user, p](Vec3 x) {
          const double point[3] = {x.x, x.y, x.z};
          double v[3] = {0, 0, 0};
          fn(int(p), point, v, user);
          return Vec3{v[0], v[1], v[2]};
        };
      }
    Settings s;
    s.density = settings[0];
    s.viscosity = settings[1];
    s.momentum_relaxation = settings[2];
    s.pressure_relaxation = settings[3];
    s.pseudo_time_scale = settings[4];
    s.monitor_interval = int(settings[5]);
    s.paired_state = settings[6] != 0;
    s.paired_state_slack_ulps = settings[7];
    s.reference_point = {settings[8], settings[9], settings[10]};
    return new Steady(std::move(mesh), conditions, s);
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return nullptr;
  }
}
int oracle_simplec_initialize(void* h, const double* u, const double* v, const double* w, const double* p) {
  try {
    auto& s = *static_cast<oracle::simplec::Steady*>(h);
    const auto n = std::size_t(s.cells());
    std::vector<oracle::Vec3> velocity(n);
    for (std::size_t c = 0; c < n; ++c) velocity[c] = {u[c], v[c], w[c]};
    s.initialize(velocity, std::vector<double>(p, p + n));
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
void oracle_simplec_free(void* h) { delete static_cast<oracle::simplec::Steady*>(h); }
// options[11] and the threshold as oracle_solver_create.
int oracle_simplec_set_pressure_options(void* h, const int* options, double threshold) {
  using oracle::AmgCycle;
  using oracle::AmgOptions;
  using oracle::NullSpace;
  using oracle::Structure;
  try {
    AmgOptions o;
    o.pre_sweeps = options[0]; o.post_sweeps = options[1]; o.coarse_sweeps = options[2];
    o.coarse_rows = options[3]; o.max_levels = options[4]; o.aggregation_passes = options[5];
    o.nullspace = options[6] ? NullSpace::Constant : NullSpace::None;
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    ORACLE_REQUIRE((options[8] == 0 || options[8] == 1) && options[9] >= 0 && options[10] >= 0,
                   "AMG: invalid options");
    o.cycle = options[8] ? AmgCycle::K : AmgCycle::V;
    o.k_depth = options[9];
    o.k_min_rows = options[10];
    o.strength_threshold = threshold;
    static_cast<oracle::simplec::Steady*>(h)->set_pressure_options(o);
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
int oracle_simplec_iterate(void* h, int count) {
  try {
    static_cast<oracle::simplec::Steady*>(h)->iterate(count);
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
// Report layout: iteration, u raw/scaled/fluent scaled/normalization, v (4),
// w (4), continuity (4), corrected mass imbalance, maximum velocity, pseudo
// time step, momentum iterations x3, pressure iterations, pressure final
// residual, pressure converged.
int oracle_simplec_report(void* h, int run_monitor, double* out) {
  try {
    auto* s = static_cast<oracle::simplec::Steady*>(h);
    const auto r = run_monitor ? s->monitor() : s->report();
    int k = 0;
    out[k++] = r.iteration;
    for (const auto* v : {&r.u, &r.v, &r.w, &r.continuity}) {
      out[k++] = v->raw;
      out[k++] = v->scaled;
      out[k++] = v->fluent_scaled;
      out[k++] = v->fluent_normalization;
    }
    out[k++] = r.corrected_mass_imbalance;
    out[k++] = r.maximum_velocity;
    out[k++] = r.pseudo_time_step;
    for (const auto& m : r.momentum) out[k++] = m.iterations;
    out[k++] = r.pressure.iterations;
    out[k++] = r.pressure.final_residual;
    out[k++] = r.pressure.converged;
    for (const auto* v : {&r.k, &r.omega}) {
      out[k++] = v->raw;
      out[k++] = v->scaled;
      out[k++] = v->fluent_scaled;
      out[k++] = v->fluent_normalization;
    }
    out[k++] = r.turbulence[0].iterations;
    out[k++] = r.turbulence[1].iterations;
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
// Copies array `name` into out (if non-null); returns its length, -1 if unknown.
long oracle_simplec_array(void* h, const char* name, float* out) {
  const auto* v = static_cast<oracle::simplec::Steady*>(h)->array(name);
  if (!v) return -1;
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(float));
  return long(v->size());
}
// forces() over the wall patches named in `names` (comma separated):
// out[10] = pressure xyz, viscous xyz, area vector xyz, reference pressure.
int oracle_simplec_forces(void* h, const char* names, double reference_pressure, double* out) {
  try {
    std::vector<std::string> patches;
    std::string all(names), item;
    for (char ch : all + ",") {
      if (ch == ',') {
        if (!item.empty()) patches.push_back(item);
        item.clear();
      } else {
        item += ch;
      }
    }
    const auto r = static_cast<oracle::simplec::Steady*>(h)->forces(patches, reference_pressure);
    const double v[]{r.pressure.x, r.pressure.y, r.pressure.z, r.viscous.x, r.viscous.y, r.viscous.z,
                     r.area_vector.x, r.area_vector.y, r.area_vector.z, r.reference_pressure};
    std::memcpy(out, v, sizeof v);
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
int oracle_simplec_sst_update(void* h, const double* options, const float* const* inputs) {
  try {
    static_cast<oracle::simplec::Steady*>(h)->sst_update(options, inputs);
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
// A solver on a general mesh (the PolyMesh arrays of mesh/mesh.hpp:36-68;
// patch names separated by '\n') with per-patch conditions: kind[p] as
// oracle_simplec_create_open, velocity[3p..3p+2] (a mass-flow inlet's
// direction), pressure[p] (a mass-flow inlet's kg/s), total[p], flags[p]
// (bit 0: a mass-flow direction is given), turbulence[p] (-1 none, 0 k and
// omega, 1 intensity and length, 2 intensity and viscosity ratio) with its
// pair turbulence_ab[2p..2p+1]. settings[12]: those of oracle_simplec_create
// and the fixed pseudo time step (0: automatic). sst[29] or null (laminar):
// the SstOptions model[19], then k second order (0/1), omega second order,
// k relaxation, omega relaxation, linear rel_tol, abs_tol, max_iter, steady
// tolerance, compensated transport (0/1), pseudo transient (0/1).
void* oracle_simplec_create_mesh(int n_points, const double* points, int n_faces, const int* face_offsets,
                                 const int* face_nodes, const int* owner, int n_internal, const int* neighbour,
                                 int n_cells, int n_patches, const char* patch_names, const int* patch_start,
                                 const int* patch_size, const int* kind, const double* velocity,
                                 const double* pressure, const int* total, const int* flags, const int* turbulence,
                                 const double* turbulence_ab, const double* settings, const double* sst) {
  try {
    using namespace oracle;
    using namespace oracle::simplec;
    ORACLE_REQUIRE(n_points > 0 && n_faces > 0 && n_internal >= 0 && n_internal <= n_faces && n_cells > 0 &&
                       n_patches > 0,
                   "FV mesh: empty or inconsistent sizes");
    PolyMesh mesh;
    mesh.points.resize(n_points);
    for (int i = 0; i < n_points; ++i) mesh.points[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    mesh.face_offsets.assign(face_offsets, face_offsets + n_faces + 1);
    mesh.face_nodes.assign(face_nodes, face_nodes + face_offsets[n_faces]);
    mesh.owner.assign(owner, owner + n_faces);
    mesh.neighbour.assign(neighbour, neighbour + n_internal);
    mesh.n_cells = n_cells;
    std::string names(patch_names), item;
    std::vector<std::string> split;
    for (char ch : names + "\n") {
      if (ch == '\n') {
        split.push_back(item);
        item.clear();
      } else {
        item += ch;
      }
    }
    ORACLE_REQUIRE(int(split.size()) == n_patches, "FV mesh: one name per patch");
    for (int p = 0; p < n_patches; ++p) mesh.patches.push_back({split[p], patch_start[p], patch_size[p]});
    std::vector<Condition> conditions(n_patches);
    for (int p = 0; p < n_patches; ++p) {
      auto& c = conditions[p];
      c.kind = kind[p];
      c.velocity = {velocity[3 * p], velocity[3 * p + 1], velocity[3 * p + 2]};
      c.pressure = pressure[p];
      c.total = total[p];
      if (c.kind == 5) {
        c.mass_flow = c.pressure;
        c.direction = c.velocity;
        c.has_direction = (flags[p] & 1) != 0;
        c.velocity = {};
        c.pressure = 0;
      }
      c.turbulence = turbulence[p];
      c.turbulence_a = turbulence_ab[2 * p];
      c.turbulence_b = turbulence_ab[2 * p + 1];
    }
    Settings s;
    s.density = settings[0];
    s.viscosity = settings[1];
    s.momentum_relaxation = settings[2];
    s.pressure_relaxation = settings[3];
    s.pseudo_time_scale = settings[4];
    s.monitor_interval = int(settings[5]);
    s.paired_state = settings[6] != 0;
    s.paired_state_slack_ulps = settings[7];
    s.reference_point = {settings[8], settings[9], settings[10]};
    s.pseudo_time_step = settings[11];
    if (sst) {
      s.sst = true;
      auto& t = s.sst_settings;
      for (int k = 0; k < 19; ++k) t.model[k] = sst[k];
      t.k_second_order = sst[19] != 0;
      t.omega_second_order = sst[20] != 0;
      t.k_relaxation = sst[21];
      t.omega_relaxation = sst[22];
      t.linear = {sst[23], sst[24], int(sst[25]), 8};
      t.steady_tolerance = sst[26];
      t.compensated_transport = sst[27] != 0;
      t.pseudo_transient = sst[28] != 0;
    }
    return new Steady(std::move(mesh), conditions, s);
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return nullptr;
  }
}
// initialize_fields: per-cell values in the solver's cell order; k and omega
// null without SST.
int oracle_simplec_initialize_fields(void* h, const double* u, const double* v, const double* w, const double* p,
                                     const double* k, const double* omega) {
  try {
    auto& s = *static_cast<oracle::simplec::Steady*>(h);
    const auto n = std::size_t(s.cells());
    std::vector<oracle::Vec3> velocity(n);
    for (std::size_t c = 0; c < n; ++c) velocity[c] = {u[c], v[c], w[c]};
    s.initialize_fields(velocity, std::vector<double>(p, p + n),
                        k ? std::vector<double>(k, k + n) : std::vector<double>{},
                        omega ? std::vector<double>(omega, omega + n) : std::vector<double>{});
    return 0;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return 1;
  }
}
int oracle_simplec_reference(void* h) { return static_cast<oracle::simplec::Steady*>(h)->reference(); }
// upload_mesh's float array `name` for make_box_hex/tet(nx, ny, nz) after RCM,
// patches marked Wall where wall[p] != 0; copies it into out when non-null
// and returns its length, -1 for an unknown name or an error.
long oracle_upload_mesh(int nx, int ny, int nz, int tet, const int* wall, const char* name, float* out) {
  try {
    using namespace oracle;
    using namespace oracle::simplec;
    auto m = make_box(nx, ny, nz, {0, 0, 0}, {1, 1, 1}, tet != 0);
    renumber_cells_rcm(m);
    const auto g = compute_geometry(m);
    std::vector<bool> is_wall(m.patches.size());
    for (std::size_t p = 0; p < is_wall.size(); ++p) is_wall[p] = wall[p] != 0;
    const auto d = upload_mesh(m, g, is_wall);
    const std::map<std::string, const std::vector<float>*> arrays = {
        {"area_x", &d.area.x}, {"area_y", &d.area.y}, {"area_z", &d.area.z},
        {"owner_to_face_x", &d.owner_to_face.x}, {"owner_to_face_y", &d.owner_to_face.y},
        {"owner_to_face_z", &d.owner_to_face.z}, {"neighbour_to_face_x", &d.neighbour_to_face.x},
        {"neighbour_to_face_y", &d.neighbour_to_face.y}, {"neighbour_to_face_z", &d.neighbour_to_face.z},
        {"owner_to_neighbour_x", &d.owner_to_neighbour.x}, {"owner_to_neighbour_y", &d.owner_to_neighbour.y},
        {"owner_to_neighbour_z", &d.owner_to_neighbour.z}, {"weight", &d.weight}, {"volume", &d.volume},
        {"sign", &d.sign}, {"wall_normal_distance", &d.wall_normal_distance},
        {"wall_distance", &d.wall_distance}};
    const auto it = arrays.find(name);
    if (it == arrays.end()) return -1;
    if (out) std::memcpy(out, it->second->data(), it->second->size() * sizeof(float));
    return long(it->second->size());
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return -1;
  }
}

// The least-squares gradient (fv_gradient with boundary_least_squares' fit,
// fv/operators.cpp:166-211, 296-302) of the linear field
// phi = c[0] + c[1] x + c[2] y + c[3] z on a box (hex or tet, optionally
// perturbed, RCM-renumbered), each patch observed as obs[p] (0 omitted,
// 1 face value, 2 normal-foot value). Cell values are phi at the centroids;
// boundary values are phi where the fit places them (the face centroid, or
// the owner centroid's normal foot on the face plane). Writes the gradient
// (3 * cells floats, x y z per cell) and the centroids (3 * cells doubles).
int oracle_lsq_gradient(int nx, int ny, int nz, int tet, double amplitude, unsigned long long seed,
                        const int* obs, const double* c, float* gradient, double* centroid) {
  try {
    using namespace oracle;
    using namespace oracle::simplec;
    auto m = make_box(nx, ny, nz, {0, 0, 0}, {1, 1, 1}, tet != 0);
    if (amplitude > 0) perturb_box(m, nx, ny, nz, {0, 0, 0}, {1, 1, 1}, amplitude, seed);
    renumber_cells_rcm(m);
    const auto g = compute_geometry(m);
    const auto st = make_stencil(m, g);
    const auto cf = build_cell_faces(m);
    const Index ni = m.n_internal_faces(), nb = m.n_faces() - ni;
    std::vector<Observation> observations(nb);
    for (std::size_t p = 0; p < m.patches.size(); ++p)
      for (Index f = m.patches[p].start; f < m.patches[p].start + m.patches[p].size; ++f)
        observations[f - ni] = Observation(obs[p]);
    const auto fit = boundary_least_squares(m, g, observations);
    const auto phi = [&](Vec3 x) { return c[0] + c[1] * x.x + c[2] * x.y + c[3] * x.z; };
    std::vector<float> cell(m.n_cells), bc(nb, 0.0f), gx(m.n_cells), gy(m.n_cells), gz(m.n_cells);
    for (Index i = 0; i < m.n_cells; ++i) cell[i] = float(phi(g.cell_centroid[i]));
    for (Index f = ni; f < m.n_faces(); ++f) {
      const auto o = m.owner[f];
      Vec3 at = g.face_centroid[f];
      if (observations[f - ni] == Observation::NormalFootValue) {
        const auto s = g.face_area[f];
        const auto d = g.face_centroid[f] - g.cell_centroid[o];
        const auto foot = s * (dot(d, s) / dot(s, s));
        at = {g.cell_centroid[o].x + foot.x, g.cell_centroid[o].y + foot.y, g.cell_centroid[o].z + foot.z};
      }
      bc[f - ni] = float(phi(at));
    }
    fv("fv_gradient", m.n_cells, cf.offsets.data(), st.adjacent.data(), st.lsq_x.data(), st.lsq_y.data(),
       st.lsq_z.data(), fit.boundary_delta.x.data(), fit.boundary_delta.y.data(), fit.boundary_delta.z.data(),
       fit.inverse.data(), cell.data(), bc.data(), gx.data(), gy.data(), gz.data(), m.n_cells);
    for (Index i = 0; i < m.n_cells; ++i) {
      gradient[3 * i] = gx[i];
      gradient[3 * i + 1] = gy[i];
      gradient[3 * i + 2] = gz[i];
      centroid[3 * i] = g.cell_centroid[i].x;
      centroid[3 * i + 1] = g.cell_centroid[i].y;
      centroid[3 * i + 2] = g.cell_centroid[i].z;
    }
    return m.n_cells;
  } catch (const std::exception& e) {
    g_simplec_error = e.what();
    return -1;
  }
}
// Cell centroids of the renumbered mesh, x y z per cell (3 * cells doubles).
int oracle_simplec_centroids(void* h, double* out) {
  const auto& g = static_cast<oracle::simplec::Steady*>(h)->geometry();
  for (std::size_t c = 0; c < g.cell_centroid.size(); ++c) {
    out[3 * c] = g.cell_centroid[c].x;
    out[3 * c + 1] = g.cell_centroid[c].y;
    out[3 * c + 2] = g.cell_centroid[c].z;
  }
  return int(g.cell_centroid.size());
}
}
