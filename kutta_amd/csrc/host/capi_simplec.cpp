// C ABI of the steady SIMPLEC solver (include/kutta_amd.h).
#include <cmath>
#include <cstring>
#include <memory>

#include "capi_internal.hpp"
#include "kutta_amd.h"
#include "simplec.hpp"

using namespace kutta_amd;
using capi::guard;

namespace {
struct Simplec {
  HipDevice* dev;
  std::unique_ptr<SteadySolver> solver;
};
Simplec& handle(void* h) { return *static_cast<Simplec*>(h); }
}  // namespace

extern "C" {

namespace {
int create(void* dev, int nx, int ny, int nz, const double* extent, const int32_t* kinds, const double* velocity,
           const double* pressure, const int32_t* total, const int32_t* profiled, ka_boundary_profile fn,
           void* user, const double* settings, void** out) {
  return guard([&] {
    KA_REQUIRE(nx > 0 && ny > 0 && nz > 0, "box dimensions must be positive");
    const Vec3 upper = extent ? Vec3{extent[0], extent[1], extent[2]} : Vec3{1, 1, 1};
    KA_REQUIRE(upper.x > 0 && upper.y > 0 && upper.z > 0 && std::isfinite(upper.x) && std::isfinite(upper.y) &&
                   std::isfinite(upper.z),
               "box extent must be positive and finite");
    auto mesh = make_box(nx, ny, nz, {0, 0, 0}, upper, false);
    std::vector<PatchCondition> conditions;
    for (std::size_t p = 0; p < mesh.patches.size(); ++p)
      conditions.push_back({mesh.patches[p].name, kinds[p],
                            {velocity[3 * p], velocity[3 * p + 1], velocity[3 * p + 2]},
                            pressure ? pressure[p] : 0.0, total && total[p] != 0});
    for (std::size_t p = 0; p < mesh.patches.size(); ++p)
      if (kinds[p] == PatchCondition::MassFlowInlet) {  // mass flow in pressure[p], direction in velocity
        conditions[p].mass_flow = pressure ? pressure[p] : 0.0;
        conditions[p].direction = conditions[p].velocity;
        // The zero vector means no direction; anything else (NaN included) is
        // a given direction that the solver validates.
        const Vec3 d = conditions[p].direction;
        conditions[p].has_direction = !(d.x == 0 && d.y == 0 && d.z == 0);
        conditions[p].velocity = {};
        conditions[p].pressure = 0;
      }
    for (std::size_t p = 0; p < mesh.patches.size(); ++p)
      if (profiled && profiled[p]) {
        KA_REQUIRE(fn, "ka_simplec_create_profiled: profiled patch without a profile function");
        KA_REQUIRE(kinds[p] == PatchCondition::VelocityInlet || kinds[p] == PatchCondition::PressureOutlet,
                   "FV: profiles apply to velocity inlets and backflow velocities");
        conditions[p].profile = [fn, user, p](Vec3 x) {
          const double point[3] = {x.x, x.y, x.z};
          double v[3] = {0, 0, 0};
          fn(int32_t(p), point, v, user);
          return Vec3{v[0], v[1], v[2]};
        };
      }
    SolverSettings s;
    s.density = settings[0];
    s.viscosity = settings[1];
    s.momentum_relaxation = settings[2];
    s.pressure_relaxation = settings[3];
    s.pseudo_time_scale = settings[4];
    s.monitor_interval = int(settings[5]);
    s.paired_state = settings[6] != 0;
    s.paired_state_slack_ulps = settings[7];
    s.reference_point = {settings[8], settings[9], settings[10]};
    auto* h = new Simplec{static_cast<HipDevice*>(dev), nullptr};
    try {
      h->solver = std::make_unique<SteadySolver>(*h->dev, std::move(mesh), std::move(conditions), s);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
}  // namespace

int ka_simplec_create(void* dev, int nx, int ny, int nz, const int32_t* kinds, const double* wall,
                      const double* settings, void** out) {
  const int rc = guard([&] {
    for (int p = 0; p < 6; ++p)
      KA_REQUIRE(kinds[p] == PatchCondition::Wall || kinds[p] == PatchCondition::Symmetry,
                 "ka_simplec_create: walls (0) and symmetry (3) only; ka_simplec_create_open takes the rest");
  });
  if (rc != 0) return rc;
  return create(dev, nx, ny, nz, nullptr, kinds, wall, nullptr, nullptr, nullptr, nullptr, nullptr, settings, out);
}
int ka_simplec_create_open(void* dev, int nx, int ny, int nz, const double* extent, const int32_t* kinds,
                           const double* velocity, const double* pressure, const int32_t* total,
                           const double* settings, void** out) {
  return create(dev, nx, ny, nz, extent, kinds, velocity, pressure, total, nullptr, nullptr, nullptr, settings, out);
}
int ka_simplec_create_profiled(void* dev, int nx, int ny, int nz, const double* extent, const int32_t* kinds,
                               const double* velocity, const double* pressure, const int32_t* total,
                               const int32_t* profiled, ka_boundary_profile profile, void* user,
                               const double* settings, void** out) {
  return create(dev, nx, ny, nz, extent, kinds, velocity, pressure, total, profiled, profile, user, settings, out);
}
int ka_simplec_create_mesh(void* dev, int n_points, const double* points, int n_faces, const int32_t* face_offsets,
                           const int32_t* face_nodes, const int32_t* owner, int n_internal, const int32_t* neighbour,
                           int n_cells, int n_patches, const char* patch_names, const int32_t* patch_start,
                           const int32_t* patch_size, const int32_t* kinds, const double* velocity,
                           const double* pressure, const int32_t* total, const int32_t* flags,
                           const int32_t* turbulence, const double* turbulence_ab, const double* settings,
                           const double* sst, void** out) {
  return guard([&] {
    KA_REQUIRE(dev && points && face_offsets && face_nodes && owner && (n_internal == 0 || neighbour) &&
                   patch_names && patch_start && patch_size && kinds && velocity && pressure && total && flags &&
                   turbulence && turbulence_ab && settings && out,
               "ka_simplec_create_mesh: null argument");
    KA_REQUIRE(n_points > 0 && n_faces > 0 && n_internal >= 0 && n_internal <= n_faces && n_cells > 0 &&
                   n_patches > 0,
               "ka_simplec_create_mesh: empty or inconsistent sizes");
    KA_REQUIRE(face_offsets[0] == 0, "ka_simplec_create_mesh: face offsets must start at 0");
    for (int f = 0; f < n_faces; ++f)
      KA_REQUIRE(face_offsets[f + 1] - face_offsets[f] >= 3, "ka_simplec_create_mesh: a face needs three nodes");
    PolyMesh mesh;
    mesh.points.resize(n_points);
    for (int i = 0; i < n_points; ++i) mesh.points[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    mesh.face_offsets.assign(face_offsets, face_offsets + n_faces + 1);
    mesh.face_nodes.assign(face_nodes, face_nodes + face_offsets[n_faces]);
    for (auto v : mesh.face_nodes) KA_REQUIRE(v >= 0 && v < n_points, "ka_simplec_create_mesh: node out of range");
    mesh.owner.assign(owner, owner + n_faces);
    mesh.neighbour.assign(neighbour, neighbour + n_internal);
    for (auto c : mesh.owner) KA_REQUIRE(c >= 0 && c < n_cells, "ka_simplec_create_mesh: owner out of range");
    for (auto c : mesh.neighbour) KA_REQUIRE(c >= 0 && c < n_cells, "ka_simplec_create_mesh: neighbour out of range");
    mesh.n_cells = n_cells;
    std::vector<std::string> names(1);
    for (const char* c = patch_names; *c; ++c) {
      if (*c == '\n') names.emplace_back();
      else names.back() += *c;
    }
    KA_REQUIRE(int(names.size()) == n_patches, "ka_simplec_create_mesh: one name per patch");
    Index next = n_internal;
    for (int p = 0; p < n_patches; ++p) {
      KA_REQUIRE(patch_start[p] == next && patch_size[p] >= 0, "ka_simplec_create_mesh: patches must tile the boundary faces");
      mesh.patches.push_back({names[p], patch_start[p], patch_size[p]});
      next += patch_size[p];
    }
    KA_REQUIRE(next == n_faces, "ka_simplec_create_mesh: patches must tile the boundary faces");
    std::vector<PatchCondition> conditions;
    for (int p = 0; p < n_patches; ++p) {
      PatchCondition c;
      c.patch = names[p];
      c.kind = kinds[p];
      c.velocity = {velocity[3 * p], velocity[3 * p + 1], velocity[3 * p + 2]};
      c.pressure = pressure[p];
      c.total_backflow = total[p] != 0;
      if (c.kind == PatchCondition::MassFlowInlet) {
        c.mass_flow = pressure[p];
        c.direction = c.velocity;
        c.has_direction = (flags[p] & 1) != 0;
        c.velocity = {};
        c.pressure = 0;
      }
      c.turbulence = turbulence[p];
      c.turbulence_a = turbulence_ab[2 * p];
      c.turbulence_b = turbulence_ab[2 * p + 1];
      conditions.push_back(std::move(c));
    }
    SolverSettings s;
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
      KA_REQUIRE(std::isfinite(sst[23]) && sst[23] >= 0 && std::isfinite(sst[24]) && sst[24] >= 0 && sst[25] >= 1,
                 "ka_simplec_create_mesh: invalid SST linear solve control");
      t.linear = {sst[23], sst[24], int(sst[25]), 8};
      t.steady_tolerance = sst[26];
      t.compensated_transport = sst[27] != 0;
      t.pseudo_transient = sst[28] != 0;
    }
    auto* h = new Simplec{static_cast<HipDevice*>(dev), nullptr};
    try {
      h->solver = std::make_unique<SteadySolver>(*h->dev, std::move(mesh), std::move(conditions), s);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int ka_simplec_initialize_fields(void* solver, const double* u, const double* v, const double* w, const double* p,
                                 const double* k, const double* omega) {
  return guard([&] {
    auto& s = *handle(solver).solver;
    const auto n = std::size_t(s.cells());
    std::vector<Vec3> velocity(n);
    for (std::size_t c = 0; c < n; ++c) velocity[c] = {u[c], v[c], w[c]};
    s.initialize_fields(velocity, std::vector<double>(p, p + n),
                        k ? std::vector<double>(k, k + n) : std::vector<double>{},
                        omega ? std::vector<double>(omega, omega + n) : std::vector<double>{});
  });
}
int ka_simplec_sst_update(void* solver, const double* options, const float* const* inputs) {
  return guard([&] { handle(solver).solver->sst_update(options, inputs); });
}
int ka_simplec_initialize(void* solver, const double* u, const double* v, const double* w, const double* p) {
  return guard([&] {
    auto& s = *handle(solver).solver;
    const auto n = std::size_t(s.cells());
    std::vector<Vec3> velocity(n);
    for (std::size_t c = 0; c < n; ++c) velocity[c] = {u[c], v[c], w[c]};
    s.initialize(velocity, std::vector<double>(p, p + n));
  });
}
void ka_simplec_destroy(void* solver) { delete static_cast<Simplec*>(solver); }
int ka_simplec_set_graphs(void* solver, int on) {
  return guard([&] { handle(solver).solver->set_graphs(on != 0); });
}
int ka_simplec_pressure_levels(void* solver, int32_t* rows, int32_t* nnz, int capacity) {
  int count = -1;
  if (guard([&] {
        const auto levels = handle(solver).solver->pressure_levels();
        for (std::size_t i = 0; i < levels.size() && int(i) < capacity; ++i) {
          rows[i] = levels[i].first;
          nnz[i] = levels[i].second;
        }
        count = int(levels.size());
      }) != 0)
    return -1;
  return count;
}
int ka_simplec_pressure_plan(void* solver, char* out, int capacity) {
  std::string text;
  if (guard([&] { text = handle(solver).solver->pressure_plan(); }) != 0) return -1;
  if (capacity > 0) {
    std::strncpy(out, text.c_str(), std::size_t(capacity));
    out[capacity - 1] = 0;
  }
  return int(text.size());
}
int ka_simplec_set_pressure_options(void* solver, const int32_t* options, double threshold) {
  return guard([&] {
    KA_REQUIRE(options != nullptr, "SIMPLEC: null pressure options");
    KA_REQUIRE((options[8] == 0 || options[8] == 1) && options[9] >= 0 && options[10] >= 0,
               "AMG: invalid options");
    AmgOptions o;
    o.pre_sweeps = options[0];
    o.post_sweeps = options[1];
    o.coarse_sweeps = options[2];
    o.coarse_rows = options[3];
    o.max_levels = options[4];
    o.aggregation_passes = options[5];
    o.nullspace = options[6] ? NullSpace::Constant : NullSpace::None;
    o.structure = options[7] ? Structure::SymmetricMMatrix : Structure::General;
    o.cycle = options[8] ? AmgCycle::K : AmgCycle::V;
    o.k_depth = options[9];
    o.k_min_rows = options[10];
    o.strength_threshold = threshold;
    handle(solver).solver->set_pressure_options(o);
  });
}
int ka_simplec_set_batched_momentum(void* solver, int on, int32_t* used) {
  return guard([&] {
    const bool u = handle(solver).solver->set_batched_momentum(on != 0);
    if (used) *used = u ? 1 : 0;
  });
}
int ka_simplec_iterate(void* solver, int count) {
  return guard([&] { handle(solver).solver->iterate(count); });
}
int ka_simplec_report(void* solver, int run_monitor, double* out) {
  return guard([&] {
    auto& s = *handle(solver).solver;
    const IterationReport r = run_monitor ? s.monitor() : s.report();
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
  });
}
long ka_simplec_field(void* solver, const char* name, float* out) {
  std::size_t size = 0;
  const float* p = nullptr;
  const int rc = guard([&] { p = handle(solver).solver->field(name, &size); });
  if (rc != 0 || !p) return -1;
  if (out && guard([&] { ok(kh_download(handle(solver).dev->raw(), p, out, size * sizeof(float))); }) != 0) return -1;
  return long(size);
}
int ka_simplec_forces(void* solver, const char* patches, double reference_pressure, double* out) {
  return guard([&] {
    std::vector<std::string> names;
    std::string item;
    for (const char* c = patches;; ++c) {
      if (*c == ',' || *c == 0) {
        if (!item.empty()) names.push_back(item);
        item.clear();
        if (*c == 0) break;
      } else {
        item += *c;
      }
    }
    const auto r = handle(solver).solver->forces(names, reference_pressure);
    const double v[]{r.pressure.x, r.pressure.y, r.pressure.z, r.viscous.x, r.viscous.y, r.viscous.z,
                     r.area_vector.x, r.area_vector.y, r.area_vector.z, r.reference_pressure};
    std::memcpy(out, v, sizeof v);
  });
}
int ka_simplec_force_coefficient(const double* report, const double* direction, double density,
                                 double velocity, double area, double pressure, double* out) {
  return guard([&] {
    ForceReport r;
    r.pressure = {report[0], report[1], report[2]};
    r.viscous = {report[3], report[4], report[5]};
    r.area_vector = {report[6], report[7], report[8]};
    r.reference_pressure = report[9];
    *out = r.coefficient({direction[0], direction[1], direction[2]}, density, velocity, area, pressure);
  });
}
int ka_simplec_info(void* solver, int32_t* info) {
  return guard([&] {
    auto& s = *handle(solver).solver;
    info[0] = s.cells();
    info[1] = s.reference_cell();
    info[2] = s.host_waits();
    info[3] = s.report().iteration;
  });
}
int ka_simplec_centroids(void* solver, double* out) {
  return guard([&] {
    const auto& g = handle(solver).solver->geometry();
    for (std::size_t c = 0; c < g.cell_centroid.size(); ++c) {
      out[3 * c] = g.cell_centroid[c].x;
      out[3 * c + 1] = g.cell_centroid[c].y;
      out[3 * c + 2] = g.cell_centroid[c].z;
    }
  });
}
}
