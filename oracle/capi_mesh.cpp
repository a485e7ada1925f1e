// ORACLE — test infrastructure only (see oracle.hpp). C interface to the mesh
// and pressure-operator restatement: one call builds the whole pipeline and
// returns a bundle of named arrays.
#include <cstring>
#include <map>
#include <string>

#include "oracle_mesh.hpp"

using namespace oracle;

namespace {
struct Bundle {
  std::map<std::string, std::vector<int>> ints;
  std::map<std::string, std::vector<float>> floats;
  std::map<std::string, std::vector<double>> doubles;
};
thread_local std::string g_mesh_error;

void put_mesh(Bundle& b, const PolyMesh& m) {
  std::vector<double> xyz;
  for (const auto& p : m.points) {
    xyz.push_back(p.x);
    xyz.push_back(p.y);
    xyz.push_back(p.z);
  }
  b.doubles["points"] = xyz;
  b.ints["face_offsets"] = m.face_offsets;
  b.ints["face_nodes"] = m.face_nodes;
  b.ints["owner"] = m.owner;
  b.ints["neighbour"] = m.neighbour;
  std::vector<int> patches;
  for (const auto& p : m.patches) {
    patches.push_back(p.start);
    patches.push_back(p.size);
  }
  b.ints["patches"] = patches;
  b.ints["n_cells"] = {m.n_cells};
}
void put_vec3(Bundle& b, const std::string& name, const std::vector<Vec3>& v) {
  std::vector<double> out;
  for (const auto& p : v) {
    out.push_back(p.x);
    out.push_back(p.y);
    out.push_back(p.z);
  }
  b.doubles[name] = out;
}
}  // namespace

extern "C" {

const char* oracle_mesh_last_error() { return g_mesh_error.c_str(); }

// Box mesh (tet or hex) on [0,1]^3, optional seeded perturbation, then RCM,
// geometry, stencil and fv_pressure_assemble with the diffusion coefficients
// and zero flux (the pure-Neumann pressure operator when reference < 0).
void* oracle_pressure_problem(int nx, int ny, int nz, int tet, double amplitude,
                              unsigned long long seed, int reference, int renumber) {
  try {
    auto* b = new Bundle;
    PolyMesh m = make_box(nx, ny, nz, {0, 0, 0}, {1, 1, 1}, tet != 0);
    int rounds = 0;
    if (amplitude > 0) rounds = perturb_box(m, nx, ny, nz, {0, 0, 0}, {1, 1, 1}, amplitude, seed);
    b->ints["rounds"] = {rounds};
    put_mesh(*b, m);
    std::vector<double> before = b->doubles["points"];
    b->doubles["points_generated"] = before;
    b->ints["generated_face_nodes"] = m.face_nodes;
    b->ints["generated_owner"] = m.owner;
    b->ints["generated_neighbour"] = m.neighbour;
    if (renumber) {
      const auto perm = renumber_cells_rcm(m);
      b->ints["cells_new_to_old"] = perm.cells_new_to_old;
      b->ints["faces_new_to_old"] = perm.faces_new_to_old;
      put_mesh(*b, m);
    }
    const Geometry g = compute_geometry(m);
    put_vec3(*b, "face_centroid", g.face_centroid);
    put_vec3(*b, "face_area", g.face_area);
    put_vec3(*b, "cell_centroid", g.cell_centroid);
    b->doubles["cell_volume"] = g.cell_volume;
    const CellFaces cf = build_cell_faces(m);
    b->ints["cell_face_offsets"] = cf.offsets;
    b->ints["cell_faces"] = cf.faces;
    std::vector<float> sign(cf.signs.begin(), cf.signs.end());
    b->floats["cell_face_sign"] = sign;
    const Stencil s = make_stencil(m, g);
    b->ints["row_ptr"] = s.pattern.row_ptr;
    b->ints["col"] = s.pattern.col;
    b->ints["diag"] = s.pattern.diag;
    b->ints["adjacent"] = s.adjacent;
    b->ints["matrix_entry"] = s.matrix_entry;
    b->floats["diffusion"] = s.diffusion;
    b->floats["lsq_x"] = s.lsq_x;
    b->floats["lsq_y"] = s.lsq_y;
    b->floats["lsq_z"] = s.lsq_z;
    b->floats["lsq_inverse"] = s.lsq_inverse;
    b->floats["nonorth_x"] = s.nonorth_x;
    b->floats["nonorth_y"] = s.nonorth_y;
    b->floats["nonorth_z"] = s.nonorth_z;
    std::vector<float> volume(g.cell_volume.size());
    for (std::size_t c = 0; c < volume.size(); ++c) volume[c] = float(g.cell_volume[c]);
    b->floats["volume"] = volume;
    const int n = m.n_cells;
    // reference -2: the reference solver's default pin, the cell holding the
    // origin (fv/steady_solver.cpp:247, reference_point default (0,0,0)).
    if (reference == -2) reference = locate_cell(m, g, {0, 0, 0});
    b->ints["reference"] = {reference};
    std::vector<float> values(s.pattern.col.size(), 0.0f), rhs(n, 0.0f), flux(m.n_faces(), 0.0f);
    fv_pressure_assemble(n, cf.offsets.data(), cf.faces.data(), sign.data(), s.adjacent.data(),
                         s.matrix_entry.data(), s.pattern.row_ptr.data(), s.pattern.diag.data(),
                         s.diffusion.data(), flux.data(), values.data(), rhs.data(), reference);
    b->floats["values"] = values;
    b->floats["rhs"] = rhs;
    return b;
  } catch (const std::exception& e) {
    g_mesh_error = e.what();
    return nullptr;
  }
}

void oracle_bundle_free(void* h) { delete static_cast<Bundle*>(h); }

// kind: 0 int32, 1 float32, 2 float64. Returns length (or -1), copies if out.
long oracle_bundle_array(void* h, const char* name, int kind, void* out) {
  auto* b = static_cast<Bundle*>(h);
  auto copy = [&](const auto& map, std::size_t elem) -> long {
    const auto it = map.find(name);
    if (it == map.end()) return -1;
    if (out) std::memcpy(out, it->second.data(), it->second.size() * elem);
    return long(it->second.size());
  };
  if (kind == 0) return copy(b->ints, sizeof(int));
  if (kind == 1) return copy(b->floats, sizeof(float));
  return copy(b->doubles, sizeof(double));
}

// kernels/fv_flow.kern.h:164-183 on caller arrays.
void oracle_fv_pressure_assemble(int n, const int* offsets, const int* faces, const float* sign,
                                 const int* adjacent, const int* entry, const int* rows,
                                 const int* diag, const float* coefficient, const float* flux,
                                 float* values, float* rhs, int reference) {
  fv_pressure_assemble(n, offsets, faces, sign, adjacent, entry, rows, diag, coefficient, flux,
                       values, rhs, reference);
}

}  // extern "C"
