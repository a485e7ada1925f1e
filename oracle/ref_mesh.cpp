// ORACLE — test infrastructure only. C wrapper over the reference's own mesh
// construction, geometry and RCM code (io/generators.cpp, mesh/mesh.cpp,
// mesh/topology.cpp compiled from the reference tree into
// oracle/_ref/libref_mesh.so), used to cross-check the oracle's restatement.
#include <cstring>
#include <string>
#include <vector>

#include "kutta/io/mesh_io.hpp"
#include "kutta/mesh/mesh.hpp"
#include "kutta/mesh/topology.hpp"

namespace kutta {
PolyMesh make_box_hex(int nx, int ny, int nz, Vec3d lo, Vec3d hi);
PolyMesh make_box_tet(int nx, int ny, int nz, Vec3d lo, Vec3d hi);
}  // namespace kutta

namespace {
struct Handle {
  kutta::PolyMesh mesh;
  kutta::MeshPermutation perm;
  kutta::MeshGeometry geometry;
};
thread_local std::string g_error;
template <class T>
long out_vec(const std::vector<T>& v, void* out) {
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return long(v.size());
}
long out_vec3(const std::vector<kutta::Vec3d>& v, void* out) {
  if (out)
    for (std::size_t i = 0; i < v.size(); ++i) {
      static_cast<double*>(out)[3 * i] = v[i].x;
      static_cast<double*>(out)[3 * i + 1] = v[i].y;
      static_cast<double*>(out)[3 * i + 2] = v[i].z;
    }
  return long(3 * v.size());
}
}  // namespace

extern "C" {
const char* ref_mesh_last_error() { return g_error.c_str(); }

void* ref_box(int nx, int ny, int nz, int tet) {
  try {
    auto* h = new Handle;
    h->mesh = tet ? kutta::make_box_tet(nx, ny, nz, {0, 0, 0}, {1, 1, 1})
                  : kutta::make_box_hex(nx, ny, nz, {0, 0, 0}, {1, 1, 1});
    return h;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void ref_free(void* h) { delete static_cast<Handle*>(h); }
void ref_set_points(void* h, const double* xyz) {
  auto& pts = static_cast<Handle*>(h)->mesh.points;
  for (std::size_t i = 0; i < pts.size(); ++i) pts[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
}
int ref_renumber(void* h) {
  try {
    auto* p = static_cast<Handle*>(h);
    p->perm = kutta::renumber_cells_rcm(p->mesh);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}
int ref_geometry(void* h) {
  try {
    auto* p = static_cast<Handle*>(h);
    p->geometry = kutta::compute_geometry(p->mesh);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}
long ref_array(void* h, const char* name, void* out) {
  auto* p = static_cast<Handle*>(h);
  const std::string n = name;
  if (n == "face_offsets") return out_vec(p->mesh.face_offsets, out);
  if (n == "face_nodes") return out_vec(p->mesh.face_nodes, out);
  if (n == "owner") return out_vec(p->mesh.owner, out);
  if (n == "neighbour") return out_vec(p->mesh.neighbour, out);
  if (n == "cells_new_to_old") return out_vec(p->perm.cells_new_to_old, out);
  if (n == "faces_new_to_old") return out_vec(p->perm.faces_new_to_old, out);
  if (n == "points") return out_vec3(p->mesh.points, out);
  if (n == "face_centroid") return out_vec3(p->geometry.face_centroid, out);
  if (n == "face_area") return out_vec3(p->geometry.face_area, out);
  if (n == "cell_centroid") return out_vec3(p->geometry.cell_centroid, out);
  if (n == "cell_volume") return out_vec(p->geometry.cell_volume, out);
  if (n == "patches") {
    std::vector<int> v;
    for (const auto& q : p->mesh.patches) {
      v.push_back(q.start);
      v.push_back(q.size);
    }
    return out_vec(v, out);
  }
  if (n == "cell_face_offsets" || n == "cell_faces") {
    const auto cf = kutta::build_cell_faces(p->mesh);
    return out_vec(n == "cell_faces" ? cf.faces : cf.offsets, out);
  }
  return -1;
}
}
