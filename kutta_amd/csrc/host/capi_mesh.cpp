// C ABI of the mesh and pressure-operator setup (include/kutta_amd.h).
#include <cstring>
#include <string>

#include "capi_internal.hpp"
#include "kutta_amd.h"

using namespace kutta_amd;
using capi::guard;
using capi::MatrixHandle;
using capi::Problem;

namespace {
template <class T>
long copy_out(const std::vector<T>& v, void* out) {
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return long(v.size());
}
long copy_vec3(const std::vector<Vec3>& v, void* out) {
  if (out)
    for (std::size_t i = 0; i < v.size(); ++i) {
      static_cast<double*>(out)[3 * i] = v[i].x;
      static_cast<double*>(out)[3 * i + 1] = v[i].y;
      static_cast<double*>(out)[3 * i + 2] = v[i].z;
    }
  return long(3 * v.size());
}
PressureProblem& problem_of(void* h) { return *static_cast<Problem*>(static_cast<MatrixHandle*>(h))->problem; }

long setup_array(const MeshSetup& p, const std::string& n, int kind, void* out) {
  if (kind == 0) {
    if (n == "face_offsets") return copy_out(p.mesh.face_offsets, out);
    if (n == "face_nodes") return copy_out(p.mesh.face_nodes, out);
    if (n == "owner") return copy_out(p.mesh.owner, out);
    if (n == "neighbour") return copy_out(p.mesh.neighbour, out);
    if (n == "cells_new_to_old") return copy_out(p.perm.cells_new_to_old, out);
    if (n == "faces_new_to_old") return copy_out(p.perm.faces_new_to_old, out);
    if (n == "cell_face_offsets") return copy_out(p.cell_faces.offsets, out);
    if (n == "cell_faces") return copy_out(p.cell_faces.faces, out);
    if (n == "row_ptr") return copy_out(p.stencil.row_ptr, out);
    if (n == "col") return copy_out(p.stencil.col, out);
    if (n == "diag") return copy_out(p.stencil.diag, out);
    if (n == "adjacent") return copy_out(p.stencil.adjacent, out);
    if (n == "matrix_entry") return copy_out(p.stencil.matrix_entry, out);
    if (n == "rounds") return copy_out(std::vector<int32_t>{p.rounds}, out);
    if (n == "patches") {
      std::vector<int32_t> v;
      for (const auto& q : p.mesh.patches) {
        v.push_back(q.start);
        v.push_back(q.size);
      }
      return copy_out(v, out);
    }
  } else if (kind == 1) {
    if (n == "diffusion") return copy_out(p.stencil.diffusion, out);
    if (n == "lsq_x") return copy_out(p.stencil.lsq_x, out);
    if (n == "lsq_y") return copy_out(p.stencil.lsq_y, out);
    if (n == "lsq_z") return copy_out(p.stencil.lsq_z, out);
    if (n == "lsq_inverse") return copy_out(p.stencil.lsq_inverse, out);
    if (n == "nonorth_x") return copy_out(p.stencil.nonorth_x, out);
    if (n == "nonorth_y") return copy_out(p.stencil.nonorth_y, out);
    if (n == "nonorth_z") return copy_out(p.stencil.nonorth_z, out);
    if (n == "cell_face_sign") {
      std::vector<float> s(p.cell_faces.signs.begin(), p.cell_faces.signs.end());
      return copy_out(s, out);
    }
  } else if (kind == 2) {
    if (n == "points") return copy_vec3(p.mesh.points, out);
    if (n == "face_centroid") return copy_vec3(p.geometry.face_centroid, out);
    if (n == "face_area") return copy_vec3(p.geometry.face_area, out);
    if (n == "cell_centroid") return copy_vec3(p.geometry.cell_centroid, out);
    if (n == "cell_volume") return copy_out(p.geometry.cell_volume, out);
    if (n == "seconds") return copy_out(std::vector<double>(p.seconds, p.seconds + 5), out);
  }
  return -1;
}
}  // namespace

extern "C" {

int ka_pressure_problem_create(void* dev, int nx, int ny, int nz, int tet, double amplitude,
                               unsigned long long seed, int reference, void** out) {
  return guard([&] {
    auto* p = new Problem;
    try {
      p->problem = std::make_unique<PressureProblem>(*static_cast<HipDevice*>(dev), nx, ny, nz,
                                                     tet != 0, amplitude, seed, reference);
    } catch (...) {
      delete p;
      throw;
    }
    *out = static_cast<MatrixHandle*>(p);
  });
}

void ka_pressure_problem_destroy(void* h) { delete static_cast<MatrixHandle*>(h); }

void* ka_pressure_problem_csr(void* h) { return &problem_of(h).host; }

int ka_pressure_problem_assemble(void* h, int reference) {
  return guard([&] { problem_of(h).assemble(reference); });
}

// info[7]: cells, faces, internal faces, nnz, cell-face entries, redraw rounds, reference
int ka_pressure_problem_info(void* h, int32_t* info, double* seconds) {
  return guard([&] {
    const auto& p = problem_of(h);
    info[0] = p.setup.mesh.n_cells;
    info[1] = p.setup.mesh.n_faces();
    info[2] = p.setup.mesh.n_internal_faces();
    info[3] = p.matrix.nnz;
    info[4] = Index(p.setup.cell_faces.faces.size());
    info[5] = p.setup.rounds;
    info[6] = p.reference;
    if (seconds) {
      for (int k = 0; k < 5; ++k) seconds[k] = p.setup.seconds[k];
      seconds[5] = p.upload_seconds;
    }
  });
}

// kind: 0 int32, 1 float32, 2 float64. Returns length, or -1 for an unknown name.
long ka_pressure_problem_array(void* h, const char* name, int kind, void* out) {
  try {
    const auto& p = problem_of(h);
    const std::string n = name;
    if (kind == 1 && n == "values") return copy_out(p.matrix.val.download(), out);
    if (kind == 1 && n == "rhs") return copy_out(p.d_rhs.download(), out);
    return setup_array(p.setup, n, kind, out);
  } catch (const std::exception& e) {
    capi::last_error() = e.what();
    return -1;
  }
}

int ka_mesh_setup_create(int nx, int ny, int nz, int tet, double amplitude, unsigned long long seed,
                         void** out) {
  return guard([&] {
    auto* s = new capi::Setup;
    try {
      s->setup = std::make_unique<MeshSetup>(nx, ny, nz, tet != 0, amplitude, seed);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}
void ka_mesh_setup_destroy(void* h) { delete static_cast<capi::Setup*>(h); }
long ka_mesh_setup_array(void* h, const char* name, int kind, void* out) {
  try {
    return setup_array(*static_cast<capi::Setup*>(h)->setup, name, kind, out);
  } catch (const std::exception& e) {
    capi::last_error() = e.what();
    return -1;
  }
}

}  // extern "C"
