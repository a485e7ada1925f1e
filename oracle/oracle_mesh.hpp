// ORACLE — test infrastructure only (see oracle.hpp). Serial restatement of the
// reference's host mesh setup on the path: box generators, FP64 geometry,
// reverse Cuthill-McKee renumbering, cell->face incidence, the FV stencil and
// the pressure-matrix assembly kernel.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "oracle.hpp"

namespace oracle {

struct Vec3 {
  double x = 0, y = 0, z = 0;
};

// mesh/mesh.hpp:36-68 (patch kinds and zones are not on the path).
struct Patch {
  std::string name;
  Index start = 0, size = 0;
};
struct PolyMesh {
  std::vector<Vec3> points;
  std::vector<Index> face_offsets{0};
  std::vector<Index> face_nodes;
  std::vector<Index> owner, neighbour;
  Index n_cells = 0;
  std::vector<Patch> patches;
  Index n_faces() const { return Index(owner.size()); }
  Index n_internal_faces() const { return Index(neighbour.size()); }
};

// mesh/mesh.hpp:79-84
struct Geometry {
  std::vector<Vec3> face_centroid, face_area, cell_centroid;
  std::vector<double> cell_volume;
};

// io/generators.cpp:10-103
PolyMesh make_box(int nx, int ny, int nz, Vec3 lo, Vec3 hi, bool tet);
// mesh/mesh.cpp:168-171 (throws on degenerate faces/cells)
Geometry compute_geometry(const PolyMesh& m);
// The same geometry without the validity exceptions, for perturbation checks.
Geometry geometry_unchecked(const PolyMesh& m);
// mesh/mesh.cpp:188-218: structure and orientation checks
void validate(const PolyMesh& m);

// mesh/topology.hpp:10-20, topology.cpp:108-147
struct CellFaces {
  std::vector<Index> offsets, faces;
  std::vector<std::int8_t> signs;
};
CellFaces build_cell_faces(const PolyMesh& m);
struct CellNeighbours {
  std::vector<Index> offsets, cells;
};
CellNeighbours build_cell_neighbours(const PolyMesh& m);
// mesh/topology.cpp:148-248
struct MeshPermutation {
  std::vector<Index> cells_new_to_old, faces_new_to_old;
};
MeshPermutation renumber_cells_rcm(PolyMesh& m);

// Seeded interior perturbation of a box mesh (no reference counterpart; the
// benchmark's tetrahedral Poisson workload). See mesh.cpp for the procedure.
int perturb_box(PolyMesh& m, int nx, int ny, int nz, Vec3 lo, Vec3 hi, double amplitude,
                std::uint64_t seed);

// fv/operators.cpp:43-91 (pattern, incidence, LSQ rows and inverse, diffusion,
// non-orthogonal vector).
struct Stencil {
  CsrHost pattern;  // values unused
  std::vector<Index> adjacent, matrix_entry;
  std::vector<float> lsq_x, lsq_y, lsq_z;  // per cell-face entry
  std::vector<float> lsq_inverse;          // six planes of n_cells
  std::vector<float> diffusion;            // per face
  std::vector<float> nonorth_x, nonorth_y, nonorth_z;
};
Stencil make_stencil(const PolyMesh& m, const Geometry& g);

// fv/steady_solver.cpp:59-75: the first cell whose face planes all enclose the
// point, else the cell with the nearest centroid.
Index locate_cell(const PolyMesh& m, const Geometry& g, Vec3 point);

// kernels/fv_flow.kern.h:164-183, one call over all cells.
void fv_pressure_assemble(int n, const int* offsets, const int* faces, const float* sign,
                          const int* adjacent, const int* entry, const int* rows, const int* diag,
                          const float* coefficient, const float* flux, float* values, float* rhs,
                          int reference);

}  // namespace oracle
