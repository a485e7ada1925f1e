// Host mesh setup for the MI355X backend: box generators, FP64 geometry,
// deterministic reverse Cuthill-McKee renumbering, cell->face incidence and
// the FV stencil. Every output must equal the reference's bit for bit
// (io/generators.cpp, mesh/mesh.cpp, mesh/topology.cpp, fv/operators.cpp),
// because the CSR pattern, the coefficients and therefore the AMG aggregates
// are derived from them.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hip_device.hpp"

namespace kutta_amd {

struct Vec3 {
  double x = 0, y = 0, z = 0;
};

// mesh/mesh.hpp:36-68 (patch kinds, zones and face zones are not on the path)
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
struct MeshGeometry {
  std::vector<Vec3> face_centroid, face_area, cell_centroid;
  std::vector<double> cell_volume;
};

// io/generators.cpp:98-103 (box on [lo, hi]; tet = six Freudenthal tetrahedra per brick)
PolyMesh make_box(int nx, int ny, int nz, Vec3 lo, Vec3 hi, bool tet);
// mesh/mesh.cpp:168-171; strict=false skips the validity exceptions.
MeshGeometry compute_geometry(const PolyMesh& m, bool strict = true);
// mesh/mesh.cpp:188-218 orientation and closure checks on computed geometry.
void validate(const PolyMesh& m, const MeshGeometry& g);

// mesh/topology.hpp:10-31
struct CellFaces {
  std::vector<Index> offsets, faces;
  std::vector<std::int8_t> signs;
};
CellFaces build_cell_faces(const PolyMesh& m);
struct CellNeighbours {
  std::vector<Index> offsets, cells;
};
CellNeighbours build_cell_neighbours(const PolyMesh& m);
struct MeshPermutation {
  std::vector<Index> cells_new_to_old, faces_new_to_old;
};
MeshPermutation renumber_cells_rcm(PolyMesh& m);

// Seeded point perturbation of a box mesh for the tetrahedral Poisson
// workload (no reference counterpart); returns the redraw rounds.
int perturb_box(PolyMesh& m, int nx, int ny, int nz, Vec3 lo, Vec3 hi, double amplitude,
                std::uint64_t seed);

// fv/operators.cpp:43-91 and fv/operators.hpp:16-24
struct Stencil {
  std::vector<Index> row_ptr{0}, col, diag;   // CSR pattern, diagonal present
  std::vector<Index> adjacent, matrix_entry;  // per cell-face entry
  std::vector<float> lsq_x, lsq_y, lsq_z;     // LSQ rows d/|d|^2 per entry
  std::vector<float> lsq_inverse;             // six planes of n_cells
  std::vector<float> diffusion;               // |S|^2 / (S.d) per face
  std::vector<float> nonorth_x, nonorth_y, nonorth_z;
};
Stencil make_stencil(const PolyMesh& m, const MeshGeometry& g);

// fv/steady_solver.cpp:59-75: the cell whose face planes all enclose `point`,
// else the cell with the nearest centroid. The SIMPLEC solver pins this cell
// (reference_point, default the origin) in closed domains (:247).
Index locate_cell(const PolyMesh& m, const MeshGeometry& g, Vec3 point);

}  // namespace kutta_amd
