// The pressure-correction operator of a box mesh as the SIMPLEC solver builds
// it (fv/steady_solver.cpp:240-371): renumber, geometry, stencil, upload, then
// fv_pressure_assemble on the device with the face diffusion coefficients.
#pragma once

#include <memory>

#include "krylov.hpp"
#include "mesh.hpp"

namespace kutta_amd {

// Host half: mesh, renumbering, geometry and stencil (no device needed).
struct MeshSetup {
  MeshSetup(int nx, int ny, int nz, bool tet, double amplitude, std::uint64_t seed);
  PolyMesh mesh;
  MeshGeometry geometry;
  MeshPermutation perm;
  CellFaces cell_faces;
  Stencil stencil;
  int rounds = 0;
  double seconds[5] = {0, 0, 0, 0, 0};  // box, perturb, rcm, geometry, stencil
};

struct PressureProblem {
  PressureProblem(HipDevice& dev, int nx, int ny, int nz, bool tet, double amplitude,
                  std::uint64_t seed, int reference);
  // Re-runs the device assembly with the current coefficients and zero flux.
  // reference: a cell to pin, -1 for none, -2 for the cell holding the origin.
  void assemble(int reference);
  // Host copy of the assembled operator (float values held as double).
  const CsrHost& host_operator() const { return host; }

  HipDevice& dev;
  MeshSetup setup;
  double upload_seconds = 0;  // upload + device assembly + host copy
  Index reference = -1;       // pinned cell, or -1 for the pure-Neumann operator

  DeviceArray<Index> d_offsets, d_faces, d_adjacent, d_entry;
  DeviceArray<float> d_sign, d_coefficient, d_flux, d_rhs;
  OwnedCsr matrix;
  CsrHost host;
};

}  // namespace kutta_amd
