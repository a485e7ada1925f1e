#include "pressure.hpp"

#include <chrono>

namespace kutta_amd {
namespace {
double since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}
}  // namespace

MeshSetup::MeshSetup(int nx, int ny, int nz, bool tet, double amplitude, std::uint64_t seed) {
  using clock = std::chrono::steady_clock;
  const Vec3 lo{0, 0, 0}, hi{1, 1, 1};
  auto t = clock::now();
  mesh = make_box(nx, ny, nz, lo, hi, tet);
  seconds[0] = since(t);
  t = clock::now();
  if (amplitude > 0) rounds = perturb_box(mesh, nx, ny, nz, lo, hi, amplitude, seed);
  seconds[1] = since(t);
  t = clock::now();
  perm = renumber_cells_rcm(mesh);
  seconds[2] = since(t);
  t = clock::now();
  geometry = compute_geometry(mesh);
  seconds[3] = since(t);
  t = clock::now();
  cell_faces = build_cell_faces(mesh);
  stencil = make_stencil(mesh, geometry);
  seconds[4] = since(t);
}

PressureProblem::PressureProblem(HipDevice& d, int nx, int ny, int nz, bool tet, double amplitude,
                                 std::uint64_t seed, int reference)
    : dev(d), setup(nx, ny, nz, tet, amplitude, seed) {
  const auto t = std::chrono::steady_clock::now();
  const PolyMesh& mesh = setup.mesh;
  const CellFaces& cell_faces = setup.cell_faces;
  const Stencil& stencil = setup.stencil;
  // Device layout: the cell->face incidence in cell order with its float
  // signs (fv/device_mesh.cpp:224-227), the stencil's adjacency and CSR entry
  // per incidence (fv/operators.cpp:55-62), and one coefficient per face.
  d_offsets = to_device(dev, cell_faces.offsets);
  d_faces = to_device(dev, cell_faces.faces);
  d_sign = to_device(dev, std::vector<float>(cell_faces.signs.begin(), cell_faces.signs.end()));
  d_adjacent = to_device(dev, stencil.adjacent);
  d_entry = to_device(dev, stencil.matrix_entry);
  d_coefficient = to_device(dev, stencil.diffusion);
  d_flux = DeviceArray<float>(dev, mesh.n_faces());
  dev.fill(d_flux.data(), 0.0f, mesh.n_faces());
  d_rhs = DeviceArray<float>(dev, mesh.n_cells);
  CsrHost pattern;
  pattern.n_rows = mesh.n_cells;
  pattern.row_ptr = stencil.row_ptr;
  pattern.col = stencil.col;
  pattern.diag = stencil.diag;
  pattern.val.assign(stencil.col.size(), 0.0);
  matrix = OwnedCsr(dev, pattern);
  assemble(reference);
  const auto values = matrix.val.download();
  host = std::move(pattern);
  for (std::size_t j = 0; j < values.size(); ++j) host.val[j] = values[j];
  upload_seconds = since(t);
}

void PressureProblem::assemble(int pin) {
  reference = pin == -2 ? locate_cell(setup.mesh, setup.geometry, {0, 0, 0}) : pin;
  KA_REQUIRE(reference >= -1 && reference < setup.mesh.n_cells, "pressure: reference cell out of range");
  dev.launch(KH_FV_PRESSURE_ASSEMBLE, setup.mesh.n_cells,
             {d_offsets, d_faces, d_sign, d_adjacent, d_entry, matrix.row_ptr, matrix.diag,
              d_coefficient, d_flux, matrix.val, d_rhs, reference});
}

}  // namespace kutta_amd
