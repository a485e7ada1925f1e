// Handle types shared by the C ABI translation units.
#pragma once

#include <memory>
#include <string>

#include "krylov.hpp"
#include "pressure.hpp"

namespace kutta_amd::capi {

// A device CSR operator handle: an owned upload, or a problem's assembled operator.
struct MatrixHandle {
  virtual ~MatrixHandle() = default;
  virtual DeviceCsr view() const = 0;
};
struct Matrix final : MatrixHandle {
  HipDevice* dev;
  OwnedCsr csr;
  Matrix(HipDevice* d, OwnedCsr c) : dev(d), csr(std::move(c)) {}
  DeviceCsr view() const override { return csr.view(); }
};
struct Problem final : MatrixHandle {
  std::unique_ptr<PressureProblem> problem;
  DeviceCsr view() const override { return problem->matrix.view(); }
};
// Host-only mesh setup handle.
struct Setup {
  std::unique_ptr<MeshSetup> setup;
};
struct Solver {
  HipDevice* dev;
  KrylovSolver solver;
};

std::string& last_error();

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const NonfiniteError& e) {
    last_error() = e.what();
    return 2;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return 1;
  }
}

}  // namespace kutta_amd::capi
