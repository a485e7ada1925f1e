// ORACLE — test infrastructure only. Runs the reference's own upload_mesh
// (fv/device_mesh.cpp, compiled from the reference tree with mesh.cpp,
// topology.cpp and generators.cpp into oracle/_ref/libref_devmesh.so) on a
// host-memory implementation of the reference's Device interface
// (runtime/device.hpp:107-240: allocation, upload and download only), and
// returns the FP32 arrays it builds, to cross-check the oracle's restatement.
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kutta/fv/device_mesh.hpp"
#include "kutta/mesh/mesh.hpp"
#include "kutta/mesh/topology.hpp"

namespace kutta {
PolyMesh make_box_hex(int nx, int ny, int nz, Vec3d lo, Vec3d hi);
PolyMesh make_box_tet(int nx, int ny, int nz, Vec3d lo, Vec3d hi);
}  // namespace kutta

namespace {
struct HostBuffer final : kutta::BufferHandle {
  std::vector<unsigned char> bytes_;
  explicit HostBuffer(std::size_t n) : bytes_(n) {}
  std::size_t bytes() const override { return bytes_.size(); }
};
// Storage only: the mesh upload allocates, uploads and downloads.
class HostDevice final : public kutta::Device {
 public:
  kutta::Backend backend() const override { return kutta::Backend::Cpu; }
  std::string name() const override { return "host storage"; }
  double sum(const kutta::Buffer<float>&, std::size_t) override { throw std::logic_error("unused"); }
  double dot(const kutta::Buffer<float>&, const kutta::Buffer<float>&, std::size_t) override {
    throw std::logic_error("unused");
  }
  double max_abs(const kutta::Buffer<float>&, std::size_t) override { throw std::logic_error("unused"); }
  void launch(kutta::KernelId, std::size_t, std::span<const kutta::KernelArg>) override {
    throw std::logic_error("unused");
  }
  void synchronize() override {}

 protected:
  void reduce_into(const kutta::Buffer<float>&, const kutta::Buffer<float>&, const kutta::Buffer<float>&,
                   const kutta::Buffer<float>&, std::size_t, kutta::Buffer<float>&, std::size_t,
                   Reduction) override {
    throw std::logic_error("unused");
  }
  std::unique_ptr<kutta::BufferHandle> alloc_bytes(std::size_t bytes) override {
    return std::make_unique<HostBuffer>(bytes);
  }
  void upload_bytes(kutta::BufferHandle& dst, const void* src, std::size_t bytes) override {
    if (bytes) std::memcpy(static_cast<HostBuffer&>(dst).bytes_.data(), src, bytes);
  }
  void download_bytes(const kutta::BufferHandle& src, void* dst, std::size_t bytes) override {
    if (bytes) std::memcpy(dst, static_cast<const HostBuffer&>(src).bytes_.data(), bytes);
  }
  void copy_bytes(kutta::BufferHandle& dst, const kutta::BufferHandle& src, std::size_t bytes) override {
    if (bytes) std::memcpy(static_cast<HostBuffer&>(dst).bytes_.data(),
                           static_cast<const HostBuffer&>(src).bytes_.data(), bytes);
  }
  void fill_bytes(kutta::BufferHandle& dst, const void* value, std::size_t width, std::size_t count) override {
    auto& b = static_cast<HostBuffer&>(dst).bytes_;
    for (std::size_t i = 0; i < count; ++i) std::memcpy(b.data() + i * width, value, width);
  }
};
thread_local std::string g_error;
}  // namespace

extern "C" {
const char* ref_devmesh_last_error() { return g_error.c_str(); }

// upload_mesh of make_box_hex/tet(nx, ny, nz) on [0,1]^3 after RCM, patches
// marked Wall where wall[p] != 0 (generator order xmin ... zmax). Copies
// float array `name` into out (when non-null) and returns its length, -1 on
// an unknown name or an error.
long ref_upload_mesh(int nx, int ny, int nz, int tet, const int* wall, const char* name, float* out) {
  try {
    auto mesh = tet ? kutta::make_box_tet(nx, ny, nz, {0, 0, 0}, {1, 1, 1})
                    : kutta::make_box_hex(nx, ny, nz, {0, 0, 0}, {1, 1, 1});
    for (std::size_t p = 0; p < mesh.patches.size(); ++p)
      mesh.patches[p].kind = wall[p] ? kutta::PatchKind::Wall : kutta::PatchKind::Symmetry;
    kutta::renumber_cells_rcm(mesh);
    const auto geometry = kutta::compute_geometry(mesh);
    HostDevice dev;
    const auto m = kutta::upload_mesh(dev, mesh, geometry);
    const std::map<std::string, const kutta::Buffer<float>*> arrays = {
        {"area_x", &m.area.x}, {"area_y", &m.area.y}, {"area_z", &m.area.z},
        {"owner_to_face_x", &m.owner_to_face.x}, {"owner_to_face_y", &m.owner_to_face.y},
        {"owner_to_face_z", &m.owner_to_face.z}, {"neighbour_to_face_x", &m.neighbour_to_face.x},
        {"neighbour_to_face_y", &m.neighbour_to_face.y}, {"neighbour_to_face_z", &m.neighbour_to_face.z},
        {"owner_to_neighbour_x", &m.owner_to_neighbour.x}, {"owner_to_neighbour_y", &m.owner_to_neighbour.y},
        {"owner_to_neighbour_z", &m.owner_to_neighbour.z}, {"weight", &m.weight}, {"volume", &m.volume},
        {"sign", &m.cell_face_sign}, {"wall_normal_distance", &m.wall_normal_distance},
        {"wall_distance", &m.wall_distance}};
    const auto it = arrays.find(name);
    if (it == arrays.end()) return -1;
    const auto values = dev.download(*it->second);
    if (out) std::memcpy(out, values.data(), values.size() * sizeof(float));
    return long(values.size());
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}
}
