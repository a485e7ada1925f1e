// Internal declarations shared by the HIP translation units of libkutta_hip.so.
// Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "kutta_hip.h"

namespace kh {

// Threads per workgroup for one-index-per-thread kernels: four wave64s.
constexpr int kBlock = 256;
// Upper bound on workgroups of a reduction grid; also sizes the partials scratch.
constexpr int kMaxReduceBlocks = 65536;
// Workgroups of a standalone grid-stride reduction.
constexpr int kReduceGrid = 2048;
// Host reductions batched behind one wait.
constexpr int kMaxHostResults = 256;

struct ProfileSlot {
  bool enabled = false;
  std::size_t only_n = 0;  // time only launches over this many indices (0: all)
  long launches = 0;
  double total_ms = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

}  // namespace kh

struct kh_device {
  int ordinal = 0;
  std::string name;
  hipStream_t stream = nullptr;
  // Reduction scratch: per-workgroup partial sums (two doubles per workgroup
  // for the fused dot pair), consumed by a finisher launch.
  double* partials = nullptr;
  // Host reduction results: device staging and pinned host copy.
  double* results = nullptr;
  double* results_host = nullptr;
  bool capturing = false;
  std::vector<kh::ProfileSlot> profile;
  // Events for the next packet-timed launch (null when not profiling it).
  hipEvent_t packet_start = nullptr;
  hipEvent_t packet_stop = nullptr;
  std::vector<hipEvent_t> event_pool;
  // Logical-bytes accounting (kh_logical_bytes): the size of every live
  // allocation, and while counting, the running totals.
  std::map<std::uintptr_t, std::size_t> allocations;
  bool count_bytes = false;
  double logical_bytes = 0;
  long logical_launches = 0;
};

namespace kh {

void set_error(const std::string& message);
// Throws std::runtime_error with the HIP error string.
void check(hipError_t status, const char* what);

// Grid for n indices, one per thread.
inline unsigned blocks_for(std::size_t n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

// Argument access with validation (count and kind are checked by the table).
inline void* buf(const kh_arg* a, int k) { return a[k].value.ptr; }
inline float f32(const kh_arg* a, int k) { return a[k].value.f32; }
inline int i32(const kh_arg* a, int k) { return a[k].value.i32; }

// One table row per kernel id: name, signature and launcher. The signature
// has one character per parameter: 'b' buffer, 'f' float, 'i' int, 'u' uint.
using Launcher = void (*)(kh_device*, int n, const kh_arg* args);
struct KernelEntry {
  uint32_t id;
  const char* name;
  const char* signature;
  Launcher launch;
  // The launcher passes kh_device::packet_start/stop to hipExtLaunchKernelGGL,
  // so a profiled launch is timed by its own dispatch packet.
  bool packet_timed = false;
};
// Defined by each kernel translation unit.
const std::vector<KernelEntry>& linalg_kernels();
const std::vector<KernelEntry>& gmres_kernels();
const std::vector<KernelEntry>& dense_kernels();
const std::vector<KernelEntry>& learned_kernels();
const std::vector<KernelEntry>& fv_kernels();

// Reduction launchers (kh_reduce.hip).
void launch_reduce_into(kh_device* dev, int mode, const float* x, const float* y,
                        const float* u, const float* v, std::size_t n, float* dst);
void launch_reduce_host(kh_device* dev, int kind, const float* x, const float* y,
                        std::size_t n, double* result);
void launch_fill32(kh_device* dev, void* dst, uint32_t word, std::size_t count);

// Grid size for a reduction over n elements.
inline unsigned reduce_blocks(std::size_t n) {
  std::size_t b = (n + 4 * kBlock - 1) / (4 * kBlock);
  if (b < 1) b = 1;
  if (b > std::size_t(kReduceGrid)) b = kReduceGrid;
  return static_cast<unsigned>(b);
}

}  // namespace kh
