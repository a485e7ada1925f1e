// Device runtime behind the C ABI: one in-order HIP stream per device,
// hipMalloc-backed buffers, a fill kernel (graph-safe), FP64 reductions in a
// fixed order, id-based kernel dispatch, graph capture and event profiling.
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "kh_device_math.hpp"

namespace kh {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& message) { g_error = message; }

void check(hipError_t status, const char* what) {
  if (status != hipSuccess)
    throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(status));
}

namespace {

template <class F>
int guard(F&& f) {
  try {
    f();
    return KH_OK;
  } catch (const std::exception& e) {
    set_error(e.what());
    return KH_ERROR;
  }
}

void require(bool ok, const char* message) {
  if (!ok) throw std::runtime_error(message);
}

// ------------------------------------------------------------------- fill
__global__ void k_fill32(uint32_t* __restrict__ dst, uint32_t word, std::size_t count) {
  std::size_t i = std::size_t(blockIdx.x) * kBlock + threadIdx.x;
  const std::size_t stride = std::size_t(gridDim.x) * kBlock;
  for (; i < count; i += stride) dst[i] = word;
}

// -------------------------------------------------------------- reductions
// Grid-stride FP64 accumulation, deterministic for a given grid: each thread
// visits a fixed index set in increasing order, then fixed trees combine.
enum Mode { kSum = 0, kDot = 1, kDot2 = 2, kMaxAbs = 3 };

__device__ __forceinline__ double nan_max(double a, double b) {
  // NaN-propagating maximum (runtime/cpu/cpu_device.cpp:52-62).
  if (a != a) return a;
  if (b != b) return b;
  return b > a ? b : a;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, 64));
  return v;
}

// Grid-stride FP64 partial sums (or NaN-propagating maxima) per workgroup.
template <int M>
__global__ void __launch_bounds__(kBlock) k_reduce_partials(const float* __restrict__ x,
    const float* __restrict__ y, const float* __restrict__ u, const float* __restrict__ v,
    std::size_t n, double* __restrict__ partials) {
  double a = 0.0, b = 0.0;
  const std::size_t stride = std::size_t(gridDim.x) * kBlock;
  for (std::size_t i = std::size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    if constexpr (M == kSum) a += double(x[i]);
    else if constexpr (M == kMaxAbs) a = nan_max(a, fabs(double(x[i])));
    else a += double(x[i]) * double(y[i]);
    if constexpr (M == kDot2) b += double(u[i]) * double(v[i]);
  }
  __shared__ double lds[2][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (M == kMaxAbs) {
    a = wave_max(a);
    if (lane == 0) lds[0][wave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) m = nan_max(m, lds[0][w]);
      partials[blockIdx.x * 2] = m;
    }
  } else {
    double mine[2] = {a, b};
    mine[0] = block_sum(mine[0], lds[0]);
    if constexpr (M == kDot2) mine[1] = block_sum(mine[1], lds[1]);
    store_partials<2>(mine, partials);
  }
}

// One workgroup combines the partials in slot order; writes pair(s) to dst
// (asynchronous reductions) or the double to result (host reductions).
template <int M>
__global__ void __launch_bounds__(kBlock) k_reduce_finish(const double* __restrict__ partials,
    unsigned count, float* dst, double* result) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (M == kMaxAbs) {
    __shared__ double lds[kBlock / 64];
    double m = 0.0;
    for (unsigned k = threadIdx.x; k < count; k += kBlock) m = nan_max(m, partials[k * 2]);
    m = wave_max(m);
    if (lane == 0) lds[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) t = nan_max(t, lds[w]);
      *result = t;
    }
  } else {
    double total[2];
    sum_partials<2>(partials, count, total);
    if (threadIdx.x != 0) return;
    if (result) {
      *result = total[0];
    } else {
      pstore(dst, 0, to_pair(total[0]));
      if constexpr (M == kDot2) pstore(dst, 2, to_pair(total[1]));
    }
  }
}

template <int M>
void reduce_on(hipStream_t stream, double* partials, const float* x, const float* y, const float* u,
               const float* v, std::size_t n, float* dst, double* result) {
  const unsigned g = reduce_blocks(n);
  k_reduce_partials<M><<<dim3(g), dim3(kBlock), 0, stream>>>(x, y, u, v, n, partials);
  k_reduce_finish<M><<<dim3(1), dim3(kBlock), 0, stream>>>(partials, g, dst, result);
}
template <int M>
void reduce(kh_device* d, const float* x, const float* y, const float* u, const float* v,
            std::size_t n, float* dst, double* result) {
  reduce_on<M>(d->stream, d->partials, x, y, u, v, n, dst, result);
}

// kh_reduce_into_static keeps its per-workgroup partials here instead of in a
// device's hipMalloc'd buffer: 2 doubles per workgroup of the largest
// reduction grid (32 KiB). Each GPU holds its own copy for the life of the
// process; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[2 * kReduceGrid];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxScratchDevices] = {};

// The current device's copy of `scratch`, resolved on its first use.
double* device_scratch() {
  int ordinal = 0;
  check(hipGetDevice(&ordinal), "hipGetDevice");
  require(ordinal >= 0 && ordinal < kMaxScratchDevices, "kh_reduce_into_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[ordinal]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    require(p != nullptr, "kh_reduce_into_static: null scratch address");
    g_scratch[ordinal] = static_cast<double*>(p);
  }
  return g_scratch[ordinal];
}

}  // namespace

namespace {
void fill32_on(hipStream_t stream, void* dst, uint32_t word, std::size_t count) {
  if (count == 0) return;
  std::size_t blocks = (count + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  k_fill32<<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(static_cast<uint32_t*>(dst), word, count);
}
}  // namespace

void launch_fill32(kh_device* d, void* dst, uint32_t word, std::size_t count) {
  fill32_on(d->stream, dst, word, count);
}

void launch_reduce_into(kh_device* d, int mode, const float* x, const float* y, const float* u,
                        const float* v, std::size_t n, float* dst) {
  if (mode == KH_INTO_SUM) reduce<kSum>(d, x, y, u, v, n, dst, nullptr);
  else if (mode == KH_INTO_DOT) reduce<kDot>(d, x, y, u, v, n, dst, nullptr);
  else reduce<kDot2>(d, x, y, u, v, n, dst, nullptr);
}

void launch_reduce_host(kh_device* d, int kind, const float* x, const float* y, std::size_t n,
                        double* result) {
  if (kind == KH_REDUCE_SUM) reduce<kSum>(d, x, y, x, y, n, nullptr, result);
  else if (kind == KH_REDUCE_DOT) reduce<kDot>(d, x, y, x, y, n, nullptr, result);
  else reduce<kMaxAbs>(d, x, y, x, y, n, nullptr, result);
}

namespace {

const KernelEntry* find_kernel(uint32_t id) {
  static const std::unordered_map<uint32_t, const KernelEntry*> index = [] {
    std::unordered_map<uint32_t, const KernelEntry*> m;
    for (const auto& e : linalg_kernels()) m[e.id] = &e;
    for (const auto& e : fv_kernels()) m[e.id] = &e;
    for (const auto& e : gmres_kernels()) m[e.id] = &e;
    for (const auto& e : dense_kernels()) m[e.id] = &e;
    for (const auto& e : learned_kernels()) m[e.id] = &e;
    return m;
  }();
  const auto it = index.find(id);
  return it == index.end() ? nullptr : it->second;
}

hipEvent_t take_event(kh_device* d) {
  if (!d->event_pool.empty()) {
    hipEvent_t e = d->event_pool.back();
    d->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

void no_capture(kh_device* d, const char* what) {
  if (d->capturing) throw std::runtime_error(std::string(what) + ": not allowed during graph capture");
}

}  // namespace
}  // namespace kh

struct kh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace kh;

extern "C" {

const char* kh_last_error(void) { return g_error.c_str(); }

int kh_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

int kh_create(int ordinal, kh_device** out) {
  return guard([&] {
    require(out != nullptr, "kh_create: null output");
    int count = 0;
    check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    require(ordinal >= 0 && ordinal < count, "kh_create: no such HIP device");
    check(hipSetDevice(ordinal), "hipSetDevice");
    auto* d = new kh_device;
    d->ordinal = ordinal;
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, ordinal), "hipGetDeviceProperties");
    d->name = prop.name[0] ? std::string(prop.name) : std::string();
    if (d->name.empty() || d->name.find("gfx") == std::string::npos)
      d->name += (d->name.empty() ? "" : " ") + std::string(prop.gcnArchName);
    check(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipMalloc(&d->partials, sizeof(double) * 2 * kMaxReduceBlocks), "hipMalloc partials");
    check(hipMalloc(&d->results, sizeof(double) * kMaxHostResults), "hipMalloc results");
    check(hipHostMalloc(&d->results_host, sizeof(double) * kMaxHostResults, hipHostMallocDefault),
          "hipHostMalloc results");
    d->profile.resize(KH_KERNEL_COUNT_MARK);
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    *out = d;
  });
}

void kh_destroy(kh_device* d) {
  if (!d) return;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  for (auto& slot : d->profile)
    for (auto& p : slot.pending) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  for (auto e : d->event_pool) (void)hipEventDestroy(e);
  (void)hipFree(d->partials);
  (void)hipFree(d->results);
  (void)hipHostFree(d->results_host);
  (void)hipStreamDestroy(d->stream);
  delete d;
}

const char* kh_name(const kh_device* d) { return d ? d->name.c_str() : ""; }

int kh_alloc(kh_device* d, size_t bytes, void** out) {
  return guard([&] {
    require(d && out, "kh_alloc: null argument");
    no_capture(d, "kh_alloc");
    void* p = nullptr;
    // Zero-byte buffers still get a distinct handle (device.hpp:40-45).
    check(hipMalloc(&p, bytes == 0 ? 4 : bytes), "hipMalloc");
    d->allocations[reinterpret_cast<std::uintptr_t>(p)] = bytes == 0 ? 4 : bytes;
    *out = p;
  });
}

namespace {
// The size of the allocation holding p (0 for memory not from kh_alloc).
std::size_t allocation_bytes(const kh_device* d, const void* p) {
  const auto key = reinterpret_cast<std::uintptr_t>(p);
  auto it = d->allocations.upper_bound(key);
  if (it == d->allocations.begin()) return 0;
  --it;
  return key < it->first + it->second ? it->second : 0;
}
void count_logical(kh_device* d, double bytes) {
  if (!d->count_bytes || d->capturing) return;
  d->logical_bytes += bytes;
  ++d->logical_launches;
}
}  // namespace

int kh_logical_bytes(kh_device* d, int enable, double* bytes, long* launches) {
  return guard([&] {
    require(d != nullptr, "kh_logical_bytes: null device");
    if (bytes) *bytes = d->logical_bytes;
    if (launches) *launches = d->logical_launches;
    d->logical_bytes = 0;
    d->logical_launches = 0;
    d->count_bytes = enable != 0;
  });
}

int kh_free(kh_device* d, void* ptr) {
  return guard([&] {
    require(d != nullptr, "kh_free: null device");
    no_capture(d, "kh_free");
    if (!ptr) return;
    // Queued work may still read the buffer: finish it first (cuda_device.cu:21-23).
    check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    check(hipFree(ptr), "hipFree");
    d->allocations.erase(reinterpret_cast<std::uintptr_t>(ptr));
  });
}

int kh_upload(kh_device* d, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    require(d != nullptr, "kh_upload: null device");
    no_capture(d, "kh_upload");
    if (bytes == 0) return;
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d->stream), "upload");
    check(hipStreamSynchronize(d->stream), "upload wait");
  });
}

int kh_download(kh_device* d, const void* src, void* dst, size_t bytes) {
  return guard([&] {
    require(d != nullptr, "kh_download: null device");
    no_capture(d, "kh_download");
    if (bytes == 0) return;
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, d->stream), "download");
    check(hipStreamSynchronize(d->stream), "download wait");
  });
}

int kh_copy(kh_device* d, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    require(d != nullptr, "kh_copy: null device");
    if (bytes == 0 || dst == src) return;
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, d->stream), "copy");
  });
}

int kh_fill32(kh_device* d, void* dst, uint32_t word, size_t count) {
  return guard([&] {
    require(d != nullptr, "kh_fill32: null device");
    launch_fill32(d, dst, word, count);
    check(hipGetLastError(), "fill32 launch");
  });
}

int kh_synchronize(kh_device* d) {
  return guard([&] {
    require(d != nullptr, "kh_synchronize: null device");
    no_capture(d, "kh_synchronize");
    check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  });
}

const char* kh_kernel_name(uint32_t kernel) {
  const auto* e = find_kernel(kernel);
  return e ? e->name : nullptr;
}

int kh_launch(kh_device* d, uint32_t kernel, size_t n, const kh_arg* args, size_t nargs) {
  return guard([&] {
    require(d != nullptr, "kh_launch: null device");
    const auto* e = find_kernel(kernel);
    require(e != nullptr, "kh_launch: bad kernel id");
    require(n <= size_t(INT_MAX), "kh_launch: n exceeds INT_MAX");
    const std::size_t expected = std::strlen(e->signature);
    if (nargs != expected)
      throw std::runtime_error(std::string("kh_launch: ") + e->name + " expects " +
                               std::to_string(expected) + " arguments");
    for (std::size_t k = 0; k < nargs; ++k) {
      const bool is_buffer = e->signature[k] == 'b';
      if (is_buffer ? args[k].kind != KH_ARG_BUFFER
                    : (args[k].kind != KH_ARG_SCALAR || args[k].size != 4))
        throw std::runtime_error(std::string("kh_launch: ") + e->name + " argument " +
                                 std::to_string(k) + " has the wrong kind");
      if (is_buffer && args[k].value.ptr == nullptr)
        throw std::runtime_error(std::string("kh_launch: ") + e->name + " argument " +
                                 std::to_string(k) + " is a null buffer");
    }
    if (n == 0) return;
    if (d->count_bytes && !d->capturing) {
      double bytes = 0;
      for (std::size_t k = 0; k < nargs; ++k)
        if (args[k].kind == KH_ARG_BUFFER) bytes += double(allocation_bytes(d, args[k].value.ptr));
      count_logical(d, bytes);
    }
    auto& slot = d->profile[kernel < d->profile.size() ? kernel : 0];
    const bool timed = kernel < d->profile.size() && slot.enabled && !d->capturing &&
                       (slot.only_n == 0 || slot.only_n == n);
    hipEvent_t start = nullptr, stop = nullptr;
    if (timed) {
      start = take_event(d);
      stop = take_event(d);
      if (e->packet_timed) {
        d->packet_start = start;
        d->packet_stop = stop;
      } else {
        check(hipEventRecord(start, d->stream), "hipEventRecord");
      }
    }
    e->launch(d, int(n), args);
    d->packet_start = d->packet_stop = nullptr;
    check(hipGetLastError(), e->name);
    if (timed) {
      if (!e->packet_timed) check(hipEventRecord(stop, d->stream), "hipEventRecord");
      slot.pending.emplace_back(start, stop);
    }
  });
}

int kh_time_launches(kh_device* d, uint32_t kernel, size_t n, const kh_arg* args, size_t nargs, int repeats,
                     double* total_ms) {
  return guard([&] {
    require(d != nullptr && total_ms != nullptr, "kh_time_launches: null argument");
    require(repeats >= 1, "kh_time_launches: repeats must be positive");
    no_capture(d, "kh_time_launches");
    hipEvent_t start = nullptr, stop = nullptr;
    check(hipEventCreate(&start), "hipEventCreate");
    check(hipEventCreate(&stop), "hipEventCreate");
    int status = 0;
    try {
      check(hipEventRecord(start, d->stream), "hipEventRecord");
      for (int k = 0; k < repeats && status == 0; ++k) status = kh_launch(d, kernel, n, args, nargs);
      check(hipEventRecord(stop, d->stream), "hipEventRecord");
      check(hipEventSynchronize(stop), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
      *total_ms = ms;
    } catch (...) {
      (void)hipEventDestroy(start);
      (void)hipEventDestroy(stop);
      throw;
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (status != 0) throw std::runtime_error(kh_last_error());
  });
}

int kh_reduce_host(kh_device* d, const kh_reduction* req, size_t count, double* out) {
  return guard([&] {
    require(d != nullptr, "kh_reduce_host: null device");
    no_capture(d, "kh_reduce_host");
    for (size_t base = 0; base < count; base += kMaxHostResults) {
      const size_t chunk = count - base < size_t(kMaxHostResults) ? count - base : size_t(kMaxHostResults);
      for (size_t k = 0; k < chunk; ++k) {
        const auto& r = req[base + k];
        require(r.kind <= KH_REDUCE_MAX_ABS, "kh_reduce_host: bad kind");
        count_logical(d, double(r.n) * (r.kind == KH_REDUCE_DOT ? 8 : 4));
        if (r.n == 0) {
          launch_fill32(d, d->results + k, 0u, 2);  // +0.0 as a double
        } else {
          launch_reduce_host(d, int(r.kind), r.x, r.kind == KH_REDUCE_DOT ? r.y : r.x, r.n,
                             d->results + k);
        }
        check(hipGetLastError(), "reduce launch");
      }
      check(hipMemcpyAsync(d->results_host, d->results, chunk * sizeof(double),
                           hipMemcpyDeviceToHost, d->stream), "reduce download");
      check(hipStreamSynchronize(d->stream), "reduce wait");
      std::memcpy(out + base, d->results_host, chunk * sizeof(double));
    }
  });
}

int kh_reduce_into(kh_device* d, int mode, const float* x, const float* y, const float* u,
                   const float* v, size_t n, float* dst, size_t offset) {
  return guard([&] {
    require(d != nullptr && dst != nullptr, "kh_reduce_into: null argument");
    require(mode >= KH_INTO_SUM && mode <= KH_INTO_DOT2, "kh_reduce_into: bad mode");
    require(dst != x && dst != y && dst != u && dst != v, "kh_reduce_into: output aliases input");
    float* out = dst + offset;
    count_logical(d, double(n) * (mode == KH_INTO_SUM ? 4 : mode == KH_INTO_DOT ? 8 : 16));
    if (n == 0) {
      launch_fill32(d, out, 0u, mode == KH_INTO_DOT2 ? 4 : 2);
    } else {
      launch_reduce_into(d, mode, x, mode == KH_INTO_SUM ? x : y, mode == KH_INTO_DOT2 ? u : x,
                         mode == KH_INTO_DOT2 ? v : x, n, out);
    }
    check(hipGetLastError(), "reduce_into launch");
  });
}

int kh_reduce_into_static(void* stream, int mode, const float* x, const float* y, const float* u,
                          const float* v, size_t n, float* dst, size_t offset) {
  return guard([&] {
    require(dst != nullptr, "kh_reduce_into_static: null output");
    require(mode >= KH_INTO_SUM && mode <= KH_INTO_DOT2, "kh_reduce_into_static: bad mode");
    require(n <= size_t(KH_STATIC_REDUCE_MAX_N), "kh_reduce_into_static: n exceeds KH_STATIC_REDUCE_MAX_N");
    require(n == 0 || (x != nullptr && (mode == KH_INTO_SUM || y != nullptr) &&
                       (mode != KH_INTO_DOT2 || (u != nullptr && v != nullptr))),
            "kh_reduce_into_static: null input");
    require(dst != x && dst != y && dst != u && dst != v, "kh_reduce_into_static: output aliases input");
    const auto s = static_cast<hipStream_t>(stream);
    float* out = dst + offset;
    if (n == 0) {
      fill32_on(s, out, 0u, mode == KH_INTO_DOT2 ? 4 : 2);
    } else {
      double* partials = device_scratch();
      const float* yy = mode == KH_INTO_SUM ? x : y;
      const float* uu = mode == KH_INTO_DOT2 ? u : x;
      const float* vv = mode == KH_INTO_DOT2 ? v : x;
      if (mode == KH_INTO_SUM) reduce_on<kSum>(s, partials, x, yy, uu, vv, n, out, nullptr);
      else if (mode == KH_INTO_DOT) reduce_on<kDot>(s, partials, x, yy, uu, vv, n, out, nullptr);
      else reduce_on<kDot2>(s, partials, x, yy, uu, vv, n, out, nullptr);
    }
    check(hipGetLastError(), "reduce_into_static launch");
  });
}

int kh_graph_begin(kh_device* d) {
  return guard([&] {
    require(d != nullptr, "kh_graph_begin: null device");
    require(!d->capturing, "kh_graph_begin: capture already active");
    check(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    d->capturing = true;
  });
}

int kh_graph_end(kh_device* d, kh_graph** out) {
  return guard([&] {
    require(d != nullptr && out != nullptr, "kh_graph_end: null argument");
    require(d->capturing, "kh_graph_end: no capture active");
    d->capturing = false;
    auto* g = new kh_graph;
    hipError_t st = hipStreamEndCapture(d->stream, &g->graph);
    if (st != hipSuccess) {
      delete g;
      check(st, "hipStreamEndCapture");
    }
    st = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (st != hipSuccess) {
      (void)hipGraphDestroy(g->graph);
      delete g;
      check(st, "hipGraphInstantiate");
    }
    *out = g;
  });
}

int kh_graph_launch(kh_device* d, kh_graph* g) {
  return guard([&] {
    require(d != nullptr && g != nullptr, "kh_graph_launch: null argument");
    no_capture(d, "kh_graph_launch");
    check(hipGraphLaunch(g->exec, d->stream), "hipGraphLaunch");
  });
}

void kh_graph_destroy(kh_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

int kh_profile_enable(kh_device* d, uint32_t kernel, int enable) {
  return guard([&] {
    require(d != nullptr, "kh_profile_enable: null device");
    require(kernel < d->profile.size() && find_kernel(kernel), "kh_profile_enable: bad kernel id");
    d->profile[kernel].enabled = enable != 0;
  });
}

int kh_profile_filter(kh_device* d, uint32_t kernel, size_t n) {
  return guard([&] {
    require(d != nullptr, "kh_profile_filter: null device");
    require(kernel < d->profile.size() && find_kernel(kernel), "kh_profile_filter: bad kernel id");
    d->profile[kernel].only_n = n;
  });
}

int kh_profile_read(kh_device* d, uint32_t kernel, long* launches, double* total_ms) {
  return guard([&] {
    require(d != nullptr, "kh_profile_read: null device");
    require(kernel < d->profile.size(), "kh_profile_read: bad kernel id");
    no_capture(d, "kh_profile_read");
    check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    auto& slot = d->profile[kernel];
    for (auto& p : slot.pending) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, p.first, p.second), "hipEventElapsedTime");
      slot.total_ms += ms;
      ++slot.launches;
      d->event_pool.push_back(p.first);
      d->event_pool.push_back(p.second);
    }
    slot.pending.clear();
    if (launches) *launches = slot.launches;
    if (total_ms) *total_ms = slot.total_ms;
  });
}

int kh_profile_reset(kh_device* d) {
  return guard([&] {
    require(d != nullptr, "kh_profile_reset: null device");
    check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    for (auto& slot : d->profile) {
      for (auto& p : slot.pending) {
        d->event_pool.push_back(p.first);
        d->event_pool.push_back(p.second);
      }
      slot.pending.clear();
      slot.launches = 0;
      slot.total_ms = 0;
    }
  });
}

}  // extern "C"
