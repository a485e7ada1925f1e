// C++ view of the kutta_hip C ABI for host code: the device, typed device
// buffers and a variadic launch helper. This is the shape of the HipDevice
// adapter a Kutta build places behind kutta::Device (runtime/device.hpp:107-240);
// errors surface as exceptions carrying the ABI's message.
#pragma once

#include <cstdint>
#include <cstring>
#include <unistd.h>  // environ
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "kutta_hip.h"

namespace kutta_amd {

using Index = std::int32_t;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NonfiniteError : Error {
  using Error::Error;
};

#define KA_REQUIRE(cond, msg)                                                        \
  do {                                                                               \
    if (!(cond)) throw ::kutta_amd::Error(std::string(__FILE__).substr(              \
        std::string(__FILE__).find_last_of('/') + 1) + ":" + std::to_string(__LINE__) + \
        ": " + (msg));                                                               \
  } while (0)

inline void ok(int status) {
  if (status != KH_OK) throw Error(kh_last_error());
}

class HipDevice;

// Move-only typed device array (device.hpp:51-70), extended to double and
// 16-bit storage, which the backend keeps private.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  DeviceArray(HipDevice& dev, std::size_t n);
  DeviceArray(DeviceArray&& o) noexcept : dev_(o.dev_), ptr_(o.ptr_), n_(o.n_) {
    o.dev_ = nullptr; o.ptr_ = nullptr; o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      dev_ = o.dev_; ptr_ = o.ptr_; n_ = o.n_;
      o.dev_ = nullptr; o.ptr_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() { release(); }

  T* data() const { return ptr_; }
  std::size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  void upload(const std::vector<T>& host);
  std::vector<T> download() const;

 private:
  void release();
  HipDevice* dev_ = nullptr;
  T* ptr_ = nullptr;
  std::size_t n_ = 0;
};

// Whether a ROCm profiler (rocprofv3, which exports ROCPROF*/ROCPROFILER*
// variables into the traced process) is attached. A first graph replay under
// its kernel tracing ended a process with a host SIGSEGV inside
// hipGraphLaunch, so graph replay is switched off while one is attached:
// profiled runs launch the same kernels one by one (identical iterates).
inline bool profiler_attached() {
  for (char** e = ::environ; e && *e; ++e)
    if (std::strncmp(*e, "ROCPROF", 7) == 0) return true;
  return false;
}

// One argument for kh_launch.
struct Arg {
  kh_arg a{};
  template <class T>
  Arg(T* p) { a.kind = KH_ARG_BUFFER; a.size = 0; a.value.ptr = const_cast<void*>(static_cast<const void*>(p)); }
  template <class T>
  Arg(const DeviceArray<T>& b) { a.kind = KH_ARG_BUFFER; a.size = 0; a.value.ptr = b.data(); }
  Arg(float v) { a.kind = KH_ARG_SCALAR; a.size = 4; a.value.f32 = v; }
  Arg(int v) { a.kind = KH_ARG_SCALAR; a.size = 4; a.value.i32 = v; }
};

class HipDevice {
 public:
  explicit HipDevice(int ordinal = 0) { ok(kh_create(ordinal, &dev_)); }
  ~HipDevice() { kh_destroy(dev_); }
  HipDevice(const HipDevice&) = delete;
  HipDevice& operator=(const HipDevice&) = delete;

  kh_device* raw() const { return dev_; }
  std::string name() const { return kh_name(dev_); }

  void launch(uint32_t kernel, std::size_t n, std::initializer_list<Arg> args) {
    std::vector<kh_arg> raw;
    raw.reserve(args.size());
    for (const auto& a : args) raw.push_back(a.a);
    ok(kh_launch(dev_, kernel, n, raw.data(), raw.size()));
  }
  void fill(void* p, float v, std::size_t count) {
    uint32_t w;
    static_assert(sizeof w == sizeof v);
    std::memcpy(&w, &v, 4);
    ok(kh_fill32(dev_, p, w, count));
  }
  void fill_int(void* p, int32_t v, std::size_t count) { ok(kh_fill32(dev_, p, uint32_t(v), count)); }
  void copy(void* dst, const void* src, std::size_t bytes) { ok(kh_copy(dev_, dst, src, bytes)); }
  void synchronize() { ok(kh_synchronize(dev_)); }
  // Host reductions behind one wait (device.hpp:162-184).
  std::vector<double> reduce_many(const std::vector<kh_reduction>& requests) {
    std::vector<double> out(requests.size());
    if (!requests.empty()) ok(kh_reduce_host(dev_, requests.data(), requests.size(), out.data()));
    return out;
  }
  static kh_reduction sum(const float* x, std::size_t n) { return {KH_REDUCE_SUM, 0, x, x, n}; }
  static kh_reduction dot(const float* x, const float* y, std::size_t n) { return {KH_REDUCE_DOT, 0, x, y, n}; }
  static kh_reduction max_abs(const float* x, std::size_t n) { return {KH_REDUCE_MAX_ABS, 0, x, x, n}; }
  void reduce_into(int mode, const float* x, const float* y, const float* u, const float* v,
                   std::size_t n, float* dst, std::size_t offset = 0) {
    ok(kh_reduce_into(dev_, mode, x, y, u, v, n, dst, offset));
  }

  // Records everything enqueued by `body` into a replayable graph.
  template <class F>
  kh_graph* capture(F&& body) {
    ok(kh_graph_begin(dev_));
    try {
      body();
    } catch (...) {
      kh_graph* g = nullptr;
      if (kh_graph_end(dev_, &g) == KH_OK) kh_graph_destroy(g);
      throw;
    }
    kh_graph* g = nullptr;
    ok(kh_graph_end(dev_, &g));
    return g;
  }
  void replay(kh_graph* g) { ok(kh_graph_launch(dev_, g)); }

 private:
  kh_device* dev_ = nullptr;
};

// Owning handle of a captured graph.
class Graph {
 public:
  Graph() = default;
  explicit Graph(kh_graph* g) : g_(g) {}
  Graph(Graph&& o) noexcept : g_(o.g_) { o.g_ = nullptr; }
  Graph& operator=(Graph&& o) noexcept {
    if (this != &o) {
      reset();
      g_ = o.g_;
      o.g_ = nullptr;
    }
    return *this;
  }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  ~Graph() { reset(); }
  kh_graph* get() const { return g_; }
  explicit operator bool() const { return g_ != nullptr; }
  void reset() {
    if (g_) kh_graph_destroy(g_);
    g_ = nullptr;
  }

 private:
  kh_graph* g_ = nullptr;
};

template <class T>
DeviceArray<T>::DeviceArray(HipDevice& dev, std::size_t n) : dev_(&dev), n_(n) {
  void* p = nullptr;
  ok(kh_alloc(dev.raw(), n * sizeof(T), &p));
  ptr_ = static_cast<T*>(p);
}
template <class T>
void DeviceArray<T>::release() {
  if (dev_ && ptr_) kh_free(dev_->raw(), ptr_);
  ptr_ = nullptr;
}
template <class T>
void DeviceArray<T>::upload(const std::vector<T>& host) {
  KA_REQUIRE(host.size() == n_, "upload size mismatch");
  ok(kh_upload(dev_->raw(), ptr_, host.data(), n_ * sizeof(T)));
}
template <class T>
std::vector<T> DeviceArray<T>::download() const {
  std::vector<T> host(n_);
  ok(kh_download(dev_->raw(), ptr_, host.data(), n_ * sizeof(T)));
  return host;
}

template <class T>
DeviceArray<T> to_device(HipDevice& dev, const std::vector<T>& host) {
  DeviceArray<T> d(dev, host.size());
  d.upload(host);
  return d;
}

}  // namespace kutta_amd
