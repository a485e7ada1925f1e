// C ABI of the learned coarse correction network (include/kutta_amd.h).
#include <memory>

#include "capi_internal.hpp"
#include "kutta_amd.h"
#include "learned.hpp"

using namespace kutta_amd;
using capi::guard;

namespace {
LearnedNetwork& net(void* h) { return *static_cast<LearnedNetwork*>(h); }
}  // namespace

extern "C" {

int64_t ka_learned_parameter_count(int layers) {
  if (layers < 1 || layers > KH_LEARNED_MAX_LAYERS) return -1;
  return int64_t(learned_parameter_count(layers));
}

int ka_learned_create(void* dev, int layers, uint64_t seed, void** out) {
  return guard([&] { *out = new LearnedNetwork(*static_cast<HipDevice*>(dev), layers, seed); });
}
void ka_learned_destroy(void* h) { delete static_cast<LearnedNetwork*>(h); }

int ka_learned_set_graph(void* h, void* matrix) {
  return guard([&] { net(h).set_graph(static_cast<capi::MatrixHandle*>(matrix)->view()); });
}

int ka_learned_set_parameters(void* h, const float* theta, int64_t count) {
  return guard([&] {
    KA_REQUIRE(count == int64_t(net(h).parameter_count()), "learned: parameter count mismatch");
    net(h).set_parameters(std::vector<float>(theta, theta + count));
  });
}

int ka_learned_parameters(void* h, float* theta, int64_t count) {
  return guard([&] {
    KA_REQUIRE(count == int64_t(net(h).parameter_count()), "learned: parameter count mismatch");
    const auto v = net(h).parameters();
    std::copy(v.begin(), v.end(), theta);
  });
}

int ka_learned_infer(void* h, const float* r, float* e) {
  return guard([&] { net(h).infer(r, e); });
}

int ka_learned_gradient(void* h, const float* r, const float* target, float* grad, float* outputs,
                        double* loss) {
  return guard([&] {
    const double l = net(h).gradient(r, target);
    if (loss) {
      loss[0] = l;
      loss[1] = net(h).target_energy();
    }
    if (grad) {
      const auto g = net(h).gradient_host();
      std::copy(g.begin(), g.end(), grad);
    }
    if (outputs) {
      const auto y = net(h).outputs();
      std::copy(y.begin(), y.end(), outputs);
    }
  });
}

int ka_learned_train(void* h, const float* r, const float* target, int pairs, int steps, float lr,
                     double* losses, double* relative) {
  return guard([&] {
    const auto l = net(h).train(r, target, pairs, steps, lr);
    if (losses) std::copy(l.loss.begin(), l.loss.end(), losses);
    if (relative) std::copy(l.relative.begin(), l.relative.end(), relative);
  });
}

}  // extern "C"
