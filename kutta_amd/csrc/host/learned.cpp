// Host orchestration of the learned coarse correction's kernels
// (kh_learned.hip): buffer sizing, the forward pass, backpropagation into
// per-slot partials and the fused reduction + Adam + bf16 packing launch.
#include "learned.hpp"

#include <cmath>

namespace kutta_amd {

namespace {
constexpr int kH = KH_LEARNED_HIDDEN;
constexpr int kF = 4;
constexpr std::size_t kInput = kF * kH + kH;
constexpr std::size_t kLayer = 2 * kH * kH + kH;
constexpr std::size_t kPacked = 2 * kH * kH;  // bf16 entries of [Ws; Wn]

// splitmix64 stream and Box-Muller normals for the initial weights.
struct Normal {
  std::uint64_t state;
  std::uint64_t next() {
    std::uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double operator()() {
    const double u = uniform(), v = uniform();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
  }
};
}  // namespace

std::size_t learned_parameter_count(int layers) {
  return kInput + std::size_t(layers) * kLayer + kH + 1;
}

LearnedNetwork::LearnedNetwork(HipDevice& dev, int layers, std::uint64_t seed)
    : dev_(dev), layers_(layers), count_(learned_parameter_count(layers)) {
  KA_REQUIRE(layers >= 1 && layers <= KH_LEARNED_MAX_LAYERS,
             "learned: layer count must be in [1, " + std::to_string(KH_LEARNED_MAX_LAYERS) + "]");
  theta_ = DeviceArray<float>(dev, count_);
  adam_m_ = DeviceArray<float>(dev, count_);
  adam_v_ = DeviceArray<float>(dev, count_);
  grad_ = DeviceArray<float>(dev, count_);
  partials_ = DeviceArray<float>(dev, std::size_t(KH_LEARNED_SLOTS) * (count_ + 2));
  scratch_ = DeviceArray<double>(dev, KH_LEARNED_SLOTS);
  loss_ = DeviceArray<double>(dev, 2);
  gstats_ = DeviceArray<float>(dev, 1);
  stats_ = DeviceArray<float>(dev, 2);
  wf_ = DeviceArray<std::uint16_t>(dev, std::size_t(layers) * kPacked);
  wb_ = DeviceArray<std::uint16_t>(dev, std::size_t(layers) * kPacked);
  // He-style initial weights (variance 2 / fan-in for the input layer,
  // 1 / fan-in elsewhere), zero biases.
  Normal normal{seed};
  std::vector<float> theta(count_, 0.0f);
  std::size_t k = 0;
  for (int i = 0; i < kF * kH; ++i) theta[k++] = float(normal() * std::sqrt(2.0 / kF));
  k += kH;
  for (int l = 0; l < layers; ++l) {
    for (int i = 0; i < 2 * kH * kH; ++i) theta[k++] = float(normal() * std::sqrt(1.0 / kH));
    k += kH;
  }
  for (int i = 0; i < kH; ++i) theta[k++] = float(normal() * std::sqrt(1.0 / kH));
  set_parameters(theta);
}

std::size_t LearnedNetwork::layer_offset(int l) const { return kInput + std::size_t(l) * kLayer; }
std::size_t LearnedNetwork::head_offset() const { return kInput + std::size_t(layers_) * kLayer; }

void LearnedNetwork::size_buffers(Index n) {
  if (features_.size() == 2 * std::size_t(n)) return;
  const std::size_t nh = std::size_t(n) * kH;
  features_ = DeviceArray<float>(dev_, 2 * std::size_t(n));
  x_ = DeviceArray<float>(dev_, 4 * std::size_t(n));
  y_ = DeviceArray<float>(dev_, n);
  dy_ = DeviceArray<float>(dev_, n);
  for (auto& b : p_) b = DeviceArray<float>(dev_, nh);
  for (auto& b : dm_) b = DeviceArray<float>(dev_, nh);
  h_.clear();
  m_.clear();
  for (int l = 0; l <= layers_; ++l) h_.emplace_back(dev_, nh);
  for (int l = 0; l < layers_; ++l) m_.emplace_back(dev_, nh);
  dz_ = DeviceArray<std::uint16_t>(dev_, nh);
}

void LearnedNetwork::set_graph(const DeviceCsr& a) {
  KA_REQUIRE(a.n_rows >= 1 && std::int64_t(a.n_rows) * kH < (std::int64_t(1) << 31),
             "learned: the graph must have between 1 and 2^25 nodes");
  if (ws_.size() != std::size_t(a.nnz) || graph_.n_rows != a.n_rows) {
    ws_ = DeviceArray<float>(dev_, a.nnz);
    wt_ = DeviceArray<float>(dev_, a.nnz);
  }
  size_buffers(a.n_rows);
  graph_ = a;
  dev_.launch(KH_LEARNED_GRAPH, a.n_rows,
              {a.row_ptr, a.col, a.val, features_, ws_, wt_, gstats_, scratch_});
}

void LearnedNetwork::set_parameters(const std::vector<float>& theta) {
  KA_REQUIRE(theta.size() == count_, "learned: expected " + std::to_string(count_) + " parameters");
  for (float v : theta) KA_REQUIRE(std::isfinite(v), "learned: non-finite parameter");
  theta_.upload(theta);
  dev_.fill(adam_m_.data(), 0.0f, count_);
  dev_.fill(adam_v_.data(), 0.0f, count_);
  step_ = 0;
  dev_.launch(KH_LEARNED_UPDATE, count_ + 2,
              {0, partials_, theta_, adam_m_, adam_v_, grad_, loss_, layers_, 0, 0.0f, 1, wf_, wb_});
}

std::vector<float> LearnedNetwork::parameters() const { return theta_.download(); }

void LearnedNetwork::forward(const float* r) {
  KA_REQUIRE(graph_.n_rows > 0, "learned: set_graph before use");
  const Index n = graph_.n_rows;
  dev_.launch(KH_LEARNED_NORM, n, {r, scratch_});
  dev_.launch(KH_LEARNED_INPUT, n, {r, scratch_, features_, gstats_, theta_, x_, h_[0], stats_});
  for (int l = 0; l < layers_; ++l)
    dev_.launch(KH_LEARNED_LAYER, n,
                {graph_.row_ptr, graph_.col, ws_, h_[l], wf_.data() + l * kPacked,
                 theta_.data() + layer_offset(l) + 2 * kH * kH, h_[l + 1], m_[l]});
}

void LearnedNetwork::infer(const float* r, float* e) {
  forward(r);
  dev_.launch(KH_LEARNED_HEAD, graph_.n_rows, {h_[layers_], theta_.data() + head_offset(), stats_, e});
}

// Head, loss and backpropagation into the slot partials, then the update
// launch in `mode` (1: gradient only, 2: Adam step); loss[0..1] receive the
// mean squared error and the mean squared normalised target.
void LearnedNetwork::backward(const float* target, int mode, float lr, double* loss) {
  const Index n = graph_.n_rows;
  const int width = int(count_ + 2);
  float* head = theta_.data() + head_offset();
  dev_.launch(KH_LEARNED_HEAD_LOSS, n,
              {h_[layers_], head, stats_, target, y_, dy_, partials_.data() + head_offset(), width});
  int cur = 0;
  for (int l = layers_ - 1; l >= 0; --l) {
    const std::size_t base = layer_offset(l);
    dev_.launch(KH_LEARNED_BACK, n,
                {l == layers_ - 1 ? 1 : 0, dy_, head, p_[cur], dm_[cur], graph_.row_ptr, graph_.col, wt_,
                 h_[l + 1], wb_.data() + l * kPacked, dz_, p_[1 - cur], dm_[1 - cur],
                 partials_.data() + base + 2 * kH * kH, width});
    dev_.launch(KH_LEARNED_WEIGHT_GRAD, n, {h_[l], m_[l], dz_, partials_.data() + base, width});
    cur = 1 - cur;
  }
  dev_.launch(KH_LEARNED_BACK_INPUT, n,
              {p_[cur], dm_[cur], graph_.row_ptr, graph_.col, wt_, h_[0], x_, partials_, width});
  if (mode == 2) ++step_;
  dev_.launch(KH_LEARNED_UPDATE, count_ + 2,
              {mode, partials_, theta_, adam_m_, adam_v_, grad_, loss, layers_, step_, lr, int(n), wf_, wb_});
}

double LearnedNetwork::gradient(const float* r, const float* target) {
  forward(r);
  backward(target, 1, 0.0f, loss_.data());
  const auto l = loss_.download();
  energy_ = l[1];
  return l[0];
}

std::vector<float> LearnedNetwork::gradient_host() const { return grad_.download(); }
std::vector<float> LearnedNetwork::outputs() const { return y_.download(); }

TrainResult LearnedNetwork::train(const float* r, const float* target, int pairs, int steps, float lr) {
  KA_REQUIRE(pairs >= 1 && steps >= 0, "learned: train needs at least one pair");
  KA_REQUIRE(std::isfinite(lr) && lr > 0, "learned: learning rate must be positive");
  TrainResult out;
  if (steps == 0) return out;
  const std::size_t n = std::size_t(graph_.n_rows);
  DeviceArray<double> sums(dev_, 2 * std::size_t(steps));
  for (int s = 0; s < steps; ++s) {
    const std::size_t k = std::size_t(s % pairs) * n;
    forward(r + k);
    backward(target + k, 2, lr, sums.data() + 2 * s);
  }
  const auto host = sums.download();
  for (int s = 0; s < steps; ++s) {
    out.loss.push_back(host[2 * s]);
    out.relative.push_back(host[2 * s + 1] > 0 ? host[2 * s] / host[2 * s + 1] : 0.0);
  }
  return out;
}

}  // namespace kutta_amd
