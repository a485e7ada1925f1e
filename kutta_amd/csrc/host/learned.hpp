// Learned coarse correction (no reference counterpart): the host side of the
// message-passing network in kh_learned.hip. The network lives on one coarse
// AMG level's graph; inference maps a restricted residual r to a correction
// e of the same length, and training fits e to exact coarse corrections with
// Adam, all on the device.
#pragma once

#include <cstdint>
#include <vector>

#include "krylov.hpp"

namespace kutta_amd {

// Parameters of a network with `layers` message-passing layers (hidden width
// KH_LEARNED_HIDDEN, 4 input features).
std::size_t learned_parameter_count(int layers);

// Per training step: the mean squared error before the step, and that error
// over the mean squared normalised target (the unexplained fraction).
struct TrainResult {
  std::vector<double> loss, relative;
};

class LearnedNetwork {
 public:
  LearnedNetwork(HipDevice& dev, int layers, std::uint64_t seed = 0);
  int layers() const { return layers_; }
  std::size_t parameter_count() const { return count_; }
  Index nodes() const { return graph_.n_rows; }

  // The level operator (symmetric, positive diagonal; its values may change
  // between calls, its pattern may not once buffers are sized): node features
  // and neighbour weights. Must be called before the first use and after
  // every change of the operator's values.
  void set_graph(const DeviceCsr& a);

  // Parameters in the layout of the kernels (W0, b0, per layer Ws, Wn, b,
  // then wo, bo); setting them clears the Adam state.
  void set_parameters(const std::vector<float>& theta);
  std::vector<float> parameters() const;

  // e = network correction of the residual r (device vectors of nodes()).
  void infer(const float* r, float* e);
  // Loss and gradient at the current parameters for (r, target);
  // target_energy() is then the mean squared normalised target.
  double gradient(const float* r, const float* target);
  double target_energy() const { return energy_; }
  std::vector<float> gradient_host() const;
  // Normalised network output y of the last gradient() or train() step.
  std::vector<float> outputs() const;
  // `steps` Adam steps cycling over `pairs` (r, target) pairs stored
  // contiguously, nodes() floats each.
  TrainResult train(const float* r, const float* target, int pairs, int steps, float lr);

 private:
  void size_buffers(Index n);
  void forward(const float* r);
  void backward(const float* target, int mode, float lr, double* loss);
  std::size_t head_offset() const;
  std::size_t layer_offset(int l) const;

  HipDevice& dev_;
  int layers_;
  std::size_t count_;
  int step_ = 0;
  double energy_ = 0;
  DeviceCsr graph_;
  DeviceArray<float> features_, ws_, wt_, gstats_, stats_, x_, y_, dy_, p_[2], dm_[2];
  DeviceArray<float> theta_, adam_m_, adam_v_, grad_, partials_;
  DeviceArray<double> scratch_, loss_;
  std::vector<DeviceArray<std::uint16_t>> h_, m_;
  DeviceArray<std::uint16_t> dz_, wf_, wb_;
};

}  // namespace kutta_amd
