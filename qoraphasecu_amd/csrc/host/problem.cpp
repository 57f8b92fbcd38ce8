// Problem construction: make_problem, make_problems, autocorrelation_support,
// simulate_intensity, half_support.  Contract (argument checks and their messages, the
// amplitude, the support rule): src/core/problem.cpp:15-127.  The transforms, and the
// roll, |.|, maximum and cut around them, run on the GPU through the qpch C ABI; the
// autocorrelation supports of several problems of one shape take one device pass.
#include "qoraphasecu/problem.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>

#include "abi.hpp"

namespace qpcu {

std::size_t numel(const Shape& shape) {
  return std::accumulate(shape.begin(), shape.end(), std::size_t{1}, std::multiplies<>{});
}

namespace {

std::string count_mismatch(const char* what, std::size_t got, std::size_t want) {
  return std::string(what) + " has " + std::to_string(got) + " elements, shape has " + std::to_string(want);
}

// The reference's argument checks, in its order, with its messages.
std::size_t checked_size(const Shape& shape) {
  if (shape.size() != 2 && shape.size() != 3) {
    throw std::invalid_argument("shape must be 2D or 3D, got ndim=" + std::to_string(shape.size()));
  }
  if (std::find(shape.begin(), shape.end(), std::size_t{0}) != shape.end()) {
    throw std::invalid_argument("shape has a zero-length axis");
  }
  return numel(shape);
}

void check_input(const ProblemInput& in, std::size_t n) {
  if (in.intensity.size() != n) throw std::invalid_argument(count_mismatch("intensity", in.intensity.size(), n));
  if (!in.support.empty() && in.support.size() != n) {
    throw std::invalid_argument(count_mismatch("support", in.support.size(), n));
  }
  if (!in.mask.empty() && in.mask.size() != n) throw std::invalid_argument(count_mismatch("mask", in.mask.size(), n));
  for (float v : in.intensity) {
    if (!std::isfinite(v)) throw std::invalid_argument("intensity contains NaN or Inf");
  }
  if (!in.mask.empty() && std::none_of(in.mask.begin(), in.mask.end(), [](std::uint8_t m) { return m == 0; })) {
    throw std::invalid_argument("mask marks every pixel as bad; nothing left to fit");
  }
}

std::vector<std::uint8_t> as_flags(std::span<const std::uint8_t> v, std::size_t n) {
  std::vector<std::uint8_t> out(n, 0);
  for (std::size_t i = 0; i < v.size(); ++i) out[i] = v[i] != 0;
  return out;
}

// Autocorrelation supports of `count` same-shape intensities laid out back to back.
std::vector<std::uint8_t> device_supports(const float* intensities, int count, const Shape& shape, float threshold) {
  std::vector<std::uint8_t> out(numel(shape) * static_cast<std::size_t>(count));
  detail::qpch_check(qpch_autocorrelation_support(intensities, count, static_cast<int>(shape.size()), shape.data(),
                                                  threshold, out.data()));
  return out;
}

}  // namespace

std::vector<Problem> make_problems(std::span<const ProblemInput> inputs, float support_threshold) {
  std::vector<Problem> problems(inputs.size());
  std::map<Shape, std::vector<std::size_t>> need_support;  // shape -> problems without a given support
  for (std::size_t k = 0; k < inputs.size(); ++k) {
    const ProblemInput& in = inputs[k];
    const std::size_t n = checked_size(in.shape);
    check_input(in, n);
    Problem& p = problems[k];
    p.shape = in.shape;
    p.amplitude.resize(n);
    std::transform(in.intensity.begin(), in.intensity.end(), p.amplitude.begin(),
                   [](float v) { return std::sqrt(std::max(v, 0.0f)); });
    p.mask = as_flags(in.mask, n);
    if (in.support.empty()) {
      need_support[in.shape].push_back(k);
    } else {
      p.support = as_flags(in.support, n);
    }
  }
  constexpr std::size_t kMaxPerPass = 65535;  // problems per device pass (its grid's y extent)
  for (const auto& [shape, members] : need_support) {
    const std::size_t n = numel(shape);
    for (std::size_t first = 0; first < members.size(); first += kMaxPerPass) {
      const std::size_t count = std::min(kMaxPerPass, members.size() - first);
      std::vector<float> packed(n * count);
      for (std::size_t j = 0; j < count; ++j) {
        const auto& src = inputs[members[first + j]].intensity;
        std::copy(src.begin(), src.end(), packed.begin() + static_cast<std::ptrdiff_t>(j * n));
      }
      const auto flags = device_supports(packed.data(), static_cast<int>(count), shape, support_threshold);
      for (std::size_t j = 0; j < count; ++j) {
        const auto from = flags.begin() + static_cast<std::ptrdiff_t>(j * n);
        problems[members[first + j]].support.assign(from, from + static_cast<std::ptrdiff_t>(n));
      }
    }
  }
  return problems;
}

Problem make_problem(std::span<const float> intensity, Shape shape, std::span<const std::uint8_t> support,
                     std::span<const std::uint8_t> mask, float support_threshold) {
  const ProblemInput in{intensity, std::move(shape), support, mask};
  return std::move(make_problems(std::span<const ProblemInput>(&in, 1), support_threshold).front());
}

std::vector<std::uint8_t> autocorrelation_support(std::span<const float> intensity, const Shape& shape,
                                                  float threshold) {
  const std::size_t n = checked_size(shape);
  if (intensity.size() != n) throw std::invalid_argument("intensity does not match shape");
  return device_supports(intensity.data(), 1, shape, threshold);
}

std::vector<float> simulate_intensity(std::span<const Complex> object, const Shape& shape) {
  const std::size_t n = checked_size(shape);
  if (object.size() != n) throw std::invalid_argument("object does not match shape");
  std::vector<float> out(n);
  detail::qpch_check(qpch_simulate_intensity(reinterpret_cast<const float*>(object.data()), 1,
                                             static_cast<int>(shape.size()), shape.data(), out.data()));
  return out;
}

std::vector<std::uint8_t> half_support(std::span<const std::uint8_t> support, const Shape& shape, int axis) {
  checked_size(shape);
  if (axis < 0 || axis >= static_cast<int>(shape.size())) throw std::invalid_argument("detwin axis out of range");
  std::size_t stride = 1;
  for (std::size_t a = static_cast<std::size_t>(axis) + 1; a < shape.size(); ++a) stride *= shape[a];
  const std::size_t len = shape[static_cast<std::size_t>(axis)];
  // centre of mass along the axis: a sum of integer coordinates (exact in double here)
  double sum = 0.0, count = 0.0;
  for (std::size_t i = 0; i < support.size(); ++i) {
    if (support[i]) {
      sum += static_cast<double>((i / stride) % len);
      count += 1.0;
    }
  }
  const double centre = sum / std::max(count, 1.0);
  std::vector<std::uint8_t> out(support.size());
  for (std::size_t i = 0; i < support.size(); ++i) {
    out[i] = support[i] && static_cast<double>((i / stride) % len) < centre;
  }
  return out;
}

}  // namespace qpcu
