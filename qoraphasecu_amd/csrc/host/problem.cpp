// make_problem and its helpers.  Reference behaviour: src/core/problem.cpp:15-127; the
// transforms run on the GPU through the qpch C ABI.
#include "qoraphasecu/problem.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string>

#include "abi.hpp"
#include "layout.hpp"

namespace qpcu {

std::size_t numel(const Shape& shape) {
  return std::accumulate(shape.begin(), shape.end(), std::size_t{1}, std::multiplies<>{});
}

namespace {

void require_shape(const Shape& shape) {
  if (shape.size() != 2 && shape.size() != 3) {
    throw std::invalid_argument("shape must be 2D or 3D, got ndim=" + std::to_string(shape.size()));
  }
  for (std::size_t n : shape) {
    if (n == 0) throw std::invalid_argument("shape has a zero-length axis");
  }
}

template <class T>
void require_size(std::span<const T> a, std::size_t n, const char* what) {
  if (!a.empty() && a.size() != n) {
    throw std::invalid_argument(std::string(what) + " has " + std::to_string(a.size()) +
                                " elements, shape has " + std::to_string(n));
  }
}

// unitary transform of an FFT-order complex array on the device
void device_fft(std::vector<Complex>& data, const Shape& shape, bool forward) {
  detail::qpch_check(qpch_fft(reinterpret_cast<float*>(data.data()), static_cast<int>(shape.size()),
                              shape.data(), forward ? 1 : 0));
}

}  // namespace

Problem make_problem(std::span<const float> intensity, Shape shape,
                     std::span<const std::uint8_t> support, std::span<const std::uint8_t> mask,
                     float support_threshold) {
  require_shape(shape);
  const std::size_t n = numel(shape);
  if (intensity.size() != n) {
    throw std::invalid_argument("intensity has " + std::to_string(intensity.size()) +
                                " elements, shape has " + std::to_string(n));
  }
  require_size(support, n, "support");
  require_size(mask, n, "mask");
  for (float v : intensity) {
    if (!std::isfinite(v)) throw std::invalid_argument("intensity contains NaN or Inf");
  }
  if (!mask.empty() && std::all_of(mask.begin(), mask.end(), [](std::uint8_t m) { return m != 0; })) {
    throw std::invalid_argument("mask marks every pixel as bad; nothing left to fit");
  }
  Problem p;
  p.amplitude.resize(n);
  for (std::size_t i = 0; i < n; ++i) p.amplitude[i] = std::sqrt(std::max(intensity[i], 0.0f));
  p.mask.assign(n, 0);
  for (std::size_t i = 0; i < mask.size(); ++i) p.mask[i] = mask[i] != 0;
  if (support.empty()) {
    p.support = autocorrelation_support(intensity, shape, support_threshold);
  } else {
    p.support.resize(n);
    for (std::size_t i = 0; i < n; ++i) p.support[i] = support[i] != 0;
  }
  p.shape = std::move(shape);
  return p;
}

std::vector<std::uint8_t> autocorrelation_support(std::span<const float> intensity,
                                                  const Shape& shape, float threshold) {
  require_shape(shape);
  const std::size_t n = numel(shape);
  if (intensity.size() != n) throw std::invalid_argument("intensity does not match shape");
  const auto ordered = detail::to_fft_order(intensity, shape);
  std::vector<Complex> buf(ordered.begin(), ordered.end());
  device_fft(buf, shape, false);
  std::vector<float> ac(n);
  for (std::size_t i = 0; i < n; ++i) ac[i] = std::abs(buf[i]);  // hypot, as the reference
  const auto centered = detail::to_centered<float>(ac, shape);
  const float level = threshold * *std::max_element(centered.begin(), centered.end());
  std::vector<std::uint8_t> out(n);
  for (std::size_t i = 0; i < n; ++i) out[i] = centered[i] > level;
  return out;
}

std::vector<float> simulate_intensity(std::span<const Complex> object, const Shape& shape) {
  require_shape(shape);
  if (object.size() != numel(shape)) throw std::invalid_argument("object does not match shape");
  auto wave = detail::to_fft_order(object, shape);
  device_fft(wave, shape, true);
  const auto centered = detail::to_centered<Complex>(wave, shape);
  std::vector<float> out(centered.size());
  for (std::size_t i = 0; i < out.size(); ++i) out[i] = std::norm(centered[i]);
  return out;
}

std::vector<std::uint8_t> half_support(std::span<const std::uint8_t> support, const Shape& shape,
                                       int axis) {
  require_shape(shape);
  if (axis < 0 || axis >= static_cast<int>(shape.size())) {
    throw std::invalid_argument("detwin axis out of range");
  }
  std::size_t stride = 1;
  for (std::size_t a = static_cast<std::size_t>(axis) + 1; a < shape.size(); ++a) stride *= shape[a];
  const std::size_t len = shape[static_cast<std::size_t>(axis)];
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
