// Single projections on centre-origin arrays.  Reference: src/cpu/operators.cpp:11-39 and
// src/cuda/operators.cu:62-130; here on the GPU through the qpch C ABI.
#include "qoraphasecu/operators.hpp"

#include <cmath>
#include <stdexcept>

#include "abi.hpp"
#include "layout.hpp"

namespace qpcu::hip {

float modulus_project(const Problem& problem, std::span<Complex> obj) {
  const std::size_t n = numel(problem.shape);
  if (obj.size() != n || problem.amplitude.size() != n || problem.mask.size() != n) {
    throw std::invalid_argument("obj does not match the problem shape");
  }
  const auto amplitude = detail::to_fft_order<float>(problem.amplitude, problem.shape);
  const auto mask = detail::to_fft_order<std::uint8_t>(problem.mask, problem.shape);
  auto x = detail::to_fft_order<Complex>(obj, problem.shape);
  double numerator = 0.0;
  detail::qpch_check(qpch_modulus_project(reinterpret_cast<float*>(x.data()), amplitude.data(),
                                          mask.data(), static_cast<int>(problem.shape.size()),
                                          problem.shape.data(), &numerator));
  double energy = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    if (!mask[i]) energy += static_cast<double>(amplitude[i]) * amplitude[i];
  }
  const auto back = detail::to_centered<Complex>(x, problem.shape);
  std::copy(back.begin(), back.end(), obj.begin());
  return static_cast<float>(std::sqrt(numerator / (energy + 1e-12)));
}

void support_project(std::span<const std::uint8_t> support, std::span<Complex> obj) {
  if (obj.size() != support.size()) throw std::invalid_argument("support does not match obj");
  detail::qpch_check(qpch_support_project(reinterpret_cast<float*>(obj.data()), support.data(), obj.size()));
}

std::vector<float> gaussian_blur(std::span<const float> image, const Shape& shape, float sigma) {
  if (shape.size() != 2 && shape.size() != 3) throw std::invalid_argument("shape must be 2D or 3D");
  if (image.size() != numel(shape)) throw std::invalid_argument("image does not match shape");
  // a circular convolution commutes with the centring roll: no reordering needed
  std::vector<float> out(image.size());
  detail::qpch_check(qpch_gaussian_blur(image.data(), out.data(), static_cast<int>(shape.size()),
                                        shape.data(), sigma));
  return out;
}

}  // namespace qpcu::hip

namespace qpcu {
namespace {
[[noreturn]] void not_here(const char* backend) {
  throw std::runtime_error(std::string("qpcu::") + backend +
                           " operators belong to the reference's backends; this library provides qpcu::hip");
}
}  // namespace

namespace cpu {
float modulus_project(const Problem&, std::span<Complex>) { not_here("cpu"); }
void support_project(std::span<const std::uint8_t>, std::span<Complex>) { not_here("cpu"); }
std::vector<float> gaussian_blur(std::span<const float>, const Shape&, float) { not_here("cpu"); }
}  // namespace cpu

namespace cuda {
float modulus_project(const Problem&, std::span<Complex>) { not_here("cuda"); }
void support_project(std::span<const std::uint8_t>, std::span<Complex>) { not_here("cuda"); }
std::vector<float> gaussian_blur(std::span<const float>, const Shape&, float) { not_here("cuda"); }
}  // namespace cuda
}  // namespace qpcu
