// Reconstruction inputs (reference: include/qoraphasecu/problem.hpp:9-49, same API).
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <span>
#include <vector>

namespace qpcu {

using Complex = std::complex<float>;

// Row-major extents, slowest axis first; 2D or 3D.
using Shape = std::vector<std::size_t>;

std::size_t numel(const Shape& shape);

// Fixed inputs of a reconstruction.  Arrays are centre-origin: index n / 2 of every axis
// is q = 0 (detector) or r = 0 (object).
struct Problem {
  Shape shape;
  std::vector<float> amplitude;       // sqrt(max(intensity, 0))
  std::vector<std::uint8_t> mask;     // 1 = unusable pixel, left unconstrained
  std::vector<std::uint8_t> support;  // 1 = inside the object
};

// Validates the inputs (std::invalid_argument) and derives the amplitude.  An empty
// `support` becomes autocorrelation_support(intensity, shape, support_threshold); an
// empty `mask` marks every pixel usable.
Problem make_problem(std::span<const float> intensity, Shape shape,
                     std::span<const std::uint8_t> support = {},
                     std::span<const std::uint8_t> mask = {}, float support_threshold = 0.1f);

// Inputs of one problem for make_problems (spans that must outlive the call).
struct ProblemInput {
  std::span<const float> intensity;
  Shape shape;
  std::span<const std::uint8_t> support = {};  // empty: the autocorrelation support
  std::span<const std::uint8_t> mask = {};     // empty: every pixel usable
};

// make_problem for several problems at once (no reference counterpart: the entry point of
// the multi-problem reconstruct).  Each problem is validated and built as make_problem
// builds it; the autocorrelation supports of all problems of one shape come from one
// device pass.
std::vector<Problem> make_problems(std::span<const ProblemInput> inputs, float support_threshold = 0.1f);

// |unitary IFFT(intensity)| above `threshold` times its maximum (computed on the GPU).
std::vector<std::uint8_t> autocorrelation_support(std::span<const float> intensity,
                                                  const Shape& shape, float threshold = 0.1f);

// |centred unitary FFT(object)|^2, the noiseless detector intensity (computed on the GPU).
std::vector<float> simulate_intensity(std::span<const Complex> object, const Shape& shape);

// The part of `support` strictly below its centre of mass along `axis`.
std::vector<std::uint8_t> half_support(std::span<const std::uint8_t> support, const Shape& shape,
                                       int axis = 0);

}  // namespace qpcu
