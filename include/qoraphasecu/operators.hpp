// Single projections on centre-origin arrays (reference: include/qoraphasecu/operators.hpp
// :9-33).  qpcu::hip runs them on the GPU; qpcu::cpu and qpcu::cuda are the reference's
// backends and throw std::runtime_error in this library.
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "qoraphasecu/problem.hpp"

namespace qpcu::hip {

// Replaces `obj` by P_M obj and returns the Fourier error of the input.
float modulus_project(const Problem& problem, std::span<Complex> obj);

void support_project(std::span<const std::uint8_t> support, std::span<Complex> obj);

// Gaussian blur with transfer function exp(-2 (pi sigma f)^2).
std::vector<float> gaussian_blur(std::span<const float> image, const Shape& shape, float sigma);

}  // namespace qpcu::hip

namespace qpcu::cpu {
float modulus_project(const Problem& problem, std::span<Complex> obj);
void support_project(std::span<const std::uint8_t> support, std::span<Complex> obj);
std::vector<float> gaussian_blur(std::span<const float> image, const Shape& shape, float sigma);
}  // namespace qpcu::cpu

namespace qpcu::cuda {
float modulus_project(const Problem& problem, std::span<Complex> obj);
void support_project(std::span<const std::uint8_t> support, std::span<Complex> obj);
std::vector<float> gaussian_blur(std::span<const float> image, const Shape& shape, float sigma);
}  // namespace qpcu::cuda
