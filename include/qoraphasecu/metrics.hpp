// Reference: include/qoraphasecu/metrics.hpp:7-15 (same API; transforms on the GPU).
#pragma once

#include <span>

#include "qoraphasecu/problem.hpp"

namespace qpcu {

// ||a - b|| / ||b|| after aligning `a` to `reference` by an integer translation, the
// conjugate flip conj(a(-r)) and a global phase.
float aligned_error(std::span<const Complex> a, std::span<const Complex> reference,
                    const Shape& shape);

}  // namespace qpcu
