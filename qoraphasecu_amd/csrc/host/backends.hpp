// Backend plug points (reference: src/core/backends.hpp:7-12, plus make_hip_solver).
#pragma once

#include <memory>

#include "qoraphasecu/solver.hpp"

namespace qpcu::detail {

std::unique_ptr<Solver> make_hip_solver(const Problem& problem, const Options& options);

// Several problems, stacked per shape and advanced by one graph launch per chunk.
std::unique_ptr<MultiSolver> make_hip_ensemble(std::span<const Problem> problems, const Options& options);

}  // namespace qpcu::detail
