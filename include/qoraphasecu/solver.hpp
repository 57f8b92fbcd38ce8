// Solvers (reference: include/qoraphasecu/solver.hpp:14-100).  Same API, plus
// Backend::hip (this library), hip_available() / hip_device_info(), the bf16 variants and
// the multi-problem reconstruct().
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "qoraphasecu/problem.hpp"
#include "qoraphasecu/recipe.hpp"

namespace qpcu {

enum class Backend { cpu, cuda, hip };

// Execution variants.  On Backend::hip:
//   naive      unfused: per-axis transforms, element-wise modulus, element-wise update
//   fused      fused passes (modulus in the frequency-axis transform, update in the
//              real-axis transform), one start per launch, eager launches
//   graph      fused passes replayed from hipGraphs, one start per launch
//   batched    graph with `batch` starts per launch (0 = all)
//   half       batched with x and the wave stored in fp16 (power-of-two extents)
//   callback   graph (the fused epilogues are this build's form of FFT callbacks)
//   mixed      half for HIO/RAAR, fp32 for ER
//   bf16       half with bfloat16 storage
//   mixed_bf16 mixed with bfloat16 storage
enum class Variant { naive, fused, graph, batched, half, callback, mixed, bf16, mixed_bf16 };

std::string_view to_string(Backend backend);
std::string_view to_string(Variant variant);
Backend parse_backend(std::string_view name);
Variant parse_variant(std::string_view name);

// Called every `sync_every` iterations with the iteration count so far and the latest
// Fourier error of each start.
using Progress = std::function<void(int iteration, std::span<const float> errors)>;

struct Options {
  Backend backend = Backend::cpu;
  Variant variant = Variant::batched;
  int starts = 1;
  // Starts per launch in the batched and reduced-precision variants; 0 means all.
  // Must divide `starts`.
  int batch = 0;
  std::uint64_t seed = 0;
  // Starting objects, centre-origin, one per start back to back, each cut to the
  // support.  Empty draws random phases from `seed`.
  std::vector<Complex> start{};
  // Iterations per graph launch, and the progress interval.
  int sync_every = 10;
  Progress progress{};
};

// One reconstructed object, centre-origin, defined up to a global phase, a translation
// and a conjugate flip.
struct Reconstruction {
  std::vector<Complex> object;
  std::vector<std::uint8_t> support;
  std::vector<float> errors;  // Fourier error per iteration

  float error() const;  // final Fourier error, NaN before any iteration
};

// Holds the state of every start; run() may be called repeatedly and continues.
class Solver {
 public:
  virtual ~Solver() = default;
  virtual void run(const Recipe& recipe) = 0;
  virtual std::vector<Reconstruction> results() const = 0;
};

std::unique_ptr<Solver> make_solver(const Problem& problem, const Options& options = {});

std::vector<Reconstruction> reconstruct(const Problem& problem,
                                        const Recipe& recipe = default_recipe(),
                                        const Options& options = {});

// The state of several independent problems (any mix of shapes), each with
// options.starts starts; run() continues like Solver::run.  results()[p] holds problem p's
// starts.  On Backend::hip with a batched or reduced-precision variant and power-of-two
// extents, problems of one shape share one device solver and every shape advances in the
// same graph launch; otherwise the problems run one after another.
class MultiSolver {
 public:
  virtual ~MultiSolver() = default;
  virtual void run(const Recipe& recipe) = 0;
  virtual std::vector<std::vector<Reconstruction>> results() const = 0;
};

std::unique_ptr<MultiSolver> make_multi_solver(std::span<const Problem> problems, const Options& options = {});

// make_multi_solver + run + results.
std::vector<std::vector<Reconstruction>> reconstruct(std::span<const Problem> problems,
                                                     const Recipe& recipe = default_recipe(),
                                                     const Options& options = {});

// Index of the run with the lowest final error (NaN never wins).
std::size_t best_index(std::span<const Reconstruction> runs);

bool cuda_available();
bool hip_available();

struct DeviceInfo {
  std::string name;
  int sm_major = 0;  // HIP: gfx major, e.g. 9 for gfx950
  int sm_minor = 0;  // HIP: the remaining digits, e.g. 50
  std::size_t memory_bytes = 0;
  std::size_t l2_bytes = 0;
  double peak_bandwidth_gbs = 0.0;  // 2 x memory clock x bus width
};

// std::runtime_error without a device of that kind.
DeviceInfo cuda_device_info();
DeviceInfo hip_device_info();

}  // namespace qpcu
