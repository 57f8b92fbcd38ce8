// Names, validation and backend dispatch.  Reference behaviour: src/core/solver.cpp:13-103.
#include "qoraphasecu/solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>

#include "abi.hpp"
#include "backends.hpp"

namespace qpcu {

namespace {

constexpr Variant kVariants[] = {Variant::naive,    Variant::fused, Variant::graph,
                                 Variant::batched,  Variant::half,  Variant::callback,
                                 Variant::mixed,    Variant::bf16,  Variant::mixed_bf16};

}  // namespace

std::string_view to_string(Backend backend) {
  switch (backend) {
    case Backend::cpu: return "cpu";
    case Backend::cuda: return "cuda";
    case Backend::hip: return "hip";
  }
  return "?";
}

std::string_view to_string(Variant variant) {
  switch (variant) {
    case Variant::naive: return "naive";
    case Variant::fused: return "fused";
    case Variant::graph: return "graph";
    case Variant::batched: return "batched";
    case Variant::half: return "half";
    case Variant::callback: return "callback";
    case Variant::mixed: return "mixed";
    case Variant::bf16: return "bf16";
    case Variant::mixed_bf16: return "mixed_bf16";
  }
  return "?";
}

Backend parse_backend(std::string_view name) {
  for (Backend b : {Backend::cpu, Backend::cuda, Backend::hip}) {
    if (name == to_string(b)) return b;
  }
  throw std::invalid_argument("unknown backend '" + std::string(name) + "'");
}

Variant parse_variant(std::string_view name) {
  for (Variant v : kVariants) {
    if (name == to_string(v)) return v;
  }
  throw std::invalid_argument("unknown variant '" + std::string(name) + "'");
}

float Reconstruction::error() const {
  return errors.empty() ? std::numeric_limits<float>::quiet_NaN() : errors.back();
}

std::unique_ptr<Solver> make_solver(const Problem& problem, const Options& options) {
  if (options.starts < 1) throw std::invalid_argument("starts must be >= 1");
  if (options.sync_every < 1) throw std::invalid_argument("sync_every must be >= 1");
  if (options.batch < 0 || (options.batch > 0 && options.starts % options.batch != 0)) {
    throw std::invalid_argument("batch must divide starts");
  }
  if (problem.shape.size() != 2 && problem.shape.size() != 3) {
    throw std::invalid_argument("problem shape must be 2D or 3D");
  }
  const std::size_t n = numel(problem.shape);
  if (problem.amplitude.size() != n || problem.mask.size() != n || problem.support.size() != n) {
    throw std::invalid_argument("problem arrays do not match its shape; build it with make_problem");
  }
  if (!options.start.empty() && options.start.size() != n * static_cast<std::size_t>(options.starts)) {
    throw std::invalid_argument("start must hold one array of the problem's size per start");
  }
  switch (options.backend) {
    case Backend::hip:
      return detail::make_hip_solver(problem, options);
    case Backend::cpu:
      throw std::runtime_error(
          "this library implements Backend::hip; Backend::cpu is the reference's CPU backend");
    case Backend::cuda:
      throw std::runtime_error("this library was built without CUDA; use Backend::hip");
  }
  throw std::invalid_argument("unknown backend");
}

std::vector<Reconstruction> reconstruct(const Problem& problem, const Recipe& recipe,
                                        const Options& options) {
  validate(recipe);  // before any device work, so a bad recipe fails the same way everywhere
  auto solver = make_solver(problem, options);
  solver->run(recipe);
  return solver->results();
}

namespace {

// One solver per problem, run one after another.
class SequentialMultiSolver final : public MultiSolver {
 public:
  SequentialMultiSolver(std::span<const Problem> problems, const Options& options) {
    for (const Problem& p : problems) solvers_.push_back(make_solver(p, options));
  }
  void run(const Recipe& recipe) override {
    validate(recipe);
    for (auto& s : solvers_) s->run(recipe);
  }
  std::vector<std::vector<Reconstruction>> results() const override {
    std::vector<std::vector<Reconstruction>> out;
    for (const auto& s : solvers_) out.push_back(s->results());
    return out;
  }

 private:
  std::vector<std::unique_ptr<Solver>> solvers_;
};

}  // namespace

std::unique_ptr<MultiSolver> make_multi_solver(std::span<const Problem> problems, const Options& options) {
  if (problems.empty()) throw std::invalid_argument("no problems");
  if (options.starts < 1) throw std::invalid_argument("starts must be >= 1");
  if (options.sync_every < 1) throw std::invalid_argument("sync_every must be >= 1");
  bool ensemble = options.backend == Backend::hip;
  switch (options.variant) {
    case Variant::naive:
    case Variant::fused:
    case Variant::graph:
    case Variant::callback:
      ensemble = false;  // one start per launch: the problems run one after another
      break;
    default:
      break;
  }
  for (const Problem& p : problems) {
    const std::size_t n = numel(p.shape);
    if (p.shape.size() != 2 && p.shape.size() != 3) throw std::invalid_argument("problem shape must be 2D or 3D");
    if (p.amplitude.size() != n || p.mask.size() != n || p.support.size() != n) {
      throw std::invalid_argument("problem arrays do not match its shape; build it with make_problem");
    }
    // shapes the fused passes cover: power-of-two axes up to 4096, other axes up to 2048
    for (std::size_t a : p.shape) {
      const bool pow2 = a >= 2 && (a & (a - 1)) == 0 && a <= 4096;
      ensemble = ensemble && (pow2 || (a >= 2 && a <= 2048));
    }
  }
  if (!ensemble) {
    if (!options.start.empty()) throw std::invalid_argument("start is not supported for several problems here");
    return std::make_unique<SequentialMultiSolver>(problems, options);
  }
  if (options.batch < 0) throw std::invalid_argument("batch must be >= 0");
  return detail::make_hip_ensemble(problems, options);
}

std::vector<std::vector<Reconstruction>> reconstruct(std::span<const Problem> problems, const Recipe& recipe,
                                                     const Options& options) {
  validate(recipe);
  auto solver = make_multi_solver(problems, options);
  solver->run(recipe);
  return solver->results();
}

std::size_t best_index(std::span<const Reconstruction> runs) {
  if (runs.empty()) throw std::invalid_argument("no runs");
  // compared in double; a NaN error never displaces a finite one
  std::size_t best = 0;
  for (std::size_t i = 1; i < runs.size(); ++i) {
    const double e = runs[i].error();
    const double b = runs[best].error();
    if (e < b || std::isnan(b)) best = i;
  }
  return best;
}

bool cuda_available() { return false; }

DeviceInfo cuda_device_info() {
  throw std::runtime_error("this library was built without CUDA; use hip_device_info()");
}

bool hip_available() {
  int count = 0;
  return qpch_device_count(&count) == QPCH_OK && count > 0;
}

DeviceInfo hip_device_info() {
  qpch_device_info_t raw{};
  detail::qpch_check(qpch_device_info(&raw));
  DeviceInfo info;
  info.name = raw.name;
  const std::string arch = raw.arch;  // gfx950[:features]
  if (info.name.empty()) info.name = "AMD GPU " + arch.substr(0, arch.find(':'));
  if (arch.rfind("gfx", 0) == 0 && arch.size() >= 5) {
    const std::string digits = arch.substr(3, arch.find(':') - 3);
    info.sm_major = digits.empty() ? 0 : digits[0] - '0';
    info.sm_minor = digits.size() > 1 ? std::atoi(digits.substr(1).c_str()) : 0;
  }
  info.memory_bytes = raw.memory_bytes;
  info.l2_bytes = raw.l2_bytes;
  info.peak_bandwidth_gbs = raw.peak_bandwidth_gbs;
  return info;
}

}  // namespace qpcu
