// Backend::hip: the recipe walk, progress and result layout on the host; every array
// operation behind the qpch C ABI.  Reference counterpart: the CudaSolver of
// src/cuda/cuda_solver.cu:58-500 (run 145-163, results 165-191, apply 236-276,
// iterate 300-338, collect_errors/report 423-453).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

#include "abi.hpp"
#include "backends.hpp"
#include "layout.hpp"
#include "trace.hpp"

namespace qpcu::detail {

namespace {

constexpr double kEps = 1e-12;  // src/core/pointwise.hpp:16

// Shapes the engine runs with fused passes (power-of-two axes up to 4096, other axes up to
// 2048 in the generic passes), whose launch configuration the autotuner can choose.
bool is_fusable(const Shape& shape) {
  for (std::size_t a : shape) {
    const bool pow2 = a >= 2 && a <= 4096 && (a & (a - 1)) == 0;
    if (!pow2 && (a < 2 || a > 2048)) return false;
  }
  return true;
}

struct Plan {
  int schedule;
  int storage;
  int mixed;
  bool batched;
};

// Launch configuration from the engine's autotuner (starts per launch, concurrent launch
// groups, XCD-affine numbering, persistent or per-pass launches); QPCH_TUNE=0 keeps the
// defaults.  The configuration does not change results.
void tune(qpch_problem& q) {
  if (const char* v = std::getenv("QPCH_TUNE"); v && v[0] == '0') {
    if (q.batch == 0) q.batch = q.starts * (q.problems > 1 ? q.problems : 1);
    return;
  }
  qpch_tuning t{};
  {
    const TraceRange range("qpch autotune");
    qpch_check(qpch_autotune(&q, &t));
  }
  q.batch = t.batch;
  q.plane = t.plane;
  q.xcd = t.xcd;
  q.lanes = t.lanes;
  q.radix = t.radix;
  q.persist = t.persist;
}

Plan plan_for(Variant v) {
  switch (v) {
    case Variant::naive: return {QPCH_UNFUSED, QPCH_F32, 0, false};
    case Variant::fused: return {QPCH_FUSED, QPCH_F32, 0, false};
    case Variant::graph: return {QPCH_GRAPH, QPCH_F32, 0, false};
    case Variant::callback: return {QPCH_GRAPH, QPCH_F32, 0, false};
    case Variant::batched: return {QPCH_GRAPH, QPCH_F32, 0, true};
    case Variant::half: return {QPCH_GRAPH, QPCH_F16, 0, true};
    case Variant::mixed: return {QPCH_GRAPH, QPCH_F16, 1, true};
    case Variant::bf16: return {QPCH_GRAPH, QPCH_BF16, 0, true};
    case Variant::mixed_bf16: return {QPCH_GRAPH, QPCH_BF16, 1, true};
  }
  throw std::invalid_argument("unknown variant");
}

// x <- S ? P_M x : a x + b P_M x; src/core/pointwise.hpp:41-45
struct Coefficients {
  float a, b;
};

// The recipe as the engine's operation list for a whole-recipe graph (qpch_group_run).
std::vector<qpch_op> engine_ops(const Recipe& recipe) {
  std::vector<qpch_op> out;
  out.reserve(recipe.ops.size());
  auto iterate = [&](int n, Coefficients c, bool refine) {
    qpch_op o{};
    o.kind = QPCH_OP_ITERATE;
    o.iterations = n;
    o.a = c.a;
    o.b = c.b;
    o.refine = refine ? 1 : 0;
    out.push_back(o);
  };
  for (const Op& op : recipe.ops) {
    std::visit(
        [&](const auto& v) {
          using T = std::decay_t<decltype(v)>;
          qpch_op o{};
          if constexpr (std::is_same_v<T, ER>) {
            iterate(v.n, {0.0f, 0.0f}, true);
          } else if constexpr (std::is_same_v<T, HIO>) {
            iterate(v.n, {1.0f, -v.beta}, false);
          } else if constexpr (std::is_same_v<T, RAAR>) {
            iterate(v.n, {v.beta, 1.0f - 2.0f * v.beta}, false);
          } else if constexpr (std::is_same_v<T, Shrinkwrap>) {
            o.kind = QPCH_OP_SHRINKWRAP;
            o.threshold = v.threshold;
            o.sigma = v.sigma;
            o.method = v.method == Threshold::max ? 0 : 1;
            out.push_back(o);
          } else if constexpr (std::is_same_v<T, DetwinBegin>) {
            o.kind = QPCH_OP_DETWIN_BEGIN;
            o.axis = v.axis;
            out.push_back(o);
          } else {
            o.kind = QPCH_OP_DETWIN_END;
            out.push_back(o);
          }
        },
        op);
  }
  return out;
}

class HipSolver final : public Solver {
 public:
  HipSolver(const Problem& problem, const Options& options)
      : shape_(problem.shape), n_(numel(problem.shape)), options_(options), starts_(options.starts),
        errors_(static_cast<std::size_t>(options.starts)) {
    const Plan plan = plan_for(options.variant);
    const auto amplitude = to_fft_order<float>(problem.amplitude, shape_);
    const auto mask = to_fft_order<std::uint8_t>(problem.mask, shape_);
    const auto support = to_fft_order<std::uint8_t>(problem.support, shape_);
    std::vector<Complex> start;
    if (!options.start.empty()) {
      start.reserve(options.start.size());
      for (int s = 0; s < starts_; ++s) {
        const auto part = to_fft_order<Complex>(
            std::span(options.start).subspan(static_cast<std::size_t>(s) * n_, n_), shape_);
        start.insert(start.end(), part.begin(), part.end());
      }
    }
    qpch_problem p{};
    p.ndim = static_cast<int>(shape_.size());
    for (std::size_t a = 0; a < shape_.size(); ++a) p.shape[a] = shape_[a];
    p.amplitude = amplitude.data();
    p.mask = mask.data();
    p.support = support.data();
    p.starts = starts_;
    p.batch = plan.batched ? options.batch : 1;
    p.seed = options.seed;
    p.start = start.empty() ? nullptr : reinterpret_cast<const float*>(start.data());
    p.storage = plan.storage;
    p.mixed = plan.mixed;
    p.schedule = plan.schedule;
    p.lanes = plan.batched ? 0 : 1;  // one start per launch stays one launch at a time
    if (plan.schedule == QPCH_GRAPH && is_fusable(shape_)) tune(p);
    qpch_check(qpch_solver_create(&p, &h_));
    energy_ = qpch_solver_energy(h_);
    // Without a progress callback the batched variants replay one graph per recipe stage
    // (measured on MI355X: as fast as 10-iteration chunks on the line passes, and up to 11%
    // faster for the whole-image kernel, which then runs a stage per launch).  The plain
    // variants keep the reference's sync_every chunks.
    stage_chunks_ = plan.batched;
  }

  ~HipSolver() override { qpch_solver_destroy(h_); }

  HipSolver(const HipSolver&) = delete;
  HipSolver& operator=(const HipSolver&) = delete;

  void run(const Recipe& recipe) override {
    validate(recipe);
    const TraceRange range("iterations");
    const int capacity = recipe.iterations();
    run_iterations_ = 0;
    if (capacity > 0) qpch_check(qpch_solver_begin_run(h_, capacity));
    for (const Op& op : recipe.ops) {
      std::visit([this](const auto& o) { apply(o); }, op);
    }
    if (run_iterations_ > 0) collect_errors();
    qpch_check(qpch_solver_synchronize(h_));
  }

  std::vector<Reconstruction> results() const override {
    std::vector<Complex> x(n_ * static_cast<std::size_t>(starts_));
    std::vector<std::uint8_t> s(n_ * static_cast<std::size_t>(starts_));
    qpch_check(qpch_solver_read_state(h_, reinterpret_cast<float*>(x.data())));
    qpch_check(qpch_solver_get_support(h_, s.data()));
    std::vector<Reconstruction> out;
    out.reserve(static_cast<std::size_t>(starts_));
    for (int k = 0; k < starts_; ++k) {
      const std::size_t off = static_cast<std::size_t>(k) * n_;
      out.push_back({to_centered<Complex>(std::span(x).subspan(off, n_), shape_),
                     to_centered<std::uint8_t>(std::span(s).subspan(off, n_), shape_),
                     errors_[static_cast<std::size_t>(k)]});
    }
    return out;
  }

 private:
  void apply(const ER& o) { iterate(o.n, {0.0f, 0.0f}, true); }
  void apply(const HIO& o) { iterate(o.n, {1.0f, -o.beta}, false); }
  void apply(const RAAR& o) { iterate(o.n, {o.beta, 1.0f - 2.0f * o.beta}, false); }

  void apply(const Shrinkwrap& o) {
    qpch_check(qpch_solver_shrinkwrap(h_, o.threshold, o.sigma, o.method == Threshold::max ? 0 : 1));
  }

  // The support cut at its centre of mass, on the device: the reference's host round trip
  // (every support to the host, half_support per start, back) costs two transfers of all
  // supports and a CPU pass over them.
  void apply(const DetwinBegin& o) {
    qpch_check(qpch_solver_save_support(h_));
    qpch_check(qpch_solver_half_support(h_, o.axis));
  }

  void apply(const DetwinEnd&) { qpch_check(qpch_solver_restore_support(h_)); }

  // Chunks end on multiples of sync_every of the solver's own iteration count, so the
  // progress schedule does not depend on how the recipe splits into stages.
  void iterate(int n, Coefficients c, bool refine) {
    const bool whole_stage = stage_chunks_ && !options_.progress;
    for (int done = 0; done < n;) {
      const int len = whole_stage ? n - done : std::min(n - done, options_.sync_every - iteration_ % options_.sync_every);
      qpch_check(qpch_solver_iterate(h_, len, c.a, c.b, refine ? 1 : 0));
      done += len;
      iteration_ += len;
      run_iterations_ += len;
      if (options_.progress && iteration_ % options_.sync_every == 0) report();
    }
  }

  float to_error(double numerator) const {
    return static_cast<float>(std::sqrt(numerator / (energy_ + kEps)));
  }

  void collect_errors() {
    std::vector<double> host(static_cast<std::size_t>(run_iterations_) * starts_);
    qpch_check(qpch_solver_read_errors(h_, 0, run_iterations_, host.data()));
    for (int t = 0; t < run_iterations_; ++t) {
      for (int k = 0; k < starts_; ++k) {
        errors_[static_cast<std::size_t>(k)].push_back(
            to_error(host[static_cast<std::size_t>(t) * starts_ + k]));
      }
    }
  }

  void report() {
    std::vector<double> host(static_cast<std::size_t>(starts_));
    qpch_check(qpch_solver_read_errors(h_, run_iterations_ - 1, 1, host.data()));
    std::vector<float> latest(static_cast<std::size_t>(starts_));
    for (int k = 0; k < starts_; ++k) latest[static_cast<std::size_t>(k)] = to_error(host[static_cast<std::size_t>(k)]);
    options_.progress(iteration_, latest);
  }

  Shape shape_;
  std::size_t n_;
  Options options_;
  int starts_;
  qpch_solver* h_ = nullptr;
  bool stage_chunks_ = false;  // a graph per recipe stage when there is no progress callback
  double energy_ = 0.0;
  std::vector<std::vector<float>> errors_;
  int iteration_ = 0;
  int run_iterations_ = 0;
};

// Several problems (any mix of shapes) in one engine: problems of one shape are stacked in
// one device solver, and the solvers of the different shapes advance together, every
// chunk of iterations one graph launch with a branch per shape.  The recipe walk, progress
// and result layout follow HipSolver.
class HipEnsemble final : public MultiSolver {
 public:
  HipEnsemble(std::span<const Problem> problems, const Options& options)
      : options_(options), problems_(problems.size()), errors_(problems.size()) {
    const Plan plan = plan_for(options.variant);
    std::map<Shape, std::vector<std::size_t>> by_shape;
    for (std::size_t k = 0; k < problems.size(); ++k) by_shape[problems[k].shape].push_back(k);
    std::size_t start_offset = 0;
    std::vector<std::size_t> offsets(problems.size());
    for (std::size_t k = 0; k < problems.size(); ++k) {
      offsets[k] = start_offset;
      start_offset += numel(problems[k].shape) * static_cast<std::size_t>(options.starts);
    }
    if (!options.start.empty() && options.start.size() != start_offset) {
      throw std::invalid_argument("start must hold one array per start of every problem");
    }
    for (auto& [shape, members] : by_shape) {
      Class c;
      c.shape = shape;
      c.n = numel(shape);
      c.members = members;
      std::vector<float> amp;
      std::vector<std::uint8_t> mask, sup;
      std::vector<Complex> start;
      for (std::size_t k : members) {
        const Problem& p = problems[k];
        const auto a = to_fft_order<float>(p.amplitude, shape);
        const auto m = to_fft_order<std::uint8_t>(p.mask, shape);
        const auto su = to_fft_order<std::uint8_t>(p.support, shape);
        amp.insert(amp.end(), a.begin(), a.end());
        mask.insert(mask.end(), m.begin(), m.end());
        sup.insert(sup.end(), su.begin(), su.end());
        if (!options.start.empty()) {
          for (int s = 0; s < options.starts; ++s) {
            const auto part = to_fft_order<Complex>(
                std::span(options.start).subspan(offsets[k] + static_cast<std::size_t>(s) * c.n, c.n), shape);
            start.insert(start.end(), part.begin(), part.end());
          }
        }
      }
      qpch_problem q{};
      q.ndim = static_cast<int>(shape.size());
      for (std::size_t a = 0; a < shape.size(); ++a) q.shape[a] = shape[a];
      q.amplitude = amp.data();
      q.mask = mask.data();
      q.support = sup.data();
      q.starts = options.starts;
      q.batch = options.batch;
      q.seed = options.seed;
      q.start = start.empty() ? nullptr : reinterpret_cast<const float*>(start.data());
      q.storage = plan.storage;
      q.mixed = plan.mixed;
      q.schedule = QPCH_GRAPH;
      q.problems = static_cast<int>(members.size());
      // a member's launch groups run one after another on its branch of the group's graph
      // (the shapes already run side by side), so concurrent lanes are not on offer
      q.lanes = 1;
      tune(q);
      qpch_check(qpch_solver_create(&q, &c.h));
      c.energies.resize(members.size());
      qpch_check(qpch_solver_energies(c.h, c.energies.data()));
      classes_.push_back(std::move(c));
    }
    std::vector<qpch_solver*> hs;
    for (auto& c : classes_) hs.push_back(c.h);
    qpch_check(qpch_group_create(hs.data(), static_cast<int>(hs.size()), &group_));
  }

  ~HipEnsemble() override {
    if (group_) qpch_group_destroy(group_);
    for (auto& c : classes_) qpch_solver_destroy(c.h);
  }

  HipEnsemble(const HipEnsemble&) = delete;
  HipEnsemble& operator=(const HipEnsemble&) = delete;

  void run(const Recipe& recipe) override {
    validate(recipe);
    const TraceRange range("iterations");
    const int capacity = recipe.iterations();
    run_iterations_ = 0;
    if (capacity > 0) {
      for (auto& c : classes_) qpch_check(qpch_solver_begin_run(c.h, capacity));
    }
    if (!options_.progress) {
      // the whole recipe in one graph launch: every shape's launch groups are branches
      const auto ops = engine_ops(recipe);
      qpch_check(qpch_group_run(group_, ops.data(), static_cast<int>(ops.size())));
      iteration_ += capacity;
      run_iterations_ = capacity;
    } else {
      for (const Op& op : recipe.ops) {
        std::visit([this](const auto& o) { apply(o); }, op);
      }
    }
    if (run_iterations_ > 0) collect_errors();
    qpch_check(qpch_group_synchronize(group_));
  }

  std::vector<std::vector<Reconstruction>> results() const override {
    std::vector<std::vector<Reconstruction>> out(problems_);
    const auto starts = static_cast<std::size_t>(options_.starts);
    for (const auto& c : classes_) {
      const std::size_t total = c.members.size() * starts;
      std::vector<Complex> x(c.n * total);
      std::vector<std::uint8_t> s(c.n * total);
      qpch_check(qpch_solver_read_state(c.h, reinterpret_cast<float*>(x.data())));
      qpch_check(qpch_solver_get_support(c.h, s.data()));
      for (std::size_t q = 0; q < c.members.size(); ++q) {
        const std::size_t k = c.members[q];
        for (std::size_t t = 0; t < starts; ++t) {
          const std::size_t off = (q * starts + t) * c.n;
          out[k].push_back({to_centered<Complex>(std::span(x).subspan(off, c.n), c.shape),
                            to_centered<std::uint8_t>(std::span(s).subspan(off, c.n), c.shape),
                            errors_[k][t]});
        }
      }
    }
    return out;
  }

 private:
  struct Class {
    Shape shape;
    std::size_t n = 0;
    std::vector<std::size_t> members;
    qpch_solver* h = nullptr;
    std::vector<double> energies;
  };

  void apply(const ER& o) { iterate(o.n, {0.0f, 0.0f}, true); }
  void apply(const HIO& o) { iterate(o.n, {1.0f, -o.beta}, false); }
  void apply(const RAAR& o) { iterate(o.n, {o.beta, 1.0f - 2.0f * o.beta}, false); }

  void apply(const Shrinkwrap& o) {
    for (auto& c : classes_) {
      qpch_check(qpch_solver_shrinkwrap(c.h, o.threshold, o.sigma, o.method == Threshold::max ? 0 : 1));
    }
  }

  void apply(const DetwinBegin& o) {
    for (auto& c : classes_) {
      qpch_check(qpch_solver_save_support(c.h));
      qpch_check(qpch_solver_half_support(c.h, o.axis));
    }
  }

  void apply(const DetwinEnd&) {
    for (auto& c : classes_) qpch_check(qpch_solver_restore_support(c.h));
  }

  void iterate(int n, Coefficients cf, bool refine) {
    for (int done = 0; done < n;) {
      const int len = std::min(n - done, options_.sync_every - iteration_ % options_.sync_every);
      qpch_check(qpch_group_iterate(group_, len, cf.a, cf.b, refine ? 1 : 0));
      done += len;
      iteration_ += len;
      run_iterations_ += len;
      if (options_.progress && iteration_ % options_.sync_every == 0) report();
    }
  }

  // per problem p, start t: errors of class start (q * starts + t)
  void collect_errors() {
    const int starts = options_.starts;
    for (auto& c : classes_) {
      const int total = static_cast<int>(c.members.size()) * starts;
      std::vector<double> host(static_cast<std::size_t>(run_iterations_) * total);
      qpch_check(qpch_solver_read_errors(c.h, 0, run_iterations_, host.data()));
      for (std::size_t q = 0; q < c.members.size(); ++q) {
        auto& errs = errors_[c.members[q]];
        errs.resize(static_cast<std::size_t>(starts));
        for (int t = 0; t < run_iterations_; ++t) {
          for (int s = 0; s < starts; ++s) {
            const double num = host[static_cast<std::size_t>(t) * total + q * starts + s];
            errs[static_cast<std::size_t>(s)].push_back(
                static_cast<float>(std::sqrt(num / (c.energies[q] + kEps))));
          }
        }
      }
    }
  }

  void report() {
    const int starts = options_.starts;
    std::vector<float> latest(problems_ * static_cast<std::size_t>(starts));
    for (auto& c : classes_) {
      const int total = static_cast<int>(c.members.size()) * starts;
      std::vector<double> host(static_cast<std::size_t>(total));
      qpch_check(qpch_solver_read_errors(c.h, run_iterations_ - 1, 1, host.data()));
      for (std::size_t q = 0; q < c.members.size(); ++q) {
        for (int s = 0; s < starts; ++s) {
          latest[c.members[q] * starts + s] =
              static_cast<float>(std::sqrt(host[q * starts + s] / (c.energies[q] + kEps)));
        }
      }
    }
    options_.progress(iteration_, latest);
  }

  Options options_;
  std::size_t problems_;
  std::vector<Class> classes_;
  qpch_group* group_ = nullptr;
  std::vector<std::vector<std::vector<float>>> errors_;  // [problem][start][iteration]
  int iteration_ = 0;
  int run_iterations_ = 0;
};

}  // namespace

std::unique_ptr<MultiSolver> make_hip_ensemble(std::span<const Problem> problems, const Options& options) {
  if (!hip_available()) throw std::runtime_error("no HIP device available");
  return std::make_unique<HipEnsemble>(problems, options);
}

std::unique_ptr<Solver> make_hip_solver(const Problem& problem, const Options& options) {
  if (!hip_available()) throw std::runtime_error("no HIP device available");
  return std::make_unique<HipSolver>(problem, options);
}

}  // namespace qpcu::detail
