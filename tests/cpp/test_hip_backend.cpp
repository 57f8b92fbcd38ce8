// Backend::hip through the public C++ API (include/qoraphasecu/*.hpp), the drop-in
// boundary itself: the solver semantics the reference pins in C++ only
// (tests/test_cuda.cpp:121-137, tests/test_recover.cpp:78-107 and the validation cases).
// A self-contained runner: each case prints PASS / FAIL, the exit status is the number of
// failed cases.  Needs a HIP device.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "qoraphasecu/qoraphasecu.hpp"

namespace {

int g_failed = 0;
int g_checks = 0;

#define CHECK(cond)                                                           \
  do {                                                                        \
    ++g_checks;                                                               \
    if (!(cond)) {                                                            \
      std::printf("    check failed at line %d: %s\n", __LINE__, #cond);      \
      throw std::runtime_error("check failed");                               \
    }                                                                         \
  } while (0)

template <class Ex, class F>
bool throws(F&& f) {
  try {
    f();
  } catch (const Ex&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

void run_case(const char* name, const std::function<void()>& body) {
  try {
    body();
    std::printf("PASS %s\n", name);
  } catch (const std::exception& e) {
    ++g_failed;
    std::printf("FAIL %s: %s\n", name, e.what());
  }
  std::fflush(stdout);
}

// A non-centrosymmetric right triangle of legs m in an n x n grid, amplitude with a dip
// and a phase ramp, and its support: a test object with a unique reconstruction up to
// the trivial ambiguities.
struct Triangle {
  qpcu::Shape shape;
  std::vector<qpcu::Complex> object;
  std::vector<std::uint8_t> support;
  std::vector<float> intensity;
};

Triangle triangle(std::size_t n, std::size_t m) {
  Triangle t;
  t.shape = {n, n};
  t.object.assign(n * n, qpcu::Complex(0.0f, 0.0f));
  t.support.assign(n * n, 0);
  const std::size_t y0 = (n - m) / 2, x0 = (n - m) / 2;
  for (std::size_t y = 0; y < m; ++y) {
    for (std::size_t x = 0; x + y < m; ++x) {
      const float a = 1.0f - 0.5f * std::exp(-0.1f * static_cast<float>((y - m / 3) * (y - m / 3) + x * x));
      const float ph = 0.3f * static_cast<float>(x) / static_cast<float>(m);
      t.object[(y0 + y) * n + x0 + x] = std::polar(a, ph);
      t.support[(y0 + y) * n + x0 + x] = 1;
    }
  }
  t.intensity = qpcu::simulate_intensity(t.object, t.shape);
  return t;
}

qpcu::Options hip_options(qpcu::Variant v, int starts, int sync_every = 10) {
  qpcu::Options o;
  o.backend = qpcu::Backend::hip;
  o.variant = v;
  o.starts = starts;
  o.sync_every = sync_every;
  return o;
}

}  // namespace

int main() {
  using namespace qpcu;
  if (!hip_available()) {
    std::printf("FAIL no HIP device\n");
    return 1;
  }
  const Triangle s = triangle(32, 10);
  const Problem p = make_problem(s.intensity, s.shape, s.support);

  run_case("device_reports_gfx950", [] {
    const DeviceInfo d = hip_device_info();
    CHECK(d.sm_major == 9 && d.sm_minor == 50);
    CHECK(d.peak_bandwidth_gbs > 10.0);
  });

  run_case("names_round_trip", [] {
    for (Variant v : {Variant::naive, Variant::fused, Variant::graph, Variant::batched, Variant::half,
                      Variant::callback, Variant::mixed, Variant::bf16, Variant::mixed_bf16}) {
      CHECK(parse_variant(to_string(v)) == v);
    }
    CHECK(parse_backend("hip") == Backend::hip);
  });

  // tests/test_cuda.cpp:121-137: progress fires every sync_every iterations with the
  // history's own values, for every variant
  for (Variant v : {Variant::naive, Variant::fused, Variant::graph, Variant::batched, Variant::callback,
                    Variant::half, Variant::mixed, Variant::bf16, Variant::mixed_bf16}) {
    const std::string name = "progress_matches_the_error_history[" + std::string(to_string(v)) + "]";
    run_case(name.c_str(), [&] {
      std::vector<int> seen;
      std::vector<float> reported;
      Options o = hip_options(v, 3, 5);
      o.progress = [&](int it, std::span<const float> errors) {
        CHECK(errors.size() == 3u);
        seen.push_back(it);
        reported.push_back(errors[2]);
      };
      const auto runs = reconstruct(p, recipe(HIO{7}, ER{8}), o);
      CHECK((seen == std::vector<int>{5, 10, 15}));
      CHECK(runs.size() == 3u && runs[2].errors.size() == 15u);
      for (std::size_t k = 0; k < seen.size(); ++k) {
        CHECK(reported[k] == runs[2].errors[static_cast<std::size_t>(seen[k] - 1)]);
      }
    });
  }

  // tests/test_recover.cpp:78-89
  run_case("progress_fires_every_sync_interval", [&] {
    std::vector<int> seen;
    Options o = hip_options(Variant::graph, 2, 4);
    o.progress = [&](int it, std::span<const float> errors) {
      CHECK(errors.size() == 2u);
      seen.push_back(it);
    };
    reconstruct(p, recipe(HIO{6}, ER{6}), o);
    CHECK((seen == std::vector<int>{4, 8, 12}));
  });

  // tests/test_recover.cpp:97-107: a run continues from the current state, bit for bit
  for (Variant v : {Variant::graph, Variant::batched, Variant::mixed}) {
    const std::string name = "run_continues_from_the_current_state[" + std::string(to_string(v)) + "]";
    run_case(name.c_str(), [&] {
      const Options o = hip_options(v, 2);
      auto solver = make_solver(p, o);
      solver->run(recipe(HIO{10}));
      solver->run(recipe(ER{5}));
      const auto split = solver->results();
      const auto whole = reconstruct(p, recipe(HIO{10}, ER{5}), o);
      for (std::size_t k = 0; k < 2; ++k) {
        CHECK(split[k].errors == whole[k].errors);
        CHECK(split[k].object == whole[k].object);
        CHECK(split[k].support == whole[k].support);
      }
    });
  }

  // axes that are not powers of two run the generic fused passes: a run still continues
  // from the current state bit for bit, and ER still descends once the iterate is in the
  // support set (from the second ER iteration on)
  const Triangle t24 = triangle(24, 8);
  const Problem q24 = make_problem(t24.intensity, t24.shape, t24.support);
  for (Variant v : {Variant::graph, Variant::batched}) {
    const std::string name = "other_sizes_continue_and_descend[" + std::string(to_string(v)) + "]";
    run_case(name.c_str(), [&] {
      const Options o = hip_options(v, 2);
      auto solver = make_solver(q24, o);
      solver->run(recipe(HIO{10}));
      solver->run(recipe(ER{30}));
      const auto split = solver->results();
      const auto whole = reconstruct(q24, recipe(HIO{10}, ER{30}), o);
      for (std::size_t k = 0; k < 2; ++k) {
        CHECK(split[k].errors == whole[k].errors);
        CHECK(split[k].object == whole[k].object);
        CHECK(whole[k].errors.size() == 40u);
        for (std::size_t t = 12; t < whole[k].errors.size(); ++t) {
          CHECK(whole[k].errors[t] <= whole[k].errors[t - 1] * (1.0f + 1e-4f));
        }
      }
    });
  }

  // tests/test_recover.cpp:53-62: ER never raises the error (from the support-cut random
  // start every iterate lies in the support set, where ER is a descent)
  run_case("error_reduction_never_raises_the_error", [&] {
    for (Variant v : {Variant::graph, Variant::batched}) {
      const auto runs = reconstruct(p, recipe(ER{100}), hip_options(v, 3));
      for (const auto& r : runs) {
        CHECK(r.errors.size() == 100u);
        CHECK(r.errors.back() < r.errors.front());
        for (std::size_t t = 1; t < r.errors.size(); ++t) CHECK(r.errors[t] <= r.errors[t - 1] * (1.0f + 1e-4f));
      }
    }
  });

  // tests/test_recover.cpp:91-95 and the argument checks of src/core/solver.cpp:61-86
  run_case("rejects_bad_arguments", [&] {
    Options o = hip_options(Variant::batched, 4);
    o.batch = 3;
    CHECK(throws<std::invalid_argument>([&] { make_solver(p, o); }));
    CHECK(throws<std::invalid_argument>([&] { reconstruct(p, recipe(ER{0}), hip_options(Variant::graph, 1)); }));
    CHECK(throws<std::invalid_argument>([&] { reconstruct(p, recipe(DetwinBegin{}, ER{2}), hip_options(Variant::graph, 1)); }));
    const Triangle odd = triangle(24, 8);
    const Problem q = make_problem(odd.intensity, {24, 24}, odd.support);
    CHECK(throws<std::invalid_argument>([&] { make_solver(q, hip_options(Variant::half, 1)); }));
  });

  // a given start is used as is: ER from the true object stays at the truth
  run_case("given_start_is_a_fixed_point_of_er", [&] {
    Options o = hip_options(Variant::graph, 1);
    o.start = s.object;
    const auto runs = reconstruct(p, recipe(ER{5}), o);
    CHECK(runs[0].error() < 1e-5f);
  });

  // best_index: the lowest final error; a NaN never displaces a finite one
  run_case("best_index_skips_nan", [] {
    std::vector<Reconstruction> r(3);
    r[0].errors = {std::nanf("")};
    r[1].errors = {0.5f};
    r[2].errors = {0.25f};
    CHECK(best_index(r) == 2u);
  });

  std::printf("%d case(s) failed, %d checks\n", g_failed, g_checks);
  return g_failed;
}
