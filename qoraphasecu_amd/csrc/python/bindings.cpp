// Python module qoraphasecu_amd._core (pybind11) over the qpcu:: C++ API.
//
// Entry points and argument meaning are those of the reference's module (python/
// bindings.cpp): reconstruct(intensity, support|None, mask|None, ops, starts, batch, seed,
// start|None, support_threshold, backend, variant, sync_every) -> [(object, support,
// errors)], with ops the 7-field records (kind, n, beta, threshold, sigma, method, axis),
// plus autocorrelation_support and aligned_error.  Added here: the "hip" backend, the
// multi-problem reconstruct_many, Solver / MultiSolver objects that outlive one run (the
// benchmark's construct, warm-up, time protocol), make_problem, simulate_intensity,
// half_support and the device queries.  Long calls release the GIL.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "qoraphasecu/qoraphasecu.hpp"

namespace py = pybind11;

namespace {

template <class T>
using In = py::array_t<T, py::array::c_style | py::array::forcecast>;
using Flags = std::optional<In<std::uint8_t>>;
using FlagList = std::optional<std::vector<Flags>>;
using Record = std::tuple<std::string, int, float, float, float, std::string, int>;

qpcu::Shape extents(const py::array& a) {
  return qpcu::Shape(a.shape(), a.shape() + a.ndim());
}

template <class T>
std::span<const T> span_of(const In<T>& a) {
  return std::span<const T>(a.data(), static_cast<std::size_t>(a.size()));
}

template <class T>
std::span<const T> span_of(const std::optional<In<T>>& a) {
  return a ? span_of(*a) : std::span<const T>{};
}

// A fresh numpy array holding a copy of `values`.
template <class T>
py::array_t<T> numpy_copy(const std::vector<T>& values, const qpcu::Shape& shape) {
  py::array_t<T> out(std::vector<py::ssize_t>(shape.begin(), shape.end()));
  if (!values.empty()) std::memcpy(out.mutable_data(), values.data(), values.size() * sizeof(T));
  return out;
}

py::tuple pack(const qpcu::Reconstruction& r, const qpcu::Shape& shape) {
  return py::make_tuple(numpy_copy(r.object, shape), numpy_copy(r.support, shape),
                        numpy_copy(r.errors, qpcu::Shape{r.errors.size()}));
}

py::list pack_all(const std::vector<qpcu::Reconstruction>& runs, const qpcu::Shape& shape) {
  py::list out;
  for (const auto& r : runs) out.append(pack(r, shape));
  return out;
}

qpcu::Threshold level_of(const std::string& method) {
  return method == "rms" ? qpcu::Threshold::rms : qpcu::Threshold::max;
}

// One engine operation from its record; unknown kinds are an argument error.
qpcu::Op parse(const Record& rec) {
  using Make = std::function<qpcu::Op(const Record&)>;
  static const std::map<std::string, Make> table = {
      {"er", [](const Record& r) -> qpcu::Op { return qpcu::ER{std::get<1>(r)}; }},
      {"hio", [](const Record& r) -> qpcu::Op { return qpcu::HIO{std::get<1>(r), std::get<2>(r)}; }},
      {"raar", [](const Record& r) -> qpcu::Op { return qpcu::RAAR{std::get<1>(r), std::get<2>(r)}; }},
      {"shrinkwrap",
       [](const Record& r) -> qpcu::Op {
         return qpcu::Shrinkwrap{std::get<3>(r), std::get<4>(r), level_of(std::get<5>(r))};
       }},
      {"detwin_begin", [](const Record& r) -> qpcu::Op { return qpcu::DetwinBegin{std::get<6>(r)}; }},
      {"detwin_end", [](const Record&) -> qpcu::Op { return qpcu::DetwinEnd{}; }},
  };
  const auto it = table.find(std::get<0>(rec));
  if (it == table.end()) throw std::invalid_argument("unknown recipe operation '" + std::get<0>(rec) + "'");
  return it->second(rec);
}

qpcu::Recipe recipe_of(const std::vector<Record>& records) {
  qpcu::Recipe recipe;
  recipe.ops.reserve(records.size());
  for (const auto& rec : records) recipe.ops.push_back(parse(rec));
  return recipe;
}

qpcu::Options options_of(const std::string& backend, const std::string& variant, int starts, int batch,
                         std::uint64_t seed, int sync_every,
                         const std::optional<In<std::complex<float>>>& start = std::nullopt) {
  qpcu::Options o;
  o.backend = qpcu::parse_backend(backend);
  o.variant = qpcu::parse_variant(variant);
  o.starts = starts;
  o.batch = batch;
  o.seed = seed;
  o.sync_every = sync_every;
  if (start) {
    const auto s = span_of(*start);
    o.start.assign(s.begin(), s.end());
  }
  return o;
}

// The problems of a multi-problem call; the inputs' arrays stay alive in the caller.
std::vector<qpcu::Problem> problems_of(const std::vector<In<float>>& intensities, const FlagList& supports,
                                       const FlagList& masks, const std::vector<qpcu::Shape>& shapes,
                                       float support_threshold) {
  auto nth = [](const FlagList& list, std::size_t k) -> std::span<const std::uint8_t> {
    return list && k < list->size() ? span_of((*list)[k]) : std::span<const std::uint8_t>{};
  };
  std::vector<qpcu::ProblemInput> inputs;
  for (std::size_t k = 0; k < intensities.size(); ++k) {
    inputs.push_back({span_of(intensities[k]), shapes[k], nth(supports, k), nth(masks, k)});
  }
  return qpcu::make_problems(inputs, support_threshold);
}

py::list reconstruct(const In<float>& intensity, const Flags& support, const Flags& mask,
                     const std::vector<Record>& ops, int starts, int batch, std::uint64_t seed,
                     const std::optional<In<std::complex<float>>>& start, float support_threshold,
                     const std::string& backend, const std::string& variant, int sync_every) {
  const qpcu::Shape shape = extents(intensity);
  const qpcu::Recipe recipe = recipe_of(ops);
  const qpcu::Options options = options_of(backend, variant, starts, batch, seed, sync_every, start);
  std::vector<qpcu::Reconstruction> runs;
  {
    py::gil_scoped_release unlocked;
    const qpcu::Problem problem =
        qpcu::make_problem(span_of(intensity), shape, span_of(support), span_of(mask), support_threshold);
    runs = qpcu::reconstruct(problem, recipe, options);
  }
  return pack_all(runs, shape);
}

py::list reconstruct_many(const std::vector<In<float>>& intensities, const FlagList& supports,
                          const FlagList& masks, const std::vector<Record>& ops, int starts, int batch,
                          std::uint64_t seed, float support_threshold, const std::string& backend,
                          const std::string& variant, int sync_every) {
  std::vector<qpcu::Shape> shapes;
  for (const auto& a : intensities) shapes.push_back(extents(a));
  const qpcu::Recipe recipe = recipe_of(ops);
  const qpcu::Options options = options_of(backend, variant, starts, batch, seed, sync_every);
  std::vector<std::vector<qpcu::Reconstruction>> runs;
  {
    py::gil_scoped_release unlocked;
    const auto problems = problems_of(intensities, supports, masks, shapes, support_threshold);
    runs = qpcu::reconstruct(std::span<const qpcu::Problem>(problems), recipe, options);
  }
  py::list out;
  for (std::size_t k = 0; k < runs.size(); ++k) out.append(pack_all(runs[k], shapes[k]));
  return out;
}

// A solver kept across runs.
class Solver {
 public:
  Solver(const In<float>& intensity, const Flags& support, const Flags& mask, int starts, int batch,
         std::uint64_t seed, const std::optional<In<std::complex<float>>>& start, float support_threshold,
         const std::string& backend, const std::string& variant, int sync_every)
      : shape_(extents(intensity)) {
    const qpcu::Options options = options_of(backend, variant, starts, batch, seed, sync_every, start);
    py::gil_scoped_release unlocked;
    const qpcu::Problem problem =
        qpcu::make_problem(span_of(intensity), shape_, span_of(support), span_of(mask), support_threshold);
    engine_ = qpcu::make_solver(problem, options);
  }

  void run(const std::vector<Record>& ops) {
    const qpcu::Recipe recipe = recipe_of(ops);
    py::gil_scoped_release unlocked;
    engine_->run(recipe);
  }

  py::list results() const {
    std::vector<qpcu::Reconstruction> runs;
    {
      py::gil_scoped_release unlocked;
      runs = engine_->results();
    }
    return pack_all(runs, shape_);
  }

 private:
  qpcu::Shape shape_;
  std::unique_ptr<qpcu::Solver> engine_;
};

// Several problems kept across runs (every shape advances in the same graph launch).
class MultiSolver {
 public:
  MultiSolver(const std::vector<In<float>>& intensities, const FlagList& supports, const FlagList& masks, int starts,
              int batch, std::uint64_t seed, float support_threshold, const std::string& backend,
              const std::string& variant, int sync_every) {
    for (const auto& a : intensities) shapes_.push_back(extents(a));
    const qpcu::Options options = options_of(backend, variant, starts, batch, seed, sync_every);
    py::gil_scoped_release unlocked;
    const auto problems = problems_of(intensities, supports, masks, shapes_, support_threshold);
    engine_ = qpcu::make_multi_solver(problems, options);
  }

  void run(const std::vector<Record>& ops) {
    const qpcu::Recipe recipe = recipe_of(ops);
    py::gil_scoped_release unlocked;
    engine_->run(recipe);
  }

  py::list results() const {
    std::vector<std::vector<qpcu::Reconstruction>> runs;
    {
      py::gil_scoped_release unlocked;
      runs = engine_->results();
    }
    py::list out;
    for (std::size_t k = 0; k < runs.size(); ++k) out.append(pack_all(runs[k], shapes_[k]));
    return out;
  }

 private:
  std::vector<qpcu::Shape> shapes_;
  std::unique_ptr<qpcu::MultiSolver> engine_;
};

py::dict device_info() {
  const auto i = qpcu::hip_device_info();
  py::dict d;
  d["name"] = i.name;
  d["major"] = i.sm_major;
  d["minor"] = i.sm_minor;
  d["memory_bytes"] = i.memory_bytes;
  d["l2_bytes"] = i.l2_bytes;
  d["peak_bandwidth_gbs"] = i.peak_bandwidth_gbs;
  return d;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  py::register_exception<std::invalid_argument>(m, "InvalidArgument", PyExc_ValueError);

  m.def("reconstruct", &reconstruct, py::arg("intensity"), py::arg("support").none(true), py::arg("mask").none(true),
        py::arg("ops"), py::arg("starts"), py::arg("batch"), py::arg("seed"), py::arg("start").none(true),
        py::arg("support_threshold"), py::arg("backend"), py::arg("variant"), py::arg("sync_every"));
  m.def("reconstruct_many", &reconstruct_many, py::arg("intensities"), py::arg("supports").none(true),
        py::arg("masks").none(true), py::arg("ops"), py::arg("starts"), py::arg("batch"), py::arg("seed"),
        py::arg("support_threshold"), py::arg("backend"), py::arg("variant"), py::arg("sync_every"));

  py::class_<Solver>(m, "Solver")
      .def(py::init<const In<float>&, const Flags&, const Flags&, int, int, std::uint64_t,
                    const std::optional<In<std::complex<float>>>&, float, const std::string&, const std::string&,
                    int>(),
           py::arg("intensity"), py::arg("support").none(true), py::arg("mask").none(true), py::arg("starts"),
           py::arg("batch"), py::arg("seed"), py::arg("start").none(true), py::arg("support_threshold"),
           py::arg("backend"), py::arg("variant"), py::arg("sync_every"))
      .def("run", &Solver::run, py::arg("ops"))
      .def("results", &Solver::results);
  py::class_<MultiSolver>(m, "MultiSolver")
      .def(py::init<const std::vector<In<float>>&, const FlagList&, const FlagList&, int, int, std::uint64_t, float,
                    const std::string&, const std::string&, int>(),
           py::arg("intensities"), py::arg("supports").none(true), py::arg("masks").none(true), py::arg("starts"),
           py::arg("batch"), py::arg("seed"), py::arg("support_threshold"), py::arg("backend"), py::arg("variant"),
           py::arg("sync_every"))
      .def("run", &MultiSolver::run, py::arg("ops"))
      .def("results", &MultiSolver::results);

  m.def(
      "make_problem",
      [](const In<float>& intensity, const Flags& support, const Flags& mask, float support_threshold) {
        const qpcu::Shape shape = extents(intensity);
        qpcu::Problem p;
        {
          py::gil_scoped_release unlocked;
          p = qpcu::make_problem(span_of(intensity), shape, span_of(support), span_of(mask), support_threshold);
        }
        return py::make_tuple(numpy_copy(p.amplitude, shape), numpy_copy(p.mask, shape), numpy_copy(p.support, shape));
      },
      py::arg("intensity"), py::arg("support").none(true), py::arg("mask").none(true),
      py::arg("support_threshold") = 0.1f);
  m.def(
      "autocorrelation_support",
      [](const In<float>& intensity, float threshold) {
        const qpcu::Shape shape = extents(intensity);
        return numpy_copy(qpcu::autocorrelation_support(span_of(intensity), shape, threshold), shape);
      },
      py::arg("intensity"), py::arg("threshold"));
  m.def(
      "simulate_intensity",
      [](const In<std::complex<float>>& object) {
        const qpcu::Shape shape = extents(object);
        return numpy_copy(qpcu::simulate_intensity(span_of(object), shape), shape);
      },
      py::arg("object"));
  m.def(
      "aligned_error",
      [](const In<std::complex<float>>& a, const In<std::complex<float>>& reference) {
        return qpcu::aligned_error(span_of(a), span_of(reference), extents(a));
      },
      py::arg("a"), py::arg("reference"));
  m.def(
      "half_support",
      [](const In<std::uint8_t>& support, int axis) {
        const qpcu::Shape shape = extents(support);
        return numpy_copy(qpcu::half_support(span_of(support), shape, axis), shape);
      },
      py::arg("support"), py::arg("axis"));
  m.def("cuda_available", &qpcu::cuda_available);
  m.def("hip_available", &qpcu::hip_available);
  m.def("hip_device_info", &device_info);
}
