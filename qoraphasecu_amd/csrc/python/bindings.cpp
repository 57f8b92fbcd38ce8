// Python module qoraphasecu_amd._core (pybind11).  Same entry points and argument
// meaning as the reference's nanobind module (python/bindings.cpp:17-123), with the
// "hip" backend added.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "qoraphasecu/qoraphasecu.hpp"

namespace py = pybind11;

namespace {

template <class T>
using Array = py::array_t<T, py::array::c_style | py::array::forcecast>;

// (kind, n, beta, threshold, sigma, method, axis), flattened on the Python side
using OpTuple = std::tuple<std::string, int, float, float, float, std::string, int>;

qpcu::Shape shape_of(const py::array& a) {
  qpcu::Shape s;
  for (py::ssize_t i = 0; i < a.ndim(); ++i) s.push_back(static_cast<std::size_t>(a.shape(i)));
  return s;
}

template <class T>
std::span<const T> view(const Array<T>& a) {
  return {a.data(), static_cast<std::size_t>(a.size())};
}

template <class T>
py::array_t<T> to_numpy(std::vector<T>&& v, const qpcu::Shape& shape) {
  auto* owned = new std::vector<T>(std::move(v));
  py::capsule owner(owned, [](void* p) { delete static_cast<std::vector<T>*>(p); });
  std::vector<py::ssize_t> dims(shape.begin(), shape.end());
  return py::array_t<T>(dims, owned->data(), owner);
}

qpcu::Recipe to_recipe(const std::vector<OpTuple>& ops) {
  qpcu::Recipe r;
  for (const auto& [kind, n, beta, threshold, sigma, method, axis] : ops) {
    if (kind == "er") {
      r.ops.emplace_back(qpcu::ER{n});
    } else if (kind == "hio") {
      r.ops.emplace_back(qpcu::HIO{n, beta});
    } else if (kind == "raar") {
      r.ops.emplace_back(qpcu::RAAR{n, beta});
    } else if (kind == "shrinkwrap") {
      r.ops.emplace_back(qpcu::Shrinkwrap{threshold, sigma, method == "rms" ? qpcu::Threshold::rms : qpcu::Threshold::max});
    } else if (kind == "detwin_begin") {
      r.ops.emplace_back(qpcu::DetwinBegin{axis});
    } else if (kind == "detwin_end") {
      r.ops.emplace_back(qpcu::DetwinEnd{});
    } else {
      throw std::invalid_argument("unknown recipe operation '" + kind + "'");
    }
  }
  return r;
}

py::list reconstruct(const Array<float>& intensity, std::optional<Array<std::uint8_t>> support,
                     std::optional<Array<std::uint8_t>> mask, const std::vector<OpTuple>& ops, int starts,
                     int batch, std::uint64_t seed, std::optional<Array<std::complex<float>>> start,
                     float support_threshold, const std::string& backend, const std::string& variant,
                     int sync_every) {
  const qpcu::Shape shape = shape_of(intensity);
  const qpcu::Recipe recipe = to_recipe(ops);
  qpcu::Options options;
  options.backend = qpcu::parse_backend(backend);
  options.variant = qpcu::parse_variant(variant);
  options.starts = starts;
  options.batch = batch;
  options.seed = seed;
  if (start) options.start.assign(start->data(), start->data() + start->size());
  options.sync_every = sync_every;
  std::vector<qpcu::Reconstruction> runs;
  {
    py::gil_scoped_release release;
    const auto problem = qpcu::make_problem(view(intensity), shape,
                                            support ? view(*support) : std::span<const std::uint8_t>{},
                                            mask ? view(*mask) : std::span<const std::uint8_t>{},
                                            support_threshold);
    runs = qpcu::reconstruct(problem, recipe, options);
  }
  py::list out;
  for (auto& run : runs) {
    const qpcu::Shape history{run.errors.size()};
    out.append(py::make_tuple(to_numpy(std::move(run.object), shape), to_numpy(std::move(run.support), shape),
                              to_numpy(std::move(run.errors), history)));
  }
  return out;
}

// Several problems in one engine run (qpcu::reconstruct over a span of problems).
py::list reconstruct_many(const std::vector<Array<float>>& intensities,
                          std::optional<std::vector<std::optional<Array<std::uint8_t>>>> supports,
                          std::optional<std::vector<std::optional<Array<std::uint8_t>>>> masks,
                          const std::vector<OpTuple>& ops, int starts, int batch, std::uint64_t seed,
                          float support_threshold, const std::string& backend, const std::string& variant,
                          int sync_every) {
  const qpcu::Recipe recipe = to_recipe(ops);
  qpcu::Options options;
  options.backend = qpcu::parse_backend(backend);
  options.variant = qpcu::parse_variant(variant);
  options.starts = starts;
  options.batch = batch;
  options.seed = seed;
  options.sync_every = sync_every;
  std::vector<qpcu::Shape> shapes;
  for (const auto& a : intensities) shapes.push_back(shape_of(a));
  std::vector<std::vector<qpcu::Reconstruction>> runs;
  {
    py::gil_scoped_release release;
    std::vector<qpcu::Problem> problems;
    for (std::size_t k = 0; k < intensities.size(); ++k) {
      std::span<const std::uint8_t> sup, msk;
      if (supports && k < supports->size() && (*supports)[k]) sup = view(*(*supports)[k]);
      if (masks && k < masks->size() && (*masks)[k]) msk = view(*(*masks)[k]);
      problems.push_back(qpcu::make_problem(view(intensities[k]), shapes[k], sup, msk, support_threshold));
    }
    runs = qpcu::reconstruct(std::span<const qpcu::Problem>(problems), recipe, options);
  }
  py::list out;
  for (std::size_t k = 0; k < runs.size(); ++k) {
    py::list per;
    for (auto& run : runs[k]) {
      const qpcu::Shape history{run.errors.size()};
      per.append(py::make_tuple(to_numpy(std::move(run.object), shapes[k]), to_numpy(std::move(run.support), shapes[k]),
                                to_numpy(std::move(run.errors), history)));
    }
    out.append(per);
  }
  return out;
}

// The multi-problem state that outlives one run.
class PyMultiSolver {
 public:
  PyMultiSolver(const std::vector<Array<float>>& intensities,
                std::optional<std::vector<std::optional<Array<std::uint8_t>>>> supports,
                std::optional<std::vector<std::optional<Array<std::uint8_t>>>> masks, int starts, int batch,
                std::uint64_t seed, float support_threshold, const std::string& backend, const std::string& variant,
                int sync_every) {
    qpcu::Options options;
    options.backend = qpcu::parse_backend(backend);
    options.variant = qpcu::parse_variant(variant);
    options.starts = starts;
    options.batch = batch;
    options.seed = seed;
    options.sync_every = sync_every;
    for (const auto& a : intensities) shapes_.push_back(shape_of(a));
    py::gil_scoped_release release;
    std::vector<qpcu::Problem> problems;
    for (std::size_t k = 0; k < intensities.size(); ++k) {
      std::span<const std::uint8_t> sup, msk;
      if (supports && k < supports->size() && (*supports)[k]) sup = view(*(*supports)[k]);
      if (masks && k < masks->size() && (*masks)[k]) msk = view(*(*masks)[k]);
      problems.push_back(qpcu::make_problem(view(intensities[k]), shapes_[k], sup, msk, support_threshold));
    }
    solver_ = qpcu::make_multi_solver(problems, options);
  }

  void run(const std::vector<OpTuple>& ops) {
    const qpcu::Recipe recipe = to_recipe(ops);
    py::gil_scoped_release release;
    solver_->run(recipe);
  }

  py::list results() const {
    std::vector<std::vector<qpcu::Reconstruction>> runs;
    {
      py::gil_scoped_release release;
      runs = solver_->results();
    }
    py::list out;
    for (std::size_t k = 0; k < runs.size(); ++k) {
      py::list per;
      for (auto& run : runs[k]) {
        const qpcu::Shape history{run.errors.size()};
        per.append(py::make_tuple(to_numpy(std::move(run.object), shapes_[k]),
                                  to_numpy(std::move(run.support), shapes_[k]), to_numpy(std::move(run.errors), history)));
      }
      out.append(per);
    }
    return out;
  }

 private:
  std::vector<qpcu::Shape> shapes_;
  std::unique_ptr<qpcu::MultiSolver> solver_;
};

// A solver that outlives one run: the benchmark's construct / warm-up / timed-run protocol.
class PySolver {
 public:
  PySolver(const Array<float>& intensity, std::optional<Array<std::uint8_t>> support,
           std::optional<Array<std::uint8_t>> mask, int starts, int batch, std::uint64_t seed,
           std::optional<Array<std::complex<float>>> start, float support_threshold, const std::string& backend,
           const std::string& variant, int sync_every)
      : shape_(shape_of(intensity)) {
    qpcu::Options options;
    options.backend = qpcu::parse_backend(backend);
    options.variant = qpcu::parse_variant(variant);
    options.starts = starts;
    options.batch = batch;
    options.seed = seed;
    if (start) options.start.assign(start->data(), start->data() + start->size());
    options.sync_every = sync_every;
    py::gil_scoped_release release;
    const auto problem = qpcu::make_problem(view(intensity), shape_,
                                            support ? view(*support) : std::span<const std::uint8_t>{},
                                            mask ? view(*mask) : std::span<const std::uint8_t>{},
                                            support_threshold);
    solver_ = qpcu::make_solver(problem, options);
  }

  void run(const std::vector<OpTuple>& ops) {
    const qpcu::Recipe recipe = to_recipe(ops);
    py::gil_scoped_release release;
    solver_->run(recipe);
  }

  py::list results() const {
    std::vector<qpcu::Reconstruction> runs;
    {
      py::gil_scoped_release release;
      runs = solver_->results();
    }
    py::list out;
    for (auto& run : runs) {
      const qpcu::Shape history{run.errors.size()};
      out.append(py::make_tuple(to_numpy(std::move(run.object), shape_), to_numpy(std::move(run.support), shape_),
                                to_numpy(std::move(run.errors), history)));
    }
    return out;
  }

 private:
  qpcu::Shape shape_;
  std::unique_ptr<qpcu::Solver> solver_;
};

}  // namespace

PYBIND11_MODULE(_core, m) {
  py::class_<PySolver>(m, "Solver")
      .def(py::init<const Array<float>&, std::optional<Array<std::uint8_t>>, std::optional<Array<std::uint8_t>>,
                    int, int, std::uint64_t, std::optional<Array<std::complex<float>>>, float, const std::string&,
                    const std::string&, int>(),
           py::arg("intensity"), py::arg("support").none(true), py::arg("mask").none(true), py::arg("starts"),
           py::arg("batch"), py::arg("seed"), py::arg("start").none(true), py::arg("support_threshold"),
           py::arg("backend"), py::arg("variant"), py::arg("sync_every"))
      .def("run", &PySolver::run, py::arg("ops"))
      .def("results", &PySolver::results);
  py::class_<PyMultiSolver>(m, "MultiSolver")
      .def(py::init<const std::vector<Array<float>>&, std::optional<std::vector<std::optional<Array<std::uint8_t>>>>,
                    std::optional<std::vector<std::optional<Array<std::uint8_t>>>>, int, int, std::uint64_t, float,
                    const std::string&, const std::string&, int>(),
           py::arg("intensities"), py::arg("supports").none(true), py::arg("masks").none(true), py::arg("starts"),
           py::arg("batch"), py::arg("seed"), py::arg("support_threshold"), py::arg("backend"), py::arg("variant"),
           py::arg("sync_every"))
      .def("run", &PyMultiSolver::run, py::arg("ops"))
      .def("results", &PyMultiSolver::results);
  m.def("reconstruct_many", &reconstruct_many, py::arg("intensities"), py::arg("supports").none(true),
        py::arg("masks").none(true), py::arg("ops"), py::arg("starts"), py::arg("batch"), py::arg("seed"),
        py::arg("support_threshold"), py::arg("backend"), py::arg("variant"), py::arg("sync_every"));
  m.def(
      "make_problem",
      [](const Array<float>& intensity, std::optional<Array<std::uint8_t>> support,
         std::optional<Array<std::uint8_t>> mask, float support_threshold) {
        const auto shape = shape_of(intensity);
        qpcu::Problem p;
        {
          py::gil_scoped_release release;
          p = qpcu::make_problem(view(intensity), shape, support ? view(*support) : std::span<const std::uint8_t>{},
                                 mask ? view(*mask) : std::span<const std::uint8_t>{}, support_threshold);
        }
        return py::make_tuple(to_numpy(std::move(p.amplitude), shape), to_numpy(std::move(p.mask), shape),
                              to_numpy(std::move(p.support), shape));
      },
      py::arg("intensity"), py::arg("support").none(true), py::arg("mask").none(true),
      py::arg("support_threshold") = 0.1f);
  py::register_exception<std::invalid_argument>(m, "InvalidArgument", PyExc_ValueError);
  m.def("reconstruct", &reconstruct, py::arg("intensity"), py::arg("support").none(true),
        py::arg("mask").none(true), py::arg("ops"), py::arg("starts"), py::arg("batch"), py::arg("seed"),
        py::arg("start").none(true), py::arg("support_threshold"), py::arg("backend"), py::arg("variant"),
        py::arg("sync_every"));
  m.def(
      "autocorrelation_support",
      [](const Array<float>& intensity, float threshold) {
        const auto shape = shape_of(intensity);
        return to_numpy(qpcu::autocorrelation_support(view(intensity), shape, threshold), shape);
      },
      py::arg("intensity"), py::arg("threshold"));
  m.def(
      "simulate_intensity",
      [](const Array<std::complex<float>>& object) {
        const auto shape = shape_of(object);
        return to_numpy(qpcu::simulate_intensity(view(object), shape), shape);
      },
      py::arg("object"));
  m.def(
      "aligned_error",
      [](const Array<std::complex<float>>& a, const Array<std::complex<float>>& reference) {
        return qpcu::aligned_error(view(a), view(reference), shape_of(a));
      },
      py::arg("a"), py::arg("reference"));
  m.def(
      "half_support",
      [](const Array<std::uint8_t>& support, int axis) {
        const auto shape = shape_of(support);
        return to_numpy(qpcu::half_support(view(support), shape, axis), shape);
      },
      py::arg("support"), py::arg("axis"));
  m.def("cuda_available", &qpcu::cuda_available);
  m.def("hip_available", &qpcu::hip_available);
  m.def("hip_device_info", [] {
    const auto i = qpcu::hip_device_info();
    py::dict d;
    d["name"] = i.name;
    d["major"] = i.sm_major;
    d["minor"] = i.sm_minor;
    d["memory_bytes"] = i.memory_bytes;
    d["l2_bytes"] = i.l2_bytes;
    d["peak_bandwidth_gbs"] = i.peak_bandwidth_gbs;
    return d;
  });
}
