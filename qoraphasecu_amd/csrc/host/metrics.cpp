// aligned_error.  Reference behaviour: src/core/metrics.cpp:13-86; the three transforms
// of each alignment run on the GPU.
#include "qoraphasecu/metrics.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "abi.hpp"
#include "layout.hpp"

namespace qpcu {

namespace {

void transform(std::vector<Complex>& v, const Shape& shape, bool forward) {
  detail::qpch_check(qpch_fft(reinterpret_cast<float*>(v.data()), static_cast<int>(shape.size()),
                              shape.data(), forward ? 1 : 0));
}

// Shift b by the peak of the circular cross-correlation with ref, then fit a global phase.
float translated_error(const std::vector<Complex>& b, std::span<const Complex> ref, const Shape& shape) {
  const std::size_t n = b.size();
  std::vector<Complex> fr(ref.begin(), ref.end());
  std::vector<Complex> fb = b;
  transform(fr, shape, true);
  transform(fb, shape, true);
  for (std::size_t i = 0; i < n; ++i) fr[i] *= std::conj(fb[i]);
  transform(fr, shape, false);
  std::size_t peak = 0;
  float best = std::norm(fr[0]);
  for (std::size_t i = 1; i < n; ++i) {  // first maximum
    const float v = std::norm(fr[i]);
    if (v > best) {
      best = v;
      peak = i;
    }
  }
  const auto d = detail::dims3(shape);
  const std::size_t s0 = peak / (d[1] * d[2]), s1 = (peak / d[2]) % d[1], s2 = peak % d[2];
  std::complex<double> overlap = 0.0;
  double ref_norm = 0.0;
  std::vector<Complex> moved(n);
  std::size_t i = 0;
  for (std::size_t z = 0; z < d[0]; ++z)
    for (std::size_t y = 0; y < d[1]; ++y)
      for (std::size_t x = 0; x < d[2]; ++x, ++i) {
        moved[i] = b[(((z + d[0] - s0) % d[0]) * d[1] + (y + d[1] - s1) % d[1]) * d[2] + (x + d[2] - s2) % d[2]];
      }
  for (std::size_t k = 0; k < n; ++k) {
    overlap += std::complex<double>(ref[k]) * std::conj(std::complex<double>(moved[k]));
    ref_norm += std::norm(std::complex<double>(ref[k]));
  }
  const double mag = std::abs(overlap);
  const std::complex<double> rot = mag > 0.0 ? overlap / mag : std::complex<double>(1.0);
  double diff = 0.0;
  for (std::size_t k = 0; k < n; ++k) diff += std::norm(rot * std::complex<double>(moved[k]) - std::complex<double>(ref[k]));
  return static_cast<float>(std::sqrt(diff / std::max(ref_norm, std::numeric_limits<double>::min())));
}

}  // namespace

float aligned_error(std::span<const Complex> a, std::span<const Complex> reference, const Shape& shape) {
  if (shape.size() != 2 && shape.size() != 3) throw std::invalid_argument("shape must be 2D or 3D");
  if (a.size() != numel(shape) || reference.size() != a.size()) {
    throw std::invalid_argument("arrays do not match shape");
  }
  const auto d = detail::dims3(shape);
  std::vector<Complex> direct(a.begin(), a.end());
  std::vector<Complex> twin(a.size());  // conj(a(-r))
  std::size_t i = 0;
  for (std::size_t z = 0; z < d[0]; ++z)
    for (std::size_t y = 0; y < d[1]; ++y)
      for (std::size_t x = 0; x < d[2]; ++x, ++i) {
        twin[i] = std::conj(a[(((d[0] - z) % d[0]) * d[1] + (d[1] - y) % d[1]) * d[2] + (d[2] - x) % d[2]]);
      }
  return std::min(translated_error(direct, reference, shape), translated_error(twin, reference, shape));
}

}  // namespace qpcu
