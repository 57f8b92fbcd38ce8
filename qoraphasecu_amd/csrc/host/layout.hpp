// Centre-origin <-> FFT-order reordering and the 3-axis view of a shape.
// Reference semantics: src/core/layout.hpp:12-47 (numpy ifftshift / fftshift).
#pragma once

#include <array>
#include <cstddef>
#include <span>
#include <vector>

#include "qoraphasecu/problem.hpp"

namespace qpcu::detail {

// {1, ny, nx} for 2D, {nz, ny, nx} for 3D.
inline std::array<std::size_t, 3> dims3(const Shape& shape) {
  if (shape.size() == 2) return {1, shape[0], shape[1]};
  return {shape[0], shape[1], shape[2]};
}

// out[i] = in[(i + s) mod n] along each axis, s = n / 2 into FFT order and n - n / 2 back.
template <class T>
std::vector<T> shifted(std::span<const T> in, const Shape& shape, bool into_fft_order) {
  const auto d = dims3(shape);
  std::array<std::size_t, 3> s{};
  for (int a = 0; a < 3; ++a) s[a] = into_fft_order ? d[a] / 2 : d[a] - d[a] / 2;
  std::vector<T> out(in.size());
  std::size_t o = 0;
  for (std::size_t z = 0; z < d[0]; ++z) {
    const std::size_t iz = (z + s[0]) % d[0];
    for (std::size_t y = 0; y < d[1]; ++y) {
      const std::size_t row = (iz * d[1] + (y + s[1]) % d[1]) * d[2];
      for (std::size_t x = 0; x < d[2]; ++x) out[o++] = in[row + (x + s[2]) % d[2]];
    }
  }
  return out;
}

template <class T>
std::vector<T> to_fft_order(std::span<const T> in, const Shape& shape) {
  return shifted(in, shape, true);
}

template <class T>
std::vector<T> to_centered(std::span<const T> in, const Shape& shape) {
  return shifted(in, shape, false);
}

}  // namespace qpcu::detail
