// Array orderings of the host API.  Public arrays are centre-origin (index n / 2 of each
// axis is the origin); the engine works in FFT order (index 0 is the origin).  Moving
// between them is numpy's ifftshift / fftshift, the convention of the reference's
// src/core/layout.hpp.
#pragma once

#include <array>
#include <cstddef>
#include <span>
#include <vector>

#include "qoraphasecu/problem.hpp"

namespace qpcu::detail {

// Extents as (planes, rows, columns): a 2D shape has one plane.
inline std::array<std::size_t, 3> dims3(const Shape& shape) {
  return shape.size() == 2 ? std::array<std::size_t, 3>{1, shape[0], shape[1]}
                           : std::array<std::size_t, 3>{shape[0], shape[1], shape[2]};
}

// For one axis of length `len`: the input coordinate each output coordinate reads.
// Into FFT order output i reads input (i + len / 2) mod len; back to centre-origin it
// reads (i + len - len / 2) mod len (the two differ on odd lengths).
inline std::vector<std::size_t> source_coordinates(std::size_t len, bool into_fft_order) {
  const std::size_t lead = into_fft_order ? len / 2 : len - len / 2;
  std::vector<std::size_t> src(len);
  for (std::size_t i = 0; i < len; ++i) src[i] = i + lead < len ? i + lead : i + lead - len;
  return src;
}

template <class T>
std::vector<T> reorder(std::span<const T> in, const Shape& shape, bool into_fft_order) {
  const auto ext = dims3(shape);
  const auto planes = source_coordinates(ext[0], into_fft_order);
  const auto rows = source_coordinates(ext[1], into_fft_order);
  const auto cols = source_coordinates(ext[2], into_fft_order);
  std::vector<T> out;
  out.reserve(in.size());
  for (std::size_t p : planes) {
    for (std::size_t r : rows) {
      const T* line = in.data() + (p * ext[1] + r) * ext[2];
      for (std::size_t c : cols) out.push_back(line[c]);
    }
  }
  return out;
}

template <class T>
std::vector<T> to_fft_order(std::span<const T> in, const Shape& shape) {
  return reorder(in, shape, true);
}

template <class T>
std::vector<T> to_centered(std::span<const T> in, const Shape& shape) {
  return reorder(in, shape, false);
}

}  // namespace qpcu::detail
