// Launch helpers of smooth_pass (smooth.hpp), shared by the translation units that
// instantiate it (smooth.hip: R2 in {3, 5, 15}; smooth_sq.hip: R2 in {9, 25, 27, 45}).
#pragma once

#include <hip/hip_runtime.h>

#include "smooth.hpp"

namespace qpch::smooth_detail {

template <int R1, int R2, bool COL, dev::Mode M>
void launch(const dev::PassArgs& a, long long lines, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((dev::smooth_pass<R1, R2, COL, M>), grid, dim3(dev::SmoothGeom<R1, R2, COL>::kThreads), 0, s,
                     a, lines);
}

template <int R1, int R2, bool COL>
void launch_mode(dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid, hipStream_t s) {
  switch (m) {
    case dev::Mode::kFwd: return launch<R1, R2, COL, dev::Mode::kFwd>(a, lines, grid, s);
    case dev::Mode::kInv: return launch<R1, R2, COL, dev::Mode::kInv>(a, lines, grid, s);
    case dev::Mode::kModulus: return launch<R1, R2, COL, dev::Mode::kModulus>(a, lines, grid, s);
    default: return launch<R1, R2, COL, dev::Mode::kUpdate>(a, lines, grid, s);
  }
}

// smooth_pass<R1, R2> for a strided (col) or row axis in mode m (not Mode::kGrad).
template <int R1, int R2>
void launch_pair(bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid, hipStream_t s) {
  if (col) {
    launch_mode<R1, R2, true>(m, a, lines, grid, s);
  } else {
    launch_mode<R1, R2, false>(m, a, lines, grid, s);
  }
}

}  // namespace qpch::smooth_detail
