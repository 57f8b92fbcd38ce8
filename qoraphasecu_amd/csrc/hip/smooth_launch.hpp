// Launch helpers of smooth_pass (smooth.hpp), shared by the translation units that
// instantiate it (smooth.hip: R2 in {3, 5, 15}; smooth_sq.hip: R2 in {9, 25, 27, 45}).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

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

// twopass_smooth<R1, R2> for `len` iterations with `workers` work-groups (0: as many per XCD
// as the busiest XCD has tiles per pass, at most what fits the device at once).
template <int R1, int R2>
hipError_t launch_twopass(const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::SyncArgs& q, int len,
                          int workers, hipStream_t s) {
  constexpr int kThreads = dev::SmoothGeom<R1, R2, true>::kThreads;
  int per_cu = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::twopass_smooth<R1, R2>, kThreads, 0);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const int cap = std::max(1, per_cu) * cus / q.xcds * q.xcds;
  const int items = q.xcds * dev::xcd_starts(p1.starts, 0, q.xcds) * std::max(q.tiles0, q.tiles1);
  const int w = workers > 0 ? workers : std::max(q.xcds, std::min(items, cap));
  const long long lines = dev::SmoothGeom<R1, R2, true>::kN;  // per start, both passes (square image)
  hipLaunchKernelGGL((dev::twopass_smooth<R1, R2>), dim3(w), dim3(kThreads), 0, s, p0, p1, lines, lines, q, len);
  return hipGetLastError();
}

}  // namespace qpch::smooth_detail
