// The smooth-length passes (smooth.hpp), instantiated per factor pair, axis kind and mode.
#include "smooth_pass.hpp"

#include "smooth.hpp"

#include <algorithm>

namespace qpch {
namespace {

// The factor pairs: R1 a power of two from 2 to 512 times R2 in {3, 5, 15}, up to 2048.
#define QPCH_FOR_EACH_SMOOTH(X)                                                                  \
  X(2, 3) X(2, 5) X(2, 15) X(4, 3) X(4, 5) X(4, 15) X(8, 3) X(8, 5) X(8, 15) X(16, 3) X(16, 5)   \
  X(16, 15) X(32, 3) X(32, 5) X(32, 15) X(64, 3) X(64, 5) X(64, 15) X(128, 3) X(128, 5) X(128, 15)        \
  X(256, 3) X(256, 5) X(512, 3)
// Square images with a persistent launch: R1 up to 16, and 96 = 32 x 3 and 160 = 32 x 5
// (us per iteration of 8 starts, per pass against persistent: 96^2 10.3 / 7.8, 160^2
// 11.8 / 8.4, 240^2 16.9 / 15.2).  The others measured faster per pass: 192^2 13.8 / 16.8,
// 480^2 39 / 64.
#define QPCH_FOR_EACH_SMOOTH_PERSISTENT(X)                                                       \
  X(2, 3) X(2, 5) X(2, 15) X(4, 3) X(4, 5) X(4, 15) X(8, 3) X(8, 5) X(8, 15) X(16, 3) X(16, 5)   \
  X(16, 15) X(32, 3) X(32, 5)

bool persistent_pair(int r1, int r2) { return r1 <= 16 || (r1 == 32 && r2 <= 5); }

bool factors(int N, int& r1, int& r2) {
  for (int r : {15, 5, 3}) {
    if (N % r) continue;
    const int q = N / r;
    if (q >= 2 && q <= 512 && (q & (q - 1)) == 0 && N <= dev::kGenericMaxN) {
      r1 = q;
      r2 = r;
      return true;
    }
  }
  return false;
}

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

template <int R1, int R2>
void launch_pair(bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid, hipStream_t s) {
  if (col) {
    launch_mode<R1, R2, true>(m, a, lines, grid, s);
  } else {
    launch_mode<R1, R2, false>(m, a, lines, grid, s);
  }
}

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

}  // namespace

bool twopass_smooth_supported(int ny, int nx) {
  int r1 = 0, r2 = 0;
  return ny == nx && factors(ny, r1, r2) && persistent_pair(r1, r2);
}

hipError_t twopass_smooth_launch(int N, const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::SyncArgs& q,
                                 int len, int workers, hipStream_t s) {
  int r1 = 0, r2 = 0;
  if (!factors(N, r1, r2) || !persistent_pair(r1, r2)) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

bool smooth_axis(int N) {
  int r1 = 0, r2 = 0;
  return factors(N, r1, r2);
}

hipError_t smooth_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                         hipStream_t s) {
  int r1 = 0, r2 = 0;
  if (!factors(N, r1, r2) || m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) {    \
    launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();  \
  }
  QPCH_FOR_EACH_SMOOTH(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

}  // namespace qpch
