// The smooth-length passes (smooth.hpp), instantiated per factor pair, axis kind and mode.
#include "smooth_pass.hpp"

#include "smooth.hpp"

#include <algorithm>

namespace qpch {
namespace {

// The factor pairs: R1 a power of two from 2 to 256 times R2 in {3, 5, 15}, up to 2048,
// except 64 x 3 and 512 x 3.  Those two (192 and 1536) measured slower than the generic
// kernel: layout A keeps only 3 of every 16 threads of a line busy through the radix-R1
// stage (192^2 x 8 starts: 25.2 against 22.4 us per iteration; 1536^2 x 2: 294 against
// 236).
#define QPCH_FOR_EACH_SMOOTH(X)                                                                  \
  X(2, 3) X(2, 5) X(2, 15) X(4, 3) X(4, 5) X(4, 15) X(8, 3) X(8, 5) X(8, 15) X(16, 3) X(16, 5)   \
  X(16, 15) X(32, 3) X(32, 5) X(32, 15) X(64, 5) X(64, 15) X(128, 3) X(128, 5) X(128, 15)        \
  X(256, 3) X(256, 5)
// Square images with a persistent launch: R1 up to 16, and 32 x 5.  Above about 256 points
// the per-pass launches measured faster (480^2 x 8: 39 against 64 us per iteration; 160^2
// x 8: 14.0 against 11.1 the other way), and the persistent 32 x 3 kernel spills.
#define QPCH_FOR_EACH_SMOOTH_PERSISTENT(X)                                                       \
  X(2, 3) X(2, 5) X(2, 15) X(4, 3) X(4, 5) X(4, 15) X(8, 3) X(8, 5) X(8, 15) X(16, 3) X(16, 5)   \
  X(16, 15) X(32, 5)

bool persistent_pair(int r1, int r2) { return r1 <= 16 || (r1 == 32 && r2 == 5); }

bool factors(int N, int& r1, int& r2) {
  for (int r : {15, 5, 3}) {
    if (N % r) continue;
    const int q = N / r;
    if (q >= 2 && q <= 256 && (q & (q - 1)) == 0 && N <= dev::kGenericMaxN && !(r == 3 && q == 64)) {
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

template <int R1, int R2>
void launch_pair(bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid, hipStream_t s) {
  if (col) {
    if (m == dev::Mode::kModulus) {
      launch<R1, R2, true, dev::Mode::kModulus>(a, lines, grid, s);
    } else {
      launch<R1, R2, true, dev::Mode::kUpdate>(a, lines, grid, s);
    }
  } else {
    if (m == dev::Mode::kModulus) {
      launch<R1, R2, false, dev::Mode::kModulus>(a, lines, grid, s);
    } else {
      launch<R1, R2, false, dev::Mode::kUpdate>(a, lines, grid, s);
    }
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
  if (!factors(N, r1, r2) || (m != dev::Mode::kModulus && m != dev::Mode::kUpdate)) return hipErrorInvalidValue;
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
