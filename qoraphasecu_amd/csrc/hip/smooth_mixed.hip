// The smooth-length passes (smooth.hpp) for lengths of 2, 3 and 5 whose power-of-two part
// is small beside the odd part (45, 100, 180, 360, 1000, ...): N = R1 R2 with R1 not a
// power of two (or R2 even), so that both layouts keep most of a line's threads busy
// (100 = 10 x 10 instead of 4 x 25), in a translation unit of their own.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

#define QPCH_FOR_EACH_SMOOTH_MIXED(X)                                                                  \
  X(9, 5) X(10, 5) X(6, 9) X(15, 5) X(6, 15) X(10, 10) X(12, 9) X(25, 5) X(15, 9) X(10, 15) X(12, 15)  \
  X(15, 15) X(10, 25) X(18, 15) X(20, 15) X(24, 15) X(25, 15) X(30, 15) X(20, 25) X(36, 15) X(24, 25) \
  X(27, 25) X(30, 25) X(36, 25) X(40, 25) X(45, 25) X(48, 25) X(9, 9) X(18, 9) X(18, 18) X(36, 18) X(27, 27)  \
  X(36, 27) X(36, 36) X(25, 25) X(45, 30) X(48, 30) X(45, 40) X(45, 45)

// Square images with a persistent launch: 45, 50, 54 and 100 (from 75 = 15 x 5 on, the
// others held 256 VGPRs and AGPRs besides in the loop over both passes).
#define QPCH_FOR_EACH_SMOOTH_MIXED_PERSISTENT(X) X(9, 5) X(10, 5) X(6, 9) X(10, 10)

hipError_t twopass_smooth_mixed_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                       const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return smooth_detail::launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_MIXED_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return twopass_smooth_mixed7_launch(r1, r2, p0, p1, q, len, workers, s);
}

hipError_t smooth_mixed_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                               dim3 grid, hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_MIXED(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return smooth_mixed7_launch(r1, r2, col, m, a, lines, grid, s);  // the pairs with a prime from 7 to 13
}

}  // namespace qpch
