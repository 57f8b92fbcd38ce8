// The smooth-length passes (smooth.hpp) for lengths with a prime from 7 to 13 split as R1 R2
// with R1 not a power of two (42 = 6 x 7, 196 = 28 x 7, 441 = 21 x 21, ...), which would
// otherwise run the chirp-z passes, in a translation unit of their own.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

#define QPCH_FOR_EACH_SMOOTH_MIXED7(X) \
  X(6, 7) X(14, 7) X(12, 11) X(28, 7) X(14, 15) X(28, 9) X(35, 9) X(30, 13) X(28, 15) X(21, 21) \
  X(42, 11) X(28, 25) X(12, 7) X(10, 7) X(11, 6) X(13, 6) X(11, 10) X(14, 9) X(13, 10) X(14, 10) \
  X(14, 11) X(13, 12) X(14, 13) X(18, 11) X(20, 11) X(18, 13) X(20, 13) X(22, 14) X(26, 14) X(22, 20) \
  X(26, 18) X(24, 21) X(26, 20) X(28, 21) X(28, 22) X(30, 22) X(28, 24) X(28, 26) X(30, 26) X(30, 28) \
  X(40, 22) X(33, 28) X(35, 28)

// Square images with a persistent launch: 42 and 70 (66, 78 and 84 held 256 VGPRs and
// AGPRs besides in the loop over both passes).
#define QPCH_FOR_EACH_SMOOTH_MIXED7_PERSISTENT(X) X(6, 7) X(10, 7)

hipError_t twopass_smooth_mixed7_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                        const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return smooth_detail::launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_MIXED7_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

hipError_t smooth_mixed7_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                dim3 grid, hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_MIXED7(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return smooth_mixed13_launch(r1, r2, col, m, a, lines, grid, s);  // smooth_mixed13.hip
}

}  // namespace qpch
