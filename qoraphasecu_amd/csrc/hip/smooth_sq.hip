// The smooth-length passes (smooth.hpp) whose odd factor R2 holds a squared prime: 9, 25,
// 27 and 45 (lengths such as 36, 144, 200, 216, 576, 720), in a translation unit of their
// own so that it compiles beside smooth.hip.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

// 9 with R1 up to 128 (from 32 on, the line engine); 25 and 27 with R1 8 and 16; 45 with
// R1 16 (the pairs that ran faster than generic_pass).
#define QPCH_FOR_EACH_SMOOTH_SQ(X)                                                       \
  X(2, 9) X(4, 9) X(8, 9) X(16, 9) X(32, 9) X(64, 9) X(128, 9) X(8, 25) X(16, 25) X(8, 27) \
  X(16, 27) X(16, 45)

// Square images with a persistent launch: 18, 36 and 72.  From 144 = 16 x 9 on, the loop
// over both passes' tiles held 256 VGPRs and AGPRs besides (occupancy 1), and 400, 432
// and 720 spilled.
#define QPCH_FOR_EACH_SMOOTH_SQ_PERSISTENT(X) X(2, 9) X(4, 9) X(8, 9)

hipError_t twopass_smooth_sq_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                    const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return smooth_detail::launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_SQ_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

hipError_t smooth_sq_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                            hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_SQ(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

}  // namespace qpch
