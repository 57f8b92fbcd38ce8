// The smooth-length passes (smooth.hpp) for further lengths with primes up to 13, odd ones
// among them (63 = 9 x 7, 121 = 11 x 11, 275 = 25 x 11, 484 = 22 x 22, ...), split as
// R1 R2 with R1 not a power of two; without them they run the chirp-z passes.  A
// translation unit of their own, so that it compiles beside the others.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

#define QPCH_FOR_EACH_SMOOTH_MIXED13(X) \
  X(9, 7) X(11, 7) X(13, 7) X(11, 9) X(15, 7) X(13, 9) X(11, 11) X(13, 11) X(21, 7) X(15, 11) \
  X(13, 13) X(25, 7) X(21, 9) X(15, 13) X(21, 11) X(22, 11) X(35, 7) X(21, 13) X(25, 11) X(22, 13) \
  X(21, 14) X(27, 11) X(25, 13) X(22, 15) X(26, 13) X(25, 14) X(27, 13) X(33, 11) X(21, 18) X(35, 11) \
  X(22, 18) X(33, 13) X(35, 13) X(22, 22) X(35, 14) X(33, 15) X(39, 13)

hipError_t smooth_mixed13_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                 dim3 grid, hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_MIXED13(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return smooth_mixed1k_launch(r1, r2, col, m, a, lines, grid, s);  // smooth_mixed1k.hip
}

}  // namespace qpch
