// The smooth-length passes (smooth.hpp) for lengths from 525 to 1014 with primes up to 13
// (525 = 25 x 21, 676 = 26 x 26, 1014 = 39 x 26, ...), split as R1 R2 with R1 not a power
// of two; without them they run the chirp-z passes.  A translation unit of their own, so
// that it compiles beside the others.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

#define QPCH_FOR_EACH_SMOOTH_MIXED1K(X) \
  X(25, 21) X(26, 21) X(25, 22) X(27, 21) X(26, 22) X(39, 15) X(27, 22) X(30, 21) X(26, 25) X(26, 26) \
  X(33, 21) X(27, 26) X(33, 22) X(35, 21) X(28, 27) X(35, 22) X(33, 24) X(39, 21) X(33, 25) X(33, 26) \
  X(35, 25) X(42, 21) X(33, 27) X(35, 26) X(36, 26) X(35, 27) X(44, 22) X(39, 25) X(33, 30) X(39, 26)

hipError_t smooth_mixed1k_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                 dim3 grid, hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_MIXED1K(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

}  // namespace qpch
