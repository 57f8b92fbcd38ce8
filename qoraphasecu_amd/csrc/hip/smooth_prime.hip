// The smooth-length passes (smooth.hpp) whose odd factor R2 holds a prime from 7 to 23: 7,
// 11, 13, 17, 19 and 23 (dft_prime), 49 (Cooley-Tukey) and 21, 33, 35, 39 and 63 (prime
// factor forms), in a translation unit of their own so that it compiles beside
// smooth.hip.  Without them these lengths run the chirp-z passes (chirp.hpp), whose
// transforms are at least 256 points long.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {

// 7 with R1 up to 256 (from 32 on, the line engine at radix 8), 11 and 13 with R1 up to
// 128 (radix 16), 17 and 19 with R1 up to 16, 21, 23, 33, 35, 39 and 49 with R1 8 and 16,
// 63 with R1 16.
#define QPCH_FOR_EACH_SMOOTH_PRIME(X)                                                              \
  X(2, 7) X(4, 7) X(8, 7) X(16, 7) X(32, 7) X(64, 7) X(128, 7) X(256, 7) X(2, 11) X(4, 11) X(8, 11) \
  X(16, 11) X(32, 11) X(64, 11) X(128, 11) X(2, 13) X(4, 13) X(8, 13) X(16, 13) X(32, 13) X(64, 13) \
  X(128, 13) X(8, 21) X(16, 21) X(8, 33) X(16, 33) X(8, 35) X(16, 35) X(8, 39) X(16, 39) X(2, 17) X(4, 17) \
  X(8, 17) X(16, 17) X(2, 19) X(4, 19) X(8, 19) X(16, 19) X(8, 23) X(16, 23) X(8, 49) X(16, 49) X(16, 63)

// Square images with a persistent launch: 14, 28, 56, 22, 44, 26 and 52 (88 = 8 x 11,
// 104 = 8 x 13 and the radix-17 and -19 squares held 256 VGPRs and AGPRs besides in the
// loop over both passes).
#define QPCH_FOR_EACH_SMOOTH_PRIME_PERSISTENT(X) X(2, 7) X(4, 7) X(8, 7) X(2, 11) X(4, 11) X(2, 13) X(4, 13)

hipError_t twopass_smooth_prime_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                       const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return smooth_detail::launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_PRIME_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

hipError_t smooth_prime_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                               dim3 grid, hipStream_t s) {
  if (m == dev::Mode::kGrad) return hipErrorInvalidValue;
#define QPCH_SMOOTH_CASE(A, B)                                   \
  if (r1 == A && r2 == B) {                                      \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();                                    \
  }
  QPCH_FOR_EACH_SMOOTH_PRIME(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

}  // namespace qpch
