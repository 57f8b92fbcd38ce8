// The smooth-length passes (smooth.hpp), instantiated per factor pair, axis kind and mode.
#include "smooth_pass.hpp"

#include "smooth.hpp"
#include "smooth_launch.hpp"

namespace qpch {
namespace {

// The factor pairs: R1 a power of two from 2 to 512 times R2 in {3, 5, 15}, up to 2048
// (R2 in {9, 25, 27, 45}: smooth_sq.hip; 7, 11, 13, 21, 33, 35, 39: smooth_prime.hip).
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

bool persistent_pair(int r1, int r2) {
  if (!dev::smooth_is_pow2(r1) || r2 % 2 == 0) {  // 45, 50, 54, 100 (smooth_mixed.hip), 42, 70 (smooth_mixed7.hip)
    return (r1 == 9 && r2 == 5) || (r1 == 10 && r2 == 5) || (r1 == 6 && r2 == 9) || (r1 == 10 && r2 == 10) ||
           (r1 == 6 && r2 == 7) || (r1 == 10 && r2 == 7);
  }
  if (r2 == 3 || r2 == 5 || r2 == 15) return r1 <= 16 || (r1 == 32 && r2 <= 5);
  if (r2 == 7) return r1 <= 8;  // 14, 28, 56 (smooth_prime.hip)
  if (r2 == 11 || r2 == 13) return r1 <= 4;  // 22, 44, 26, 52
  return r2 == 9 && r1 <= 8;  // 18, 36 and 72 (smooth_sq.hip)
}

// N = R1 R2 with R2 its odd part: a pair with a kernel.  R2 in {3, 5, 15}: R1 from 2 to
// 512; 7: R1 from 2 to 256; 9, 11 and 13: R1 from 2 to 128; 17 and 19: R1 from 2 to 16;
// 21, 23, 25, 27, 33, 35, 39 and 49: R1 8 and 16; 45 and 63: R1 16.  The pairs left out ran
// slower than generic_pass (us per iteration of 8 starts, generic / smooth: 100^2 16.1 /
// 20.1, 108^2 19.2 / 20.9, 360^2 53.3 / 55.2; R2 = 75 at every R1, 600^2 131 / 318) or than
// chirp_pass (46^2 20.4 / 22.2, 92^2 20.4 / 21.6, 504^2 83.5 / 108): with R1 threads of R2
// registers in layout B, a small R1 leaves most of a line's threads idle in the radix-R2
// stage.
// Lengths of 2, 3 and 5 whose power-of-two part is small beside the odd part take a more
// even split, with a factor that is not a power of two (smooth_mixed.hip).
struct MixedPair {
  int n, r1, r2;
};
constexpr MixedPair kMixed[] = {{45, 9, 5},      {50, 10, 5},     {54, 6, 9},      {75, 15, 5},     {90, 6, 15},
                                {100, 10, 10},   {108, 12, 9},    {125, 25, 5},    {135, 15, 9},    {150, 10, 15},
                                {180, 12, 15},   {225, 15, 15},   {250, 10, 25},   {270, 18, 15},   {300, 20, 15},
                                {360, 24, 15},   {375, 25, 15},   {450, 30, 15},   {500, 20, 25},   {540, 36, 15},
                                {600, 24, 25},   {675, 27, 25},   {750, 30, 25},   {900, 36, 25},   {1000, 40, 25},
                                {1125, 45, 25},  {1200, 48, 25},
                                {81, 9, 9},      {162, 18, 9},    {324, 18, 18},   {648, 36, 18},   {729, 27, 27},
                                {972, 36, 27},   {1296, 36, 36},  {625, 25, 25},   {1350, 45, 30},  {1440, 48, 30},
                                {1800, 45, 40},  {2025, 45, 45},
                                // lengths with a prime from 7 to 13 (otherwise chirp_pass)
                                {42, 6, 7},      {98, 14, 7},     {132, 12, 11},   {196, 28, 7},    {210, 14, 15},
                                {252, 28, 9},    {315, 35, 9},    {390, 30, 13},   {420, 28, 15},   {441, 21, 21},
                                {462, 42, 11},   {700, 28, 25},   {84, 12, 7},     {70, 10, 7},     {66, 11, 6},
                                {78, 13, 6},     {110, 11, 10},   {126, 14, 9},    {130, 13, 10},   {140, 14, 10},
                                {154, 14, 11},   {156, 13, 12},   {182, 14, 13},   {198, 18, 11},   {220, 20, 11},
                                {234, 18, 13},   {260, 20, 13},   {308, 22, 14},   {364, 26, 14},   {440, 22, 20},
                                {468, 26, 18},   {504, 24, 21},   {520, 26, 20},   {588, 28, 21},   {616, 28, 22},
                                {660, 30, 22},   {672, 28, 24},   {728, 28, 26},   {780, 30, 26},   {840, 30, 28},
                                {880, 40, 22},   {924, 33, 28},   {980, 35, 28},
                                // more lengths with primes up to 13, odd ones among them (smooth_mixed13.hip)
                                {63, 9, 7},      {77, 11, 7},     {91, 13, 7},     {99, 11, 9},     {105, 15, 7},
                                {117, 13, 9},    {121, 11, 11},   {143, 13, 11},   {147, 21, 7},    {165, 15, 11},
                                {169, 13, 13},   {175, 25, 7},    {189, 21, 9},    {195, 15, 13},   {231, 21, 11},
                                {242, 22, 11},   {245, 35, 7},    {273, 21, 13},   {275, 25, 11},   {286, 22, 13},
                                {294, 21, 14},   {297, 27, 11},   {325, 25, 13},   {330, 22, 15},   {338, 26, 13},
                                {350, 25, 14},   {351, 27, 13},   {363, 33, 11},   {378, 21, 18},   {385, 35, 11},
                                {396, 22, 18},   {429, 33, 13},   {455, 35, 13},   {484, 22, 22},   {490, 35, 14},
                                {495, 33, 15},   {507, 39, 13},
                                // 525 to 1014 with primes up to 13 (smooth_mixed1k.hip)
                                {525, 25, 21},   {546, 26, 21},   {550, 25, 22},   {567, 27, 21},   {572, 26, 22},
                                {585, 39, 15},   {594, 27, 22},   {630, 30, 21},   {650, 26, 25},   {676, 26, 26},
                                {693, 33, 21},   {702, 27, 26},   {726, 33, 22},   {735, 35, 21},   {756, 28, 27},
                                {770, 35, 22},   {792, 33, 24},   {819, 39, 21},   {825, 33, 25},   {858, 33, 26},
                                {875, 35, 25},   {882, 42, 21},   {891, 33, 27},   {910, 35, 26},   {936, 36, 26},
                                {945, 35, 27},   {968, 44, 22},   {975, 39, 25},   {990, 33, 30},   {1014, 39, 26}};

bool factors(int N, int& r1, int& r2) {
  if (N < 2 || N > dev::kGenericMaxN) return false;
  for (const MixedPair& m : kMixed) {
    if (m.n == N) {
      r1 = m.r1;
      r2 = m.r2;
      return true;
    }
  }
  int q = 1;
  while ((N / q) % 2 == 0) q *= 2;
  const int r = N / q;
  if (q < 2) return false;
  const bool ok = r == 3 || r == 5 || r == 15 || r == 7 || ((r == 9 || r == 11 || r == 13) && q <= 128) ||
                  ((r == 17 || r == 19) && q <= 16) ||
                  ((r == 25 || r == 27 || r == 21 || r == 23 || r == 33 || r == 35 || r == 39 || r == 49) &&
                   (q == 8 || q == 16)) ||
                  ((r == 45 || r == 63) && q == 16);
  if (!ok) return false;
  r1 = q;
  r2 = r;
  return true;
}

bool sq_factor(int r2) { return r2 == 9 || r2 == 25 || r2 == 27 || r2 == 45; }
bool prime_factor(int r2) {
  return r2 == 7 || r2 == 11 || r2 == 13 || r2 == 17 || r2 == 19 || r2 == 23 || r2 == 21 || r2 == 33 || r2 == 35 ||
         r2 == 39 || r2 == 49 || r2 == 63;
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
  if (!dev::smooth_is_pow2(r1) || r2 % 2 == 0) return twopass_smooth_mixed_launch(r1, r2, p0, p1, q, len, workers, s);
  if (sq_factor(r2)) return twopass_smooth_sq_launch(r1, r2, p0, p1, q, len, workers, s);
  if (prime_factor(r2)) return twopass_smooth_prime_launch(r1, r2, p0, p1, q, len, workers, s);
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) return smooth_detail::launch_twopass<A, B>(p0, p1, q, len, workers, s);
  QPCH_FOR_EACH_SMOOTH_PERSISTENT(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

bool smooth_axis(int N) {
  int r1 = 0, r2 = 0;
  return factors(N, r1, r2);
}

int smooth_lines(int N) {
  int r1 = 0, r2 = 0;
  return factors(N, r1, r2) ? dev::smooth_lines_per_block(r1, r2) : 0;
}

hipError_t smooth_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                         hipStream_t s) {
  int r1 = 0, r2 = 0;
  if (!factors(N, r1, r2) || m == dev::Mode::kGrad) return hipErrorInvalidValue;
  if (!dev::smooth_is_pow2(r1) || r2 % 2 == 0) return smooth_mixed_launch(r1, r2, col, m, a, lines, grid, s);
  if (sq_factor(r2)) return smooth_sq_launch(r1, r2, col, m, a, lines, grid, s);
  if (prime_factor(r2)) return smooth_prime_launch(r1, r2, col, m, a, lines, grid, s);
#define QPCH_SMOOTH_CASE(A, B) \
  if (r1 == A && r2 == B) {    \
    smooth_detail::launch_pair<A, B>(col, m, a, lines, grid, s); \
    return hipGetLastError();  \
  }
  QPCH_FOR_EACH_SMOOTH(QPCH_SMOOTH_CASE)
#undef QPCH_SMOOTH_CASE
  return hipErrorInvalidValue;
}

}  // namespace qpch
