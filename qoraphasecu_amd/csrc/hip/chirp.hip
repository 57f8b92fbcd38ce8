// The chirp-z passes (chirp.hpp), instantiated per transform length M, axis kind and mode.
#include "chirp_pass.hpp"

#include "chirp.hpp"
#include "smooth_pass.hpp"

namespace qpch {
namespace {

template <int M, bool COL, dev::Mode MODE>
void launch(const dev::PassArgs& a, long long lines, int n, dim3 grid, const float2* chirp, const float2* bhat,
            hipStream_t s) {
  hipLaunchKernelGGL((dev::chirp_pass<M, COL, MODE>), grid, dim3(dev::ChirpGeom<M, COL>::kThreads), 0, s, a, lines,
                     n, chirp, bhat);
}

template <int M, bool COL>
void launch_mode(dev::Mode m, const dev::PassArgs& a, long long lines, int n, dim3 grid, const float2* chirp,
                 const float2* bhat, hipStream_t s) {
  switch (m) {
    case dev::Mode::kFwd: return launch<M, COL, dev::Mode::kFwd>(a, lines, n, grid, chirp, bhat, s);
    case dev::Mode::kInv: return launch<M, COL, dev::Mode::kInv>(a, lines, n, grid, chirp, bhat, s);
    case dev::Mode::kModulus: return launch<M, COL, dev::Mode::kModulus>(a, lines, n, grid, chirp, bhat, s);
    default: return launch<M, COL, dev::Mode::kUpdate>(a, lines, n, grid, chirp, bhat, s);
  }
}

template <int M>
void launch_m(bool col, dev::Mode m, const dev::PassArgs& a, long long lines, int n, dim3 grid, const float2* chirp,
              const float2* bhat, hipStream_t s) {
  if (col) {
    launch_mode<M, true>(m, a, lines, n, grid, chirp, bhat, s);
  } else {
    launch_mode<M, false>(m, a, lines, n, grid, chirp, bhat, s);
  }
}

}  // namespace

int chirp_length(int N) {
  int m = 256;  // the kernels' smallest transform (a longer one serves as well)
  while (m < 2 * N - 1) m *= 2;
  return m;
}

bool chirp_axis(int N) {
  if (N < 32 || N > dev::kGenericMaxN || (N & (N - 1)) == 0 || smooth_axis(N)) return false;
  int rest = N;
  for (int f : {2, 3, 5}) {
    while (rest % f == 0) rest /= f;
  }
  // a prime factor of 7 or more: the generic kernel's stage for it is an O(r^2) sum (above 64
  // a plain DFT per line, or Bluestein through LDS).  Lengths of 2, 3 and 5 alone keep the
  // generic kernel's short forms, which measured faster (100^2 x 8: 16.1 against 20.8 us per
  // iteration; 1000 x 64: 24.2 against 25.8), as do lengths below 32.
  return rest > 1;
}

int chirp_lines(int N) { return dev::chirp_lines_per_block(chirp_length(N)); }

hipError_t chirp_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                        const void* chirp, const void* bhat, hipStream_t s) {
  if (!chirp_axis(N) || m == dev::Mode::kGrad) return hipErrorInvalidValue;
  const auto* c = static_cast<const float2*>(chirp);
  const auto* b = static_cast<const float2*>(bhat);
  switch (chirp_length(N)) {
    case 256: launch_m<256>(col, m, a, lines, N, grid, c, b, s); break;
    case 512: launch_m<512>(col, m, a, lines, N, grid, c, b, s); break;
    case 1024: launch_m<1024>(col, m, a, lines, N, grid, c, b, s); break;
    case 2048: launch_m<2048>(col, m, a, lines, N, grid, c, b, s); break;
    case 4096: launch_m<4096>(col, m, a, lines, N, grid, c, b, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qpch
