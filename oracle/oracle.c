/*
 * CPU oracle for the reconstruction path: TEST INFRASTRUCTURE, never part of the
 * product.  See oracle.h for what it restates; oracle_impl.h holds the type-generic
 * body, instantiated here for float and double.
 */
#include "oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846
/* kTwoPi, src/core/pointwise.hpp:17 */
#define OR_TWO_PI 6.28318530717958647692
/* kEps, src/core/pointwise.hpp:16 */
#define OR_EPS 1e-12
#define OR_MAXR 64
#define OR_V 8 /* lines per SIMD block of the FFT */
#define OR_BLOCK 4096

static int g_threads = 0;

void or_set_threads(int threads) { g_threads = threads; }

static int or_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* src/core/pointwise.hpp:19-24 (splitmix64 finaliser) */
uint64_t or_mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* src/core/pointwise.hpp:29-32: top 24 bits of a keyed hash, times 2^-24 */
float or_uniform(uint64_t seed, uint64_t start, uint64_t index) {
  const uint64_t key = or_mix64(seed ^ or_mix64(start + 1));
  return (float)(or_mix64(key + index) >> 40) * 0x1p-24f;
}

/* src/core/layout.hpp:13-16 */
static void or_dims3(int ndim, const size_t* shape, size_t* d) {
  if (ndim == 2) {
    d[0] = 1;
    d[1] = shape[0];
    d[2] = shape[1];
  } else {
    d[0] = shape[0];
    d[1] = shape[1];
    d[2] = shape[2];
  }
}

/* src/core/layout.hpp:19-35: out[i] = in[(i + off) mod n] per axis, off = n/2 towards FFT
 * order (ifftshift) and n - n/2 back (fftshift). */
void or_roll(const void* in, void* out, size_t elem, int ndim, const size_t* shape,
             int to_fft_order) {
  size_t d[3], off[3];
  or_dims3(ndim, shape, d);
  for (int a = 0; a < 3; ++a) off[a] = to_fft_order ? d[a] / 2 : d[a] - d[a] / 2;
  const unsigned char* src = (const unsigned char*)in;
  unsigned char* dst = (unsigned char*)out;
  size_t i = 0;
  for (size_t z = 0; z < d[0]; ++z) {
    const size_t sz = (z + off[0]) % d[0];
    for (size_t y = 0; y < d[1]; ++y) {
      const size_t sy = (y + off[1]) % d[1];
      const size_t row = (sz * d[1] + sy) * d[2];
      for (size_t x = 0; x < d[2]; ++x) {
        memcpy(dst + elem * i++, src + elem * (row + (x + off[2]) % d[2]), elem);
      }
    }
  }
}

/* src/core/problem.cpp:101-127: keep the support strictly below its centre of mass along
 * `axis` (centre-origin input). */
void or_half_support(const uint8_t* support, uint8_t* out, int ndim, const size_t* shape,
                     int axis) {
  size_t inner = 1, n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
  for (int a = axis + 1; a < ndim; ++a) inner *= shape[a];
  const size_t len = shape[axis];
  double moment = 0.0, weight = 0.0;
  for (size_t i = 0; i < n; ++i) {
    if (support[i]) {
      moment += (double)((i / inner) % len);
      weight += 1.0;
    }
  }
  const double centre = moment / (weight > 1.0 ? weight : 1.0);
  for (size_t i = 0; i < n; ++i) out[i] = support[i] && (double)((i / inner) % len) < centre;
}

#define R float
#define S f
#define OR_ROUND_ERROR(v) ((double)(float)(v))
#include "oracle_impl.h"
#undef R
#undef S
#undef OR_ROUND_ERROR

#define R double
#define S d
#define OR_ROUND_ERROR(v) (v)
#include "oracle_impl.h"
#undef R
#undef S
#undef OR_ROUND_ERROR
