// Set-up, conversion and Shrinkwrap kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "passes.hpp"
#include "storage.hpp"

namespace qpch::dev {

// splitmix64 finaliser and the counter-based uniform draw of the reference
// (src/core/pointwise.hpp:19-32): any element of any start can be drawn on its own.
__device__ __forceinline__ std::uint64_t mix64(std::uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float uniform01(std::uint64_t seed, std::uint64_t start,
                                           std::uint64_t index) {
  const std::uint64_t key = mix64(seed ^ mix64(start + 1));
  return static_cast<float>(mix64(key + index) >> 40) * 0x1p-24f;
}

// w = A e^{i 2 pi u}: the spectrum of a random start before its unitary inverse
// transform (src/cpu/kernels.cpp:104-111).
// Start g of the launch belongs to problem (first + g) / per_problem and is that
// problem's start (first + g) % per_problem; each problem has its own amplitude row.
__global__ void random_wave(float2* w, const float* amp, long long n, std::uint64_t seed,
                            std::uint64_t first_start, int per_problem) {
  const std::uint64_t global = first_start + blockIdx.y;
  const std::uint64_t start = global % static_cast<std::uint64_t>(per_problem);
  amp += static_cast<long long>(global / static_cast<std::uint64_t>(per_problem)) * n;
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float phase = 6.28318530717958647692f * uniform01(seed, start, static_cast<std::uint64_t>(k));
    const float a = amp[k];
    w[soff + k] = {a * cosf(phase), a * sinf(phase)};
  }
}

// out = scale * in, zeroed outside the support when one is given.
template <class Sin, class Sout>
__global__ void convert(const Sin* in, Sout* out, long long total, float scale,
                        const std::uint8_t* support) {
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < total;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    float2 v = Storage<Sin>::ld(in, k);
    v = {v.x * scale, v.y * scale};
    if (support && !support[k]) v = {0.0f, 0.0f};
    Storage<Sout>::st(out, k, v);
  }
}

__global__ void fill_int(int* p, int count, int value) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
    p[k] = value;
  }
}

__global__ void fill_double(double* p, long long count, double value) {
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < count;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    p[k] = value;
  }
}

__global__ void fill_u64(unsigned long long* p, int count) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) p[k] = 0ull;
}

// --- Detwin: the support cut at its centre of mass along one axis ----------------------
// Reference: src/core/problem.cpp:101-127, applied to the centre-origin support; here the
// support is in FFT order, where index f along the axis is centre-origin coordinate
// (f + len / 2) % len.  The moment and the count are sums of integers, exact in 64-bit
// integers and in the host's doubles alike, so the centre moment / max(count, 1) is the
// host's bit for bit whatever the order of the additions.

__global__ void __launch_bounds__(256)
    support_moments(const std::uint8_t* support, long long n, long long inner, int len,
                    unsigned long long* sums /* [start][2]: moment, count */) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  unsigned long long moment = 0, count = 0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (support[soff + e]) {
      const int f = static_cast<int>((e / inner) % len);
      moment += static_cast<unsigned long long>((f + len / 2) % len);
      count += 1;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    moment += __shfl_down(moment, o, 64);
    count += __shfl_down(count, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && count) {
    atomicAdd(sums + 2 * blockIdx.y, moment);
    atomicAdd(sums + 2 * blockIdx.y + 1, count);
  }
}

__global__ void __launch_bounds__(256)
    support_cut(std::uint8_t* support, long long n, long long inner, int len, const unsigned long long* sums) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  const double count = static_cast<double>(sums[2 * blockIdx.y + 1]);
  const double centre = static_cast<double>(sums[2 * blockIdx.y]) / (count > 1.0 ? count : 1.0);
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int f = static_cast<int>((e / inner) % len);
    const int c = (f + len / 2) % len;
    support[soff + e] = (support[soff + e] != 0 && static_cast<double>(c) < centre) ? 1 : 0;
  }
}

// --- Shrinkwrap ------------------------------------------------------------------------
// The Gaussian transfer exp(-2 pi^2 sigma^2 |f|^2) of the reference (src/cpu/kernels.cpp:
// 52-75) factorises over the axes, so the blur is a separable circular convolution with
// the exact per-axis kernel h[d] = IDFT_N(exp(-2 pi^2 sigma^2 f^2))[d] (computed on the
// host in double).  taps[t] multiplies in[(k + t - J) mod N] (J = (T - 1) / 2), or
// in[(k + t) mod N] when the kernel spans the whole axis (full = 1, T = N).

struct BlurAxis {
  int N;        // axis length
  int inner;    // element stride of the axis
  int T;        // number of taps
  int J;        // offset of tap 0
  const float* taps;
};

template <class St, bool kFirst>
__global__ void blur_axis(const void* in_, float* out, long long n, BlurAxis ax) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>((e / ax.inner) % ax.N);
    const long long line0 = e - static_cast<long long>(k) * ax.inner;
    float acc = 0.0f;
    for (int t = 0; t < ax.T; ++t) {
      int kk = k + t - ax.J;
      kk %= ax.N;
      if (kk < 0) kk += ax.N;
      const long long at = soff + line0 + static_cast<long long>(kk) * ax.inner;
      float v;
      if constexpr (kFirst) {
        const float2 c = Storage<St>::ld(static_cast<const St*>(in_), at);
        v = sqrtf(c.x * c.x + c.y * c.y);
      } else {
        v = static_cast<const float*>(in_)[at];
      }
      acc += ax.taps[t] * v;
    }
    out[soff + e] = acc;
  }
}

// Per-block statistic of the blurred magnitude: max (method 0) or the FP64 sum of squares
// and count over the current support (method 1, the rms of src/cpu/kernels.cpp:84-98).
__global__ void __launch_bounds__(256)
    shrinkwrap_level(const float* blurred, const std::uint8_t* support, long long n, int method,
                     double* partials /* [start][blocks][2] */) {
  __shared__ double red[16];
  __shared__ float redf[16];
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  double sum = 0.0, count = 0.0;
  float peak = 0.0f;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float v = blurred[soff + e];
    if (method == 0) {
      peak = fmaxf(peak, v);
    } else if (support[soff + e]) {
      sum += static_cast<double>(v) * static_cast<double>(v);
      count += 1.0;
    }
  }
  double* out = partials + (static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x) * 2;
  if (method == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) peak = fmaxf(peak, __shfl_down(peak, o, 64));
    if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = peak;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = redf[0];
      for (int k = 1; k < static_cast<int>((blockDim.x + 63) >> 6); ++k) m = fmaxf(m, redf[k]);
      out[0] = m;
      out[1] = 0.0;
    }
  } else {
    const double s = block_sum(sum, red);
    __syncthreads();
    const double c = block_sum(count, red);
    if (threadIdx.x == 0) {
      out[0] = s;
      out[1] = c;
    }
  }
}

// support = blurred > threshold * level.  Every block reduces its start's partials to the
// level itself, in a fixed order (lane-strided sums, then a fixed tree), so all blocks of a
// start agree on it bit for bit.
__global__ void __launch_bounds__(256)
    shrinkwrap_threshold(const float* blurred, const double* partials, int blocks, long long n,
                         int method, float threshold, std::uint8_t* support) {
  __shared__ double red[16];
  __shared__ float redf[16];
  __shared__ float cut;
  const double* part = partials + static_cast<long long>(blockIdx.y) * blocks * 2;
  double s = 0.0, c = 0.0;
  float m = 0.0f;
  for (int k = threadIdx.x; k < blocks; k += blockDim.x) {
    if (method == 0) {
      m = fmaxf(m, static_cast<float>(part[2 * k]));
    } else {
      s += part[2 * k];
      c += part[2 * k + 1];
    }
  }
  if (method == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      float level = redf[0];
      for (int k = 1; k < static_cast<int>((blockDim.x + 63) >> 6); ++k) level = fmaxf(level, redf[k]);
      cut = threshold * level;
    }
  } else {
    const double ts = block_sum(s, red);
    __syncthreads();
    const double tc = block_sum(c, red);
    if (threadIdx.x == 0) cut = threshold * static_cast<float>(sqrt(ts / fmax(tc, 1.0)));
  }
  __syncthreads();
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  const float cv = cut;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    support[soff + e] = blurred[soff + e] > cv;
  }
}


// --- make_problem on the device (src/core/problem.cpp:72-99) ----------------------------
// Arrays of `count` problems of one shape, back to back (blockIdx.y = problem).  Shape
// (d0, d1, d2) with d0 = 1 in 2D.  The centre-origin index c of an axis of length N sits at
// FFT-order index (c + N - N / 2) % N (numpy's ifftshift), and FFT-order index p at
// centre-origin index (p + N / 2) % N (fftshift).

struct Grid3 {
  int d0, d1, d2;
};

__device__ __forceinline__ long long centred_to_fft(long long c, Grid3 g) {
  const int x = static_cast<int>(c % g.d2);
  const int y = static_cast<int>((c / g.d2) % g.d1);
  const int z = static_cast<int>(c / (static_cast<long long>(g.d2) * g.d1));
  const int fx = (x + g.d2 - g.d2 / 2) % g.d2;
  const int fy = (y + g.d1 - g.d1 / 2) % g.d1;
  const int fz = (z + g.d0 - g.d0 / 2) % g.d0;
  return (static_cast<long long>(fz) * g.d1 + fy) * g.d2 + fx;
}

__device__ __forceinline__ long long fft_to_centred(long long p, Grid3 g) {
  const int x = static_cast<int>(p % g.d2);
  const int y = static_cast<int>((p / g.d2) % g.d1);
  const int z = static_cast<int>(p / (static_cast<long long>(g.d2) * g.d1));
  const int cx = (x + g.d2 / 2) % g.d2;
  const int cy = (y + g.d1 / 2) % g.d1;
  const int cz = (z + g.d0 / 2) % g.d0;
  return (static_cast<long long>(cz) * g.d1 + cy) * g.d2 + cx;
}

// w[p] = in[c]: a centre-origin real (kComplex = false) or complex array moved to FFT order
// as complex values.
template <bool kComplex>
__global__ void __launch_bounds__(256)
    centred_to_wave(const float* in, float2* w, long long n, Grid3 g) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  for (long long c = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long p = centred_to_fft(c, g);
    if constexpr (kComplex) {
      const float2 v = reinterpret_cast<const float2*>(in)[soff + c];
      w[soff + p] = v;
    } else {
      w[soff + p] = {in[soff + c], 0.0f};
    }
  }
}

// |w| of the autocorrelation (std::abs, i.e. hypot: the root of the sum of squares taken
// in double and rounded once to float), stored at centre-origin positions, and the
// largest value of each block (a maximum is exact in any order).
__global__ void __launch_bounds__(256)
    autocorrelation_magnitude(const float2* w, float* mag, float* block_max, long long n, Grid3 g) {
  __shared__ float red[256];
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  float best = 0.0f;
  for (long long p = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; p < n;
       p += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float2 v = w[soff + p];
    const double re = v.x, im = v.y;
    const float m = static_cast<float>(sqrt(re * re + im * im));
    mag[soff + fft_to_centred(p, g)] = m;
    best = fmaxf(best, m);
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (static_cast<int>(threadIdx.x) < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) block_max[static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// support = mag > threshold * max (strict; the level is the float product, as the
// reference's `threshold * max`).  Every block first takes the maximum of the `blocks`
// block maxima of its problem.
__global__ void __launch_bounds__(256)
    autocorrelation_cut(const float* mag, const float* block_max, int blocks, long long n, float threshold,
                        std::uint8_t* support) {
  __shared__ float red[256];
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  float best = 0.0f;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    best = fmaxf(best, block_max[static_cast<long long>(blockIdx.y) * blocks + b]);
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (static_cast<int>(threadIdx.x) < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const float level = __fmul_rn(threshold, red[0]);
  for (long long c = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<long long>(gridDim.x) * blockDim.x) {
    support[soff + c] = mag[soff + c] > level ? 1 : 0;
  }
}

// intensity[c] = |w[p]|^2 = re * re + im * im in float, each product rounded (std::norm),
// at centre-origin positions.
__global__ void __launch_bounds__(256)
    wave_to_intensity(const float2* w, float* intensity, long long n, Grid3 g) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  for (long long p = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; p < n;
       p += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float2 v = w[soff + p];
    intensity[soff + fft_to_centred(p, g)] = __fadd_rn(__fmul_rn(v.x, v.x), __fmul_rn(v.y, v.y));
  }
}

}  // namespace qpch::dev
