// Shrinkwrap's Gaussian blur through the FFT passes (power-of-two shapes): the same
// FFT -> multiply by exp(-2 pi^2 sigma^2 |f|^2) -> inverse FFT as the reference
// (src/cpu/kernels.cpp:52-75), fused into three kinds of line pass over an fp32 scratch:
//   kMag      first transformed axis: load x (any storage), |x| as a real line, forward FFT
//   kGauss    axis 0: forward FFT, multiply by the separable transfer factors of every
//             axis (one table per axis), inverse FFT
//   kReal     last axis: inverse FFT, real part times 1/n, store the blurred magnitude and
//             the block's Shrinkwrap statistic (max, or sum of squares and count over the
//             current support)
// 3D inserts plain y forward / inverse passes (line_pass kFwd/kInv on the scratch).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_engine.hpp"
#include "passes.hpp"
#include "storage.hpp"

namespace qpch::dev {

enum class Blur : int { kMag = 0, kGauss = 1, kReal = 2 };

struct BlurArgs {
  int outer, inner;
  long long n;
  const void* x;          // kMag: the iterate (St)
  float2* w;              // scratch spectrum [start][n]
  float* out;             // kReal: blurred magnitude [start][n]
  const std::uint8_t* support;  // kReal, rms: current support
  const float* g0;        // per-axis transfer factors (axis 0, 1, 2; FFT-order index)
  const float* g1;
  const float* g2;
  int d1, d2;             // extents of axes 1 and 2 (1 when absent)
  int method;             // 0 max, 1 rms
  float scale;            // kReal: 1/n
  double* partials;       // kReal: [start][tiles][2]
  int starts;
  int slots;
};

template <int N, bool COL, Blur M, class St>
__global__ void __launch_bounds__((PassGeom<N, COL, 8>::kThreads))
    blur_pass(BlurArgs p) {
  using G = PassGeom<N, COL, 8>;
  constexpr int kT = G::kT;
  constexpr int kR = G::kR;
  constexpr int kB = G::kB;
  __shared__ float2 lds[G::kTile + N];
  __shared__ double red[16];
  __shared__ float redf[16];
  float2* tw = lds + G::kTile;

  int tile, start;
  if (p.slots > 0) {
    start = blockIdx.x % p.slots;
    tile = blockIdx.x / p.slots;
    if (start >= p.starts) return;
  } else {
    const int tiles = gridDim.x / p.starts;
    start = blockIdx.x / tiles;
    tile = blockIdx.x % tiles;
  }
  const int tid = threadIdx.x;
  int b, i;
  if constexpr (COL) {
    b = tid % kB;
    i = tid / kB;
  } else {
    i = tid % kT;
    b = tid / kT;
  }
  long long base;
  bool valid;
  int c = 0;
  if constexpr (COL) {
    const int groups = (p.inner + kB - 1) / kB;
    const int o = tile / groups;
    c = (tile % groups) * kB + b;
    valid = c < p.inner;
    base = static_cast<long long>(o) * N * p.inner + c;
  } else {
    const int o = tile * kB + b;
    valid = o < p.outer;
    base = static_cast<long long>(o) * N;
  }
  const long long es = COL ? p.inner : 1;
  const long long soff = static_cast<long long>(start) * p.n;
  // unconditional loads (all in flight at once); lanes past a ragged tile read line 0
  const long long lb = soff + (valid ? base : 0);
  float2* line = lds + b * G::kStride;
  float2 v[kR];

  if constexpr (M == Blur::kMag) {
    const St* x = static_cast<const St*>(p.x);
#pragma unroll
    for (int m = 0; m < kR; ++m) {
      float2 e = Storage<St>::ld(x, lb + (i + m * kT) * es);
      v[m] = {sqrtf(e.x * e.x + e.y * e.y), 0.0f};
    }
    load_twiddles<N>(tw);
    __syncthreads();
    fft_line<N, 16, true>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) p.w[soff + base + (i + m * kT) * es] = v[m];
    }
  } else if constexpr (M == Blur::kGauss) {
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = p.w[lb + (i + m * kT) * es];
    load_twiddles<N>(tw);
    __syncthreads();
    fft_line<N, 16, true>(v, i, line, tw);
    // column c of axis 0 is (axis-1 index, axis-2 index) = (c / d2, c % d2)
    const int cc = valid ? c : 0;
    const float gc = p.g1[cc / p.d2] * (p.g2 ? p.g2[cc % p.d2] : 1.0f);
#pragma unroll
    for (int m = 0; m < kR; ++m) {
      const float g = gc * p.g0[i + m * kT];
      v[m] = {v[m].x * g, v[m].y * g};
    }
    fft_line<N, 16, false>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) p.w[soff + base + (i + m * kT) * es] = v[m];
    }
  } else {  // kReal
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = p.w[lb + (i + m * kT) * es];
    bool inside[kR];
    if (p.method != 0) {
#pragma unroll
      for (int m = 0; m < kR; ++m) inside[m] = p.support[lb + (i + m * kT) * es] != 0;
    }
    load_twiddles<N>(tw);
    __syncthreads();
    fft_line<N, 16, false>(v, i, line, tw);
    double sum = 0.0, count = 0.0;
    float peak = 0.0f;
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) {
        const long long at = lb + (i + m * kT) * es;
        const float r = v[m].x * p.scale;
        p.out[at] = r;
        if (p.method == 0) {
          peak = fmaxf(peak, r);
        } else if (inside[m]) {
          sum += static_cast<double>(r) * static_cast<double>(r);
          count += 1.0;
        }
      }
    }
    double* part = p.partials + (static_cast<long long>(start) * (gridDim.x / (p.slots > 0 ? p.slots : p.starts)) +
                                 tile) * 2;
    if (p.method == 0) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) peak = fmaxf(peak, __shfl_down(peak, o, 64));
      if ((tid & 63) == 0) redf[tid >> 6] = peak;
      __syncthreads();
      if (tid == 0) {
        float mx = redf[0];
        for (int k = 1; k < static_cast<int>((blockDim.x + 63) >> 6); ++k) mx = fmaxf(mx, redf[k]);
        part[0] = mx;
        part[1] = 0.0;
      }
    } else {
      const double s = block_sum(sum, red);
      __syncthreads();
      const double cn = block_sum(count, red);
      if (tid == 0) {
        part[0] = s;
        part[1] = cn;
      }
    }
  }
}

}  // namespace qpch::dev
