// Device FFT engine for gfx950: power-of-two lines of length N, radix-R Stockham with
// the line's elements held in registers and exchanged through LDS between stages.
//
// Register layout (the contract every fused pass relies on): the thread with butterfly
// index i in [0, N/R) holds v[m] = line[i + m * N/R], m in [0, R), both before and after
// a transform.  Loads and stores are therefore strided by N/R within a line, which makes
// consecutive threads touch consecutive elements (coalesced along the line) and lets a
// forward transform, an element-wise epilogue and an inverse transform chain without any
// LDS round trip between them.
//
// Stage s of radix r with sub-transform length L: butterfly i' reads x[i' + q N/r],
// twiddles by W_{L r}^{q (i' mod L)}, and writes y[(i' - i' mod L) r + i' mod L + q L]
// (the autosort formulation).  The last stage has L r = N, so its outputs land on the
// thread's own register positions: a transform costs (stages - 1) LDS round trips.
// N = R^k * R2 with R2 < R handled as a final stage of R/R2 radix-R2 butterflies per
// thread, whose inputs and outputs are again the thread's own registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qpch::dev {

// ---------------------------------------------------------------------------------------
// complex helpers (float2 = re, im)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// a * (-i) and a * (+i)
__device__ __forceinline__ float2 mul_mi(float2 a) { return {a.y, -a.x}; }
__device__ __forceinline__ float2 mul_pi(float2 a) { return {-a.y, a.x}; }

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// cos / sin of 2 pi t / 16, t in [0, 16): exact textbook values rounded to float.
__device__ __forceinline__ constexpr float cos16(int t) {
  constexpr float c[16] = {1.0f,
                           0.92387953251128675613f,
                           0.70710678118654752440f,
                           0.38268343236508977173f,
                           0.0f,
                           -0.38268343236508977173f,
                           -0.70710678118654752440f,
                           -0.92387953251128675613f,
                           -1.0f,
                           -0.92387953251128675613f,
                           -0.70710678118654752440f,
                           -0.38268343236508977173f,
                           0.0f,
                           0.38268343236508977173f,
                           0.70710678118654752440f,
                           0.92387953251128675613f};
  return c[t & 15];
}
__device__ __forceinline__ constexpr float sin16(int t) { return cos16(t + 12); }

// v * W_R^t with W_R = exp(-+ 2 pi i / R), R a power of two <= 16.
template <int R, bool FWD>
__device__ __forceinline__ float2 twiddle_const(float2 v, int t) {
  const int u = (t * (16 / R)) & 15;  // index in sixteenths of a turn
  if (u == 0) return v;
  if (u == 4) return FWD ? mul_mi(v) : mul_pi(v);
  if (u == 8) return {-v.x, -v.y};
  if (u == 12) return FWD ? mul_pi(v) : mul_mi(v);
  const float c = cos16(u);
  const float s = FWD ? -sin16(u) : sin16(u);
  return {v.x * c - v.y * s, v.x * s + v.y * c};
}

// In-register DFT of size R (power of two <= 16): natural order in, natural order out.
// Radix-2 decimation in frequency; the bit reversal is a compile-time renaming.
template <int R, bool FWD>
__device__ __forceinline__ void dft_regs(float2* v) {
  if constexpr (R == 1) {
    return;
  } else {
#pragma unroll
    for (int span = R / 2; span >= 1; span /= 2) {
#pragma unroll
      for (int start = 0; start < R; start += 2 * span) {
#pragma unroll
        for (int j = 0; j < span; ++j) {
          const float2 a = v[start + j];
          const float2 b = v[start + j + span];
          v[start + j] = cadd(a, b);
          // twiddle W_{2 span}^j = W_R^{j R / (2 span)}
          v[start + j + span] = twiddle_const<R, FWD>(csub(a, b), j * (R / (2 * span)));
        }
      }
    }
    float2 t[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int r = 0;
#pragma unroll
      for (int b = 0; b < ilog2(R); ++b) r |= ((k >> b) & 1) << (ilog2(R) - 1 - b);
      t[k] = v[r];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = t[k];
  }
}

// LDS position of line element p: one pad slot per R elements keeps the stage writes
// (stride R in the first stage) off a single bank.
template <int R>
__device__ __forceinline__ int lds_pos(int p) {
  return p + (p >> ilog2(R));
}

template <int N, int R>
struct Shape1 {
  static constexpr int kR = R < N ? R : N;
  static constexpr int kT = N / kR;  // threads (butterflies) per line
  static constexpr int kLogR = ilog2(kR);
  static constexpr int kFull = [] {
    int s = 0, n = N;
    while (n % kR == 0 && n > 1) {
      n /= kR;
      ++s;
    }
    return s;
  }();                                         // number of radix-kR stages
  static constexpr int kR2 = N / [] {
    int p = 1;
    for (int s = 0; s < kFull; ++s) p *= kR;
    return p;
  }();                                         // last radix (1 = none)
  static constexpr int kLine = N + N / kR;     // padded LDS line length (float2)
};

// Transform of one line.  v: the thread's registers (layout above); i: its butterfly
// index; line: the line's LDS region (padded, kLine float2); tw: LDS table of W_N^j,
// j in [0, N), for the forward direction (the inverse conjugates).  Every thread of the
// work-group must call this (it contains barriers); threads without a line pass any
// valid LDS region and discard the result.
template <int N, int R, bool FWD>
__device__ __forceinline__ void fft_line(float2* v, int i, float2* line, const float2* tw) {
  using S = Shape1<N, R>;
  constexpr int kR = S::kR;
  constexpr int kT = S::kT;
  if constexpr (N == 1) return;
  int L = 1;
#pragma unroll
  for (int s = 0; s < S::kFull; ++s) {
    const int k = i & (L - 1);
    if (s > 0) {
      // twiddle x_q by W_{L R}^{q k} = W_N^{q k N / (L R)}
      const int step = k * (N / (L * kR));
#pragma unroll
      for (int q = 1; q < kR; ++q) {
        const float2 w = tw[q * step];
        v[q] = FWD ? cmul(v[q], w) : cmulc(v[q], w);
      }
    }
    dft_regs<kR, FWD>(v);
    const bool last = (s == S::kFull - 1) && S::kR2 == 1;
    if (!last) {
      const int base = (i - k) * kR + k;
#pragma unroll
      for (int q = 0; q < kR; ++q) line[lds_pos<kR>(base + q * L)] = v[q];
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kR; ++m) v[m] = line[lds_pos<kR>(i + m * kT)];
      __syncthreads();
    }
    L *= kR;
  }
  if constexpr (S::kR2 > 1) {
    // final stage: R / R2 butterflies of radix R2; butterfly j = i + j kT reads
    // registers j + q (R / R2) and is twiddled by W_N^{q (i + j kT)}.
    constexpr int kR2 = S::kR2;
    constexpr int kG = kR / kR2;
#pragma unroll
    for (int j = 0; j < kG; ++j) {
      float2 u[kR2];
      const int ip = i + j * kT;
#pragma unroll
      for (int q = 0; q < kR2; ++q) {
        u[q] = v[j + q * kG];
        if (q > 0) {
          const float2 w = tw[q * ip];
          u[q] = FWD ? cmul(u[q], w) : cmulc(u[q], w);
        }
      }
      dft_regs<kR2, FWD>(u);
#pragma unroll
      for (int q = 0; q < kR2; ++q) v[j + q * kG] = u[q];
    }
  }
}

}  // namespace qpch::dev
