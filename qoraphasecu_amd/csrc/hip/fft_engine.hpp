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

template <int R>
__host__ __device__ constexpr int bitrev(int k) {
  int r = 0;
  for (int b = 0; b < ilog2(R); ++b) r |= ((k >> b) & 1) << (ilog2(R) - 1 - b);
  return r;
}

// The butterfly (a, b) -> (a + w b, a - w b) with w = (c, s).  a + w b takes four FMAs
// (each product lands in one), and a - w b = 2a - (a + w b) one more per component: six
// instructions where a separate product and sum take eight.
__device__ __forceinline__ void bfly_w(float2& a, float2& b, float c, float s) {
  const float2 y0 = {fmaf(b.x, c, fmaf(-b.y, s, a.x)), fmaf(b.x, s, fmaf(b.y, c, a.y))};
  b = {fmaf(2.0f, a.x, -y0.x), fmaf(2.0f, a.y, -y0.y)};
  a = y0;
}

// In-register DFT of size R (power of two <= 16) with W_R = exp(-+ 2 pi i / R): natural
// order in and out.  Radix-2 decimation in time over the bit-reversed input (the bit
// reversal is a compile-time renaming), so every twiddle multiplies a butterfly's second
// input and fuses into it (bfly_w).  With TW the inputs are first multiplied by w[q]
// (conjugated for the inverse), q >= 1, input 0 untwiddled: the products of the first
// stage's second inputs fuse the same way.
template <int R, bool FWD, bool TW = false>
__device__ __forceinline__ void dft_regs(float2* v, const float2* w = nullptr) {
  if constexpr (R == 1) {
    return;
  } else {
    float2 u[R];
#pragma unroll
    for (int k = 0; k < R; ++k) u[k] = v[bitrev<R>(k)];
    // span 1: the pairs (v[q], v[q + R/2]), q = bitrev(k)
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      float2 a = u[k], b = u[k + 1];
      if constexpr (TW) {
        const int q = bitrev<R>(k);
        if (q != 0) a = FWD ? cmul(a, w[q]) : cmulc(a, w[q]);
        const float2 t = w[q + R / 2];
        bfly_w(a, b, t.x, FWD ? t.y : -t.y);
      } else {
        const float2 y0 = cadd(a, b);
        b = csub(a, b);
        a = y0;
      }
      u[k] = a;
      u[k + 1] = b;
    }
#pragma unroll
    for (int span = 2; span < R; span *= 2) {
#pragma unroll
      for (int start = 0; start < R; start += 2 * span) {
#pragma unroll
        for (int j = 0; j < span; ++j) {
          // twiddle W_{2 span}^j, in sixteenths of a turn
          const int t = j * (16 / (2 * span));
          float2& a = u[start + j];
          float2& b = u[start + j + span];
          if (t == 0) {
            const float2 y0 = cadd(a, b);
            b = csub(a, b);
            a = y0;
          } else if (t == 4) {
            const float2 wb = FWD ? mul_mi(b) : mul_pi(b);
            b = csub(a, wb);
            a = cadd(a, wb);
          } else {
            bfly_w(a, b, cos16(t), FWD ? -sin16(t) : sin16(t));
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = u[k];
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
      // x_q twiddled by W_{L R}^{q k} = W_N^{q k N / (L R)}, inside the DFT's first stage
      const int step = k * (N / (L * kR));
      float2 w[kR];
#pragma unroll
      for (int q = 1; q < kR; ++q) w[q] = tw[q * step];
      dft_regs<kR, FWD, true>(v, w);
    } else {
      dft_regs<kR, FWD>(v);
    }
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
      float2 w[kR2];
      const int ip = i + j * kT;
#pragma unroll
      for (int q = 0; q < kR2; ++q) {
        u[q] = v[j + q * kG];
        if (q > 0) w[q] = tw[q * ip];
      }
      dft_regs<kR2, FWD, true>(u, w);
#pragma unroll
      for (int q = 0; q < kR2; ++q) v[j + q * kG] = u[q];
    }
  }
}

}  // namespace qpch::dev
