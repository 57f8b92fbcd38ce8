// Fused passes along axes of smooth length N = R1 * R2 (R1 a power of two, R2 an odd factor
// of smooth_odd_factor; the pairs with a kernel are listed in smooth.hip, smooth_sq.hip
// and smooth_prime.hip),
// with the line held in registers: a radix-R1 transform (in registers up to 16, the
// radix-16 line engine from 32 on), one LDS exchange, one in-register radix-R2 stage,
// where generic_pass (generic.hpp) runs one radix <= 5 stage per LDS round trip and
// barrier (240 = 16 * 15: two stages instead of four each way).
//
// Register layouts of a line (thread i of the line's T threads):
//   A: thread i < R2 holds x[i + m R2], m < R1
//   B: thread i < R1 holds x[i + m R1], m < R2
// The forward transform takes A to B: the radix-R1 DFT of A's registers (stage 1 of the
// autosort form, L = 1) goes to LDS at i R1 + q; thread j then reads j + q R1, twiddles
// input q by W_N^(q j) and takes the radix-R2 DFT, whose outputs are the thread's own
// layout-B positions (L R2 = N).  The inverse takes B to A the same way with the radices
// in the other order.  So a modulus pass loads and stores in layout A (projection in B)
// and an update pass loads the wave in B and the iterate in A (update in A).
//
// Radix 15 is the prime-factor (Good-Thomas) form of 3 x 5: input n1, n2 at
// (5 n1 + 3 n2) mod 15, five radix-3 and three radix-5 short forms without twiddles,
// output k1, k2 at (10 k1 + 6 k2) mod 15 (10 = 5 * (5^-1 mod 3), 6 = 3 * (3^-1 mod 5)).
//
// The pass semantics are generic_tile's (lines, scales, one FP64 error partial per
// work-group, the fixed-order sum in the update pass, block numbering), so the two kernels
// are interchangeable per axis; the arithmetic (the transform's rounding) is this
// kernel's own.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "fft_engine.hpp"
#include "generic.hpp"
#include "passes.hpp"
#include "persist.hpp"
#include "storage.hpp"

namespace qpch::dev {

__host__ __device__ constexpr int smooth_pow2_ceil(int v) { return v <= 1 ? 1 : 2 * smooth_pow2_ceil((v + 1) / 2); }

// R1 <= 16, or R1 not a power of two (up to 48): the R1 stage is one in-register DFT
// (dft_any; layouts A and B above).  R1 a power of two >= 32
// (the line engine): layout A is R2 interleaved subsequences x[i + p R2], p < R1, each
// transformed by the line engine (fft_line, fft_engine.hpp) at radix RA on R1/RA threads of
// RA registers: thread (i, i') holds x[i + (i' + m R1/RA) R2], m < RA.  The exchange to
// layout B (thread j < R1 holds X[j + m R1], m < R2) goes through LDS at k R2 + i, and
// thread j twiddles S_i[j] by W_N^(i j) before the radix-R2 DFT over i.  The inverse
// mirrors it.  RA shrinks with R2 so that more of a line's R1 threads work in layout A:
// 16 for R2 = 15 (15 of 16 busy), 8 for 5 (5 of 8), 4 for 3 (3 of 4).  With radix 16 for
// every R2, 192 = 64 x 3 and 1536 = 512 x 3 were slower than the generic kernel; radix 8
// for R2 = 3 measured 5-13% slower than 4.
__host__ __device__ constexpr int smooth_engine_radix(int r2) {
  return r2 >= 9 ? 16 : r2 >= 5 ? 8 : 4;
}
// The odd factors R2: 3, 5 and 15; 9, 25, 27 and 45 (lengths of 2, 3 and 5 with a squared
// odd prime); the primes 7 to 23, and 21, 33, 35, 39, 49 and 63.  Only R2 <= 15 runs on the
// line engine (R1 >= 32), whose exchange needs RA + 1 >= R2 (7: RA 8, 7 of 8 threads busy;
// 9, 11, 13: RA 16).
__host__ __device__ constexpr bool smooth_odd_factor(int r2) {
  return r2 == 3 || r2 == 5 || r2 == 15 || r2 == 9 || r2 == 25 || r2 == 27 || r2 == 45 || r2 == 7 || r2 == 11 ||
         r2 == 13 || r2 == 21 || r2 == 33 || r2 == 35 || r2 == 39 || r2 == 17 || r2 == 19 || r2 == 23 || r2 == 49 ||
         r2 == 63;
}
__host__ __device__ constexpr bool smooth_is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
__host__ __device__ constexpr int smooth_odd_part(int v) { return v % 2 == 0 ? smooth_odd_part(v / 2) : v; }
// A radix one thread transforms in registers (dft_any): a power of two up to 16, an odd
// factor above, or their product (the prime-factor form).
__host__ __device__ constexpr bool smooth_reg_radix(int r) {
  return r >= 2 && r / smooth_odd_part(r) <= 16 && (smooth_odd_part(r) == 1 || smooth_odd_factor(smooth_odd_part(r)));
}
// Whether R1 runs on the line engine: powers of two from 32 on.
__host__ __device__ constexpr bool smooth_engine(int r1) { return r1 > 16 && smooth_is_pow2(r1); }
// Threads per line (layout A's busy threads, or R1, rounded up to a power of two), and
// lines per work-group: generic_pass's count for a power of two R1 and R2 in {3, 5, 15}
// (whole waves for every such pair); otherwise 256 threads (generic_pass's count does not
// always make whole waves, and with up to 768 threads the radix-45 and -75 stages spilled
// registers).
__host__ __device__ constexpr int smooth_threads_per_line(int r1, int r2) {
  const int ta = r2 * (smooth_engine(r1) ? r1 / smooth_engine_radix(r2) : 1);
  return smooth_pow2_ceil(ta > r1 ? ta : r1);
}
__host__ __device__ constexpr int smooth_lines_per_block(int r1, int r2) {
  const int t = smooth_threads_per_line(r1, r2);
  return smooth_is_pow2(r1) && (r2 == 3 || r2 == 5 || r2 == 15) ? generic_lines_per_block(r1 * r2)
                                                                 : (t >= 256 ? 1 : 256 / t);
}

template <int R1, int R2, bool COL>
struct SmoothGeom {
  static_assert((R1 >= 2 && R1 <= 512 && smooth_is_pow2(R1)) || (smooth_reg_radix(R1) && R1 <= 48),
                "R1: a power of two from 2 to 512, or a register radix up to 48");
  static_assert(smooth_reg_radix(R2) && R2 <= 63, "R2: a register radix up to 63");
  static constexpr int kN = R1 * R2;
  static_assert(kN <= kGenericMaxN, "the generic kernels' length limit");
  static constexpr bool kEngine = smooth_engine(R1);           // the line engine for R1
  static_assert(!kEngine || (R2 <= 15 && R2 % 2 == 1), "the line engine's exchange holds R2 <= RA + 1 subsequences");
  static constexpr int kRA = smooth_engine_radix(R2);          // its radix
  static constexpr int kSub = kEngine ? R1 / kRA : 1;          // threads per subsequence (layout A)
  static constexpr int kA = kEngine ? kRA : R1;                // registers in layout A
  static constexpr int kTA = R2 * kSub;                        // threads holding layout A
  static constexpr int kV = kA > R2 ? kA : R2;                 // registers per thread
  static constexpr int kT = smooth_threads_per_line(R1, R2);  // threads per line
  static constexpr int kB = smooth_lines_per_block(R1, R2);    // lines per work-group
  static constexpr int kThreads = kB * kT;
  static constexpr int kWaves = (kThreads + 63) / 64;
  // line scratch: R1 <= 16, the line with one pad slot per R1 elements; the engine, one
  // padded fft_line region per subsequence slot i = t / kSub (all kT / kSub = RA of them, so
  // that threads without a subsequence transform in a region of their own; RA + 1 >= R2, so
  // they also hold the N-element exchange)
  static constexpr int kRegion = R1 + R1 / kRA;
  static constexpr int kLine = kEngine ? (kT / kSub) * kRegion : kN + kN / R1;
  static constexpr int kStride = COL ? (kLine | 1) : kLine;    // odd for column tiles (banks)
  static constexpr int kTwP = kEngine ? R1 : 0;                // W_R1 table of the engine
  static_assert(kThreads % 64 == 0, "whole waves (the error partial's shuffles)");
  static_assert(!kEngine || kT == R1, "one thread per layout-B column");
  static_assert(!kEngine || kLine >= kN, "the exchange fits the engine regions");
  // line position held in layout-A register m by thread t (t < kTA), and in layout B (t < R1)
  __device__ static int posA(int t, int m) {
    if constexpr (kEngine) {
      return t / kSub + (t % kSub + m * kSub) * R2;
    } else {
      return t + m * R2;
    }
  }
  __device__ static int posB(int t, int m) { return t + m * R1; }
};

// W_N^j = exp(-2 pi i j / N), j < N, into LDS: evaluated in double, rounded once (the
// values of the generic kernels' tables).
template <int N, int kThreads>
__device__ __forceinline__ void smooth_twiddles(float2* tw) {
  for (int j = threadIdx.x; j < N; j += kThreads) {
    double sn, cs;
    sincospi(2.0 * static_cast<double>(j) / static_cast<double>(N), &sn, &cs);
    tw[j] = {static_cast<float>(cs), static_cast<float>(-sn)};
  }
}

// LDS position of line element p: stage-1 writes of the forward transform go to i R1 + q,
// a stride of R1 across the lanes; the pad makes it R1 + 1.
template <int R1>
__device__ __forceinline__ int smooth_pos(int p) {
  return p + p / R1;
}

template <bool INV>
__device__ __forceinline__ void dft15(float2* v) {
  float2 u[3][5];
#pragma unroll
  for (int n1 = 0; n1 < 3; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) u[n1][n2] = v[(5 * n1 + 3 * n2) % 15];
  }
#pragma unroll
  for (int n2 = 0; n2 < 5; ++n2) {
    float2 t[3] = {u[0][n2], u[1][n2], u[2][n2]};
    dft_small<3, INV>(t);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) u[k1][n2] = t[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) dft_small<5, INV>(u[k1]);
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
#pragma unroll
    for (int k2 = 0; k2 < 5; ++k2) v[(10 * k1 + 6 * k2) % 15] = u[k1][k2];
  }
}

// sin and cos of x, |x| <= pi, in double at compile time: Taylor series to 30 terms (the
// last one below 1e-18 of the sum), for the twiddle constants of the stages below.
constexpr double ce_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 30; ++k) {
    term *= -x * x / static_cast<double>((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 30; ++k) {
    term *= -x * x / static_cast<double>((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
// The angle 2 pi j / R reduced to (-pi, pi].
constexpr double ce_angle(int j, int R) {
  const int r = j % R;
  return 2.0 * 3.14159265358979323846 * static_cast<double>(2 * r > R ? r - R : r) / static_cast<double>(R);
}
constexpr int ce_inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x) {
    if ((a * x) % m == 1) return x;
  }
  return 0;
}

template <int... I, class F>
__device__ __forceinline__ void unroll_seq(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, i>) for i < N: indices usable in constant expressions
template <int N, class F>
__device__ __forceinline__ void unroll(F&& f) {
  unroll_seq(std::make_integer_sequence<int, N>{}, f);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_odd(float2* v);
template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v);

// Cooley-Tukey R = P Q in registers: with n = n1 + P n2 and k = Q k1 + k2,
//   Y[n1][k2] = DFT_Q over n2 of x[n1 + P n2],  Y[n1][k2] *= W_R^(n1 k2),
//   X[Q k1 + k2] = DFT_P over n1 of Y[n1][k2];
// the twiddles are compile-time constants (W_R^j rounded once from double).
template <int P, int Q, bool INV>
__device__ __forceinline__ void dft_ct(float2* v) {
  constexpr int R = P * Q;
  float2 y[P][Q];
#pragma unroll
  for (int n1 = 0; n1 < P; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) y[n1][n2] = v[n1 + P * n2];
    dft_odd<Q, INV>(y[n1]);
  }
  unroll<P>([&](auto n1) {
    unroll<Q>([&](auto k2) {
      constexpr int j = decltype(n1)::value * decltype(k2)::value;
      if constexpr (j % R != 0) {
        constexpr float c = static_cast<float>(ce_cos(ce_angle(j, R)));
        constexpr float sn = static_cast<float>(ce_sin(ce_angle(j, R)));
        y[n1][k2] = cmul(y[n1][k2], float2{c, INV ? sn : -sn});
      }
    });
  });
#pragma unroll
  for (int k2 = 0; k2 < Q; ++k2) {
    float2 t[P];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) t[n1] = y[n1][k2];
    dft_odd<P, INV>(t);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) v[Q * k1 + k2] = t[k1];
  }
}

// Prime-factor (Good-Thomas) form of N = N1 N2, gcd 1: input n1, n2 at (N2 n1 + N1 n2) mod N,
// DFT_N1 over n1, DFT_N2 over n2, no twiddles, output k1, k2 at (a k1 + b k2) mod N with
// a = N2 (N2^-1 mod N1), b = N1 (N1^-1 mod N2) (dft15 is this form for 3 x 5).
template <int N1, int N2, bool INV>
__device__ __forceinline__ void dft_pfa(float2* v) {
  constexpr int N = N1 * N2;
  constexpr int kA = N2 * ce_inv_mod(N2 % N1, N1), kB = N1 * ce_inv_mod(N1 % N2, N2);
  static_assert((kA % N1) == 1 && (kA % N2) == 0 && (kB % N2) == 1 && (kB % N1) == 0, "coprime factors");
  float2 u[N1][N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) u[n1][n2] = v[(N2 * n1 + N1 * n2) % N];
  }
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    float2 t[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) t[n1] = u[n1][n2];
    dft_any<N1, INV>(t);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) u[k1][n2] = t[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) dft_any<N2, INV>(u[k1]);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) v[(kA * k1 + kB * k2) % N] = u[k1][k2];
  }
}

// DFT of an odd prime P in registers by the symmetric short form (dft_small's for 3 and 5,
// in general): with s_j = x_j + x_(P-j), d_j = x_j - x_(P-j), j = 1 .. (P-1)/2,
//   X_0 = x_0 + sum s_j,
//   X_k, X_(P-k) = x_0 + sum_j cos(2 pi j k / P) s_j  -+ i sum_j sin(2 pi j k / P) d_j
// (upper signs forward), the constants evaluated at compile time.
template <int P, bool INV>
__device__ __forceinline__ void dft_prime(float2* v) {
  constexpr int H = (P - 1) / 2;
  const float2 a = v[0];
  float2 sm[H], df[H];
  float2 x0 = a;
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    sm[j - 1] = {v[j].x + v[P - j].x, v[j].y + v[P - j].y};
    df[j - 1] = {v[j].x - v[P - j].x, v[j].y - v[P - j].y};
    x0 = {x0.x + sm[j - 1].x, x0.y + sm[j - 1].y};
  }
  unroll<H>([&](auto kk) {
    constexpr int k = decltype(kk)::value + 1;
    float2 re = a, im = {0.0f, 0.0f};
    unroll<H>([&](auto jj) {
      constexpr int j = decltype(jj)::value + 1;
      constexpr float c = static_cast<float>(ce_cos(ce_angle(j * k, P)));
      constexpr float sn = static_cast<float>(ce_sin(ce_angle(j * k, P)));
      re = {re.x + c * sm[j - 1].x, re.y + c * sm[j - 1].y};
      im = {im.x + sn * df[j - 1].x, im.y + sn * df[j - 1].y};
    });
    const float2 u = rot<INV>(im);
    v[k] = {re.x + u.x, re.y + u.y};
    v[P - k] = {re.x - u.x, re.y - u.y};
  });
  v[0] = x0;
}

// In-register DFT of the odd radices of the smooth passes: 3 and 5 (short forms), 7 to 23
// (dft_prime), 15 (prime factor 3 x 5), 9 = 3 x 3, 25 = 5 x 5, 27 = 3 x 9 and 49 = 7 x 7
// (Cooley-Tukey), 21, 33, 35, 39, 45 = 9 x 5 and 63 = 9 x 7 (prime factor).
template <int R, bool INV>
__device__ __forceinline__ void dft_odd(float2* v) {
  if constexpr (R == 15) {
    dft15<INV>(v);
  } else if constexpr (R == 45) {
    dft_pfa<9, 5, INV>(v);
  } else if constexpr (R == 7 || R == 11 || R == 13 || R == 17 || R == 19 || R == 23) {
    dft_prime<R, INV>(v);
  } else if constexpr (R == 49) {
    dft_ct<7, 7, INV>(v);
  } else if constexpr (R == 63) {
    dft_pfa<9, 7, INV>(v);
  } else if constexpr (R == 21 || R == 33 || R == 39) {
    dft_pfa<3, R / 3, INV>(v);
  } else if constexpr (R == 35) {
    dft_pfa<5, 7, INV>(v);
  } else if constexpr (R == 9 || R == 27) {
    dft_ct<3, R / 3, INV>(v);
  } else if constexpr (R == 25) {
    dft_ct<5, 5, INV>(v);
  } else {
    dft_small<R, INV>(v);
  }
}

// In-register DFT of any register radix R = 2^a o (smooth_reg_radix): dft_regs for a power
// of two, dft_odd for an odd R, and the prime-factor form of 2^a x o otherwise.
template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v) {
  constexpr int O = smooth_odd_part(R);
  if constexpr (O == 1) {
    dft_regs<R, !INV>(v);
  } else if constexpr (O == R) {
    dft_odd<R, INV>(v);
  } else {
    dft_pfa<R / O, O, INV>(v);
  }
}

// Layout A -> layout B (forward).  t: the thread's index within its line; line: the line's
// scratch; tw: W_N^j, j < N (forward table); twP: W_R1^j, j < R1 (the engine's table, R1 >=
// 32).  Every thread of the work-group calls it.
template <int R1, int R2>
__device__ __forceinline__ void smooth_fwd(float2* v, int t, float2* line, const float2* tw, const float2* twP) {
  using G = SmoothGeom<R1, R2, false>;
  if constexpr (!G::kEngine) {
    const int i = t;
    if (i < R2) {
      dft_any<R1, false>(v);
#pragma unroll
      for (int q = 0; q < R1; ++q) line[smooth_pos<R1>(i * R1 + q)] = v[q];
    }
    __syncthreads();
    if (i < R1) {
      float2 w[R2];
#pragma unroll
      for (int q = 0; q < R2; ++q) {
        v[q] = line[smooth_pos<R1>(i + q * R1)];
        w[q] = tw[q * i];
      }
#pragma unroll
      for (int q = 1; q < R2; ++q) {
        const float2 u = cmul(v[q], w[q]);
        if (i != 0) v[q] = u;  // W^0 = 1: the input stays as it is
      }
      dft_any<R2, false>(v);
    }
    __syncthreads();
  } else {
    constexpr int kSub = G::kSub;
    const int i = t / kSub, ip = t % kSub;
    fft_line<R1, G::kRA, true>(v, ip, line + i * G::kRegion, twP);  // S_i, natural order (ends past its last LDS read)
    if (t < G::kTA) {
#pragma unroll
      for (int m = 0; m < G::kRA; ++m) line[(ip + m * kSub) * R2 + i] = v[m];
    }
    __syncthreads();
    if (t < R1) {
      float2 w[R2];
#pragma unroll
      for (int q = 0; q < R2; ++q) {
        v[q] = line[t * R2 + q];
        w[q] = tw[q * t];
      }
#pragma unroll
      for (int q = 1; q < R2; ++q) {
        const float2 u = cmul(v[q], w[q]);
        if (t != 0) v[q] = u;
      }
      dft_odd<R2, false>(v);
    }
    __syncthreads();
  }
}

// Layout B -> layout A (inverse, unnormalised).
template <int R1, int R2>
__device__ __forceinline__ void smooth_inv(float2* v, int t, float2* line, const float2* tw, const float2* twP) {
  using G = SmoothGeom<R1, R2, false>;
  if constexpr (!G::kEngine) {
    const int i = t;
    if (i < R1) {
      dft_any<R2, true>(v);
#pragma unroll
      for (int q = 0; q < R2; ++q) line[smooth_pos<R1>(i * R2 + q)] = v[q];
    }
    __syncthreads();
    if (i < R2) {
      float2 w[R1];
#pragma unroll
      for (int q = 0; q < R1; ++q) {
        v[q] = line[smooth_pos<R1>(i + q * R2)];
        w[q] = tw[q * i];
      }
#pragma unroll
      for (int q = 1; q < R1; ++q) {
        const float2 u = cmulc(v[q], w[q]);
        if (i != 0) v[q] = u;
      }
      dft_any<R1, true>(v);
    }
    __syncthreads();
  } else {
    constexpr int kSub = G::kSub;
    if (t < R1) {
      dft_odd<R2, true>(v);  // output i of column t: T_i[t] before its twiddle
#pragma unroll
      for (int i = 0; i < R2; ++i) {
        const float2 u = cmulc(v[i], tw[i * t]);
        line[t * R2 + i] = (t != 0 && i != 0) ? u : v[i];
      }
    }
    __syncthreads();
    const int i = t / kSub, ip = t % kSub;
    if (t < G::kTA) {
#pragma unroll
      for (int m = 0; m < G::kRA; ++m) v[m] = line[(ip + m * kSub) * R2 + i];
    }
    __syncthreads();
    fft_line<R1, G::kRA, false>(v, ip, line + i * G::kRegion, twP);
  }
}

// The modulus or update pass over one tile (kB lines) of start `start`: generic_tile's
// semantics for these two modes (fp32).  lds: kB line slots of kStride float2; tw: the
// twiddle table W_N^j (kN float2), filled here with OWN_TW (while the loads are in
// flight), else already filled and visible; red: one double per wave.  `slot`: the error
// history slot of the update pass (kSlotFromCounter: *p.counter - 1).  COH: the wave, the
// iterate and the partials may have been written by other CUs in the same launch (the
// persistent launches): device-scope loads.
template <int R1, int R2, bool COL, Mode M, bool OWN_TW = true, bool COH = false>
__device__ __forceinline__ void smooth_tile(const PassArgs& p, long long lines, int start, int tile, float2* lds,
                                            float2* tw, float2* twP, double* red, int slot) {
  static_assert(M != Mode::kGrad, "no gradient mode");
  using G = SmoothGeom<R1, R2, COL>;
  constexpr int N = G::kN;
  using S = Storage<float2>;
  const int tid = threadIdx.x;
  int b, t;  // the thread's line in the tile, and its index within the line
  if constexpr (COL) {
    b = tid % G::kB;
    t = tid / G::kB;
  } else {
    t = tid % G::kT;
    b = tid / G::kT;
  }
  const long long L = static_cast<long long>(tile) * G::kB + b;
  const bool valid = L < lines;
  const bool inA = t < G::kTA, inB = t < R1;
  // Loads are unconditional (all in flight at once): lanes past the last line read line 0
  // of the start, idle lanes of a layout read its thread 0's positions; neither stores.
  const long long Lc = valid ? L : 0;
  const int inner = p.inner;
  const long long base = (Lc / inner) * N * inner + Lc % inner;  // element 0 of the line within the start
  const int ta = inA ? t : 0, tb = inB ? t : 0;
  const long long soff = static_cast<long long>(start) * p.n;
  float2* w = static_cast<float2*>(p.w) + soff;
  float2* x = static_cast<float2*>(p.x) + soff;
  float2* line = lds + b * G::kStride;
  float2 v[G::kV];
  if constexpr (M == Mode::kFwd || M == Mode::kInv) {
    // the transforms of the prologue and of the 3D line passes: read src (or the wave),
    // transform, write the wave scaled; forward A -> B, inverse B -> A
    const float2* in = p.src ? static_cast<const float2*>(p.src) + soff : w;
    constexpr bool kF = M == Mode::kFwd;
#pragma unroll
    for (int m = 0; m < (kF ? G::kA : R2); ++m) {
      const int q = kF ? G::posA(ta, m) : G::posB(tb, m);
      v[m] = in[base + static_cast<long long>(q) * inner];
    }
    if constexpr (OWN_TW) {
      smooth_twiddles<N, G::kThreads>(tw);
      if constexpr (G::kEngine) smooth_twiddles<R1, G::kThreads>(twP);
    }
    __syncthreads();
    if constexpr (kF) {
      smooth_fwd<R1, R2>(v, t, line, tw, twP);
    } else {
      smooth_inv<R1, R2>(v, t, line, tw, twP);
    }
    if (valid && (kF ? inB : inA)) {
#pragma unroll
      for (int m = 0; m < (kF ? R2 : G::kA); ++m) {
        const int q = kF ? G::posB(t, m) : G::posA(t, m);
        w[base + static_cast<long long>(q) * inner] = {v[m].x * p.out_scale, v[m].y * p.out_scale};
      }
    }
  } else if constexpr (M == Mode::kModulus) {
    const long long arow = p.amp_stride * ((p.err_offset + start) / (p.amp_group > 1 ? p.amp_group : 1));
    float amp[R2];
#pragma unroll
    for (int m = 0; m < G::kA; ++m) {
      v[m] = S::ld<COH>(w, static_cast<size_t>(base + static_cast<long long>(G::posA(ta, m)) * inner));
    }
#pragma unroll
    for (int m = 0; m < R2; ++m) amp[m] = p.amp[arow + base + static_cast<long long>(G::posB(tb, m)) * inner];
    if constexpr (OWN_TW) {  // while the loads are in flight
      smooth_twiddles<N, G::kThreads>(tw);
      if constexpr (G::kEngine) smooth_twiddles<R1, G::kThreads>(twP);
    }
    __syncthreads();
    smooth_fwd<R1, R2>(v, t, line, tw, twP);
    double acc = 0.0;
    if (valid && inB) {
      const float io = p.in_scale * p.out_scale;
#pragma unroll
      for (int m = 0; m < R2; ++m) v[m] = project_scaled(v[m], amp[m], p.in_scale, io, acc);
    }
    const double ws = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = ws;
    if (tid == 0 && tile == 0 && start == 0 && p.counter) atomicAdd(p.counter, 1);
    smooth_inv<R1, R2>(v, t, line, tw, twP);  // its barriers also publish red
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < G::kWaves; ++k) t += red[k];
      p.partials[static_cast<long long>(start) * p.partial_stride + tile] = t;
    }
    if (valid && inA) {
#pragma unroll
      for (int m = 0; m < G::kA; ++m) w[base + static_cast<long long>(G::posA(t, m)) * inner] = v[m];  // out_scale in io
    }
  } else {
    const std::uint8_t* sup = p.support + soff;
    float2 xo[G::kA];
    bool inside[G::kA];
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      v[m] = S::ld<COH>(w, static_cast<size_t>(base + static_cast<long long>(G::posB(tb, m)) * inner));
    }
#pragma unroll
    for (int m = 0; m < G::kA; ++m) {
      const long long e = base + static_cast<long long>(G::posA(ta, m)) * inner;
      xo[m] = S::ld<COH>(x, static_cast<size_t>(e));
      inside[m] = sup[e] != 0;
    }
    if constexpr (OWN_TW) {
      smooth_twiddles<N, G::kThreads>(tw);
      if constexpr (G::kEngine) smooth_twiddles<R1, G::kThreads>(twP);
    }
    __syncthreads();
    smooth_inv<R1, R2>(v, t, line, tw, twP);
    if (valid && inA) {
#pragma unroll
      for (int m = 0; m < G::kA; ++m) {
        const float2 pm = {v[m].x * p.in_scale, v[m].y * p.in_scale};
        float2 nx = pm;
        if (!inside[m]) nx = {p.a * xo[m].x + p.b * pm.x, p.a * xo[m].y + p.b * pm.y};
        x[base + static_cast<long long>(G::posA(t, m)) * inner] = nx;
        v[m] = nx;
      }
    }
    smooth_fwd<R1, R2>(v, t, line, tw, twP);
    if (valid && inB) {
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        w[base + static_cast<long long>(G::posB(t, m)) * inner] = {v[m].x * p.out_scale, v[m].y * p.out_scale};
      }
    }
    if (tile == 0 && tid < 64 && p.errs) {
      // fixed-order sum of the previous modulus pass's partials of this start
      const double t = partial_sum<COH>(p.partials + static_cast<long long>(start) * p.partial_stride,
                                        p.partial_stride, tid);
      if (tid == 0) {
        const int e = slot == kSlotFromCounter ? *p.counter - 1 : slot;
        if (e >= 0 && e < p.err_cap) p.errs[static_cast<long long>(e) * p.err_stride + p.err_offset + start] = t;
      }
    }
  }
}

// One smooth_tile per work-group; block numbering as generic_pass's.
template <int R1, int R2, bool COL, Mode M>
__global__ void __launch_bounds__((SmoothGeom<R1, R2, COL>::kThreads))
    smooth_pass(PassArgs p, long long lines) {
  using G = SmoothGeom<R1, R2, COL>;
  __shared__ __align__(16) float2 lds[G::kB * G::kStride + G::kN + G::kTwP + G::kWaves];
  int tile, start;
  if (p.slots > 0) {
    start = blockIdx.x % p.slots;
    tile = blockIdx.x / p.slots;
    if (start >= p.starts) return;  // padding slot: the whole block leaves together
  } else {
    start = blockIdx.y;
    tile = blockIdx.x;
  }
  float2* tw = lds + G::kB * G::kStride;
  float2* twP = tw + G::kN;
  smooth_tile<R1, R2, COL, M>(p, lines, start, tile, lds, tw, twP, reinterpret_cast<double*>(twP + G::kTwP),
                              kSlotFromCounter);
}

// `len` iterations of every start of a launch group on an N x N image whose axis length is
// smooth, in one persistent launch: persistent_loop (persist.hpp) over the tiles of the
// column modulus pass and the row update pass, each a smooth_tile (the per-pass kernels'
// body; the same work-group size for both, as the axes are equal), with device-scope
// loads of what the launch itself writes.  Both tables are filled once.
template <int R1, int R2>
__global__ void __launch_bounds__((SmoothGeom<R1, R2, true>::kThreads))
    twopass_smooth(PassArgs p0, PassArgs p1, long long lines0, long long lines1, SyncArgs q, int len) {
  using GC = SmoothGeom<R1, R2, true>;
  using GR = SmoothGeom<R1, R2, false>;
  static_assert(GC::kThreads == GR::kThreads, "one work-group size for both passes");
  constexpr int kSlots = GC::kB * GC::kStride > GR::kB * GR::kStride ? GC::kB * GC::kStride : GR::kB * GR::kStride;
  __shared__ __align__(16) float2 lds[kSlots + GC::kN + GC::kTwP + GC::kWaves];
  __shared__ int claim;
  float2* tw = lds + kSlots;
  float2* twP = tw + GC::kN;
  double* red = reinterpret_cast<double*>(twP + GC::kTwP);
  smooth_twiddles<GC::kN, GC::kThreads>(tw);  // both axes: one table (and one engine table)
  if constexpr (GC::kEngine) smooth_twiddles<R1, GC::kThreads>(twP);
  persistent_loop(
      p1, q, len, claim,
      [&](int start, int tile) {
        smooth_tile<R1, R2, true, Mode::kModulus, false, true>(p0, lines0, start, tile, lds, tw, twP, red, 0);
      },
      [&](int start, int tile, int slot) {
        smooth_tile<R1, R2, false, Mode::kUpdate, false, true>(p1, lines1, start, tile, lds, tw, twP, red, slot);
      });
}

}  // namespace qpch::dev
