// Fused passes along axes whose length N (32 to 2048) has a prime factor of 7 or more and is
// not a smooth length (smooth.hpp takes odd parts with primes up to 23 where it measured
// faster), by
// Bluestein's chirp-z transform with the power-of-two transforms of length M = 2^k >=
// max(256, 2N - 1) held in registers by the radix-16 line engine (fft_line,
// fft_engine.hpp).  generic_pass runs such factors as O(r^2) stages up to 64, and above as
// the same algorithm with one radix-4/2 stage per LDS round trip (M = 2048: six each way)
// or, below 256, a plain O(N^2) DFT per line.
//
// A line lives in the engine's layout throughout: thread t < M/16 holds positions
// p = t + m M/16, m < 16; positions p >= N hold zero.  The length-N DFT (unnormalised,
// direction by INV) of the line is
//   a_p = x_p c_p (p < N),  A = FFT_M(a),  A_k *= bhat_k,  conv = IFFT_M(A) (unnormalised),
//   X_p = c_p conv_p / M (p < N),
// with c_p = exp(-i pi p^2 / N) and bhat = FFT_M of the chirp filter (the tables of
// generic_plan, made on the host in double); the inverse takes both tables conjugated.
// The modulus pass projects X in registers and transforms back; the update pass transforms
// the wave back, updates in registers and transforms forward.  The pass semantics are
// generic_tile's (scales, one FP64 error partial per work-group, the fixed-order sum in the
// update pass, block numbering), with this kernel's own lines per work-group.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_engine.hpp"
#include "generic.hpp"
#include "passes.hpp"

namespace qpch::dev {

template <int M, bool COL>
struct ChirpGeom {
  static_assert(M >= 256 && M <= 4096 && (M & (M - 1)) == 0, "M: a power of two from 256 to 4096");
  static constexpr int kT = M / 16;                       // threads per line
  static constexpr int kB = kT >= 256 ? 1 : 256 / kT;      // lines per work-group
  static constexpr int kThreads = kB * kT;                 // 256
  static constexpr int kWaves = kThreads / 64;
  static constexpr int kLine = M + M / 16;                 // the engine's padded line
  static constexpr int kStride = COL ? (kLine | 1) : kLine;
  static constexpr int kMaxN = M / 2;                      // M >= 2N - 1
  // LDS (float2): the lines, W_M, the chirp (N), bhat (M), one double per wave
  static constexpr int kLds = kB * kStride + M + kMaxN + M + kWaves;
};

// Host side: the chirp kernel's lines per work-group for a transform of length M.
__host__ __device__ constexpr int chirp_lines_per_block(int M) { return M / 16 >= 256 ? 1 : 256 / (M / 16); }

template <bool INV>
__device__ __forceinline__ float2 chirp_mul(float2 a, float2 w) {
  return INV ? cmulc(a, w) : cmul(a, w);
}

// The length-N DFT of the thread's line in registers (positions p = t + m M/16; inputs at
// p >= N are ignored, outputs at p >= N are zero).  Every thread of the work-group calls
// it.  ch / bh: the chirp and bhat tables in LDS; tw: W_M in LDS.
template <int M, bool INV>
__device__ __forceinline__ void chirp_dft(float2* v, int t, int n, float2* line, const float2* tw, const float2* ch,
                                          const float2* bh) {
  constexpr int kT = M / 16;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int p = t + m * kT;
    const float2 u = chirp_mul<INV>(v[m], ch[p < n ? p : 0]);
    v[m] = p < n ? u : float2{0.0f, 0.0f};
  }
  fft_line<M, 16, true>(v, t, line, tw);
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = chirp_mul<INV>(v[m], bh[t + m * kT]);
  fft_line<M, 16, false>(v, t, line, tw);
  const float inv_m = 1.0f / static_cast<float>(M);  // a power of two: exact
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int p = t + m * kT;
    const float2 u = chirp_mul<INV>(v[m], ch[p < n ? p : 0]);
    v[m] = p < n ? float2{u.x * inv_m, u.y * inv_m} : float2{0.0f, 0.0f};
  }
}

// The modulus or update pass over one tile (kB lines) of start `start`.
template <int M, bool COL, Mode MODE>
__device__ __forceinline__ void chirp_tile(const PassArgs& p, long long lines, int n, const float2* chirp,
                                           const float2* bhat, int start, int tile, float2* lds) {
  static_assert(MODE != Mode::kGrad, "no gradient mode");
  using G = ChirpGeom<M, COL>;
  float2* tw = lds + G::kB * G::kStride;
  float2* ch = tw + M;
  float2* bh = ch + G::kMaxN;
  double* red = reinterpret_cast<double*>(bh + M);
  const int tid = threadIdx.x;
  int b, t;  // the thread's line in the tile, and its butterfly index within the line
  if constexpr (COL) {
    b = tid % G::kB;
    t = tid / G::kB;
  } else {
    t = tid % G::kT;
    b = tid / G::kT;
  }
  const long long L = static_cast<long long>(tile) * G::kB + b;
  const bool valid = L < lines;
  // Loads are unconditional (all in flight at once): lanes past the last line read line 0
  // of the start, positions past the line's end read its element 0; neither is used.
  const long long Lc = valid ? L : 0;
  const int inner = p.inner;
  const long long base = (Lc / inner) * n * inner + Lc % inner;
  auto elem = [&](int m) {
    const int q = t + m * G::kT;
    return base + static_cast<long long>(q < n ? q : 0) * inner;
  };
  const long long soff = static_cast<long long>(start) * p.n;
  float2* w = static_cast<float2*>(p.w) + soff;
  float2* x = static_cast<float2*>(p.x) + soff;
  float2* line = lds + b * G::kStride;
  float2 v[16];
  auto tables = [&]() {  // while the loads are in flight
    for (int j = tid; j < M; j += G::kThreads) {
      double sn, cs;
      sincospi(2.0 * static_cast<double>(j) / static_cast<double>(M), &sn, &cs);
      tw[j] = {static_cast<float>(cs), static_cast<float>(-sn)};
      bh[j] = bhat[j];
    }
    for (int j = tid; j < n; j += G::kThreads) ch[j] = chirp[j];
  };
  if constexpr (MODE == Mode::kFwd || MODE == Mode::kInv) {
    // the transforms of the prologue and of the 3D line passes: read src (or the wave),
    // transform, write the wave scaled
    const float2* in = p.src ? static_cast<const float2*>(p.src) + soff : w;
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = in[elem(m)];
    tables();
    __syncthreads();
    chirp_dft<M, MODE == Mode::kInv>(v, t, n, line, tw, ch, bh);
    if (valid) {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if (t + m * G::kT < n) w[elem(m)] = {v[m].x * p.out_scale, v[m].y * p.out_scale};
      }
    }
  } else if constexpr (MODE == Mode::kModulus) {
    const long long arow = p.amp_stride * ((p.err_offset + start) / (p.amp_group > 1 ? p.amp_group : 1));
    float amp[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      v[m] = w[elem(m)];
      amp[m] = p.amp[arow + elem(m)];
    }
    tables();
    __syncthreads();
    chirp_dft<M, false>(v, t, n, line, tw, ch, bh);
    double acc = 0.0;
    if (valid) {
      const float io = p.in_scale * p.out_scale;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if (t + m * G::kT < n) v[m] = project_scaled(v[m], amp[m], p.in_scale, io, acc);
      }
    }
    const double ws = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = ws;
    if (tid == 0 && tile == 0 && start == 0 && p.counter) atomicAdd(p.counter, 1);
    chirp_dft<M, true>(v, t, n, line, tw, ch, bh);  // its barriers also publish red
    if (tid == 0) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < G::kWaves; ++k) s += red[k];
      p.partials[static_cast<long long>(start) * p.partial_stride + tile] = s;
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if (t + m * G::kT < n) w[elem(m)] = v[m];  // out_scale in io
      }
    }
  } else {
    const std::uint8_t* sup = p.support + soff;
    float2 xo[16];
    bool inside[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      v[m] = w[elem(m)];
      inside[m] = sup[elem(m)] != 0;
    }
    // ER (a = b = 0) skips the old iterate (outside the support the update is zero)
    if (p.a != 0.0f || p.b != 0.0f) {
#pragma unroll
      for (int m = 0; m < 16; ++m) xo[m] = x[elem(m)];
    } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) xo[m] = float2{0.0f, 0.0f};
    }
    tables();
    __syncthreads();
    chirp_dft<M, true>(v, t, n, line, tw, ch, bh);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const float2 pm = {v[m].x * p.in_scale, v[m].y * p.in_scale};
      float2 nx = pm;
      if (!inside[m]) nx = {p.a * xo[m].x + p.b * pm.x, p.a * xo[m].y + p.b * pm.y};
      if (valid && t + m * G::kT < n) x[elem(m)] = nx;
      v[m] = nx;
    }
    chirp_dft<M, false>(v, t, n, line, tw, ch, bh);
    if (valid) {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if (t + m * G::kT < n) w[elem(m)] = {v[m].x * p.out_scale, v[m].y * p.out_scale};
      }
    }
    if (tile == 0 && tid < 64 && p.errs) {
      // fixed-order sum of the previous modulus pass's partials of this start
      const double s = partial_sum<false>(p.partials + static_cast<long long>(start) * p.partial_stride,
                                          p.partial_stride, tid);
      if (tid == 0) {
        const int e = *p.counter - 1;
        if (e >= 0 && e < p.err_cap) p.errs[static_cast<long long>(e) * p.err_stride + p.err_offset + start] = s;
      }
    }
  }
}

// One chirp_tile per work-group; block numbering as generic_pass's (with this kernel's
// tiles).
template <int M, bool COL, Mode MODE>
__global__ void __launch_bounds__((ChirpGeom<M, COL>::kThreads))
    chirp_pass(PassArgs p, long long lines, int n, const float2* chirp, const float2* bhat) {
  __shared__ __align__(16) float2 lds[ChirpGeom<M, COL>::kLds];
  int tile, start;
  if (p.slots > 0) {
    start = blockIdx.x % p.slots;
    tile = blockIdx.x / p.slots;
    if (start >= p.starts) return;  // padding slot: the whole block leaves together
  } else {
    start = blockIdx.y;
    tile = blockIdx.x;
  }
  chirp_tile<M, COL, MODE>(p, lines, n, chirp, bhat, start, tile, lds);
}

}  // namespace qpch::dev
