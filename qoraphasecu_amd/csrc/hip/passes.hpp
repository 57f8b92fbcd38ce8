// Fused line passes of the projection iteration (power-of-two axes).
//
// One launch transforms every line of one axis of a batch of starts.  A line is N
// elements spaced `inner` apart; lines are numbered (o, c) with o in [0, outer) and c in
// [0, inner), element k of line (o, c) of start s at s*n + o*N*inner + k*inner + c.
// inner == 1 is a row pass (lines contiguous, B lines per work-group); inner > 1 is a
// column pass (B adjacent columns per work-group, so every load of the group covers B
// consecutive elements of a row).
//
// Modes (what a work-group does with each line):
//   kFwd / kInv  transform (optionally reading `src` instead of `w`), scale, store to w
//   kModulus     forward, Fourier-modulus projection with the error partial, inverse:
//                the frequency-axis pass; reference src/cuda/device.cuh:17-29 and
//                src/cpu/kernels.cpp:25-42
//   kUpdate      inverse -> P_M x, x = S ? P_M x : a x + b P_M x, forward of the new x:
//                the real-axis pass; reference src/cpu/kernels.cpp:44-50,
//                src/core/pointwise.hpp:34-45
//   kGrad        inverse -> P_M x, out = 2 (x - P_M x): the last pass of the amplitude
//                loss gradient (no reference counterpart)
// The error numerator of a start is reduced in a fixed order: the modulus pass writes one
// FP64 partial per wave, the next update pass (work-group 0 of each start) sums them in
// index order and stores errs[slot][start]; slot = *counter - 1, the modulus pass having
// advanced the counter.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "fft_engine.hpp"
#include "storage.hpp"

namespace qpch::dev {

enum class Mode : int { kFwd = 0, kInv = 1, kModulus = 2, kUpdate = 3, kGrad = 4 };

struct PassArgs {
  int outer;
  int inner;
  long long n;  // elements per start
  void* w;      // wave, St*
  const void* src;  // kFwd/kInv input when not null (St_in = St)
  void* x;          // kUpdate / kGrad: iterate, St*
  void* out;        // kGrad: gradient, St*
  const std::uint8_t* support;  // kUpdate: [start][n]
  const float* amp;             // kModulus: signed amplitude [n], -1 = masked
  long long amp_stride;         // 0: one amplitude for every start; n: one per problem
  int amp_group;                // consecutive starts per problem (0 or 1: one per start)
  double* partials;             // [start][partial_stride]: modulus pass, one per wave
  int partial_stride;
  double* errs;                 // [slot][err_stride]
  int err_stride;
  int err_offset;               // global index of this launch's first start
  int err_cap;                  // slots available
  int* counter;                 // error slot counter (null: no history, as in the loss)
  int* done;                    // image_iterations: work-groups finished in this launch (0 between launches)
  int starts;                   // starts in this launch
  int slots;                    // > 0: XCD-affine grid, block = tile * slots + start
  int no_xcd;                   // launcher input: 1 keeps the plain block numbering
  int radix;                    // launcher input: 8 selects the radix-8 kernels where built
  float a, b;                   // update coefficients
  float in_scale;               // applied after the first transform of the pass
  float out_scale;              // applied before the store (kModulus: after projection)
};

// Work-group geometry of a pass, shared by the kernel and its launcher.  EB is the
// storage size of one complex element: column passes take at least B adjacent columns so
// that one load instruction of a row covers 64 contiguous bytes (the two halves of a
// 128-byte line go to work-groups of the same start, hence of the same XCD's L2; halving
// the work-group from 32 to 16 columns at 256 x fp16 made the pass 5% faster).  R is the FFT radix: 16 holds 16
// elements per thread (fewer LDS exchanges), 8 holds 8 (twice the waves in flight).
// BF > 0 fixes the lines per work-group (the persistent launches size row tiles to the
// column pass's work-group; the lines per work-group never change a line's arithmetic).
template <int N, bool COL, int EB, int R = 16, int BF = 0>
struct PassGeom {
  using S = Shape1<N, R>;
  static constexpr int kT = S::kT;
  static constexpr int kR = S::kR;
  static constexpr int kLdsLimit = 96 * 1024 / 8 - N;  // float2 slots for the line tiles
  static constexpr int pow2_floor(int v) { return v <= 1 ? 1 : 2 * pow2_floor(v / 2); }
  static constexpr int clampB(int b) {
    int lim = kLdsLimit / (S::kLine + 1);
    if (1024 / kT < lim) lim = 1024 / kT;
    return pow2_floor(b < lim ? b : lim);
  }
  static constexpr int kColB = 64 / EB;
  static constexpr int kB =
      BF > 0 ? BF : clampB(COL ? (256 / kT > kColB ? 256 / kT : kColB) : (256 / kT > 1 ? 256 / kT : 1));
  static constexpr int kThreads = kB * kT;
  static constexpr int kWaves = (kThreads + 63) / 64;  // error partials per work-group
  // Column passes put consecutive lanes on different lines: an odd line stride (in float2)
  // spreads them over the LDS banks; a row pass's lanes walk along one line.
  static constexpr int kStride = COL ? (S::kLine | 1) : S::kLine;
  static constexpr int kTile = kB * kStride;  // float2
};

// Sum over the work-group in a fixed order; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int waves = (blockDim.x + 63) >> 6;
    for (int k = 0; k < waves; ++k) t += scratch[k];
  }
  return t;
}

// W_N^j = exp(-2 pi i j / N) into LDS, evaluated in double and rounded once.
template <int N>
__device__ __forceinline__ void load_twiddles(float2* tw) {
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    double s, c;
    sincospi(2.0 * static_cast<double>(j) / static_cast<double>(N), &s, &c);
    tw[j] = {static_cast<float>(c), static_cast<float>(-s)};
  }
}

// One element of the modulus projection: w is the (scaled) spectrum value, a its measured
// amplitude or -1 when masked.  Usable pixels add (|w| - a)^2 to acc and take magnitude
// a; masked ones keep their own.  Reference semantics: src/cpu/kernels.cpp:29-39.
// The magnitude and the ratio use the hardware square root and reciprocal (each within
// about one ulp) rather than the correctly rounded sequences, which cost as many
// instructions as a radix-16 stage; the selects keep the lanes branch-free.
__device__ __forceinline__ float2 project(float2 w, float a, double& acc) {
  const float mag = __builtin_amdgcn_sqrtf(w.x * w.x + w.y * w.y);
  const bool usable = a >= 0.0f;
  const float d = usable ? mag - a : 0.0f;
  acc += static_cast<double>(d) * static_cast<double>(d);
  const float target = usable ? a : mag;
  const float f = target * __builtin_amdgcn_rcpf(mag + 1e-12f);
  return {w.x * f, w.y * f};
}

// One thread's strided view of one start's array: register m is element e0 + m * em.
// Column passes (BUF) address it through a buffer resource: the stride em is a runtime
// value that the SGPR offset carries, so a load costs no per-lane 64-bit address
// arithmetic.  Row passes use the pointer: there em is a compile-time constant that folds
// into the instructions' immediate offsets.
template <class T, bool BUF>
struct View {
  T* p;
  Rsrc r;
  int e0, em;
  __device__ __forceinline__ View(T* base, long long count, int first, int step)
      : p(base), r(make_rsrc(base, count * static_cast<long long>(sizeof(T)))), e0(first), em(step) {}
};

// COH: a device-scope load, for data another CU may have written in the same launch.
template <bool COH = false, class St, bool BUF>
__device__ __forceinline__ float2 vld(const View<St, BUF>& v, int m) {
  constexpr int EB = static_cast<int>(sizeof(St));
  using T = std::remove_const_t<St>;
  if constexpr (BUF) {
    return Storage<T>::template bld<COH>(v.r, v.e0 * EB, m * v.em * EB);
  } else {
    return Storage<T>::template ld<COH>(v.p, static_cast<size_t>(v.e0 + m * v.em));
  }
}
template <class St, bool BUF>
__device__ __forceinline__ void vst(const View<St, BUF>& v, int m, float2 x) {
  constexpr int EB = static_cast<int>(sizeof(St));
  if constexpr (BUF) {
    Storage<St>::bst(v.r, v.e0 * EB, m * v.em * EB, x);
  } else {
    Storage<St>::st(v.p, static_cast<size_t>(v.e0 + m * v.em), x);
  }
}
template <bool BUF>
__device__ __forceinline__ float vld_f32(const View<const float, BUF>& v, int m) {
  if constexpr (BUF) {
    return buf_f32(v.r, v.e0 * 4, m * v.em * 4);
  } else {
    return v.p[v.e0 + m * v.em];
  }
}
template <bool BUF>
__device__ __forceinline__ bool vld_flag(const View<const std::uint8_t, BUF>& v, int m) {
  if constexpr (BUF) {
    return buf_u8(v.r, v.e0, m * v.em) != 0;
  } else {
    return v.p[v.e0 + m * v.em] != 0;
  }
}

// The modulus projection of a pass whose input and output carry scales: v is the unscaled
// spectrum value and the projection acts on u = in * v, so |u| = in |v|; the result,
// out * P(u), is v times one real factor (io = in * out, uniform), which saves the separate
// scalings of u and of the output.  Semantics as project().
__device__ __forceinline__ float2 project_scaled(float2 v, float a, float in, float io, double& acc) {
  const float mag = in * __builtin_amdgcn_sqrtf(v.x * v.x + v.y * v.y);
  const bool usable = a >= 0.0f;
  const float d = usable ? mag - a : 0.0f;
  acc += static_cast<double>(d) * static_cast<double>(d);
  const float target = usable ? a : mag;
  const float g = io * target * __builtin_amdgcn_rcpf(mag + 1e-12f);
  return {v.x * g, v.y * g};
}

// Sum over the 64 lanes of a wave in a fixed order; the result is valid in lane 0.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// The error numerator of one start from its partials, in a fixed order: lane l adds
// part[l], part[l + 64], ... in turn, then the wave's shuffle tree; valid in lane 0.  The
// loads go out eight at a time, so a long row (2048 partials per start at 128^3) costs a
// few round trips rather than one per element, with the same order of additions.
template <bool COH = false>
__device__ __forceinline__ double partial_sum(const double* part, int count, int lane) {
  constexpr int kU = 8;
  auto ld = [&](int i) {
    if constexpr (COH) {
      return load_coherent(part + i);
    } else {
      return part[i];
    }
  };
  double s = 0.0;
  int k = lane;
  for (; k + (kU - 1) * 64 < count; k += kU * 64) {
    double a[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) a[u] = ld(k + u * 64);
#pragma unroll
    for (int u = 0; u < kU; ++u) s += a[u];
  }
  for (; k < count; k += 64) s += ld(k);
  return wave_sum(s);
}

// Slot argument of line_tile meaning "the error slot is *p.counter - 1" (the modulus pass
// of the same iteration advanced the counter in an earlier launch).
constexpr int kSlotFromCounter = -2147483647 - 1;

// One tile of a pass: B lines of start `start` (launch-relative), tile index `tile`.
// lds holds the tile's line scratch (G::kTile float2); tw the axis's twiddle table W_N^j.
// With OWN_TW the tile fills tw itself while its loads are in flight and then synchronises
// the work-group; without, tw is already filled and visible to every thread.  `slot` is
// the error-history slot the update pass writes (kSlotFromCounter: read the counter).
// Every thread of the work-group must call this (the transforms hold barriers).
// COH: the tile's inputs may have been written by other CUs in the same launch (the
// persistent launches): loads of the wave, the iterate and the partials are device scope.
template <int N, bool COL, Mode M, class St, int R, bool OWN_TW, bool COH = false, int BF = 0>
__device__ __forceinline__ void line_tile(const PassArgs& p, int start, int tile, float2* lds, float2* tw,
                                          int slot) {
  using G = PassGeom<N, COL, sizeof(St), R, BF>;
  constexpr int kT = G::kT;
  constexpr int kR = G::kR;
  constexpr int kB = G::kB;
  constexpr int kWaves = G::kWaves;
  const int tid = threadIdx.x;
  int b, i;
  if constexpr (COL) {
    b = tid % kB;
    i = tid / kB;
  } else {
    i = tid % kT;
    b = tid / kT;
  }
  long long base;  // element 0 of this thread's line, relative to its start
  bool valid;
  if constexpr (COL) {
    const int groups = (p.inner + kB - 1) / kB;
    const int o = tile / groups;
    const int c = (tile % groups) * kB + b;
    valid = c < p.inner;
    base = static_cast<long long>(o) * N * p.inner + c;
  } else {
    const int o = tile * kB + b;
    valid = o < p.outer;
    base = static_cast<long long>(o) * N;
  }
  // Register m of this thread is element e0 + m * em of its start.  Loads are
  // unconditional so that all of a thread's loads are in flight at once (a load under a
  // branch gets its own wait): a lane past the last line of a ragged tile reads line 0 of
  // its start instead, and its results are never stored.
  const int es = COL ? p.inner : 1;
  const int e0 = static_cast<int>(valid ? base : 0) + i * es;
  const int em = kT * es;
  const long long soff = static_cast<long long>(start) * p.n;
  float2* line = lds + b * G::kStride;
  using SV = View<St, COL>;
  const SV vw(static_cast<St*>(p.w) + soff, p.n, e0, em);
  float2 v[kR];

  if constexpr (M == Mode::kFwd || M == Mode::kInv) {
    const SV vin(p.src ? static_cast<St*>(const_cast<void*>(p.src)) + soff : vw.p, p.n, e0, em);
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = vld<COH>(vin, m);
    if constexpr (OWN_TW) {
      load_twiddles<N>(tw);  // while the loads are in flight
      __syncthreads();
    }
    fft_line<N, R, M == Mode::kFwd>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) vst(vw, m, float2{v[m].x * p.out_scale, v[m].y * p.out_scale});
    }
  } else if constexpr (M == Mode::kModulus) {
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = vld<COH>(vw, m);
    float amp[kR];
    const long long arow = p.amp_stride * ((p.err_offset + start) / (p.amp_group > 1 ? p.amp_group : 1));
    const View<const float, COL> va(p.amp + arow, p.n, e0, em);
#pragma unroll
    for (int m = 0; m < kR; ++m) amp[m] = vld_f32(va, m);
    if constexpr (OWN_TW) {
      load_twiddles<N>(tw);
      __syncthreads();
    }
    fft_line<N, R, true>(v, i, line, tw);
    double acc = 0.0;
    if (valid) {
      const float io = p.in_scale * p.out_scale;
#pragma unroll
      for (int m = 0; m < kR; ++m) v[m] = project_scaled(v[m], amp[m], p.in_scale, io, acc);
    }
    // Each wave's share of the error numerator goes out now, one partial per wave: its
    // shuffles overlap the inverse transform, and no block-wide reduction ends the pass.
    const double wsum = wave_sum(acc);
    if ((tid & 63) == 0) {
      p.partials[static_cast<long long>(start) * p.partial_stride + tile * kWaves + (tid >> 6)] = wsum;
    }
    if (tid == 0 && tile == 0 && start == 0 && p.counter) atomicAdd(p.counter, 1);
    fft_line<N, R, false>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) vst(vw, m, v[m]);
    }
  } else if constexpr (M == Mode::kGrad) {
    const SV vx(static_cast<St*>(p.x) + soff, p.n, e0, em);
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = vld<COH>(vw, m);
    float2 xo[kR];
#pragma unroll
    for (int m = 0; m < kR; ++m) xo[m] = vld<COH>(vx, m);
    if constexpr (OWN_TW) {
      load_twiddles<N>(tw);
      __syncthreads();
    }
    fft_line<N, R, false>(v, i, line, tw);
    if (valid) {
      const SV vo(static_cast<St*>(p.out) + soff, p.n, e0, em);
#pragma unroll
      for (int m = 0; m < kR; ++m) {
        const float2 pm = {v[m].x * p.in_scale, v[m].y * p.in_scale};
        vst(vo, m, float2{2.0f * (xo[m].x - pm.x), 2.0f * (xo[m].y - pm.y)});
      }
    }
  } else {  // kUpdate
    const SV vx(static_cast<St*>(p.x) + soff, p.n, e0, em);
    const View<const std::uint8_t, COL> vs(p.support + soff, p.n, e0, em);
#pragma unroll
    for (int m = 0; m < kR; ++m) v[m] = vld<COH>(vw, m);
    // the old iterate and the support do not depend on the transform: fetch them now.  ER
    // (a = b = 0) skips the iterate: outside the support a x + b P_M x is zero whatever x
    // holds (only the sign of that zero can differ), and the pass moves 8 (fp32) or 4
    // (fp16) bytes less per element.
    float2 xo[kR];
    bool inside[kR];
#pragma unroll
    for (int m = 0; m < kR; ++m) inside[m] = vld_flag(vs, m);
    if (p.a != 0.0f || p.b != 0.0f) {
#pragma unroll
      for (int m = 0; m < kR; ++m) xo[m] = vld<COH>(vx, m);
    } else {
#pragma unroll
      for (int m = 0; m < kR; ++m) xo[m] = float2{0.0f, 0.0f};
    }
    if constexpr (OWN_TW) {
      load_twiddles<N>(tw);
      __syncthreads();
    }
    fft_line<N, R, false>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) {
        const float2 pm = {v[m].x * p.in_scale, v[m].y * p.in_scale};
        float2 nx = pm;
        if (!inside[m]) nx = {p.a * xo[m].x + p.b * pm.x, p.a * xo[m].y + p.b * pm.y};
        vst(vx, m, nx);
        v[m] = nx;
      }
    }
    fft_line<N, R, true>(v, i, line, tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kR; ++m) vst(vw, m, float2{v[m].x * p.out_scale, v[m].y * p.out_scale});
    }
    if (tile == 0 && tid < 64 && p.errs) {
      // fixed-order sum of the previous modulus pass's partials of this start
      const double s = partial_sum<COH>(p.partials + static_cast<long long>(start) * p.partial_stride,
                                   p.partial_stride, tid);
      if (tid == 0) {
        const int e = slot == kSlotFromCounter ? *p.counter - 1 : slot;
        if (e >= 0 && e < p.err_cap) {
          p.errs[static_cast<long long>(e) * p.err_stride + p.err_offset + start] = s;
        }
      }
    }
  }
}

// Block numbering: with p.slots > 0 (a multiple of 8) block L handles start L % slots and
// tile L / slots, so every block of a start lands on the same XCD (blocks are dealt to the
// eight XCDs round-robin) and the start's arrays stay in that XCD's L2 from one pass to the
// next; otherwise block L handles tile L % tiles of start L / tiles.
template <int N, bool COL, Mode M, class St, int R>
__global__ void __launch_bounds__((PassGeom<N, COL, sizeof(St), R>::kThreads))
    line_pass(PassArgs p) {
  using G = PassGeom<N, COL, sizeof(St), R>;
  __shared__ float2 lds[G::kTile + N];
  int tile, start;
  if (p.slots > 0) {
    start = blockIdx.x % p.slots;
    tile = blockIdx.x / p.slots;
    if (start >= p.starts) return;  // padding slot: the whole block leaves together
  } else {
    const int tiles = gridDim.x / p.starts;
    start = blockIdx.x / tiles;
    tile = blockIdx.x % tiles;
  }
  line_tile<N, COL, M, St, R, true>(p, start, tile, lds, lds + G::kTile, kSlotFromCounter);
}

}  // namespace qpch::dev
