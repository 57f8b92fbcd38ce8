// Kernels for axes that are not powers of two (any length up to 2048): the fused passes of
// line_pass in a generic mixed-radix form (generic_pass, used by the fused schedules on
// those axes), the plain line transform (generic_fft, for the standalone operators, the
// unfused schedule and the FP64 adjoint), and the element-wise kernels of the unfused
// iteration.  Power-of-two axes up to 4096 use the radix-16 passes of passes.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "fft_engine.hpp"
#include "passes.hpp"

namespace qpch::dev {

constexpr int kGenericMaxN = 2048;
// Every generic kernel is launched with exactly kGenericThreads threads, and the
// work-group loops stride by this constant: blockDim.x is a vector load from the launch's
// implicit arguments, and before each stage loop it cost a wait on all outstanding loads.
constexpr int kGenericThreads = 256;
// generic_pass keeps a block's elements (at most 2048: one line up to kGenericMaxN, or up
// to 16 lines of fewer than 1024) in 2048 / kGenericThreads registers per thread.
static_assert(kGenericMaxN <= 2048 && 2048 % kGenericThreads == 0, "generic_pass register slots");

struct GenericPlan {
  int n;
  int nfac;  // 0: one plain DFT of the whole line (or Bluestein when m > 0)
  int fac[32];
  // Bluestein (chirp-z) for lengths with a prime factor above 64 from kBluesteinMinN on:
  // the line times the chirp c_j = exp(-i pi j^2 / n), zero-padded to m = 2^k >= 2n - 1,
  // transformed (radix-4/2 stages, factors mfac), multiplied by the transformed chirp
  // filter bhat, transformed back and multiplied by the chirp again.  The inverse
  // transform conjugates both tables.  m = 0: not used.
  int m;
  int nmfac;
  int mfac[8];
  const void* chirp;  // n values (float2 or double2, the kernel's precision)
  const void* bhat;   // m values
};

constexpr int kBluesteinMinN = 256;

template <class Real>
struct Cplx;
template <>
struct Cplx<float> {
  using T = float2;
};
template <>
struct Cplx<double> {
  using T = double2;
};

template <class C>
__device__ __forceinline__ C gadd(C a, C b) {
  return {a.x + b.x, a.y + b.y};
}
template <class C>
__device__ __forceinline__ C gmul(C a, C b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Lines per work-group for a line of length N: enough that short lines keep the threads
// of a stage busy (a radix-r stage has N/r butterflies per line).
__host__ __device__ constexpr int generic_lines_per_block(int N) {
  return N >= 1024 ? 1 : (1024 / N < 16 ? 1024 / N : 16);
}

// Elements of one line slot in LDS, and of the twiddle table: the padded length m of a
// Bluestein plan, else the line length.
__host__ __device__ constexpr int generic_slot(const GenericPlan& plan) { return plan.m > 0 ? plan.m : plan.n; }

// Dynamic LDS of generic_fft: the twiddle table, two line buffers per line, the line bases.
template <class C>
__host__ __device__ constexpr size_t generic_lds_bytes(const GenericPlan& plan) {
  const int lpb = generic_lines_per_block(plan.n);
  return sizeof(C) * static_cast<size_t>(generic_slot(plan)) * (1 + 2 * lpb) + sizeof(long long) * lpb;
}

// a / b for 0 <= a < 2^20 and 1 <= b: the quotient from the float reciprocal is off by at
// most one at these sizes, and one step corrects it (exact integer result, a few
// instructions instead of an integer division sequence).
__device__ __forceinline__ int small_div(int a, int b) {
  int q = static_cast<int>(static_cast<float>(a) * __builtin_amdgcn_rcpf(static_cast<float>(b)));
  const int r = a - q * b;
  q += (r >= b) - (r < 0);
  return q;
}

// Twiddle j of a table, conjugated for an inverse transform read from a forward table
// (negation is exact, so the values equal those of an inverse table).
template <bool INV, class C>
__device__ __forceinline__ C twid(const C* tw, int j) {
  const C w = tw[j];
  return INV ? C{w.x, -w.y} : w;
}

// a * (-i) for a forward transform, a * (+i) for an inverse one (exact)
template <bool INV, class C>
__device__ __forceinline__ C rot(C a) {
  return INV ? C{-a.y, a.x} : C{a.y, -a.x};
}

// In-register DFT of size R in {2, 3, 4, 5} with W_R = exp(-+ 2 pi i / R), natural order in
// and out.  Radix 3 and 5 use the textbook short forms: with s = x1 + x2, d = x1 - x2,
//   X0 = x0 + s,  X1,2 = (x0 - s/2) -+ i sin(2pi/3) d;
// and with s1 = x1 + x4, d1 = x1 - x4, s2 = x2 + x3, d2 = x2 - x3,
//   X0 = x0 + s1 + s2,
//   X1,4 = x0 + cos(2pi/5) s1 + cos(4pi/5) s2 -+ i (sin(2pi/5) d1 + sin(4pi/5) d2),
//   X2,3 = x0 + cos(4pi/5) s1 + cos(2pi/5) s2 -+ i (sin(4pi/5) d1 - sin(2pi/5) d2)
// (upper signs forward).  Four and two are the split radix-2 butterflies.
template <int R, bool INV, class C>
__device__ __forceinline__ void dft_small(C* v) {
  using Re = decltype(C{}.x);
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  } else if constexpr (R == 4) {
    const C s0 = {v[0].x + v[2].x, v[0].y + v[2].y}, d0 = {v[0].x - v[2].x, v[0].y - v[2].y};
    const C s1 = {v[1].x + v[3].x, v[1].y + v[3].y}, d1 = rot<INV>(C{v[1].x - v[3].x, v[1].y - v[3].y});
    v[0] = {s0.x + s1.x, s0.y + s1.y};
    v[2] = {s0.x - s1.x, s0.y - s1.y};
    v[1] = {d0.x + d1.x, d0.y + d1.y};
    v[3] = {d0.x - d1.x, d0.y - d1.y};
  } else if constexpr (R == 3) {
    constexpr Re kS = Re(0.86602540378443864676);  // sin(2 pi / 3)
    const C a = v[0];
    const C s = {v[1].x + v[2].x, v[1].y + v[2].y};
    const C d = rot<INV>(C{kS * (v[1].x - v[2].x), kS * (v[1].y - v[2].y)});
    const C t = {a.x - Re(0.5) * s.x, a.y - Re(0.5) * s.y};
    v[0] = {a.x + s.x, a.y + s.y};
    v[1] = {t.x + d.x, t.y + d.y};
    v[2] = {t.x - d.x, t.y - d.y};
  } else {
    static_assert(R == 5, "short forms exist for radix 2, 3, 4 and 5");
    constexpr Re kC1 = Re(0.30901699437494742410);   // cos(2 pi / 5)
    constexpr Re kC2 = Re(-0.80901699437494742410);  // cos(4 pi / 5)
    constexpr Re kS1 = Re(0.95105651629515357212);   // sin(2 pi / 5)
    constexpr Re kS2 = Re(0.58778525229247312917);   // sin(4 pi / 5)
    const C a = v[0];
    const C s1 = {v[1].x + v[4].x, v[1].y + v[4].y}, d1 = {v[1].x - v[4].x, v[1].y - v[4].y};
    const C s2 = {v[2].x + v[3].x, v[2].y + v[3].y}, d2 = {v[2].x - v[3].x, v[2].y - v[3].y};
    const C t1 = {a.x + kC1 * s1.x + kC2 * s2.x, a.y + kC1 * s1.y + kC2 * s2.y};
    const C t2 = {a.x + kC2 * s1.x + kC1 * s2.x, a.y + kC2 * s1.y + kC1 * s2.y};
    const C u1 = rot<INV>(C{kS1 * d1.x + kS2 * d2.x, kS1 * d1.y + kS2 * d2.y});
    const C u2 = rot<INV>(C{kS2 * d1.x - kS1 * d2.x, kS2 * d1.y - kS1 * d2.y});
    v[0] = {a.x + s1.x + s2.x, a.y + s1.y + s2.y};
    v[1] = {t1.x + u1.x, t1.y + u1.y};
    v[4] = {t1.x - u1.x, t1.y - u1.y};
    v[2] = {t2.x + u2.x, t2.y + u2.y};
    v[3] = {t2.x - u2.x, t2.y - u2.y};
  }
}

// One autosort stage of radix r over the nl lines of a block.  Output j of butterfly t
// (k = t mod L) is  sum_q W_N^(q k m) W_r^(j q) src[t + q N/r]  (m = N / (L r)), written to
// dst[(t - k) r + k + j L].  R in {2, 3, 4, 5} fixes the radix at compile time (inputs
// held in registers, the short forms of dft_small); R = 0 takes r at run time and sums the
// O(r^2) terms with W_r from the table, advancing (j q) mod r by j per term instead of
// dividing.  INV conjugates the table's twiddles.
template <int R, bool INV, class C>
__device__ __forceinline__ void generic_stage(const C* src, C* dst, const C* tw, int N, int r, int L,
                                              int nl) {
  const int rr = R > 0 ? R : r;
  const int nr = N / rr;
  const int m = N / (L * rr);
  for (int t = threadIdx.x; t < nl * nr; t += kGenericThreads) {
    const int l = small_div(t, nr);
    const int tl = t - l * nr;
    const int k = tl - small_div(tl, L) * L;
    const int outb = (tl - k) * rr + k;
    const C* sl = src + l * N;
    C* dl = dst + l * N;
    if constexpr (R > 0) {
      // every input and twiddle read issued before the first use: no branch around the
      // twiddle reads (entry 0 is always there), the product selected where tq != 0
      C v[R], w[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        v[q] = sl[tl + q * nr];
        if (q > 0) w[q] = twid<INV>(tw, q * k * m);  // stage twiddle W_{L r}^{q k}
      }
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const C t = gmul(v[q], w[q]);
        if (k != 0) v[q] = t;  // tq = q k m is zero exactly when k is
      }
      dft_small<R, INV>(v);
#pragma unroll
      for (int j = 0; j < R; ++j) dl[outb + j * L] = v[j];
    } else {
      for (int j = 0; j < rr; ++j) {
        C acc = {0, 0};
        int e = 0;  // (j q) mod r
        for (int q = 0; q < rr; ++q) {
          C v = sl[tl + q * nr];
          const int tq = q * k * m;
          if (tq) v = gmul(v, twid<INV>(tw, tq));
          acc = gadd(acc, e ? gmul(v, twid<INV>(tw, e * nr)) : v);
          e += j;
          if (e >= rr) e -= rr;
        }
        dl[outb + j * L] = acc;
      }
    }
  }
}

// The unnormalised length-n DFT of the nl lines in src (line l at l*n) by Bluestein's chirp-z
// convolution (plan.m > 0).  Both buffers hold nl slots of m elements; tw is the forward
// W_m table.  The result is in src on return (line l at l*n).  Ends with a barrier.
template <bool INV, class C>
__device__ __forceinline__ void bluestein(C*& src, C*& dst, const C* tw, const GenericPlan& plan, int nl) {
  const int N = plan.n, Mm = plan.m;
  const C* chirp = static_cast<const C*>(plan.chirp);
  const C* bhat = static_cast<const C*>(plan.bhat);
  // a_j = x_j c_j, zero-padded to m
  for (int t = threadIdx.x; t < nl * Mm; t += kGenericThreads) {
    const int l = small_div(t, Mm);
    const int j = t - l * Mm;
    dst[t] = j < N ? gmul(src[l * N + j], twid<INV>(chirp, j)) : C{0, 0};
  }
  __syncthreads();
  C* cur = dst;
  C* oth = src;
  auto stages = [&](auto inverse) {
    constexpr bool kInv = decltype(inverse)::value;
    int L = 1;
    for (int s = 0; s < plan.nmfac; ++s) {
      const int r = plan.mfac[s];
      if (r == 4) {
        generic_stage<4, kInv>(cur, oth, tw, Mm, r, L, nl);
      } else {
        generic_stage<2, kInv>(cur, oth, tw, Mm, r, L, nl);
      }
      __syncthreads();
      C* t = cur;
      cur = oth;
      oth = t;
      L *= r;
    }
  };
  stages(std::false_type{});
  for (int t = threadIdx.x; t < nl * Mm; t += kGenericThreads) {
    const int j = t - small_div(t, Mm) * Mm;
    cur[t] = gmul(cur[t], twid<INV>(bhat, j));
  }
  __syncthreads();
  stages(std::true_type{});
  // X_k = c_k conv_k / m (1/m is a power of two: exact)
  using R = decltype(C{}.x);
  const R inv_m = R(1) / static_cast<R>(Mm);
  for (int t = threadIdx.x; t < nl * N; t += kGenericThreads) {
    const int l = small_div(t, N);
    const int k = t - l * N;
    const C v = gmul(cur[l * Mm + k], twid<INV>(chirp, k));
    oth[t] = {v.x * inv_m, v.y * inv_m};
  }
  __syncthreads();
  src = oth;
  dst = cur;
}

// The unnormalised DFT of the nl lines in src (LDS, line l at l*N), by the plan's stages or,
// for nfac == 0, one plain DFT per line; src and dst swap roles once per stage, so the
// result is in src on return.  Ends with a barrier.
template <bool INV, class C>
__device__ __forceinline__ void generic_transform(C*& src, C*& dst, const C* tw, const GenericPlan& plan, int nl) {
  const int N = plan.n;
  if (plan.m > 0) {
    bluestein<INV>(src, dst, tw, plan, nl);
    return;
  }
  if (plan.nfac == 0) {
    for (int t = threadIdx.x; t < nl * N; t += kGenericThreads) {
      const int l = small_div(t, N);
      const int j = t - l * N;
      const C* sl = src + l * N;
      C acc = {0, 0};
      int idx = 0;  // (j q) mod N
      for (int q = 0; q < N; ++q) {
        acc = gadd(acc, gmul(sl[q], twid<INV>(tw, idx)));
        idx += j;
        if (idx >= N) idx -= N;
      }
      dst[t] = acc;
    }
    __syncthreads();
    C* t = src;
    src = dst;
    dst = t;
    return;
  }
  int L = 1;
  for (int s = 0; s < plan.nfac; ++s) {
    const int r = plan.fac[s];
    switch (r) {
      case 2: generic_stage<2, INV>(src, dst, tw, N, r, L, nl); break;
      case 3: generic_stage<3, INV>(src, dst, tw, N, r, L, nl); break;
      case 4: generic_stage<4, INV>(src, dst, tw, N, r, L, nl); break;
      case 5: generic_stage<5, INV>(src, dst, tw, N, r, L, nl); break;
      default: generic_stage<0, INV>(src, dst, tw, N, r, L, nl); break;
    }
    __syncthreads();
    C* tt = src;
    src = dst;
    dst = tt;
    L *= r;
  }
}

// generic_lines_per_block(N) lines per work-group, autosort mixed-radix stages ping-ponging
// between two LDS buffers; radix-r butterflies are evaluated as O(r^2) sums with the
// line's own W_N table (evaluated in double, rounded once to Real).  Lines are addressed
// like the fused passes: element k of line L = o*inner + c of start s at
// s*n + o*N*inner + k*inner + c.  Loads and stores walk adjacent lines first when the
// axis is strided (inner > 1) and along the line otherwise, so that neighbouring threads
// touch neighbouring addresses.  Real = float (the unfused schedule) or double (the FP64
// adjoint).
template <class Real>
__global__ void __launch_bounds__(kGenericThreads)
    generic_fft(typename Cplx<Real>::T* data, long long n_per_start, int inner, long long lines,
                GenericPlan plan, int forward, Real scale) {
  using C = typename Cplx<Real>::T;
  extern __shared__ __align__(16) unsigned char sm_raw[];
  C* sm = reinterpret_cast<C*>(sm_raw);
  const int N = plan.n;
  const int lpb = generic_lines_per_block(N);
  const int slot = generic_slot(plan);
  C* tw = sm;
  C* bufa = sm + slot;
  C* bufb = bufa + lpb * slot;
  long long* lbase = reinterpret_cast<long long*>(bufb + lpb * slot);
  // the forward table (W_N, or W_m of a Bluestein plan); an inverse transform conjugates
  // its entries as it reads them (the stages' template argument carries the direction)
  for (int j = threadIdx.x; j < slot; j += kGenericThreads) {
    double s, c;
    sincospi(2.0 * static_cast<double>(j) / static_cast<double>(slot), &s, &c);
    tw[j] = {static_cast<Real>(c), static_cast<Real>(-s)};
  }
  const long long line0 = static_cast<long long>(blockIdx.x) * lpb;
  const int nl = static_cast<int>(lines - line0 < lpb ? lines - line0 : lpb);
  if (static_cast<int>(threadIdx.x) < nl) {
    const long long L = line0 + threadIdx.x;
    lbase[threadIdx.x] = static_cast<long long>(blockIdx.y) * n_per_start + (L / inner) * N * inner + L % inner;
  }
  __syncthreads();
  const bool strided = inner > 1;
  for (int e = threadIdx.x; e < nl * N; e += kGenericThreads) {
    const int l = strided ? e - small_div(e, nl) * nl : small_div(e, N);
    const int k = strided ? small_div(e, nl) : e - small_div(e, N) * N;
    bufa[l * N + k] = data[lbase[l] + static_cast<long long>(k) * inner];
  }
  __syncthreads();
  C* src = bufa;
  C* dst = bufb;
  if (forward) {
    generic_transform<false>(src, dst, tw, plan, nl);
  } else {
    generic_transform<true>(src, dst, tw, plan, nl);
  }
  for (int e = threadIdx.x; e < nl * N; e += kGenericThreads) {
    const int l = strided ? e - small_div(e, nl) * nl : small_div(e, N);
    const int k = strided ? small_div(e, nl) : e - small_div(e, N) * N;
    const C v = src[l * N + k];
    data[lbase[l] + static_cast<long long>(k) * inner] = {v.x * scale, v.y * scale};
  }
}

// Dynamic LDS of generic_pass: the forward twiddle table, two line buffers per line, the
// line bases and one FP64 slot per wave.
__host__ __device__ constexpr size_t generic_pass_lds_bytes(const GenericPlan& plan) {
  return generic_lds_bytes<float2>(plan) + sizeof(double) * (kGenericThreads / 64);
}

// One tile of the fused passes of line_pass (passes.hpp) for an axis of any length up to
// kGenericMaxN, fp32 storage: the same modes, arguments and scale conventions, with
// generic_lines_per_block(N) lines of start `start` (launch-relative) per work-group.  The
// modulus pass leaves one FP64 error partial per tile (its waves' sums added in order); the
// update pass's tile 0 sums the start's partial_stride partials in order, as line_pass
// does, into history slot `eslot` (kSlotFromCounter: *p.counter - 1).  LDS: tw (the
// plan's forward table, filled here with OWN_TW, else already filled and visible), two
// line buffers of generic_lines_per_block(N) x generic_slot(plan) elements, the line bases
// and one FP64 slot per wave.  COH: inputs written by other CUs in the same launch (the
// persistent launches) are read with device-scope loads.  Every thread of the work-group
// must call this (the transforms hold barriers).
template <Mode M, bool OWN_TW, bool COH = false>
__device__ __forceinline__ void generic_tile(const PassArgs& p, const GenericPlan& plan, long long lines, int start,
                                             int tile, float2* tw, float2* bufa, float2* bufb, long long* lbase,
                                             double* red, int eslot) {
  const int N = plan.n;
  const int lpb = generic_lines_per_block(N);
  const int slot = generic_slot(plan);
  const long long soff = static_cast<long long>(start) * p.n;
  const long long line0 = static_cast<long long>(tile) * lpb;
  const int nl = static_cast<int>(lines - line0 < lpb ? lines - line0 : lpb);
  const int inner = p.inner;
  if (static_cast<int>(threadIdx.x) < nl) {
    const long long L = line0 + threadIdx.x;
    lbase[threadIdx.x] = (L / inner) * N * inner + L % inner;  // element 0 of the line, within the start
  }
  __syncthreads();
  // Element e = threadIdx.x + u * kGenericThreads of the block's work is line l, position k
  // (adjacent lines first on a strided axis).  A block holds at most 2048 elements, so a
  // thread has at most kPer.  Every slot loads (slots past the end re-read the last
  // element), so that all of a thread's global loads are in flight before the first use;
  // only slots below the end are written back.
  constexpr int kPer = 2048 / kGenericThreads;
  const int total = nl * N;
  const bool strided = inner > 1;
  float2* w = static_cast<float2*>(p.w) + soff;
  const float2* in = (M == Mode::kFwd || M == Mode::kInv) && p.src ? static_cast<const float2*>(p.src) + soff : w;
  float2* x = static_cast<float2*>(p.x) + soff;
  const std::uint8_t* sup = p.support + soff;
  const long long arow =
      M == Mode::kModulus ? p.amp_stride * ((p.err_offset + start) / (p.amp_group > 1 ? p.amp_group : 1)) : 0;
  long long gi[kPer];  // element index within the start
  int li[kPer];        // LDS index l * N + k
  float2 v0[kPer];
  float amp[kPer];
  float2 xo[kPer];
  bool inside[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int e0 = threadIdx.x + u * kGenericThreads;
    const int e = e0 < total ? e0 : total - 1;
    int l, k;
    if (strided) {
      k = small_div(e, nl);
      l = e - k * nl;
    } else {
      l = small_div(e, N);
      k = e - l * N;
    }
    li[u] = l * N + k;
    gi[u] = lbase[l] + static_cast<long long>(k) * inner;
    v0[u] = Storage<float2>::ld<COH>(in, static_cast<size_t>(gi[u]));
    if constexpr (M == Mode::kModulus) amp[u] = p.amp[arow + gi[u]];
    if constexpr (M == Mode::kUpdate) {
      xo[u] = Storage<float2>::ld<COH>(x, static_cast<size_t>(gi[u]));
      inside[u] = sup[gi[u]] != 0;
    }
  }
  if constexpr (OWN_TW) {
    // the twiddle table while the loads are in flight (forward W_N^j, or W_m^j of a
    // Bluestein plan)
    for (int j = threadIdx.x; j < slot; j += kGenericThreads) {
      double sn, cs;
      sincospi(2.0 * static_cast<double>(j) / static_cast<double>(slot), &sn, &cs);
      tw[j] = {static_cast<float>(cs), static_cast<float>(-sn)};
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (static_cast<int>(threadIdx.x) + u * kGenericThreads < total) bufa[li[u]] = v0[u];
  }
  __syncthreads();
  float2* src = bufa;
  float2* dst = bufb;
  float out_scale = p.out_scale;
  if constexpr (M == Mode::kFwd) {
    generic_transform<false>(src, dst, tw, plan, nl);
  } else if constexpr (M == Mode::kInv) {
    generic_transform<true>(src, dst, tw, plan, nl);
  } else if constexpr (M == Mode::kModulus) {
    generic_transform<false>(src, dst, tw, plan, nl);
    const float io = p.in_scale * p.out_scale;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (static_cast<int>(threadIdx.x) + u * kGenericThreads < total) {
        src[li[u]] = project_scaled(src[li[u]], amp[u], p.in_scale, io, acc);
      }
    }
    const double ws = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ws;
    if (threadIdx.x == 0 && tile == 0 && start == 0 && p.counter) atomicAdd(p.counter, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int v = 0; v < kGenericThreads / 64; ++v) t += red[v];
      p.partials[static_cast<long long>(start) * p.partial_stride + tile] = t;
    }
    generic_transform<true>(src, dst, tw, plan, nl);
    out_scale = 1.0f;  // folded into the projection
  } else {  // kUpdate
    generic_transform<true>(src, dst, tw, plan, nl);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (static_cast<int>(threadIdx.x) + u * kGenericThreads < total) {
        const float2 v = src[li[u]];
        const float2 pm = {v.x * p.in_scale, v.y * p.in_scale};
        float2 nx = pm;
        if (!inside[u]) nx = {p.a * xo[u].x + p.b * pm.x, p.a * xo[u].y + p.b * pm.y};
        x[gi[u]] = nx;
        src[li[u]] = nx;
      }
    }
    __syncthreads();
    generic_transform<false>(src, dst, tw, plan, nl);
    if (tile == 0 && threadIdx.x < 64 && p.errs) {
      const double t = partial_sum<COH>(p.partials + static_cast<long long>(start) * p.partial_stride,
                                        p.partial_stride, threadIdx.x);
      if (threadIdx.x == 0) {
        const int e = eslot == kSlotFromCounter ? *p.counter - 1 : eslot;
        if (e >= 0 && e < p.err_cap) {
          p.errs[static_cast<long long>(e) * p.err_stride + p.err_offset + start] = t;
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (static_cast<int>(threadIdx.x) + u * kGenericThreads < total) {
      const float2 v = src[li[u]];
      w[gi[u]] = {v.x * out_scale, v.y * out_scale};
    }
  }
}

// The fused passes along an axis that is not a power of two: one generic_tile per
// work-group.
template <Mode M>
__global__ void __launch_bounds__(kGenericThreads)
    generic_pass(PassArgs p, GenericPlan plan, long long lines) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  float2* sm = reinterpret_cast<float2*>(sm_raw);
  const int lpb = generic_lines_per_block(plan.n);
  const int slot = generic_slot(plan);
  float2* tw = sm;
  float2* bufa = sm + slot;
  float2* bufb = bufa + lpb * slot;
  long long* lbase = reinterpret_cast<long long*>(bufb + lpb * slot);
  double* red = reinterpret_cast<double*>(lbase + lpb);
  // Block numbering: with p.slots > 0 (a 1D grid, slots a multiple of 8) block B runs tile
  // B / slots of start B % slots, so all of a start's blocks share one XCD's L2 (a strided
  // axis's blocks read adjacent pieces of the same 128-byte lines); otherwise the grid is
  // (tiles, starts).
  int tile, start;
  if (p.slots > 0) {
    start = blockIdx.x % p.slots;
    tile = blockIdx.x / p.slots;
    if (start >= p.starts) return;  // padding slot: the whole block leaves together
  } else {
    start = blockIdx.y;
    tile = blockIdx.x;
  }
  generic_tile<M, true>(p, plan, lines, start, tile, tw, bufa, bufb, lbase, red, kSlotFromCounter);
}

// The non-template kernels below have internal linkage: this header is included by more
// than one translation unit of the library, and each uses its own copy (or none).

// Element-wise modulus projection over whole starts (unfused iteration).
static __global__ void __launch_bounds__(256)
    modulus_elementwise(float2* w, const float* amp, long long n, float in_scale,
                        float out_scale, double* partials, int* counter) {
  __shared__ double red[16];
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  double acc = 0.0;
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float2 v = w[soff + k];
    const float2 q = project({v.x * in_scale, v.y * in_scale}, amp[k], acc);
    w[soff + k] = {q.x * out_scale, q.y * out_scale};
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) {
    partials[static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x] = t;
    if (blockIdx.x == 0 && blockIdx.y == 0) atomicAdd(counter, 1);
  }
}

// Element-wise update (unfused iteration), plus the fixed-order error reduction.
static __global__ void __launch_bounds__(256)
    update_elementwise(float2* x, const float2* pm, const std::uint8_t* support, long long n,
                       float a, float b, float in_scale, const double* partials,
                       int partial_stride, double* errs, int err_stride, int err_offset,
                       int err_cap, const int* counter) {
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float2 v = pm[soff + k];
    const float2 pv = {v.x * in_scale, v.y * in_scale};
    if (support[soff + k]) {
      x[soff + k] = pv;
    } else {
      const float2 xo = x[soff + k];
      x[soff + k] = {a * xo.x + b * pv.x, a * xo.y + b * pv.y};
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 64 && errs) {
    const double s = partial_sum(partials + static_cast<long long>(blockIdx.y) * partial_stride,
                                 partial_stride, threadIdx.x);
    if (threadIdx.x == 0) {
      const int slot = *counter - 1;
      if (slot >= 0 && slot < err_cap) {
        errs[static_cast<long long>(slot) * err_stride + err_offset + blockIdx.y] = s;
      }
    }
  }
}

// --- amplitude loss (FP64 and non-power-of-two path) ---------------------------------
// w holds the unnormalised spectrum of x; afterwards it holds the input of an unnormalised
// inverse transform that lands on P_M x, and partials[start][block] the FP64 sums of
// (|F x| - A)^2 over usable pixels (amp < 0 marks a masked pixel).
template <class Real>
__global__ void __launch_bounds__(256)
    loss_project(typename Cplx<Real>::T* w, const Real* amp, long long amp_stride, long long n, Real scale,
                 double* partials) {
  __shared__ double red[16];
  const long long soff = static_cast<long long>(blockIdx.y) * n;
  double acc = 0.0;
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const auto v = w[soff + k];
    const Real re = v.x * scale, im = v.y * scale;
    const Real mag = sqrt(re * re + im * im);
    const Real a = amp[blockIdx.y * amp_stride + k];
    Real target = mag;
    if (a >= Real(0)) {
      const Real d = mag - a;
      acc += static_cast<double>(d) * static_cast<double>(d);
      target = a;
    }
    const Real f = target / (mag + Real(1e-12)) * scale;
    w[soff + k] = {re * f, im * f};
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) partials[static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

// grad = 2 (x - P_M x), P_M x = in_scale * p
template <class Real>
__global__ void grad_from(const typename Cplx<Real>::T* x, const typename Cplx<Real>::T* p,
                          typename Cplx<Real>::T* grad, long long total, Real in_scale) {
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < total;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const auto xv = x[k];
    const auto pv = p[k];
    grad[k] = {Real(2) * (xv.x - in_scale * pv.x), Real(2) * (xv.y - in_scale * pv.y)};
  }
}

// loss[s] = sum of the `count` partials of start s, in order.
static __global__ void sum_partials(const double* partials, int count, double* loss, int starts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= starts) return;
  double t = 0.0;
  for (int k = 0; k < count; ++k) t += partials[static_cast<long long>(s) * count + k];
  loss[s] = t;
}

}  // namespace qpch::dev
