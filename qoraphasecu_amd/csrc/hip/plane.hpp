// The real-space pass of a 3D iteration on whole xy-planes: one work-group holds a
// NY x NX plane (up to 128 x 128, 141 KB of LDS), so the y and x inverse transforms, the
// update and the y and x forward transforms all happen between one read and one write of
// the plane.  With the z-line modulus pass this makes a 3D iteration two HBM passes
// (53 B per element in fp32, 29 B in fp16), like 2D.
//
// Thread layouts (radix 16, each thread holds 16 elements of one line):
//   row layout:    thread t -> row y = t / (NX/16), butterfly i = t % (NX/16), x = i + m NX/16
//   column layout: thread t -> column x = t % NX, butterfly j = t / NX, y = j + m NY/16
// The plane moves between the two through an LDS transpose.  The wave is read and written
// in the row layout (runs of NX/16 consecutive elements per row); the iterate and the
// support in the column layout (each wave of 64 lanes covers 64 consecutive x).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_engine.hpp"
#include "passes.hpp"
#include "storage.hpp"

namespace qpch::dev {

template <int NY, int NX>
struct PlaneGeom {
  using SX = Shape1<NX, 16>;
  using SY = Shape1<NY, 16>;
  static_assert(SX::kR == 16 && SY::kR == 16, "plane passes need both axes >= 16");
  static constexpr int kThreads = NY * NX / 16;
  static constexpr int kTX = NX / 16;
  static constexpr int kTY = NY / 16;
  // y-lines are handled by consecutive lanes: an odd stride spreads them over the banks
  static constexpr int kColStride = SY::kLine | 1;
  // transpose rows: 8 lanes write a row, 8 rows a wave; a row stride of 8 float2 more than
  // a multiple of 32 puts four rows side by side on the 64 banks
  static constexpr int kTStride = NX >= 32 ? NX + 8 : NX + 1;
  static constexpr int kRowLines = NY * SX::kLine;  // x-line scratch
  static constexpr int kColLines = NX * kColStride;  // y-line scratch
  static constexpr int kTranspose = NY * kTStride;  // padded plane
  static constexpr int kMax1 = kRowLines > kColLines ? kRowLines : kColLines;
  static constexpr int kScratch = kMax1 > kTranspose ? kMax1 : kTranspose;
  static constexpr int kLds = kScratch + NX + NY;  // + both twiddle tables (float2)
};

// x-layout registers -> column layout through the padded LDS plane
template <int NY, int NX>
__device__ __forceinline__ void rows_to_cols(float2* v, int y, int i, int x, int j, float2* t) {
  using G = PlaneGeom<NY, NX>;
#pragma unroll
  for (int m = 0; m < 16; ++m) t[y * G::kTStride + i + m * G::kTX] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = t[(j + m * G::kTY) * G::kTStride + x];
  __syncthreads();
}

template <int NY, int NX>
__device__ __forceinline__ void cols_to_rows(float2* v, int y, int i, int x, int j, float2* t) {
  using G = PlaneGeom<NY, NX>;
#pragma unroll
  for (int m = 0; m < 16; ++m) t[(j + m * G::kTY) * G::kTStride + x] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = t[y * G::kTStride + i + m * G::kTX];
  __syncthreads();
}

// The plane update pass: the same argument block as the line passes (outer = planes per
// start, n = elements per start); block numbering as in line_pass, one tile per plane.
template <int NY, int NX, class St>
__global__ void __launch_bounds__((PlaneGeom<NY, NX>::kThreads))
    plane_update(PassArgs p) {
  using G = PlaneGeom<NY, NX>;
  __shared__ float2 lds[G::kLds];
  float2* twx = lds + G::kScratch;
  float2* twy = twx + NX;

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
  const int t = threadIdx.x;
  const int y = t / G::kTX, i = t % G::kTX;  // row layout
  const int x = t % NX, j = t / NX;          // column layout
  const long long plane = static_cast<long long>(start) * p.n + static_cast<long long>(tile) * NY * NX;
  // The block's plane as buffer resources: a 32-bit lane offset plus compile-time
  // per-register offsets (the column layout steps by kTY rows, beyond the reach of a global
  // load's immediate offset), so no access needs 64-bit address arithmetic.
  constexpr int kN = NY * NX;
  const View<St, true> wrow(static_cast<St*>(p.w) + plane, kN, y * NX + i, G::kTX);
  const View<St, true> xcol(static_cast<St*>(p.x) + plane, kN, j * NX + x, G::kTY * NX);
  const View<const std::uint8_t, true> scol(p.support + plane, kN, j * NX + x, G::kTY * NX);

  float2 v[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = vld(wrow, m);
  load_twiddles<NX>(twx);
  load_twiddles<NY>(twy);
  __syncthreads();

  // P_M x: x inverse, y inverse
  fft_line<NX, 16, false>(v, i, lds + y * G::SX::kLine, twx);
  rows_to_cols<NY, NX>(v, y, i, x, j, lds);
  fft_line<NY, 16, false>(v, j, lds + x * G::kColStride, twy);
  // (the old iterate and the support are read here rather than before the transforms: at
  // 128 x 128 the 1024-thread block has 128 registers per thread, holding them across the
  // transforms spills with fp32 storage, and with fp16 storage it measured no faster.  They
  // are read eight elements at a time, unconditionally, so that sixteen loads are in flight
  // together; sixteen at a time spills at 128 x 128.)
  // ER (a = b = 0) skips the old iterate, as the line passes do: outside the support the
  // update is zero whatever it holds.
  constexpr int kG = 8;
  const bool need_x = p.a != 0.0f || p.b != 0.0f;
#pragma unroll
  for (int h = 0; h < 16; h += kG) {
    float2 xo[kG];
    bool inside[kG];
#pragma unroll
    for (int q = 0; q < kG; ++q) inside[q] = vld_flag(scol, h + q);
    if (need_x) {
#pragma unroll
      for (int q = 0; q < kG; ++q) xo[q] = vld(xcol, h + q);
    } else {
#pragma unroll
      for (int q = 0; q < kG; ++q) xo[q] = float2{0.0f, 0.0f};
    }
#pragma unroll
    for (int q = 0; q < kG; ++q) {
      const int m = h + q;
      const float2 pm = {v[m].x * p.in_scale, v[m].y * p.in_scale};
      float2 nx = pm;
      if (!inside[q]) nx = {p.a * xo[q].x + p.b * pm.x, p.a * xo[q].y + p.b * pm.y};
      vst(xcol, m, nx);
      v[m] = nx;
    }
  }
  // the next iteration's partial spectrum: y forward, x forward
  fft_line<NY, 16, true>(v, j, lds + x * G::kColStride, twy);
  cols_to_rows<NY, NX>(v, y, i, x, j, lds);
  fft_line<NX, 16, true>(v, i, lds + y * G::SX::kLine, twx);
#pragma unroll
  for (int m = 0; m < 16; ++m) vst(wrow, m, float2{v[m].x * p.out_scale, v[m].y * p.out_scale});

  if (tile == 0 && t < 64 && p.errs) {
    const double s = partial_sum(p.partials + static_cast<long long>(start) * p.partial_stride,
                                 p.partial_stride, t);
    if (t == 0) {
      const int slot = *p.counter - 1;
      if (slot >= 0 && slot < p.err_cap) {
        p.errs[static_cast<long long>(slot) * p.err_stride + p.err_offset + start] = s;
      }
    }
  }
}


// Whole iterations of a 2D image that one work-group holds (up to 8192 elements: the
// iterate, support and amplitude of a start stay in registers, 16 elements each per
// thread, which 128 x 128 would spill).  Per iteration: x and y forward transforms,
// projection with the error numerator summed in a fixed order (waves' shuffle trees, then
// the waves in index order), y and x inverse transforms, update; the transposes and line
// exchanges are those of plane_update.  The iterate is read and written once per chunk of
// `len` iterations and no pass boundary separates them; the amplitude is read once.
// fp32 storage; in_scale = out_scale = 1/sqrt(n), as in the modulus pass, so the inverse
// lands on P_M x.  Iteration it writes error slot *counter + it, and the last work-group
// of the launch to finish advances the counter by len.  One work-group per start of the
// launch (blockIdx.x).
template <int NY, int NX>
__global__ void __launch_bounds__((PlaneGeom<NY, NX>::kThreads))
    image_iterations(PassArgs p, int len) {
  using G = PlaneGeom<NY, NX>;
  static_assert(NY * NX <= 8192, "the image must fit the registers of one work-group");
  __shared__ float2 lds[G::kLds];
  __shared__ double red[G::kThreads / 64];
  float2* twx = lds + G::kScratch;
  float2* twy = twx + NX;
  const int start = blockIdx.x;
  const int t = threadIdx.x;
  const int y = t / G::kTX, i = t % G::kTX;  // row layout
  const int x = t % NX, j = t / NX;          // column layout
  constexpr int kN = NY * NX;
  const long long soff = static_cast<long long>(start) * p.n;
  const View<float2, true> xrow(static_cast<float2*>(p.x) + soff, kN, y * NX + i, G::kTX);
  const View<const std::uint8_t, true> srow(p.support + soff, kN, y * NX + i, G::kTX);
  const long long arow = p.amp_stride * ((p.err_offset + start) / (p.amp_group > 1 ? p.amp_group : 1));
  const View<const float, true> acol(p.amp + arow, kN, j * NX + x, G::kTY * NX);
  float2 xr[16];
  bool inside[16];
  float amp[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    xr[m] = vld(xrow, m);
    inside[m] = vld_flag(srow, m);
    amp[m] = vld_f32(acol, m);
  }
  load_twiddles<NX>(twx);
  load_twiddles<NY>(twy);
  const int slot0 = p.counter ? *p.counter : 0;
  __syncthreads();
  const float io = p.in_scale * p.out_scale;
  for (int it = 0; it < len; ++it) {
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = xr[m];
    fft_line<NX, 16, true>(v, i, lds + y * G::SX::kLine, twx);
    rows_to_cols<NY, NX>(v, y, i, x, j, lds);
    fft_line<NY, 16, true>(v, j, lds + x * G::kColStride, twy);
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = project_scaled(v[m], amp[m], p.in_scale, io, acc);
    const double ws = wave_sum(acc);
    if ((t & 63) == 0) red[t >> 6] = ws;
    __syncthreads();
    if (t == 0 && p.errs) {
      double e = 0.0;
      for (int w = 0; w < G::kThreads / 64; ++w) e += red[w];
      const int slot = slot0 + it;
      if (slot >= 0 && slot < p.err_cap) {
        p.errs[static_cast<long long>(slot) * p.err_stride + p.err_offset + start] = e;
      }
    }
    fft_line<NY, 16, false>(v, j, lds + x * G::kColStride, twy);
    cols_to_rows<NY, NX>(v, y, i, x, j, lds);  // (its barriers also order red's next writes)
    fft_line<NX, 16, false>(v, i, lds + y * G::SX::kLine, twx);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      xr[m] = inside[m] ? v[m] : float2{p.a * xr[m].x + p.b * v[m].x, p.a * xr[m].y + p.b * v[m].y};
    }
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) vst(xrow, m, xr[m]);
  // Every work-group read the counter at its start and used the slot (thread 0 wrote the
  // errors with it) before thread 0 takes a ticket here, so the last ticket's holder can
  // advance the counter without any work-group of this launch reading the new value.
  if (p.counter && t == 0) {
    if (atomicAdd(p.done, 1) == static_cast<int>(gridDim.x) - 1) {
      *p.done = 0;
      *p.counter = slot0 + len;
    }
  }
}

}  // namespace qpch::dev
