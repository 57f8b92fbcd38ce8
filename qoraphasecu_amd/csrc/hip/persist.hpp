// Whole chunks of 2D iterations in one launch: the frequency-axis (column modulus) and
// real-axis (row update) passes of every iteration as work items of one persistent
// kernel, so that no launch boundary separates the passes.
//
// Each XCD runs its own starts (start s on XCD s mod X, X the device's XCD count: 8 on a
// whole MI355X) from its own queue: the work-group reads the XCD it runs on from the
// hardware (HW_REG_XCC_ID).  A queue's items are the
// tiles of the two passes of its starts, ordered (iteration, pass, start, tile) and claimed
// in that order from one counter.  A row tile of start s in iteration t needs every column
// tile of (t, s) done; a column tile of (t, s) needs every row tile of (t - 1, s) done.
// Each pass of each start counts its finished tiles.  A worker claims an item only after
// every earlier item of its queue has been claimed, by a worker that is already running,
// and an item waits only on earlier items of its queue: the earliest unfinished item can
// always proceed, whatever the number of workers and whatever else shares the device, so
// no wait can deadlock.  Each wait still has a time bound (a stuck wait sets a sticky
// abort flag, which the host reports, and every later wait gives up at once), and the last
// worker checks that every queue with items was emptied (an XCD without workers would
// leave its starts undone: also reported through the flag).
//
// Every tile runs line_tile, the arithmetic of the two-pass launches, on the same lines
// with the same partial layout, so results equal the two-pass schedule bit for bit.
// Producer and consumer of a start's data share one XCD, hence one L2: a finished tile's
// waves drain their stores to the L2 (s_waitcnt) before thread 0 counts the tile, and a
// worker reads data that other CUs wrote in this launch with device-scope loads, which do
// not hit in its CU's L1.
// Agent-scope fences (an L2 write-back and invalidate per tile) measured 10-50 us per tile
// round of 512 work-groups, against about 1 us for this protocol.
#pragma once

#include <hip/hip_runtime.h>

#include "generic.hpp"
#include "passes.hpp"

namespace qpch::dev {

// Device state of one launch group's persistent launches.  next, done and count are zero
// between launches (the last worker of a launch resets them); abort is sticky.
constexpr int kXcds = 8;       // XCDs of a whole MI355X (the most a gfx950 device has)
constexpr int kLineInts = 32;  // one 128-byte line
struct SyncArgs {
  int* next;    // [kXcds][kLineInts]: next item of each XCD's queue (one line each)
  int* done;    // workers that have finished
  int* count;   // [2][starts][kLineInts]: finished tiles of the column and the row pass, per
                // start, one line each (the waiters poll them)
  int* abort;   // set when a wait ran past its bound or a queue was left with items
  int tiles0;   // column-pass tiles per start
  int tiles1;   // row-pass tiles per start
  int xcds;     // XCDs of the device (1, 2, 4 or 8: a partitioned MI355X has fewer)
  long long wait_limit;  // bound of one wait in wall_clock64 ticks
};

// The XCD this wave runs on, numbered within the device (the hardware's XCC number modulo
// the device's XCD count: a partition's XCDs are consecutive and aligned).
__device__ __forceinline__ int xcd_id(int xcds) {
  return static_cast<int>(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & (xcds - 1));
}

// Starts of a launch on XCD x of `xcds`: s = x, x + xcds, ...
__host__ __device__ __forceinline__ int xcd_starts(int starts, int x, int xcds) {
  return starts > x ? (starts - x + xcds - 1) / xcds : 0;
}

__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0 of a worker: wait until *c >= target, or give up after `limit` ticks of the
// device's constant-rate clock (setting abort), or at once once abort is set.  A wait on a
// healthy launch lasts a few tile durations (tens of microseconds); the bound is time, not
// polls, because a poll's latency grows with the number of waiters.
__device__ __forceinline__ void wait_count(const int* c, int target, int* abort, long long limit) {
  const long long t0 = static_cast<long long>(wall_clock64());
  int polls = 0;
  while (load_agent(c) < target) {
    if (static_cast<long long>(wall_clock64()) - t0 > limit) {
      __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    if ((++polls & 63) == 0 && load_agent(abort) != 0) return;
    __builtin_amdgcn_s_sleep(2);
  }
}

// Work-group shape of the persistent launch: the column pass's (its tiles and error
// partials exactly those of the two-pass schedule), with row tiles of as many lines as
// fill the same work-group.
template <int N0, int N1, class St, int R>
struct TwoPassGeom {
  using G0 = PassGeom<N0, true, sizeof(St), R>;
  static constexpr int kThreads = G0::kThreads;
  static constexpr int kT1 = Shape1<N1, R>::kT;
  static constexpr int kB1 = kThreads / kT1 > 0 ? kThreads / kT1 : 1;
  using G1 = PassGeom<N1, false, sizeof(St), R, kB1>;
  static constexpr int kTile = G0::kTile > G1::kTile ? G0::kTile : G1::kTile;
  static constexpr int kLdsBytes = (kTile + N0 + N1) * 8 + 16;
  static constexpr bool kValid = G1::kThreads == kThreads && kLdsBytes <= 160 * 1024;
};

// The work-item loop of a persistent launch, shared by its kernels: `col(start, tile)` runs
// a column (modulus) tile, `row(start, tile, slot)` a row (update) tile that writes history
// slot `slot`; both must be called by every thread of the work-group, and the twiddle
// tables they use must be filled before the call (the loop's first barrier publishes them).
// p1: the update pass's argument block (starts, counter); `claim`: a work-group shared int.
// The history slot of iteration t is *p1.counter + t, and the counter advances by len
// when the launch ends.
template <class ColTile, class RowTile>
__device__ __forceinline__ void persistent_loop(const PassArgs& p1, const SyncArgs& q, int len, int& claim,
                                                ColTile&& col, RowTile&& row) {
  const int tid = threadIdx.x;
  const int S = p1.starts;
  const int xcd = __builtin_amdgcn_readfirstlane(xcd_id(q.xcds));
  const int mine = xcd_starts(S, xcd, q.xcds);
  const int per0 = mine * q.tiles0;
  const int per_iter = per0 + mine * q.tiles1;
  const int total = per_iter * len;
  int* next = q.next + xcd * kLineInts;
  int* count0 = q.count;                           // + start * kLineInts
  int* count1 = q.count + S * kLineInts;
  // every worker reads the counter before it takes its ticket, so the ticket's last holder
  // can advance it without another worker of this launch reading the new value
  const int slot0 = p1.counter ? *p1.counter : 0;
  // Thread 0 counts the previous item and claims the next one in ONE block at the top of
  // the loop, so that every path from one thread-0 block to the next crosses a barrier.
  // With the count at the bottom and the claim at the top, the compiler threaded the other
  // lanes from the one past the other, back to the claim's barrier: a second back edge that,
  // under the exec masks, sent every lane but thread 0 round the loop again on the old
  // claim, for ever.
  int* done_tile = nullptr;  // the count the previous item increments
  for (;;) {
    if (tid == 0) {
      if (done_tile) __hip_atomic_fetch_add(done_tile, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      claim = __hip_atomic_fetch_add(next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    // readfirstlane: one value for the whole work-group, so the exit and every branch on
    // the item are uniform for the compiler too
    const int item = __builtin_amdgcn_readfirstlane(claim);
    if (item >= total) break;
    const int it = item / per_iter;
    int r = item - it * per_iter;
    const bool is_row = r >= per0;
    int start, tile, target;
    const int* dep;
    if (!is_row) {
      const int j = r / q.tiles0;
      tile = r - j * q.tiles0;
      start = xcd + q.xcds * j;
      dep = count1 + start * kLineInts;
      target = it * q.tiles1;
      done_tile = count0 + start * kLineInts;
    } else {
      r -= per0;
      const int j = r / q.tiles1;
      tile = r - j * q.tiles1;
      start = xcd + q.xcds * j;
      dep = count0 + start * kLineInts;
      target = (it + 1) * q.tiles0;
      done_tile = count1 + start * kLineInts;
    }
    if (tid == 0 && target > 0) wait_count(dep, target, q.abort, q.wait_limit);
    __syncthreads();  // (also: every thread has read `claim` before it is claimed again)
    if (!is_row) {
      col(start, tile);
    } else {
      row(start, tile, slot0 + it);
    }
    // every wave's stores reach the L2 before thread 0 counts the tile at the top
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // the last worker to finish checks the queues and resets the launch state
  __syncthreads();
  if (tid == 0) {
    const int t = __hip_atomic_fetch_add(q.done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    claim = t == static_cast<int>(gridDim.x) - 1 ? 1 : 0;
  }
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(claim)) {
    for (int k = tid; k < 2 * S; k += blockDim.x) q.count[k * kLineInts] = 0;
    if (tid == 0) {
      for (int x = 0; x < q.xcds; ++x) {
        const int items = xcd_starts(S, x, q.xcds) * (q.tiles0 + q.tiles1) * len;
        if (items > 0 && load_agent(q.next + x * kLineInts) < items) {
          __hip_atomic_store(q.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        q.next[x * kLineInts] = 0;
      }
      *q.done = 0;
      if (p1.counter) *p1.counter = slot0 + len;
    }
  }
}

// `len` iterations of every start of a launch group on an N0 x N1 image (axis 0 = the
// column pass, axis 1 = the row pass), power-of-two axes.  p0 / p1: the argument blocks of
// the two passes as the two-pass schedule builds them (p0.counter unused).
template <int N0, int N1, class St, int R>
__global__ void __launch_bounds__((TwoPassGeom<N0, N1, St, R>::kThreads))
    twopass_iterations(PassArgs p0, PassArgs p1, SyncArgs q, int len) {
  using TG = TwoPassGeom<N0, N1, St, R>;
  static_assert(TG::kValid, "both passes must use work-groups of one size");
  __shared__ float2 lds[TG::kTile + N0 + N1];
  __shared__ int claim;
  float2* tw0 = lds + TG::kTile;
  float2* tw1 = tw0 + N0;
  load_twiddles<N0>(tw0);
  load_twiddles<N1>(tw1);
  persistent_loop(
      p1, q, len, claim,
      [&](int start, int tile) { line_tile<N0, true, Mode::kModulus, St, R, false, true>(p0, start, tile, lds, tw0, 0); },
      [&](int start, int tile, int slot) {
        line_tile<N1, false, Mode::kUpdate, St, R, false, true, TG::kB1>(p1, start, tile, lds, tw1, slot);
      });
}

// Static LDS of twopass_generic: the claim word, padded to the 16-byte alignment of the
// dynamic LDS that follows it.
constexpr size_t kGenericStaticLds = 16;

// Dynamic LDS of twopass_generic: both axes' twiddle tables, two line buffers as large as
// either axis's tile needs, the line bases and one FP64 slot per wave.
__host__ __device__ constexpr size_t twopass_generic_lds_bytes(const GenericPlan& a, const GenericPlan& b) {
  const int la = generic_lines_per_block(a.n), lb = generic_lines_per_block(b.n);
  const int buf = la * generic_slot(a) > lb * generic_slot(b) ? la * generic_slot(a) : lb * generic_slot(b);
  return sizeof(float2) * (static_cast<size_t>(generic_slot(a)) + generic_slot(b) + 2 * static_cast<size_t>(buf)) +
         sizeof(long long) * (la > lb ? la : lb) + sizeof(double) * (kGenericThreads / 64);
}

// The same launch for images whose two axes are both not powers of two (fp32): generic_tile
// (generic.hpp, the per-pass kernels' body) for the tiles of both passes.  (A template, so
// that each translation unit including this header may instantiate it.)
// The two plans come through device memory (plans[0]: axis 0, plans[1]: axis 1): passed by
// value, the two structures' fields kept 266 scalar registers spilled.
template <int kThreads = kGenericThreads>
__global__ void __launch_bounds__(kThreads)
    twopass_generic(PassArgs p0, PassArgs p1, const GenericPlan* __restrict__ plans, long long lines0, long long lines1,
                    SyncArgs q, int len) {
  // 16-byte aligned: without it the dynamic LDS began 4 bytes in, after `claim`, and every
  // float2 and double access in it was misaligned (the tiles ran 2-3 times slower)
  extern __shared__ __align__(16) unsigned char sm_raw[];
  __shared__ int claim;
  const GenericPlan& plan0 = plans[0];
  const GenericPlan& plan1 = plans[1];
  const int s0 = generic_slot(plan0), s1 = generic_slot(plan1);
  const int l0 = generic_lines_per_block(plan0.n), l1 = generic_lines_per_block(plan1.n);
  const int buf = l0 * s0 > l1 * s1 ? l0 * s0 : l1 * s1;
  float2* tw0 = reinterpret_cast<float2*>(sm_raw);
  float2* tw1 = tw0 + s0;
  float2* bufa = tw1 + s1;
  float2* bufb = bufa + buf;
  long long* lbase = reinterpret_cast<long long*>(bufb + buf);
  double* red = reinterpret_cast<double*>(lbase + (l0 > l1 ? l0 : l1));
  for (int j = threadIdx.x; j < s0 + s1; j += kThreads) {
    const int m = j < s0 ? s0 : s1;
    const int k = j < s0 ? j : j - s0;
    double sn, cs;
    sincospi(2.0 * static_cast<double>(k) / static_cast<double>(m), &sn, &cs);
    tw0[j] = {static_cast<float>(cs), static_cast<float>(-sn)};
  }
  persistent_loop(
      p1, q, len, claim,
      [&](int start, int tile) {
        generic_tile<Mode::kModulus, false, true>(p0, plan0, lines0, start, tile, tw0, bufa, bufb, lbase, red, 0);
      },
      [&](int start, int tile, int slot) {
        generic_tile<Mode::kUpdate, false, true>(p1, plan1, lines1, start, tile, tw1, bufa, bufb, lbase, red, slot);
      });
}

}  // namespace qpch::dev
