// Host interface of the smooth-length passes (smooth.hpp, instantiated in smooth.hip, its
// own translation unit so that it compiles beside the engine).
#pragma once

#include <hip/hip_runtime.h>

#include "passes.hpp"
#include "persist.hpp"

namespace qpch {

// Whether an axis of length N runs its fused passes (all but the gradient) in smooth_pass:
// N = R1 R2 <= 2048 with R1 a power of two and R2 its odd part: R2 in {3, 5, 15} with R1
// from 2 to 512, 7 with R1 from 2 to 256, 9, 11 and 13 with R1 from 2 to 128, 17 and 19
// with R1 from 2 to 16, 21, 23, 25, 27, 33, 35, 39 and 49 with R1 8 or 16, 45 and 63 with
// R1 16, and the mixed pairs (smooth_mixed.hip).
bool smooth_axis(int N);

// Lines per work-group of smooth_pass on an axis of length N (smooth_axis(N)): generic_pass's
// count for R2 in {3, 5, 15}, else as many lines as fill 256 threads.
int smooth_lines(int N);

// Launches smooth_pass for a forward, inverse, modulus or update pass (not Mode::kGrad) of
// an axis of length N with `lines` lines per start, on the grid generic_pass would use
// (block numbering from a.slots).  col: the axis is strided (inner > 1).  Returns the
// launch status; hipErrorInvalidValue when no kernel exists.
hipError_t smooth_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                         hipStream_t s);

// smooth_launch for R2 in {9, 25, 27, 45} (smooth_sq.hip): the pair (r1, r2) as
// smooth_axis factors it.
hipError_t smooth_sq_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                            hipStream_t s);

// smooth_launch for the pairs whose R1 is not a power of two or whose R2 is even
// (smooth_mixed.hip).
hipError_t smooth_mixed_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                               dim3 grid, hipStream_t s);

// smooth_mixed_launch for the mixed pairs of lengths with a prime from 7 to 13
// (smooth_mixed7.hip).
hipError_t smooth_mixed7_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                dim3 grid, hipStream_t s);

// smooth_mixed13_launch for the pairs of smooth_mixed1k.hip (525 to 1014).
hipError_t smooth_mixed1k_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                 dim3 grid, hipStream_t s);

// smooth_mixed7_launch for the further mixed pairs of smooth_mixed13.hip.
hipError_t smooth_mixed13_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                                 dim3 grid, hipStream_t s);

// smooth_launch for R2 in {7, 11, 13, 17, 19, 21, 23, 33, 35, 39, 49, 63} (smooth_prime.hip).
hipError_t smooth_prime_launch(int r1, int r2, bool col, dev::Mode m, const dev::PassArgs& a, long long lines,
                               dim3 grid, hipStream_t s);

// twopass_smooth_launch for the mixed pairs of smooth_mixed7.hip.
hipError_t twopass_smooth_mixed7_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                        const dev::SyncArgs& q, int len, int workers, hipStream_t s);

// twopass_smooth_launch for the mixed pairs (smooth_mixed.hip, then smooth_mixed7.hip).
hipError_t twopass_smooth_mixed_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                       const dev::SyncArgs& q, int len, int workers, hipStream_t s);

// twopass_smooth_launch for R2 in {7, 11, 13} (smooth_prime.hip).
hipError_t twopass_smooth_prime_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                       const dev::SyncArgs& q, int len, int workers, hipStream_t s);

// twopass_smooth_launch for R2 in {9, 25, 27, 45} (smooth_sq.hip).
hipError_t twopass_smooth_sq_launch(int r1, int r2, const dev::PassArgs& p0, const dev::PassArgs& p1,
                                    const dev::SyncArgs& q, int len, int workers, hipStream_t s);

// The persistent launch for square images N x N of smooth axis length (twopass_smooth; the
// lengths R1 R2 with R2 in {3, 5, 15} and R1 up to 16, 96, 160, and 18, 36, 72, 14, 28,
// 56, 22, 44, 26, 52, 45, 50, 54, 100, 42 and 70):
// whether a kernel exists, and the launch of `len` iterations with `workers` work-groups
// (0: as many per XCD as the busiest XCD has tiles per pass, at most what fits the device
// at once).  q.tiles0 / q.tiles1: the tiles per start of the two passes.
bool twopass_smooth_supported(int ny, int nx);
hipError_t twopass_smooth_launch(int N, const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::SyncArgs& q,
                                 int len, int workers, hipStream_t s);

}  // namespace qpch
