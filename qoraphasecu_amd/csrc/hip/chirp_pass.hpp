// Host interface of the chirp-z passes (chirp.hpp, instantiated in chirp.hip, its own
// translation unit so that it compiles beside the engine).
#pragma once

#include <hip/hip_runtime.h>

#include "passes.hpp"

namespace qpch {

// The power of two M >= max(256, 2N - 1) of Bluestein's algorithm for a length-N transform.
int chirp_length(int N);

// Whether an axis of length N runs its fused passes (all but the gradient) in chirp_pass: 32 <= N <=
// 2048, not a power of two nor a smooth length, with a prime factor of 7 or more.
bool chirp_axis(int N);

// Lines per work-group of chirp_pass for an axis of length N (its tiles and error partials).
int chirp_lines(int N);

// Launches chirp_pass for a forward, inverse, modulus or update pass (not Mode::kGrad) of an
// axis of length N with `lines` lines per start on `grid` (tiles of chirp_lines(N) lines;
// block numbering from a.slots).  col: the axis is strided (inner > 1).  chirp / bhat: the
// length's Bluestein tables (fp32, N and chirp_length(N) entries).  Returns the launch
// status; hipErrorInvalidValue when no kernel exists.
hipError_t chirp_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                        const void* chirp, const void* bhat, hipStream_t s);

}  // namespace qpch
