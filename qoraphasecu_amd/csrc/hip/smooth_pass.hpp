// Host interface of the smooth-length passes (smooth.hpp, instantiated in smooth.hip, its
// own translation unit so that it compiles beside the engine).
#pragma once

#include <hip/hip_runtime.h>

#include "passes.hpp"
#include "persist.hpp"

namespace qpch {

// Whether an axis of length N runs its fused passes (all but the gradient) in smooth_pass: N = R1 R2
// <= 2048 with R1 a power of two from 2 to 512 and R2 in {3, 5, 15} (R2 = 15 preferred).
bool smooth_axis(int N);

// Launches smooth_pass for a forward, inverse, modulus or update pass (not Mode::kGrad) of
// an axis of length N with `lines` lines per start, on the grid generic_pass would use
// (block numbering from a.slots).  col: the axis is strided (inner > 1).  Returns the
// launch status; hipErrorInvalidValue when no kernel exists.
hipError_t smooth_launch(int N, bool col, dev::Mode m, const dev::PassArgs& a, long long lines, dim3 grid,
                         hipStream_t s);

// The persistent launch for square images N x N of smooth axis length (twopass_smooth; the
// lengths R1 R2 with R1 up to 16, and 96 and 160):
// whether a kernel exists, and the launch of `len` iterations with `workers` work-groups
// (0: as many per XCD as the busiest XCD has tiles per pass, at most what fits the device
// at once).  q.tiles0 / q.tiles1: the tiles per start of the two passes.
bool twopass_smooth_supported(int ny, int nx);
hipError_t twopass_smooth_launch(int N, const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::SyncArgs& q,
                                 int len, int workers, hipStream_t s);

}  // namespace qpch
