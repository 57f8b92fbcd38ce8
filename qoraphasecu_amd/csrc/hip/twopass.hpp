// Host interface of the persistent two-pass launches (twopass.hip, its own translation
// unit so that it compiles beside the engine).
#pragma once

#include <hip/hip_runtime.h>

#include "passes.hpp"
#include "persist.hpp"

namespace qpch {

// storage: 0 complex64, 1 half2, 2 bfloat162 (the iterate's storage, as QPCH_F32/F16/BF16)
bool twopass_supported(int ny, int nx, int storage);

// Row-pass tiles per start of the persistent kernel for this shape (its row tiles hold as
// many lines as fill the column pass's work-group); 0 when there is no kernel.
int twopass_row_tiles(int ny, int nx, int storage);

// XCDs of the current device for the persistent launches: its compute units in groups of
// 32 (an MI355X XCD), when that gives 1, 2, 4 or 8; 0 when the device does not divide so,
// and the persistent launches are not used there.
int twopass_xcds();

// Launches twopass_iterations for `len` iterations on stream s with `workers` work-groups
// (0: as many per XCD as the busiest XCD has tiles per pass, at most what fits the device
// at once).  Returns the launch status; hipErrorInvalidValue when no kernel exists.
hipError_t twopass_launch(int ny, int nx, int storage, const dev::PassArgs& p0, const dev::PassArgs& p1,
                          const dev::SyncArgs& q, int len, int workers, hipStream_t s);

// The persistent launch for 2D images whose two axes are both not powers of two (fp32):
// whether its LDS (both plans' tables and tiles) fits a work-group, and the launch.
bool twopass_generic_supported(const dev::GenericPlan& plan0, const dev::GenericPlan& plan1);
// plans_dev: the two plans in device memory (axis 0, then axis 1); plan0 / plan1: the same
// plans on the host (for the LDS size).
hipError_t twopass_generic_launch(const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::GenericPlan& plan0,
                                  const dev::GenericPlan& plan1, const dev::GenericPlan* plans_dev, long long lines0,
                                  long long lines1, const dev::SyncArgs& q, int len, int workers, hipStream_t s);

}  // namespace qpch
