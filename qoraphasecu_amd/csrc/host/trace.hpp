// Named host ranges for rocprofv3 --marker-trace (roctx).  The reference brackets a warmed
// run with the NVTX range "iterations" (bench/profile.cu:27-34); this backend marks every
// run that way, and the autotuner's timing trials as "qpch autotune".  Without a profiler
// attached a push/pop is a call into an idle library.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace qpcu::detail {

class TraceRange {
 public:
  explicit TraceRange(const char* name) { roctxRangePush(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace qpcu::detail
