// The persistent two-pass kernels (persist.hpp), instantiated per image shape and storage.
#include "twopass.hpp"

#include <algorithm>

namespace qpch {
namespace {

// 2D shapes with a persistent kernel: every power of two from 64 to 1024 on both axes,
// where the work-group's LDS fits (persist.hpp, TwoPassGeom::kValid).
#define QPCH_FOR_EACH_TWOPASS_N1(X, A) X(A, 64) X(A, 128) X(A, 256) X(A, 512) X(A, 1024)
#define QPCH_FOR_EACH_TWOPASS(X)                                                                  \
  QPCH_FOR_EACH_TWOPASS_N1(X, 64) QPCH_FOR_EACH_TWOPASS_N1(X, 128) QPCH_FOR_EACH_TWOPASS_N1(X, 256) \
  QPCH_FOR_EACH_TWOPASS_N1(X, 512) QPCH_FOR_EACH_TWOPASS_N1(X, 1024)

template <int N0, int N1, class St>
constexpr bool valid() {
  return dev::TwoPassGeom<N0, N1, St, 16>::kValid;
}

template <int N0, int N1, class St>
int row_tiles() {
  if constexpr (valid<N0, N1, St>()) {
    return (N0 + dev::TwoPassGeom<N0, N1, St, 16>::kB1 - 1) / dev::TwoPassGeom<N0, N1, St, 16>::kB1;
  } else {
    return 0;
  }
}

// Work-groups that fit the device at once (resident per CU x CUs), rounded down to a
// multiple of the XCD count; 0 when the query fails.
template <int N0, int N1, class St>
int capacity() {
  static const int cap = [] {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::twopass_iterations<N0, N1, St, 16>,
                                                     dev::TwoPassGeom<N0, N1, St, 16>::kThreads, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      return 0;
    }
    return std::max(1, per_cu) * cus;
  }();
  return cap;
}

template <int N0, int N1, class St>
hipError_t launch(const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::SyncArgs& q, int len, int workers,
                  hipStream_t s) {
  if constexpr (valid<N0, N1, St>()) {
    using TG = dev::TwoPassGeom<N0, N1, St, 16>;
    // work-groups are dealt to the XCDs round-robin: as many per XCD as the busiest XCD
    // has tiles per pass, at most what fits the device at once (a multiple of the XCDs)
    const int items = q.xcds * dev::xcd_starts(p1.starts, 0, q.xcds) * std::max(q.tiles0, q.tiles1);
    const int cap = capacity<N0, N1, St>() / q.xcds * q.xcds;
    if (cap <= 0) return hipErrorInvalidDeviceFunction;
    const int w = workers > 0 ? workers : std::max(q.xcds, std::min(items, cap));
    hipLaunchKernelGGL((dev::twopass_iterations<N0, N1, St, 16>), dim3(w), dim3(TG::kThreads), 0, s, p0, p1, q, len);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

}  // namespace

int twopass_xcds() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    return 0;
  }
  const int x = cus / 32;
  return cus % 32 == 0 && (x == 1 || x == 2 || x == 4 || x == 8) ? x : 0;
}

bool twopass_supported(int ny, int nx, int storage) {
#define QPCH_CASE(A, B)                                                               \
  if (ny == A && nx == B) {                                                           \
    return storage == 0   ? valid<A, B, float2>()                                     \
           : storage == 1 ? valid<A, B, __half2>()                                    \
                          : valid<A, B, __hip_bfloat162>();                           \
  }
  QPCH_FOR_EACH_TWOPASS(QPCH_CASE)
#undef QPCH_CASE
  return false;
}

int twopass_row_tiles(int ny, int nx, int storage) {
#define QPCH_CASE(A, B)                                                                     \
  if (ny == A && nx == B) {                                                                 \
    return storage == 0   ? row_tiles<A, B, float2>()                                       \
           : storage == 1 ? row_tiles<A, B, __half2>()                                      \
                          : row_tiles<A, B, __hip_bfloat162>();                             \
  }
  QPCH_FOR_EACH_TWOPASS(QPCH_CASE)
#undef QPCH_CASE
  return 0;
}

hipError_t twopass_launch(int ny, int nx, int storage, const dev::PassArgs& p0, const dev::PassArgs& p1,
                          const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
#define QPCH_CASE(A, B)                                                                          \
  if (ny == A && nx == B) {                                                                      \
    return storage == 0   ? launch<A, B, float2>(p0, p1, q, len, workers, s)                     \
           : storage == 1 ? launch<A, B, __half2>(p0, p1, q, len, workers, s)                    \
                          : launch<A, B, __hip_bfloat162>(p0, p1, q, len, workers, s);           \
  }
  QPCH_FOR_EACH_TWOPASS(QPCH_CASE)
#undef QPCH_CASE
  return hipErrorInvalidValue;
}

bool twopass_generic_supported(const dev::GenericPlan& plan0, const dev::GenericPlan& plan1) {
  return dev::twopass_generic_lds_bytes(plan0, plan1) + dev::kGenericStaticLds <= 160 * 1024;
}

hipError_t twopass_generic_launch(const dev::PassArgs& p0, const dev::PassArgs& p1, const dev::GenericPlan& plan0,
                                  const dev::GenericPlan& plan1, const dev::GenericPlan* plans_dev, long long lines0,
                                  long long lines1, const dev::SyncArgs& q, int len, int workers, hipStream_t s) {
  const size_t lds = dev::twopass_generic_lds_bytes(plan0, plan1);
  if (lds + dev::kGenericStaticLds > 160 * 1024) return hipErrorInvalidValue;
  int per_cu = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::twopass_generic<>, dev::kGenericThreads,
                                                              lds);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const int cap = std::max(1, per_cu) * cus / q.xcds * q.xcds;
  const int items = q.xcds * dev::xcd_starts(p1.starts, 0, q.xcds) * std::max(q.tiles0, q.tiles1);
  const int w = workers > 0 ? workers : std::max(q.xcds, std::min(items, cap));
  hipLaunchKernelGGL(dev::twopass_generic<>, dim3(w), dim3(dev::kGenericThreads), lds, s, p0, p1, plans_dev, lines0,
                     lines1, q, len);
  return hipGetLastError();
}

}  // namespace qpch
