// The MI355X reconstruction engine behind the qpch C ABI: device buffers, the per-shape
// pass schedule, hipGraph replay of iteration chunks, Shrinkwrap and the operators.
//
// Data layout in HBM (per solver): x and the wave are [starts][n] in FFT order, row-major
// (complex64, or half2 / bfloat162 storage); the support is [starts][n] uint8; the signed
// amplitude (-1 = masked) and the true amplitude are [n] float shared by every start;
// the error history is [slot][starts] FP64 numerators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "qpch.h"
#include "generic.hpp"
#include "misc.hpp"
#include "passes.hpp"
#include "plane.hpp"
#include "chirp_pass.hpp"
#include "smooth_pass.hpp"
#include "twopass.hpp"
#include "blur.hpp"

namespace qpch {

using dev::Mode;
using dev::PassArgs;

namespace {

thread_local std::string g_error;

struct InvalidArgument : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define QPCH_HIP(expr)                                                                       \
  do {                                                                                     \
    const hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                                \
      throw ::qpch::DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" +       \
                        __FILE__ + ":" + std::to_string(__LINE__) + ")");                 \
    }                                                                                      \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return QPCH_OK;
  } catch (const InvalidArgument& e) {
    g_error = e.what();
    return QPCH_EINVAL;
  } catch (const std::exception& e) {
    g_error = e.what();
    return QPCH_EDEVICE;
  } catch (...) {
    g_error = "unknown error";
    return QPCH_EDEVICE;
  }
}

template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t count) { resize(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~DevBuf() { if (p_) (void)hipFree(p_); }
  void resize(size_t count) {
    if (p_) { (void)hipFree(p_); p_ = nullptr; }
    n_ = count;
    if (count) QPCH_HIP(hipMalloc(&p_, count * sizeof(T)));
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Host <-> device copies go through a stream of the caller's own (non-blocking) and wait
// for it: never through the legacy stream, which a graph capture running in another
// thread forbids (and which would not order the copy after that stream's queued work).
void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (!bytes) return;
  QPCH_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  QPCH_HIP(hipStreamSynchronize(s));
}

template <class T>
DevBuf<T> upload(const T* host, size_t count, hipStream_t s) {
  DevBuf<T> d(count);
  copy_sync(d.get(), host, count * sizeof(T), hipMemcpyHostToDevice, s);
  return d;
}

// A non-blocking stream for one synchronous entry point, drained and destroyed on exit.
// Declared before the entry point's device buffers, so they are freed first.
struct OwnStream {
  hipStream_t s{};
  OwnStream() { QPCH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~OwnStream() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
  OwnStream(const OwnStream&) = delete;
  OwnStream& operator=(const OwnStream&) = delete;
};

bool is_pow2(long long v) { return v >= 1 && (v & (v - 1)) == 0; }

// ---------------------------------------------------------------------------------------
// geometry

struct Dims {
  int nd = 0;
  int d[3] = {1, 1, 1};
  long long n = 0;
  // axis a: lines of length d[a], `inner` apart, `outer` blocks of them
  long long inner(int a) const {
    long long v = 1;
    for (int b = a + 1; b < nd; ++b) v *= d[b];
    return v;
  }
  long long outer(int a) const {
    long long v = 1;
    for (int b = 0; b < a; ++b) v *= d[b];
    return v;
  }
};

Dims make_dims(int ndim, const size_t* shape) {
  if (ndim != 2 && ndim != 3) throw InvalidArgument("shape must be 2D or 3D");
  Dims d;
  d.nd = ndim;
  d.n = 1;
  for (int a = 0; a < ndim; ++a) {
    if (shape[a] == 0) throw InvalidArgument("shape has a zero-length axis");
    if (shape[a] > (1u << 20)) throw InvalidArgument("axis too long");
    d.d[a] = static_cast<int>(shape[a]);
    d.n *= static_cast<long long>(shape[a]);
  }
  return d;
}

constexpr int kMaxFused = 4096;

dev::Grid3 grid3(const Dims& d) {
  return d.nd == 2 ? dev::Grid3{1, d.d[0], d.d[1]} : dev::Grid3{d.d[0], d.d[1], d.d[2]};
}

bool eager_launches() {
  const char* v = std::getenv("QPCH_EAGER");
  return v && v[0] == '1';
}

// The fused passes address one start's arrays with 32-bit byte offsets (buffer
// resources): a start of complex64 must stay below 2 GiB.
bool fusable(const Dims& d) {
  for (int a = 0; a < d.nd; ++a) {
    if (!is_pow2(d.d[a]) || d.d[a] < 2 || d.d[a] > kMaxFused) return false;
  }
  return d.n * static_cast<long long>(sizeof(float2)) < (1LL << 31);
}

// The fused passes also cover fp32 shapes whose other axes are any length up to
// kGenericMaxN: those axes run the generic pass kernel (generic.hpp).
bool generic_fusable(const Dims& d) {
  for (int a = 0; a < d.nd; ++a) {
    const int N = d.d[a];
    const bool pow2_ok = is_pow2(N) && N >= 2 && N <= kMaxFused;
    if (!pow2_ok && (N < 2 || N > dev::kGenericMaxN)) return false;
  }
  return d.n * static_cast<long long>(sizeof(float2)) < (1LL << 31);
}

bool generic_axis(int N) { return !is_pow2(N) || N > kMaxFused; }

int grid_for(long long work, int threads = 256, int cap = 2048) {
  long long g = (work + threads - 1) / threads;
  return static_cast<int>(std::max<long long>(1, std::min<long long>(g, cap)));
}

// ---------------------------------------------------------------------------------------
// fused line-pass dispatch

// Radix-8 kernels exist for these lengths (shorter lines already fit one or two radix-16
// stages; longer ones would need a fourth LDS exchange).
constexpr bool has_radix8(int N) { return N >= 64 && N <= 2048; }
int radix_for(int N, int radix) { return radix == 8 && has_radix8(N) ? 8 : 16; }

// Blocks per start of a pass.
template <int N, bool COL, int EB, int R = 16>
int pass_tiles_t(int outer, int inner) {
  using G = dev::PassGeom<N, COL, EB, R>;
  return COL ? outer * ((inner + G::kB - 1) / G::kB) : (outer + G::kB - 1) / G::kB;
}

// Starts of a launch go XCD-affine (a start's blocks all on one XCD) when they split
// evenly over the eight XCDs; a single start keeps its blocks spread over the chip.
int xcd_slots(int starts) { return (starts >= 8 && starts % 8 == 0) ? starts : 0; }

template <int N, bool COL, Mode M, class St, int R>
void launch_pass_r(PassArgs a, int starts, hipStream_t s) {
  using G = dev::PassGeom<N, COL, sizeof(St), R>;
  const int tiles = pass_tiles_t<N, COL, sizeof(St), R>(a.outer, a.inner);
  a.starts = starts;
  a.slots = a.no_xcd ? 0 : xcd_slots(starts);
  const int gx = tiles * (a.slots > 0 ? a.slots : starts);
  hipLaunchKernelGGL((dev::line_pass<N, COL, M, St, R>), dim3(gx), dim3(G::kThreads), 0, s, a);
  QPCH_HIP(hipGetLastError());
}

template <int N, bool COL, Mode M, class St>
void launch_pass_t(PassArgs a, int starts, hipStream_t s) {
  if constexpr (has_radix8(N) && M != Mode::kGrad) {
    if (a.radix == 8) return launch_pass_r<N, COL, M, St, 8>(a, starts, s);
  }
  launch_pass_r<N, COL, M, St, 16>(a, starts, s);
}

template <int N, class St>
void launch_pass_n(bool col, Mode m, const PassArgs& a, int starts, hipStream_t s) {
  if (col) {
    switch (m) {
      case Mode::kFwd: return launch_pass_t<N, true, Mode::kFwd, St>(a, starts, s);
      case Mode::kInv: return launch_pass_t<N, true, Mode::kInv, St>(a, starts, s);
      case Mode::kModulus: return launch_pass_t<N, true, Mode::kModulus, St>(a, starts, s);
      case Mode::kUpdate: return launch_pass_t<N, true, Mode::kUpdate, St>(a, starts, s);
      case Mode::kGrad: return launch_pass_t<N, true, Mode::kGrad, St>(a, starts, s);
    }
  } else {
    switch (m) {
      case Mode::kFwd: return launch_pass_t<N, false, Mode::kFwd, St>(a, starts, s);
      case Mode::kInv: return launch_pass_t<N, false, Mode::kInv, St>(a, starts, s);
      case Mode::kModulus: return launch_pass_t<N, false, Mode::kModulus, St>(a, starts, s);
      case Mode::kUpdate: return launch_pass_t<N, false, Mode::kUpdate, St>(a, starts, s);
      case Mode::kGrad: return launch_pass_t<N, false, Mode::kGrad, St>(a, starts, s);
    }
  }
}

#define QPCH_FOR_EACH_N(X) \
  X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

template <class St>
void launch_pass(int N, bool col, Mode m, const PassArgs& a, int starts, hipStream_t s) {
  switch (N) {
#define QPCH_CASE(V) \
  case V:            \
    return launch_pass_n<V, St>(col, m, a, starts, s);
    QPCH_FOR_EACH_N(QPCH_CASE)
#undef QPCH_CASE
    default:
      throw InvalidArgument("no fused pass for axis length " + std::to_string(N));
  }
}

// Error partials per start of a modulus pass: one per wave of every work-group.
template <int N, bool COL, int EB, int R = 16>
int pass_partials_t(int outer, int inner) {
  return pass_tiles_t<N, COL, EB, R>(outer, inner) * dev::PassGeom<N, COL, EB, R>::kWaves;
}

template <int EB>
int pass_partials(int N, bool col, int outer, int inner, int radix = 16) {
  const bool r8 = radix_for(N, radix) == 8;
  switch (N) {
#define QPCH_CASE(V)                                                                                 \
  case V:                                                                                            \
    if (r8) return col ? pass_partials_t<V, true, EB, 8>(outer, inner) : pass_partials_t<V, false, EB, 8>(outer, inner); \
    return col ? pass_partials_t<V, true, EB>(outer, inner) : pass_partials_t<V, false, EB>(outer, inner);
    QPCH_FOR_EACH_N(QPCH_CASE)
#undef QPCH_CASE
    default:
      throw InvalidArgument("no fused pass for axis length " + std::to_string(N));
  }
}

template <int EB>
int pass_tiles(int N, bool col, int outer, int inner, int radix = 16) {
  const bool r8 = radix_for(N, radix) == 8;
  switch (N) {
#define QPCH_CASE(V)                                                                              \
  case V:                                                                                         \
    if (r8) return col ? pass_tiles_t<V, true, EB, 8>(outer, inner) : pass_tiles_t<V, false, EB, 8>(outer, inner); \
    return col ? pass_tiles_t<V, true, EB>(outer, inner) : pass_tiles_t<V, false, EB>(outer, inner);
    QPCH_FOR_EACH_N(QPCH_CASE)
#undef QPCH_CASE
    default:
      throw InvalidArgument("no fused pass for axis length " + std::to_string(N));
  }
}

// --- 3D plane pass: (ny, nx) planes of 1024 to 16384 elements, both axes >= 16 ---------

#define QPCH_FOR_EACH_PLANE(X) \
  X(16, 64) X(16, 128) X(32, 32) X(32, 64) X(32, 128) X(64, 16) X(64, 32) X(64, 64) X(64, 128) \
  X(128, 16) X(128, 32) X(128, 64) X(128, 128)

bool plane_supported(int ny, int nx) {
#define QPCH_CASE(A, B) \
  if (ny == A && nx == B) return true;
  QPCH_FOR_EACH_PLANE(QPCH_CASE)
#undef QPCH_CASE
  return false;
}

// --- whole 2D images in one work-group (image_iterations): the plane shapes up to 8192 ----

#define QPCH_FOR_EACH_IMAGE(X) \
  X(16, 64) X(16, 128) X(32, 32) X(32, 64) X(32, 128) X(64, 16) X(64, 32) X(64, 64) X(64, 128) \
  X(128, 16) X(128, 32) X(128, 64)

bool image_supported(int ny, int nx) {
#define QPCH_CASE(A, B) \
  if (ny == A && nx == B) return true;
  QPCH_FOR_EACH_IMAGE(QPCH_CASE)
#undef QPCH_CASE
  return false;
}

void launch_image(int ny, int nx, const PassArgs& p, int starts, int len, hipStream_t s) {
#define QPCH_CASE(A, B)                                                                          \
  if (ny == A && nx == B) {                                                                      \
    hipLaunchKernelGGL((dev::image_iterations<A, B>), dim3(starts), dim3(dev::PlaneGeom<A, B>::kThreads), 0, s, p, \
                       len);                                                                     \
    QPCH_HIP(hipGetLastError());                                                                 \
    return;                                                                                      \
  }
  QPCH_FOR_EACH_IMAGE(QPCH_CASE)
#undef QPCH_CASE
  throw InvalidArgument("no whole-image kernel for " + std::to_string(ny) + " x " + std::to_string(nx));
}

template <int NY, int NX, class St>
void launch_plane_t(PassArgs a, int starts, hipStream_t s) {
  using G = dev::PlaneGeom<NY, NX>;
  a.starts = starts;
  a.slots = a.no_xcd ? 0 : xcd_slots(starts);
  const int gx = a.outer * (a.slots > 0 ? a.slots : starts);
  hipLaunchKernelGGL((dev::plane_update<NY, NX, St>), dim3(gx), dim3(G::kThreads), 0, s, a);
  QPCH_HIP(hipGetLastError());
}

template <class St>
void launch_plane(int ny, int nx, const PassArgs& a, int starts, hipStream_t s) {
#define QPCH_CASE(A, B)                                  \
  if (ny == A && nx == B) return launch_plane_t<A, B, St>(a, starts, s);
  QPCH_FOR_EACH_PLANE(QPCH_CASE)
#undef QPCH_CASE
  throw InvalidArgument("no plane pass for this plane shape");
}

// --- Shrinkwrap blur passes ---------------------------------------------------------------

template <int N, bool COL, dev::Blur M, class St>
void launch_blur_t(dev::BlurArgs a, int starts, hipStream_t s) {
  using G = dev::PassGeom<N, COL, 8>;
  const int tiles = pass_tiles_t<N, COL, 8>(a.outer, a.inner);
  a.starts = starts;
  a.slots = xcd_slots(starts);
  hipLaunchKernelGGL((dev::blur_pass<N, COL, M, St>), dim3(tiles * (a.slots > 0 ? a.slots : starts)),
                     dim3(G::kThreads), 0, s, a);
  QPCH_HIP(hipGetLastError());
}

template <int N, class St>
void launch_blur_n(bool col, dev::Blur m, const dev::BlurArgs& a, int starts, hipStream_t s) {
  if (m == dev::Blur::kMag) {
    return col ? launch_blur_t<N, true, dev::Blur::kMag, St>(a, starts, s)
               : launch_blur_t<N, false, dev::Blur::kMag, St>(a, starts, s);
  }
  if (m == dev::Blur::kGauss) {
    return col ? launch_blur_t<N, true, dev::Blur::kGauss, float2>(a, starts, s)
               : launch_blur_t<N, false, dev::Blur::kGauss, float2>(a, starts, s);
  }
  return col ? launch_blur_t<N, true, dev::Blur::kReal, float2>(a, starts, s)
             : launch_blur_t<N, false, dev::Blur::kReal, float2>(a, starts, s);
}

template <class St>
void launch_blur(int N, bool col, dev::Blur m, const dev::BlurArgs& a, int starts, hipStream_t s) {
  switch (N) {
#define QPCH_CASE(V) \
  case V:            \
    return launch_blur_n<V, St>(col, m, a, starts, s);
    QPCH_FOR_EACH_N(QPCH_CASE)
#undef QPCH_CASE
    default:
      throw InvalidArgument("no blur pass for axis length " + std::to_string(N));
  }
}

// ---------------------------------------------------------------------------------------
// generic (any length) transforms

// --- Bluestein tables: the chirp c_j = exp(-i pi j^2 / n) (j < n) and the transformed chirp
// filter bhat = DFT_m(b), b_j = conj(c_|j|) for |j| < n (indices mod m), zero elsewhere;
// evaluated on the host in double (j^2 reduced mod 2n first), rounded once to the kernel's
// precision, uploaded once per device and length, and kept until the process ends.
struct BluesteinTables {
  const void* chirp = nullptr;
  const void* bhat = nullptr;
};

// The padded length of Bluestein's algorithm: a power of two >= 2n - 1, and at least 256
// (the smallest transform of the chirp-z passes; the generic kernel takes Bluestein from
// 256 on, where 2n - 1 is larger anyway).  The tables of a length serve both kernels.
int bluestein_length(int n) { return chirp_length(n); }

// In-place forward DFT of a power-of-two length in double (radix 2, twiddles evaluated
// directly at every use).
void host_fft_pow2(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  const double pi = 3.14159265358979323846264338327950288;
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t k = 0; k < len / 2; ++k) {
      const double ang = -2.0 * pi * static_cast<double>(k) / static_cast<double>(len);
      const std::complex<double> w(std::cos(ang), std::sin(ang));
      for (size_t i = 0; i < n; i += len) {
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

std::mutex g_bluestein_mu;
std::map<std::tuple<int, int, bool>, BluesteinTables> g_bluestein;  // (device, n, fp64)

bool needs_bluestein(int n);

const BluesteinTables* bluestein_lookup(int n, bool fp64) {
  int dev = 0;
  QPCH_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_bluestein_mu);
  auto it = g_bluestein.find(std::make_tuple(dev, n, fp64));
  return it == g_bluestein.end() ? nullptr : &it->second;
}

const BluesteinTables& bluestein_tables(int n, bool fp64) {
  int dev = 0;
  QPCH_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_bluestein_mu);
  const auto key = std::make_tuple(dev, n, fp64);
  auto it = g_bluestein.find(key);
  if (it != g_bluestein.end()) return it->second;
  const int m = bluestein_length(n);
  const double pi = 3.14159265358979323846264338327950288;
  std::vector<std::complex<double>> chirp(n), filt(m, 0.0);
  for (int j = 0; j < n; ++j) {
    const long long q = (static_cast<long long>(j) * j) % (2LL * n);
    const double ang = -pi * static_cast<double>(q) / static_cast<double>(n);
    chirp[j] = std::complex<double>(std::cos(ang), std::sin(ang));
  }
  filt[0] = std::conj(chirp[0]);
  for (int j = 1; j < n; ++j) filt[j] = filt[m - j] = std::conj(chirp[j]);
  host_fft_pow2(filt);
  BluesteinTables t;
  const OwnStream st;
  auto put = [&](const std::vector<std::complex<double>>& v) -> const void* {
    void* d = nullptr;
    if (fp64) {
      QPCH_HIP(hipMalloc(&d, sizeof(double2) * v.size()));
      std::vector<double2> h(v.size());
      for (size_t k = 0; k < v.size(); ++k) h[k] = {v[k].real(), v[k].imag()};
      copy_sync(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, st.s);
    } else {
      QPCH_HIP(hipMalloc(&d, sizeof(float2) * v.size()));
      std::vector<float2> h(v.size());
      for (size_t k = 0; k < v.size(); ++k) h[k] = {static_cast<float>(v[k].real()), static_cast<float>(v[k].imag())};
      copy_sync(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, st.s);
    }
    return d;
  };
  t.chirp = put(chirp);
  t.bhat = put(filt);
  return g_bluestein.emplace(key, t).first->second;
}

// The plan of a generic line transform of length n in the kernel's precision; lengths with a
// prime factor above 64 from dev::kBluesteinMinN on take Bluestein's algorithm (their
// tables are made on first use: outside any graph capture, which solver set-up ensures).
dev::GenericPlan generic_plan(int n, bool fp64 = false) {
  dev::GenericPlan p{};
  p.n = n;
  int rest = n, nf = 0;
  bool ok = true;
  while (rest % 4 == 0) { p.fac[nf++] = 4; rest /= 4; }
  while (rest % 2 == 0) { p.fac[nf++] = 2; rest /= 2; }
  for (int f = 3; rest > 1 && ok; f += 2) {
    while (rest % f == 0) {
      if (f > 64 || nf >= 32) { ok = false; break; }
      p.fac[nf++] = f;
      rest /= f;
    }
    if (ok && rest > 1 && f * f > rest) {
      if (rest > 64 || nf >= 32) { ok = false; break; }
      p.fac[nf++] = rest;
      rest = 1;
    }
  }
  p.nfac = ok ? nf : 0;
  if (p.nfac == 0 && n >= dev::kBluesteinMinN) {
    p.m = bluestein_length(n);
    int rest2 = p.m, k = 0;
    while (rest2 % 4 == 0) { p.mfac[k++] = 4; rest2 /= 4; }
    if (rest2 == 2) p.mfac[k++] = 2;
    p.nmfac = k;
    const BluesteinTables& t = bluestein_tables(n, fp64);
    p.chirp = t.chirp;
    p.bhat = t.bhat;
  }
  return p;
}

bool needs_bluestein(int n) {
  if (n < dev::kBluesteinMinN || n > dev::kGenericMaxN || (n & (n - 1)) == 0) return false;
  int rest = n;
  for (int f = 2; f <= 64 && rest > 1; ++f) {
    while (rest % f == 0) rest /= f;
  }
  return rest > 1;  // a prime factor above 64 remains
}

// Dynamic LDS above the default 64 KiB is allowed up to the device's 160 KiB; a plan that
// needed more would fail its launch with an error.
void check_lds(size_t bytes) {
  if (bytes > 160 * 1024) throw InvalidArgument("generic transform needs " + std::to_string(bytes) + " bytes of LDS");
}

void launch_generic(float2* data, long long n_per_start, int batch, int N, long long outer,
                    long long inner, bool forward, float scale, hipStream_t s) {
  if (N > dev::kGenericMaxN) {
    throw InvalidArgument("axis length " + std::to_string(N) +
                          " is neither a power of two <= 4096 nor <= 2048");
  }
  const dev::GenericPlan plan = generic_plan(N);
  const long long lines = outer * inner;
  const long long lpb = dev::generic_lines_per_block(N);
  check_lds(dev::generic_lds_bytes<float2>(plan));
  hipLaunchKernelGGL(dev::generic_fft<float>, dim3(static_cast<unsigned>((lines + lpb - 1) / lpb), batch),
                     dim3(dev::kGenericThreads), dev::generic_lds_bytes<float2>(plan), s, data, n_per_start,
                     static_cast<int>(inner), lines, plan, forward ? 1 : 0, scale);
  QPCH_HIP(hipGetLastError());
}

// A fused pass (line_pass semantics) along an axis of any length up to kGenericMaxN, fp32.
void launch_generic_pass(int N, Mode m, PassArgs a, int starts, hipStream_t s) {
  if (N > dev::kGenericMaxN) throw InvalidArgument("no generic pass for axis length " + std::to_string(N));
  const long long lines = static_cast<long long>(a.outer) * a.inner;
  // lengths with a prime factor of 7 or more run their passes in chirp_pass, whose tiles
  // hold chirp_lines(N) lines
  const bool chirp = m != Mode::kGrad && chirp_axis(N);
  const bool smooth = m != Mode::kGrad && smooth_axis(N);
  const long long lpb = chirp ? chirp_lines(N) : smooth ? smooth_lines(N) : dev::generic_lines_per_block(N);
  const unsigned tiles = static_cast<unsigned>((lines + lpb - 1) / lpb);
  // Strided axes number their blocks XCD-affine, as the line passes do: a block reads a
  // few adjacent columns, and the other pieces of the same 128-byte lines are then read by
  // blocks on the same XCD, from its L2 (240^2 x 8 starts: the column modulus pass 16.7 ->
  // 12.4 us).  Row passes read whole lines and keep the (tiles, starts) grid, which
  // measured the same or up to 3% faster.
  a.starts = starts;
  a.slots = (a.inner > 1 && !a.no_xcd) ? xcd_slots(starts) : 0;
  const dim3 grid = a.slots > 0 ? dim3(tiles * a.slots) : dim3(tiles, starts);
  // Smooth lengths (N = R1 R2, R1 a power of two up to 512, R2 in {3, 5, 15}) run their
  // passes with the line in registers (smooth.hpp): same tiles, partials and numbering
  if (smooth) {
    QPCH_HIP(smooth_launch(N, a.inner > 1, m, a, lines, grid, s));
    return;
  }
  // Bluestein's algorithm with the power-of-two transforms in registers (chirp.hpp); the
  // tables exist from solver set-up on, so this lookup allocates nothing
  if (chirp) {
    const BluesteinTables& t = bluestein_tables(N, false);
    QPCH_HIP(chirp_launch(N, a.inner > 1, m, a, lines, grid, t.chirp, t.bhat, s));
    return;
  }
  const dev::GenericPlan plan = generic_plan(N);
  const size_t lds = dev::generic_pass_lds_bytes(plan);
  check_lds(lds);
  switch (m) {
    case Mode::kFwd: hipLaunchKernelGGL(dev::generic_pass<Mode::kFwd>, grid, dim3(dev::kGenericThreads), lds, s, a, plan, lines); break;
    case Mode::kInv: hipLaunchKernelGGL(dev::generic_pass<Mode::kInv>, grid, dim3(dev::kGenericThreads), lds, s, a, plan, lines); break;
    case Mode::kModulus: hipLaunchKernelGGL(dev::generic_pass<Mode::kModulus>, grid, dim3(dev::kGenericThreads), lds, s, a, plan, lines); break;
    case Mode::kUpdate: hipLaunchKernelGGL(dev::generic_pass<Mode::kUpdate>, grid, dim3(dev::kGenericThreads), lds, s, a, plan, lines); break;
    default: throw InvalidArgument("the generic pass has no gradient mode");
  }
  QPCH_HIP(hipGetLastError());
}

// Error partials per start of the modulus pass along axis 0 (line_pass: one per wave;
// generic_pass: one per work-group).
template <int EB>
int axis0_partials(const Dims& d, int radix) {
  const int o = static_cast<int>(d.outer(0)), in = static_cast<int>(d.inner(0));
  if (generic_axis(d.d[0])) {
    const long long lpb = chirp_axis(d.d[0])    ? chirp_lines(d.d[0])
                          : smooth_axis(d.d[0]) ? smooth_lines(d.d[0])
                                                : dev::generic_lines_per_block(d.d[0]);
    return static_cast<int>((static_cast<long long>(o) * in + lpb - 1) / lpb);
  }
  return pass_partials<EB>(d.d[0], in > 1, o, in, radix);
}

// Unnormalised FP64 transform along every axis (generic kernel, lines up to 1024).
void device_fft_f64(double2* data, const Dims& d, int batch, bool forward, hipStream_t s) {
  for (int a = d.nd - 1; a >= 0; --a) {
    const int N = d.d[a];
    if (N == 1) continue;
    if (N > 1024) throw InvalidArgument("the FP64 transform supports axes up to 1024");
    const dev::GenericPlan plan = generic_plan(N, true);
    const long long lines = d.outer(a) * d.inner(a);
    const long long lpb = dev::generic_lines_per_block(N);
    check_lds(dev::generic_lds_bytes<double2>(plan));
    hipLaunchKernelGGL(dev::generic_fft<double>, dim3(static_cast<unsigned>((lines + lpb - 1) / lpb), batch),
                       dim3(dev::kGenericThreads), dev::generic_lds_bytes<double2>(plan), s, data, d.n,
                       static_cast<int>(d.inner(a)), lines, plan, forward ? 1 : 0, 1.0);
    QPCH_HIP(hipGetLastError());
  }
}

void scale_inplace(float2* data, long long total, float scale, hipStream_t s) {
  hipLaunchKernelGGL((dev::convert<float2, float2>), dim3(grid_for(total)), dim3(256), 0, s, data,
                     data, total, scale, nullptr);
  QPCH_HIP(hipGetLastError());
}

// Unitary (scale = 1/sqrt(n)) or unnormalised (scale = 1) n-D transform in place, fp32.
void device_fft(float2* data, const Dims& d, int batch, bool forward, float scale,
                hipStream_t s) {
  bool scaled = false;
  for (int a = d.nd - 1; a >= 0; --a) {
    const int N = d.d[a];
    if (N == 1) continue;
    const bool last_axis = [&] {
      for (int b = a - 1; b >= 0; --b)
        if (d.d[b] > 1) return false;
      return true;
    }();
    const float sc = last_axis ? scale : 1.0f;
    scaled = scaled || last_axis;
    const long long inner = d.inner(a), outer = d.outer(a);
    if ((is_pow2(N) && N <= kMaxFused) || smooth_axis(N)) {
      // power-of-two lines in line_pass; smooth lengths in smooth_pass (tables made in the
      // kernel: nothing is allocated, so the no-allocation entry points may take this path)
      PassArgs p{};
      p.outer = static_cast<int>(outer);
      p.inner = static_cast<int>(inner);
      p.n = d.n;
      p.w = data;
      p.out_scale = sc;
      if (is_pow2(N)) {
        launch_pass<float2>(N, inner > 1, forward ? Mode::kFwd : Mode::kInv, p, batch, s);
      } else {
        launch_generic_pass(N, forward ? Mode::kFwd : Mode::kInv, p, batch, s);
      }
    } else {
      launch_generic(data, d.n, batch, N, outer, inner, forward, sc, s);
    }
  }
  if (!scaled && scale != 1.0f) scale_inplace(data, d.n * batch, scale, s);
}

// ---------------------------------------------------------------------------------------
// Shrinkwrap blur taps: h[d] = (1/N) sum_j exp(k f_j^2) cos(2 pi j d / N), the inverse
// DFT of the axis's Gaussian transfer (real and symmetric), k = -2 pi^2 sigma^2.

struct Taps {
  std::vector<float> taps;
  int J = 0;  // offset of tap 0
};

Taps blur_taps(int N, float sigma) {
  const double pi = 3.14159265358979323846;
  const double k = -2.0 * pi * pi * static_cast<double>(sigma) * static_cast<double>(sigma);
  std::vector<double> g(N), h(N);
  for (int j = 0; j < N; ++j) {
    const int s = j < (N + 1) / 2 ? j : j - N;
    const double f = static_cast<double>(s) / N;
    g[j] = std::exp(k * f * f);
  }
  for (int dd = 0; dd < N; ++dd) {
    double acc = 0.0;
    for (int j = 0; j < N; ++j) acc += g[j] * std::cos(2.0 * pi * static_cast<double>((static_cast<long long>(j) * dd) % N) / N);
    h[dd] = acc / N;
  }
  // keep |h[d]| > 1e-12 h[0] around d = 0; the rest cannot change a float sum
  int J = 0;
  for (int dd = 1; dd <= N / 2; ++dd) {
    if (std::fabs(h[dd]) > 1e-12 * std::fabs(h[0]) || std::fabs(h[N - dd]) > 1e-12 * std::fabs(h[0])) J = dd;
  }
  Taps t;
  if (2 * J + 1 >= N) {  // the kernel spans the axis: full circular convolution
    t.J = 0;
    t.taps.resize(N);
    for (int dd = 0; dd < N; ++dd) t.taps[dd] = static_cast<float>(h[(N - dd) % N]);
    // tap t multiplies in[(k + t) mod N]: weight h[(k - (k + t)) mod N] = h[(-t) mod N]
  } else {
    t.J = J;
    t.taps.resize(2 * J + 1);
    for (int q = 0; q < 2 * J + 1; ++q) t.taps[q] = static_cast<float>(h[(J - q + N) % N]);
  }
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// the solver

struct SolverImpl {
  Dims dims;
  long long n = 0;
  int starts = 1, batch = 1, groups = 1;
  int storage = QPCH_F32;
  bool mixed = false;
  int schedule = QPCH_GRAPH;
  bool fused = true;  // fused passes (generic ones on axes that are not powers of two)
  bool pow2 = true;   // every axis a power of two <= 4096: the FFT blur, fp16 / bf16 storage
  hipStream_t stream = nullptr;
  hipStream_t on = nullptr;   // where run_pass and chunk launch: the solver's stream or a lane
  int lanes = 1;              // launch groups in flight at once (each on its own stream)
  std::vector<hipStream_t> lane_streams;
  std::vector<hipEvent_t> lane_done;
  hipEvent_t lane_fork = nullptr;
  float amp_scale = 1.0f;  // problem 0's
  double energy = 0.0;     // problem 0's
  int problems = 1, per_problem = 1;
  std::vector<float> amp_scales;
  std::vector<double> energies;

  DevBuf<float> amp_signed, amp_true;
  DevBuf<std::uint8_t> supp, saved;
  DevBuf<unsigned long long> detwin_sums;  // [start][2]: support moment and count along the detwin axis
  DevBuf<float2> x32, w32;
  DevBuf<std::uint32_t> x16, w16;  // half2 / bfloat162 storage, 4 bytes per element
  DevBuf<double> partials, errs, sw_part;
  DevBuf<int> counters;
  DevBuf<float> sw_a, sw_b;
  DevBuf<float2> sw_c;  // the blur's spectrum scratch
  int err_cap = 0;
  bool low_active = false;
  bool wave_valid = false;
  int generation = 0;  // bumped when buffers that graphs capture are reallocated
  std::map<std::tuple<int, float, float>, std::pair<int, DevBuf<float>>> taps_cache;  // (N, sigma) -> (J, taps)
  std::map<std::tuple<int, int, float, float, bool>, hipGraphExec_t> graphs;

  ~SolverImpl() {
    // Drain first: graph execs and buffers must not go while a launch may still use them.
    for (hipStream_t l : lane_streams) (void)hipStreamSynchronize(l);
    if (stream) (void)hipStreamSynchronize(stream);
    clear_graphs();
    for (hipStream_t l : lane_streams) (void)hipStreamDestroy(l);
    for (hipEvent_t e : lane_done) (void)hipEventDestroy(e);
    if (lane_fork) (void)hipEventDestroy(lane_fork);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void clear_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  bool low() const { return storage != QPCH_F32; }

  template <class St>
  St* xp(int first) const {
    if constexpr (std::is_same_v<St, float2>) return x32.get() + first * n;
    else return reinterpret_cast<St*>(x16.get()) + first * n;
  }
  template <class St>
  St* wp(int first) const {
    if constexpr (std::is_same_v<St, float2>) return w32.get() + first * n;
    else return reinterpret_cast<St*>(w16.get()) + first * n;
  }

  // --- fused schedule: the passes of one iteration and of the prologue -------------------
  struct PassSpec {
    int axis;
    Mode mode;
    bool from_x;  // prologue: read x instead of the wave
    float in_scale, out_scale;
    bool plane = false;  // 3D: the update on whole xy-planes (axes 1 and 2 at once)
  };
  bool plane3d = false;  // 3D iterations use the plane pass
  bool resident = false;  // 2D fp32 iterations run whole chunks in image_iterations
  int xcd_mode = 0;  // 0 XCD-affine by start, 2 plain numbering, 3 plain for the modulus pass only
  int radix = 16;        // FFT radix of the line passes (8 or 16)
  bool persist = false;  // 2D chunks of iterations in one persistent launch where a kernel exists
  // Images whose two axes are not powers of two run persistent launches by default only
  // when they are square of a smooth length (twopass_smooth); the others (twopass_generic)
  // only when asked (qpch_problem.persist = 1): there the per-pass launches are faster.
  bool persist_generic = false;
  DevBuf<dev::GenericPlan> generic_plans;  // twopass_generic: the plans of axes 0 and 1 in device memory
  DevBuf<int> psync;  // a line holding the abort flag, then per launch group: next[8 lines], done, count[2][batch]
  long long psync_group() const {
    return static_cast<long long>(dev::kXcds + 1 + 2 * batch) * dev::kLineInts;  // (kXcds: room for the most)
  }
  long long wait_ticks = 0;  // bound of one persistent wait: half a second of wall_clock64
  int persist_xcds = 0;      // XCDs the persistent launches spread starts over
  int persist_workers = 0;   // work-groups per persistent launch (0: sized by the launcher)

  std::vector<PassSpec> iteration_passes(bool lowp) const {
    const int nd = dims.nd;
    std::vector<PassSpec> v;
    const float s = static_cast<float>(1.0 / std::sqrt(static_cast<double>(n)));
    auto ax = [&](int a) { return lowp ? static_cast<float>(1.0 / std::sqrt(static_cast<double>(dims.d[a]))) : 1.0f; };
    const float ms = lowp ? ax(0) : s;
    v.push_back({0, Mode::kModulus, false, ms, ms});
    if (nd == 3 && plane3d) {
      const float ps = lowp ? ax(1) * ax(2) : 1.0f;
      v.push_back({2, Mode::kUpdate, false, ps, ps, true});
      return v;
    }
    if (nd == 3) v.push_back({1, Mode::kInv, false, 1.0f, ax(1)});
    v.push_back({nd - 1, Mode::kUpdate, false, ax(nd - 1), ax(nd - 1)});
    if (nd == 3) v.push_back({1, Mode::kFwd, false, 1.0f, ax(1)});
    return v;
  }

  std::vector<PassSpec> prologue_passes(bool lowp) const {
    const int nd = dims.nd;
    auto ax = [&](int a) { return lowp ? static_cast<float>(1.0 / std::sqrt(static_cast<double>(dims.d[a]))) : 1.0f; };
    std::vector<PassSpec> v;
    v.push_back({nd - 1, Mode::kFwd, true, 1.0f, ax(nd - 1)});
    if (nd == 3) v.push_back({1, Mode::kFwd, false, 1.0f, ax(1)});
    return v;
  }

  // The argument block of one pass of launch group g (count starts).
  template <class St>
  PassArgs pass_args(const PassSpec& ps, int g, int count, float a, float b) const {
    const int first = g * batch;
    PassArgs p{};
    const long long inner = dims.inner(ps.axis), outer = dims.outer(ps.axis);
    p.outer = static_cast<int>(outer);
    p.inner = static_cast<int>(inner);
    p.n = n;
    p.w = wp<St>(first);
    p.src = ps.from_x ? static_cast<const void*>(xp<St>(first)) : nullptr;
    p.x = xp<St>(first);
    p.support = supp.get() + first * n;
    p.amp = amp_signed.get();
    p.amp_stride = problems > 1 ? n : 0;
    p.amp_group = per_problem;
    p.partial_stride = axis0_partials<sizeof(St)>(dims, radix);
    p.partials = partials.get() + static_cast<long long>(first) * p.partial_stride;  // own rows per group
    p.errs = errs.get();
    p.err_stride = starts;
    p.err_offset = first;
    p.err_cap = err_cap;
    p.counter = counters.get() + g;
    p.a = a;
    p.b = b;
    p.no_xcd = (xcd_mode == 2 || (xcd_mode == 3 && ps.mode == Mode::kModulus)) ? 1 : 0;
    p.radix = radix;
    p.in_scale = ps.in_scale;
    p.out_scale = ps.out_scale;
    p.starts = count;
    return p;
  }

  template <class St>
  void run_pass(const PassSpec& ps, int g, int count, float a, float b) {
    PassArgs p = pass_args<St>(ps, g, count, a, b);
    const long long inner = dims.inner(ps.axis);
    if (ps.plane) {
      p.outer = dims.d[0];  // one tile per plane
      launch_plane<St>(dims.d[1], dims.d[2], p, count, on ? on : stream);
      return;
    }
    if (generic_axis(dims.d[ps.axis])) {
      if constexpr (!std::is_same_v<St, float2>) {
        throw InvalidArgument("fp16/bf16 storage needs power-of-two extents");
      } else {
        launch_generic_pass(dims.d[ps.axis], ps.mode, p, count, on ? on : stream);
        return;
      }
    }
    launch_pass<St>(dims.d[ps.axis], inner > 1, ps.mode, p, count, on ? on : stream);
  }

  template <class St>
  void fused_prologue(int g) {
    if (std::is_same_v<St, float2> && resident) return;  // image_iterations keeps no wave
    for (const auto& ps : prologue_passes(!std::is_same_v<St, float2>)) run_pass<St>(ps, g, batch, 0, 0);
  }

  // A chunk of fp32 iterations of every start of group g in image_iterations (which also
  // advances the group's error-slot counter past the chunk).
  void resident_iterations(int g, int len, float a, float b) {
    const int first = g * batch;
    PassArgs p{};
    p.n = n;
    p.x = x32.get() + static_cast<long long>(first) * n;
    p.support = supp.get() + static_cast<long long>(first) * n;
    p.amp = amp_signed.get();
    p.amp_stride = problems > 1 ? n : 0;
    p.amp_group = per_problem;
    p.errs = errs.get();
    p.err_stride = starts;
    p.err_offset = first;
    p.err_cap = err_cap;
    p.counter = counters.get() + g;
    p.done = counters.get() + groups + g;  // tickets: the second half of counters
    p.a = a;
    p.b = b;
    p.in_scale = p.out_scale = static_cast<float>(1.0 / std::sqrt(static_cast<double>(n)));
    launch_image(dims.d[0], dims.d[1], p, batch, len, on ? on : stream);
  }

  template <class St>
  static constexpr int storage_code() {
    return std::is_same_v<St, float2> ? QPCH_F32 : std::is_same_v<St, __half2> ? QPCH_F16 : QPCH_BF16;
  }

  // Whether chunks of this storage run as one persistent launch: twopass_iterations
  // (power-of-two axes) or twopass_generic (fp32, both axes not powers of two, persist = 1).
  template <class St>
  bool persistent() const {
    if (!persist || dims.nd != 2) return false;
    if (radix == 16 && twopass_supported(dims.d[0], dims.d[1], storage_code<St>())) return true;
    if constexpr (std::is_same_v<St, float2>) {
      if (twopass_smooth_supported(dims.d[0], dims.d[1])) return true;
      if (persist_generic && generic_axis(dims.d[0]) && generic_axis(dims.d[1]) && !smooth_axis(dims.d[0]) &&
          !smooth_axis(dims.d[1]) && !chirp_axis(dims.d[0]) && !chirp_axis(dims.d[1])) {
        return twopass_generic_supported(generic_plan(dims.d[0]), generic_plan(dims.d[1]));
      }
    }
    return false;
  }

  // A chunk of iterations of launch group g in one persistent launch; false where the
  // shape and storage have no persistent kernel.
  template <class St>
  bool persistent_iterations(int g, int len, float a, float b) {
    if (!persistent<St>()) return false;
    const auto passes = iteration_passes(!std::is_same_v<St, float2>);
    PassArgs p0 = pass_args<St>(passes[0], g, batch, a, b);
    PassArgs p1 = pass_args<St>(passes[1], g, batch, a, b);
    p0.counter = nullptr;  // the launch advances the slot counter itself
    dev::SyncArgs q{};
    int* base = psync.get() + static_cast<long long>(g) * psync_group() + dev::kLineInts;
    q.next = base;
    q.done = base + dev::kXcds * dev::kLineInts;
    q.count = q.done + dev::kLineInts;
    q.abort = psync.get();
    q.xcds = persist_xcds;
    q.wait_limit = wait_ticks;
    hipStream_t st = on ? on : stream;
    if (generic_axis(dims.d[0])) {
      if constexpr (std::is_same_v<St, float2>) {
        if (twopass_smooth_supported(dims.d[0], dims.d[1])) {
          const int N = dims.d[0];
          q.tiles0 = q.tiles1 = (N + smooth_lines(N) - 1) / smooth_lines(N);
          QPCH_HIP(twopass_smooth_launch(N, p0, p1, q, len, persist_workers, st));
          return true;
        }
        const dev::GenericPlan g0 = generic_plan(dims.d[0]), g1 = generic_plan(dims.d[1]);
        const long long lines0 = dims.d[1], lines1 = dims.d[0];  // axis 0: columns; axis 1: rows
        q.tiles0 = static_cast<int>((lines0 + dev::generic_lines_per_block(dims.d[0]) - 1) /
                                    dev::generic_lines_per_block(dims.d[0]));
        q.tiles1 = static_cast<int>((lines1 + dev::generic_lines_per_block(dims.d[1]) - 1) /
                                    dev::generic_lines_per_block(dims.d[1]));
        QPCH_HIP(twopass_generic_launch(p0, p1, g0, g1, generic_plans.get(), lines0, lines1, q, len, persist_workers, st));
        return true;
      }
    }
    q.tiles0 = pass_tiles<sizeof(St)>(dims.d[0], true, 1, dims.d[1], 16);
    q.tiles1 = twopass_row_tiles(dims.d[0], dims.d[1], storage_code<St>());
    QPCH_HIP(twopass_launch(dims.d[0], dims.d[1], storage_code<St>(), p0, p1, q, len, persist_workers, st));
    return true;
  }

  template <class St>
  void fused_iterations(int g, int len, float a, float b) {
    if constexpr (std::is_same_v<St, float2>) {
      if (resident) return resident_iterations(g, len, a, b);
    }
    if (persistent_iterations<St>(g, len, a, b)) return;
    const auto passes = iteration_passes(!std::is_same_v<St, float2>);
    for (int it = 0; it < len; ++it) {
      for (const auto& ps : passes) run_pass<St>(ps, g, batch, a, b);
    }
  }

  // --- unfused schedule (any shape, fp32) -------------------------------------------------
  void unfused_iterations(int g, int len, float a, float b) {
    const int first = g * batch;
    float2* x = x32.get() + first * n;
    float2* w = w32.get() + first * n;
    const float s = static_cast<float>(1.0 / std::sqrt(static_cast<double>(n)));
    const int mgrid = grid_for(n, 256, 1024);
    for (int it = 0; it < len; ++it) {
      QPCH_HIP(hipMemcpyAsync(w, x, sizeof(float2) * n * batch, hipMemcpyDeviceToDevice, stream));
      device_fft(w, dims, batch, true, 1.0f, stream);
      hipLaunchKernelGGL(dev::modulus_elementwise, dim3(mgrid, batch), dim3(256), 0, stream, w,
                         amp_signed.get(), n, s, s, partials.get(), counters.get() + g);
      QPCH_HIP(hipGetLastError());
      device_fft(w, dims, batch, false, 1.0f, stream);
      hipLaunchKernelGGL(dev::update_elementwise, dim3(grid_for(n, 256, 1024), batch), dim3(256), 0,
                         stream, x, w, supp.get() + first * n, n, a, b, 1.0f, partials.get(), mgrid,
                         errs.get(), starts, first, err_cap, counters.get() + g);
      QPCH_HIP(hipGetLastError());
    }
  }

  // --- precision switching -----------------------------------------------------------------
  // x between fp32 and the reduced storage, for starts [first, first + count), on stream s.
  void convert_range(bool to_low, int first, int count, hipStream_t s) {
    const long long total = n * count;
    const long long off = n * first;
    float2* x32p = x32.get() + off;
    std::uint32_t* x16p = x16.get() + off;
    if (to_low) {
      if (storage == QPCH_F16) {
        hipLaunchKernelGGL((dev::convert<float2, __half2>), dim3(grid_for(total)), dim3(256), 0, s, x32p,
                           reinterpret_cast<__half2*>(x16p), total, 1.0f, nullptr);
      } else {
        hipLaunchKernelGGL((dev::convert<float2, __hip_bfloat162>), dim3(grid_for(total)), dim3(256), 0, s, x32p,
                           reinterpret_cast<__hip_bfloat162*>(x16p), total, 1.0f, nullptr);
      }
    } else {
      if (storage == QPCH_F16) {
        hipLaunchKernelGGL((dev::convert<__half2, float2>), dim3(grid_for(total)), dim3(256), 0, s,
                           reinterpret_cast<const __half2*>(x16p), x32p, total, 1.0f, nullptr);
      } else {
        hipLaunchKernelGGL((dev::convert<__hip_bfloat162, float2>), dim3(grid_for(total)), dim3(256), 0, s,
                           reinterpret_cast<const __hip_bfloat162*>(x16p), x32p, total, 1.0f, nullptr);
      }
    }
    QPCH_HIP(hipGetLastError());
  }

  void convert_x(bool to_low) { convert_range(to_low, 0, starts, stream); }

  void set_precision(bool lowp) {
    if (lowp == low_active) return;
    convert_x(lowp);
    low_active = lowp;
    wave_valid = false;
  }

  // --- iteration chunk ---------------------------------------------------------------------
  template <class St>
  void chunk(int g, int len, float a, float b, bool need_prologue) {
    if (need_prologue) fused_prologue<St>(g);
    if (schedule != QPCH_GRAPH) {
      fused_iterations<St>(g, len, a, b);
      return;
    }
    const auto key = std::make_tuple(g, len, a, b, !std::is_same_v<St, float2>);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      QPCH_HIP(hipStreamBeginCapture(on, hipStreamCaptureModeThreadLocal));
      try {
        fused_iterations<St>(g, len, a, b);
      } catch (...) {
        hipGraph_t dummy{};
        (void)hipStreamEndCapture(on, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
      }
      hipGraph_t graph{};
      QPCH_HIP(hipStreamEndCapture(on, &graph));
      hipGraphExec_t exec{};
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      QPCH_HIP(e);
      it = graphs.emplace(key, exec).first;
    }
    QPCH_HIP(hipGraphLaunch(it->second, on));
  }

  // Launch groups side by side in one graph per chunk: a branch per lane, forked from and
  // joined to the solver's stream inside the capture, so one launch starts every group.
  template <class St>
  void lanes_chunk(int L, int len, float a, float b, bool need_prologue) {
    if (need_prologue) {
      for (int g = 0; g < groups; ++g) fused_prologue<St>(g);
    }
    const auto key = std::make_tuple(-L, len, a, b, !std::is_same_v<St, float2>);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      QPCH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      hipGraph_t graph{};
      try {
        QPCH_HIP(hipEventRecord(lane_fork, stream));
        for (int l = 0; l < L; ++l) QPCH_HIP(hipStreamWaitEvent(lane_streams[l], lane_fork, 0));
        for (int g = 0; g < groups; ++g) {
          on = lane_streams[g % L];
          fused_iterations<St>(g, len, a, b);
        }
        on = stream;
        for (int l = 0; l < L; ++l) {
          QPCH_HIP(hipEventRecord(lane_done[l], lane_streams[l]));
          QPCH_HIP(hipStreamWaitEvent(stream, lane_done[l], 0));
        }
      } catch (...) {
        on = stream;
        hipGraph_t dummy{};
        (void)hipStreamEndCapture(stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
      }
      QPCH_HIP(hipStreamEndCapture(stream, &graph));
      hipGraphExec_t exec{};
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      QPCH_HIP(e);
      it = graphs.emplace(key, exec).first;
    }
    QPCH_HIP(hipGraphLaunch(it->second, stream));
  }

  // --- group support: the precision switch and prologue run eagerly; the chunk's passes
  // for every launch group are then enqueued on `stream` (inside a capture by the group).
  bool prepare(bool refine) {
    if (!fused) throw InvalidArgument("grouped solvers need the fused schedule");
    if (err_cap <= 0) throw InvalidArgument("qpch_solver_begin_run must precede iterate");
    const bool lowp = low() && !(mixed && refine);
    set_precision(lowp);
    if (!wave_valid) {
      for (int g = 0; g < groups; ++g) {
        if (!lowp) fused_prologue<float2>(g);
        else if (storage == QPCH_F16) fused_prologue<__half2>(g);
        else fused_prologue<__hip_bfloat162>(g);
      }
    }
    return lowp;
  }

  void enqueue(int len, float a, float b, bool lowp) {
    for (int g = 0; g < groups; ++g) {
      if (!lowp) fused_iterations<float2>(g, len, a, b);
      else if (storage == QPCH_F16) fused_iterations<__half2>(g, len, a, b);
      else fused_iterations<__hip_bfloat162>(g, len, a, b);
    }
    wave_valid = true;
  }

  void iterate(int len, float a, float b, bool refine) {
    if (len < 1) throw InvalidArgument("iterations must be >= 1");
    if (err_cap <= 0) throw InvalidArgument("qpch_solver_begin_run must precede iterate");
    const bool lowp = low() && !(mixed && refine);
    set_precision(lowp);
    const bool need_prologue = !wave_valid;
    const int L = fused ? std::min<int>(static_cast<int>(lane_streams.size()), groups) : 0;
    if (L > 1 && schedule == QPCH_GRAPH) {
      if (!lowp) {
        lanes_chunk<float2>(L, len, a, b, need_prologue);
      } else if (storage == QPCH_F16) {
        lanes_chunk<__half2>(L, len, a, b, need_prologue);
      } else {
        lanes_chunk<__hip_bfloat162>(L, len, a, b, need_prologue);
      }
      wave_valid = true;
      return;
    }
    if (L > 1) {
      // launch groups run side by side on the lanes, forked from and joined to the stream
      QPCH_HIP(hipEventRecord(lane_fork, stream));
      for (int l = 0; l < L; ++l) QPCH_HIP(hipStreamWaitEvent(lane_streams[l], lane_fork, 0));
    }
    for (int g = 0; g < groups; ++g) {
      if (!fused) {
        unfused_iterations(g, len, a, b);
        continue;
      }
      on = L > 1 ? lane_streams[g % L] : stream;
      if (!lowp) {
        chunk<float2>(g, len, a, b, need_prologue);
      } else if (storage == QPCH_F16) {
        chunk<__half2>(g, len, a, b, need_prologue);
      } else {
        chunk<__hip_bfloat162>(g, len, a, b, need_prologue);
      }
    }
    on = stream;
    if (L > 1) {
      for (int l = 0; l < L; ++l) {
        QPCH_HIP(hipEventRecord(lane_done[l], lane_streams[l]));
        QPCH_HIP(hipStreamWaitEvent(stream, lane_done[l], 0));
      }
    }
    wave_valid = fused;
  }

  void begin_run(int capacity) {
    if (capacity < 1) throw InvalidArgument("capacity must be >= 1");
    const size_t need = static_cast<size_t>(capacity) * starts;
    if (errs.size() < need) {
      clear_graphs();  // captured graphs hold the history's address
      errs.resize(need);
      ++generation;
    }
    err_cap = capacity;
    hipLaunchKernelGGL(dev::fill_double, dim3(grid_for(static_cast<long long>(errs.size()))), dim3(256), 0,
                       stream, errs.get(), static_cast<long long>(errs.size()), 0.0);
    hipLaunchKernelGGL(dev::fill_int, dim3(1), dim3(256), 0, stream, counters.get(), 2 * groups, 0);
    QPCH_HIP(hipGetLastError());
  }

  // --- Shrinkwrap ------------------------------------------------------------------------
  const std::pair<int, DevBuf<float>>& taps_for(int N, float sigma) {
    const auto key = std::make_tuple(N, sigma, 0.0f);
    auto it = taps_cache.find(key);
    if (it == taps_cache.end()) {
      Taps t = blur_taps(N, sigma);
      DevBuf<float> d = upload(t.taps.data(), t.taps.size(), stream);
      it = taps_cache.emplace(key, std::make_pair(t.J, std::move(d))).first;
    }
    return it->second;
  }

  // exp(-2 pi^2 sigma^2 f^2) for every FFT-order index of an axis of length N
  const float* transfer_for(int N, float sigma) {
    const auto key = std::make_tuple(N, sigma, 1.0f);
    auto it = taps_cache.find(key);
    if (it == taps_cache.end()) {
      const double pi = 3.14159265358979323846;
      const double k = -2.0 * pi * pi * static_cast<double>(sigma) * static_cast<double>(sigma);
      std::vector<float> g(N);
      for (int j = 0; j < N; ++j) {
        const int sj = j < (N + 1) / 2 ? j : j - N;
        const double f = static_cast<double>(sj) / N;
        g[j] = static_cast<float>(std::exp(k * f * f));
      }
      DevBuf<float> d = upload(g.data(), g.size(), stream);
      it = taps_cache.emplace(key, std::make_pair(0, std::move(d))).first;
    }
    return it->second.second.get();
  }

  // Buffers and per-axis tables of a Shrinkwrap with this sigma: made before any capture
  // (the uploads synchronise), sized for every start.
  void shrinkwrap_prepare(float sigma) {
    if (!(sigma >= 0.0f)) throw InvalidArgument("Shrinkwrap needs threshold > 0, sigma >= 0");
    const long long total = n * starts;
    if (sw_a.size() < static_cast<size_t>(total)) sw_a.resize(total);
    if (fused && pow2) {
      if (sw_c.size() < static_cast<size_t>(total)) sw_c.resize(total);
      const int last = dims.nd - 1;
      const size_t need = static_cast<size_t>(pass_tiles<8>(dims.d[last], false, static_cast<int>(dims.outer(last)), 1)) *
                          starts * 2;
      if (sw_part.size() < need) sw_part.resize(need);
      for (int a = 0; a < dims.nd; ++a) (void)transfer_for(dims.d[a], sigma);
      return;
    }
    if (sw_b.size() < static_cast<size_t>(total)) sw_b.resize(total);
    for (int a = 0; a < dims.nd; ++a) (void)taps_for(dims.d[a], sigma);
    const int blocks = grid_for(n, 256, 512);
    if (sw_part.size() < static_cast<size_t>(blocks) * starts * 2) sw_part.resize(static_cast<size_t>(blocks) * starts * 2);
  }

  // The blur through the FFT passes (power-of-two shapes) for starts [first, first + count).
  template <class St>
  void shrinkwrap_fft(float threshold, float sigma, int method, int first, int count, hipStream_t s) {
    const int nd = dims.nd;
    const int last = nd - 1;
    const long long off = n * first;
    const int tiles = pass_tiles<8>(dims.d[last], false, static_cast<int>(dims.outer(last)), 1);
    dev::BlurArgs a{};
    a.n = n;
    a.x = xp<St>(first);
    a.w = sw_c.get() + off;
    a.out = sw_a.get() + off;
    a.support = supp.get() + off;
    a.g0 = transfer_for(dims.d[0], sigma);
    a.g1 = transfer_for(dims.d[1], sigma);
    a.g2 = nd == 3 ? transfer_for(dims.d[2], sigma) : nullptr;
    a.d1 = dims.d[1];
    a.d2 = nd == 3 ? dims.d[2] : 1;
    a.method = method;
    a.scale = static_cast<float>(1.0 / static_cast<double>(n));
    a.partials = sw_part.get() + static_cast<long long>(first) * tiles * 2;
    auto geometry = [&](int axis) {
      a.outer = static_cast<int>(dims.outer(axis));
      a.inner = static_cast<int>(dims.inner(axis));
    };
    auto line = [&](int axis, Mode mode) {
      PassArgs p{};
      p.outer = static_cast<int>(dims.outer(axis));
      p.inner = static_cast<int>(dims.inner(axis));
      p.n = n;
      p.w = sw_c.get() + off;
      p.out_scale = 1.0f;
      launch_pass<float2>(dims.d[axis], p.inner > 1, mode, p, count, s);
    };
    geometry(last);
    launch_blur<St>(dims.d[last], false, dev::Blur::kMag, a, count, s);
    if (nd == 3) line(1, Mode::kFwd);
    geometry(0);
    launch_blur<St>(dims.d[0], true, dev::Blur::kGauss, a, count, s);
    if (nd == 3) line(1, Mode::kInv);
    geometry(last);
    launch_blur<St>(dims.d[last], false, dev::Blur::kReal, a, count, s);
    hipLaunchKernelGGL(dev::shrinkwrap_threshold, dim3(grid_for(n, 256, 2048), count), dim3(256), 0, s,
                       sw_a.get() + off, sw_part.get() + static_cast<long long>(first) * tiles * 2, tiles, n, method,
                       threshold, supp.get() + off);
    QPCH_HIP(hipGetLastError());
  }

  // Shrinkwrap of starts [first, first + count) on stream s, x read in the reduced storage
  // when lowx; shrinkwrap_prepare(sigma) must have run.
  void shrinkwrap_range(float threshold, float sigma, int method, int first, int count, bool lowx, hipStream_t s) {
    if (!(threshold > 0.0f) || !(sigma >= 0.0f)) throw InvalidArgument("Shrinkwrap needs threshold > 0, sigma >= 0");
    if (fused && pow2) {
      if (!lowx) {
        shrinkwrap_fft<float2>(threshold, sigma, method, first, count, s);
      } else if (storage == QPCH_F16) {
        shrinkwrap_fft<__half2>(threshold, sigma, method, first, count, s);
      } else {
        shrinkwrap_fft<__hip_bfloat162>(threshold, sigma, method, first, count, s);
      }
      return;
    }
    const long long off = n * first;
    const int nd = dims.nd;
    float* cur = nullptr;
    bool first_axis = true;
    for (int a = nd - 1; a >= 0; --a) {
      const int N = dims.d[a];
      const auto& tp = taps_for(N, sigma);
      dev::BlurAxis ax{N, static_cast<int>(dims.inner(a)), static_cast<int>(tp.second.size()), tp.first,
                       tp.second.get()};
      float* out = (cur == sw_a.get() + off) ? sw_b.get() + off : sw_a.get() + off;
      const dim3 grid(grid_for(n, 256, 2048), count);
      if (first_axis) {
        if (!lowx) {
          hipLaunchKernelGGL((dev::blur_axis<float2, true>), grid, dim3(256), 0, s, x32.get() + off, out, n, ax);
        } else if (storage == QPCH_F16) {
          hipLaunchKernelGGL((dev::blur_axis<__half2, true>), grid, dim3(256), 0, s, x16.get() + off, out, n, ax);
        } else {
          hipLaunchKernelGGL((dev::blur_axis<__hip_bfloat162, true>), grid, dim3(256), 0, s, x16.get() + off, out, n,
                             ax);
        }
      } else {
        hipLaunchKernelGGL((dev::blur_axis<float2, false>), grid, dim3(256), 0, s, cur, out, n, ax);
      }
      QPCH_HIP(hipGetLastError());
      cur = out;
      first_axis = false;
    }
    const int blocks = grid_for(n, 256, 512);
    double* part = sw_part.get() + static_cast<long long>(first) * blocks * 2;
    hipLaunchKernelGGL(dev::shrinkwrap_level, dim3(blocks, count), dim3(256), 0, s, cur, supp.get() + off, n, method,
                       part);
    hipLaunchKernelGGL(dev::shrinkwrap_threshold, dim3(grid_for(n, 256, 2048), count), dim3(256), 0, s, cur, part,
                       blocks, n, method, threshold, supp.get() + off);
    QPCH_HIP(hipGetLastError());
  }

  void shrinkwrap(float threshold, float sigma, int method) {
    if (!(threshold > 0.0f) || !(sigma >= 0.0f)) throw InvalidArgument("Shrinkwrap needs threshold > 0, sigma >= 0");
    shrinkwrap_prepare(sigma);
    shrinkwrap_range(threshold, sigma, method, 0, starts, low_active, stream);
  }

  // --- detwin over a range of starts (save, cut at the centre of mass; restore) -------------
  void detwin_prepare() {
    if (detwin_sums.size() < static_cast<size_t>(2 * starts)) detwin_sums.resize(2 * starts);
  }

  void half_support_range(int axis, int first, int count, hipStream_t s) {
    if (axis < 0 || axis >= dims.nd) throw InvalidArgument("detwin axis out of range");
    const long long inner = dims.inner(axis);
    const int len = dims.d[axis];
    const long long off = n * first;
    const dim3 grid(static_cast<unsigned>(std::min<long long>((n + 255) / 256, 256)), count);
    hipLaunchKernelGGL(dev::fill_u64, dim3(1), dim3(256), 0, s, detwin_sums.get() + 2 * first, 2 * count);
    hipLaunchKernelGGL(dev::support_moments, grid, dim3(256), 0, s, supp.get() + off, n, inner, len,
                       detwin_sums.get() + 2 * first);
    hipLaunchKernelGGL(dev::support_cut, grid, dim3(256), 0, s, supp.get() + off, n, inner, len,
                       detwin_sums.get() + 2 * first);
    QPCH_HIP(hipGetLastError());
  }

  void copy_support_range(bool save, int first, int count, hipStream_t s) {
    const long long off = n * first;
    std::uint8_t* dst = save ? saved.get() + off : supp.get() + off;
    const std::uint8_t* src = save ? supp.get() + off : saved.get() + off;
    QPCH_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(n * count), hipMemcpyDeviceToDevice, s));
  }

  // --- a whole recipe for one launch group (qpch_group_run) ---------------------------------
  // Runs `ops` for the starts of launch group g on stream s: precision switches, prologue,
  // iterations, Shrinkwrap and detwin, exactly the kernels the step-by-step entry points
  // launch for those starts.  `lowp` and `valid` are the group's storage precision and wave
  // validity, updated as the recipe goes.  Launches only: safe inside a capture, once
  // shrinkwrap_prepare / detwin_prepare have run.
  void sequence(int g, const qpch_op* ops, int count, hipStream_t s, bool& lowp, bool& valid) {
    const int first = g * batch;
    hipStream_t saved_on = on;
    on = s;
    try {
      for (int k = 0; k < count; ++k) {
        const qpch_op& op = ops[k];
        switch (op.kind) {
          case QPCH_OP_ITERATE: {
            const bool want = low() && !(mixed && op.refine);
            if (want != lowp) {
              convert_range(want, first, batch, s);
              lowp = want;
              valid = false;
            }
            if (!want) {
              if (!valid) fused_prologue<float2>(g);
              fused_iterations<float2>(g, op.iterations, op.a, op.b);
            } else if (storage == QPCH_F16) {
              if (!valid) fused_prologue<__half2>(g);
              fused_iterations<__half2>(g, op.iterations, op.a, op.b);
            } else {
              if (!valid) fused_prologue<__hip_bfloat162>(g);
              fused_iterations<__hip_bfloat162>(g, op.iterations, op.a, op.b);
            }
            valid = true;
            break;
          }
          case QPCH_OP_SHRINKWRAP:
            shrinkwrap_range(op.threshold, op.sigma, op.method, first, batch, lowp, s);
            break;
          case QPCH_OP_DETWIN_BEGIN:
            copy_support_range(true, first, batch, s);
            half_support_range(op.axis, first, batch, s);
            break;
          case QPCH_OP_DETWIN_END:
            copy_support_range(false, first, batch, s);
            break;
          default:
            throw InvalidArgument("unknown operation kind");
        }
      }
    } catch (...) {
      on = saved_on;
      throw;
    }
    on = saved_on;
  }

  // --- set-up ------------------------------------------------------------------------------
  void init(const qpch_problem& p) {
    dims = make_dims(p.ndim, p.shape);
    n = dims.n;
    if (p.starts < 1) throw InvalidArgument("starts must be >= 1");
    problems = p.problems > 1 ? p.problems : 1;
    per_problem = p.starts;
    starts = per_problem * problems;  // every problem's starts, problem-major
    batch = p.batch ? p.batch : starts;
    if (batch < 1 || starts % batch != 0) throw InvalidArgument("batch must divide starts");
    groups = starts / batch;
    storage = p.storage;
    if (storage != QPCH_F32 && storage != QPCH_F16 && storage != QPCH_BF16) throw InvalidArgument("unknown storage");
    mixed = p.mixed != 0 && storage != QPCH_F32;
    schedule = p.schedule;
    if (schedule < QPCH_UNFUSED || schedule > QPCH_GRAPH) throw InvalidArgument("unknown schedule");
    // QPCH_EAGER=1 launches the graph schedule's kernels one by one (for kernel tracers
    // that do not follow graph launches); the kernels and results are unchanged.
    if (schedule == QPCH_GRAPH && eager_launches()) schedule = QPCH_FUSED;
    pow2 = fusable(dims);
    fused = (pow2 || (storage == QPCH_F32 && generic_fusable(dims))) && schedule != QPCH_UNFUSED;
    plane3d = fused && dims.nd == 3 && plane_supported(dims.d[1], dims.d[2]) && p.plane != 2;
    resident = fused && dims.nd == 2 && image_supported(dims.d[0], dims.d[1]) && p.plane != 2;
    if (p.xcd != 0 && p.xcd != 2 && p.xcd != 3) throw InvalidArgument("xcd must be 0, 2 or 3");
    xcd_mode = p.xcd;
    if (p.radix != 0 && p.radix != 8 && p.radix != 16) throw InvalidArgument("radix must be 0, 8 or 16");
    radix = p.radix == 8 ? 8 : 16;
    if (storage != QPCH_F32 && !fusable(dims)) {
      throw InvalidArgument("fp16/bf16 storage needs power-of-two extents between 2 and 4096");
    }
    if (storage != QPCH_F32 && schedule == QPCH_UNFUSED) throw InvalidArgument("fp16/bf16 storage needs a fused schedule");
    if (problems > 1 && !fused) throw InvalidArgument("several problems in one solver need the fused schedule");
    for (int a = 0; a < dims.nd; ++a) {
      if (!(is_pow2(dims.d[a]) && dims.d[a] <= kMaxFused) && dims.d[a] > dev::kGenericMaxN) {
        throw InvalidArgument("axis length " + std::to_string(dims.d[a]) + " is not supported");
      }
    }
    if (!p.amplitude || !p.mask || !p.support) throw InvalidArgument("amplitude, mask and support are required");
    // Bluestein tables of long prime axes now, so that no graph capture ever makes them
    for (int a = 0; a < dims.nd; ++a) {
      if (needs_bluestein(dims.d[a]) || chirp_axis(dims.d[a])) (void)bluestein_tables(dims.d[a], false);
    }
    if (p.lanes < 0 || p.lanes > 8) throw InvalidArgument("lanes must be between 0 and 8");
    if (p.plane < 0 || p.plane > 2) throw InvalidArgument("plane must be 0, 1 or 2");
    if (p.persist < 0 || p.persist > 2) throw InvalidArgument("persist must be 0, 1 or 2");
    QPCH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    on = stream;
    lanes = std::max(1, p.lanes);
    if (fused && lanes > 1 && groups > 1) {
      for (int l = 0; l < std::min(lanes, groups); ++l) {
        hipStream_t st{};
        QPCH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        lane_streams.push_back(st);
        hipEvent_t e{};
        QPCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        lane_done.push_back(e);
      }
      QPCH_HIP(hipEventCreateWithFlags(&lane_fork, hipEventDisableTiming));
    }

    const long long np = n * problems;
    std::vector<float> amp(p.amplitude, p.amplitude + np);
    amp_scales.assign(problems, 1.0f);
    energies.assign(problems, 0.0);
    std::vector<float> signed_amp(np);
    for (int q = 0; q < problems; ++q) {
      float* a = amp.data() + q * n;
      if (storage != QPCH_F32) {
        // keep the unnormalised spectra inside fp16 range: peak amplitude 1 per problem
        const float peak = *std::max_element(a, a + n);
        amp_scales[q] = peak > 0.0f ? peak : 1.0f;
        for (long long i = 0; i < n; ++i) a[i] /= amp_scales[q];
      }
      double e = 0.0;
      for (long long i = 0; i < n; ++i) {
        const bool masked = p.mask[q * n + i] != 0;
        signed_amp[q * n + i] = masked ? -1.0f : a[i];
        if (!masked) e += static_cast<double>(a[i]) * a[i];
      }
      energies[q] = e;
    }
    energy = energies[0];
    amp_scale = amp_scales[0];
    amp_true = upload(amp.data(), np, stream);
    amp_signed = upload(signed_amp.data(), np, stream);
    std::vector<std::uint8_t> sup(static_cast<size_t>(n) * starts);
    for (int s = 0; s < starts; ++s) {
      const std::uint8_t* src = p.support + (s / per_problem) * n;
      for (long long i = 0; i < n; ++i) sup[s * n + i] = src[i] != 0;
    }
    supp = upload(sup.data(), sup.size(), stream);
    saved.resize(sup.size());
    x32.resize(static_cast<size_t>(n) * starts);
    w32.resize(static_cast<size_t>(n) * starts);
    if (storage != QPCH_F32) {
      x16.resize(static_cast<size_t>(n) * starts);
      w16.resize(static_cast<size_t>(n) * starts);
    }
    counters.resize(2 * static_cast<size_t>(groups));  // slot counters, then image_iterations' tickets
    persist_xcds = twopass_xcds();
    persist = fused && dims.nd == 2 && !resident && p.persist != 2 && persist_xcds > 0;
    persist_generic = persist && p.persist == 1;
    if (persist) {
      int dev = 0, khz = 0;
      QPCH_HIP(hipGetDevice(&dev));
      QPCH_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      wait_ticks = static_cast<long long>(std::max(khz, 1000)) * 500;  // 0.5 s
      // Test hooks for the safety nets (tests/test_gpu_persistent.py): a wait bound in
      // clock ticks, and fewer work-groups than XCDs (queues left with items).  Both end in
      // the abort flag and a DeviceError, never in a hang.
      if (const char* v = std::getenv("QPCH_TEST_WAIT_TICKS")) wait_ticks = std::max(0LL, std::atoll(v));
      if (const char* v = std::getenv("QPCH_TEST_PERSIST_WORKERS")) persist_workers = std::max(0, std::atoi(v));
      if (generic_axis(dims.d[0]) && generic_axis(dims.d[1])) {
        const dev::GenericPlan gp[2] = {generic_plan(dims.d[0]), generic_plan(dims.d[1])};
        generic_plans = upload(gp, 2, stream);
      }
      const long long cnt = dev::kLineInts + static_cast<long long>(groups) * psync_group();
      psync.resize(static_cast<size_t>(cnt));
      hipLaunchKernelGGL(dev::fill_int, dim3(grid_for(cnt)), dim3(256), 0, stream, psync.get(), static_cast<int>(cnt), 0);
      QPCH_HIP(hipGetLastError());
    }
    // partial sums of the error: one per block of the modulus pass of each start
    int pb = grid_for(n, 256, 1024);
    if (fused) pb = std::max({pb, axis0_partials<8>(dims, radix), axis0_partials<4>(dims, radix)});
    partials.resize(static_cast<size_t>(pb) * starts);  // groups may run at once: rows of their own

    // starting objects
    if (p.start) {
      std::vector<float> host(p.start, p.start + 2 * n * starts);
      for (int s = 0; s < starts; ++s) {
        const float sc = amp_scales[s / per_problem];
        if (sc != 1.0f) for (long long i = 0; i < 2 * n; ++i) host[2 * n * s + i] /= sc;
      }
      copy_sync(w32.get(), host.data(), sizeof(float2) * n * starts, hipMemcpyHostToDevice, stream);
    } else {
      for (int g = 0; g < groups; ++g) {
        float2* w = w32.get() + static_cast<long long>(g) * batch * n;
        hipLaunchKernelGGL(dev::random_wave, dim3(grid_for(n, 256, 2048), batch), dim3(256), 0, stream, w,
                           amp_true.get(), n, p.seed, static_cast<std::uint64_t>(g) * batch, per_problem);
        QPCH_HIP(hipGetLastError());
        device_fft(w, dims, batch, false, static_cast<float>(1.0 / std::sqrt(static_cast<double>(n))), stream);
      }
    }
    hipLaunchKernelGGL((dev::convert<float2, float2>), dim3(grid_for(n * starts)), dim3(256), 0, stream,
                       w32.get(), x32.get(), n * starts, 1.0f, supp.get());
    QPCH_HIP(hipGetLastError());
    low_active = false;
    wave_valid = false;
    QPCH_HIP(hipStreamSynchronize(stream));
  }

  // Average duration of each pass of one iteration (group 0), each replayed `reps` times
  // between two events on the solver's stream.  Leaves the iterate changed.
  int time_passes(bool lowp, int reps, float* ms, int cap) {
    if (!fused) throw InvalidArgument("pass timing needs the fused schedule");
    if (lowp && !low()) throw InvalidArgument("this solver has no reduced-precision storage");
    if (err_cap <= 0) begin_run(1);
    set_precision(lowp);
    if (!wave_valid) {
      if (!lowp) fused_prologue<float2>(0);
      else if (storage == QPCH_F16) fused_prologue<__half2>(0);
      else fused_prologue<__hip_bfloat162>(0);
    }
    const auto passes = iteration_passes(lowp);
    hipEvent_t e0{}, e1{};
    QPCH_HIP(hipEventCreate(&e0));
    QPCH_HIP(hipEventCreate(&e1));
    int k = 0;
    for (const auto& ps : passes) {
      if (k >= cap) break;
      QPCH_HIP(hipEventRecord(e0, stream));
      for (int r = 0; r < reps; ++r) {
        if (!lowp) run_pass<float2>(ps, 0, batch, 0.9f, -0.8f);
        else if (storage == QPCH_F16) run_pass<__half2>(ps, 0, batch, 0.9f, -0.8f);
        else run_pass<__hip_bfloat162>(ps, 0, batch, 0.9f, -0.8f);
      }
      QPCH_HIP(hipEventRecord(e1, stream));
      QPCH_HIP(hipEventSynchronize(e1));
      float t = 0.0f;
      QPCH_HIP(hipEventElapsedTime(&t, e0, e1));
      ms[k++] = t / static_cast<float>(reps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    wave_valid = false;
    return k;
  }

  // Average time of one iteration of every launch group on the production path: chunks
  // of `len` iterations replayed from their hipGraphs (graph schedule) or launched eagerly
  // (fused schedule), `reps` chunks after one warm-up chunk, between two events on the
  // solver's stream.  Restarts the run (its error history) and changes the iterate.
  float time_iterations(bool lowp, int len, int reps) {
    if (!fused) throw InvalidArgument("iteration timing needs a fused schedule");
    if (lowp && !low()) throw InvalidArgument("this solver has no reduced-precision storage");
    // the precision iterate() picks is low() && !(mixed && refine)
    const bool refine = !lowp;
    if (!lowp && low() && !mixed) throw InvalidArgument("a reduced-precision solver without mixed runs no fp32 iterations");
    begin_run(len * (reps + 1));
    iterate(len, 0.9f, -0.8f, refine);
    hipEvent_t e0{}, e1{};
    QPCH_HIP(hipEventCreate(&e0));
    QPCH_HIP(hipEventCreate(&e1));
    QPCH_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; ++r) iterate(len, 0.9f, -0.8f, refine);
    QPCH_HIP(hipEventRecord(e1, stream));
    QPCH_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    QPCH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms * 1e3f / static_cast<float>(reps * len);
  }

  // A persistent launch whose wait ran past its bound, or that left a queue with items (an
  // XCD without work-groups), sets the abort flag: its results are not valid, and the
  // solver reports a device failure from then on.
  void check_persistent() {
    if (!persist) return;
    int flag = 0;
    copy_sync(&flag, psync.get(), sizeof(int), hipMemcpyDeviceToHost, stream);
    if (flag != 0) {
      throw DeviceError("a persistent iteration launch did not complete its work items (a wait ran past its bound "
                        "or an XCD's queue was left with items); its results are not valid");
    }
  }

  void read_state(float* host) {
    if (low_active) convert_x(false);
    QPCH_HIP(hipStreamSynchronize(stream));
    if (low_active) {
      // x32 now mirrors the low-precision authority; it stays valid until the next
      // low-precision iteration converts back.
    }
    std::vector<float> tmp;
    copy_sync(host, x32.get(), sizeof(float2) * n * starts, hipMemcpyDeviceToHost, stream);
    for (int s = 0; s < starts; ++s) {
      const float sc = amp_scales[s / per_problem];
      if (sc != 1.0f) for (long long i = 0; i < 2 * n; ++i) host[2 * n * s + i] *= sc;
    }
  }
};

}  // namespace qpch

struct qpch_solver {
  qpch::SolverImpl impl;
};

// Several solvers advanced by one graph launch: each solver's passes are one branch of the
// graph, forked from and joined back to the group's own stream.
struct qpch_group {
  std::vector<qpch_solver*> solvers;
  hipStream_t main = nullptr;
  std::vector<hipEvent_t> done;  // one per solver
  hipEvent_t fork = nullptr, joined = nullptr;
  struct Entry {
    hipGraphExec_t exec;
    std::vector<int> generations;
  };
  std::map<std::tuple<int, float, float, int>, Entry> graphs;

  // whole-recipe graphs (qpch_group_run), keyed by the recipe and the solvers' starting
  // precision state
  struct RunEntry {
    hipGraphExec_t exec;
    std::vector<int> generations;
    std::vector<std::pair<bool, bool>> final_state;  // per solver: (low_active, wave_valid)
  };
  std::map<std::string, RunEntry> runs;
  std::vector<hipEvent_t> branch_done;  // one per branch stream of a run

  ~qpch_group() {
    if (main) (void)hipStreamSynchronize(main);
    for (auto& kv : runs) (void)hipGraphExecDestroy(kv.second.exec);
    for (hipEvent_t e : branch_done) (void)hipEventDestroy(e);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
    for (hipEvent_t e : done) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (joined) (void)hipEventDestroy(joined);
    if (main) (void)hipStreamDestroy(main);
  }

  // an event recorded on the group stream now (it orders nothing it did not already)
  hipEvent_t fork_ready() {
    QPCH_HIP(hipEventRecord(fork, main));
    return fork;
  }

  // The streams the branches of solver k run on: its lanes when it has them, else its own
  // stream (launch group g goes to lane g mod L).
  static std::vector<hipStream_t> branch_streams(qpch::SolverImpl& m) {
    if (m.lane_streams.size() > 1) return m.lane_streams;
    return {m.stream};
  }

  void run(const qpch_op* ops, int count) {
    if (count < 0 || (count > 0 && !ops)) throw qpch::InvalidArgument("bad recipe");
    long long iterations = 0;
    bool detwin = false;
    for (int k = 0; k < count; ++k) {
      const qpch_op& op = ops[k];
      if (op.kind == QPCH_OP_ITERATE) {
        if (op.iterations < 1) throw qpch::InvalidArgument("iterations must be >= 1");
        iterations += op.iterations;
      } else if (op.kind == QPCH_OP_SHRINKWRAP) {
        if (!(op.threshold > 0.0f) || !(op.sigma >= 0.0f)) {
          throw qpch::InvalidArgument("Shrinkwrap needs threshold > 0, sigma >= 0");
        }
        if (op.method != 0 && op.method != 1) throw qpch::InvalidArgument("Shrinkwrap method must be 0 or 1");
      } else if (op.kind == QPCH_OP_DETWIN_BEGIN || op.kind == QPCH_OP_DETWIN_END) {
        detwin = true;
      } else {
        throw qpch::InvalidArgument("unknown operation kind");
      }
    }
    if (count == 0) return;
    // buffers and tables first: nothing inside the capture may allocate or upload
    std::string key(reinterpret_cast<const char*>(ops), sizeof(qpch_op) * static_cast<size_t>(count));
    std::vector<int> gens;
    for (auto* sv : solvers) {
      auto& m = sv->impl;
      if (m.err_cap < iterations) throw qpch::InvalidArgument("qpch_solver_begin_run capacity is below the recipe's iterations");
      for (int k = 0; k < count; ++k) {
        if (ops[k].kind == QPCH_OP_SHRINKWRAP) m.shrinkwrap_prepare(ops[k].sigma);
        if (ops[k].kind == QPCH_OP_DETWIN_BEGIN && (ops[k].axis < 0 || ops[k].axis >= m.dims.nd)) {
          throw qpch::InvalidArgument("detwin axis out of range");
        }
      }
      if (detwin) m.detwin_prepare();
      key += m.low_active ? 'L' : 'F';
      key += m.wave_valid ? 'v' : 'i';
      gens.push_back(m.generation);
    }
    // the run starts after every solver's queued work (begin_run's resets among it)
    for (size_t k = 0; k < solvers.size(); ++k) {
      QPCH_HIP(hipEventRecord(done[k], solvers[k]->impl.stream));
      QPCH_HIP(hipStreamWaitEvent(main, done[k], 0));
    }
    auto it = runs.find(key);
    if (it != runs.end() && it->second.generations != gens) {
      (void)hipGraphExecDestroy(it->second.exec);
      runs.erase(it);
      it = runs.end();
    }
    const bool eager = qpch::eager_launches();
    std::vector<std::pair<bool, bool>> final_state;
    if (it == runs.end() || eager) {
      size_t branches = 0;
      for (auto* sv : solvers) branches += branch_streams(sv->impl).size();
      while (branch_done.size() < branches) {
        hipEvent_t e{};
        QPCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        branch_done.push_back(e);
      }
      if (!eager) QPCH_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
      hipGraph_t graph{};
      try {
        QPCH_HIP(hipEventRecord(fork, main));
        size_t b = 0;
        for (auto* sv : solvers) {
          auto& m = sv->impl;
          const auto streams = branch_streams(m);
          for (hipStream_t st : streams) QPCH_HIP(hipStreamWaitEvent(st, fork, 0));
          bool lowp = m.low_active, valid = m.wave_valid;
          for (int g = 0; g < m.groups; ++g) {
            bool l = m.low_active, v = m.wave_valid;  // every group starts from the solver's state
            m.sequence(g, ops, count, streams[static_cast<size_t>(g) % streams.size()], l, v);
            lowp = l;
            valid = v;
          }
          final_state.emplace_back(lowp, valid);
          for (hipStream_t st : streams) {
            QPCH_HIP(hipEventRecord(branch_done[b], st));
            QPCH_HIP(hipStreamWaitEvent(main, branch_done[b], 0));
            ++b;
          }
        }
      } catch (...) {
        if (!eager) {
          (void)hipStreamEndCapture(main, &graph);
          if (graph) (void)hipGraphDestroy(graph);
        }
        throw;
      }
      if (!eager) {
        QPCH_HIP(hipStreamEndCapture(main, &graph));
        hipGraphExec_t exec{};
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        QPCH_HIP(e);
        it = runs.emplace(key, RunEntry{exec, gens, final_state}).first;
      }
    }
    if (!eager) {
      QPCH_HIP(hipGraphLaunch(it->second.exec, main));
      final_state = it->second.final_state;
    }
    QPCH_HIP(hipEventRecord(joined, main));
    for (size_t k = 0; k < solvers.size(); ++k) {
      auto& m = solvers[k]->impl;
      QPCH_HIP(hipStreamWaitEvent(m.stream, joined, 0));
      m.low_active = final_state[k].first;
      m.wave_valid = final_state[k].second;
    }
  }

  void iterate(int len, float a, float b, bool refine) {
    if (len < 1) throw qpch::InvalidArgument("iterations must be >= 1");
    int lowmask = 0;
    for (size_t k = 0; k < solvers.size(); ++k) {
      if (solvers[k]->impl.prepare(refine)) lowmask |= 1 << (k % 30);
      QPCH_HIP(hipEventRecord(done[k], solvers[k]->impl.stream));
      QPCH_HIP(hipStreamWaitEvent(main, done[k], 0));
    }
    if (qpch::eager_launches()) {
      for (size_t k = 0; k < solvers.size(); ++k) {
        auto& m = solvers[k]->impl;
        QPCH_HIP(hipStreamWaitEvent(m.stream, fork_ready(), 0));
        m.enqueue(len, a, b, (lowmask >> (k % 30)) & 1);
      }
      return;
    }
    std::vector<int> gens;
    for (auto* s : solvers) gens.push_back(s->impl.generation);
    const auto key = std::make_tuple(len, a, b, lowmask);
    auto it = graphs.find(key);
    if (it != graphs.end() && it->second.generations != gens) {
      (void)hipGraphExecDestroy(it->second.exec);
      graphs.erase(it);
      it = graphs.end();
    }
    if (it == graphs.end()) {
      QPCH_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
      hipGraph_t graph{};
      try {
        QPCH_HIP(hipEventRecord(fork, main));
        for (size_t k = 0; k < solvers.size(); ++k) {
          auto& m = solvers[k]->impl;
          QPCH_HIP(hipStreamWaitEvent(m.stream, fork, 0));
          m.enqueue(len, a, b, (lowmask >> (k % 30)) & 1);
          QPCH_HIP(hipEventRecord(done[k], m.stream));
          QPCH_HIP(hipStreamWaitEvent(main, done[k], 0));
        }
      } catch (...) {
        (void)hipStreamEndCapture(main, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      QPCH_HIP(hipStreamEndCapture(main, &graph));
      hipGraphExec_t exec{};
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      QPCH_HIP(e);
      it = graphs.emplace(key, Entry{exec, gens}).first;
    }
    QPCH_HIP(hipGraphLaunch(it->second.exec, main));
    QPCH_HIP(hipEventRecord(joined, main));
    for (auto* s : solvers) {
      QPCH_HIP(hipStreamWaitEvent(s->impl.stream, joined, 0));
      s->impl.wave_valid = true;
    }
  }
};

// Device scratch of qpch_amplitude_loss_noalloc: the wave and loss partials of one call.
// Every GPU the library runs on holds its own copy from module load until the process ends;
// all calls on one GPU share it.  Host code never passes this name to a kernel: it is a
// host-side stand-in, and the device address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t{32} << 20;  // 32 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace qpch {

// --- the amplitude loss and its gradient ---------------------------------------------------
// Scratch of one call: the wave (the batch's transforms) and the FP64 loss partials, carved
// from one block at 256-byte aligned offsets.
struct LossPlan {
  bool fused = false;  // fp32 with power-of-two extents: the fused line passes
  int parts = 0;       // loss partials per object
  size_t wave_bytes = 0, part_offset = 0, bytes = 0;
};

struct LossCall {
  const void* x;
  const void* amplitude;
  long long amplitude_stride;
  void* grad;
  double* loss;
  Dims d;
  int batch, precision;
  hipStream_t st;
  LossPlan plan;
};

size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

LossCall loss_call(const void* x, const void* amplitude, long long amplitude_stride, void* grad, double* loss,
                   int ndim, const size_t* shape, int batch, int precision, void* stream) {
  LossCall c{x, amplitude, amplitude_stride, grad, loss, make_dims(ndim, shape), batch, precision,
             static_cast<hipStream_t>(stream), {}};
  const Dims& d = c.d;
  if (batch < 1) throw InvalidArgument("batch must be >= 1");
  if (precision != 0 && precision != 1) throw InvalidArgument("precision must be 0 (fp32) or 1 (fp64)");
  if (!x || !amplitude || !grad || !loss) throw InvalidArgument("null pointer");
  if (amplitude_stride != 0 && amplitude_stride != d.n) {
    throw InvalidArgument("amplitude_stride must be 0 (shared) or the element count (one per object)");
  }
  LossPlan& pl = c.plan;
  pl.fused = precision == 0 && fusable(d);
  if (pl.fused) {
    pl.parts = pass_partials<8>(d.d[0], d.inner(0) > 1, static_cast<int>(d.outer(0)), static_cast<int>(d.inner(0)));
  } else {
    pl.parts = grid_for(d.n, 256, 512);
  }
  pl.wave_bytes = (precision == 0 ? sizeof(float2) : sizeof(double2)) * static_cast<size_t>(d.n) * batch;
  pl.part_offset = round256(pl.wave_bytes);
  pl.bytes = pl.part_offset + round256(sizeof(double) * static_cast<size_t>(pl.parts) * batch);
  return c;
}

// Runs the loss and gradient of a validated call with its scratch at `base` (plan.bytes,
// 256-byte aligned).  Every scratch element is written before it is read.
void amplitude_loss_run(const LossCall& c, void* base) {
  const Dims& d = c.d;
  const int batch = c.batch;
  const hipStream_t st = c.st;
  const long long total = d.n * batch;
  const double s = 1.0 / std::sqrt(static_cast<double>(d.n));
  auto* partials = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + c.plan.part_offset);
  const int parts = c.plan.parts;
  if (c.plan.fused) {
    auto* w = static_cast<float2*>(base);
    auto pass = [&](int axis, dev::Mode mode, const void* src, float in_scale, float out_scale) {
      dev::PassArgs p{};
      p.outer = static_cast<int>(d.outer(axis));
      p.inner = static_cast<int>(d.inner(axis));
      p.n = d.n;
      p.w = w;
      p.src = src;
      p.x = const_cast<void*>(c.x);
      p.out = c.grad;
      p.amp = static_cast<const float*>(c.amplitude);
      p.amp_stride = c.amplitude_stride;
      p.partials = partials;
      p.partial_stride = parts;
      p.in_scale = in_scale;
      p.out_scale = out_scale;
      launch_pass<float2>(d.d[axis], p.inner > 1, mode, p, batch, st);
    };
    const int last = d.nd - 1;
    pass(last, dev::Mode::kFwd, c.x, 1.0f, 1.0f);
    if (d.nd == 3) pass(1, dev::Mode::kFwd, nullptr, 1.0f, 1.0f);
    pass(0, dev::Mode::kModulus, nullptr, static_cast<float>(s), static_cast<float>(s));
    if (d.nd == 3) pass(1, dev::Mode::kInv, nullptr, 1.0f, 1.0f);
    pass(last, dev::Mode::kGrad, nullptr, 1.0f, 1.0f);
  } else if (c.precision == 0) {
    auto* w = static_cast<float2*>(base);
    QPCH_HIP(hipMemcpyAsync(w, c.x, sizeof(float2) * total, hipMemcpyDeviceToDevice, st));
    device_fft(w, d, batch, true, 1.0f, st);
    hipLaunchKernelGGL(dev::loss_project<float>, dim3(parts, batch), dim3(256), 0, st, w,
                       static_cast<const float*>(c.amplitude), c.amplitude_stride, d.n, static_cast<float>(s),
                       partials);
    device_fft(w, d, batch, false, 1.0f, st);
    hipLaunchKernelGGL(dev::grad_from<float>, dim3(grid_for(total)), dim3(256), 0, st,
                       static_cast<const float2*>(c.x), w, static_cast<float2*>(c.grad), total, 1.0f);
  } else {
    auto* w = static_cast<double2*>(base);
    QPCH_HIP(hipMemcpyAsync(w, c.x, sizeof(double2) * total, hipMemcpyDeviceToDevice, st));
    device_fft_f64(w, d, batch, true, st);
    hipLaunchKernelGGL(dev::loss_project<double>, dim3(parts, batch), dim3(256), 0, st, w,
                       static_cast<const double*>(c.amplitude), c.amplitude_stride, d.n, s, partials);
    device_fft_f64(w, d, batch, false, st);
    hipLaunchKernelGGL(dev::grad_from<double>, dim3(grid_for(total)), dim3(256), 0, st,
                       static_cast<const double2*>(c.x), w, static_cast<double2*>(c.grad), total, 1.0);
  }
  QPCH_HIP(hipGetLastError());
  hipLaunchKernelGGL(dev::sum_partials, dim3((batch + 63) / 64), dim3(64), 0, st, partials, parts, c.loss, batch);
  QPCH_HIP(hipGetLastError());
}

// The current device's copy of `scratch`, looked up once per device with that device
// current (each GPU holds the array at its own address).
void* noalloc_scratch() {
  static std::mutex mu;
  static std::map<int, void*> address;  // device -> its copy
  int dev = 0;
  QPCH_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = address.find(dev);
  if (it != address.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    throw DeviceError(std::string("the no-allocation scratch has no device address: ") + hipGetErrorString(e));
  }
  address.emplace(dev, ptr);
  return ptr;
}

}  // namespace qpch

using qpch::guarded;

extern "C" {

const char* qpch_last_error(void) { return qpch::g_error.c_str(); }

int qpch_device_count(int* count) {
  return guarded([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
  });
}

int qpch_device_info(qpch_device_info_t* info) {
  return guarded([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) throw qpch::DeviceError("no HIP device available");
    int dev = 0;
    QPCH_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    QPCH_HIP(hipGetDeviceProperties(&prop, dev));
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s", prop.name);
    if (info->name[0] == '\0') (void)hipDeviceGetName(info->name, sizeof(info->name), dev);
    if (info->name[0] == '\0') {
      std::string arch(prop.gcnArchName);
      std::snprintf(info->name, sizeof(info->name), "AMD GPU %s", arch.substr(0, arch.find(':')).c_str());
    }
    std::snprintf(info->arch, sizeof(info->arch), "%s", prop.gcnArchName);
    info->compute_units = prop.multiProcessorCount;
    info->memory_bytes = prop.totalGlobalMem;
    info->l2_bytes = static_cast<size_t>(prop.l2CacheSize);
    int clock_khz = 0;
    QPCH_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeMemoryClockRate, dev));
    // 2 x memory clock x bus width (the reference's formula, src/cuda/operators.cu:18-36)
    // undercounts HBM3E: on gfx950 the runtime's clock and width give 4096 GB/s, half of
    // the 8 TB/s AMD publishes for MI355X (and that the fused passes' measured traffic
    // already exceeds).  Known devices report their published peak.
    const std::string arch(prop.gcnArchName);
    if (arch.rfind("gfx950", 0) == 0) {
      info->peak_bandwidth_gbs = 8000.0;
    } else {
      info->peak_bandwidth_gbs = 2.0 * clock_khz * 1e3 * (prop.memoryBusWidth / 8.0) / 1e9;
    }
  });
}

int qpch_solver_create(const qpch_problem* problem, qpch_solver** out) {
  *out = nullptr;
  return guarded([&] {
    if (!problem) throw qpch::InvalidArgument("null problem");
    auto s = std::make_unique<qpch_solver>();
    s->impl.init(*problem);
    *out = s.release();
  });
}

void qpch_solver_destroy(qpch_solver* solver) {
  if (!solver) return;
  if (solver->impl.stream) (void)hipStreamSynchronize(solver->impl.stream);
  delete solver;
}

int qpch_solver_begin_run(qpch_solver* s, int capacity) {
  return guarded([&] { s->impl.begin_run(capacity); });
}

int qpch_solver_iterate(qpch_solver* s, int iterations, float a, float b, int refine) {
  return guarded([&] { s->impl.iterate(iterations, a, b, refine != 0); });
}

int qpch_solver_shrinkwrap(qpch_solver* s, float threshold, float sigma, int method) {
  return guarded([&] { s->impl.shrinkwrap(threshold, sigma, method); });
}

int qpch_solver_get_support(qpch_solver* s, uint8_t* host) {
  return guarded([&] {
    auto& m = s->impl;
    qpch::copy_sync(host, m.supp.get(), m.supp.size(), hipMemcpyDeviceToHost, m.stream);
  });
}

int qpch_solver_set_support(qpch_solver* s, const uint8_t* host) {
  return guarded([&] {
    auto& m = s->impl;
    qpch::copy_sync(m.supp.get(), host, m.supp.size(), hipMemcpyHostToDevice, m.stream);
  });
}

int qpch_solver_save_support(qpch_solver* s) {
  return guarded([&] {
    auto& m = s->impl;
    m.copy_support_range(true, 0, m.starts, m.stream);
  });
}

int qpch_solver_half_support(qpch_solver* s, int axis) {
  return guarded([&] {
    auto& m = s->impl;
    m.detwin_prepare();
    m.half_support_range(axis, 0, m.starts, m.stream);
  });
}

int qpch_solver_restore_support(qpch_solver* s) {
  return guarded([&] {
    auto& m = s->impl;
    m.copy_support_range(false, 0, m.starts, m.stream);
  });
}

int qpch_solver_read_errors(qpch_solver* s, int first, int count, double* host) {
  return guarded([&] {
    auto& m = s->impl;
    if (first < 0 || count < 0 || first + count > m.err_cap) throw qpch::InvalidArgument("error slots out of range");
    QPCH_HIP(hipStreamSynchronize(m.stream));
    m.check_persistent();
    qpch::copy_sync(host, m.errs.get() + static_cast<size_t>(first) * m.starts, sizeof(double) * count * m.starts,
                    hipMemcpyDeviceToHost, m.stream);
  });
}

int qpch_solver_read_state(qpch_solver* s, float* host) {
  return guarded([&] { s->impl.read_state(host); });
}

double qpch_solver_energy(const qpch_solver* s) { return s->impl.energy; }

int qpch_solver_energies(const qpch_solver* s, double* out) {
  return guarded([&] {
    for (int q = 0; q < s->impl.problems; ++q) out[q] = s->impl.energies[q];
  });
}

int qpch_group_create(qpch_solver* const* solvers, int count, qpch_group** out) {
  *out = nullptr;
  return guarded([&] {
    if (count < 1) throw qpch::InvalidArgument("a group needs at least one solver");
    auto g = std::make_unique<qpch_group>();
    for (int k = 0; k < count; ++k) {
      if (!solvers[k]) throw qpch::InvalidArgument("null solver");
      if (!solvers[k]->impl.fused) throw qpch::InvalidArgument("grouped solvers need the fused schedule");
      g->solvers.push_back(solvers[k]);
      hipEvent_t e{};
      QPCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      g->done.push_back(e);
    }
    QPCH_HIP(hipStreamCreateWithFlags(&g->main, hipStreamNonBlocking));
    QPCH_HIP(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    QPCH_HIP(hipEventCreateWithFlags(&g->joined, hipEventDisableTiming));
    *out = g.release();
  });
}

int qpch_group_iterate(qpch_group* g, int iterations, float a, float b, int refine) {
  return guarded([&] { g->iterate(iterations, a, b, refine != 0); });
}

int qpch_group_run(qpch_group* g, const qpch_op* ops, int count) {
  return guarded([&] { g->run(ops, count); });
}

int qpch_group_synchronize(qpch_group* g) {
  return guarded([&] {
    QPCH_HIP(hipStreamSynchronize(g->main));
    for (auto* s : g->solvers) QPCH_HIP(hipStreamSynchronize(s->impl.stream));
  });
}

void qpch_group_destroy(qpch_group* g) {
  if (!g) return;
  (void)hipStreamSynchronize(g->main);
  delete g;
}

int qpch_solver_time_passes(qpch_solver* s, int low, int reps, float* ms, int cap, int* count) {
  return guarded([&] {
    if (reps < 1) throw qpch::InvalidArgument("reps must be >= 1");
    *count = s->impl.time_passes(low != 0, reps, ms, cap);
  });
}

int qpch_solver_time_iterations(qpch_solver* s, int low, int len, int reps, float* us_per_iteration) {
  return guarded([&] {
    if (len < 1 || reps < 1) throw qpch::InvalidArgument("len and reps must be >= 1");
    if (!us_per_iteration) throw qpch::InvalidArgument("null output");
    *us_per_iteration = s->impl.time_iterations(low != 0, len, reps);
  });
}

int qpch_solver_pass_info(qpch_solver* s, int low, int* axes, int* modes, int cap, int* count) {
  return guarded([&] {
    const auto passes = s->impl.iteration_passes(low != 0);
    int k = 0;
    for (const auto& ps : passes) {
      if (k >= cap) break;
      axes[k] = ps.axis;
      modes[k] = static_cast<int>(ps.mode);
      ++k;
    }
    *count = k;
  });
}

int qpch_solver_synchronize(qpch_solver* s) {
  return guarded([&] { QPCH_HIP(hipStreamSynchronize(s->impl.stream)); });
}

// --- operators ---------------------------------------------------------------------------

int qpch_fft(float* host, int ndim, const size_t* shape, int forward) {
  return guarded([&] {
    const qpch::Dims d = qpch::make_dims(ndim, shape);
    const qpch::OwnStream st;
    qpch::DevBuf<float2> buf = qpch::upload(reinterpret_cast<const float2*>(host), d.n, st.s);
    qpch::device_fft(buf.get(), d, 1, forward != 0, static_cast<float>(1.0 / std::sqrt(static_cast<double>(d.n))), st.s);
    qpch::copy_sync(host, buf.get(), sizeof(float2) * d.n, hipMemcpyDeviceToHost, st.s);
  });
}

int qpch_modulus_project(float* host_x, const float* amplitude, const uint8_t* mask, int ndim,
                         const size_t* shape, double* numerator) {
  return guarded([&] {
    const qpch::Dims d = qpch::make_dims(ndim, shape);
    std::vector<float> signed_amp(d.n);
    for (long long i = 0; i < d.n; ++i) signed_amp[i] = mask[i] ? -1.0f : amplitude[i];
    const qpch::OwnStream st;
    qpch::DevBuf<float> amp = qpch::upload(signed_amp.data(), d.n, st.s);
    qpch::DevBuf<float2> buf = qpch::upload(reinterpret_cast<const float2*>(host_x), d.n, st.s);
    const int grid = qpch::grid_for(d.n, 256, 1024);
    qpch::DevBuf<double> part(grid);
    qpch::DevBuf<int> counter(1);
    const float s = static_cast<float>(1.0 / std::sqrt(static_cast<double>(d.n)));
    qpch::device_fft(buf.get(), d, 1, true, 1.0f, st.s);
    hipLaunchKernelGGL(qpch::dev::modulus_elementwise, dim3(grid, 1), dim3(256), 0, st.s, buf.get(), amp.get(),
                       d.n, s, s, part.get(), counter.get());
    QPCH_HIP(hipGetLastError());
    qpch::device_fft(buf.get(), d, 1, false, 1.0f, st.s);
    std::vector<double> hp(grid);
    qpch::copy_sync(hp.data(), part.get(), sizeof(double) * grid, hipMemcpyDeviceToHost, st.s);
    double num = 0.0;
    for (double v : hp) num += v;
    *numerator = num;
    qpch::copy_sync(host_x, buf.get(), sizeof(float2) * d.n, hipMemcpyDeviceToHost, st.s);
  });
}

int qpch_support_project(float* host_x, const uint8_t* support, size_t n) {
  return guarded([&] {
    const long long nn = static_cast<long long>(n);
    const qpch::OwnStream st;
    qpch::DevBuf<float2> buf = qpch::upload(reinterpret_cast<const float2*>(host_x), n, st.s);
    qpch::DevBuf<std::uint8_t> sup = qpch::upload(support, n, st.s);
    hipLaunchKernelGGL((qpch::dev::convert<float2, float2>), dim3(qpch::grid_for(nn)), dim3(256), 0, st.s,
                       buf.get(), buf.get(), nn, 1.0f, sup.get());
    QPCH_HIP(hipGetLastError());
    qpch::copy_sync(host_x, buf.get(), sizeof(float2) * n, hipMemcpyDeviceToHost, st.s);
  });
}

int qpch_gaussian_blur(const float* host_in, float* host_out, int ndim, const size_t* shape, float sigma) {
  return guarded([&] {
    if (!(sigma >= 0.0f)) throw qpch::InvalidArgument("sigma must be >= 0");
    const qpch::Dims d = qpch::make_dims(ndim, shape);
    const qpch::OwnStream st;
    qpch::DevBuf<float> a = qpch::upload(host_in, d.n, st.s);
    qpch::DevBuf<float> b(d.n);
    float* cur = a.get();
    float* other = b.get();
    std::vector<qpch::DevBuf<float>> keep;
    for (int ax = d.nd - 1; ax >= 0; --ax) {
      const qpch::Taps t = qpch::blur_taps(d.d[ax], sigma);
      keep.push_back(qpch::upload(t.taps.data(), t.taps.size(), st.s));
      qpch::dev::BlurAxis bx{d.d[ax], static_cast<int>(d.inner(ax)), static_cast<int>(t.taps.size()), t.J,
                             keep.back().get()};
      hipLaunchKernelGGL((qpch::dev::blur_axis<float2, false>), dim3(qpch::grid_for(d.n), 1), dim3(256), 0,
                         st.s, cur, other, d.n, bx);
      QPCH_HIP(hipGetLastError());
      std::swap(cur, other);
    }
    qpch::copy_sync(host_out, cur, sizeof(float) * d.n, hipMemcpyDeviceToHost, st.s);
  });
}

int qpch_autocorrelation_support(const float* intensity, int count, int ndim, const size_t* shape,
                                 float threshold, uint8_t* support) {
  return guarded([&] {
    if (count < 1 || count > 65535) throw qpch::InvalidArgument("count must be between 1 and 65535");
    if (!intensity || !support) throw qpch::InvalidArgument("intensity and support are required");
    const qpch::Dims d = qpch::make_dims(ndim, shape);
    const qpch::dev::Grid3 g = qpch::grid3(d);
    const long long total = d.n * count;
    const qpch::OwnStream st;
    qpch::DevBuf<float> buf = qpch::upload(intensity, static_cast<size_t>(total), st.s);
    qpch::DevBuf<float2> w(static_cast<size_t>(total));
    const int blocks = qpch::grid_for(d.n, 256, 1024);
    qpch::DevBuf<float> block_max(static_cast<size_t>(blocks) * count);
    qpch::DevBuf<std::uint8_t> out(static_cast<size_t>(total));
    hipLaunchKernelGGL(qpch::dev::centred_to_wave<false>, dim3(blocks, count), dim3(256), 0, st.s, buf.get(),
                       w.get(), d.n, g);
    QPCH_HIP(hipGetLastError());
    qpch::device_fft(w.get(), d, count, false, static_cast<float>(1.0 / std::sqrt(static_cast<double>(d.n))), st.s);
    // the magnitudes overwrite the intensities (read by the first kernel only)
    hipLaunchKernelGGL(qpch::dev::autocorrelation_magnitude, dim3(blocks, count), dim3(256), 0, st.s, w.get(),
                       buf.get(), block_max.get(), d.n, g);
    hipLaunchKernelGGL(qpch::dev::autocorrelation_cut, dim3(blocks, count), dim3(256), 0, st.s, buf.get(),
                       block_max.get(), blocks, d.n, threshold, out.get());
    QPCH_HIP(hipGetLastError());
    qpch::copy_sync(support, out.get(), static_cast<size_t>(total), hipMemcpyDeviceToHost, st.s);
  });
}

int qpch_simulate_intensity(const float* object, int count, int ndim, const size_t* shape, float* intensity) {
  return guarded([&] {
    if (count < 1 || count > 65535) throw qpch::InvalidArgument("count must be between 1 and 65535");
    if (!object || !intensity) throw qpch::InvalidArgument("object and intensity are required");
    const qpch::Dims d = qpch::make_dims(ndim, shape);
    const qpch::dev::Grid3 g = qpch::grid3(d);
    const long long total = d.n * count;
    const qpch::OwnStream st;
    qpch::DevBuf<float> obj = qpch::upload(object, static_cast<size_t>(2 * total), st.s);
    qpch::DevBuf<float2> w(static_cast<size_t>(total));
    qpch::DevBuf<float> out(static_cast<size_t>(total));
    const int blocks = qpch::grid_for(d.n, 256, 1024);
    hipLaunchKernelGGL(qpch::dev::centred_to_wave<true>, dim3(blocks, count), dim3(256), 0, st.s, obj.get(), w.get(),
                       d.n, g);
    QPCH_HIP(hipGetLastError());
    qpch::device_fft(w.get(), d, count, true, static_cast<float>(1.0 / std::sqrt(static_cast<double>(d.n))), st.s);
    hipLaunchKernelGGL(qpch::dev::wave_to_intensity, dim3(blocks, count), dim3(256), 0, st.s, w.get(), out.get(), d.n,
                       g);
    QPCH_HIP(hipGetLastError());
    qpch::copy_sync(intensity, out.get(), sizeof(float) * static_cast<size_t>(total), hipMemcpyDeviceToHost, st.s);
  });
}

// --- autotuner -----------------------------------------------------------------------

}  // extern "C"

namespace {
std::mutex g_tune_mu;
std::map<std::string, qpch_tuning> g_tune_cache;
std::string g_tune_log;

std::string tune_key(const qpch_problem& p) {
  std::string k = std::to_string(p.ndim);
  for (int a = 0; a < p.ndim; ++a) k += "x" + std::to_string(p.shape[a]);
  k += "/s" + std::to_string(p.starts * (p.problems > 1 ? p.problems : 1)) + "/st" + std::to_string(p.storage) +
       "/m" + std::to_string(p.mixed) + "/b" + std::to_string(p.batch) + "/l" + std::to_string(p.lanes) + "/r" +
       std::to_string(p.radix) + "/p" + std::to_string(p.persist);
  return k;
}
}  // namespace

extern "C" {

const char* qpch_autotune_log(void) {
  thread_local std::string copy;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  copy = g_tune_log;
  return copy.c_str();
}

int qpch_autotune(const qpch_problem* problem, qpch_tuning* out) {
  return guarded([&] {
    if (!problem) throw qpch::InvalidArgument("null problem");
    const std::string key = tune_key(*problem);
    {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        *out = it->second;
        return;
      }
    }
    const qpch::Dims d = qpch::make_dims(problem->ndim, problem->shape);
    const int total = problem->starts * (problem->problems > 1 ? problem->problems : 1);
    std::vector<int> batches;
    if (problem->batch > 0) {
      batches.push_back(problem->batch);  // fixed by the caller: tune the rest
    } else {
      for (int b = total; b >= 1 && batches.size() < 4; b /= 2) {  // all, 1/2, 1/4, 1/8
        if (total % b == 0) batches.push_back(b);
      }
    }
    // Plane vs line passes (3D) change the order of the arithmetic and are not tuned: the
    // plane pass where supported (it was the faster one in every measurement), unless the
    // caller asks for line passes.
    const bool plane_ok = d.nd == 3 && qpch::plane_supported(d.d[1], d.d[2]);
    const std::vector<int> planes{problem->plane != 0 ? problem->plane : (plane_ok ? 1 : 0)};
    // The radix changes the arithmetic (rounding), so it is not tuned: results must not
    // depend on timings, on the split of the starts into launches (batched == looped) or
    // on what the process ran before.  Radix 16 everywhere unless the caller fixes it in
    // problem->radix (radix 8 was never 2% faster on the workloads with all their starts
    // in one launch: 128^2 to 1024^2 and 128^3, fp32 and fp16).
    const int tuned_radix = problem->radix != 0 ? problem->radix : 16;
    std::vector<int> radices{tuned_radix};
    std::string line = "qpch autotune " + key + ":";
    qpch_tuning best{};
    best.us_per_iteration = 1e300;
    const bool lowp = problem->storage != QPCH_F32;
    auto measure = [&](const qpch_tuning& t) -> double {
      qpch_problem q = *problem;
      q.batch = t.batch;
      q.lanes = t.lanes;
      q.plane = t.plane;
      q.xcd = t.xcd;
      q.radix = t.radix;
      q.persist = t.persist;
      q.start = nullptr;
      q.schedule = QPCH_GRAPH;
      auto s = std::make_unique<qpch_solver>();
      s->impl.init(q);
      auto& m = s->impl;
      m.begin_run(160);
      hipEvent_t e0{}, e1{};
      QPCH_HIP(hipEventCreate(&e0));
      QPCH_HIP(hipEventCreate(&e1));
      // Per iteration: the fastest of three rounds of two chunks of ten (the reference's
      // default sync_every, the shortest chunk the C++ host launches), after a warm-up chunk
      // that captures the graph (candidates a few percent apart must not swap places on a
      // hiccup).  Chunks of four made the persistent launch look slower than the per-pass
      // launches at 256^2 x 32 (21.4 against 19.6 us per iteration), where chunks of ten
      // and more run it faster (13 and 12.3 against 16.2).
      auto per_iteration = [&](bool refine) {
        m.iterate(10, 0.9f, -0.8f, refine);
        float best_ms = 1e30f;
        for (int round = 0; round < 3; ++round) {
          QPCH_HIP(hipEventRecord(e0, m.stream));
          for (int r = 0; r < 2; ++r) m.iterate(10, 0.9f, -0.8f, refine);
          QPCH_HIP(hipEventRecord(e1, m.stream));
          QPCH_HIP(hipEventSynchronize(e1));
          float ms = 0.0f;
          QPCH_HIP(hipEventElapsedTime(&ms, e0, e1));
          best_ms = std::min(best_ms, ms);
        }
        return best_ms * 1e3 / 20.0;
      };
      double us = per_iteration(!lowp);
      // A mixed solver also runs its ER stages in fp32 with the same launch choices:
      // weigh the two precisions 2 : 1, as the recipes split RAAR / HIO and ER.
      if (lowp && q.mixed) us = (2.0 * us + per_iteration(true)) / 3.0;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return us;
    };
    // images whose two axes are not powers of two, other than square smooth ones: persistent
    // launches only with persist = 1
    const bool generic2d = d.nd == 2 && qpch::generic_axis(d.d[0]) && qpch::generic_axis(d.d[1]) &&
                           !qpch::twopass_smooth_supported(d.d[0], d.d[1]);
    auto trial = [&](const qpch_tuning& t) {
      qpch_tuning r = t;
      r.us_per_iteration = measure(t);
      line += " [batch " + std::to_string(t.batch) + (t.lanes > 1 ? " x" + std::to_string(t.lanes) + " lanes" : "") +
              (t.plane ? (t.plane == 1 ? " plane" : " lines") : "") + (t.xcd == 2 ? " xcd-off" : t.xcd == 3 ? " xcd-off-modulus" : "") +
              (t.radix == 8 ? " radix-8" : "") + (t.persist == 2 ? " pass-launches" : "") +
              (t.persist == 1 && generic2d ? " persistent" : "") + ": " +
              std::to_string(r.us_per_iteration) + " us/it]";
      // A candidate must beat the incumbent by 2%: timings a few tenths of a percent apart
      // are noise, and the first candidate is the default (all starts in one launch, one
      // lane, XCD-affine numbering), which then stays unless another is clearly faster.
      if (r.us_per_iteration < 0.98 * best.us_per_iteration) best = r;
    };
    // first the split of the starts into launches (default kernels), then the kernel
    // choices for the best split
    for (int b : batches) {
      std::vector<int> lane_options{1};
      if (problem->lanes != 1 && total / b > 1) lane_options.push_back(std::min(4, total / b));
      for (int ln : lane_options) trial(qpch_tuning{b, planes[0], 0, 0.0, ln, radices[0], problem->persist});
    }
    const qpch_tuning split = best;
    // The persistent launches place every start on an XCD by the hardware's XCD number,
    // whatever the block numbering: nothing to try there.
    const bool image = d.nd == 2 && qpch::image_supported(d.d[0], d.d[1]) && problem->plane != 2;
    auto persistent_for = [&](int storage) {
      if (d.nd != 2 || problem->persist == 2 || qpch::twopass_xcds() == 0) return false;
      if (storage == QPCH_F32 && qpch::generic_axis(d.d[0]) && qpch::generic_axis(d.d[1])) {
        // (smooth axes run smooth_pass: the persistent launch of square images of a smooth
        // length runs it too, twopass_generic does not)
        if (qpch::twopass_smooth_supported(d.d[0], d.d[1])) return true;
        return !qpch::smooth_axis(d.d[0]) && !qpch::smooth_axis(d.d[1]) && !qpch::chirp_axis(d.d[0]) &&
               !qpch::chirp_axis(d.d[1]) &&
               qpch::twopass_generic_supported(qpch::generic_plan(d.d[0]), qpch::generic_plan(d.d[1]));
      }
      return tuned_radix == 16 && qpch::twopass_supported(d.d[0], d.d[1], storage) && !(storage == QPCH_F32 && image);
    };
    const bool can_persist =
        persistent_for(problem->storage) && (!lowp || !problem->mixed || persistent_for(QPCH_F32));
    // whether the default candidates (problem->persist) already run persistent launches
    const bool persistent = can_persist && (!generic2d || problem->persist == 1);
    for (int pl : planes) {
      for (int xcd : {0, 2, 3}) {
        if (xcd != 0 && persistent) continue;
        if (xcd != 0 && !(split.batch >= 8 && split.batch % 8 == 0)) continue;  // numbering only differs then
        for (int rx : radices) {
          if (pl == split.plane && xcd == split.xcd && rx == split.radix) continue;
          trial(qpch_tuning{split.batch, pl, xcd, 0.0, split.lanes, rx, split.persist});
        }
      }
    }
    // Persistent or per-pass launches (the same arithmetic, bit for bit): the persistent
    // kernel saves each pass's launch ramp and drain and keeps an XCD's starts in its L2,
    // but pays a queue claim and a wait per tile, which loses where a launch already holds
    // many tiles per CU (512^2 x 64 fp16: 290 against 267 us per iteration).  When the
    // per-pass launches win, their XCD numbering is tried too.
    if (can_persist && generic2d && problem->persist == 0) {
      trial(qpch_tuning{best.batch, best.plane, 0, 0.0, best.lanes, best.radix, 1});
    } else if (persistent && problem->persist == 0) {
      const qpch_tuning keep = best;
      for (int xcd : {0, 2, 3}) {
        if (xcd != 0 && !(best.batch >= 8 && best.batch % 8 == 0)) continue;
        trial(qpch_tuning{keep.batch, keep.plane, xcd, 0.0, keep.lanes, keep.radix, 2});
        if (best.persist != 2) break;  // numbering only matters once the pass launches win
      }
    }
    line += " -> batch " + std::to_string(best.batch) + " lanes " + std::to_string(best.lanes) + " plane " +
            std::to_string(best.plane) + " xcd " + std::to_string(best.xcd) + " radix " + std::to_string(best.radix) +
            (best.persist == 2 ? " pass-launches" : "") + (best.persist == 1 && generic2d ? " persistent" : "") +
            "\n";
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = best;
    g_tune_log += line;
    if (const char* v = std::getenv("QPCH_TUNE_LOG"); v && v[0] == '1') std::fputs(line.c_str(), stderr);
    *out = best;
  });
}

int qpch_amplitude_loss(const void* x, const void* amplitude, long long amplitude_stride, void* grad,
                        double* loss, int ndim, const size_t* shape, int batch, int precision, void* stream) {
  return guarded([&] {
    const qpch::LossCall call =
        qpch::loss_call(x, amplitude, amplitude_stride, grad, loss, ndim, shape, batch, precision, stream);
    // stream-ordered scratch: safe for callers that interleave several streams
    void* base = nullptr;
    QPCH_HIP(hipMallocAsync(&base, call.plan.bytes, call.st));
    try {
      qpch::amplitude_loss_run(call, base);
    } catch (...) {
      (void)hipFreeAsync(base, call.st);
      throw;
    }
    QPCH_HIP(hipFreeAsync(base, call.st));
  });
}

int qpch_amplitude_loss_noalloc(const void* x, const void* amplitude, long long amplitude_stride, void* grad,
                                double* loss, int ndim, const size_t* shape, int batch, int precision, void* stream) {
  return guarded([&] {
    const qpch::LossCall call =
        qpch::loss_call(x, amplitude, amplitude_stride, grad, loss, ndim, shape, batch, precision, stream);
    for (int a = 0; a < call.d.nd; ++a) {
      if (!call.plan.fused && qpch::needs_bluestein(call.d.d[a]) &&
          !qpch::bluestein_lookup(call.d.d[a], call.precision == 1)) {
        throw qpch::InvalidArgument("axis length " + std::to_string(call.d.d[a]) +
                                    " needs Bluestein tables, made by a first qpch_amplitude_loss call of this shape");
      }
    }
    if (call.plan.bytes > kNoAllocScratchBytes) {
      throw qpch::InvalidArgument("this call needs " + std::to_string(call.plan.bytes) +
                                  " bytes of scratch; qpch_amplitude_loss_noalloc holds " +
                                  std::to_string(kNoAllocScratchBytes) +
                                  " (use qpch_amplitude_loss, which allocates)");
    }
    qpch::amplitude_loss_run(call, qpch::noalloc_scratch());
  });
}

}  // extern "C"
