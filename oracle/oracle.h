/*
 * CPU oracle for the phase-retrieval reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the reference's CPU
 * backend (src/cpu/kernels.cpp, src/cpu/cpu_solver.cpp, src/cpu/fft.cpp semantics,
 * src/core/{pointwise.hpp,layout.hpp,problem.cpp,recipe.cpp,metrics.cpp}).  It is the
 * checker the HIP path is compared against; only the tests, the smoke check and the
 * benchmark's CPU-baseline leg load it.  Nothing in the product links it.
 *
 * Every function exists twice: `_f` computes in float (the reference's own precision)
 * and `_d` in double (the FP64 twin used to state the error of every config).
 *
 * Arrays: complex data is interleaved (re, im); shapes are row-major, slowest axis
 * first, 2D or 3D.  "FFT order" is numpy's ifftshift of a centre-origin array.
 */
#ifndef QPC_ORACLE_H
#define QPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Recipe operations, the same 7-tuple the reference's Python binding flattens to
 * (python/bindings.cpp:22-23). */
enum { OR_ER = 0, OR_HIO = 1, OR_RAAR = 2, OR_SHRINKWRAP = 3, OR_DETWIN_BEGIN = 4, OR_DETWIN_END = 5 };
enum { OR_MAX = 0, OR_RMS = 1 };

typedef struct {
  int kind;
  int n;
  float beta;
  float threshold;
  float sigma;
  int method;
  int axis;
} or_op;

/* hash RNG, src/core/pointwise.hpp:19-32 */
uint64_t or_mix64(uint64_t z);
float or_uniform(uint64_t seed, uint64_t start, uint64_t index);

/* centre-origin <-> FFT order, src/core/layout.hpp:19-47; elem = bytes per element */
void or_roll(const void* in, void* out, size_t elem, int ndim, const size_t* shape, int to_fft_order);

/* src/core/problem.cpp:101-127 */
void or_half_support(const uint8_t* support, uint8_t* out, int ndim, const size_t* shape, int axis);

/* OpenMP threads used by every parallel loop below (0 = library default). */
void or_set_threads(int threads);

#define OR_DECLARE(R, S)                                                                          \
  /* unitary n-D FFT in place, src/cpu/fft.cpp:12-35 */                                           \
  void or_fft_##S(R* data, int ndim, const size_t* shape, int forward);                           \
  /* src/cpu/kernels.cpp:16-23 */                                                                 \
  double or_usable_energy_##S(const R* amplitude, const uint8_t* mask, size_t n);                 \
  /* src/cpu/kernels.cpp:25-42; returns the Fourier error (float-rounded in _f) */                \
  double or_modulus_##S(R* wave, const R* amplitude, const uint8_t* mask, size_t n,               \
                        double energy);                                                           \
  /* Same projection, numerator only: sum over usable pixels of (|w| - A)^2. */                   \
  double or_modulus_num_##S(R* wave, const R* amplitude, const uint8_t* mask, size_t n);          \
  /* src/cpu/kernels.cpp:44-50 */                                                                 \
  void or_update_##S(R* x, const R* projected, const uint8_t* support, size_t n, float a,         \
                     float b);                                                                    \
  /* src/cpu/kernels.cpp:52-75 */                                                                 \
  void or_blur_##S(const R* image, R* out, int ndim, const size_t* shape, float sigma);           \
  /* src/cpu/kernels.cpp:77-102 */                                                                \
  void or_shrinkwrap_##S(const R* x, int ndim, const size_t* shape, float threshold,              \
                         float sigma, int method, uint8_t* support);                              \
  /* src/cpu/kernels.cpp:104-117 (FFT-order amplitude and support) */                             \
  void or_random_start_##S(const R* amplitude, const uint8_t* support, int ndim,                  \
                           const size_t* shape, uint64_t seed, uint64_t start, R* x);             \
  /* src/core/problem.cpp:72-88 (centre-origin in and out) */                                     \
  void or_autocorrelation_support_##S(const float* intensity, int ndim, const size_t* shape,      \
                                      float threshold, uint8_t* support);                         \
  /* src/core/problem.cpp:90-99 */                                                                \
  void or_simulate_intensity_##S(const R* object, int ndim, const size_t* shape, R* intensity);   \
  /* src/core/metrics.cpp:17-86 */                                                                \
  double or_aligned_error_##S(const R* a, const R* reference, int ndim, const size_t* shape);     \
  /* The solver, src/cpu/cpu_solver.cpp:18-110.  Inputs are centre-origin like Problem. */        \
  void* or_solver_create_##S(int ndim, const size_t* shape, const float* amplitude,               \
                             const uint8_t* mask, const uint8_t* support, int starts,             \
                             uint64_t seed, const float* start);                                  \
  /* src/core/recipe.cpp:30-61 + cpu_solver.cpp:45-50; returns 0, or -1 for a bad recipe */      \
  int or_solver_run_##S(void* solver, const or_op* ops, int nops);                                \
  int or_solver_iterations_##S(const void* solver);                                               \
  /* cpu_solver.cpp:52-59: centre-origin objects, supports, and the error history              \
   * errors[start * iterations + t] */                                                            \
  void or_solver_results_##S(const void* solver, R* objects, uint8_t* supports, double* errors);  \
  /* FFT-order state access for step-by-step comparisons */                                       \
  void or_solver_get_state_##S(const void* solver, R* x, uint8_t* supports);                      \
  void or_solver_set_state_##S(void* solver, const R* x, const uint8_t* supports);                \
  void or_solver_destroy_##S(void* solver);

OR_DECLARE(float, f)
OR_DECLARE(double, d)

#undef OR_DECLARE

#ifdef __cplusplus
}
#endif

#endif
