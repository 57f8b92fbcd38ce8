/*
 * qpch: the C ABI of the MI355X (gfx950) reconstruction engine.
 *
 * Everything the C++ host (qpcu::Backend::hip) and the Python/torch adapter need from the
 * GPU goes through these entry points: plain pointers and sizes, integer status codes, no
 * C++ types and no exceptions across the boundary.  Arrays are FFT order (numpy
 * ifftshift of the centre-origin arrays), row-major, slowest axis first; complex values
 * are interleaved float pairs (complex64).
 *
 * Reference interfaces replaced (NaNIsaac/QoraPhaseCU v0.1.0):
 *   qpch_solver_*          the CudaSolver behind detail::make_cuda_solver
 *                          (src/cuda/cuda_solver.cu:58-500, src/core/backends.hpp:9-10)
 *   qpch_solver_iterate    CudaSolver::iterate/graph/step_device (cuda_solver.cu:300-421)
 *   qpch_solver_shrinkwrap CudaSolver::apply(Shrinkwrap) (cuda_solver.cu:240-254)
 *   qpch_solver_*_support  the device side of apply(DetwinBegin/End) (cuda_solver.cu:256-276)
 *   qpch_solver_read_*     collect_errors/report/results (cuda_solver.cu:165-191,423-453)
 *   qpch_fft / qpch_modulus_project / qpch_support_project / qpch_gaussian_blur
 *                          qpcu::cuda operators (include/qoraphasecu/operators.hpp:25-33,
 *                          src/cuda/operators.cu:62-130) and the cuFFT plans they use
 *                          (src/cuda/common.cuh:84-140)
 *   qpch_device_count / qpch_device_info
 *                          cuda_available / cuda_device_info (src/cuda/operators.cu:13-36)
 *   qpch_amplitude_loss    no counterpart: the adjoint of the amplitude loss for the torch
 *                          autograd Function.
 *
 * Conventions: every function returns QPCH_OK or an error code and leaves a message for
 * qpch_last_error() (thread-local).  One solver handle must not be used from two threads
 * at once; different handles may run in different threads at the same time.  Solver calls
 * run on the solver's own non-blocking stream and return after the work is queued, except
 * those that hand data back to the host.  No call touches the null stream: host copies go
 * through the solver's stream (the standalone operators create one of their own), so one
 * thread's copies never wait on, or break, another thread's graph capture.
 */
#ifndef QPCH_H
#define QPCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPCH_OK 0
#define QPCH_EINVAL 1  /* bad argument: maps to std::invalid_argument / ValueError */
#define QPCH_EDEVICE 2 /* no device or a device failure: std::runtime_error */

/* storage precision of the iterate and the wave between passes */
#define QPCH_F32 0
#define QPCH_F16 1
#define QPCH_BF16 2

/* execution schedules, the reference's Variant ladder (include/qoraphasecu/solver.hpp:21-29) */
#define QPCH_UNFUSED 0 /* one kernel per operation: transforms per axis, modulus, update */
#define QPCH_FUSED 1   /* two fused passes per iteration (three or four in 3D), eager launches;
                          axes that are not powers of two (fp32, up to 2048) run the same
                          passes in a generic mixed-radix kernel */
#define QPCH_GRAPH 2   /* fused passes replayed from hipGraphs */

typedef struct qpch_solver qpch_solver;

typedef struct {
  int ndim;          /* 2 or 3 */
  size_t shape[3];   /* slowest first */
  const float* amplitude;  /* problems * n, FFT order: sqrt of the clipped intensity */
  const uint8_t* mask;     /* problems * n, FFT order, 1 = unusable pixel */
  const uint8_t* support;  /* problems * n, FFT order, the initial support of each problem's starts */
  int starts;          /* per problem */
  int batch;           /* starts per launch group, must divide problems * starts (0 = all) */
  uint64_t seed;       /* random starting phases: start k of a problem draws (seed, k) */
  const float* start;  /* problems * starts * n complex, FFT order, or NULL for random starts */
  int storage;         /* QPCH_F32 / QPCH_F16 / QPCH_BF16 */
  int mixed;           /* with F16/BF16: run refine (ER) iterations in F32 */
  int schedule;        /* QPCH_UNFUSED / QPCH_FUSED / QPCH_GRAPH */
  int problems;        /* independent problems of this shape stacked in one solver (0 = 1) */
  int plane;           /* 0 = automatic, 1 = plane kernels (3D: xy-plane update pass; 2D fp32 images
                          up to 8192 elements: whole iterations in one work-group per start),
                          2 = line passes only */
  int xcd;             /* 0 = XCD-affine block numbering when starts per launch are a multiple of 8
                          (a start's blocks share one XCD's L2), 2 = plain numbering (a tile of
                          every start on one XCD), 3 = plain for the frequency-axis pass only
                          (its amplitude is shared by all starts) */
  int lanes;           /* launch groups in flight at once, each on its own stream (0 or 1 = one
                          after another; at most 8) */
  int radix;           /* FFT radix of the line passes: 0 or 16 = radix 16, 8 = radix 8 for axes
                          of 64 to 2048 (other axes stay radix 16) */
  int persist;         /* 0 = 2D chunks of iterations as one persistent launch where a kernel
                          exists for power-of-two axes (64 to 1024, radix 16) and for square
                          fp32 images of a smooth length (R1 R2, R1 in 2..16 a power of two,
                          R2 in {3, 5, 15}; 96, 160; 18, 36, 72; 14, 28, 56, 22, 44, 26, 52;
                          45, 50, 54, 100, 42, 70),
                          1 = also for the fp32 images whose two axes run the generic kernel,
                          2 = a launch per pass; every choice is the same arithmetic, bit for
                          bit */
} qpch_problem;

/* Launch configuration chosen by qpch_autotune. */
typedef struct {
  int batch;  /* starts per launch */
  int plane;  /* as qpch_problem.plane */
  int xcd;    /* as qpch_problem.xcd */
  double us_per_iteration;
  int lanes;  /* as qpch_problem.lanes */
  int radix;  /* as qpch_problem.radix */
  int persist;  /* as qpch_problem.persist */
} qpch_tuning;

typedef struct {
  char name[256];
  char arch[64];            /* gcnArchName, e.g. gfx950:sramecc+:xnack- */
  int compute_units;
  size_t memory_bytes;
  size_t l2_bytes;
  double peak_bandwidth_gbs; /* peak HBM bandwidth: the published figure for gfx950 (MI355X,
                                8000 GB/s), otherwise 2 x memory clock x bus width */
} qpch_device_info_t;

const char* qpch_last_error(void);

int qpch_device_count(int* count);
int qpch_device_info(qpch_device_info_t* info);

/* --- solver ---------------------------------------------------------------------------- */

int qpch_solver_create(const qpch_problem* problem, qpch_solver** out);
void qpch_solver_destroy(qpch_solver* solver);

/* Sizes the error history for `capacity` more iterations and resets its slot counter.
 * Must precede qpch_solver_iterate; the history holds the FP64 numerators
 * sum (|F x| - A)^2 of every start per iteration. */
int qpch_solver_begin_run(qpch_solver* solver, int capacity);

/* `iterations` iterations of x <- S ? P_M x : a x + b P_M x for every start.  refine = 1
 * marks an ER stage (run in F32 by a mixed solver). */
int qpch_solver_iterate(qpch_solver* solver, int iterations, float a, float b, int refine);

/* method 0 = max, 1 = rms over the current support */
int qpch_solver_shrinkwrap(qpch_solver* solver, float threshold, float sigma, int method);

int qpch_solver_get_support(qpch_solver* solver, uint8_t* host /* starts * n */);
int qpch_solver_set_support(qpch_solver* solver, const uint8_t* host);
int qpch_solver_save_support(qpch_solver* solver);
int qpch_solver_restore_support(qpch_solver* solver);
/* Detwin on the device: every start's support cut at its centre of mass along `axis`
 * (centre-origin coordinates), bit-identical to the host computation of the reference's
 * half_support (src/core/problem.cpp:101-127).  Replaces the reference's host round trip
 * (get support, half_support per start, set support: src/cpu/cpu_solver.cpp:70-80,
 * src/cuda/cuda_solver.cu:256-276); save the support first to restore it at the end. */
int qpch_solver_half_support(qpch_solver* solver, int axis);

/* errors[t * starts + s], t in [first, first + count) of the current run */
int qpch_solver_read_errors(qpch_solver* solver, int first, int count, double* host);
/* objects in FFT order, rescaled to the data (starts * n complex) */
int qpch_solver_read_state(qpch_solver* solver, float* host);
/* usable energy sum A^2 in the solver's scale: the error denominator (problem 0) */
double qpch_solver_energy(const qpch_solver* solver);
/* the same for every problem of the solver (problems doubles) */
int qpch_solver_energies(const qpch_solver* solver, double* out);
int qpch_solver_synchronize(qpch_solver* solver);

/* Measurement: the average duration (ms) of each pass of one iteration of the first
 * launch group, every pass replayed `reps` times between two hipEvents on the solver's
 * stream (low = 1: the reduced-precision passes).  Changes the iterate: time a solver
 * whose state is not needed afterwards.  No reference counterpart; the reference's
 * equivalent is its qpcu_profile tool (bench/profile.cu:15-36). */
int qpch_solver_time_passes(qpch_solver* solver, int low, int reps, float* ms, int cap, int* count);
/* Measurement: microseconds per iteration of all starts on the solver's own schedule,
 * i.e. chunks of `len` iterations replayed from their hipGraphs under QPCH_GRAPH, timed
 * with two hipEvents around `reps` chunks after a warm-up chunk (low = 1: the
 * reduced-precision iteration, low = 0: fp32, which a reduced-precision solver runs only
 * when mixed).  Restarts the error history and changes the iterate.  No reference
 * counterpart; it lets kernel timings taken from eager launches be checked against the
 * graph-replayed path. */
int qpch_solver_time_iterations(qpch_solver* solver, int low, int len, int reps, float* us_per_iteration);
/* the passes of one iteration: axis and mode (0 fwd, 1 inv, 2 modulus, 3 update) */
int qpch_solver_pass_info(qpch_solver* solver, int low, int* axes, int* modes, int cap, int* count);

/* --- autotuner --------------------------------------------------------------------------
 * Times candidate launch configurations of `problem` (starts per launch, concurrent launch
 * groups, XCD-affine or plain block numbering) on trial solvers with random starts, a few
 * graph-replayed iterations each, and returns the fastest (a candidate replaces the best so
 * far only when it is 2% faster, the first candidate being the default: all starts in one
 * launch, one lane, XCD-affine numbering).  Every candidate gives the same results, bit
 * for bit; only the time differs.  The choices that change rounding are not tuned: the FFT
 * radix is 16 unless problem->radix fixes it, and the 3D plane / line choice is
 * problem->plane (the plane pass where supported when 0), so results never depend on
 * timings or on what the process ran before.  Cached per process by shape, total starts
 * and storage; qpch_autotune_log() returns one line per tuned key (also printed to stderr
 * when QPCH_TUNE_LOG=1).  No reference counterpart: the reference leaves Options::batch to
 * the user (include/qoraphasecu/solver.hpp:44-46).  A problem->batch > 0 is kept and only
 * the other choices are tuned; problem->lanes == 1 keeps launch groups sequential. */
int qpch_autotune(const qpch_problem* problem, qpch_tuning* out);
/* The tuner's log so far: a copy owned by the calling thread, valid until that thread calls
 * qpch_autotune_log() again.  Tuners in other threads may append meanwhile. */
const char* qpch_autotune_log(void);

/* --- groups: several solvers (e.g. one per problem shape) advanced by one graph launch --
 * qpch_group_iterate captures, per chunk length and coefficients, one hipGraph in which
 * every solver's passes form an independent branch, and replays it: the solvers'
 * iterations run concurrently from a single launch.  Each solver must have had
 * qpch_solver_begin_run; Shrinkwrap, supports and read-back stay per solver. */
typedef struct qpch_group qpch_group;
int qpch_group_create(qpch_solver* const* solvers, int count, qpch_group** out);
int qpch_group_iterate(qpch_group* group, int iterations, float a, float b, int refine);
int qpch_group_synchronize(qpch_group* group);
void qpch_group_destroy(qpch_group* group);

/* One step of a recipe for qpch_group_run: the flat recipe of the reference
 * (include/qoraphasecu/recipe.hpp:44-80) with each iteration stage as its coefficients. */
#define QPCH_OP_ITERATE 0       /* `iterations` of x <- S ? P_M x : a x + b P_M x */
#define QPCH_OP_SHRINKWRAP 1    /* threshold, sigma, method (0 max, 1 rms) */
#define QPCH_OP_DETWIN_BEGIN 2  /* save the support, cut it at its centre of mass along `axis` */
#define QPCH_OP_DETWIN_END 3    /* restore the saved support */
typedef struct {
  int kind;
  int iterations;
  float a, b;
  int refine;  /* ITERATE: an ER stage (run in fp32 by a mixed solver) */
  float threshold, sigma;
  int method;
  int axis;
} qpch_op;

/* Every solver of the group runs the whole recipe `ops` in ONE graph launch: each launch
 * group of each solver (its `batch` starts) is an independent branch that carries its
 * starts through every iteration, Shrinkwrap, detwin and precision switch, on the
 * solver's lane streams, from one fork to one join.  The graph is captured on the first
 * run of a recipe from a given precision state and replayed after that.  The kernels and
 * results are those of the step-by-step calls (qpch_group_iterate / qpch_solver_iterate,
 * qpch_solver_shrinkwrap, save/half/restore_support), bit for bit; what goes is the launch
 * and host round trip of every chunk and step.  Each solver needs qpch_solver_begin_run
 * with a capacity of at least the recipe's iterations.  Replaces the reference's graph per
 * sync_every chunk with eager stage steps in between (src/cuda/cuda_solver.cu:300-421). */
int qpch_group_run(qpch_group* group, const qpch_op* ops, int count);

/* --- operators on host arrays (synchronous) --------------------------------------------- */

/* unitary n-D transform in place (complex, FFT order) */
int qpch_fft(float* host, int ndim, const size_t* shape, int forward);
/* x <- P_M x on an FFT-order object; *numerator = sum over usable pixels (|F x| - A)^2.
 * amplitude is FFT order with masked pixels marked by mask. */
int qpch_modulus_project(float* host_x, const float* amplitude, const uint8_t* mask, int ndim,
                         const size_t* shape, double* numerator);
int qpch_support_project(float* host_x, const uint8_t* support, size_t n);
/* Gaussian blur with transfer exp(-2 pi^2 sigma^2 |f|^2) (real image, any origin) */
int qpch_gaussian_blur(const float* host_in, float* host_out, int ndim, const size_t* shape,
                       float sigma);

/* make_problem's autocorrelation support for `count` problems of one shape at once
 * (src/core/problem.cpp:72-88, which does one problem on the host around one pocketfft
 * call): each centre-origin intensity (count * n floats, back to back) is moved to FFT
 * order, inverse transformed (unitary), |.| taken, moved back to centre-origin and cut at
 * threshold * its maximum (strict >).  support: count * n bytes, centre-origin, 0 / 1.
 * The whole chain runs on the device; one copy in, one copy out. */
int qpch_autocorrelation_support(const float* intensity, int count, int ndim, const size_t* shape,
                                 float threshold, uint8_t* support);
/* simulate_intensity for `count` objects of one shape (src/core/problem.cpp:90-99):
 * |centred unitary F(object)|^2.  object: count * n complex (interleaved floats),
 * centre-origin; intensity: count * n floats, centre-origin. */
int qpch_simulate_intensity(const float* object, int count, int ndim, const size_t* shape, float* intensity);

/* --- device-pointer entry points (caller's stream) --------------------------------------- */

/* Amplitude loss and its gradient for `batch` objects x (device pointers, complex):
 *   loss[b] = sum over usable pixels (|F x_b| - A)^2   (F unitary, FP64 accumulation)
 *   grad_b  = 2 (x_b - P_M x_b) = dL/dRe x + i dL/dIm x  (the gradient torch expects)
 * The amplitude A is in FFT order (frequency origin at index 0); the real-space origin of
 * x does not matter (a circular shift changes only spectral phases).  A negative amplitude
 * marks a masked pixel.  amplitude_stride: 0 = one amplitude for every object, n = one
 * amplitude array per object (back to back).  precision 0: complex64 x/grad, float A (fused passes for
 * power-of-two extents); 1: complex128 x/grad, double A (axes up to 1024).  loss is a
 * device array of `batch` doubles.  Work is queued on `stream` (a hipStream_t, NULL for the
 * null stream) with stream-ordered scratch memory. */
int qpch_amplitude_loss(const void* x, const void* amplitude, long long amplitude_stride, void* grad,
                        double* loss, int ndim, const size_t* shape, int batch, int precision, void* stream);

/* qpch_amplitude_loss without any allocation, for callers that cannot allocate device
 * memory (code under hipGraph / torch CUDA-graph capture, tight refinement loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments, results and
 * kernels as qpch_amplitude_loss (identical bits); its scratch (the batch's transforms and
 * the FP64 loss partials: 8 or 16 bytes per element of x plus 8 bytes per partial, each
 * part rounded up to 256 bytes) is a 32 MiB device array of the library, one per GPU, held
 * from load until the process ends.  A call needing more scratch fails with QPCH_EINVAL
 * before any work is queued (fp32: x of just under 4.2 million elements in all, e.g.
 * 32 x 256^2 or 1 x 128^3; fp64: half that).  Every call on one GPU uses that GPU's array: calls that overlap on
 * one GPU (two streams, two threads) race and must be ordered by the caller.  Nothing in
 * the array needs clearing between calls.  An axis whose length has a prime factor above
 * 64 (from 256 on) also needs the transform's Bluestein tables, which one call of
 * qpch_amplitude_loss for that shape makes; without them the call fails with QPCH_EINVAL. */
int qpch_amplitude_loss_noalloc(const void* x, const void* amplitude, long long amplitude_stride, void* grad,
                                double* loss, int ndim, const size_t* shape, int batch, int precision,
                                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QPCH_H */
