/*
 * Type-generic body of the CPU oracle (test infrastructure, never shipped in the
 * product).  oracle.c includes this file twice, with R = float / S = f and with
 * R = double / S = d.  <tgmath.h> picks sqrt/exp/cos/sin/hypot by the type of R.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 */

#define OR_CAT_(a, b) a##_##b
#define OR_CAT(a, b) OR_CAT_(a, b)
#define F(name) OR_CAT(name, S)

/* ---------------------------------------------------------------------------------------
 * 1D mixed-radix Stockham FFT.  pocketfft (the reference's CPU FFT, src/cpu/fft.cpp:34)
 * is a third-party dependency absent here; this is the textbook mixed-radix
 * Cooley-Tukey DFT in autosort (Stockham) order, with twiddles evaluated in double and
 * rounded once to R.  It is pinned against numpy.fft (pocketfft C++) fixtures and a
 * brute-force DFT in the tests.
 * ------------------------------------------------------------------------------------- */

typedef struct {
  size_t n;
  int nfac;      /* 0 with n > 1: no factorisation into radices <= OR_MAXR, plain DFT */
  int fac[64];
  R* tw;         /* 2n values: W_n^j = exp(-2 pi i j / n), j in [0, n) */
} F(plan_t);

static F(plan_t) * F(plan_create)(size_t n) {
  F(plan_t)* p = (F(plan_t)*)calloc(1, sizeof(F(plan_t)));
  p->n = n;
  p->tw = (R*)malloc(sizeof(R) * 2 * (n ? n : 1));
  for (size_t j = 0; j < n; ++j) {
    const double a = 2.0 * OR_PI * (double)j / (double)n;
    p->tw[2 * j] = (R)cos(a);
    p->tw[2 * j + 1] = (R)(-sin(a));
  }
  size_t rest = n;
  int nf = 0;
  int ok = 1;
  while (rest % 4 == 0) { p->fac[nf++] = 4; rest /= 4; }
  while (rest % 2 == 0) { p->fac[nf++] = 2; rest /= 2; }
  for (size_t f = 3; rest > 1 && ok; f += 2) {
    while (rest % f == 0) {
      if (f > OR_MAXR) { ok = 0; break; }
      p->fac[nf++] = (int)f;
      rest /= f;
    }
    if (ok && rest > 1 && f * f > rest) { /* what is left is prime */
      if (rest > OR_MAXR) { ok = 0; break; }
      p->fac[nf++] = (int)rest;
      rest = 1;
    }
  }
  if (!ok) nf = 0;
  p->nfac = nf;
  return p;
}

static void F(plan_destroy)(F(plan_t) * p) {
  if (!p) return;
  free(p->tw);
  free(p);
}

/* Radix-r DFT of (vr, vi) into (yr, yi); W_r comes from the line's own table. */
static void F(dft_small)(const F(plan_t) * p, int r, int forward, const R* vr, const R* vi, R* yr,
                         R* yi) {
  if (r == 2) {
    yr[0] = vr[0] + vr[1]; yi[0] = vi[0] + vi[1];
    yr[1] = vr[0] - vr[1]; yi[1] = vi[0] - vi[1];
    return;
  }
  if (r == 4) {
    const R t0r = vr[0] + vr[2], t0i = vi[0] + vi[2];
    const R t1r = vr[0] - vr[2], t1i = vi[0] - vi[2];
    const R t2r = vr[1] + vr[3], t2i = vi[1] + vi[3];
    const R t3r = vr[1] - vr[3], t3i = vi[1] - vi[3];
    yr[0] = t0r + t2r; yi[0] = t0i + t2i;
    yr[2] = t0r - t2r; yi[2] = t0i - t2i;
    if (forward) { /* W_4 = -i */
      yr[1] = t1r + t3i; yi[1] = t1i - t3r;
      yr[3] = t1r - t3i; yi[3] = t1i + t3r;
    } else {
      yr[1] = t1r - t3i; yi[1] = t1i + t3r;
      yr[3] = t1r + t3i; yi[3] = t1i - t3r;
    }
    return;
  }
  const size_t step = p->n / (size_t)r;
  for (int j = 0; j < r; ++j) {
    R sr = vr[0], si = vi[0];
    for (int q = 1; q < r; ++q) {
      const size_t t = ((size_t)j * (size_t)q % (size_t)r) * step;
      const R c = p->tw[2 * t];
      const R s = forward ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
      sr += vr[q] * c - vi[q] * s;
      si += vr[q] * s + vi[q] * c;
    }
    yr[j] = sr;
    yi[j] = si;
  }
}

/* Unnormalised DFT of one contiguous line x (2n values); w is scratch of the same size.
 * Stage with radix r after sub-transforms of length L: butterfly i reads x[i + q n/r],
 * twiddles by W_{L r}^{q (i mod L)} and writes y[(i - k) r + k + q L]. */
static void F(fft_line)(const F(plan_t) * p, R* x, R* w, int forward) {
  const size_t n = p->n;
  if (n <= 1) return;
  if (p->nfac == 0) { /* plain DFT for a large prime factor */
    for (size_t j = 0; j < n; ++j) {
      R sr = 0, si = 0;
      for (size_t q = 0; q < n; ++q) {
        const size_t t = (j * q) % n;
        const R c = p->tw[2 * t];
        const R s = forward ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
        sr += x[2 * q] * c - x[2 * q + 1] * s;
        si += x[2 * q] * s + x[2 * q + 1] * c;
      }
      w[2 * j] = sr;
      w[2 * j + 1] = si;
    }
    memcpy(x, w, sizeof(R) * 2 * n);
    return;
  }
  R* src = x;
  R* dst = w;
  size_t L = 1;
  R vr[OR_MAXR], vi[OR_MAXR], yr[OR_MAXR], yi[OR_MAXR];
  for (int s = 0; s < p->nfac; ++s) {
    const int r = p->fac[s];
    const size_t m = n / (L * (size_t)r);
    const size_t nr = n / (size_t)r;
    for (size_t i = 0; i < nr; ++i) {
      const size_t k = i % L;
      for (int q = 0; q < r; ++q) {
        const R a = src[2 * (i + (size_t)q * nr)];
        const R b = src[2 * (i + (size_t)q * nr) + 1];
        const size_t t = (size_t)q * k * m;
        if (t == 0) {
          vr[q] = a;
          vi[q] = b;
        } else {
          const R c = p->tw[2 * t];
          const R sn = forward ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
          vr[q] = a * c - b * sn;
          vi[q] = a * sn + b * c;
        }
      }
      F(dft_small)(p, r, forward, vr, vi, yr, yi);
      const size_t base = (i - k) * (size_t)r + k;
      for (int q = 0; q < r; ++q) {
        dst[2 * (base + (size_t)q * L)] = yr[q];
        dst[2 * (base + (size_t)q * L) + 1] = yi[q];
      }
    }
    R* t = src;
    src = dst;
    dst = t;
    L *= (size_t)r;
  }
  if (src != x) memcpy(x, src, sizeof(R) * 2 * n);
}

/* The same transform on OR_V lines at once, stored lane-minor: value k of line v at
 * x[2 (k OR_V + v)] (re) and + 1 (im).  Every lane runs exactly the scalar fft_line's
 * operations in the same order (the loops over v only repeat them), so the results are
 * bit-identical; the lane loops let the compiler use SIMD across lines, as pocketfft's
 * multi-line iteration does.  Plans without a factorisation (nfac == 0) are not passed here. */
static void F(dft_small_v)(const F(plan_t) * p, int r, int forward, const R* vr, const R* vi, R* yr, R* yi) {
  /* vr[q * OR_V + v] */
  if (r == 2) {
    for (int v = 0; v < OR_V; ++v) {
      yr[v] = vr[v] + vr[OR_V + v]; yi[v] = vi[v] + vi[OR_V + v];
      yr[OR_V + v] = vr[v] - vr[OR_V + v]; yi[OR_V + v] = vi[v] - vi[OR_V + v];
    }
    return;
  }
  if (r == 4) {
    for (int v = 0; v < OR_V; ++v) {
      const R t0r = vr[v] + vr[2 * OR_V + v], t0i = vi[v] + vi[2 * OR_V + v];
      const R t1r = vr[v] - vr[2 * OR_V + v], t1i = vi[v] - vi[2 * OR_V + v];
      const R t2r = vr[OR_V + v] + vr[3 * OR_V + v], t2i = vi[OR_V + v] + vi[3 * OR_V + v];
      const R t3r = vr[OR_V + v] - vr[3 * OR_V + v], t3i = vi[OR_V + v] - vi[3 * OR_V + v];
      yr[v] = t0r + t2r; yi[v] = t0i + t2i;
      yr[2 * OR_V + v] = t0r - t2r; yi[2 * OR_V + v] = t0i - t2i;
      if (forward) {
        yr[OR_V + v] = t1r + t3i; yi[OR_V + v] = t1i - t3r;
        yr[3 * OR_V + v] = t1r - t3i; yi[3 * OR_V + v] = t1i + t3r;
      } else {
        yr[OR_V + v] = t1r - t3i; yi[OR_V + v] = t1i + t3r;
        yr[3 * OR_V + v] = t1r + t3i; yi[3 * OR_V + v] = t1i - t3r;
      }
    }
    return;
  }
  const size_t step = p->n / (size_t)r;
  for (int j = 0; j < r; ++j) {
    R sr[OR_V], si[OR_V];
    for (int v = 0; v < OR_V; ++v) { sr[v] = vr[v]; si[v] = vi[v]; }
    for (int q = 1; q < r; ++q) {
      const size_t t = ((size_t)j * (size_t)q % (size_t)r) * step;
      const R c = p->tw[2 * t];
      const R s = forward ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
      for (int v = 0; v < OR_V; ++v) {
        sr[v] += vr[q * OR_V + v] * c - vi[q * OR_V + v] * s;
        si[v] += vr[q * OR_V + v] * s + vi[q * OR_V + v] * c;
      }
    }
    for (int v = 0; v < OR_V; ++v) { yr[j * OR_V + v] = sr[v]; yi[j * OR_V + v] = si[v]; }
  }
}

static void F(fft_lines_v)(const F(plan_t) * p, R* x, R* w, int forward) {
  const size_t n = p->n;
  R* src = x;
  R* dst = w;
  size_t L = 1;
  R vr[OR_MAXR * OR_V], vi[OR_MAXR * OR_V], yr[OR_MAXR * OR_V], yi[OR_MAXR * OR_V];
  for (int s = 0; s < p->nfac; ++s) {
    const int r = p->fac[s];
    const size_t m = n / (L * (size_t)r);
    const size_t nr = n / (size_t)r;
    for (size_t i = 0; i < nr; ++i) {
      const size_t k = i % L;
      for (int q = 0; q < r; ++q) {
        const R* in = src + 2 * (i + (size_t)q * nr) * OR_V;
        const size_t t = (size_t)q * k * m;
        if (t == 0) {
          for (int v = 0; v < OR_V; ++v) { vr[q * OR_V + v] = in[2 * v]; vi[q * OR_V + v] = in[2 * v + 1]; }
        } else {
          const R c = p->tw[2 * t];
          const R sn = forward ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
          for (int v = 0; v < OR_V; ++v) {
            const R a = in[2 * v], b = in[2 * v + 1];
            vr[q * OR_V + v] = a * c - b * sn;
            vi[q * OR_V + v] = a * sn + b * c;
          }
        }
      }
      F(dft_small_v)(p, r, forward, vr, vi, yr, yi);
      const size_t base = (i - k) * (size_t)r + k;
      for (int q = 0; q < r; ++q) {
        R* out = dst + 2 * (base + (size_t)q * L) * OR_V;
        for (int v = 0; v < OR_V; ++v) { out[2 * v] = yr[q * OR_V + v]; out[2 * v + 1] = yi[q * OR_V + v]; }
      }
    }
    R* t = src;
    src = dst;
    dst = t;
    L *= (size_t)r;
  }
  if (src != x) memcpy(x, src, sizeof(R) * 2 * n * OR_V);
}

/* Unnormalised transform along every axis, then one unitary scale 1/sqrt(N), like the
 * reference's pocketfft::c2c(..., fct = 1/sqrt(N)) over all axes (src/cpu/fft.cpp:12-35). */
void F(or_fft)(R* data, int ndim, const size_t* shape, int forward) {
  size_t total = 1;
  for (int a = 0; a < ndim; ++a) total *= shape[a];
  for (int a = 0; a < ndim; ++a) {
    const size_t len = shape[a];
    if (len <= 1) continue;
    size_t inner = 1;
    for (int b = a + 1; b < ndim; ++b) inner *= shape[b];
    const size_t outer = total / (len * inner);
    F(plan_t)* plan = F(plan_create)(len);
    const long long lines = (long long)(outer * inner);
    /* Lines go in blocks of OR_V through fft_lines_v where the plan factorises: OR_V
     * adjacent columns (inner > 1, within one outer block) or OR_V consecutive rows; the
     * rest one at a time through fft_line.  Both give the same bits per line. */
    const int blocked = plan->nfac > 0;
    const long long cols_v = inner > 1 ? (long long)(inner / OR_V) : 0;  /* column blocks per outer */
    const long long nblocks = blocked ? (inner > 1 ? (long long)outer * cols_v : lines / OR_V) : 0;
#pragma omp parallel num_threads(or_threads())
    {
      R* line = (R*)malloc(sizeof(R) * 2 * len * OR_V);
      R* work = (R*)malloc(sizeof(R) * 2 * len * OR_V);
#pragma omp for schedule(static)
      for (long long bl = 0; bl < nblocks; ++bl) {
        if (inner > 1) {
          const size_t o = (size_t)(bl / cols_v);
          const size_t in = (size_t)(bl % cols_v) * OR_V;
          R* base = data + 2 * (o * len * inner + in);
          for (size_t k = 0; k < len; ++k) memcpy(line + 2 * k * OR_V, base + 2 * k * inner, sizeof(R) * 2 * OR_V);
          F(fft_lines_v)(plan, line, work, forward);
          for (size_t k = 0; k < len; ++k) memcpy(base + 2 * k * inner, line + 2 * k * OR_V, sizeof(R) * 2 * OR_V);
        } else {
          R* base = data + 2 * (size_t)bl * OR_V * len;
          for (int v = 0; v < OR_V; ++v) {
            for (size_t k = 0; k < len; ++k) {
              line[2 * (k * OR_V + v)] = base[2 * ((size_t)v * len + k)];
              line[2 * (k * OR_V + v) + 1] = base[2 * ((size_t)v * len + k) + 1];
            }
          }
          F(fft_lines_v)(plan, line, work, forward);
          for (int v = 0; v < OR_V; ++v) {
            for (size_t k = 0; k < len; ++k) {
              base[2 * ((size_t)v * len + k)] = line[2 * (k * OR_V + v)];
              base[2 * ((size_t)v * len + k) + 1] = line[2 * (k * OR_V + v) + 1];
            }
          }
        }
      }
      /* the lines no block covered */
#pragma omp for schedule(static)
      for (long long l = 0; l < lines; ++l) {
        const size_t o = (size_t)l / inner;
        const size_t in = (size_t)l % inner;
        if (blocked && (inner > 1 ? (long long)in < cols_v * OR_V : l < nblocks * OR_V)) continue;
        R* base = data + 2 * (o * len * inner + in);
        if (inner == 1) {
          memcpy(line, base, sizeof(R) * 2 * len);
        } else {
          for (size_t k = 0; k < len; ++k) {
            line[2 * k] = base[2 * k * inner];
            line[2 * k + 1] = base[2 * k * inner + 1];
          }
        }
        F(fft_line)(plan, line, work, forward);
        if (inner == 1) {
          memcpy(base, line, sizeof(R) * 2 * len);
        } else {
          for (size_t k = 0; k < len; ++k) {
            base[2 * k * inner] = line[2 * k];
            base[2 * k * inner + 1] = line[2 * k + 1];
          }
        }
      }
      free(line);
      free(work);
    }
    F(plan_destroy)(plan);
  }
  const R scale = (R)1 / sqrt((R)total);
  const long long n2 = (long long)(2 * total);
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long i = 0; i < n2; ++i) data[i] *= scale;
}

/* ---------------------------------------------------------------------------------------
 * Element-wise kernels, src/cpu/kernels.cpp.  Reductions are summed in fixed blocks and
 * then in block order, so the result does not depend on the thread count.
 * ------------------------------------------------------------------------------------- */

/* src/cpu/kernels.cpp:16-23 */
double F(or_usable_energy)(const R* amplitude, const uint8_t* mask, size_t n) {
  const long long nb = (long long)((n + OR_BLOCK - 1) / OR_BLOCK);
  double* part = (double*)calloc((size_t)(nb ? nb : 1), sizeof(double));
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long b = 0; b < nb; ++b) {
    double s = 0.0;
    const size_t end = ((size_t)b + 1) * OR_BLOCK < n ? ((size_t)b + 1) * OR_BLOCK : n;
    for (size_t i = (size_t)b * OR_BLOCK; i < end; ++i) {
      if (!mask[i]) s += (double)amplitude[i] * (double)amplitude[i];
    }
    part[b] = s;
  }
  double sum = 0.0;
  for (long long b = 0; b < nb; ++b) sum += part[b];
  free(part);
  return sum;
}

/* src/cpu/kernels.cpp:25-42: mag = sqrt(|w|^2); usable pixels add (mag - A)^2 to the
 * error and take the magnitude A; masked ones keep their own.  w *= target/(mag + 1e-12). */
double F(or_modulus_num)(R* wave, const R* amplitude, const uint8_t* mask, size_t n) {
  const long long nb = (long long)((n + OR_BLOCK - 1) / OR_BLOCK);
  double* part = (double*)calloc((size_t)(nb ? nb : 1), sizeof(double));
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long b = 0; b < nb; ++b) {
    double s = 0.0;
    const size_t end = ((size_t)b + 1) * OR_BLOCK < n ? ((size_t)b + 1) * OR_BLOCK : n;
    for (size_t i = (size_t)b * OR_BLOCK; i < end; ++i) {
      const R re = wave[2 * i], im = wave[2 * i + 1];
      const R mag = sqrt(re * re + im * im);
      R target = mag;
      if (!mask[i]) {
        const R d = mag - amplitude[i];
        s += (double)d * (double)d;
        target = amplitude[i];
      }
      const R f = target / (mag + (R)OR_EPS);
      wave[2 * i] = re * f;
      wave[2 * i + 1] = im * f;
    }
    part[b] = s;
  }
  double sum = 0.0;
  for (long long b = 0; b < nb; ++b) sum += part[b];
  free(part);
  return sum;
}

double F(or_modulus)(R* wave, const R* amplitude, const uint8_t* mask, size_t n, double energy) {
  const double num = F(or_modulus_num)(wave, amplitude, mask, n);
  return OR_ROUND_ERROR(sqrt(num / (energy + OR_EPS)));
}

/* src/cpu/kernels.cpp:44-50 with the coefficients of src/core/pointwise.hpp:34-45 */
void F(or_update)(R* x, const R* projected, const uint8_t* support, size_t n, float a, float b) {
  const R ca = (R)a, cb = (R)b;
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    if (support[i]) {
      x[2 * i] = projected[2 * i];
      x[2 * i + 1] = projected[2 * i + 1];
    } else {
      x[2 * i] = ca * x[2 * i] + cb * projected[2 * i];
      x[2 * i + 1] = ca * x[2 * i + 1] + cb * projected[2 * i + 1];
    }
  }
}

/* src/core/pointwise.hpp:48-51 */
static R F(fftfreq)(size_t k, size_t n) {
  const long s = k < (n + 1) / 2 ? (long)k : (long)k - (long)n;
  return (R)s / (R)n;
}

/* src/cpu/kernels.cpp:52-75: unitary FFT, multiply by exp(-2 pi^2 sigma^2 |f|^2), unitary
 * inverse, real part. */
void F(or_blur)(const R* image, R* out, int ndim, const size_t* shape, float sigma) {
  size_t d[3];
  or_dims3(ndim, shape, d);
  const size_t n = d[0] * d[1] * d[2];
  R* buf = (R*)malloc(sizeof(R) * 2 * n);
  for (size_t i = 0; i < n; ++i) {
    buf[2 * i] = image[i];
    buf[2 * i + 1] = 0;
  }
  F(or_fft)(buf, ndim, shape, 1);
  const R pi = (R)OR_TWO_PI / 2;
  const R k = (R)-2 * pi * pi * (R)sigma * (R)sigma;
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long z = 0; z < (long long)d[0]; ++z) {
    const R fz = F(fftfreq)((size_t)z, d[0]);
    size_t i = (size_t)z * d[1] * d[2];
    for (size_t y = 0; y < d[1]; ++y) {
      const R fy = F(fftfreq)(y, d[1]);
      for (size_t x = 0; x < d[2]; ++x, ++i) {
        const R fx = F(fftfreq)(x, d[2]);
        const R g = exp(k * (fz * fz + fy * fy + fx * fx));
        buf[2 * i] *= g;
        buf[2 * i + 1] *= g;
      }
    }
  }
  F(or_fft)(buf, ndim, shape, 0);
  for (size_t i = 0; i < n; ++i) out[i] = buf[2 * i];
  free(buf);
}

/* src/cpu/kernels.cpp:77-102: level = max of the blurred magnitude, or its rms over the
 * current support (summed in double); support = blurred > threshold * level (strict). */
void F(or_shrinkwrap)(const R* x, int ndim, const size_t* shape, float threshold, float sigma,
                      int method, uint8_t* support) {
  size_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
  R* mag = (R*)malloc(sizeof(R) * n);
  for (size_t i = 0; i < n; ++i) mag[i] = sqrt(x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
  F(or_blur)(mag, mag, ndim, shape, sigma);
  R level = 0;
  if (method == OR_MAX) {
    level = mag[0];
    for (size_t i = 1; i < n; ++i) level = mag[i] > level ? mag[i] : level;
  } else {
    const long long nb = (long long)((n + OR_BLOCK - 1) / OR_BLOCK);
    double sum = 0.0, count = 0.0;
    for (long long b = 0; b < nb; ++b) {
      double s = 0.0, c = 0.0;
      const size_t end = ((size_t)b + 1) * OR_BLOCK < n ? ((size_t)b + 1) * OR_BLOCK : n;
      for (size_t i = (size_t)b * OR_BLOCK; i < end; ++i) {
        if (support[i]) {
          s += (double)mag[i] * (double)mag[i];
          c += 1.0;
        }
      }
      sum += s;
      count += c;
    }
    level = (R)sqrt(sum / (count > 1.0 ? count : 1.0));
  }
  const R cut = (R)threshold * level;
  for (size_t i = 0; i < n; ++i) support[i] = mag[i] > cut;
  free(mag);
}

/* src/cpu/kernels.cpp:104-117: A e^{i 2 pi u}, unitary inverse FFT, cut to the support. */
void F(or_random_start)(const R* amplitude, const uint8_t* support, int ndim, const size_t* shape,
                        uint64_t seed, uint64_t start, R* x) {
  size_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
#pragma omp parallel for num_threads(or_threads()) schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    const R phase = (R)OR_TWO_PI * (R)or_uniform(seed, start, (uint64_t)i);
    x[2 * i] = amplitude[i] * cos(phase);
    x[2 * i + 1] = amplitude[i] * sin(phase);
  }
  F(or_fft)(x, ndim, shape, 0);
  for (size_t i = 0; i < n; ++i) {
    if (!support[i]) x[2 * i] = x[2 * i + 1] = 0;
  }
}

/* src/core/problem.cpp:72-88: |unitary IFFT(intensity)| (std::abs, i.e. hypot) thresholded
 * at threshold * max. */
void F(or_autocorrelation_support)(const float* intensity, int ndim, const size_t* shape,
                                   float threshold, uint8_t* support) {
  size_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
  float* ordered = (float*)malloc(sizeof(float) * n);
  or_roll(intensity, ordered, sizeof(float), ndim, shape, 1);
  R* buf = (R*)malloc(sizeof(R) * 2 * n);
  for (size_t i = 0; i < n; ++i) {
    buf[2 * i] = (R)ordered[i];
    buf[2 * i + 1] = 0;
  }
  F(or_fft)(buf, ndim, shape, 0);
  R* ac = (R*)malloc(sizeof(R) * n);
  for (size_t i = 0; i < n; ++i) ac[i] = hypot(buf[2 * i], buf[2 * i + 1]);
  R* centered = (R*)malloc(sizeof(R) * n);
  or_roll(ac, centered, sizeof(R), ndim, shape, 0);
  R peak = centered[0];
  for (size_t i = 1; i < n; ++i) peak = centered[i] > peak ? centered[i] : peak;
  const R level = (R)threshold * peak;
  for (size_t i = 0; i < n; ++i) support[i] = centered[i] > level;
  free(ordered);
  free(buf);
  free(ac);
  free(centered);
}

/* src/core/problem.cpp:90-99: |centred unitary FFT(object)|^2 */
void F(or_simulate_intensity)(const R* object, int ndim, const size_t* shape, R* intensity) {
  size_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
  R* wave = (R*)malloc(sizeof(R) * 2 * n);
  or_roll(object, wave, 2 * sizeof(R), ndim, shape, 1);
  F(or_fft)(wave, ndim, shape, 1);
  R* centered = (R*)malloc(sizeof(R) * 2 * n);
  or_roll(wave, centered, 2 * sizeof(R), ndim, shape, 0);
  for (size_t i = 0; i < n; ++i) {
    intensity[i] = centered[2 * i] * centered[2 * i] + centered[2 * i + 1] * centered[2 * i + 1];
  }
  free(wave);
  free(centered);
}

/* ---------------------------------------------------------------------------------------
 * aligned_error, src/core/metrics.cpp:17-86
 * ------------------------------------------------------------------------------------- */

static double F(translated_error)(const R* b, const R* ref, int ndim, const size_t* shape) {
  size_t d[3];
  or_dims3(ndim, shape, d);
  const size_t n = d[0] * d[1] * d[2];
  R* fr = (R*)malloc(sizeof(R) * 2 * n);
  R* fb = (R*)malloc(sizeof(R) * 2 * n);
  memcpy(fr, ref, sizeof(R) * 2 * n);
  memcpy(fb, b, sizeof(R) * 2 * n);
  F(or_fft)(fr, ndim, shape, 1);
  F(or_fft)(fb, ndim, shape, 1);
  for (size_t i = 0; i < n; ++i) { /* fr *= conj(fb) */
    const R ar = fr[2 * i], ai = fr[2 * i + 1];
    const R br = fb[2 * i], bi = -fb[2 * i + 1];
    fr[2 * i] = ar * br - ai * bi;
    fr[2 * i + 1] = ar * bi + ai * br;
  }
  F(or_fft)(fr, ndim, shape, 0);
  size_t peak = 0;
  R best = fr[0] * fr[0] + fr[1] * fr[1];
  for (size_t i = 1; i < n; ++i) { /* first maximum of |.|^2, like std::ranges::max_element */
    const R v = fr[2 * i] * fr[2 * i] + fr[2 * i + 1] * fr[2 * i + 1];
    if (v > best) {
      best = v;
      peak = i;
    }
  }
  const size_t s0 = peak / (d[1] * d[2]);
  const size_t s1 = (peak / d[2]) % d[1];
  const size_t s2 = peak % d[2];
  double ovr = 0.0, ovi = 0.0, ref_norm = 0.0;
  R* shifted = fb;
  size_t i = 0;
  for (size_t z = 0; z < d[0]; ++z)
    for (size_t y = 0; y < d[1]; ++y)
      for (size_t x = 0; x < d[2]; ++x, ++i) {
        const size_t src = (((z + d[0] - s0) % d[0]) * d[1] + (y + d[1] - s1) % d[1]) * d[2] +
                           (x + d[2] - s2) % d[2];
        shifted[2 * i] = b[2 * src];
        shifted[2 * i + 1] = b[2 * src + 1];
      }
  for (size_t k = 0; k < n; ++k) { /* overlap += ref * conj(shifted) */
    const double rr = ref[2 * k], ri = ref[2 * k + 1];
    const double sr = shifted[2 * k], si = shifted[2 * k + 1];
    ovr += rr * sr + ri * si;
    ovi += ri * sr - rr * si;
    ref_norm += rr * rr + ri * ri;
  }
  const double mag = sqrt(ovr * ovr + ovi * ovi);
  const double cr = mag > 0.0 ? ovr / mag : 1.0;
  const double ci = mag > 0.0 ? ovi / mag : 0.0;
  double diff = 0.0;
  for (size_t k = 0; k < n; ++k) {
    const double sr = shifted[2 * k], si = shifted[2 * k + 1];
    const double er = cr * sr - ci * si - (double)ref[2 * k];
    const double ei = cr * si + ci * sr - (double)ref[2 * k + 1];
    diff += er * er + ei * ei;
  }
  free(fr);
  free(fb);
  return OR_ROUND_ERROR(sqrt(diff / (ref_norm > DBL_MIN ? ref_norm : DBL_MIN)));
}

double F(or_aligned_error)(const R* a, const R* reference, int ndim, const size_t* shape) {
  size_t d[3];
  or_dims3(ndim, shape, d);
  const size_t n = d[0] * d[1] * d[2];
  R* twin = (R*)malloc(sizeof(R) * 2 * n);
  size_t i = 0;
  for (size_t z = 0; z < d[0]; ++z)
    for (size_t y = 0; y < d[1]; ++y)
      for (size_t x = 0; x < d[2]; ++x, ++i) {
        const size_t src =
            (((d[0] - z) % d[0]) * d[1] + (d[1] - y) % d[1]) * d[2] + (d[2] - x) % d[2];
        twin[2 * i] = a[2 * src];
        twin[2 * i + 1] = -a[2 * src + 1];
      }
  const double e0 = F(translated_error)(a, reference, ndim, shape);
  const double e1 = F(translated_error)(twin, reference, ndim, shape);
  free(twin);
  return e0 < e1 ? e0 : e1;
}

/* ---------------------------------------------------------------------------------------
 * The solver, src/cpu/cpu_solver.cpp:18-110
 * ------------------------------------------------------------------------------------- */

typedef struct {
  int ndim;
  size_t shape[3];
  size_t n;
  int starts;
  R* amplitude; /* FFT order */
  uint8_t* mask;
  double energy;
  R* x;          /* starts * n complex, FFT order */
  uint8_t* supp; /* starts * n */
  uint8_t* saved;
  R* wave;
  double* errors; /* starts * cap */
  int iterations;
  int cap;
} F(solver_t);

void* F(or_solver_create)(int ndim, const size_t* shape, const float* amplitude,
                          const uint8_t* mask, const uint8_t* support, int starts, uint64_t seed,
                          const float* start) {
  F(solver_t)* s = (F(solver_t)*)calloc(1, sizeof(F(solver_t)));
  s->ndim = ndim;
  s->n = 1;
  for (int a = 0; a < ndim; ++a) {
    s->shape[a] = shape[a];
    s->n *= shape[a];
  }
  const size_t n = s->n;
  s->starts = starts;
  float* tmp = (float*)malloc(sizeof(float) * n);
  or_roll(amplitude, tmp, sizeof(float), ndim, shape, 1);
  s->amplitude = (R*)malloc(sizeof(R) * n);
  for (size_t i = 0; i < n; ++i) s->amplitude[i] = (R)tmp[i];
  free(tmp);
  s->mask = (uint8_t*)malloc(n);
  or_roll(mask, s->mask, 1, ndim, shape, 1);
  s->energy = F(or_usable_energy)(s->amplitude, s->mask, n);
  s->supp = (uint8_t*)malloc(n * (size_t)starts);
  s->saved = (uint8_t*)malloc(n * (size_t)starts);
  s->x = (R*)malloc(sizeof(R) * 2 * n * (size_t)starts);
  s->wave = (R*)malloc(sizeof(R) * 2 * n);
  uint8_t* sup = (uint8_t*)malloc(n);
  or_roll(support, sup, 1, ndim, shape, 1);
  float* given = start ? (float*)malloc(sizeof(float) * 2 * n) : NULL;
  for (int k = 0; k < starts; ++k) {
    uint8_t* sk = s->supp + (size_t)k * n;
    R* xk = s->x + 2 * (size_t)k * n;
    memcpy(sk, sup, n);
    if (!start) {
      F(or_random_start)(s->amplitude, sup, ndim, shape, seed, (uint64_t)k, xk);
      continue;
    }
    or_roll(start + 2 * (size_t)k * n, given, 2 * sizeof(float), ndim, shape, 1);
    for (size_t i = 0; i < n; ++i) {
      xk[2 * i] = sup[i] ? (R)given[2 * i] : 0;
      xk[2 * i + 1] = sup[i] ? (R)given[2 * i + 1] : 0;
    }
  }
  free(sup);
  free(given);
  return s;
}

/* src/core/recipe.cpp:30-61 */
static int F(validate)(const or_op* ops, int nops) {
  if (nops <= 0) return -1;
  int in_detwin = 0;
  for (int i = 0; i < nops; ++i) {
    const or_op* o = &ops[i];
    switch (o->kind) {
      case OR_ER:
      case OR_HIO:
      case OR_RAAR:
        if (o->n < 1) return -1;
        break;
      case OR_SHRINKWRAP:
        if (!(o->threshold > 0.0f) || !(o->sigma >= 0.0f)) return -1;
        break;
      case OR_DETWIN_BEGIN:
        if (in_detwin) return -1;
        in_detwin = 1;
        break;
      case OR_DETWIN_END:
        if (!in_detwin) return -1;
        in_detwin = 0;
        break;
      default:
        return -1;
    }
  }
  return in_detwin ? -1 : 0;
}

static void F(grow)(F(solver_t) * s, int more) {
  const int need = s->iterations + more;
  if (need <= s->cap) return;
  int cap = s->cap ? s->cap : 64;
  while (cap < need) cap *= 2;
  double* e = (double*)calloc((size_t)cap * (size_t)s->starts, sizeof(double));
  for (int k = 0; k < s->starts && s->iterations > 0; ++k) {  /* no history yet: errors is NULL */
    memcpy(e + (size_t)k * cap, s->errors + (size_t)k * s->cap, sizeof(double) * s->iterations);
  }
  free(s->errors);
  s->errors = e;
  s->cap = cap;
}

/* cpu_solver.cpp:82-93: per iteration and start: forward FFT, modulus (error of the
 * iterate entering the iteration), inverse FFT, update. */
static void F(iterate)(F(solver_t) * s, int iters, float a, float b) {
  F(grow)(s, iters);
  const size_t n = s->n;
  for (int it = 0; it < iters; ++it) {
    for (int k = 0; k < s->starts; ++k) {
      R* xk = s->x + 2 * (size_t)k * n;
      memcpy(s->wave, xk, sizeof(R) * 2 * n);
      F(or_fft)(s->wave, s->ndim, s->shape, 1);
      s->errors[(size_t)k * s->cap + s->iterations] =
          F(or_modulus)(s->wave, s->amplitude, s->mask, n, s->energy);
      F(or_fft)(s->wave, s->ndim, s->shape, 0);
      F(or_update)(xk, s->wave, s->supp + (size_t)k * n, n, a, b);
    }
    s->iterations++;
  }
}

int F(or_solver_run)(void* solver, const or_op* ops, int nops) {
  F(solver_t)* s = (F(solver_t)*)solver;
  if (F(validate)(ops, nops) != 0) return -1;
  const size_t n = s->n;
  for (int i = 0; i < nops; ++i) {
    const or_op* o = &ops[i];
    switch (o->kind) {
      case OR_ER:
        F(iterate)(s, o->n, 0.0f, 0.0f);
        break;
      case OR_HIO:
        F(iterate)(s, o->n, 1.0f, -o->beta);
        break;
      case OR_RAAR:
        F(iterate)(s, o->n, o->beta, 1.0f - 2.0f * o->beta);
        break;
      case OR_SHRINKWRAP:
        for (int k = 0; k < s->starts; ++k) {
          F(or_shrinkwrap)(s->x + 2 * (size_t)k * n, s->ndim, s->shape, o->threshold, o->sigma,
                           o->method, s->supp + (size_t)k * n);
        }
        break;
      case OR_DETWIN_BEGIN: { /* cpu_solver.cpp:70-76 */
        uint8_t* centered = (uint8_t*)malloc(n);
        uint8_t* half = (uint8_t*)malloc(n);
        memcpy(s->saved, s->supp, n * (size_t)s->starts);
        for (int k = 0; k < s->starts; ++k) {
          or_roll(s->supp + (size_t)k * n, centered, 1, s->ndim, s->shape, 0);
          or_half_support(centered, half, s->ndim, s->shape, o->axis);
          or_roll(half, s->supp + (size_t)k * n, 1, s->ndim, s->shape, 1);
        }
        free(centered);
        free(half);
        break;
      }
      case OR_DETWIN_END:
        memcpy(s->supp, s->saved, n * (size_t)s->starts);
        break;
    }
  }
  return 0;
}

int F(or_solver_iterations)(const void* solver) { return ((const F(solver_t)*)solver)->iterations; }

void F(or_solver_results)(const void* solver, R* objects, uint8_t* supports, double* errors) {
  const F(solver_t)* s = (const F(solver_t)*)solver;
  const size_t n = s->n;
  for (int k = 0; k < s->starts; ++k) {
    if (objects) {
      or_roll(s->x + 2 * (size_t)k * n, objects + 2 * (size_t)k * n, 2 * sizeof(R), s->ndim,
              s->shape, 0);
    }
    if (supports) {
      or_roll(s->supp + (size_t)k * n, supports + (size_t)k * n, 1, s->ndim, s->shape, 0);
    }
    if (errors) {
      memcpy(errors + (size_t)k * s->iterations, s->errors + (size_t)k * s->cap,
             sizeof(double) * s->iterations);
    }
  }
}

void F(or_solver_get_state)(const void* solver, R* x, uint8_t* supports) {
  const F(solver_t)* s = (const F(solver_t)*)solver;
  if (x) memcpy(x, s->x, sizeof(R) * 2 * s->n * (size_t)s->starts);
  if (supports) memcpy(supports, s->supp, s->n * (size_t)s->starts);
}

void F(or_solver_set_state)(void* solver, const R* x, const uint8_t* supports) {
  F(solver_t)* s = (F(solver_t)*)solver;
  if (x) memcpy(s->x, x, sizeof(R) * 2 * s->n * (size_t)s->starts);
  if (supports) memcpy(s->supp, supports, s->n * (size_t)s->starts);
}

void F(or_solver_destroy)(void* solver) {
  F(solver_t)* s = (F(solver_t)*)solver;
  if (!s) return;
  free(s->amplitude);
  free(s->mask);
  free(s->x);
  free(s->supp);
  free(s->saved);
  free(s->wave);
  free(s->errors);
  free(s);
}

#undef F
#undef OR_CAT
#undef OR_CAT_
