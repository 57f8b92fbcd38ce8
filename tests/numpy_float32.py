"""A second float32 implementation of the reference's CPU iteration, for the GPU tests'
rounding yardsticks: numpy.fft (pocketfft, the FFT library of the reference's CPU backend,
src/cpu/fft.cpp) with the element-wise steps of src/cpu/kernels.cpp:25-102 and the recipe
walk of src/cpu/cpu_solver.cpp:61-93, all in float32 (sums in float64).

It is as valid a float computation as the oracle's own float run; how far the two land
from each other, and from the FP64 oracle, is the scale that float rounding alone reaches on
a given run.  Test infrastructure only."""

from __future__ import annotations

import numpy as np

import oracle
import recipes as R


def _iterate(x, amp, usable, energy, support, kind, n, beta, errs):
    a, b = (np.float32(c) for c in R.coefficients(kind, beta))
    for _ in range(n):
        w = np.fft.fftn(x, norm="ortho").astype(np.complex64)
        mag = np.sqrt(w.real * w.real + w.imag * w.imag)
        d = (mag - amp)[usable].astype(np.float64)
        errs.append(np.sqrt(np.sum(d * d) / (energy + 1e-12)))
        target = np.where(usable, amp, mag)
        p = np.fft.ifftn(w * (target / (mag + np.float32(1e-12))), norm="ortho").astype(np.complex64)
        x = np.where(support, p, a * x + b * p).astype(np.complex64)
    return x


def _shrinkwrap(x, support, threshold, sigma, method, f2):
    mag = np.sqrt(x.real * x.real + x.imag * x.imag)
    k = np.float32(-2.0 * np.pi**2 * sigma * sigma)
    spectrum = np.fft.fftn(mag.astype(np.complex64), norm="ortho") * np.exp(k * f2)
    blurred = np.fft.ifftn(spectrum, norm="ortho").real.astype(np.float32)
    if method == "max":
        level = blurred.max()
    else:
        inside = blurred.astype(np.float64)[support]
        level = np.float32(np.sqrt(np.sum(inside * inside) / max(int(support.sum()), 1)))
    return blurred > np.float32(threshold) * level


def run(amplitude, mask, support, starts_fft, ops):
    """Error histories (float64 arrays, one per start) of a recipe of 7-field records, from
    the given FFT-order starting iterates (cut to the support), on centre-origin
    amplitude / mask / support."""
    f = oracle.to_fft_order
    amp = f(np.asarray(amplitude, np.float32))
    usable = f(np.asarray(mask)) == 0
    energy = float(np.sum(amp.astype(np.float64)[usable] ** 2))
    freqs = np.meshgrid(*[np.fft.fftfreq(n).astype(np.float32) for n in amp.shape], indexing="ij")
    f2 = sum(q * q for q in freqs).astype(np.float32)
    out = []
    for x0 in starts_fft:
        s = f(np.asarray(support)).astype(bool)
        x = np.where(s, x0, 0).astype(np.complex64)
        saved, errs = None, []
        for kind, n, beta, threshold, sigma, method, axis in ops:
            if kind in ("er", "hio", "raar"):
                x = _iterate(x, amp, usable, energy, s, kind, n, beta, errs)
            elif kind == "shrinkwrap":
                s = _shrinkwrap(x, s, threshold, sigma, method, f2)
            elif kind == "detwin_begin":
                saved = s.copy()
                s = f(oracle.half_support(oracle.to_centered(s.astype(np.uint8)), axis)).astype(bool)
            elif kind == "detwin_end":
                s = saved
        out.append(np.asarray(errs, np.float64))
    return out


def random_starts(amplitude, support, starts, seed=0):
    """The oracle's random starting iterates (FFT order), as its solver draws them."""
    f = oracle.to_fft_order
    return [oracle.random_start(f(amplitude), f(support), seed, k) for k in range(starts)]
