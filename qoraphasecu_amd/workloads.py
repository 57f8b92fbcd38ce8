"""Seeded synthetic inputs of the benchmark configurations (the role of the reference's
bench/common.hpp:28-50 bench_problem, at the sizes and statistics of the five configs).

Every generator is deterministic for a seed (numpy PCG64), returns centre-origin arrays
and never touches the GPU; intensities are formed with numpy's FFT because they are
inputs, not part of the reconstruction path.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def _smooth_phase(rng, shape, sigma):
    """White noise low-passed by a Gaussian of `sigma` pixels, scaled to [-pi, pi]."""
    noise = rng.standard_normal(shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    f2 = sum(f * f for f in freqs)
    smooth = np.fft.ifftn(np.fft.fftn(noise) * np.exp(-2.0 * math.pi**2 * sigma**2 * f2)).real
    peak = np.max(np.abs(smooth))
    return smooth / (peak if peak > 0 else 1.0) * math.pi


def _soft_ellipsoids(rng, shape, box, count, amp_lo, amp_hi, edge=1.0):
    """Amplitude of `count` random soft-edged ellipsoids inside a centred box."""
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    amp = np.zeros(shape)
    for _ in range(count):
        centre = [n / 2 + rng.uniform(-box / 4, box / 4) for n in shape]
        axes = [rng.uniform(box / 10, box / 4) for _ in shape]
        rho = np.sqrt(sum(((g - c) / a) ** 2 for g, c, a in zip(grids, centre, axes)))
        # distance outside the boundary in pixels, erf edge of ~1 px
        dist = (rho - 1.0) * min(axes)
        inside = 0.5 * (1.0 - erf(dist / edge))
        amp = np.maximum(amp, rng.uniform(amp_lo, amp_hi) * inside)
    lo = [int(n // 2 - box // 2) for n in shape]
    keep = np.zeros(shape, bool)
    keep[tuple(slice(l, l + int(box)) for l in lo)] = True
    return np.where(keep, amp, 0.0)


def _intensity(obj):
    """|centred unitary FFT|^2 (numpy, double)."""
    f = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(obj), norm="ortho"))
    return np.abs(f) ** 2


def beamstop(shape, radius):
    """Mask (1 = unusable) of a centred disc of `radius` pixels."""
    grids = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    r2 = sum(g.astype(np.float64) ** 2 for g in grids)
    return (r2 <= radius * radius).astype(np.uint8)


def object_2d(n=256, seed=0, box=None):
    """Config 2 object: 3-6 soft ellipses in a centred box (96 px at 256^2, 3/8 of the
    grid otherwise), amplitude U[0.3, 1], smooth phase (sigma 8 px)."""
    rng = np.random.default_rng(seed)
    box = 96 if box is None else box
    count = int(rng.integers(3, 7))
    amp = _soft_ellipsoids(rng, (n, n), box, count, 0.3, 1.0)
    phase = _smooth_phase(rng, (n, n), 8.0)
    return (amp * np.exp(1j * phase)).astype(np.complex64)


def problem_2d(n=256, seed=0, photons=1e7, beamstop_radius=6, box=None):
    """Config 2: Poisson intensity at `photons` total counts, beamstop mask.
    Returns (object, intensity f32, mask u8)."""
    obj = object_2d(n, seed, box)
    rng = np.random.default_rng(10_000 + seed)
    inten = _intensity(obj.astype(np.complex128))
    inten *= photons / inten.sum()
    counts = rng.poisson(inten).astype(np.float32)
    return obj, counts, beamstop((n, n), beamstop_radius)


def problem_3d(n=128, seed=0, box=48, mask_fraction=0.01):
    """Config 4: random ellipsoids in a 48^3 box, amplitude U[0.4, 1], smooth phase,
    noiseless intensity, Bernoulli(0.01) mask.  Returns (object, intensity, mask)."""
    rng = np.random.default_rng(seed)
    count = int(rng.integers(3, 7))
    amp = _soft_ellipsoids(rng, (n, n, n), box, count, 0.4, 1.0)
    phase = _smooth_phase(rng, (n, n, n), 4.0)
    obj = (amp * np.exp(1j * phase)).astype(np.complex64)
    inten = _intensity(obj.astype(np.complex128)).astype(np.float32)
    mask = (rng.random((n, n, n)) < mask_fraction).astype(np.uint8)
    return obj, inten, mask


def problem_disc(n=64, radius=16, seed=0):
    """Config 1: disc of radius 16 px, amplitude U[0.5, 1], phase U[-pi, pi], noiseless."""
    rng = np.random.default_rng(seed)
    y, x = np.indices((n, n)) - n // 2
    inside = x * x + y * y <= radius * radius
    amp = rng.uniform(0.5, 1.0, (n, n))
    phase = rng.uniform(-math.pi, math.pi, (n, n))
    obj = np.where(inside, amp * np.exp(1j * phase), 0).astype(np.complex64)
    return obj, _intensity(obj.astype(np.complex128)).astype(np.float32)


def ensemble_uneven(per_size=16, sizes=(128, 256, 512), seed=0):
    """Config 3: `per_size` problems at each size, box 3/8 of the grid, photons
    log-uniform in [1e6, 1e8], beamstop radius U{4..10}.  [(object, intensity, mask)]."""
    rng = np.random.default_rng(seed)
    out = []
    k = 0
    for n in sizes:
        for _ in range(per_size):
            photons = 10 ** rng.uniform(6, 8)
            radius = int(rng.integers(4, 11))
            out.append(problem_2d(n, 1000 + k, photons, radius, box=3 * n // 8))
            k += 1
    return out


def raar_shrinkwrap_recipe(repeats, raar_n, er_n, method="max", beta=0.9):
    """repeat(repeats, RAAR(raar_n, beta), Shrinkwrap(method)) then ER(er_n)."""
    from .recipe import ER, RAAR, Recipe, Shrinkwrap, repeat

    return Recipe(repeat(repeats, RAAR(raar_n, beta), Shrinkwrap(method=method)), ER(er_n))


def config(name):
    """The benchmark workloads by name, as dicts (shape, starts, intensity, mask, recipe,
    desc; "problems" for the uneven ensemble):
      c1  64^2 disc, 8 starts, default_recipe, noiseless, no mask
      c2  256^2 ellipses, 32 starts, repeat(10, RAAR(50, 0.9), Shrinkwrap max) + ER(200),
          Poisson 1e7 photons, 6 px beamstop
      c3  48 problems (16 each at 128^2, 256^2, 512^2), 8 starts each, default_recipe
      c4  128^3 ellipsoids, 8 starts, repeat(8, RAAR(50, 0.9), Shrinkwrap rms) + ER(200),
          noiseless, 1% random-pixel mask
    Every problem's support comes from the autocorrelation (threshold 0.1)."""
    from .recipe import default_recipe

    if name == "c1":
        obj, inten = problem_disc(64, radius=16, seed=1000)
        return dict(name="c1", shape=(64, 64), starts=8, object=obj, intensity=inten, mask=None,
                    recipe=default_recipe(),
                    desc="64^2 disc (radius 16), 8 starts, default_recipe, noiseless, no mask, "
                         "autocorrelation support")
    if name == "c2":
        obj, inten, mask = problem_2d(256, seed=1000, photons=1e7, beamstop_radius=6)
        return dict(name="c2", shape=(256, 256), starts=32, object=obj, intensity=inten, mask=mask,
                    recipe=raar_shrinkwrap_recipe(10, 50, 200, method="max"),
                    desc="256^2, 32 starts, repeat(10, RAAR(50, 0.9), Shrinkwrap(max)) + ER(200), "
                         "Poisson 1e7 photons, 6 px beamstop, autocorrelation support")
    if name == "c4":
        obj, inten, mask = problem_3d(128, seed=2000)
        return dict(name="c4", shape=(128, 128, 128), starts=8, object=obj, intensity=inten, mask=mask,
                    recipe=raar_shrinkwrap_recipe(8, 50, 200, method="rms"),
                    desc="128^3, 8 starts, repeat(8, RAAR(50, 0.9), Shrinkwrap(rms)) + ER(200), "
                         "noiseless, 1% random-pixel mask, autocorrelation support")
    if name == "c3":
        problems = ensemble_uneven(per_size=16, sizes=(128, 256, 512), seed=0)
        return dict(name="c3", shape=None, starts=8, problems=problems, recipe=default_recipe(),
                    desc="48 problems (16 each at 128^2, 256^2, 512^2), 8 starts each, default_recipe, "
                         "Poisson 1e6-1e8 photons, 4-10 px beamstops, autocorrelation supports")
    raise ValueError(f"unknown workload {name!r}")
