"""Known answers for the oracle's usable energy, Shrinkwrap level and threshold, random
start and autocorrelation support (src/cpu/kernels.cpp:16-23,77-117,
src/core/problem.cpp:72-88).

Each expectation is computed here from the mathematics, not by retyping the oracle's
formula: a numpy sum, the Gaussian blur that tests/test_oracle.py pins to the separable
circular convolution, a brute-force circular autocorrelation, and the spectrum a unitary
inverse transform must have.  The cases are chosen so that a plausible slip fails one of
them: squaring the wrong thing, `>=` for `>` (a pixel sits exactly on the cut), the rms
over every pixel instead of the current support, a forward transform for the inverse,
dropped or conjugated phases, a roll in the wrong direction on odd lengths.
"""

import numpy as np
import pytest

import oracle

# Relative distance from the cut that float rounding cannot cross (the oracle's blur and
# transforms are within ~1e-6 of exact; cases below assert they keep this margin).
MARGIN = 1e-4


def _mag(x):
    """|x| as the oracle and the reference evaluate it: sqrt(re*re + im*im) in float."""
    re = x.real.astype(np.float32)
    im = x.imag.astype(np.float32)
    return np.sqrt(re * re + im * im)


def _field(shape, seed, prec=np.complex64):
    """Random complex values under a Gaussian envelope about an off-centre point, so that
    a blurred-magnitude cut falls inside the grid."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    r2 = sum(((g - 0.4 * s) / (0.25 * s)) ** 2 for g, s in zip(grids, shape))
    noise = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return (np.exp(-0.5 * r2) * (1.0 + 0.5 * noise)).astype(prec)


def _margin(values, cut):
    """Smallest relative distance of any value from the cut."""
    return float(np.min(np.abs(values.astype(np.float64) - float(cut))) / abs(float(cut)))


# --- usable energy (src/cpu/kernels.cpp:16-23) ---------------------------------------------


@pytest.mark.parametrize("prec", ["f", "d"])
def test_usable_energy_is_the_unmasked_sum_of_squares(prec):
    rng = np.random.default_rng(1)
    a = rng.uniform(0.0, 3.0, 5000).astype(np.float32 if prec == "f" else np.float64)
    m = (rng.random(5000) < 0.3).astype(np.uint8)
    want = float(np.sum(a.astype(np.float64) ** 2 * (m == 0)))
    got = oracle.usable_energy(a, m, prec)
    assert abs(got - want) <= 1e-12 * want
    # a masked pixel contributes nothing, an unmasked one its squared amplitude
    assert oracle.usable_energy(a, np.ones_like(m), prec) == 0.0
    one = np.zeros_like(a)
    one[17] = 3.0
    assert oracle.usable_energy(one, np.zeros_like(m), prec) == 9.0


# --- Shrinkwrap (src/cpu/kernels.cpp:77-102) -------------------------------------------------


SW_SHAPES = [(12, 10), (9, 7), (6, 5, 4), (16, 16)]


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_shrinkwrap_max_level_is_the_blurred_maximum(shape):
    x = _field(shape, 2)
    support = (np.random.default_rng(3).random(shape) < 0.4).astype(np.uint8)
    blurred = oracle.blur(_mag(x), 1.0)
    tau = np.float32(0.3)
    cut = tau * blurred.max()
    assert _margin(blurred, cut) > MARGIN
    want = (blurred > cut).astype(np.uint8)
    assert 0 < want.sum() < want.size
    got = oracle.shrinkwrap(x, support, threshold=0.3, sigma=1.0, method="max")
    assert np.array_equal(got, want)
    # the level ignores the current support: a different one gives the same answer
    assert np.array_equal(oracle.shrinkwrap(x, 1 - support, 0.3, 1.0, "max"), want)
    # and it is a magnitude blur: the blurred |x|^2 would give another support
    sq = oracle.blur(_mag(x) ** 2, 1.0)
    assert not np.array_equal(want, (sq > tau * sq.max()).astype(np.uint8))


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_shrinkwrap_cut_is_strict(shape):
    """threshold 1 puts the blurred maximum exactly on the cut: a strict comparison keeps
    nothing (the maximum), `>=` would keep the maximum's pixels."""
    x = _field(shape, 4)
    support = np.ones(shape, np.uint8)
    got = oracle.shrinkwrap(x, support, threshold=1.0, sigma=1.0, method="max")
    assert got.sum() == 0


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_shrinkwrap_rms_is_over_the_current_support(shape):
    x = _field(shape, 5)
    # the current support: the brighter 40% of the pixels, so its rms is well above the
    # rms over the whole grid
    support = (_mag(x) > np.quantile(_mag(x), 0.6)).astype(np.uint8)
    blurred = oracle.blur(_mag(x), 1.0)
    b = blurred.astype(np.float64)
    level = np.float32(np.sqrt(np.sum(b[support == 1] ** 2) / support.sum()))
    cut = np.float32(0.8) * level
    assert _margin(blurred, cut) > MARGIN
    want = (blurred > cut).astype(np.uint8)
    assert 0 < want.sum() < want.size
    got = oracle.shrinkwrap(x, support, threshold=0.8, sigma=1.0, method="rms")
    assert np.array_equal(got, want)
    # the rms over every pixel would cut elsewhere
    level_all = np.float32(np.sqrt(np.mean(b**2)))
    assert not np.array_equal(want, (blurred > np.float32(0.8) * level_all).astype(np.uint8))


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_shrinkwrap_rms_of_one_pixel_is_that_pixel(shape):
    """A one-pixel support's rms is that pixel's blurred value, exactly (the float value
    squared in double is exact, and so is its square root).  At threshold 1 the pixel sits
    on the cut: the strict comparison drops it and keeps exactly the larger values."""
    x = _field(shape, 7)
    blurred = oracle.blur(_mag(x), 1.0)
    i = int(np.argsort(blurred, axis=None)[blurred.size // 2])  # the median pixel
    support = np.zeros(shape, np.uint8)
    support.flat[i] = 1
    got = oracle.shrinkwrap(x, support, threshold=1.0, sigma=1.0, method="rms")
    want = (blurred > blurred.flat[i]).astype(np.uint8)
    assert got.flat[i] == 0
    assert np.array_equal(got, want)
    assert 0 < got.sum() < got.size


def test_shrinkwrap_rms_of_an_empty_support_is_zero():
    """An empty support's count is taken as 1 (std::max(count, 1.0)): the level is 0 and
    every pixel with a positive blurred magnitude is kept."""
    x = _field((12, 10), 8)
    blurred = oracle.blur(_mag(x), 1.0)
    got = oracle.shrinkwrap(x, np.zeros((12, 10), np.uint8), threshold=0.3, sigma=1.0, method="rms")
    assert np.array_equal(got, (blurred > 0).astype(np.uint8))
    assert got.sum() == got.size


# --- random start (src/cpu/kernels.cpp:104-117) -----------------------------------------------


@pytest.mark.parametrize("shape", [(8, 6), (7, 9), (5, 4, 3)])
@pytest.mark.parametrize("seed,start", [(0, 0), (3, 5), (2**40 + 7, 1)])
def test_random_start_has_the_amplitude_and_hashed_phases(shape, seed, start):
    """Before the cut, x0 is the unitary inverse transform of A e^{i 2 pi u}, u the
    reference's hash (pinned against its compiled pointwise.hpp in test_oracle.py): its
    unitary forward transform gives back exactly that spectrum."""
    n = int(np.prod(shape))
    rng = np.random.default_rng(9)
    amp = rng.uniform(0.2, 2.0, shape)
    ones = np.ones(shape, np.uint8)
    u = np.array([oracle.uniform(seed, start, i) for i in range(n)], np.float64).reshape(shape)
    want = amp * np.exp(2j * np.pi * u)
    for prec, tol in (("d", 1e-6), ("f", 2e-5)):
        x0 = oracle.random_start(amp, ones, seed, start, prec)
        spec = np.fft.fftn(x0.astype(np.complex128), norm="ortho")
        # u is float (24 bits), the phase 2 pi u is then evaluated at the oracle's precision
        assert np.max(np.abs(spec - want)) < tol * amp.max(), (prec, np.max(np.abs(spec - want)))
        assert np.allclose(np.abs(spec), amp, rtol=tol)
    # the cut: zero outside the support, unchanged inside
    sup = (rng.random(shape) < 0.5).astype(np.uint8)
    full = oracle.random_start(amp, ones, seed, start, "d")
    cut = oracle.random_start(amp, sup, seed, start, "d")
    assert np.array_equal(cut, np.where(sup == 1, full, 0))


def test_random_starts_differ_by_start_and_seed():
    amp = np.ones((8, 8))
    ones = np.ones((8, 8), np.uint8)
    a = oracle.random_start(amp, ones, 0, 0, "d")
    assert not np.allclose(a, oracle.random_start(amp, ones, 0, 1, "d"))
    assert not np.allclose(a, oracle.random_start(amp, ones, 1, 0, "d"))


# --- autocorrelation support (src/core/problem.cpp:72-88) -------------------------------------


def _brute_autocorrelation(obj):
    """|sum_j o[j + d] conj(o[j])| over circular displacements d, centred (d = 0 at index
    n // 2 of each axis), by explicit sums in double."""
    o = obj.astype(np.complex128)
    shape = o.shape
    n = o.size
    out = np.empty(shape)
    for c in np.ndindex(*shape):
        d = tuple((ci - s // 2) % s for ci, s in zip(c, shape))
        shifted = np.roll(o, tuple(-di for di in d), axis=tuple(range(o.ndim)))  # shifted[j] = o[j + d]
        out[c] = abs(np.sum(shifted * np.conj(o)))
    assert n == out.size
    return out


def _object(shape, seed):
    """A random complex object inside a box of about half the grid (so the autocorrelation
    has structure), centre-origin."""
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape, np.complex128)
    box = tuple(slice(s // 4, s // 4 + max(2, s // 2)) for s in shape)
    sub = obj[box]
    obj[box] = rng.uniform(0.3, 1.0, sub.shape) * np.exp(1j * rng.uniform(-np.pi, np.pi, sub.shape))
    return obj


@pytest.mark.parametrize("shape,seed", [((6, 5), 1), ((7, 8), 2), ((9, 7), 3), ((5, 4, 3), 4), ((3, 5, 7), 5)])
def test_autocorrelation_support_matches_brute_force(shape, seed):
    obj = _object(shape, seed)
    # |centred unitary F(object)|^2, in double, handed over as float like an intensity
    inten = np.abs(np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(obj), norm="ortho"))) ** 2
    inten = inten.astype(np.float32)
    ac = _brute_autocorrelation(obj)
    for tau in (0.1, 0.35):
        cut = tau * ac.max()
        assert _margin(ac, cut) > MARGIN, (shape, tau)
        want = (ac > cut).astype(np.uint8)
        got = oracle.autocorrelation_support(inten, tau)
        assert np.array_equal(got, want), (shape, tau)
        # not symmetric under a one-pixel roll on odd axes: the centring direction matters
        for ax, s in enumerate(shape):
            if s % 2:
                assert not np.array_equal(got, np.roll(want, 1, axis=ax))
    # threshold 1: the peak (zero displacement) sits exactly on the cut and is dropped
    assert oracle.autocorrelation_support(inten, 1.0).sum() == 0
