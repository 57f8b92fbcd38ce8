"""Shrinkwrap on the device against the oracle (src/cpu/kernels.cpp:77-102), both level
methods (max, rms over the current support), and the reference's CUDA-vs-CPU recipe check
(tests/test_cuda.cpp:66-87, supports EXPECT_EQ) with an rms Shrinkwrap.

Direct check: a device state after a few HIO iterations, a current support, one device
Shrinkwrap; the oracle's Shrinkwrap of the same (read back) iterate and support must give
the same mask.  The device and the oracle blur in float with different transforms (about
1e-7 apart), so a pixel whose blurred magnitude lies within 1e-5 relative of the cut may
fall either way: every pixel outside that band must match exactly, and the band itself is
reported.  In the recipe check the supports must match exactly, pixel for pixel."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

pytestmark = pytest.mark.gpu

BAND = 1e-5
IMAGE_SHAPES = [(16, 64), (16, 128), (32, 32), (32, 64), (32, 128), (64, 16), (64, 32), (64, 64), (64, 128),
                (128, 16), (128, 32), (128, 64)]


def _mag(x):
    re = x.real.astype(np.float32)
    im = x.imag.astype(np.float32)
    return np.sqrt(re * re + im * im)


def _cut(blurred, support, threshold, method):
    """The cut as the oracle computes it: float max, or the rms over the current support
    summed in double (src/cpu/kernels.cpp:88-100)."""
    if method == "max":
        level = blurred.max()
    else:
        b = blurred.astype(np.float64)
        level = np.float32(np.sqrt(np.sum(b[support != 0] ** 2) / max(1, int(np.count_nonzero(support)))))
    return np.float32(threshold) * np.float32(level)


POW2_SHAPES = [(64, 64), (256, 256), (128, 64), (16, 64, 64), (32, 32, 32), (8, 128, 128)]
GENERIC_SHAPES = [(48, 40), (12, 20, 18)]  # fp32 only: fp16 storage needs power-of-two extents


@pytest.mark.parametrize("shape,storage", [(s, "f32") for s in POW2_SHAPES + GENERIC_SHAPES] +
                         [(s, "f16") for s in POW2_SHAPES])
@pytest.mark.parametrize("method,threshold", [("max", 0.3), ("rms", 0.3), ("rms", 0.8)])
def test_device_shrinkwrap_matches_oracle(qpch, shape, method, threshold, storage):
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, seed=5, storage=storage, schedule="graph")
    s.begin_run(8)
    a, b = R.coefficients("hio", 0.9)
    s.iterate(8, a, b)
    # the current support: start 0 keeps the object's, start 1 takes a looser one (the
    # oracle's max Shrinkwrap of its own iterate at 0.05)
    cur = s.get_support()
    x = s.read_state()
    cur[1] = oracle.shrinkwrap(x[1], cur[1], threshold=0.05, sigma=1.0, method="max")
    s.set_support(cur)
    s.shrinkwrap(threshold, 1.0, method)
    got = s.get_support()
    s.close()
    band = 0
    for k in range(2):
        want = oracle.shrinkwrap(x[k], cur[k], threshold=threshold, sigma=1.0, method=method)
        blurred = oracle.blur(_mag(x[k]), 1.0)
        cut = _cut(blurred, cur[k], threshold, method)
        near = np.abs(blurred.astype(np.float64) - float(cut)) <= BAND * abs(float(cut))
        band += int(near.sum())
        diff = got[k].astype(bool) != want.astype(bool)
        assert not np.any(diff & ~near), (shape, method, k, int((diff & ~near).sum()))
        assert 0 < int(want.sum()) < want.size, (shape, method, k)
    print(f"{shape} {storage} {method} {threshold}: supports equal outside the band; {band} pixel(s) within "
          f"{BAND:g} of the cut")


@pytest.mark.parametrize("shape,schedule", [((64, 64), "graph"), ((64, 48), "graph"), ((16, 64, 64), "graph"),
                                            ((32, 32, 32), "graph"), ((64, 64), "unfused"),
                                            ((16, 64, 64), "fused")] + [(s, "graph") for s in IMAGE_SHAPES
                                                                        if s != (64, 64)])
def test_agrees_with_oracle_rms_recipe(qpch, shape, schedule):
    """tests/test_cuda.cpp:66-87 with Shrinkwrap(rms): HIO15, Shrinkwrap(0.3, 1, rms), RAAR10,
    detwin(HIO5), Shrinkwrap(0.2, 1, rms), ER10, 2 starts; every Fourier error within 1e-3
    and the supports equal.  Grids of the reference test's size (64 x 48) and the other
    shapes of up to 65536 pixels: on much larger grids a pixel whose blurred magnitude
    lies within rounding of the cut becomes likely, and one flipped pixel changes the rest
    of the run (the direct check above covers 256^2 with that band)."""
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    ops = R.flat(R.hio(15), R.shrinkwrap(method="rms"), R.raar(10), R.detwin(R.hio(5)),
                 R.shrinkwrap(0.2, 1.0, "rms"), R.er(10))
    o = oracle.Solver(amp, mask, sup, starts=2, seed=3)
    o.run(ops)
    ores = o.results()
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, seed=3, schedule=schedule)
    errs = R.run_abi(s, ops, sync_every=4)
    xs = s.read_state()
    sups = s.get_support()
    s.close()
    worst = 0.0
    for k in range(2):
        e_o = ores[k][2]
        d = np.abs(errs[:, k] - e_o) / e_o
        worst = max(worst, float(d.max()))
        assert np.all(d <= 1e-3), (k, d.max())
        assert np.linalg.norm(oracle.to_centered(xs[k]) - ores[k][0]) < 1e-3 * np.linalg.norm(ores[k][0])
        assert np.array_equal(oracle.to_centered(sups[k]).astype(bool), ores[k][1].astype(bool)), k
    print(f"{shape} {schedule}: worst per-iteration relative error difference {worst:.2e}; supports equal")
