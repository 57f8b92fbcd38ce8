"""Problem set-up and the object metric on the device against the oracle: the
autocorrelation support and simulate_intensity (src/core/problem.cpp:72-99, now a device
kernel chain: roll, transform, |.|, maximum, cut), and aligned_error
(src/core/metrics.cpp:17-86), on the benchmark workloads' own inputs.

The device and the oracle transform in float with different algorithms (about 1e-7
apart relative to the largest value).  An autocorrelation value within 1e-5 relative of
the cut may therefore fall on either side; every other pixel must match exactly, and on
these inputs the band turns out to be empty or nearly so (it is printed)."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BAND = 1e-5


def _exact_autocorrelation(intensity):
    """|unitary inverse FFT of the FFT-order intensity|, centre-origin, in double."""
    ac = np.fft.ifftn(np.fft.ifftshift(intensity.astype(np.float64)), norm="ortho")
    return np.fft.fftshift(np.abs(ac))


def _workload_intensities():
    from qoraphasecu_amd import workloads

    out = []
    for name in ("c1", "c2", "c4"):
        w = workloads.config(name)
        out.append((name, w["intensity"]))
    c3 = workloads.config("c3")["problems"]
    for k in (0, 16, 32):
        out.append((f"c3 problem {k} ({c3[k][1].shape[0]}^2)", c3[k][1]))
    return out


@pytest.fixture(scope="module")
def intensities():
    return _workload_intensities()


def test_autocorrelation_support_matches_oracle(qpch, intensities):
    import qoraphasecu_amd as qpc

    for name, inten in intensities:
        for threshold in (0.1, 0.02):
            got = qpc.autocorrelation_support(inten, threshold).astype(np.uint8)
            want = oracle.autocorrelation_support(inten, threshold)
            ac = _exact_autocorrelation(inten)
            cut = threshold * ac.max()
            near = np.abs(ac - cut) <= BAND * cut
            diff = got != want
            print(f"{name} threshold {threshold}: {int(want.sum())} pixels in the support, {int(near.sum())} within "
                  f"{BAND:g} of the cut, {int(diff.sum())} differ")
            assert not np.any(diff & ~near), (name, threshold, int((diff & ~near).sum()))
            assert 0 < int(want.sum()) < want.size


def test_batched_autocorrelation_support_equals_one_by_one(qpch):
    """The 16 problems of one size in one device pass give each problem's own support, bit
    for bit."""
    from qoraphasecu_amd import workloads

    probs = [p[1] for p in workloads.config("c3")["problems"][16:32]]
    stacked = qpch.autocorrelation_support(np.stack(probs), 0.1, count=len(probs))
    for k, inten in enumerate(probs):
        assert np.array_equal(stacked[k], qpch.autocorrelation_support(inten, 0.1)), k


def test_make_problem_support_is_the_device_autocorrelation(qpch):
    """make_problem (the C++ API) without a support takes the device autocorrelation."""
    from qoraphasecu_amd import _core, workloads

    w = workloads.config("c2")
    amp, mask, sup = _core.make_problem(w["intensity"], None, w["mask"], 0.1)
    assert np.array_equal(sup, qpch.autocorrelation_support(w["intensity"], 0.1))
    assert np.array_equal(amp, np.sqrt(np.maximum(w["intensity"], 0)).astype(np.float32))
    assert np.array_equal(mask, w["mask"])


@pytest.mark.parametrize("shape", [(64, 64), (256, 256), (21, 15), (17, 24), (15, 16, 17), (32, 32, 32)])
def test_simulate_intensity_matches_oracle(shape):
    import qoraphasecu_amd as qpc

    rng = np.random.default_rng(sum(shape))
    obj = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = qpc.simulate_intensity(obj)
    want = oracle.simulate_intensity(obj.astype(np.complex128), "d")
    rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    assert rel <= 1e-6, (shape, rel)
    assert float(np.linalg.norm(oracle.simulate_intensity(obj) - want) / np.linalg.norm(want)) <= 1e-6


def test_simulate_intensity_of_the_workload_objects():
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    for name in ("c1", "c2", "c4"):
        obj = workloads.config(name)["object"]
        got = qpc.simulate_intensity(obj)
        want = oracle.simulate_intensity(obj.astype(np.complex128), "d")
        rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        print(f"{name}: simulate_intensity within {rel:.1e} of the FP64 oracle")
        assert rel <= 1e-6, (name, rel)


def _moved(obj, shift, angle, twin):
    """obj shifted circularly, rotated by a global phase and (optionally) replaced by its
    twin conj(obj(-r))."""
    out = obj
    if twin:
        out = np.conj(np.flip(out, axis=tuple(range(out.ndim))))
        out = np.roll(out, 1, axis=tuple(range(out.ndim)))  # conj(obj[(-r) mod n])
    return (np.exp(1j * angle) * np.roll(out, shift, axis=tuple(range(out.ndim)))).astype(np.complex64)


@pytest.mark.parametrize("shift,angle,twin", [((0, 0), 0.0, False), ((5, -7), 0.7, False), ((3, 11), -2.1, True),
                                              ((0, 0), 1.3, True)])
@pytest.mark.parametrize("noise", [0.0, 0.05])
def test_aligned_error_matches_oracle(shift, angle, twin, noise):
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    obj = workloads.config("c2")["object"]
    a = _moved(obj, shift, angle, twin)
    if noise:
        rng = np.random.default_rng(4)
        a = (a + noise * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape)) * np.abs(obj).max()
             ).astype(np.complex64)
    got = qpc.aligned_error(a, obj)
    want = oracle.aligned_error(a.astype(np.complex128), obj.astype(np.complex128), "d")
    print(f"shift {shift} angle {angle} twin {twin} noise {noise}: device {got:.7e}, FP64 oracle {want:.7e}")
    assert abs(got - want) <= 1e-6, (got, want)
    if not noise:
        assert got <= 1e-6


@pytest.mark.parametrize("shape,shift,twin", [((48, 40), (7, -3), False), ((48, 40), (-5, 9), True),
                                         ((12, 20, 18), (1, -4, 6), True)])
def test_aligned_error_generic_shapes_match_oracle(shape, shift, twin):
    """Axes that are not powers of two (the generic transforms), a random complex object with
    a compact support, noise on top: aligned_error against the FP64 oracle's."""
    import qoraphasecu_amd as qpc

    rng = np.random.default_rng(sum(shape))
    obj = np.zeros(shape, np.complex64)
    inner = tuple(slice(s // 4, s // 4 + s // 2) for s in shape)
    obj[inner] = (rng.uniform(0.3, 1.0, obj[inner].shape) * np.exp(1j * rng.uniform(-np.pi, np.pi, obj[inner].shape)))
    a = _moved(obj, shift, 0.9, twin)
    a = (a + 0.02 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)
    got = qpc.aligned_error(a, obj)
    want = oracle.aligned_error(a.astype(np.complex128), obj.astype(np.complex128), "d")
    print(f"{shape} shift {shift} twin {twin}: device {got:.7e}, FP64 oracle {want:.7e}")
    assert abs(got - want) <= 1e-6, (got, want)


def test_aligned_error_3d_matches_oracle():
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    obj = workloads.config("c4")["object"]
    a = _moved(obj, (2, -3, 4), 0.4, True)
    got = qpc.aligned_error(a, obj)
    want = oracle.aligned_error(a.astype(np.complex128), obj.astype(np.complex128), "d")
    assert abs(got - want) <= 1e-6, (got, want)
