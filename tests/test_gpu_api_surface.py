"""The Python API's contract, case by case as the reference's python/tests/test_api.py:1-100
states it (return types, shapes and dtypes, the negative-intensity warning, the derived
support, given starts), on the HIP backend."""

import numpy as np
import pytest

N, M = 64, 24


@pytest.fixture(scope="module")
def triangle():
    """A 24-pixel triangle with a smooth amplitude dip and phase bump in a 64 x 64 field
    (the object of the reference's python/tests/conftest.py:7-16)."""
    y, x = np.indices((M, M)) - (M - 1) / 2
    inside = np.indices((M, M)).sum(0) < M
    block = (1 - 0.5 * np.exp(-(x ** 2 + y ** 2) / 20)) * np.exp(1j * 1.2 * np.exp(-((x + 4) ** 2 + (y + 4) ** 2) / 30))
    obj = np.zeros((N, N), np.complex64)
    obj[N // 2 - M // 2:N // 2 + M // 2, N // 2 - M // 2:N // 2 + M // 2] = block * inside
    intensity = np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(obj), norm="ortho"))) ** 2
    return obj, np.abs(obj) > 0, intensity.astype(np.float32)


@pytest.fixture(scope="module")
def qpc():
    import qoraphasecu_amd as qpc

    assert qpc.hip_available(), "no HIP device visible"
    return qpc


@pytest.mark.gpu
def test_single_start_returns_a_reconstruction(qpc, triangle):
    obj, support, intensity = triangle
    result = qpc.reconstruct(intensity, support=support, recipe=qpc.Recipe(qpc.HIO(30), qpc.ER(20)), backend="hip")
    assert isinstance(result, qpc.Reconstruction)
    assert result.object.shape == obj.shape and result.object.dtype == np.complex64
    assert result.support.dtype == bool and result.support.shape == obj.shape
    assert result.errors.shape == (50,)
    assert result.error == result.errors[-1]


@pytest.mark.gpu
def test_ranking_and_best(qpc, triangle):
    _, support, intensity = triangle
    ens = qpc.reconstruct(intensity, support=support, recipe=qpc.Recipe(qpc.HIO(20), qpc.ER(10)), starts=5,
                          backend="hip")
    assert isinstance(ens, qpc.Ensemble) and len(ens) == 5
    errs = [r.error for r in ens]
    assert list(ens.ranking) == sorted(range(5), key=lambda i: errs[i])
    assert ens.best is ens[ens.ranking[0]]


@pytest.mark.gpu
def test_negative_intensity_warns(qpc, triangle):
    _, support, intensity = triangle
    noisy = intensity - 0.5 * intensity.max()
    with pytest.warns(UserWarning, match="negative"):
        qpc.reconstruct(noisy, support=support, recipe=qpc.Recipe(qpc.ER(1)), backend="hip")


@pytest.mark.gpu
def test_autocorrelation_support_covers_the_object(qpc, triangle):
    _, support, intensity = triangle
    derived = qpc.autocorrelation_support(intensity)
    assert derived.dtype == bool and derived.shape == support.shape
    assert derived.sum() > support.sum()


@pytest.mark.gpu
def test_init_replaces_the_random_start(qpc, triangle):
    obj, support, intensity = triangle
    er = qpc.Recipe(qpc.ER(5))
    exact = qpc.reconstruct(intensity, support=support, recipe=er, init=obj, backend="hip")
    assert exact.error < 1e-5
    np.testing.assert_allclose(exact.object, obj, atol=1e-5)

    rng = np.random.default_rng(0)
    starts = (rng.random((2, *obj.shape)) - 0.5).astype(np.complex64)
    a = qpc.reconstruct(intensity, support=support, recipe=er, starts=2, seed=1, init=starts, backend="hip")
    b = qpc.reconstruct(intensity, support=support, recipe=er, starts=2, seed=2, init=starts, backend="hip")
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.object, y.object)  # the seed no longer matters
    shared = qpc.reconstruct(intensity, support=support, recipe=er, starts=2, init=starts[0], backend="hip")
    np.testing.assert_array_equal(shared[0].object, shared[1].object)
    with pytest.raises(ValueError, match=r"init.shape"):
        qpc.reconstruct(intensity, support=support, recipe=er, starts=3, init=starts, backend="hip")


@pytest.mark.gpu
def test_auto_backend_is_hip(qpc, triangle):
    _, support, intensity = triangle
    recipe = qpc.Recipe(qpc.HIO(10), qpc.ER(5))
    auto = qpc.reconstruct(intensity, support=support, recipe=recipe, seed=3)
    hip = qpc.reconstruct(intensity, support=support, recipe=recipe, seed=3, backend="hip")
    np.testing.assert_array_equal(auto.errors, hip.errors)


@pytest.mark.gpu
def test_two_threads_run_independent_solvers(qpc, triangle):
    """The API releases the GIL for a run, and each solver keeps its own stream, graphs and
    error state, so two threads may reconstruct at once (one solver per thread, as the
    reference's contract states): the results equal the sequential runs bit for bit."""
    import threading

    _, support, intensity = triangle
    recipe = qpc.Recipe(qpc.HIO(30), qpc.Shrinkwrap(), qpc.ER(20))
    kw = dict(support=support, recipe=recipe, starts=4, backend="hip")
    want = [qpc.reconstruct(intensity, seed=s, variant=v, **kw) for s, v in ((1, "batched"), (2, "mixed"))]
    got = [None, None]
    errors = []

    def work(k, seed, variant):
        try:
            got[k] = qpc.reconstruct(intensity, seed=seed, variant=variant, **kw)
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(0, 1, "batched")), threading.Thread(target=work, args=(1, 2, "mixed"))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for g, w in zip(got, want):
        for a, b in zip(g, w):
            np.testing.assert_array_equal(a.object, b.object)
            np.testing.assert_array_equal(a.errors, b.errors)
