"""Symmetries every correct implementation keeps (the covariance checks of the reference's
python/tests/test_equivalence.py:119-148): a global phase on the start carries through
the run, and the twin start conj(x(-r)) with the mirrored support gives the mirrored
run, with the same error history.  Checked on the oracle (CPU) and on the HIP backend."""

import numpy as np
import pytest

import oracle


def _case(shape, seed):
    """An ellipse of random amplitude and phase, its intensity (centre-origin), a box
    support around it and a random-phase start consistent with the intensity."""
    rng = np.random.default_rng(seed)
    grid = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    inside = sum((g / (0.22 * n)) ** 2 for g, n in zip(grid, shape)) < 1
    obj = np.where(inside, (0.6 + 0.4 * rng.random(shape)) * np.exp(2j * np.pi * rng.random(shape)), 0)
    spec = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(obj), norm="ortho"))
    intensity = (np.abs(spec) ** 2).astype(np.float32)
    support = np.zeros(shape, bool)
    support[tuple(slice(n // 4 - 2, 3 * n // 4 + 2) for n in shape)] = True
    phase = np.exp(2j * np.pi * rng.random(shape))
    start = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(np.sqrt(intensity) * phase), norm="ortho"))
    return intensity, support, start.astype(np.complex64)


def _twin(a):
    """a(-r) about the centre-origin pixel n // 2 of every axis."""
    return np.roll(a[(slice(None, None, -1),) * a.ndim], 1, axis=tuple(range(a.ndim)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle_run(ops, intensity, support, start):
    (obj, _, err), = oracle.reconstruct(intensity, ops, support=support.astype(np.uint8), starts=1,
                                       start=start[None])
    return obj, np.asarray(err)


@pytest.mark.parametrize("shape", [(32, 32), (16, 12, 20)])
def test_oracle_global_phase_and_twin(shape):
    intensity, support, start = _case(shape, 4)
    ops = [("raar", 5, 0.9, 0.0, 0.0, "max", 0), ("er", 5, 0.0, 0.0, 0.0, "max", 0)]
    turn = np.complex64(np.exp(0.7j))
    a_obj, a_err = _oracle_run(ops, intensity, support, start)
    b_obj, b_err = _oracle_run(ops, intensity, support, turn * start)
    assert _rel(b_obj, turn * a_obj) < 1e-5
    np.testing.assert_allclose(b_err, a_err, atol=1e-6)
    support = support.copy()
    support.flat[support.size // 3] = True  # so that the mirrored support differs
    a_obj, a_err = _oracle_run(ops, intensity, support, start)
    c_obj, c_err = _oracle_run(ops, intensity, _twin(support), np.conj(_twin(start)))
    assert _rel(c_obj, np.conj(_twin(a_obj))) < 1e-5
    np.testing.assert_allclose(c_err, a_err, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,recipe_kind,tol", [("graph", "hio+er", 3e-5), ("batched", "hio+er", 3e-5),
                                                     ("half", "er", 2e-2)])
@pytest.mark.parametrize("shape", [(32, 32), (16, 16, 32)])
def test_hip_global_phase_and_twin(shape, variant, recipe_kind, tol):
    """A float FFT is not exactly covariant (the rotated or mirrored input rounds
    differently) and HIO amplifies rounding, so the fp32 bar is 3e-5 over HIO(5) + ER(5)
    (the reference's 1e-5 was set on its CPU FFT; the oracle here meets it).  fp16 storage
    rounds every stored value at 2^-11 (three stores per element and iteration in 3D), so it
    is checked over contractive ER iterations at 2e-2 (measured: 4e-3 in 2D, 7e-3 in 3D).
    A symmetry bug (a wrong mirror index, a dropped conjugate) shows at O(1)."""
    import qoraphasecu_amd as qpc

    intensity, support, start = _case(shape, 5)
    recipe = qpc.Recipe(qpc.HIO(5), qpc.ER(5)) if recipe_kind == "hio+er" else qpc.Recipe(qpc.ER(10))
    turn = np.complex64(np.exp(0.7j))

    def run(sup, init):
        return qpc.reconstruct(intensity, support=sup, recipe=recipe, init=init, backend="hip", variant=variant)

    a = run(support, start)
    b = run(support, turn * start)
    assert _rel(b.object, turn * a.object) < tol
    np.testing.assert_allclose(b.errors, a.errors, atol=1e-6, rtol=tol)
    support = support.copy()
    support.flat[support.size // 3] = True  # so that the mirrored support differs
    a = run(support, start)
    c = run(_twin(support), np.conj(_twin(start)))
    assert _rel(c.object, np.conj(_twin(a.object))) < tol
    np.testing.assert_allclose(c.errors, a.errors, atol=1e-6, rtol=tol)
