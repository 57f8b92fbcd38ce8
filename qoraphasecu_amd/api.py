"""``reconstruct``: array in, reconstruction out (reference: python/qoraphasecu/api.py:1-185).

The arrays go to the C++ engine (``_core``), whose ``Backend::hip`` runs every iteration
on the GPU.  ``backend="auto"`` picks ``"hip"`` when a device is present.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Literal

import numpy as np

from . import _core, interop
from .recipe import default_recipe

NEGATIVE_WARN_FRACTION = 0.01


@dataclass(frozen=True)
class Reconstruction:
    """One reconstructed object, defined up to a global phase, a translation and a twin."""

    object: np.ndarray
    support: np.ndarray
    errors: np.ndarray

    @property
    def amplitude(self) -> np.ndarray:
        return np.abs(self.object)

    @property
    def phase(self) -> np.ndarray:
        return np.angle(self.object)

    @property
    def error(self) -> float:
        """Final Fourier error; the ranking key."""
        return float(self.errors[-1]) if self.errors.size else float("nan")

    def to_qoraphase(self):
        """The same run as a ``qoraphase.Reconstruction``; ``errors`` becomes ``metrics``."""
        return interop.reconstruction_to_qoraphase(self)


@dataclass(frozen=True)
class Ensemble:
    """The runs of a multi-start reconstruction, not mutually aligned."""

    runs: tuple

    def __len__(self) -> int:
        return len(self.runs)

    def __iter__(self):
        return iter(self.runs)

    def __getitem__(self, index):
        return self.runs[index]

    @property
    def ranking(self) -> tuple:
        """Run indices, best (lowest error) first."""
        return tuple(sorted(range(len(self.runs)), key=lambda i: self.runs[i].error))

    @property
    def best(self) -> Reconstruction:
        return self.runs[self.ranking[0]]

    def to_qoraphase(self, problem=None):
        """A ``qoraphase.Ensemble`` of the runs, for its ``align`` and ``average``; pass the
        measurement's ``qoraphase.CDI`` as *problem* to have ``average`` scored against it."""
        return interop.ensemble_to_qoraphase(self.runs, problem)


def cuda_available() -> bool:
    return _core.cuda_available()


def hip_available() -> bool:
    return _core.hip_available()


def hip_device_info() -> dict:
    return _core.hip_device_info()


def autocorrelation_support(intensity, threshold: float = 0.1) -> np.ndarray:
    """Loose support from the thresholded autocorrelation; Shrinkwrap tightens it."""
    intensity = np.ascontiguousarray(intensity, dtype=np.float32)
    return _core.autocorrelation_support(intensity, threshold).astype(bool)


def simulate_intensity(obj) -> np.ndarray:
    """|centred unitary FFT(obj)|^2."""
    return _core.simulate_intensity(np.ascontiguousarray(obj, dtype=np.complex64))


def aligned_error(a, reference) -> float:
    """``||a - reference|| / ||reference||`` after removing shift, global phase and twin."""
    a = np.ascontiguousarray(a, dtype=np.complex64)
    reference = np.ascontiguousarray(reference, dtype=np.complex64)
    if a.shape != reference.shape:
        raise ValueError(f"shapes differ: {a.shape} vs {reference.shape}")
    return _core.aligned_error(a, reference)


def _flag(name, array, shape):
    if array is None:
        return None
    array = np.ascontiguousarray(array, dtype=bool)
    if array.shape != shape:
        raise ValueError(f"{name}.shape {array.shape} != intensity.shape {shape}")
    return array.view(np.uint8)


def _start(init, shape, starts):
    if init is None:
        return None
    init = np.asarray(init, dtype=np.complex64)
    if init.shape == shape:
        init = np.broadcast_to(init, (starts, *shape))
    if init.shape != (starts, *shape):
        raise ValueError(f"init.shape {init.shape} must be the intensity's {shape} or (starts, *shape)")
    return np.ascontiguousarray(init)


def reconstruct(
    intensity,
    *,
    support=None,
    mask=None,
    recipe=None,
    starts: int = 1,
    batch: int | None = None,
    seed: int | None = None,
    init=None,
    support_threshold: float = 0.1,
    backend: Literal["auto", "cpu", "cuda", "hip"] = "auto",
    variant: Literal["naive", "fused", "graph", "batched", "half", "callback", "mixed", "bf16",
                     "mixed_bf16"] = "batched",
    sync_every: int = 10,
) -> Reconstruction | Ensemble:
    """Array in, reconstruction out.

    Args:
        intensity: measured intensity, centre-origin, 2D or 3D.
        support: boolean, ``True`` inside the object; ``None`` derives an
            autocorrelation support.
        mask: boolean, ``True`` marks *unusable* pixels.
        recipe: a :class:`Recipe` or a QoraPhase recipe (converted by
            :func:`from_qoraphase`); ``None`` uses :func:`default_recipe`.
        starts: number of random starts; more than one returns an :class:`Ensemble`.
        batch: starts per launch in the batched and reduced-precision variants;
            ``None`` puts all of them in one.
        seed: seed of the random starting phases; ``None`` means 0.
        init: starting object(s), centre-origin, instead of random phases: one array of
            the intensity's shape for every start, or ``(starts, *shape)``.
        backend: ``"auto"`` picks ``"hip"`` when a device is present.
        variant: execution variant (see ``qpcu::Variant``).
    """
    intensity = np.ascontiguousarray(intensity, dtype=np.float32)
    if intensity.ndim not in (2, 3):
        raise ValueError(f"intensity must be 2D or 3D, got ndim={intensity.ndim}")
    negative = float(np.mean(intensity < 0))
    if negative > NEGATIVE_WARN_FRACTION:
        warnings.warn(
            f"{negative:.1%} of intensity pixels are negative and will be clipped to 0; "
            "consider masking them instead",
            stacklevel=2,
        )
    if backend == "auto":
        backend = "hip" if hip_available() else "cpu"
    if recipe is None:
        recipe = default_recipe()
    elif interop.is_qoraphase(recipe):
        recipe = interop.from_qoraphase(recipe)
    runs = _core.reconstruct(
        intensity,
        _flag("support", support, intensity.shape),
        _flag("mask", mask, intensity.shape),
        recipe.ops(),
        starts,
        0 if batch is None else batch,
        0 if seed is None else seed,
        _start(init, intensity.shape, starts),
        support_threshold,
        backend,
        variant,
        sync_every,
    )
    results = tuple(Reconstruction(obj, sup.astype(bool), err) for obj, sup, err in runs)
    return results[0] if starts == 1 else Ensemble(results)


def reconstruct_many(
    intensities,
    *,
    supports=None,
    masks=None,
    recipe=None,
    starts: int = 1,
    batch: int | None = None,
    seed: int | None = None,
    support_threshold: float = 0.1,
    backend: Literal["auto", "hip"] = "auto",
    variant: str = "batched",
    sync_every: int = 10,
) -> list:
    """Several independent problems (any mix of 2D/3D shapes) with one recipe.

    On the batched and reduced-precision variants, problems of one shape share one
    stacked device solver and all shapes advance together, one graph launch per chunk of
    iterations.  Returns one :class:`Reconstruction` or :class:`Ensemble` per problem."""
    intensities = [np.ascontiguousarray(i, dtype=np.float32) for i in intensities]
    if not intensities:
        raise ValueError("no problems")
    for i in intensities:
        if i.ndim not in (2, 3):
            raise ValueError(f"intensity must be 2D or 3D, got ndim={i.ndim}")
    if backend == "auto":
        backend = "hip" if hip_available() else "cpu"
    if recipe is None:
        recipe = default_recipe()
    elif interop.is_qoraphase(recipe):
        recipe = interop.from_qoraphase(recipe)

    def flags(name, arrays):
        if arrays is None:
            return None
        if len(arrays) != len(intensities):
            raise ValueError(f"{name} needs one entry per intensity")
        return [None if a is None else _flag(name, a, i.shape) for a, i in zip(arrays, intensities)]

    runs = _core.reconstruct_many(intensities, flags("support", supports), flags("mask", masks), recipe.ops(),
                                  starts, 0 if batch is None else batch, 0 if seed is None else seed,
                                  support_threshold, backend, variant, sync_every)
    out = []
    for per in runs:
        results = tuple(Reconstruction(obj, sup.astype(bool), err) for obj, sup, err in per)
        out.append(results[0] if starts == 1 else Ensemble(results))
    return out
