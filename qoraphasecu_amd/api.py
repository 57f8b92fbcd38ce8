"""Array in, reconstruction out: `reconstruct` and its result types, with the call
signature and result shape of the reference's python/qoraphasecu/api.py, plus
`reconstruct_many` for problems of several sizes at once.

Every call goes to the C++ engine (`_core`); its Backend::hip runs the iterations on the
GPU.  backend="auto" means "hip" when a device is visible.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Literal

import numpy as np

from . import _core, interop
from .recipe import Recipe, default_recipe

# Above this fraction of negative intensity pixels the caller is warned that they will be
# clipped to zero (the reference's behaviour and threshold).
NEGATIVE_WARN_FRACTION = 0.01


@dataclass(frozen=True)
class Reconstruction:
    """One run's object (centre-origin, known up to a global phase, a shift and the
    twin), its final support and its Fourier error per iteration."""

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
        """The last Fourier error (NaN for an empty history): what runs are ranked by."""
        return float(self.errors[-1]) if len(self.errors) else math.nan

    def to_qoraphase(self):
        """This run as a qoraphase.Reconstruction (the error history becomes its metrics)."""
        return interop.reconstruction_to_qoraphase(self)


@dataclass(frozen=True)
class Ensemble:
    """All runs of a multi-start reconstruction, each in its own frame."""

    runs: tuple

    def __len__(self) -> int:
        return len(self.runs)

    def __iter__(self):
        return iter(self.runs)

    def __getitem__(self, index):
        return self.runs[index]

    @property
    def ranking(self) -> tuple:
        """Run indices by increasing final error; runs whose error is NaN come last (the
        rule of the C++ best_index)."""
        def key(i):
            e = self.runs[i].error
            return (math.isnan(e), e if not math.isnan(e) else 0.0, i)

        return tuple(sorted(range(len(self.runs)), key=key))

    @property
    def best(self) -> Reconstruction:
        return self.runs[self.ranking[0]]

    def to_qoraphase(self, problem=None):
        """The runs as a qoraphase.Ensemble, for its align and average; give the
        measurement's qoraphase.CDI as *problem* to have the average scored against it."""
        return interop.ensemble_to_qoraphase(self.runs, problem)


def cuda_available() -> bool:
    return _core.cuda_available()


def hip_available() -> bool:
    return _core.hip_available()


def hip_device_info() -> dict:
    return _core.hip_device_info()


def autocorrelation_support(intensity, threshold: float = 0.1) -> np.ndarray:
    """The loose support a reconstruction starts from: where the autocorrelation's
    magnitude exceeds threshold times its peak."""
    return _core.autocorrelation_support(np.ascontiguousarray(intensity, dtype=np.float32), threshold).astype(bool)


def simulate_intensity(obj) -> np.ndarray:
    """The noiseless intensity |centred unitary FFT(obj)|^2."""
    return _core.simulate_intensity(np.ascontiguousarray(obj, dtype=np.complex64))


def aligned_error(a, reference) -> float:
    """||a - reference|| / ||reference|| once the shift, the global phase and the twin of
    `a` that fit `reference` best are taken out."""
    a = np.ascontiguousarray(a, dtype=np.complex64)
    reference = np.ascontiguousarray(reference, dtype=np.complex64)
    if a.shape != reference.shape:
        raise ValueError(f"shapes differ: {a.shape} vs {reference.shape}")
    return _core.aligned_error(a, reference)


# --- argument normalisation -------------------------------------------------------------


def _intensity(intensity) -> np.ndarray:
    arr = np.ascontiguousarray(intensity, dtype=np.float32)
    if arr.ndim not in (2, 3):
        raise ValueError(f"intensity must be 2D or 3D, got ndim={arr.ndim}")
    return arr


def _warn_negative(intensity: np.ndarray) -> None:
    share = float(np.count_nonzero(intensity < 0)) / max(intensity.size, 1)
    if share > NEGATIVE_WARN_FRACTION:
        warnings.warn(f"{share:.1%} of intensity pixels are negative and will be clipped to 0; "
                      "consider masking them instead", stacklevel=3)


def _bool_plane(name: str, arr, shape):
    """A boolean array of the intensity's shape as uint8 (None stays None)."""
    if arr is None:
        return None
    flags = np.ascontiguousarray(arr, dtype=bool)
    if flags.shape != shape:
        raise ValueError(f"{name}.shape {flags.shape} != intensity.shape {shape}")
    return flags.view(np.uint8)


def _initial(init, shape, starts):
    """Starting objects as a (starts, *shape) complex64 array: one object is used for
    every start."""
    if init is None:
        return None
    arr = np.asarray(init, dtype=np.complex64)
    if arr.shape == shape:
        arr = np.broadcast_to(arr, (starts, *shape))
    if arr.shape != (starts, *shape):
        raise ValueError(f"init.shape {arr.shape} must be the intensity's {shape} or (starts, *shape)")
    return np.ascontiguousarray(arr)


def _resolve(recipe, backend: str):
    if backend == "auto":
        backend = "hip" if hip_available() else "cpu"
    if recipe is None:
        recipe = default_recipe()
    elif interop.is_qoraphase(recipe):
        recipe = interop.from_qoraphase(recipe)
    return recipe, backend


def _package(runs, starts):
    results = tuple(Reconstruction(obj, sup.astype(bool), err) for obj, sup, err in runs)
    return results[0] if starts == 1 else Ensemble(results)


# --- entry points -----------------------------------------------------------------------


def reconstruct(
    intensity,
    *,
    support=None,
    mask=None,
    recipe: Recipe | None = None,
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
    """Reconstruct an object from one measured intensity.

    intensity: centre-origin, 2D or 3D.  support: True inside the object (None: the
    autocorrelation support at `support_threshold`).  mask: True marks an unusable pixel.
    recipe: a Recipe, or a QoraPhase recipe or stage (None: `default_recipe()`).
    starts: random starts; more than one returns an Ensemble.  batch: starts per launch in
    the batched and reduced-precision variants (None: all in one).  seed: of the random
    starting phases (None: 0).  init: a starting object for every start, or one per start
    as (starts, *shape), instead of random phases; it is cut to the support first.
    backend: "auto" takes "hip" when a device is visible.  variant: the execution variant
    (qpcu::Variant).  sync_every: iterations between error read-backs.
    """
    arr = _intensity(intensity)
    _warn_negative(arr)
    recipe, backend = _resolve(recipe, backend)
    runs = _core.reconstruct(arr, _bool_plane("support", support, arr.shape), _bool_plane("mask", mask, arr.shape),
                             recipe.ops(), starts, batch or 0, seed or 0, _initial(init, arr.shape, starts),
                             support_threshold, backend, variant, sync_every)
    return _package(runs, starts)


def reconstruct_many(
    intensities,
    *,
    supports=None,
    masks=None,
    recipe: Recipe | None = None,
    starts: int = 1,
    batch: int | None = None,
    seed: int | None = None,
    support_threshold: float = 0.1,
    backend: Literal["auto", "hip"] = "auto",
    variant: str = "batched",
    sync_every: int = 10,
) -> list:
    """Reconstruct several independent problems (any mix of 2D and 3D shapes) with one
    recipe; one Reconstruction or Ensemble per problem, in order.

    On the batched and reduced-precision variants the problems of one shape share one
    stacked device solver, and every shape advances in the same graph launch."""
    arrays = [_intensity(i) for i in intensities]
    if not arrays:
        raise ValueError("no problems")
    recipe, backend = _resolve(recipe, backend)

    def per_problem(name, given):
        if given is None:
            return None
        if len(given) != len(arrays):
            raise ValueError(f"{name} needs one entry per intensity")
        return [_bool_plane(name, g, a.shape) for g, a in zip(given, arrays)]

    runs = _core.reconstruct_many(arrays, per_problem("support", supports), per_problem("mask", masks),
                                  recipe.ops(), starts, batch or 0, seed or 0, support_threshold, backend, variant,
                                  sync_every)
    return [_package(per, starts) for per in runs]
