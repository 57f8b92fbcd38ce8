"""QoraPhase-MI355X: the phase-retrieval reconstruction path of QoraPhaseCU on AMD
Instinct MI355X (gfx950), behind the same ``reconstruct`` API.

    import qoraphasecu_amd as qpc
    result = qpc.reconstruct(intensity, support=support, backend="hip")

Arrays are centre-origin, ``mask=True`` marks an unusable pixel, ``intensity`` is a
measured intensity.  The native parts are built in-tree by ``qoraphasecu_amd.build``.
"""

import importlib.util as _ilu

# One HIP runtime per process: torch's ROCm wheel carries its own libamdhip64, which its
# libraries load by path.  If this package's engine brought up the system runtime first,
# torch would start a second one and find no GPU.  Loading torch first (when it is
# installed) lets the engine bind to the same runtime, so both can share the device.
if _ilu.find_spec("torch") is not None:
    import torch as _torch  # noqa: F401

from .api import (
    Ensemble,
    Reconstruction,
    aligned_error,
    autocorrelation_support,
    cuda_available,
    hip_available,
    hip_device_info,
    reconstruct,
    reconstruct_many,
    simulate_intensity,
)
from .interop import from_qoraphase
from .recipe import ER, HIO, RAAR, Recipe, Shrinkwrap, default_recipe, detwin, repeat

__all__ = [
    "ER",
    "HIO",
    "RAAR",
    "Ensemble",
    "Recipe",
    "Reconstruction",
    "Shrinkwrap",
    "aligned_error",
    "autocorrelation_support",
    "cuda_available",
    "default_recipe",
    "detwin",
    "from_qoraphase",
    "hip_available",
    "hip_device_info",
    "reconstruct",
    "reconstruct_many",
    "repeat",
    "simulate_intensity",
]
