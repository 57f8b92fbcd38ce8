"""Recipes and results exchanged with QoraPhase (reference: python/qoraphasecu/interop.py:1-49,
python/qoraphasecu/api.py:37-76).

``qoraphase`` is imported only inside the functions that need it (it brings JAX), so this
package never requires it.  Conversion maps each QoraPhase stage type to the stage of the
same name here, field by field; a stage type without a counterpart, including a subclass
of a built-in stage (whose step may differ), raises ``TypeError``.
"""

from __future__ import annotations

import sys

from .recipe import ER, HIO, RAAR, Recipe, Shrinkwrap, _Detwin, _Repeat

SUPPORTED = "ER, HIO, RAAR, Shrinkwrap, Recipe, repeat and detwin"


def is_qoraphase(obj) -> bool:
    """True for any QoraPhase stage (user subclasses too).  Looks only at modules that are
    already imported: a QoraPhase object cannot exist before its package is."""
    stages = sys.modules.get("qoraphase.recipes.recipe")
    return stages is not None and isinstance(obj, stages.Stage)


def from_qoraphase(recipe) -> Recipe:
    """A :class:`Recipe` equal to a QoraPhase recipe or single stage."""
    import qoraphase as qp

    parts = recipe.stages if type(recipe) is qp.Recipe else (recipe,)
    return Recipe(*(_convert(stage, qp) for stage in parts))


def _convert(stage, qp):
    from qoraphase.recipes.recipe import Detwin, Repeat

    kind = type(stage)  # exact type: a subclass may change what a step does
    if kind is qp.ER:
        return ER(stage.n)
    if kind is qp.HIO:
        return HIO(stage.n, float(stage.beta))
    if kind is qp.RAAR:
        return RAAR(stage.n, float(stage.beta))
    if kind is qp.Shrinkwrap:
        return Shrinkwrap(float(stage.threshold), float(stage.sigma), stage.method)
    if kind is qp.Recipe:
        return from_qoraphase(stage)
    if kind is Repeat:
        return _Repeat(stage.n, from_qoraphase(stage.body))
    if kind is Detwin:
        return _Detwin(from_qoraphase(stage.body), stage.axis)
    raise TypeError(f"{kind.__module__}.{kind.__qualname__} has no QoraPhaseCU counterpart; "
                    f"supported stages: {SUPPORTED}")


def reconstruction_to_qoraphase(run):
    """``qoraphase.Reconstruction(object, support, metrics)`` of one run; the error history
    becomes the metrics."""
    import qoraphase as qp

    return qp.Reconstruction(run.object, run.support, run.errors)


def ensemble_to_qoraphase(runs, problem=None):
    """``qoraphase.Ensemble`` of the runs (for its align and average); *problem* is the
    measurement's ``qoraphase.CDI``, which its average is scored against."""
    import qoraphase as qp

    return qp.Ensemble(tuple(reconstruction_to_qoraphase(run) for run in runs), problem)
