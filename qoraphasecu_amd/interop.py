"""Exchange with QoraPhase: its recipes in, our results out (the interface of the
reference's python/qoraphasecu/interop.py and the to_qoraphase methods of its api.py).

`qoraphase` (which brings JAX) is imported only inside the functions that need it.  A
QoraPhase stage converts only when its exact type has a counterpart here; anything else,
a subclass of a built-in stage included (its step may differ), raises TypeError.
"""

from __future__ import annotations

import sys

from .recipe import ER, HIO, RAAR, Detwin, Recipe, Repeat, Shrinkwrap

SUPPORTED = "ER, HIO, RAAR, Shrinkwrap, Recipe, repeat and detwin"


def is_qoraphase(obj) -> bool:
    """Whether obj is a QoraPhase stage.  Only an already imported qoraphase is looked at:
    no such object can exist before its package is loaded."""
    module = sys.modules.get("qoraphase.recipes.recipe")
    return module is not None and isinstance(obj, module.Stage)


def _converters(qp):
    """QoraPhase stage type -> function making the equivalent stage here."""
    from qoraphase.recipes.recipe import Detwin as QpDetwin
    from qoraphase.recipes.recipe import Repeat as QpRepeat

    return {
        qp.ER: lambda st: ER(st.n),
        qp.HIO: lambda st: HIO(st.n, float(st.beta)),
        qp.RAAR: lambda st: RAAR(st.n, float(st.beta)),
        qp.Shrinkwrap: lambda st: Shrinkwrap(float(st.threshold), float(st.sigma), st.method),
        qp.Recipe: lambda st: _recipe(st, qp),
        QpRepeat: lambda st: Repeat(st.n, _recipe(st.body, qp)),
        QpDetwin: lambda st: Detwin(_recipe(st.body, qp), st.axis),
    }


def _recipe(obj, qp) -> Recipe:
    table = _converters(qp)
    parts = obj.stages if type(obj) is qp.Recipe else (obj,)
    out = []
    for stage in parts:
        make = table.get(type(stage))
        if make is None:
            kind = type(stage)
            raise TypeError(f"{kind.__module__}.{kind.__qualname__} has no QoraPhaseCU counterpart; "
                            f"supported stages: {SUPPORTED}")
        out.append(make(stage))
    return Recipe(*out)


def from_qoraphase(recipe) -> Recipe:
    """The Recipe equal to a QoraPhase recipe or to one QoraPhase stage."""
    import qoraphase as qp

    return _recipe(recipe, qp)


def reconstruction_to_qoraphase(run):
    """qoraphase.Reconstruction(object, support, metrics) of one run; the error history
    is its metrics."""
    import qoraphase as qp

    return qp.Reconstruction(run.object, run.support, run.errors)


def ensemble_to_qoraphase(runs, problem=None):
    """qoraphase.Ensemble of the runs (for its align and average); *problem*, the
    measurement's qoraphase.CDI, is what its average is scored against."""
    import qoraphase as qp

    return qp.Ensemble(tuple(reconstruction_to_qoraphase(run) for run in runs), problem)
