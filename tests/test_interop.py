"""Recipes and results exchanged with QoraPhase (the reference's python/tests/test_qoraphase.py:19-87).

``qoraphase`` (0.1.0, the reference's uv.lock:660-661) is not installed here and cannot be.
These tests run the converter against a small stand-in package with the layout the
reference's converter relies on (python/qoraphasecu/interop.py:14-49): stage classes on
the top-level module, ``Stage``, ``Repeat`` and ``Detwin`` in ``qoraphase.recipes.recipe``,
``Recipe`` being a stage that holds ``stages``, and ``Reconstruction(object, support,
metrics)`` / ``Ensemble(runs, problem)``.  The expectations are the reference test's.  Parity
with the real QoraPhase package is unpinned."""

import sys
import types
from dataclasses import dataclass

import numpy as np
import pytest

import qoraphasecu_amd as qpc
from qoraphasecu_amd import interop


def _stand_in():
    qp = types.ModuleType("qoraphase")
    recipes = types.ModuleType("qoraphase.recipes")
    recipe = types.ModuleType("qoraphase.recipes.recipe")

    class Stage:
        pass

    @dataclass(frozen=True)
    class ER(Stage):
        n: int

    @dataclass(frozen=True)
    class HIO(Stage):
        n: int
        beta: float = 0.9

    @dataclass(frozen=True)
    class RAAR(Stage):
        n: int
        beta: float = 0.9

    @dataclass(frozen=True)
    class Shrinkwrap(Stage):
        threshold: float = 0.3
        sigma: float = 1.0
        method: str = "max"

    class Recipe(Stage):
        def __init__(self, *stages):
            self.stages = tuple(stages)

    @dataclass(frozen=True)
    class Repeat(Stage):
        n: int
        body: Recipe

    @dataclass(frozen=True)
    class Detwin(Stage):
        body: Recipe
        axis: int = 0

    @dataclass
    class Reconstruction:
        object: np.ndarray
        support: np.ndarray
        metrics: np.ndarray

    @dataclass
    class Ensemble:
        runs: tuple
        problem: object = None

    recipe.Stage, recipe.Repeat, recipe.Detwin = Stage, Repeat, Detwin
    qp.ER, qp.HIO, qp.RAAR, qp.Shrinkwrap, qp.Recipe = ER, HIO, RAAR, Shrinkwrap, Recipe
    qp.repeat = lambda n, *s: Repeat(n, Recipe(*s))
    qp.detwin = lambda *s, axis=0: Detwin(Recipe(*s), axis)
    qp.default_recipe = lambda: Recipe(RAAR(50), qp.detwin(HIO(20)), RAAR(50),
                                       qp.repeat(8, RAAR(50), Shrinkwrap()), ER(200))
    qp.Reconstruction, qp.Ensemble = Reconstruction, Ensemble
    qp.recipes, recipes.recipe = recipes, recipe
    return qp, {"qoraphase": qp, "qoraphase.recipes": recipes, "qoraphase.recipes.recipe": recipe}


@pytest.fixture
def qp(monkeypatch):
    qp, modules = _stand_in()
    for name, module in modules.items():
        monkeypatch.setitem(sys.modules, name, module)
    return qp


def test_nothing_is_qoraphase_before_it_is_imported(monkeypatch):
    monkeypatch.delitem(sys.modules, "qoraphase.recipes.recipe", raising=False)
    assert not interop.is_qoraphase(qpc.ER(3))
    assert not interop.is_qoraphase(object())


def test_default_recipe_converts_to_ours(qp):
    assert interop.is_qoraphase(qp.default_recipe()) and not interop.is_qoraphase(qpc.default_recipe())
    assert qpc.from_qoraphase(qp.default_recipe()).ops() == qpc.default_recipe().ops()


def test_parameters_and_nesting_survive_conversion(qp):
    theirs = qp.Recipe(
        qp.HIO(5, beta=0.7),
        qp.detwin(qp.RAAR(3, beta=0.8), axis=1),
        qp.repeat(2, qp.ER(4), qp.Shrinkwrap(threshold=0.25, sigma=2.0, method="rms")),
    )
    ours = qpc.Recipe(
        qpc.HIO(5, beta=0.7),
        qpc.detwin(qpc.RAAR(3, beta=0.8), axis=1),
        qpc.repeat(2, qpc.ER(4), qpc.Shrinkwrap(threshold=0.25, sigma=2.0, method="rms")),
    )
    assert qpc.from_qoraphase(theirs).ops() == ours.ops()
    assert qpc.from_qoraphase(qp.ER(7)).ops() == qpc.Recipe(qpc.ER(7)).ops()
    # a recipe nested in a recipe flattens like ours
    assert qpc.from_qoraphase(qp.Recipe(qp.Recipe(qp.ER(1)), qp.HIO(2))).ops() == \
        qpc.Recipe(qpc.ER(1), qpc.HIO(2)).ops()


def test_stages_without_a_counterpart_are_rejected(qp):
    @dataclass(frozen=True)
    class DampedER(qp.ER):
        pass

    with pytest.raises(TypeError, match="DampedER has no QoraPhaseCU counterpart"):
        qpc.from_qoraphase(qp.Recipe(qp.HIO(5), qp.repeat(2, DampedER(3))))
    intensity = np.ones((8, 8), np.float32)
    # reconstruct converts before it touches a backend
    with pytest.raises(TypeError, match="no QoraPhaseCU counterpart"):
        qpc.reconstruct(intensity, support=np.ones((8, 8), bool), recipe=DampedER(3), backend="hip")


def test_results_plug_into_qoraphase(qp):
    rng = np.random.default_rng(0)
    runs = tuple(qpc.Reconstruction(rng.standard_normal((4, 4)).astype(np.complex64), np.ones((4, 4), bool),
                                    np.array([0.5, 0.1 * (k + 1)], np.float32)) for k in range(3))
    ens = qpc.Ensemble(runs)
    one = ens.best.to_qoraphase()
    assert isinstance(one, qp.Reconstruction)
    assert one.object is ens.best.object and one.metrics is ens.best.errors
    theirs = ens.to_qoraphase("measurement")
    assert isinstance(theirs, qp.Ensemble) and theirs.problem == "measurement"
    assert [r.object is o.object for r, o in zip(theirs.runs, runs)] == [True] * 3
