"""Recipes: the stage objects a caller composes, and their flattening into the records the
engine walks.

Public names, fields and defaults are those of the reference's Python API
(python/qoraphasecu/recipe.py) so that a recipe written for it runs unchanged here.  The
flattened form is a list of 7-field records (kind, n, beta, threshold, sigma, method,
axis), the format the C++ binding reads (python/bindings.cpp:22-23); `Op` is a named tuple,
so a record is also a plain 7-tuple.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Iterator, Literal, NamedTuple


class Op(NamedTuple):
    """One engine operation.  Iteration kinds use n and beta, "shrinkwrap" uses threshold,
    sigma and method, "detwin_begin" uses axis."""

    kind: str
    n: int = 0
    beta: float = 0.0
    threshold: float = 0.0
    sigma: float = 0.0
    method: str = "max"
    axis: int = 0


ITERATION_KINDS = frozenset({"er", "hio", "raar"})
_LEVELS = ("max", "rms")


class Stage:
    """Something that expands to engine operations, in execution order."""

    def records(self) -> Iterator[Op]:
        raise NotImplementedError

    def ops(self) -> list:
        return list(self.records())


@dataclass(frozen=True)
class ER(Stage):
    """Error reduction: inside the support the iterate takes its Fourier projection,
    outside it becomes zero."""

    n: int

    def records(self):
        yield Op("er", self.n)


@dataclass(frozen=True)
class HIO(Stage):
    """Hybrid input-output: the projection inside the support, x - beta * projection
    outside."""

    n: int
    beta: float = 0.9

    def records(self):
        yield Op("hio", self.n, self.beta)


@dataclass(frozen=True)
class RAAR(Stage):
    """Relaxed averaged alternating reflections: the projection inside the support,
    beta * x + (1 - 2 beta) * projection outside."""

    n: int
    beta: float = 0.9

    def records(self):
        yield Op("raar", self.n, self.beta)


@dataclass(frozen=True)
class Shrinkwrap(Stage):
    """Replace the support by the pixels whose Gaussian-blurred (sigma) magnitude exceeds
    threshold times a level: the blurred maximum ("max") or the root mean square over the
    current support ("rms")."""

    threshold: float = 0.3
    sigma: float = 1.0
    method: Literal["max", "rms"] = "max"

    def __post_init__(self):
        if self.method not in _LEVELS:
            raise ValueError(f"Shrinkwrap level {self.method!r} is unknown; use one of {_LEVELS}")

    def records(self):
        yield Op("shrinkwrap", threshold=self.threshold, sigma=self.sigma, method=self.method)


class Recipe(Stage):
    """Stages run in the order given."""

    def __init__(self, *stages: Stage):
        self.stages = tuple(stages)

    def __len__(self) -> int:
        return len(self.stages)

    def records(self):
        for stage in self.stages:
            yield from stage.records()

    def iterations(self) -> int:
        """Total iterations of the flattened recipe (the length of each error history)."""
        return sum(op.n for op in self.records() if op.kind in ITERATION_KINDS)


@dataclass(frozen=True)
class Repeat(Stage):
    """`count` copies of a body, one after another."""

    count: int
    body: Recipe

    def records(self):
        for _ in range(self.count):
            yield from self.body.records()


@dataclass(frozen=True)
class Detwin(Stage):
    """The body run on the support cut at its centre of mass along `axis`; the full support
    is put back afterwards."""

    body: Recipe
    axis: int = 0

    def records(self):
        yield Op("detwin_begin", axis=self.axis)
        yield from self.body.records()
        yield Op("detwin_end")


def repeat(n: int, *stages: Stage) -> Repeat:
    """repeat(10, RAAR(50), Shrinkwrap()) runs the pair ten times."""
    return Repeat(n, Recipe(*stages))


def detwin(*stages: Stage, axis: int = 0) -> Detwin:
    return Detwin(Recipe(*stages), axis)


def default_recipe() -> Recipe:
    """The default of the reference (src/core/recipe.cpp:26-28): 720 iterations."""
    return Recipe(RAAR(50), detwin(HIO(20)), RAAR(50), repeat(8, RAAR(50), Shrinkwrap()), ER(200))
