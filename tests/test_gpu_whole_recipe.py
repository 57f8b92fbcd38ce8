"""A whole recipe in one graph launch (qpch_group_run) against the same recipe walked step by
step through the C ABI (chunks of sync_every iterations, Shrinkwrap, detwin): the error
histories, the iterates and the supports are the same, bit for bit, for one solver and for
a group of solvers of different shapes, on the first run (captured) and on a repeat
(replayed), for fp32 (whole-image kernel, line passes, generic passes, the 3D plane pass),
fp16 / bf16 storage with ER in fp32, and launch groups on concurrent lanes."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

pytestmark = pytest.mark.gpu

RECIPE = R.flat(R.raar(12), R.detwin(R.hio(6), axis=1), R.shrinkwrap(), R.raar(9), R.shrinkwrap(0.2, 1.0, "rms"),
                R.er(7))


def _stepwise(solver, ops, sync_every=10):
    """The C++ host's chunked walk, with detwin on the device."""
    total = R.iterations(ops)
    solver.begin_run(total)
    done = 0
    for kind, n, beta, thr, sigma, method, axis in ops:
        if kind in ("er", "hio", "raar"):
            a, b = R.coefficients(kind, beta)
            k = 0
            while k < n:
                ln = min(n - k, sync_every - done % sync_every)
                solver.iterate(ln, a, b, refine=(kind == "er"))
                k += ln
                done += ln
        elif kind == "shrinkwrap":
            solver.shrinkwrap(thr, sigma, method)
        elif kind == "detwin_begin":
            solver.save_support()
            solver.half_support(axis)
        elif kind == "detwin_end":
            solver.restore_support()
    return solver.read_errors(0, total)


def _solver(qpch, shape, **kw):
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    amp, mask, sup = oracle.make_problem(inten, None, None)
    f = oracle.to_fft_order
    return qpch.Solver(f(amp), f(mask), f(sup), seed=2, schedule="graph", **kw)


def _same(a, b):
    assert np.array_equal(a.read_state(), b.read_state())
    assert np.array_equal(a.get_support(), b.get_support())


CASES = [
    ((64, 64), dict(starts=4)),  # whole-image kernel
    ((256, 256), dict(starts=4, storage="f16", mixed=True)),
    ((48, 40), dict(starts=2)),  # generic passes
    ((16, 64, 64), dict(starts=2, storage="bf16", mixed=True)),  # plane pass
    ((128, 128), dict(starts=4, batch=2, lanes=2)),  # launch groups on two lanes
    ((32, 32, 32), dict(starts=2, plane=2)),  # 3D line passes
]


@pytest.mark.parametrize("shape,kw", CASES)
def test_whole_recipe_equals_the_step_by_step_walk(qpch, shape, kw):
    a = _solver(qpch, shape, **kw)
    b = _solver(qpch, shape, **kw)
    group = qpch.Group([b])
    total = R.iterations(RECIPE)
    for rep in range(2):  # captured, then replayed
        want = _stepwise(a, RECIPE)
        b.begin_run(total)
        group.run(RECIPE)
        got = b.read_errors(0, total)
        assert np.array_equal(got, want), (shape, rep, np.max(np.abs(got - want)))
        _same(a, b)
    group.close()
    a.close()
    b.close()


def test_group_of_shapes_equals_separate_walks(qpch):
    shapes = [((64, 64), dict(starts=2)), ((256, 256), dict(starts=2, storage="f16", mixed=True)),
              ((16, 64, 64), dict(starts=2))]
    singles = [_solver(qpch, s, **kw) for s, kw in shapes]
    members = [_solver(qpch, s, **kw) for s, kw in shapes]
    group = qpch.Group(members)
    total = R.iterations(RECIPE)
    for rep in range(2):
        wants = [_stepwise(s, RECIPE) for s in singles]
        for m in members:
            m.begin_run(total)
        group.run(RECIPE)
        for m, s, want in zip(members, singles, wants):
            assert np.array_equal(m.read_errors(0, total), want), rep
            _same(m, s)
    group.close()
    for s in singles + members:
        s.close()


def test_whole_recipe_needs_room_for_its_iterations(qpch):
    s = _solver(qpch, (64, 64), starts=2)
    group = qpch.Group([s])
    s.begin_run(5)
    with pytest.raises(ValueError):
        group.run(RECIPE)
    group.close()
    s.close()
