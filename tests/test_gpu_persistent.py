"""The persistent two-pass launches (twopass_iterations: every iteration of a chunk in one
launch, the tiles of both passes claimed from per-XCD queues) against the launch-per-pass
schedule (qpch_problem.persist = 2) on the same inputs: error histories, iterates and
supports equal bit for bit, over recipes with Shrinkwrap, detwin and ER stages, for fp32 and
fp16 / bf16 storage (ER in fp32), starts that do not fill the eight XCDs evenly, shapes
whose row tiles are resized to the column pass's work-group, launch groups running as
concurrent persistent launches on lanes, images whose two axes are not powers of two
(twopass_generic: mixed-radix, plain-DFT and Bluestein axes; twopass_smooth: square images
of a smooth length, whose passes hold the line in registers), and whole recipes of several
shapes in one graph (their persistent launches run side by side).  Also against the oracle: the per-iteration
error of a plain persistent run within 1e-5 of the oracle's float run."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic
from test_gpu_whole_recipe import RECIPE, _same, _stepwise

pytestmark = pytest.mark.gpu


def _solver(qpch, shape, persist, **kw):
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    amp, mask, sup = oracle.make_problem(inten, None, None)
    f = oracle.to_fft_order
    return qpch.Solver(f(amp), f(mask), f(sup), seed=5, schedule="graph", persist=persist, **kw)


CASES = [
    ((256, 256), dict(starts=32)),  # the headline shape: 4 starts per XCD
    ((256, 256), dict(starts=12, storage="f16", mixed=True)),  # XCDs with two starts and with one
    ((256, 256), dict(starts=3, storage="bf16", mixed=True)),  # five XCDs without work
    ((256, 256), dict(starts=2, batch=1)),  # one start per launch (the plain path)
    ((128, 128), dict(starts=8)),
    ((512, 512), dict(starts=4, storage="f16", mixed=True)),  # row tiles of 16 lines
    ((128, 512), dict(starts=4)),
    ((1024, 64), dict(starts=2)),  # column work-group of 512 threads
    ((64, 1024), dict(starts=2, storage="f16")),
    ((64, 64), dict(starts=8, storage="f16", mixed=True)),  # fp16 64^2 (fp32 runs whole images)
    ((256, 256), dict(starts=16, batch=4, lanes=4)),  # four persistent launches at once
    # twopass_generic: both axes of primes 2, 3 and 5 only, neither a power of two nor a
    # smooth length (fp32)
    ((243, 243), dict(starts=8)),  # 3^5
    ((27, 486), dict(starts=3, batch=1)),
    ((405, 243), dict(starts=4)),  # 5 * 81, 3^5
    ((405, 27), dict(starts=2)),  # 27 lines of 405: ragged tiles
    ((252, 252), dict(starts=8)),  # a factor 7: the chirp-z passes, per pass either way
    ((125, 125), dict(starts=8)),  # 25 * 5: smooth passes, per pass either way
    ((100, 100), dict(starts=8)),  # 10 * 10: twopass_smooth
    ((54, 54), dict(starts=3, batch=1)),  # 6 * 9: twopass_smooth
    ((70, 70), dict(starts=8)),  # 10 * 7: twopass_smooth
    ((42, 42), dict(starts=3, batch=1)),  # 6 * 7
    ((216, 216), dict(starts=8)),  # 8 * 27: smooth passes, per pass either way
    # twopass_smooth: square images of a smooth length (smooth.hpp)
    ((240, 240), dict(starts=8)),  # 16 * 15
    ((48, 48), dict(starts=3, batch=1)),  # 16 * 3
    ((120, 120), dict(starts=12)),  # 8 * 15
    ((30, 30), dict(starts=8)),  # 2 * 15: two tiles of 16 lines, the second ragged
    ((160, 160), dict(starts=8)),  # 32 * 5: the power-of-two stage through the line engine
    ((96, 96), dict(starts=8)),  # 32 * 3: the line engine at radix 4
    ((36, 36), dict(starts=8)),  # 4 * 9 (radix 9 = 3 x 3)
    ((72, 72), dict(starts=3, batch=1)),  # 8 * 9
    ((18, 18), dict(starts=8)),  # 2 * 9
    ((56, 56), dict(starts=8)),  # 8 * 7 (the prime short form)
    ((44, 44), dict(starts=3, batch=1)),  # 4 * 11
    ((26, 26), dict(starts=8)),  # 2 * 13
    ((480, 480), dict(starts=4)),  # 32 * 15: no persistent kernel, the per-pass launches either way
    ((80, 240), dict(starts=4)),  # smooth, not square: the per-pass launches either way
]


@pytest.mark.parametrize("shape,kw", CASES)
def test_persistent_equals_pass_launches(qpch, shape, kw):
    a = _solver(qpch, shape, 2, **kw)
    b = _solver(qpch, shape, 1, **kw)  # 1: the generic shapes' persistent launches too
    for rep in range(2):  # the graphs captured, then replayed
        want = _stepwise(a, RECIPE)
        got = _stepwise(b, RECIPE)
        assert np.array_equal(got, want), (shape, kw, rep, np.argwhere(got != want)[:4])
        _same(a, b)
    a.close()
    b.close()


def test_persistent_whole_recipes_side_by_side(qpch):
    shapes = [((128, 128), dict(starts=8)), ((256, 256), dict(starts=8, storage="f16", mixed=True)),
              ((512, 512), dict(starts=4))]
    singles = [_solver(qpch, s, 2, **kw) for s, kw in shapes]
    members = [_solver(qpch, s, 0, **kw) for s, kw in shapes]
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


@pytest.mark.parametrize("kind", ["er", "hio", "raar"])
def test_persistent_per_iteration_error_matches_oracle(qpch, kind):
    """Chunks of 10 and 20 iterations (two persistent launches) from given starts: the
    first ten errors of each start within 1e-5 of the oracle's float run (beyond them,
    float rounding alone separates equally valid float runs; the whole-run bars are in
    test_gpu_per_iteration.py, whose 256 x 256 cases run these launches too)."""
    shape = (128, 128)
    obj, sup, inten = synthetic.triangle(shape, 32)
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    start = np.stack([synthetic.random_field(obj.size, 31).reshape(shape),
                      synthetic.random_field(obj.size, 32).reshape(shape)])
    ops = [(kind, 30, 0.9, 0.0, 0.0, "max", 0)]
    o = oracle.Solver(amp, mask, sup, starts=2, start=start, prec="f")
    o.run(ops)
    want = [np.asarray(r[2], np.float64) for r in o.results()]
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]), schedule="graph",
                    persist=0)
    errs = R.run_abi(s, ops, sync_every=20).astype(np.float64)
    s.close()
    for k in range(2):
        d = np.abs(errs[:10, k] - want[k][:10]) / want[k][:10]
        assert np.all(d <= 1e-5), (kind, k, d.max())


def test_a_wait_past_its_bound_is_reported(qpch, monkeypatch):
    """The safety net of the waits, forced from the host: with a bound of zero clock ticks,
    the first wait that finds its dependency unfinished gives up and sets the abort flag
    (the row tiles of a start wait on its column tiles at least once per launch); the
    launch still runs to its end, and the host reports a device failure."""
    monkeypatch.setenv("QPCH_TEST_WAIT_TICKS", "0")
    s = _solver(qpch, (256, 256), 0, starts=32)
    s.begin_run(20)
    s.iterate(20, 0.9, -0.8)
    with pytest.raises(qpch.DeviceError, match="persistent"):
        s.read_errors(0, 20)
    s.close()


def test_an_xcd_without_workers_is_reported(qpch, monkeypatch):
    """The safety net of the queues, forced from the host: one work-group (on XCD 0) runs
    its own queue's items and nothing else; the last worker finds the other XCDs' queues
    unclaimed and sets the abort flag, so starts left undone never pass for results."""
    monkeypatch.setenv("QPCH_TEST_PERSIST_WORKERS", "1")
    s = _solver(qpch, (128, 128), 0, starts=8)
    s.begin_run(10)
    s.iterate(10, 0.9, -0.8)
    with pytest.raises(qpch.DeviceError, match="persistent"):
        s.read_errors(0, 10)
    s.close()
