"""The per-iteration Fourier error over whole runs, against the oracle from the same starts
(fp32 storage): the 1e-5 relative bar on every iteration, beyond the single teacher-forced
step.

ER is contractive and HIO (on these problems) stays within 1e-5 of the oracle at every
iteration of the run.  RAAR amplifies rounding differences, as it does between the oracle's
own float and double runs, so its bar is 1e-5 over the first 40 iterations and 1e-4 over
all 50 (measured on the MI355X: within 1e-5 for 45 to 50 of them, worst 1.8e-5).  Shapes:
power-of-two 2D and 3D, a generic (48 x 40) shape, and a generic z axis before the plane
pass (24 x 64 x 64)."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (256, 256), (48, 40), (32, 32, 32), (24, 64, 64)]


def _worst(errs, ores):
    """Per iteration, the larger relative error difference of the two starts."""
    return np.max(np.stack([np.abs(errs[:, k] - ores[k][2]) / ores[k][2] for k in range(len(ores))]), axis=0)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("name,ops,bar,tail", [("ER100", [R.er(100)], 1e-5, 1e-5),
                                               ("HIO50", [R.hio(50)], 1e-5, 1e-5),
                                               ("RAAR50", [R.raar(50)], 1e-5, 1e-4)])
def test_per_iteration_error(qpch, shape, name, ops, bar, tail):
    obj, sup, inten = synthetic.triangle(shape, max(4, shape[-1] // 4))
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    start = np.stack([synthetic.random_field(obj.size, 21).reshape(shape),
                      synthetic.random_field(obj.size, 22).reshape(shape)])
    o = oracle.Solver(amp, mask, sup, starts=2, start=start)
    o.run(ops)
    ores = o.results()
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]), schedule="graph")
    errs = R.run_abi(s, ops)
    s.close()
    d = _worst(errs, ores)
    head = 40 if tail > bar else len(d)
    print(f"{shape} {name}: worst per-iteration relative difference {d.max():.2e} "
          f"(first {head}: {d[:head].max():.2e})")
    assert np.all(d[:head] <= bar), (shape, name, d[:head].max())
    assert np.all(d <= tail), (shape, name, d.max())
