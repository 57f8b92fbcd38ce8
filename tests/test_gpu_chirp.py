"""The chirp-z passes (chirp.hpp: axes of 32 to 2048 with a prime factor of 7 or more, Bluestein's algorithm
with the power-of-two transforms held in registers by the radix-16 line engine) against
the oracle: one teacher-forced HIO, RAAR and ER iteration from the same starts, error and
object within 1e-5 relative, for every transform length M (256 to 4096) as a column
(strided) and as a row axis, composite lengths with a large prime factor, beside
power-of-two and smooth passes, as 3D z axes and as 3D y axes (forward and inverse line
passes); and RAAR(30) runs, the per-iteration
error within the bar of test_gpu_per_iteration.py (1e-5, or float rounding's measured
spread)."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic
from test_gpu_abi import _problem, rel

pytestmark = pytest.mark.gpu

# N -> M: 67 -> 256, 131 -> 512, 257 -> 1024, 521 and 1021 -> 2048, 2039 -> 4096;
# 134 = 2 * 67 and 201 = 3 * 67 (M 512); factors 7 to 64: 116 = 4 * 29, 248 = 8 * 31,
# 532 = 4 * 7 * 19, 868 = 4 * 7 * 31 (M 256 to 2048)
SHAPES = [(67, 16), (16, 67), (131, 16), (16, 131), (257, 16), (16, 257), (521, 8), (8, 521), (1021, 4), (4, 1021),
          (2039, 2), (2, 2039), (134, 24), (24, 201), (67, 30), (131, 240), (67, 16, 16), (8, 4, 389), (131, 8, 8),
          (116, 64), (64, 116), (248, 248), (532, 36), (36, 532), (868, 8), (8, 868), (248, 16, 16),
          # y axes of 3D shapes (forward and inverse line passes): 248, 67, 131, 1021
          (16, 248, 116), (8, 67, 32), (4, 131, 24), (2, 1021, 4)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("schedule", ["graph", "fused"])
def test_chirp_teacher_forced_iteration(qpch, shape, schedule):
    obj, amp, mask, sup = _problem(shape, max(2, min(shape[-2:]) // 4), masked=True)
    start = np.stack([synthetic.random_field(obj.size, 61).reshape(shape),
                      synthetic.random_field(obj.size, 62).reshape(shape)])
    sf = np.stack([oracle.to_fft_order(s) for s in start])
    for kind, ops in [("hio", [R.hio(1)]), ("raar", [R.raar(1)]), ("er", [R.er(1)])]:
        o = oracle.Solver(amp, mask, sup, starts=2, start=start)
        o.run(ops)
        ores = o.results()
        s = qpch.Solver(oracle.to_fft_order(amp), oracle.to_fft_order(mask), oracle.to_fft_order(sup),
                        starts=2, start=sf, schedule=schedule)
        errs = R.run_abi(s, ops)
        xs = s.read_state()
        s.close()
        for k in range(2):
            d = abs(errs[0, k] - ores[k][2][0]) / ores[k][2][0]
            xo = oracle.to_fft_order(ores[k][0])
            print(shape, schedule, kind, k, f"error {d:.1e} object {rel(xs[k], xo):.1e}")
            assert d <= 1e-5, (shape, kind, k, d)
            assert rel(xs[k], xo) < 1e-5, (shape, kind, k, rel(xs[k], xo))


@pytest.mark.parametrize("shape", [(131, 64), (64, 1021), (67, 16, 16), (248, 116), (8, 67, 32)])
def test_chirp_per_iteration_error(qpch, shape):
    """RAAR(30) from given starts: the device's error within max(1e-5, 3 x the spread of the
    oracle's float run against its FP64 run) of the oracle's float run at every iteration."""
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    start = np.stack([synthetic.random_field(obj.size, 71).reshape(shape),
                      synthetic.random_field(obj.size, 72).reshape(shape)])
    ops = [("raar", 30, 0.9, 0.0, 0.0, "max", 0)]
    hist = {}
    for prec in ("f", "d"):
        o = oracle.Solver(amp, mask, sup, starts=2, start=start, prec=prec)
        o.run(ops)
        hist[prec] = np.stack([np.asarray(r[2], np.float64) for r in o.results()], axis=1)
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]), schedule="graph")
    errs = R.run_abi(s, ops).astype(np.float64)
    s.close()
    d = np.max(np.abs(errs - hist["f"]) / hist["f"], axis=1)
    drift = np.maximum.accumulate(np.max(np.abs(hist["f"] - hist["d"]) / hist["d"], axis=1))
    bar = np.maximum(1e-5, 3.0 * drift)
    print(shape, f"device vs float oracle worst {d.max():.2e}; float oracle vs FP64 {drift.max():.2e}")
    assert np.all(d <= bar), (shape, d.max(), int(np.argmax(d > bar)))
