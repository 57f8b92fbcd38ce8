"""Degenerate inputs through the fused kernels, against the oracle from the same starts.

- empty support: every pixel takes the feedback branch a x + b P_M x;
- full support: x = P_M x everywhere;
- every pixel masked: no usable energy, the projection keeps every magnitude, and the
  error is sqrt(0 / (0 + 1e-12)) = 0;
- zero intensity: the projection zeroes the spectrum, and the error divides by the 1e-12
  guard alone;
- the true object as the start, with its own support: the error is zero up to rounding.

Reference semantics: src/cpu/kernels.cpp:16-50 (energy, modulus, update) and
src/core/pointwise.hpp:16 (the 1e-12 guard).  Shapes cover the whole-image kernel (64 x 64),
the line passes (256 x 256), the generic passes (48 x 40) and the 3D plane pass
(16 x 64 x 64).  fp32 errors agree with the oracle within 1e-5 relative plus, where the
error itself is at the float floor, an absolute allowance of max(2e-7, 3 x the oracle's own
float-vs-FP64 distance on that iteration); iterates within 1e-5 (norm-relative, or absolute
against the start's norm when the iterate vanishes).  fp16 storage on the same
cases stays finite and, where the error is zero for fp32, zero."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (256, 256), (48, 40), (16, 64, 64)]
CASES = ["empty_support", "full_support", "all_masked", "zero_intensity", "true_object"]
OPS = [R.hio(3), R.raar(3), R.er(3)]


def _case(shape, case):
    obj, sup, inten = synthetic.triangle(shape, max(4, min(shape[-2:]) // 4))
    mask = None
    if case == "all_masked":
        mask = np.ones(shape, np.uint8)
    if case == "zero_intensity":
        inten = np.zeros(shape, np.float32)
    amp, m, s = oracle.make_problem(inten, sup, mask)
    if case == "empty_support":
        s = np.zeros(shape, np.uint8)
    elif case == "full_support":
        s = np.ones(shape, np.uint8)
    if case == "true_object":
        start = np.stack([obj, obj]).astype(np.complex64)
    else:
        start = np.stack([synthetic.random_field(obj.size, 31).reshape(shape),
                          synthetic.random_field(obj.size, 32).reshape(shape)])
    return amp, m, s, start


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("shape", SHAPES)
def test_degenerate_inputs_match_oracle(qpch, shape, case):
    amp, mask, sup, start = _case(shape, case)
    res = {}
    for prec in ("f", "d"):
        o = oracle.Solver(amp, mask, sup, starts=2, start=start, prec=prec)
        o.run(OPS)
        res[prec] = o.results()
    ores = res["f"]
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]), schedule="graph")
    errs = R.run_abi(s, OPS).astype(np.float64)
    xs = s.read_state()
    s.close()
    scale = max(float(np.linalg.norm(start[0])), 1e-30)
    for k in range(2):
        want = np.asarray(ores[k][2], np.float64)
        drift = np.abs(want - np.asarray(res["d"][k][2], np.float64))
        d = np.abs(errs[:, k] - want)
        print(f"{shape} {case} start {k}: errors {want[[0, -1]]}, largest difference {d.max():.2e} "
              f"(float oracle vs FP64 {drift.max():.2e})")
        assert np.all(d <= 1e-5 * want + np.maximum(2e-7, 3.0 * drift)), (shape, case, k, d.max())
        if case == "all_masked":
            assert np.all(errs[:, k] == 0.0)
        x_want = f(ores[k][0])
        norm = float(np.linalg.norm(x_want))
        diff = float(np.linalg.norm(xs[k] - x_want))
        assert diff <= 1e-5 * max(norm, scale * 1e-3), (shape, case, k, diff, norm)


@pytest.mark.parametrize("case", ["empty_support", "all_masked", "zero_intensity", "true_object"])
@pytest.mark.parametrize("shape", [(64, 64), (16, 64, 64)])
def test_degenerate_inputs_stay_finite_in_fp16(qpch, shape, case):
    amp, mask, sup, start = _case(shape, case)
    f = oracle.to_fft_order
    s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]), schedule="graph",
                    storage="f16", mixed=True)
    errs = R.run_abi(s, OPS)
    xs = s.read_state()
    s.close()
    assert np.all(np.isfinite(errs)) and np.all(np.isfinite(xs)), (shape, case)
    if case == "all_masked":
        assert np.all(errs == 0.0)
