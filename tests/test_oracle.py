"""Pins the CPU oracle: golden vectors (numpy's pocketfft, the reference's own compiled
pointwise/layout/recipe code), brute-force DFTs, closed forms, and the reference's
operator / problem / recipe / recovery tests (tests/test_operators.cpp,
tests/test_problem.cpp, tests/test_recipe.cpp, tests/test_recover.cpp)."""

import os

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dft_matrix(n, forward):
    k = np.arange(n)
    sign = -1.0 if forward else 1.0
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) / np.sqrt(n)


def brute_fftn(x, forward):
    """Unitary DFT by explicit matrix products along each axis (double)."""
    y = x.astype(np.complex128)
    for ax in range(y.ndim):
        y = np.moveaxis(np.tensordot(dft_matrix(y.shape[ax], forward), np.moveaxis(y, ax, 0), axes=(1, 0)), 0, ax)
    return y


# --- FFT ----------------------------------------------------------------------------------


def test_fft_matches_pocketfft_golden():
    g = np.load(os.path.join(GOLDEN, "fft_golden.npz"))
    k = 0
    while f"x{k}" in g:
        x = g[f"x{k}"]
        for fwd, key in ((True, "fwd"), (False, "inv")):
            d = oracle.fft(x, fwd, "d")
            assert rel(d, g[f"{key}64_{k}"]) < 1e-14, (x.shape, key)
            f = oracle.fft(x, fwd, "f")
            assert rel(f, g[f"{key}64_{k}"]) < 1e-6, (x.shape, key)
            assert rel(f, g[f"{key}32_{k}"]) < 1e-6, (x.shape, key)
        k += 1
    assert k >= 6


@pytest.mark.parametrize("shape", [(3, 5), (7, 4), (2, 9, 5), (6, 6), (11, 13), (67, 2)])
def test_fft_matches_brute_force_dft(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for fwd in (True, False):
        assert rel(oracle.fft(x, fwd, "d"), brute_fftn(x, fwd)) < 1e-13


def test_fft_invariants():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((24, 20)) + 1j * rng.standard_normal((24, 20))
    y = oracle.fft(x, True, "d")
    assert abs(np.linalg.norm(y) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)  # Parseval (unitary)
    assert rel(oracle.fft(y, False, "d"), x) < 1e-14  # round trip
    delta = np.zeros((8, 12), complex)
    delta[0, 0] = 1.0
    assert np.allclose(oracle.fft(delta, True, "d"), 1.0 / np.sqrt(96))  # delta <-> constant
    # a planted sign error (inverse twiddles on the forward transform) fails the golden check
    assert rel(oracle.fft(x, False, "d"), np.fft.fftn(x, norm="ortho")) > 0.1


# --- pointwise / layout / recipe: the reference's own compiled code ------------------------


def test_hash_matches_reference_build():
    g = np.load(os.path.join(GOLDEN, "ref_golden.npz"))
    for z, want in zip(g["mix64_in"], g["mix64_out"]):
        assert oracle.mix64(int(z)) == int(want)
    for (s, st, ix), want in zip(g["uniform_in"], g["uniform_out"]):
        assert np.float32(oracle.uniform(int(s), int(st), int(ix))) == want


def test_coefficients_match_reference_build():
    g = np.load(os.path.join(GOLDEN, "ref_golden.npz"))
    kinds = {0: "er", 1: "hio", 2: "raar"}
    for kind, beta, a, b in g["coefficients"]:
        ca, cb = R.coefficients(kinds[int(kind)], float(np.float32(beta)))
        assert np.float32(ca) == np.float32(a) and np.float32(cb) == np.float32(b)


def test_roll_matches_reference_build():
    g = np.load(os.path.join(GOLDEN, "ref_golden.npz"))
    for name, shape in {"roll_547": (5, 4, 7), "roll_65": (6, 5), "roll_44": (4, 4), "roll_333": (3, 3, 3)}.items():
        src = np.arange(int(np.prod(shape)), dtype=np.uint32).reshape(shape)
        assert np.array_equal(oracle.to_centered(src), g[f"{name}_0"])
        assert np.array_equal(oracle.to_fft_order(src), g[f"{name}_1"])
        assert np.array_equal(oracle.to_fft_order(src), np.fft.ifftshift(src))
        assert np.array_equal(oracle.to_centered(oracle.to_fft_order(src)), src)


def test_default_recipe_matches_reference_build():
    g = np.load(os.path.join(GOLDEN, "ref_golden.npz"))
    kinds = {"er": 0, "hio": 1, "raar": 2, "shrinkwrap": 3, "detwin_begin": 4, "detwin_end": 5}
    ops = R.default_recipe()
    assert [kinds[o[0]] for o in ops] == list(g["default_recipe_kinds"])
    assert R.iterations(ops) == int(g["default_recipe_iterations"]) == 720


def test_validate_matches_reference_build():
    g = np.load(os.path.join(GOLDEN, "ref_golden.npz"))
    cases = [
        [],
        [R.er(0)],
        [("detwin_begin", 0, 0.0, 0.0, 0.0, "max", 0), R.er(1)],
        [("detwin_end", 0, 0.0, 0.0, 0.0, "max", 0)],
        R.detwin(*R.detwin(R.er(1))),
        [R.shrinkwrap(threshold=0.0)],
        R.default_recipe(),
    ]
    obj, sup, inten = synthetic.triangle((16, 16), 4)
    amp, mask, sup = oracle.make_problem(inten, sup)
    for ops, rejects in zip(cases, g["validate_rejects"]):
        s = oracle.Solver(amp, mask, sup)
        if rejects:
            with pytest.raises(ValueError):
                s.run(ops)
        elif ops:
            s.run(ops[:1])


# --- operators (tests/test_operators.cpp) ---------------------------------------------------

SHAPES = [(32, 32), (17, 24), (16, 16, 16)]


def _op_problem(shape, masked=()):
    obj, sup, inten = synthetic.triangle(shape, 8)
    amp, mask, sup = oracle.make_problem(inten, sup)
    mask = mask.copy()
    for i in masked:
        mask.flat[i] = 1
    return obj, oracle.to_fft_order(amp), oracle.to_fft_order(mask), oracle.to_fft_order(sup)


def _project(x, af, mf):
    w, err = oracle.modulus(oracle.fft(oracle.to_fft_order(x), True), af, mf)
    return oracle.to_centered(oracle.fft(w, False)), err


@pytest.mark.parametrize("shape", SHAPES)
def test_modulus_projection_is_idempotent(shape):
    obj, af, mf, _ = _op_problem(shape)
    x = synthetic.random_field(obj.size, 1).reshape(shape)
    once, _ = _project(x, af, mf)
    twice, err = _project(once, af, mf)
    assert err < 1e-5
    assert np.max(np.abs(twice - once)) < 1e-5


@pytest.mark.parametrize("shape", SHAPES)
def test_masked_pixels_keep_their_magnitude(shape):
    n = int(np.prod(shape))
    obj, af, mf, _ = _op_problem(shape, masked=(0, n // 2))
    _, af0, mf0, _ = _op_problem(shape)
    x = synthetic.random_field(n, 2).reshape(shape)
    a, _ = _project(x, af, mf)
    b, _ = _project(x, af0, mf0)
    assert np.max(np.abs(a - b)) > 1e-4
    _, err = _project(a, af, mf)
    assert err < 1e-5


@pytest.mark.parametrize("shape", SHAPES)
def test_error_is_zero_at_the_true_object(shape):
    obj, af, mf, _ = _op_problem(shape)
    p, err = _project(obj, af, mf)
    assert err < 1e-5
    assert np.max(np.abs(p - obj)) < 1e-4


def test_modulus_closed_form():
    rng = np.random.default_rng(5)
    w = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    a = rng.uniform(0.1, 2.0, 64).astype(np.float32)
    m = (rng.random(64) < 0.2).astype(np.uint8)
    got, num = oracle.modulus_num(w, a, m)
    mag = np.abs(w.astype(np.complex128))
    want = np.where(m == 1, w, a / (mag + 1e-12) * w)
    assert rel(got, want) < 1e-6
    assert abs(num - np.sum(((mag - a) ** 2)[m == 0])) < 1e-5 * num


def test_update_closed_forms():
    rng = np.random.default_rng(6)
    x = (rng.standard_normal(50) + 1j * rng.standard_normal(50)).astype(np.complex64)
    p = (rng.standard_normal(50) + 1j * rng.standard_normal(50)).astype(np.complex64)
    s = (rng.random(50) < 0.5).astype(np.uint8)
    for kind, beta in (("er", 0.9), ("hio", 0.9), ("raar", 0.9), ("raar", 0.5)):
        a, b = R.coefficients(kind, beta)
        got = oracle.update(x, p, s, a, b)
        if kind == "er":
            want = np.where(s == 1, p, 0)
        elif kind == "hio":
            want = np.where(s == 1, p, x - beta * p)
        else:  # (beta/2)(R_S R_M x + x) + (1 - beta) P_M x outside the support
            want = np.where(s == 1, p, beta * x + (1 - 2 * beta) * p)
        assert np.max(np.abs(got - want)) < 1e-6, kind


@pytest.mark.parametrize("shape", SHAPES)
def test_support_projection_is_idempotent(shape):
    obj, sup, _ = synthetic.triangle(shape, 8)
    x = synthetic.random_field(obj.size, 3).reshape(shape)
    once = np.where(sup != 0, x, 0)
    assert np.array_equal(np.where(sup != 0, once, 0), once)
    assert np.all(once[sup == 0] == 0)


@pytest.mark.parametrize("shape", SHAPES)
def test_blur_preserves_the_sum(shape):
    image = np.zeros(shape, np.float32)
    image.flat[image.size // 2 + 3] = 1.0
    out = oracle.blur(image, 1.5)
    assert abs(float(out.astype(np.float64).sum()) - 1.0) < 1e-5
    assert out.max() < 0.5


def test_blur_is_the_separable_gaussian():
    """The FFT-domain blur equals the circular convolution with the per-axis kernels."""
    rng = np.random.default_rng(7)
    img = rng.random((12, 10))
    sigma = float(np.float32(1.3))  # the oracle takes sigma as a float

    def kernel(n):
        f = np.fft.fftfreq(n)
        return np.fft.ifft(np.exp(-2 * np.pi**2 * sigma**2 * f * f)).real

    want = img
    for ax in range(2):
        k = kernel(img.shape[ax])
        want = np.apply_along_axis(lambda v: np.real(np.fft.ifft(np.fft.fft(v) * np.fft.fft(k))), ax, want)
    assert np.max(np.abs(oracle.blur(img, sigma, "d") - want)) < 1e-13


# --- problem (tests/test_problem.cpp) -------------------------------------------------------


def test_center_moves_to_origin():
    v = np.zeros((6, 5), np.int32)
    v[3, 2] = 1
    assert oracle.to_fft_order(v)[0, 0] == 1


def test_make_problem_clips_negative_intensity():
    amp, mask, sup = oracle.make_problem(np.array([[4.0, -1.0], [9.0, 0.0]], np.float32),
                                         np.array([[1, 1], [0, 0]], np.uint8))
    assert np.array_equal(amp, np.array([[2.0, 0.0], [3.0, 0.0]], np.float32))
    assert np.array_equal(mask, np.zeros((2, 2), np.uint8))


def test_half_support_keeps_the_lower_half():
    support = np.zeros((8, 4), np.uint8)
    support[2:6, :] = 1
    half = oracle.half_support(support, 0)
    want = np.zeros((8, 4), np.uint8)
    want[2:4, :] = 1
    assert np.array_equal(half, want)
    cols = oracle.half_support(support, 1)
    assert cols[2, 1] == 1 and cols[2, 2] == 0


def test_autocorrelation_support_covers_the_object():
    obj, sup, inten = synthetic.triangle((64, 64), 16)
    ac = oracle.autocorrelation_support(inten, 0.1)
    inside = int(sup.sum())
    assert int((sup & ac).sum()) > inside // 2
    assert int(ac.sum()) > inside


def test_simulate_intensity_matches_numpy():
    obj, _, inten = synthetic.triangle((40, 36), 10)
    want = np.abs(np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(obj.astype(np.complex128)), norm="ortho"))) ** 2
    assert rel(inten, want) < 1e-6


# --- recovery (tests/test_recover.cpp) -------------------------------------------------------


def _best_aligned(shape, m, ops, starts, prec="f"):
    obj, sup, inten = synthetic.triangle(shape, m)
    runs = oracle.reconstruct(inten, ops, support=sup, starts=starts, seed=7, prec=prec)
    return min(oracle.aligned_error(r[0], obj, prec) for r in runs)


def test_aligned_error_ignores_shift_phase_and_twin():
    obj, _, _ = synthetic.triangle((32, 32), 10)
    moved = np.empty_like(obj)
    rot = np.complex64(np.exp(0.7j))
    for y in range(32):
        for x in range(32):
            moved[y, x] = rot * np.conj(obj[(32 - y) % 32, (32 - x + 3) % 32])
    assert oracle.aligned_error(moved, obj) < 1e-5
    assert oracle.aligned_error(synthetic.random_field(obj.size, 4).reshape(obj.shape), obj) > 0.5


def test_given_start_replaces_the_random_one():
    obj, sup, inten = synthetic.triangle((32, 24), 8)
    start = np.stack([obj, synthetic.random_field(obj.size, 3).reshape(obj.shape)])
    a = oracle.reconstruct(inten, [R.er(5)], support=sup, starts=2, seed=1, start=start)
    b = oracle.reconstruct(inten, [R.er(5)], support=sup, starts=2, seed=2, start=start)
    assert a[0][2][-1] < 1e-5
    assert np.max(np.abs(a[0][0] - obj)) < 1e-5
    assert np.array_equal(a[1][0], b[1][0])


def test_hio_then_er_recovers_2d():
    assert _best_aligned((64, 64), 20, [R.hio(300), R.er(100)], 4) < 1e-2


def test_er_never_increases_the_error():
    obj, sup, inten = synthetic.triangle((64, 64), 20)
    err = oracle.reconstruct(inten, [R.er(100)], support=sup)[0][2]
    assert err.size == 100 and err[-1] < err[0]
    assert np.all(err[1:] <= err[:-1] * (1 + 1e-4))


@pytest.mark.slow
def test_raar_recovers_3d():
    assert _best_aligned((24, 24, 24), 8, [R.raar(300), R.er(100)], 2) < 1e-2


def test_default_recipe_without_support():
    obj, _, inten = synthetic.triangle((64, 64), 24)
    runs = oracle.reconstruct(inten, R.default_recipe(), starts=8)
    best = min(runs, key=lambda r: r[2][-1])
    assert best[2][-1] < 1e-3
    assert oracle.aligned_error(best[0], obj) < 1e-2


def test_run_continues_from_the_current_state():
    obj, sup, inten = synthetic.triangle((32, 32), 8)
    amp, mask, sup = oracle.make_problem(inten, sup)
    s = oracle.Solver(amp, mask, sup)
    s.run([R.hio(10)])
    s.run([R.er(5)])
    split = s.results()[0]
    whole = oracle.reconstruct(inten, [R.hio(10), R.er(5)], support=sup)[0]
    assert np.array_equal(split[2], whole[2])
    assert np.array_equal(split[0], whole[0])


def test_fp64_twin_agrees_with_float_on_a_short_run():
    obj, sup, inten = synthetic.triangle((48, 40), 12)
    ops = [R.hio(15), R.shrinkwrap(), R.raar(10), R.er(10)]
    f = oracle.reconstruct(inten, ops, support=sup, starts=2, seed=3)
    d = oracle.reconstruct(inten, ops, support=sup, starts=2, seed=3, prec="d")
    for (fo, fs, fe), (do, ds, de) in zip(f, d):
        assert np.all(np.abs(fe - de) <= 1e-4 * de + 1e-7)
        assert rel(fo, do) < 1e-4
        assert np.array_equal(fs, ds)


def test_documented_cpu_accuracy_64_default_recipe():
    """docs/index.md:89,93: 64^2 triangle, default recipe, 8 starts, no support -> best
    Fourier error ~1.4e-7 on the reference's CPU backend."""
    obj, _, inten = synthetic.triangle((64, 64), 24)
    runs = oracle.reconstruct(inten, R.default_recipe(), starts=8)
    best = min(r[2][-1] for r in runs)
    assert best < 5e-7
