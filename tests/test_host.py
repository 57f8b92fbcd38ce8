"""Host-side logic of the product without a GPU: the C ABI library loads and exports what
include/qpch.h declares, the C++ API validates inputs like the reference
(python/tests/test_api.py, tests/test_problem.cpp, tests/test_recipe.cpp), and the
Python recipe flattens like the reference's."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qpch_[a-z_0-9]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    from qoraphasecu_amd import _qpch

    names = declared_functions("qpch.h")
    assert len(names) >= 20
    lib = ctypes.CDLL(_qpch.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_c_abi_reports_bad_arguments_without_a_device():
    from qoraphasecu_amd import _qpch

    with pytest.raises(ValueError):
        _qpch.fft(np.zeros((3,), np.complex64))  # 1D is rejected before any device call


def test_core_module_imports():
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import _core

    for name in ("reconstruct", "autocorrelation_support", "aligned_error", "half_support", "make_problem",
                 "hip_available", "cuda_available", "Solver"):
        assert hasattr(_core, name), name
    assert qpc.cuda_available() is False


def test_recipe_flattens_like_the_reference():
    import qoraphasecu_amd as qpc

    recipe = qpc.default_recipe()
    kinds = [op[0] for op in recipe.ops()]
    assert kinds.count("raar") == 1 + 1 + 8
    assert kinds.count("shrinkwrap") == 8
    assert kinds[1:4] == ["detwin_begin", "hio", "detwin_end"]
    assert recipe.iterations() == 720
    r = qpc.Recipe(qpc.HIO(5), qpc.repeat(3, qpc.RAAR(10), qpc.Shrinkwrap()), qpc.detwin(qpc.ER(2)))
    assert len(r.ops()) == 1 + 6 + 3 and r.iterations() == 37
    with pytest.raises(ValueError):
        qpc.Shrinkwrap(method="median")


def test_engine_errors_surface_as_value_errors():
    import qoraphasecu_amd as qpc

    inten = np.ones((16, 16), np.float32)
    sup = np.ones((16, 16), bool)
    with pytest.raises(ValueError):
        qpc.reconstruct(inten, support=sup, recipe=qpc.Recipe(qpc.ER(0)), backend="hip")
    with pytest.raises(ValueError):
        qpc.reconstruct(inten, support=sup, starts=4, batch=3, backend="hip")
    with pytest.raises(ValueError, match=r"support\.shape"):
        qpc.reconstruct(inten, support=np.ones((3, 3), bool), backend="hip")
    with pytest.raises(ValueError):
        qpc.reconstruct(np.ones(16, np.float32), backend="hip")
    with pytest.raises(ValueError):
        qpc.reconstruct(inten, support=sup, variant="warp", backend="hip")
    with pytest.raises(ValueError):
        qpc.reconstruct(inten, support=sup, backend="tpu")


def test_make_problem_validation():
    from qoraphasecu_amd import _core

    four = np.ones((2, 2), np.float32)
    with pytest.raises(ValueError):
        _core.make_problem(four, None, np.ones((2, 2), np.uint8), 0.1)  # every pixel masked
    nan = four.copy()
    nan[0, 1] = np.nan
    with pytest.raises(ValueError):
        _core.make_problem(nan, np.ones((2, 2), np.uint8), None, 0.1)
    amp, mask, sup = _core.make_problem(np.array([[4.0, -1.0], [9.0, 0.0]], np.float32),
                                        np.array([[1, 1], [0, 0]], np.uint8), None, 0.1)
    assert np.array_equal(amp, np.array([[2.0, 0.0], [3.0, 0.0]], np.float32))
    assert np.array_equal(mask, np.zeros((2, 2), np.uint8))
    assert np.array_equal(sup, np.array([[1, 1], [0, 0]], np.uint8))


def test_half_support_known_answer():
    # tests/test_problem.cpp:62-77, through the product's C++ half_support
    from qoraphasecu_amd import _core

    support = np.zeros((8, 4), np.uint8)
    support[2:6, :] = 1
    half = _core.half_support(support, 0)
    want = np.zeros((8, 4), np.uint8)
    want[2:4, :] = 1
    assert np.array_equal(half, want)
    cols = _core.half_support(support, 1)
    assert cols[2, 1] == 1 and cols[2, 2] == 0
    with pytest.raises(ValueError):
        _core.half_support(support, 2)


def test_cpu_and_cuda_backends_are_the_references():
    import qoraphasecu_amd as qpc

    inten = np.ones((16, 16), np.float32)
    for backend in ("cpu", "cuda"):
        with pytest.raises(RuntimeError):
            qpc.reconstruct(inten, support=np.ones((16, 16), bool), recipe=qpc.Recipe(qpc.ER(1)), backend=backend)


def test_bench_refuses_several_gpus():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
