"""ctypes front end of the CPU oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the CPU-baseline leg of ``bench.py``, always as the checker.  The product package
(``qoraphasecu_amd``) never imports it.

Every entry restates the reference's CPU backend; see oracle.h for the file:line
each one follows.  Arrays are numpy; complex arrays are complex64 for the ``"f"``
precision and complex128 for the FP64 twin ``"d"``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_KINDS = {
    "er": 0,
    "hio": 1,
    "raar": 2,
    "shrinkwrap": 3,
    "detwin_begin": 4,
    "detwin_end": 5,
}
_METHODS = {"max": 0, "rms": 1}


class _Op(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("n", ctypes.c_int),
        ("beta", ctypes.c_float),
        ("threshold", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("method", ctypes.c_int),
        ("axis", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (the recipe of oracle/Makefile)."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < max(
        os.path.getmtime(src),
        os.path.getmtime(os.path.join(_HERE, "oracle_impl.h")),
        os.path.getmtime(os.path.join(_HERE, "oracle.h")),
    ):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, sz, i, u64, f = (
            ctypes.c_void_p,
            ctypes.c_size_t,
            ctypes.c_int,
            ctypes.c_uint64,
            ctypes.c_float,
        )
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.or_mix64.restype = u64
        L.or_mix64.argtypes = [u64]
        L.or_uniform.restype = f
        L.or_uniform.argtypes = [u64, u64, u64]
        L.or_roll.argtypes = [vp, vp, sz, i, szp, i]
        L.or_half_support.argtypes = [vp, vp, i, szp, i]
        L.or_set_threads.argtypes = [i]
        for s in "fd":
            g = lambda name: getattr(L, f"{name}_{s}")  # noqa: E731
            g("or_fft").argtypes = [vp, i, szp, i]
            g("or_usable_energy").restype = ctypes.c_double
            g("or_usable_energy").argtypes = [vp, vp, sz]
            g("or_modulus").restype = ctypes.c_double
            g("or_modulus").argtypes = [vp, vp, vp, sz, ctypes.c_double]
            g("or_modulus_num").restype = ctypes.c_double
            g("or_modulus_num").argtypes = [vp, vp, vp, sz]
            g("or_update").argtypes = [vp, vp, vp, sz, f, f]
            g("or_blur").argtypes = [vp, vp, i, szp, f]
            g("or_shrinkwrap").argtypes = [vp, i, szp, f, f, i, vp]
            g("or_random_start").argtypes = [vp, vp, i, szp, u64, u64, vp]
            g("or_autocorrelation_support").argtypes = [vp, i, szp, f, vp]
            g("or_simulate_intensity").argtypes = [vp, i, szp, vp]
            g("or_aligned_error").restype = ctypes.c_double
            g("or_aligned_error").argtypes = [vp, vp, i, szp]
            g("or_solver_create").restype = vp
            g("or_solver_create").argtypes = [i, szp, vp, vp, vp, i, u64, vp]
            g("or_solver_run").restype = i
            g("or_solver_run").argtypes = [vp, ctypes.POINTER(_Op), i]
            g("or_solver_iterations").restype = i
            g("or_solver_iterations").argtypes = [vp]
            g("or_solver_results").argtypes = [vp, vp, vp, vp]
            g("or_solver_get_state").argtypes = [vp, vp, vp]
            g("or_solver_set_state").argtypes = [vp, vp, vp]
            g("or_solver_destroy").argtypes = [vp]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def _shape(shape):
    shape = tuple(int(s) for s in shape)
    return len(shape), (ctypes.c_size_t * len(shape))(*shape)


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _real(prec):
    return np.float32 if prec == "f" else np.float64


def _cplx(prec):
    return np.complex64 if prec == "f" else np.complex128


def _fn(name, prec):
    return getattr(lib(), f"{name}_{prec}")


# -- pointwise.hpp / layout.hpp / problem.cpp --------------------------------------------


def mix64(z: int) -> int:
    return int(lib().or_mix64(z))


def uniform(seed: int, start: int, index: int) -> float:
    return float(lib().or_uniform(seed, start, index))


def roll(a: np.ndarray, to_fft_order: bool) -> np.ndarray:
    a = np.ascontiguousarray(a)
    out = np.empty_like(a)
    nd, sh = _shape(a.shape)
    lib().or_roll(_ptr(a), _ptr(out), a.itemsize, nd, sh, int(to_fft_order))
    return out


def to_fft_order(a):
    return roll(a, True)


def to_centered(a):
    return roll(a, False)


def half_support(support: np.ndarray, axis: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(support, dtype=np.uint8)
    out = np.empty_like(s)
    nd, sh = _shape(s.shape)
    lib().or_half_support(_ptr(s), _ptr(out), nd, sh, axis)
    return out


# -- FFT and element-wise kernels (FFT order) --------------------------------------------


def fft(x: np.ndarray, forward: bool = True, prec: str = "f") -> np.ndarray:
    """Unitary n-D transform over every axis."""
    y = np.array(x, dtype=_cplx(prec), order="C", copy=True)
    nd, sh = _shape(y.shape)
    _fn("or_fft", prec)(_ptr(y), nd, sh, int(forward))
    return y


def usable_energy(amplitude, mask, prec="f") -> float:
    a = np.ascontiguousarray(amplitude, dtype=_real(prec))
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    return float(_fn("or_usable_energy", prec)(_ptr(a), _ptr(m), a.size))


def modulus(wave, amplitude, mask, energy=None, prec="f"):
    """Returns (projected wave, error) like src/cpu/kernels.cpp:25-42."""
    w = np.array(wave, dtype=_cplx(prec), order="C", copy=True)
    a = np.ascontiguousarray(amplitude, dtype=_real(prec))
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if energy is None:
        energy = usable_energy(a, m, prec)
    err = _fn("or_modulus", prec)(_ptr(w), _ptr(a), _ptr(m), w.size, energy)
    return w, float(err)


def modulus_num(wave, amplitude, mask, prec="f"):
    w = np.array(wave, dtype=_cplx(prec), order="C", copy=True)
    a = np.ascontiguousarray(amplitude, dtype=_real(prec))
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    num = _fn("or_modulus_num", prec)(_ptr(w), _ptr(a), _ptr(m), w.size)
    return w, float(num)


def update(x, projected, support, a, b, prec="f"):
    y = np.array(x, dtype=_cplx(prec), order="C", copy=True)
    p = np.ascontiguousarray(projected, dtype=_cplx(prec))
    s = np.ascontiguousarray(support, dtype=np.uint8)
    _fn("or_update", prec)(_ptr(y), _ptr(p), _ptr(s), y.size, a, b)
    return y


def blur(image, sigma, prec="f"):
    im = np.ascontiguousarray(image, dtype=_real(prec))
    out = np.empty_like(im)
    nd, sh = _shape(im.shape)
    _fn("or_blur", prec)(_ptr(im), _ptr(out), nd, sh, sigma)
    return out


def shrinkwrap(x, support, threshold=0.3, sigma=1.0, method="max", prec="f"):
    xx = np.ascontiguousarray(x, dtype=_cplx(prec))
    s = np.array(support, dtype=np.uint8, order="C", copy=True)
    nd, sh = _shape(xx.shape)
    _fn("or_shrinkwrap", prec)(_ptr(xx), nd, sh, threshold, sigma, _METHODS[method], _ptr(s))
    return s


def random_start(amplitude, support, seed, start, prec="f"):
    a = np.ascontiguousarray(amplitude, dtype=_real(prec))
    s = np.ascontiguousarray(support, dtype=np.uint8)
    x = np.empty(a.shape, dtype=_cplx(prec))
    nd, sh = _shape(a.shape)
    _fn("or_random_start", prec)(_ptr(a), _ptr(s), nd, sh, seed, start, _ptr(x))
    return x


def autocorrelation_support(intensity, threshold=0.1, prec="f"):
    i = np.ascontiguousarray(intensity, dtype=np.float32)
    out = np.empty(i.shape, np.uint8)
    nd, sh = _shape(i.shape)
    _fn("or_autocorrelation_support", prec)(_ptr(i), nd, sh, threshold, _ptr(out))
    return out


def simulate_intensity(obj, prec="f"):
    o = np.ascontiguousarray(obj, dtype=_cplx(prec))
    out = np.empty(o.shape, _real(prec))
    nd, sh = _shape(o.shape)
    _fn("or_simulate_intensity", prec)(_ptr(o), nd, sh, _ptr(out))
    return out


def aligned_error(a, reference, prec="f") -> float:
    x = np.ascontiguousarray(a, dtype=_cplx(prec))
    r = np.ascontiguousarray(reference, dtype=_cplx(prec))
    nd, sh = _shape(x.shape)
    return float(_fn("or_aligned_error", prec)(_ptr(x), _ptr(r), nd, sh))


def make_problem(intensity, support=None, mask=None, support_threshold=0.1):
    """Amplitude, mask and support like src/core/problem.cpp:40-70 (centre-origin)."""
    intensity = np.ascontiguousarray(intensity, dtype=np.float32)
    amplitude = np.sqrt(np.maximum(intensity, np.float32(0))).astype(np.float32)
    mask = (
        np.zeros(intensity.shape, np.uint8)
        if mask is None
        else (np.asarray(mask) != 0).astype(np.uint8)
    )
    if support is None:
        support = autocorrelation_support(intensity, support_threshold)
    support = (np.asarray(support) != 0).astype(np.uint8)
    return amplitude, mask, support


# -- recipes -------------------------------------------------------------------------------


def op_array(ops):
    """Recipe 7-tuples (kind, n, beta, threshold, sigma, method, axis) -> or_op[]."""
    arr = (_Op * max(1, len(ops)))()
    for k, (kind, n, beta, thr, sigma, method, axis) in enumerate(ops):
        arr[k] = _Op(_KINDS[kind], n, beta, thr, sigma, _METHODS[method], axis)
    return arr


class Solver:
    """The reference's CpuSolver (src/cpu/cpu_solver.cpp:18-110) in float or double."""

    def __init__(self, amplitude, mask, support, starts=1, seed=0, start=None, prec="f"):
        self.prec = prec
        self.shape = tuple(amplitude.shape)
        self.n = int(np.prod(self.shape))
        self.starts = starts
        self._amp = np.ascontiguousarray(amplitude, dtype=np.float32)
        self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
        self._sup = np.ascontiguousarray(support, dtype=np.uint8)
        self._start = None if start is None else np.ascontiguousarray(start, dtype=np.complex64)
        nd, sh = _shape(self.shape)
        self._h = _fn("or_solver_create", prec)(
            nd,
            sh,
            _ptr(self._amp),
            _ptr(self._mask),
            _ptr(self._sup),
            starts,
            seed,
            None if self._start is None else _ptr(self._start),
        )

    def run(self, ops):
        arr = op_array(ops)
        if _fn("or_solver_run", self.prec)(self._h, arr, len(ops)) != 0:
            raise ValueError("invalid recipe")

    @property
    def iterations(self) -> int:
        return int(_fn("or_solver_iterations", self.prec)(self._h))

    def results(self):
        """[(object centre-origin, support bool, errors float64)] per start."""
        it = self.iterations
        obj = np.empty((self.starts, *self.shape), _cplx(self.prec))
        sup = np.empty((self.starts, *self.shape), np.uint8)
        err = np.empty((self.starts, it), np.float64)
        _fn("or_solver_results", self.prec)(self._h, _ptr(obj), _ptr(sup), _ptr(err))
        return [(obj[k], sup[k].astype(bool), err[k]) for k in range(self.starts)]

    def get_state(self):
        x = np.empty((self.starts, *self.shape), _cplx(self.prec))
        s = np.empty((self.starts, *self.shape), np.uint8)
        _fn("or_solver_get_state", self.prec)(self._h, _ptr(x), _ptr(s))
        return x, s

    def set_state(self, x=None, supports=None):
        xx = None if x is None else np.ascontiguousarray(x, dtype=_cplx(self.prec))
        ss = None if supports is None else np.ascontiguousarray(supports, dtype=np.uint8)
        _fn("or_solver_set_state", self.prec)(
            self._h, None if xx is None else _ptr(xx), None if ss is None else _ptr(ss)
        )

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _fn("or_solver_destroy", self.prec)(h)
            self._h = None


def reconstruct(intensity, ops, support=None, mask=None, starts=1, seed=0, start=None,
                support_threshold=0.1, prec="f"):
    amplitude, m, s = make_problem(intensity, support, mask, support_threshold)
    solver = Solver(amplitude, m, s, starts=starts, seed=seed, start=start, prec=prec)
    solver.run(ops)
    return solver.results()
