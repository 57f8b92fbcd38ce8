"""ctypes binding of libqpch.so, the C ABI of the MI355X engine (include/qpch.h).

The library is built in-tree by ``qoraphasecu_amd.build``.  Loading fails loudly when it
is missing: there is no CPU fallback behind these functions.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqpch.so")

QPCH_OK, QPCH_EINVAL, QPCH_EDEVICE = 0, 1, 2
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
SCHEDULE = {"unfused": 0, "fused": 1, "graph": 2}


class DeviceError(RuntimeError):
    """A device failure or a missing device (std::runtime_error in the C++ API)."""


class QpchProblem(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_size_t * 3),
        ("amplitude", ctypes.c_void_p),
        ("mask", ctypes.c_void_p),
        ("support", ctypes.c_void_p),
        ("starts", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("start", ctypes.c_void_p),
        ("storage", ctypes.c_int),
        ("mixed", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("problems", ctypes.c_int),
        ("plane", ctypes.c_int),
        ("xcd", ctypes.c_int),
        ("lanes", ctypes.c_int),
        ("radix", ctypes.c_int),
        ("persist", ctypes.c_int),
    ]


class QpchTuning(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int),
        ("plane", ctypes.c_int),
        ("xcd", ctypes.c_int),
        ("us_per_iteration", ctypes.c_double),
        ("lanes", ctypes.c_int),
        ("radix", ctypes.c_int),
        ("persist", ctypes.c_int),
    ]


class QpchOp(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("a", ctypes.c_float),
        ("b", ctypes.c_float),
        ("refine", ctypes.c_int),
        ("threshold", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("method", ctypes.c_int),
        ("axis", ctypes.c_int),
    ]


OP_ITERATE, OP_SHRINKWRAP, OP_DETWIN_BEGIN, OP_DETWIN_END = 0, 1, 2, 3


def engine_ops(records):
    """The engine's operation array for a recipe of 7-field records (kind, n, beta,
    threshold, sigma, method, axis); update coefficients as src/core/pointwise.hpp:41-45,
    in float32."""
    f = np.float32
    out = (QpchOp * len(records))()
    for k, (kind, n, beta, threshold, sigma, method, axis) in enumerate(records):
        o = out[k]
        if kind in ("er", "hio", "raar"):
            o.kind, o.iterations, o.refine = OP_ITERATE, n, int(kind == "er")
            o.a, o.b = {"er": (0.0, 0.0), "hio": (1.0, float(-f(beta))),
                        "raar": (float(f(beta)), float(f(1.0) - f(2.0) * f(beta)))}[kind]
        elif kind == "shrinkwrap":
            o.kind, o.threshold, o.sigma, o.method = OP_SHRINKWRAP, threshold, sigma, 0 if method == "max" else 1
        elif kind == "detwin_begin":
            o.kind, o.axis = OP_DETWIN_BEGIN, axis
        elif kind == "detwin_end":
            o.kind = OP_DETWIN_END
        else:
            raise ValueError(f"unknown recipe operation {kind!r}")
    return out


class QpchDeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int),
        ("memory_bytes", ctypes.c_size_t),
        ("l2_bytes", ctypes.c_size_t),
        ("peak_bandwidth_gbs", ctypes.c_double),
    ]


_lib = None


def lib():
    """The loaded libqpch.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DeviceError(
                f"{LIB_PATH} is missing: build it with qoraphasecu_amd.build.build() "
                "(there is no CPU fallback)"
            )
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.qpch_last_error.restype = ctypes.c_char_p
        L.qpch_device_count.argtypes = [ctypes.POINTER(i)]
        L.qpch_device_info.argtypes = [ctypes.POINTER(QpchDeviceInfo)]
        L.qpch_solver_create.argtypes = [ctypes.POINTER(QpchProblem), ctypes.POINTER(vp)]
        L.qpch_solver_destroy.argtypes = [vp]
        L.qpch_solver_destroy.restype = None
        L.qpch_solver_begin_run.argtypes = [vp, i]
        L.qpch_solver_iterate.argtypes = [vp, i, f, f, i]
        L.qpch_solver_shrinkwrap.argtypes = [vp, f, f, i]
        L.qpch_solver_get_support.argtypes = [vp, vp]
        L.qpch_solver_set_support.argtypes = [vp, vp]
        L.qpch_solver_save_support.argtypes = [vp]
        L.qpch_solver_restore_support.argtypes = [vp]
        L.qpch_solver_half_support.argtypes = [vp, ctypes.c_int]
        L.qpch_solver_read_errors.argtypes = [vp, i, i, vp]
        L.qpch_solver_read_state.argtypes = [vp, vp]
        L.qpch_solver_energy.argtypes = [vp]
        L.qpch_solver_energy.restype = d
        L.qpch_solver_synchronize.argtypes = [vp]
        L.qpch_solver_energies.argtypes = [vp, vp]
        L.qpch_autotune.argtypes = [ctypes.POINTER(QpchProblem), ctypes.POINTER(QpchTuning)]
        L.qpch_autotune_log.restype = ctypes.c_char_p
        L.qpch_group_create.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(vp)]
        L.qpch_group_iterate.argtypes = [vp, i, f, f, i]
        L.qpch_group_synchronize.argtypes = [vp]
        L.qpch_group_run.argtypes = [vp, ctypes.POINTER(QpchOp), i]
        L.qpch_group_destroy.argtypes = [vp]
        L.qpch_group_destroy.restype = None
        L.qpch_solver_time_passes.argtypes = [vp, i, i, ctypes.POINTER(f), i, ctypes.POINTER(i)]
        L.qpch_solver_time_iterations.argtypes = [vp, i, i, i, ctypes.POINTER(f)]
        L.qpch_solver_pass_info.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i), i, ctypes.POINTER(i)]
        L.qpch_fft.argtypes = [vp, i, szp, i]
        L.qpch_modulus_project.argtypes = [vp, vp, vp, i, szp, ctypes.POINTER(d)]
        L.qpch_support_project.argtypes = [vp, vp, sz]
        L.qpch_gaussian_blur.argtypes = [vp, vp, i, szp, f]
        L.qpch_amplitude_loss.argtypes = [vp, vp, ctypes.c_longlong, vp, vp, i, szp, i, i, vp]
        L.qpch_amplitude_loss_noalloc.argtypes = [vp, vp, ctypes.c_longlong, vp, vp, i, szp, i, i, vp]
        L.qpch_autocorrelation_support.argtypes = [vp, i, i, szp, f, vp]
        L.qpch_simulate_intensity.argtypes = [vp, i, i, szp, vp]
        _lib = L
    return _lib


def check(status: int) -> None:
    if status == QPCH_OK:
        return
    msg = lib().qpch_last_error().decode(errors="replace")
    if status == QPCH_EINVAL:
        raise ValueError(msg)
    raise DeviceError(msg)


def _ptr(a):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _shape(shape):
    shape = tuple(int(s) for s in shape)
    return len(shape), (ctypes.c_size_t * len(shape))(*shape)


def device_count() -> int:
    c = ctypes.c_int(0)
    check(lib().qpch_device_count(ctypes.byref(c)))
    return c.value


def device_info() -> dict:
    info = QpchDeviceInfo()
    check(lib().qpch_device_info(ctypes.byref(info)))
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "compute_units": info.compute_units,
        "memory_bytes": info.memory_bytes,
        "l2_bytes": info.l2_bytes,
        "peak_bandwidth_gbs": info.peak_bandwidth_gbs,
    }


# --- operators on host arrays ------------------------------------------------------------


def fft(x: np.ndarray, forward: bool = True) -> np.ndarray:
    """Unitary n-D transform (FFT order, complex64) on the device."""
    y = np.array(x, dtype=np.complex64, order="C", copy=True)
    nd, sh = _shape(y.shape)
    check(lib().qpch_fft(_ptr(y), nd, sh, int(forward)))
    return y


def modulus_project(x, amplitude, mask):
    """(P_M x, numerator) on FFT-order arrays."""
    y = np.array(x, dtype=np.complex64, order="C", copy=True)
    a = np.ascontiguousarray(amplitude, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nd, sh = _shape(y.shape)
    num = ctypes.c_double(0.0)
    check(lib().qpch_modulus_project(_ptr(y), _ptr(a), _ptr(m), nd, sh, ctypes.byref(num)))
    return y, num.value


def support_project(x, support):
    y = np.array(x, dtype=np.complex64, order="C", copy=True)
    s = np.ascontiguousarray(support, dtype=np.uint8)
    check(lib().qpch_support_project(_ptr(y), _ptr(s), y.size))
    return y


def gaussian_blur(image, sigma):
    im = np.ascontiguousarray(image, dtype=np.float32)
    out = np.empty_like(im)
    nd, sh = _shape(im.shape)
    check(lib().qpch_gaussian_blur(_ptr(im), _ptr(out), nd, sh, float(sigma)))
    return out


def autocorrelation_support(intensity, threshold=0.1, count=None):
    """Autocorrelation supports on the device of centre-origin intensities: one array, or
    `count` of one shape stacked along a leading axis (one device pass for all)."""
    i = np.ascontiguousarray(intensity, dtype=np.float32)
    shape = i.shape if count is None else i.shape[1:]
    out = np.empty(i.shape, np.uint8)
    nd, sh = _shape(shape)
    check(lib().qpch_autocorrelation_support(_ptr(i), 1 if count is None else count, nd, sh, float(threshold),
                                             _ptr(out)))
    return out


def simulate_intensity(obj, count=None):
    """|centred unitary F(obj)|^2 on the device, of one object or `count` stacked ones."""
    o = np.ascontiguousarray(obj, dtype=np.complex64)
    shape = o.shape if count is None else o.shape[1:]
    out = np.empty(o.shape, np.float32)
    nd, sh = _shape(shape)
    check(lib().qpch_simulate_intensity(_ptr(o), 1 if count is None else count, nd, sh, _ptr(out)))
    return out


def amplitude_loss(x_ptr, amp_ptr, amp_stride, grad_ptr, loss_ptr, shape, batch, precision, stream):
    """Device-pointer entry: see qpch_amplitude_loss in include/qpch.h."""
    nd, sh = _shape(shape)
    check(lib().qpch_amplitude_loss(ctypes.c_void_p(x_ptr), ctypes.c_void_p(amp_ptr), amp_stride,
                                    ctypes.c_void_p(grad_ptr), ctypes.c_void_p(loss_ptr), nd, sh, batch,
                                    precision, ctypes.c_void_p(stream)))


def amplitude_loss_noalloc(x_ptr, amp_ptr, amp_stride, grad_ptr, loss_ptr, shape, batch, precision, stream):
    """Device-pointer entry that allocates nothing: see qpch_amplitude_loss_noalloc in
    include/qpch.h."""
    nd, sh = _shape(shape)
    check(lib().qpch_amplitude_loss_noalloc(ctypes.c_void_p(x_ptr), ctypes.c_void_p(amp_ptr), amp_stride,
                                            ctypes.c_void_p(grad_ptr), ctypes.c_void_p(loss_ptr), nd, sh, batch,
                                            precision, ctypes.c_void_p(stream)))


# --- solver ------------------------------------------------------------------------------


class Solver:
    """A device solver over FFT-order inputs (what the C++ HipSolver drives)."""

    def __init__(self, amplitude, mask, support, starts=1, batch=0, seed=0, start=None,
                 storage="f32", mixed=False, schedule="graph", problems=1, plane=0, xcd=0, lanes=1, radix=0, tuned=False,
                 persist=0):
        """amplitude / mask / support: one problem's FFT-order arrays, or `problems` of them
        stacked along a leading axis; `starts` per problem.  tuned: take the launch choices
        the autotuner makes for this problem (qpch_autotune) instead of the arguments."""
        shape = tuple(amplitude.shape)
        self.problems = problems
        self.shape = shape[1:] if problems > 1 else shape
        self.n = int(np.prod(self.shape))
        self.starts = starts * problems
        self.tuning = None
        self._keep = [
            np.ascontiguousarray(amplitude, dtype=np.float32),
            np.ascontiguousarray(mask, dtype=np.uint8),
            np.ascontiguousarray(support, dtype=np.uint8),
            None if start is None else np.ascontiguousarray(start, dtype=np.complex64),
        ]
        p = QpchProblem()
        p.ndim = len(self.shape)
        for k, s in enumerate(self.shape):
            p.shape[k] = s
        p.problems = problems
        p.plane = plane
        p.xcd = xcd
        p.lanes = lanes
        p.radix = radix
        p.persist = persist
        p.amplitude = _ptr(self._keep[0])
        p.mask = _ptr(self._keep[1])
        p.support = _ptr(self._keep[2])
        p.starts = starts
        p.batch = batch
        p.seed = seed
        p.start = _ptr(self._keep[3])
        p.storage = STORAGE[storage]
        p.mixed = int(mixed)
        p.schedule = SCHEDULE[schedule]
        if tuned and os.environ.get("QPCH_TUNE") != "0":  # QPCH_TUNE=0: the defaults, as in the C++ host
            t = QpchTuning()
            check(lib().qpch_autotune(ctypes.byref(p), ctypes.byref(t)))
            p.batch, p.plane, p.xcd, p.lanes, p.radix, p.persist = t.batch, t.plane, t.xcd, t.lanes, t.radix, t.persist
            self.tuning = dict(batch=t.batch, plane=t.plane, xcd=t.xcd, lanes=t.lanes, radix=t.radix,
                               persist=t.persist, us_per_iteration=t.us_per_iteration)
        h = ctypes.c_void_p()
        check(lib().qpch_solver_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.energy = lib().qpch_solver_energy(self._h)
        e = np.empty(problems, np.float64)
        check(lib().qpch_solver_energies(self._h, _ptr(e)))
        self.energies = e

    def begin_run(self, capacity):
        check(lib().qpch_solver_begin_run(self._h, capacity))

    def iterate(self, n, a, b, refine=False):
        check(lib().qpch_solver_iterate(self._h, n, a, b, int(refine)))

    def shrinkwrap(self, threshold=0.3, sigma=1.0, method="max"):
        check(lib().qpch_solver_shrinkwrap(self._h, threshold, sigma, 0 if method == "max" else 1))

    def get_support(self):
        out = np.empty((self.starts, *self.shape), np.uint8)
        check(lib().qpch_solver_get_support(self._h, _ptr(out)))
        return out

    def set_support(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        check(lib().qpch_solver_set_support(self._h, _ptr(s)))

    def save_support(self):
        check(lib().qpch_solver_save_support(self._h))

    def restore_support(self):
        check(lib().qpch_solver_restore_support(self._h))

    def half_support(self, axis):
        """Detwin on the device: each start's support cut at its centre of mass along axis."""
        check(lib().qpch_solver_half_support(self._h, axis))

    def read_errors(self, first, count):
        out = np.empty((count, self.starts), np.float64)
        check(lib().qpch_solver_read_errors(self._h, first, count, _ptr(out)))
        return out

    def read_state(self):
        out = np.empty((self.starts, *self.shape), np.complex64)
        check(lib().qpch_solver_read_state(self._h, _ptr(out)))
        return out

    def time_passes(self, low=False, reps=20):
        """[(axis, mode, ms)] for each pass of one iteration."""
        ms = (ctypes.c_float * 8)()
        axes = (ctypes.c_int * 8)()
        modes = (ctypes.c_int * 8)()
        cnt = ctypes.c_int(0)
        check(lib().qpch_solver_pass_info(self._h, int(low), axes, modes, 8, ctypes.byref(cnt)))
        check(lib().qpch_solver_time_passes(self._h, int(low), reps, ms, 8, ctypes.byref(cnt)))
        names = {0: "fwd", 1: "inv", 2: "modulus", 3: "update"}
        return [(axes[k], names[modes[k]], float(ms[k])) for k in range(cnt.value)]

    def time_iterations(self, low=False, length=10, reps=20):
        """Microseconds per iteration of all starts, graph-replayed chunks of `length`."""
        us = ctypes.c_float(0.0)
        check(lib().qpch_solver_time_iterations(self._h, int(low), length, reps, ctypes.byref(us)))
        return us.value

    def synchronize(self):
        check(lib().qpch_solver_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().qpch_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Group:
    """Solvers advanced together, one graph launch per chunk (qpch_group_*)."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        check(lib().qpch_group_create(arr, len(self.solvers), ctypes.byref(h)))
        self._h = h

    def iterate(self, n, a, b, refine=False):
        check(lib().qpch_group_iterate(self._h, n, a, b, int(refine)))

    def run(self, records):
        """A whole recipe of 7-field records for every solver in one graph launch
        (qpch_group_run); each solver needs begin_run with room for its iterations."""
        ops = engine_ops(records)
        check(lib().qpch_group_run(self._h, ops, len(records)))

    def synchronize(self):
        check(lib().qpch_group_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().qpch_group_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def autotune_log() -> str:
    return lib().qpch_autotune_log().decode()
