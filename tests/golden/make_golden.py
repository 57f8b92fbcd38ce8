#!/usr/bin/env python3
"""Writes the committed fixtures under tests/golden/ (run in the build container).

- fft_golden.npz: seeded complex inputs and their unitary n-D DFTs from numpy.fft, whose
  complex64/complex128 transforms are pocketfft C++ (the third-party FFT the reference's
  CPU backend calls at src/cpu/fft.cpp:34), on the odd and mixed-radix shapes the
  reference's tests use (17x24, 21x15, 16x20x12) and a few more.
- ref_golden.npz: outputs of the reference's own pointwise.hpp / layout.hpp / recipe.cpp,
  compiled from the reference tree into oracle/_ref/libref.so (`make -C oracle ref`):
  the splitmix hash and uniform draw, fftfreq, the update coefficients, the fftshift
  permutations, the default recipe's op kinds and the validate() verdicts.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

FFT_SHAPES = [(17, 24), (21, 15), (16, 20, 12), (15, 16, 17), (32, 32), (8, 8, 8), (64, 48), (2, 3)]


def fft_fixtures():
    rng = np.random.default_rng(20240601)
    out = {}
    for k, shape in enumerate(FFT_SHAPES):
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        out[f"x{k}"] = x
        out[f"fwd32_{k}"] = np.fft.fftn(x, norm="ortho").astype(np.complex64)
        out[f"inv32_{k}"] = np.fft.ifftn(x, norm="ortho").astype(np.complex64)
        out[f"fwd64_{k}"] = np.fft.fftn(x.astype(np.complex128), norm="ortho")
        out[f"inv64_{k}"] = np.fft.ifftn(x.astype(np.complex128), norm="ortho")
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def ref_fixtures():
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    L = ctypes.CDLL(lib_path)
    u64, f, i = ctypes.c_uint64, ctypes.c_float, ctypes.c_int
    L.ref_mix64.restype = u64
    L.ref_mix64.argtypes = [u64]
    L.ref_uniform.restype = f
    L.ref_uniform.argtypes = [u64, u64, u64]
    L.ref_fftfreq.restype = f
    L.ref_fftfreq.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.ref_coefficients.argtypes = [i, f, ctypes.POINTER(f), ctypes.POINTER(f)]
    L.ref_roll_u32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, i, ctypes.POINTER(ctypes.c_size_t), i]
    L.ref_default_recipe_kinds.argtypes = [ctypes.POINTER(i), i]
    out = {}
    zs = np.array([0, 1, 2, 3, 12345, 2**32 - 1, 2**63 + 5, 2**64 - 1], dtype=np.uint64)
    out["mix64_in"] = zs
    out["mix64_out"] = np.array([L.ref_mix64(int(z)) for z in zs], dtype=np.uint64)
    triples = np.array([(s, st, ix) for s in (0, 1, 7, 2**40 + 3) for st in (0, 1, 5, 31) for ix in (0, 1, 2, 1000, 65535, 2**21 + 7)],
                       dtype=np.uint64)
    out["uniform_in"] = triples
    out["uniform_out"] = np.array([L.ref_uniform(int(a), int(b), int(c)) for a, b, c in triples], dtype=np.float32)
    ff = np.array([(k, n) for n in (1, 2, 5, 8, 17, 256) for k in range(n)], dtype=np.uint32)
    out["fftfreq_in"] = ff
    out["fftfreq_out"] = np.array([L.ref_fftfreq(int(k), int(n)) for k, n in ff], dtype=np.float32)
    coeffs = []
    for kind in (0, 1, 2):
        for beta in (0.9, 0.5, 0.99, 0.7):
            a, b = f(), f()
            L.ref_coefficients(kind, beta, ctypes.byref(a), ctypes.byref(b))
            coeffs.append((kind, beta, a.value, b.value))
    out["coefficients"] = np.array(coeffs, dtype=np.float64)
    for name, shape in {"roll_547": (5, 4, 7), "roll_65": (6, 5), "roll_44": (4, 4), "roll_333": (3, 3, 3)}.items():
        n = int(np.prod(shape))
        src = np.arange(n, dtype=np.uint32)
        sh = (ctypes.c_size_t * len(shape))(*shape)
        for direction in (0, 1):
            dst = np.empty(n, np.uint32)
            L.ref_roll_u32(src.ctypes.data, dst.ctypes.data, len(shape), sh, direction)
            out[f"{name}_{direction}"] = dst.reshape(shape)
    kinds = (ctypes.c_int * 64)()
    count = L.ref_default_recipe_kinds(kinds, 64)
    out["default_recipe_kinds"] = np.array(kinds[:count], dtype=np.int32)
    out["default_recipe_iterations"] = np.array(L.ref_default_recipe_iterations(), dtype=np.int32)
    out["validate_rejects"] = np.array([L.ref_validate_rejects(c) for c in range(7)], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_golden.npz"), **out)


if __name__ == "__main__":
    fft_fixtures()
    ref_fixtures()
    print("written", sorted(os.listdir(HERE)))
