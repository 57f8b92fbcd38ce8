"""Test objects, restated from the reference's tests/synthetic.hpp:21-83.

``triangle`` is a right triangle (not centrosymmetric) with an amplitude dip and a
phase bump; ``corner`` is the 3D triangle with one corner cut off; ``random_field``
is the reference's LCG field.  Intensities come from the oracle's
``simulate_intensity`` in float, like ``qpcu::simulate_intensity``.
"""

from __future__ import annotations

import numpy as np

import oracle


def _dims3(shape):
    return (1, *shape) if len(shape) == 2 else tuple(shape)


def triangle(shape, m):
    """tests/synthetic.hpp:21-50 -> (object c64, support u8, intensity f32)."""
    shape = tuple(int(s) for s in shape)
    d = _dims3(shape)
    is3d = len(shape) == 3
    obj = np.zeros(d, np.complex64)
    sup = np.zeros(d, np.uint8)
    c = (m - 1) / 2.0
    md = float(m)
    z, y, x = np.meshgrid(np.arange(d[0]), np.arange(d[1]), np.arange(d[2]), indexing="ij")
    lz = z.astype(np.float64) - float(d[0] // 2 - m // 2)
    ly = y.astype(np.float64) - float(d[1] // 2 - m // 2)
    lx = x.astype(np.float64) - float(d[2] // 2 - m // 2)
    in_box = (ly >= 0) & (ly < md) & (lx >= 0) & (lx < md)
    if is3d:
        in_box &= (lz >= 0) & (lz < md)
    inside = in_box & (ly + lx < md)
    ry = ly - c
    rx = lx - c
    amp = 1.0 - 0.5 * np.exp(-(rx * rx + ry * ry) / 20.0)
    phase = 1.2 * np.exp(-((rx + 4) ** 2 + (ry + 4) ** 2) / 30.0)
    if is3d:
        phase = phase + 0.3 * (lz - c) / md
    # std::polar(float amp, float phase) = (amp cos, amp sin) in float
    a32 = amp.astype(np.float32)
    p32 = phase.astype(np.float32)
    val = (a32 * np.cos(p32)) + 1j * (a32 * np.sin(p32))
    obj[inside] = val[inside].astype(np.complex64)
    sup[inside] = 1
    obj = obj.reshape(shape)
    sup = sup.reshape(shape)
    intensity = oracle.simulate_intensity(obj).astype(np.float32)
    return obj, sup, intensity


def corner(n, m):
    """tests/synthetic.hpp:56-72."""
    obj, sup, _ = triangle((n, n, n), m)
    lo = float(n // 2 - m // 2)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cut = (z + y + x).astype(np.float64) - 3 * lo >= 1.5 * m
    obj = obj.copy()
    sup = sup.copy()
    obj[cut] = 0
    sup[cut] = 0
    return obj, sup, oracle.simulate_intensity(obj).astype(np.float32)


def random_field(n, seed):
    """tests/synthetic.hpp:74-83: LCG pairs in [-0.5, 0.5)."""
    out = np.empty(2 * n, np.float32)
    state = (seed * 2654435761 + 1) & 0xFFFFFFFF
    for k in range(2 * n):
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        out[k] = np.float32(state >> 8) / np.float32(16777216.0) - np.float32(0.5)
    return out.view(np.complex64)
