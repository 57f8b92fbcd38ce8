"""Results do not depend on process history: two processes whose first solvers of a shape
have different start counts (so the autotuner sees different first callers) return the same
bits for the same inputs.  This extends the reference's batched == looped check
(tests/test_cuda.cpp:160-180) across processes: no choice that changes rounding (FFT radix,
3D plane or line passes) is made by timing."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import hashlib, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/tests"]
import qoraphasecu_amd as qpc
import synthetic
first, variant, shape = int(sys.argv[2]), sys.argv[3], tuple(int(v) for v in sys.argv[4].split("x"))
obj, sup, inten = synthetic.triangle(shape, shape[-1] // 4)
sup = sup.astype(bool)
# the process's first solver of this shape: `first` starts (tuned on first use)
qpc.reconstruct(inten, support=sup, recipe=qpc.Recipe(qpc.RAAR(2)), starts=first, backend="hip", variant=variant)
runs = qpc.reconstruct(inten, support=sup, recipe=qpc.Recipe(qpc.RAAR(30), qpc.Shrinkwrap(), qpc.ER(10)), starts=4,
                       seed=5, backend="hip", variant=variant)
h = hashlib.sha256()
for r in runs:
    h.update(r.object.tobytes()); h.update(r.support.tobytes()); h.update(r.errors.tobytes())
print("DIGEST", h.hexdigest(), flush=True)
"""


def _digest(first, variant, shape):
    env = dict(os.environ, QPCH_TUNE="1")
    out = subprocess.run([sys.executable, "-c", SCRIPT, ROOT, str(first), variant, "x".join(map(str, shape))],
                         capture_output=True, text=True, timeout=240, env=env)
    assert out.returncode == 0, out.stderr[-2000:]
    return [l.split()[1] for l in out.stdout.splitlines() if l.startswith("DIGEST")][0]


@pytest.mark.parametrize("variant,shape", [("batched", (256, 256)), ("mixed", (256, 256)), ("graph", (128, 128)),
                                           ("mixed", (32, 64, 64))])
def test_same_bits_whatever_the_first_solver(variant, shape):
    a = _digest(1, variant, shape)
    b = _digest(16, variant, shape)
    assert a == b, (variant, shape)
