"""Write workloads_manifest.json: SHA-256 of every benchmark input this build generates
(float32 intensity, uint8 mask, and the uint8 autocorrelation support the oracle derives),
with the numpy version and CPU features they were made with.

    python tests/golden/make_manifest.py

The generators are seeded numpy (PCG64); numpy's transcendental functions may dispatch
to vector code by CPU feature, so a host with different features may produce different
last bits and therefore different hashes (the manifest names the features it saw)."""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def sha(a):
    return None if a is None else hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def entry(intensity, mask):
    import oracle

    _, _, sup = oracle.make_problem(intensity, None, mask)
    return {"intensity": sha(np.asarray(intensity, np.float32)),
            "mask": sha(None if mask is None else np.asarray(mask, np.uint8)),
            "support": sha(np.asarray(sup, np.uint8)), "shape": list(np.shape(intensity))}


def cpu_features():
    try:
        from numpy._core._multiarray_umath import __cpu_features__ as f
    except ImportError:  # pragma: no cover
        return []
    return sorted(k for k, v in f.items() if v and k.startswith("AVX"))


def manifest(names=("c1", "c2", "c4", "c3")):
    from qoraphasecu_amd import workloads

    out = {"numpy": np.__version__, "cpu_features": cpu_features(), "workloads": {}}
    for name in names:
        w = workloads.config(name)
        if "problems" in w:
            out["workloads"][name] = [entry(p[1], p[2]) for p in w["problems"]]
        else:
            out["workloads"][name] = entry(w["intensity"], w["mask"])
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "workloads_manifest.json"), "w") as fh:
        json.dump(manifest(), fh, indent=1)
        fh.write("\n")
