"""The benchmark inputs are reproducible: regenerating every workload gives the bytes the
committed manifest hashes (intensity, mask, oracle autocorrelation support)."""

import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_manifest  # noqa: E402


def _committed():
    with open(os.path.join(HERE, "golden", "workloads_manifest.json")) as fh:
        return json.load(fh)


def test_config_shapes():
    from qoraphasecu_amd import workloads

    for name, shape, starts, iters in [("c1", (64, 64), 8, 720), ("c2", (256, 256), 32, 700),
                                       ("c4", (128, 128, 128), 8, 600)]:
        w = workloads.config(name)
        assert w["shape"] == shape and w["starts"] == starts
        assert w["intensity"].shape == shape and w["intensity"].dtype == np.float32
        assert w["recipe"].iterations() == iters
    w = workloads.config("c3")
    assert [p[1].shape[0] for p in w["problems"]] == [128] * 16 + [256] * 16 + [512] * 16


def test_inputs_match_the_manifest():
    want = _committed()
    if want["numpy"] != np.__version__ or want["cpu_features"] != make_manifest.cpu_features():
        pytest.skip("numpy version or CPU vector features differ from the manifest's host")
    got = make_manifest.manifest()
    assert got["workloads"] == want["workloads"]
