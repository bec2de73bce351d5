"""Host-side pieces of bench.py: the workload table, the two-layer case (the
reference's CPU-runnable 96 x 128 configuration) and the --gpus refusal."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_workloads_are_complete():
    keys = {"name", "shape", "dh", "dt", "nt", "order", "abcn", "nshots", "f0", "seed", "per_launch"}
    for name, W in bench.WORKLOADS.items():
        assert keys <= set(W), name
        assert W["per_launch"] <= W["nshots"], name


def test_two_layer_case_matches_its_description():
    W = bench.WORKLOADS["acoustic2d_small"]
    assert W["shape"] == (96, 128) and W["dh"] == 10.0 and W["dt"] == 2e-3 and W["nt"] == 800
    vt, v0, src, rec = bench.two_layer_case(W)
    assert vt.shape == (96, 128) and vt.dtype == np.float32
    assert np.all(vt[:48] == 1500.0) and np.all(vt[48:] == 2500.0)
    assert np.all(v0 == 1500.0)
    assert src.tolist() == [[32, 48]] and rec.tolist() == [[[96, 48]]]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
