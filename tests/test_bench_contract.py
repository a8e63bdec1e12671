"""bench.py's host-side contract, without a GPU: the multi-GPU refusal, the algorithmic byte model
of one diffusion sub-step and the rocprof traffic figures it reads from profiles/."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "error" in json.loads(r.stdout.strip().splitlines()[-1])


def test_algorithmic_bytes_per_sub_step():
    # read P phase fractions + the stored concentration, write the concentration
    assert bench.alg_bytes_per_substep(False, 2) == 2 * 8 + 8 + 8
    assert bench.alg_bytes_per_substep(True, 2) == 2 * 4 + 5 + 5  # FP32 value + one-byte code
    assert bench.alg_bytes_per_substep(True, 3) == 3 * 4 + 5 + 5


def test_measured_traffic_is_read_from_profiles():
    t = bench.measured_traffic("fecmn_3d_512")
    assert set(t) >= {"FormF64", "FormMixed"}
    # at least the compulsory bytes the marching kernel moves per cell (16 / 10 B at 512^3)
    assert t["FormF64"] >= 16 * 512 ** 3 and t["FormMixed"] >= 10 * 512 ** 3
    assert t["FormF64"] > t["FormMixed"]
