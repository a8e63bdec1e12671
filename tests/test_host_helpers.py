"""CPU checks of the device helpers that compile for the host: the sweep list's row-tile order,
the 4-bit compensation code storage and the code's rounding against the oracle's independent
restatement (tests/host_helpers_main.cpp, built with hipcc as host code)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_helpers(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    lib = oracle.build()
    exe = tmp_path / "host_helpers"
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "host_helpers_main.cpp"), "-L", os.path.dirname(lib),
                    "-loracle", f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=600)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "host_helpers_main: ok" in res.stdout
