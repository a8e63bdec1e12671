"""The host model (phasefield_amd/csrc/host_model.cpp) and the oracle (oracle/pf_oracle.cpp) built
with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and run through
tests/sanitize_main.cpp: parsing, parameter updates, every writer and reader, the input error
paths, and the oracle's forms and boundary conditions.  Any finding aborts the driver."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_model_and_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_main"
    srcs = [os.path.join(ROOT, "tests", "sanitize_main.cpp"),
            os.path.join(ROOT, "phasefield_amd", "csrc", "host_model.cpp"),
            os.path.join(ROOT, "oracle", "pf_oracle.cpp")]
    flags = ["-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"),
             "-I", os.path.join(ROOT, "oracle")]
    subprocess.run(["g++", *flags, "-o", str(exe), *srcs], check=True)
    shutil.copy(os.path.join(GOLDEN, "ref_example_log_file.txt"), tmp_path / "log_file.txt")
    shutil.copy(os.path.join(GOLDEN, "ref_example_thermo_data.txt"), tmp_path / "thermo_data.txt")
    os.chmod(tmp_path / "log_file.txt", 0o644)
    os.chmod(tmp_path / "thermo_data.txt", 0o644)
    scratch = tmp_path / "out"
    scratch.mkdir()
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), str(tmp_path / "log_file.txt"), str(scratch)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "sanitize_main: ok" in r.stdout
