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


def test_byte_models():
    # plane-marching kernel: the stored concentration read and written once per cell
    assert bench.march_bytes_per_cell(False) == 8 + 8
    assert bench.march_bytes_per_cell(True) == 4.5 + 4.5  # FP32 value + 4-bit code each way
    # a listed cell: its own and six neighbours' concentration + value_phase, one write
    assert bench.listed_bytes_per_cell(False, 2) == 7 * (8 + 16) + 8
    assert bench.listed_bytes_per_cell(True, 2) == 7 * (4.5 + 8) + 4.5
    # the scope document's model per step at 200 sub-steps, K = 4, P = 2, Mn once per step
    assert bench.model_8d_bytes_per_cell_step(False, 2, 2, 200) == 6560
    assert bench.model_8d_bytes_per_cell_step(True, 2, 2, 200) == 4904


def test_measured_traffic_names_its_sources():
    t = bench.measured_traffic("fecmn_3d_512")
    assert t, "profiles/ holds no traffic line for the default workload"
    assert t["stale"] == (t.get("sources") != bench.kernel_source_hash())
    if not t["stale"]:
        # bytes per launch by kernel name, for each profiled tile shape of both forms: at least
        # nine tenths of the compulsory bytes the marching kernel moves per cell (16 / 9 B at 512^3;
        # the x ends of a lane's run come from registers, the edge rows once per tile)
        for rows in (8, 16):
            for mixed, bpc in ((False, 16), (True, 9)):
                name = bench.march_kernel_name(mixed, 2, False, rows)
                assert name in t["march"], name
                assert t["march"][name] >= 0.9 * bpc * 512 ** 3


def test_pmc_summary_reads_counter_files(tmp_path):
    d = tmp_path / "host" / "123"
    d.mkdir(parents=True)
    (d / "run_counter_collection.csv").write_text(
        "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
        "1,k_a,FETCH_SIZE,10\n2,k_a,FETCH_SIZE,30\n3,k_b,FETCH_SIZE,5\n")
    out = bench.pmc_summary(str(tmp_path))
    assert out["k_a"]["FETCH_SIZE"] == {"mean": 20.0, "dispatches": 2}
    assert out["k_b"]["FETCH_SIZE"]["mean"] == 5.0
