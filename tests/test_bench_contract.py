"""The benchmark script's contract on the CPU: one GPU only, and the committed traffic figures
consistent with the byte counts the script reports against."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "--gpus 1" in r.stderr
    assert r.stdout.strip() == ""


def test_fused_volume_bytes_counts():
    # per element at ngl = 8: 19 doubles per node, 12 ngl face-flux doubles read, 12 ngl trace
    # doubles written, and averaged over 5 stages the wave speed of the last (one double)
    assert bench.fused_volume_bytes(1, 8, 1, True, stages=5) == round(8 * (19 * 64 + 24 * 8 + 0.2))
    # the sensor adds b per node and the viscosity and its plot copy per element
    assert bench.fused_volume_bytes(1, 8, 1, True, sensor=True, stages=5) == round(8 * (20 * 64 + 24 * 8 + 2.2))
    # without side traces the update evaluates nothing of the next state
    assert bench.fused_volume_bytes(1, 8, 1, False, sensor=True) == 8 * (19 * 64 + 12 * 8)
    # the limiter inside the update reads the element size
    assert bench.fused_volume_bytes(1, 8, 1, True, sensor=True, stages=5, limiter=True) == \
        round(8 * (20 * 64 + 24 * 8 + 3.2))


def test_committed_traffic_matches_the_reported_workload():
    tr = bench.load_traffic("fused_volume")
    assert tr is not None
    alg = bench.fused_volume_bytes(262144, 8, 1, True, sensor=True, stages=5)
    assert tr["algorithmic_bytes_per_launch"] == alg
    cal = tr["calibration"]
    assert (cal["fetch_factor"], cal["write_factor"]) == (2.0, 1.0)
    read = tr["fetch_size_kib_per_dispatch"] * 1024 * cal["fetch_factor"]
    write = tr["write_size_kib_per_dispatch"] * 1024 * cal["write_factor"]
    assert tr["hbm_bytes_per_launch"] == read + write
    # HBM traffic near the compulsory bytes: a few percent below them is possible, since part of
    # the face fluxes the face launch has just written is still in the L2 caches (8 x 4 MB) when
    # the update reads them; well above them would mean re-reads
    assert 0.95 <= tr["hbm_bytes_per_launch"] / alg < 1.5


def test_committed_flops_definition():
    fl = json.load(open(os.path.join(ROOT, "profiles", "fused_volume_flops.json")))
    c = fl["wave_instructions_per_dispatch"]
    ex = 64 * (c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_MUL_F64"] + c["SQ_INSTS_VALU_TRANS_F64"]) \
        + 128 * c["SQ_INSTS_VALU_FMA_F64"]
    assert fl["executed_fp64_flops_per_launch"] == pytest.approx(ex, rel=1e-12)
    assert fl["algorithmic_fp64_flops_per_launch"] == bench.volume_flops(262144, 8)
