"""The measurement script's CPU-side contract: the algorithmic byte counts its
roofline divides by (the figures derived for the zipf-1m sizes: SpMV A
268,000,004 B, A' 272,000,004 B, phase 1 368,000,004 B, phase 2
300,000,004 B per launch; FP32 values halve the value stream), and the
refusal of more than one GPU before anything touches a device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_algorithmic_bytes_of_the_zipf1m_sizes():
    m, n, nnz = 1_000_000, 2_000_000, 20_000_000
    assert bench.spmv_bytes(m, n, nnz) == 268_000_004
    assert bench.spmv_bytes(n, m, nnz) == 272_000_004
    p1, p2 = bench.phase_bytes(m, n, nnz)
    assert (p1, p2) == (368_000_004, 300_000_004)
    # an iteration moves 668,000,008 B: 83.5 us at 8 TB/s
    assert abs((p1 + p2) / (bench.HBM_PEAK_GBPS * 1e9) - 83.5e-6) < 1e-9


def test_fp32_values_halve_the_value_stream():
    m, n, nnz = 4_000_000, 8_000_000, 120_000_000
    # zipf-4m with FP32 values and gathers: nnz * (4 + 4) + 4 (m + 1) + 4 n + 8 m
    assert bench.spmv_bytes(m, n, nnz, sv=4, sx=4) == nnz * 8 + 4 * (m + 1) + 4 * n + 8 * m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
