"""The serving-loop benchmark binary (gpu_compressstore2_amd/csx_serve)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "gpu_compressstore2_amd", "csx_serve")


def test_serve_bad_arguments():
    r = subprocess.run([SERVE, "-r", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "out of range" in r.stderr


@pytest.mark.gpu
def test_serve_short_run_verifies():
    r = subprocess.run([SERVE, "-t", "2", "-r", "2000", "-b", "22", "-B", "8", "-v", "4"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["completed"] == res["requests"] > 0
    assert res["mismatches"] == 0 and res["verified"] > 0
    assert 0 < res["p50_ms"] <= res["p99_ms"] <= res["max_ms"]


@pytest.mark.gpu
def test_serve_under_budget_pressure():
    """40000 requests/s against a 4 GiB budget: requests wait for admission,
    every one still completes and every checked one matches."""
    r = subprocess.run([SERVE, "-t", "1", "-r", "40000", "-b", "24", "-B", "4", "-v", "8"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["completed"] == res["requests"] > 0
    assert res["mismatches"] == 0 and res["verified"] > 0
    assert res["admission_waits"] > 0 and res["peak_budget_GiB"] <= 4.0
