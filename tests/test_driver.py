"""The main.cu-compatible driver (gpu_compressstore2_amd/csx_bench)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gpu_compressstore2_amd", "csx_bench")
HEADER = "benchmark;chunk_length;block_size;grid_size;time_popc;time_pss1;time_pss2;time_proc;time_total;bytes;gbps;pct_peak"


def run(args, timeout=240):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args,msg", [
    (["-b", "3"], "max block size"),
    (["-n", "2048"], "min block size"),
    (["-c", "48"], "max chunk length"),
    (["-k", "16"], "min chunk length"),
    (["-u", "-z"], "one mask type"),
    (["-w", "3"], "element width"),
])
def test_driver_argument_checks(args, msg):
    # the reference's checks (src/main.cu:210-245, 299-304) run before any GPU call
    r = run(args)
    assert r.returncode != 0 and msg in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["-l", "100003", "-i", "2", "-k", "256", "-c", "1024"],                # default: seed-42 pattern mask, p = 0.5, L = 8
    ["-u", "-s", "0.3", "-l", "200000", "-i", "2", "-k", "1024", "-c", "1024"],
    ["-z", "-l", "65536", "-i", "1", "-k", "512", "-c", "512"],
    ["-e", "-s", "0.4", "-l", "50000", "-i", "1", "-k", "32", "-c", "64"],
    ["-8", "-u", "-l", "70001", "-i", "1", "-k", "4096", "-c", "4096"],
    ["-w", "1", "-u", "-s", "0.9", "-l", "300007", "-i", "1", "-k", "1024", "-c", "1024"],
    ["-w", "2", "-p", "5", "-s", "0.6", "-l", "12345", "-i", "1", "-k", "128", "-c", "128"],
])
def test_driver_runs_and_validates(args):
    print("driver", " ".join(args), flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "validation failure" not in r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert lines[0] == HEADER
    rows = [ln.split(";") for ln in lines[1:]]
    assert rows and all(len(x) == 12 for x in rows)
    names = {x[0] for x in rows}
    for must in ("bench1_base_variant", "bench8_rocprim_flagged", "bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss",
                 "csx_three_phase"):
        assert must in names
    assert all(float(x[8]) > 0 for x in rows)
