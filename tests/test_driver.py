"""The main.cu-compatible driver (gpu_compressstore2_amd/csx_bench)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gpu_compressstore2_amd", "csx_bench")
HEADER = "benchmark;chunk_length;block_size;grid_size;time_popc;time_pss1;time_pss2;time_proc;time_total;bytes;gbps;pct_peak;launch_block;launch_grid"


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


def test_mismatch_report_format(tmp_path):
    """The driver's -r report (csx_compat::report_mismatches, the reference's
    per-failure lines of src/kernels/data_generator.cuh:136-147): index,
    expected and got bits in hex, in index order, capped with a remainder
    line; checked on host arrays with planted mismatches (no GPU)."""
    exe = tmp_path / "report"
    src = os.path.join(ROOT, "tests", "cpp", "report_mismatches.cpp")
    b = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), src, "-o",
                        str(exe)], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    e8 = [(0x0102030405060708 * (i + 1)) & 0xFF for i in range(1000)]
    e32 = [(0x0102030405060708 * (i + 1)) & 0xFFFFFFFF for i in range(1000)]
    expect = [
        f"failure: index: 0: expected: {e8[0]:x}, got: {e8[0] ^ 0x5A:x}",
        f"failure: index: 999: expected: {e8[999]:x}, got: {e8[999] ^ 0x5A:x}",
        "count 2",
        f"failure: index: 5: expected: {e32[5]:x}, got: {e32[5] ^ 0x5A:x}",
        f"failure: index: 17: expected: {e32[17]:x}, got: {e32[17] ^ 0x5A:x}",
        "... 2 more mismatching elements",
        "count 4",
        "count 0",
    ]
    assert out[:len(expect)] == expect
    # a double's bits: expected and got differ only in the low byte
    assert out[len(expect)].startswith("failure: index: 3: expected: ") and out[-1] == "count 1"


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
    assert rows and all(len(x) == 14 for x in rows)
    names = {x[0] for x in rows}
    for must in ("bench1_base_variant", "bench8_rocprim_flagged", "bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss",
                 "csx_three_phase"):
        assert must in names
    assert all(float(x[8]) > 0 for x in rows)


@pytest.mark.gpu
@pytest.mark.parametrize("w", [1, 4, 8])
def test_driver_block_sweep_sets_the_tile_workgroup(w):
    """-n 64 -b 1024 sweeps block sizes 64..1024: the LDS-staged writeout
    (bench10) and the engine's three-phase pipeline run tiles of the nearest
    built workgroup size (128, 256, 512) with one workgroup per tile of
    block x 8 16-byte vectors; the other kernels keep 256; rocPRIM's select
    reports 0;0; every row validated."""
    n = 300007
    args = ["-w", str(w), "-u", "-s", "0.5", "-l", str(n), "-i", "1", "-k", "1024", "-c", "1024", "-n", "64", "-b", "1024",
            "-x", "bench1_base_variant,bench8_rocprim_flagged,bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss,csx_three_phase,"
                  "csx_two_pass"]
    print("driver", " ".join(args), flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "validation failure" not in r.stderr
    rows = [ln.split(";") for ln in r.stdout.splitlines()[1:] if ln.strip()]
    assert len(rows) == 5 * 5  # block sizes 64, 128, 256, 512, 1024
    for x in rows:
        name, block, launch_block, launch_grid = x[0], int(x[2]), int(x[12]), int(x[13])
        if name == "bench8_rocprim_flagged":
            assert (launch_block, launch_grid) == (0, 0)
        elif name.startswith("bench10") or name == "csx_three_phase":
            want = min(max(block, 128), 512)
            te = want * 8 * (16 // w)
            assert (launch_block, launch_grid) == (want, (n + te - 1) // te), x
        else:
            assert launch_block == 256 and launch_grid > 0, x
