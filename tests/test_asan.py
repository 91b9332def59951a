"""The CPU tests again under AddressSanitizer: the oracle's C restatement
(make -C oracle asan) and the library's host code -- argument checks,
workspace planning, the tuning-table parser, the serving loop's host side
(python -m gpu_compressstore2_amd.build --asan) -- both built with clang's
ASan, loaded through CSX_ORACLE_LIBRARY / CSX_LIBRARY into a Python process
with the ASan runtime preloaded.  Any report fails the run (abort_on_error);
leak checks are off, the interpreter's own allocations being none of ours."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNTIME = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
ASAN_LIB = os.path.join(ROOT, "gpu_compressstore2_amd", "_build", "asan", "libcsx_asan.so")
ASAN_ORACLE = os.path.join(ROOT, "oracle", "_build", "libcsx_oracle_asan.so")
HOST_TESTS = ["tests/test_oracle.py", "tests/test_predicate_oracle.py", "tests/test_abi.py", "tests/test_tune.py"]


def _stale(target, sources):
    return not os.path.exists(target) or any(os.path.getmtime(s) > os.path.getmtime(target) for s in sources)


@pytest.mark.skipif(not RUNTIME, reason="clang's ASan runtime is not installed")
def test_host_code_under_address_sanitizer():
    csrc = glob.glob(os.path.join(ROOT, "gpu_compressstore2_amd", "csrc", "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    if _stale(ASAN_LIB, csrc):
        subprocess.run([sys.executable, "-m", "gpu_compressstore2_amd.build", "--asan"], cwd=ROOT, check=True, capture_output=True)
    if _stale(ASAN_ORACLE, [os.path.join(ROOT, "oracle", "csx_oracle.c"), os.path.join(ROOT, "oracle", "csx_oracle.h")]):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=RUNTIME[-1], ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
               CSX_LIBRARY=ASAN_LIB, CSX_ORACLE_LIBRARY=ASAN_ORACLE)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *HOST_TESTS], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert "AddressSanitizer" not in r.stdout + r.stderr, (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert " passed" in r.stdout
