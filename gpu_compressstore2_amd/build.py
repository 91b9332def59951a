"""Build the gfx950 compaction library in-tree.

    python -m gpu_compressstore2_amd.build      # -> gpu_compressstore2_amd/libcsx.so

Every HIP translation unit is compiled with hipcc for gfx950 only; objects go
to gpu_compressstore2_amd/_build/, the shared library next to this file.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libcsx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-Wall",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _units():
    """(source, object, extra flags) for every translation unit."""
    units = [
        ("csx_core.hip", "csx_core.o", DEVICE),
        # host code, compiled by hipcc as HIP: gfx950 only, no other device target
        ("csx_api.cpp", "csx_api.o", [f"--offload-arch={ARCH}"]),
        ("csx_server.cpp", "csx_server.o", [f"--offload-arch={ARCH}"]),
    ]
    for w in (1, 2, 4, 8):
        units.append(("csx_tile.hip", f"csx_tile_w{w}.o", DEVICE + [f"-DCSX_TILE_W={w}"]))
    for extra in ("csx_gen.hip", "csx_batch.hip", "csx_predicate.hip", "csx_noalloc.hip"):
        if os.path.exists(os.path.join(CSRC, extra)):
            units.append((extra, extra.replace(".hip", ".o"), DEVICE))
    return units


def _compile(unit):
    src, obj, flags = unit
    src_path = os.path.join(CSRC, src)
    obj_path = os.path.join(BUILD, obj)
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *lang, *COMMON, *flags, "-c", src_path, "-o", obj_path]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj_path


def build(jobs=None):
    """Compile every unit (always; no up-to-date check) and link libcsx.so."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    units = _units()
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, units))
    _link(LIB, objs)
    # the rocPRIM baseline: a separate library, not part of the engine
    _link(BASELINE_LIB, [_compile(("baseline_rocprim.hip", "baseline_rocprim.o", DEVICE))])
    _build_driver()
    return LIB


BASELINE_LIB = os.path.join(PKG, "libcsx_baseline.so")
DRIVER = os.path.join(PKG, "csx_bench")


SERVE = os.path.join(PKG, "csx_serve")


def _build_driver():
    """The main.cu-compatible benchmark driver and the serving-loop benchmark."""
    for src, target, libs in (("csx_bench.cpp", DRIVER, ["-lcsx", "-lcsx_baseline"]), ("csx_serve.cpp", SERVE, ["-lcsx"])):
        tmp = target + ".tmp"
        lang = ["-x", "hip", *DEVICE] if src.endswith(".hip") else [f"--offload-arch={ARCH}"]
        cmd = [HIPCC, *lang, *COMMON, "-o", tmp, os.path.join(PKG, "driver", src), f"-L{PKG}", *libs, "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        shutil.move(tmp, target)


ASAN_LIB = os.path.join(BUILD, "asan", "libcsx_asan.so")


def build_asan(jobs=None):
    """The library with AddressSanitizer on its host-code units (the C ABI's
    argument checks, workspace planning, tuning-table parser and the serving
    loop: csx_api.cpp, csx_server.cpp), -fsanitize=address after -Xarch_host,
    linked with the regular objects of the kernel units: for the CPU tests
    only.  Run it with the clang ASan runtime preloaded and CSX_LIBRARY
    naming it."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    units = _units()
    if not all(os.path.exists(os.path.join(BUILD, obj)) for _, obj, _ in units):
        build(jobs)
    out = os.path.dirname(ASAN_LIB)
    os.makedirs(out, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    objs = []
    for src, obj, flags in units:
        if src.endswith(".hip"):
            objs.append(os.path.join(BUILD, obj))
            continue
        o = os.path.join(out, obj)
        cmd = [HIPCC, *COMMON, "-O1", "-g", *flags, *san, "-c", os.path.join(CSRC, src), "-o", o]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(o)
    tmp = ASAN_LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *san, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, ASAN_LIB)
    return ASAN_LIB


def _link(target, objs):
    tmp = target + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, target)


if __name__ == "__main__":
    print(build_asan() if "--asan" in sys.argv[1:] else build())
    sys.exit(0)
