"""MI355X (gfx950) bitmask stream compaction -- Python binding of libcsx.so.

The compute path is the C ABI in ``include/csx.h``, implemented by
hand-written HIP kernels in ``gpu_compressstore2_amd/csrc``.  This module
only marshals torch device tensors (device memory and streams are torch's)
into those calls.  It never computes anything itself: if ``libcsx.so`` is
missing, every call raises.

Functions mirror the reference's phase launchers
(RememberOfLife/gpu_compressstore2, src/kernels/kernel_3pass.cuh):
  popc        launch_3pass_popc_none         kernel_3pass.cuh:40-57
  scan        launch_3pass_pss_gmem + pss2   kernel_3pass.cuh:107-214
  scan_tree   launch_3pass_pss_gmem          kernel_3pass.cuh:107-153
  tree_offsets launch_3pass_pss2_gmem        kernel_3pass.cuh:188-214
  writeout    launch_3pass_proc_true/none    kernel_3pass.cuh:536-693
  compact     benchN compositions            src/benchmarks.cuh:153-440
"""
import ctypes
import os

import torch

PKG = os.path.dirname(os.path.abspath(__file__))
# CSX_LIBRARY names another build of the library (the host-sanitized one of
# the CPU tests: python -m gpu_compressstore2_amd.build --asan)
LIB_PATH = os.environ.get("CSX_LIBRARY") or os.path.join(PKG, "libcsx.so")
TUNED_TABLE = os.path.join(PKG, "tuned_mi355x.txt")

MSB_FIRST = 0
LSB_FIRST = 1

WRITEOUT_STAGED = 0
WRITEOUT_DIRECT = 1
WRITEOUT_THREAD_PER_CHUNK = 2
WRITEOUT_SPARSE = 3

PIPELINE_FUSED = 0
PIPELINE_THREE_PHASE = 1
PIPELINE_TWO_PASS = 2
PIPELINE_SPARSE = 3
PIPELINE_SINGLE = 4
SINGLE_MAX_ELEMENTS = 1 << 24

MAX_ELEMENTS = 1 << 40

STATUS = {
    0: "CSX_OK",
    1: "CSX_ERR_INVALID",
    2: "CSX_ERR_WORKSPACE",
    3: "CSX_ERR_HIP",
    4: "CSX_ERR_UNSUPPORTED",
    5: "CSX_ERR_DEVICE",
}


class CsxError(RuntimeError):
    """A csx_* call returned a status other than CSX_OK."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_sz = ctypes.c_size_t

# name -> (restype, argtypes); the exported surface of include/csx.h
SIGNATURES = {
    "csx_status_string": (ctypes.c_char_p, [_int]),
    "csx_last_error": (ctypes.c_char_p, []),
    "csx_version": (_int, []),
    "csx_chunk_count": (_u64, [_u64, _u32]),
    "csx_popc": (_int, [_vp, _vp, _u64, _u32, _int, _vp]),
    "csx_scan_workspace_size": (_int, [_u64, ctypes.POINTER(_sz)]),
    "csx_scan": (_int, [_vp, _vp, _u64, _vp, _vp, _sz, _vp]),
    "csx_scan_tree": (_int, [_vp, _u64, _vp, _vp, _sz, _vp]),
    "csx_tree_offsets": (_int, [_vp, _vp, _u64, _vp]),
    "csx_writeout": (_int, [_vp, _vp, _vp, _vp, _vp, _u64, _u32, _u32, _int, _int, _vp]),
    "csx_writeout_w1": (_int, [_vp, _vp, _vp, _vp, _vp, _u64, _u32, _int, _int, _vp]),
    "csx_writeout_w2": (_int, [_vp, _vp, _vp, _vp, _vp, _u64, _u32, _int, _int, _vp]),
    "csx_writeout_w4": (_int, [_vp, _vp, _vp, _vp, _vp, _u64, _u32, _int, _int, _vp]),
    "csx_writeout_w8": (_int, [_vp, _vp, _vp, _vp, _vp, _u64, _u32, _int, _int, _vp]),
    "csx_workspace_size": (_int, [_u64, _u32, _u32, ctypes.POINTER(_sz)]),
    "csx_compact": (_int, [_vp, _vp, _vp, _u64, _u32, _u32, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_compact_w1": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_compact_w2": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_compact_w4": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_compact_w8": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_workspace_error": (_int, [_vp, ctypes.POINTER(_u32), _vp]),
    "csx_validate": (_int, [_vp, _vp, _u64, _u32, _vp, _vp]),
    "csx_scan_u32": (_int, [_vp, _vp, _u64, _vp, _vp, _sz, _vp]),
    "csx_writeout_ex": (_int, [_vp, _vp]),
    "csx_copy": (_int, [_vp, _vp, _u64, _vp]),
    "csx_batch_workspace_size": (_int, [_u32, _u64, _u32, ctypes.POINTER(_sz)]),
    "csx_compact_batch": (_int, [_vp, _u32, _u64, _u32, _int, _vp, _vp, _sz, _vp]),
    "csx_streams_workspace_size": (_int, [_u64, _u32, _u32, _u32, ctypes.POINTER(_sz)]),
    "csx_compact_streams": (_int, [_vp, _vp, _vp, _u64, _u32, _u32, _int, _vp, _u32, _vp, ctypes.POINTER(_u64), _vp, _sz]),
    "csx_gen_mask_bernoulli": (_int, [_vp, _u64, ctypes.c_double, _u64, _int, _vp]),
    "csx_gen_mask_pattern": (_int, [_vp, _u64, _u32, _u32, _int, _vp]),
    "csx_zipf_table": (_int, [ctypes.c_double, _u32, _vp, _vp]),
    "csx_gen_mask_zipf_blocks": (_int, [_vp, _u64, _u32, ctypes.c_double, _u32, _u64, _int, _vp]),
    "csx_gen_runs_workspace_size": (_int, [_u64, ctypes.POINTER(_sz)]),
    "csx_gen_mask_runs": (_int, [_vp, _u64, _u32, _u64, _int, _vp, _sz, _vp]),
    "csx_gen_values": (_int, [_vp, _u64, _u32, _int, _u64, _vp]),
    "csx_pattern_count": (_int, [_u64, _u32, _u32, ctypes.POINTER(_u64)]),
    "csx_pattern_compact": (_int, [_vp, _vp, _u64, _u32, _u32, _u32, _vp, ctypes.POINTER(_u64), _vp]),
    "csx_stream_probe": (_int, [_int, _vp, _vp, _u64, _u32, _u32, _vp]),
    "csx_profile_begin": (_int, []),
    "csx_profile_end": (_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u64)]),
    "csx_last_launch": (_int, [ctypes.POINTER(_u32), ctypes.POINTER(_u64)]),
    "csx_compact_config": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _vp, _vp, _vp, _vp, _sz, _vp]),
    "csx_compact_noalloc": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _vp, _vp, ctypes.POINTER(_u64), _vp]),
    "csx_auto_workspace_size": (_int, [_u64, _u32, ctypes.POINTER(_sz)]),
    "csx_tune_lookup": (_int, [_u32, _u64, ctypes.c_double, _vp]),
    "csx_compact_auto": (_int, [_vp, _vp, _vp, _u64, _u32, _int, ctypes.c_double, _vp, _vp, _vp, _sz, _vp]),
    "csx_autotune": (_int, [_u32, _u64, ctypes.c_double, _u32, _vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "csx_autotune_inputs": (_int, [_vp, _vp, _vp, _u64, _u32, _int, _u32, _vp, _sz, _vp, ctypes.POINTER(ctypes.c_double), _vp, _vp]),
    "csx_autotune_batch": (_int, [_vp, _u32, _u64, _u32, _int, _vp, _vp, _sz, _u32, _vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "csx_compact_batch_config": (_int, [_vp, _u32, _u64, _u32, _int, _vp, _vp, _vp, _sz, _vp]),
    "csx_tune_save": (_int, [ctypes.c_char_p]),
    "csx_tune_load": (_int, [ctypes.c_char_p]),
    "csx_tune_clear": (_int, []),
    "csx_predicate_mask": (_int, [_vp, _u64, _vp, _vp, _int, _vp, _vp]),
    "csx_predicate_workspace_size": (_int, [_u64, _int, ctypes.POINTER(_sz)]),
    "csx_compact_predicate": (_int, [_vp, _u64, _vp, _vp, _vp, _int, _vp, ctypes.POINTER(_u64), _vp, _sz, _vp]),
    # serving: struct arguments travel as void* (structs in serving.py)
    "csx_arena_create": (_int, [_u64, ctypes.POINTER(_vp)]),
    "csx_arena_destroy": (_int, [_vp]),
    "csx_arena_alloc": (_int, [_vp, _u64, ctypes.POINTER(_vp)]),
    "csx_arena_free": (_int, [_vp, _vp]),
    "csx_arena_used": (_u64, [_vp]),
    "csx_server_create": (_int, [_vp, ctypes.POINTER(_vp)]),
    "csx_server_destroy": (_int, [_vp]),
    "csx_server_submit": (_int, [_vp, _vp]),
    "csx_server_step": (_int, [_vp, ctypes.POINTER(_u32)]),
    "csx_server_poll": (_int, [_vp, _vp, _u32, ctypes.POINTER(_u32)]),
    "csx_server_queued": (_u32, [_vp]),
    "csx_server_inflight": (_u32, [_vp]),
}

_LIB = None


def lib():
    """The loaded libcsx.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gpu_compressstore2_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _LIB = L
        # the measured tuning table of this GPU model, when shipped
        if os.path.exists(TUNED_TABLE) and L.csx_tune_load(os.fsencode(TUNED_TABLE)) != 0:
            raise CsxError(1, f"{TUNED_TABLE}: {L.csx_last_error().decode()}")
    return _LIB


def _check(status):
    if status != 0:
        msg = lib().csx_last_error().decode()
        raise CsxError(status, msg)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _require_device(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("csx operates on device tensors")
        if t is not None and not t.is_contiguous():
            raise ValueError("csx needs contiguous tensors")


def chunk_count(n, chunk_len):
    return (n + chunk_len - 1) // chunk_len


def workspace_size(n, elem_size, chunk_len=1024):
    out = _sz(0)
    _check(lib().csx_workspace_size(n, elem_size, chunk_len, ctypes.byref(out)))
    return out.value


def scan_workspace_size(chunks):
    out = _sz(0)
    _check(lib().csx_scan_workspace_size(chunks, ctypes.byref(out)))
    return out.value


def alloc_workspace(nbytes, device=None):
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device or "cuda")


def popc(mask, n, chunk_len=1024, bit_order=MSB_FIRST, counts=None, stream=None):
    """Per-chunk selected counts (u32) of a packed mask of n elements."""
    _require_device(mask, counts)
    if counts is None:
        counts = torch.empty(max(chunk_count(n, chunk_len), 1), dtype=torch.int32, device=mask.device)
    _check(lib().csx_popc(_ptr(mask), _ptr(counts), n, chunk_len, bit_order, _stream(stream)))
    return counts[: chunk_count(n, chunk_len)]


def scan(counts, offsets=None, total=None, workspace=None, stream=None):
    """Exclusive prefix sum of u32 chunk counts -> (u64 offsets, u64 total tensor)."""
    _require_device(counts)
    m = counts.numel()
    dev = counts.device
    if offsets is None:
        offsets = torch.empty(max(m, 1), dtype=torch.int64, device=dev)
    if total is None:
        total = torch.empty(1, dtype=torch.int64, device=dev)
    if workspace is None:
        workspace = alloc_workspace(scan_workspace_size(m), dev)
    _check(lib().csx_scan(_ptr(counts), _ptr(offsets), m, _ptr(total), _ptr(workspace), workspace.numel(), _stream(stream)))
    return offsets[:m], total


def scan_tree(pss, total=None, workspace=None, stream=None):
    """In place: u32 counts -> the reference's up-swept tree; returns (pss, total)."""
    _require_device(pss)
    m = pss.numel()
    if total is None:
        total = torch.empty(1, dtype=torch.int64, device=pss.device)
    if workspace is None:
        workspace = alloc_workspace(scan_workspace_size(m), pss.device)
    _check(lib().csx_scan_tree(_ptr(pss), m, _ptr(total), _ptr(workspace), workspace.numel(), _stream(stream)))
    return pss, total


def tree_offsets(tree, out=None, stream=None):
    """u32 exclusive offsets from an up-swept tree (the reference's tree walk)."""
    _require_device(tree)
    m = tree.numel()
    if out is None:
        out = torch.empty(max(m, 1), dtype=torch.int32, device=tree.device)
    _check(lib().csx_tree_offsets(_ptr(tree), _ptr(out), m, _stream(stream)))
    return out[:m]


class WriteoutArgs(ctypes.Structure):
    """Layout of csx_writeout_args."""

    _fields_ = [
        ("input", _vp), ("mask", _vp), ("out", _vp), ("offsets", _vp), ("offsets_u32", _int), ("counts", _vp), ("out_base", _vp),
        ("n", _u64), ("elem_size", _u32), ("chunk_len", _u32), ("bit_order", _int), ("variant", _int), ("block_size", _u32),
    ]


def writeout(inp, mask, offsets, n, chunk_len=1024, bit_order=MSB_FIRST, variant=WRITEOUT_STAGED, counts=None, out=None, stream=None,
             out_base=None, block_size=0):
    """Phase 3: selected elements of inp to out at the chunk offsets (int64
    offsets, or int32 offsets as the reference keeps them); out_base is an
    optional one-element device tensor added to every output index;
    block_size the staged tile's workgroup size (0: 256)."""
    _require_device(inp, mask, offsets, counts, out, out_base)
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    wa = WriteoutArgs(_ptr(inp), _ptr(mask), _ptr(out), _ptr(offsets), int(offsets.element_size() == 4), _ptr(counts), _ptr(out_base),
                      n, inp.element_size(), chunk_len, bit_order, variant, block_size)
    _check(lib().csx_writeout_ex(ctypes.byref(wa), _stream(stream)))
    return out


def scan_u32(counts, offsets=None, total=None, workspace=None, stream=None):
    """Exclusive scan into 32-bit offsets (offsets may be counts itself)."""
    _require_device(counts)
    m = counts.numel()
    dev = counts.device
    if offsets is None:
        offsets = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
    if total is None:
        total = torch.empty(1, dtype=torch.int64, device=dev)
    if workspace is None:
        workspace = alloc_workspace(scan_workspace_size(m), dev)
    _check(lib().csx_scan_u32(_ptr(counts), _ptr(offsets), m, _ptr(total), _ptr(workspace), workspace.numel(), _stream(stream)))
    return offsets[:m], total


def compact(inp, mask, n=None, chunk_len=1024, bit_order=MSB_FIRST, pipeline=PIPELINE_THREE_PHASE, out=None, count=None, workspace=None,
            stream=None, host_count=True):
    """Full compaction.  Returns (out, k) with k an int (host_count) or a device tensor."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    dev = inp.device
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=dev)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=dev)
    if workspace is None:
        workspace = alloc_workspace(workspace_size(n, w, chunk_len), dev)
    h = _u64(0)
    _check(
        lib().csx_compact(
            _ptr(inp), _ptr(mask), _ptr(out), n, w, chunk_len, bit_order, pipeline, _ptr(count),
            ctypes.byref(h) if host_count else None, _ptr(workspace), workspace.numel(), _stream(stream),
        )
    )
    return out, (int(h.value) if host_count else count)


# ---------------------------------------------------------------- tuning
class Config(ctypes.Structure):
    """csx_config: pipeline, chunk length, tile vectors per thread, count-kernel
    grid bound (batched), tile workgroup size (0: 256)."""

    _fields_ = [("pipeline", _int), ("chunk_len", _u32), ("vpt", _u32), ("count_blocks", _u32), ("block_size", _u32)]

    def as_dict(self):
        return {"pipeline": self.pipeline, "chunk_len": self.chunk_len, "vpt": self.vpt, "count_blocks": self.count_blocks,
                "block_size": self.block_size or 256}


def auto_workspace_size(n, elem_size):
    b = _sz(0)
    _check(lib().csx_auto_workspace_size(n, elem_size, ctypes.byref(b)))
    return b.value


def compact_config(inp, mask, cfg, n=None, bit_order=MSB_FIRST, out=None, count=None, workspace=None, stream=None, host_count=True):
    """csx_compact with an explicit Config; returns (out, k)."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=inp.device)
    if workspace is None:
        workspace = alloc_workspace(workspace_size(n, w, cfg.chunk_len), inp.device)
    h = _u64(0)
    _check(lib().csx_compact_config(_ptr(inp), _ptr(mask), _ptr(out), n, w, bit_order, ctypes.byref(cfg), _ptr(count),
                                    ctypes.byref(h) if host_count else None, _ptr(workspace), workspace.numel(), _stream(stream)))
    return out, (int(h.value) if host_count else count)


NOALLOC_SCRATCH_BYTES = 32 << 20


def compact_noalloc(inp, mask, cfg, n=None, bit_order=MSB_FIRST, out=None, count=None, stream=None, host_count=True):
    """csx_compact_noalloc: csx_compact_config in the library's static device
    scratch (no workspace, no allocation by the call); returns (out, k)."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=inp.device)
    h = _u64(0)
    _check(lib().csx_compact_noalloc(_ptr(inp), _ptr(mask), _ptr(out), n, w, bit_order, ctypes.byref(cfg), _ptr(count),
                                     ctypes.byref(h) if host_count else None, _stream(stream)))
    return out, (int(h.value) if host_count else count)


def compact_auto(inp, mask, selectivity=-1.0, n=None, bit_order=MSB_FIRST, out=None, count=None, workspace=None, stream=None, host_count=True):
    """csx_compact with the tuned configuration for (width, n, selectivity)."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=inp.device)
    if workspace is None:
        workspace = alloc_workspace(auto_workspace_size(n, w), inp.device)
    h = _u64(0)
    _check(lib().csx_compact_auto(_ptr(inp), _ptr(mask), _ptr(out), n, w, bit_order, float(selectivity), _ptr(count),
                                  ctypes.byref(h) if host_count else None, _ptr(workspace), workspace.numel(), _stream(stream)))
    return out, (int(h.value) if host_count else count)


def tune_lookup(elem_size, n, selectivity=-1.0):
    c = Config()
    _check(lib().csx_tune_lookup(elem_size, n, float(selectivity), ctypes.byref(c)))
    return c


def autotune(elem_size, n, selectivity, reps=10, stream=None):
    """Times every candidate on device-generated inputs; returns (Config, median ms)."""
    c = Config()
    ms = ctypes.c_double(0)
    _check(lib().csx_autotune(elem_size, n, float(selectivity), reps, ctypes.byref(c), ctypes.byref(ms), _stream(stream)))
    return c, ms.value


def autotune_inputs(inp, mask, n=None, bit_order=MSB_FIRST, reps=10, out=None, workspace=None, stream=None):
    """Times every candidate on the caller's column and mask and records the
    fastest under (width, n, k/n); returns (Config, median ms, k).  `out`
    holds the compacted column afterwards."""
    _require_device(inp, mask, out)
    n = inp.numel() if n is None else n
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if workspace is None:
        workspace = alloc_workspace(auto_workspace_size(n, w), inp.device)
    c = Config()
    ms = ctypes.c_double(0)
    k = _u64(0)
    _check(lib().csx_autotune_inputs(_ptr(inp), _ptr(mask), _ptr(out), n, w, bit_order, reps, _ptr(workspace), workspace.numel(), ctypes.byref(c),
                                     ctypes.byref(ms), ctypes.byref(k), _stream(stream)))
    return c, ms.value, int(k.value)


def autotune_batch(table, elem_size, max_total, counts, workspace, reps=5, bit_order=MSB_FIRST, stream=None):
    """Tunes the batched form on a real segment table; returns (Config, median ms)."""
    c = Config()
    ms = ctypes.c_double(0)
    _check(lib().csx_autotune_batch(_ptr(table), table.shape[0], max_total, elem_size, bit_order, _ptr(counts), _ptr(workspace),
                                    workspace.numel(), reps, ctypes.byref(c), ctypes.byref(ms), _stream(stream)))
    return c, ms.value


def compact_batch_config(table, elem_size, max_total, cfg, counts=None, workspace=None, bit_order=MSB_FIRST, stream=None):
    nseg = table.shape[0]
    if counts is None:
        counts = torch.empty(max(nseg, 1), dtype=torch.int64, device=table.device)
    if workspace is None:
        workspace = alloc_workspace(batch_workspace_size(nseg, max_total, elem_size), table.device)
    _check(lib().csx_compact_batch_config(_ptr(table), nseg, max_total, elem_size, bit_order, ctypes.byref(cfg), _ptr(counts), _ptr(workspace),
                                          workspace.numel(), _stream(stream)))
    return counts


def tune_save(path):
    _check(lib().csx_tune_save(os.fsencode(path)))


def tune_load(path):
    _check(lib().csx_tune_load(os.fsencode(path)))


def tune_clear():
    _check(lib().csx_tune_clear())


# ---------------------------------------------------------------- predicate
PRED_GT, PRED_GE, PRED_LT, PRED_LE, PRED_EQ, PRED_NE, PRED_BETWEEN = range(7)
PRED_OPS = {">": PRED_GT, ">=": PRED_GE, "<": PRED_LT, "<=": PRED_LE, "==": PRED_EQ, "!=": PRED_NE, "between": PRED_BETWEEN}
# csx_type index -> numpy dtype name; torch dtype -> csx_type index
TYPE_NAMES = ["uint8", "uint16", "uint32", "uint64", "int8", "int16", "int32", "int64", "float32", "float64"]
_TORCH_TYPES = {
    torch.uint8: 0, torch.int8: 4, torch.int16: 5, torch.int32: 6, torch.int64: 7, torch.float32: 8, torch.float64: 9,
    getattr(torch, "uint16", None): 1, getattr(torch, "uint32", None): 2, getattr(torch, "uint64", None): 3,
}


class Predicate(ctypes.Structure):
    """csx_predicate: the csx_scalar unions travel as their 8 raw bytes."""

    _fields_ = [("type", _int), ("op", _int), ("a", _u64), ("b", _u64)]


def predicate(col_type, op, a, b=0):
    """csx_predicate for a column of csx_type index `col_type` (or a torch dtype)."""
    import numpy as np

    if not isinstance(col_type, int):
        col_type = _TORCH_TYPES[col_type]
    op = PRED_OPS.get(op, op)

    def bits(v):
        raw = np.array([v], dtype=TYPE_NAMES[col_type]).view(np.uint8)
        return int.from_bytes(bytes(raw) + bytes(8 - raw.size), "little")

    return Predicate(col_type, op, bits(a), bits(b))


def predicate_mask(col, op, a, b=0, n=None, col_type=None, bit_order=MSB_FIRST, mask=None, count=None, stream=None):
    """Packed mask of `col OP a` (or a <= col <= b).  Returns (mask, count tensor)."""
    _require_device(col, mask, count)
    n = col.numel() if n is None else n
    q = predicate(col.dtype if col_type is None else col_type, op, a, b)
    if mask is None:
        mask = torch.empty(max((n + 7) // 8, 1), dtype=torch.uint8, device=col.device)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=col.device)
    _check(lib().csx_predicate_mask(_ptr(col), n, ctypes.byref(q), _ptr(mask), bit_order, _ptr(count), _stream(stream)))
    return mask, count


def predicate_workspace_size(n, col_type):
    b = _sz(0)
    _check(lib().csx_predicate_workspace_size(n, col_type, ctypes.byref(b)))
    return b.value


def compact_predicate(col, op, a, b=0, n=None, col_type=None, bit_order=MSB_FIRST, out=None, mask=None, count=None, workspace=None, stream=None,
                      host_count=True):
    """Fused predicate + compaction: (out, k); the mask is written to `mask` when given."""
    _require_device(col, out, mask, count)
    n = col.numel() if n is None else n
    t = _TORCH_TYPES[col.dtype] if col_type is None else col_type
    q = predicate(t, op, a, b)
    if out is None:
        out = torch.empty(max(n, 1), dtype=col.dtype, device=col.device)
    if count is None:
        count = torch.empty(1, dtype=torch.int64, device=col.device)
    if workspace is None:
        workspace = alloc_workspace(predicate_workspace_size(n, t), col.device)
    h = _u64(0)
    _check(
        lib().csx_compact_predicate(
            _ptr(col), n, ctypes.byref(q), _ptr(out), _ptr(mask), bit_order, _ptr(count), ctypes.byref(h) if host_count else None,
            _ptr(workspace), workspace.numel(), _stream(stream),
        )
    )
    return out, (int(h.value) if host_count else count)


def validate(expected, actual, length, failures=None, stream=None):
    """Device count of positions < length where expected and actual differ."""
    _require_device(expected, actual)
    if failures is None:
        failures = torch.zeros(1, dtype=torch.int64, device=expected.device)
    _check(lib().csx_validate(_ptr(expected), _ptr(actual), length, expected.element_size(), _ptr(failures), _stream(stream)))
    return failures


def pattern_count(n, pattern, pattern_length):
    k = _u64(0)
    _check(lib().csx_pattern_count(n, pattern, pattern_length, ctypes.byref(k)))
    return k.value


def pattern_compact(inp, pattern, pattern_length, n=None, out=None, count=None, stream=None):
    """Compaction by a periodic MSB-first pattern of pattern_length bits (no mask)."""
    _require_device(inp, out, count)
    n = inp.numel() if n is None else n
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    k = _u64(0)
    _check(lib().csx_pattern_compact(_ptr(inp), _ptr(out), n, inp.element_size(), pattern, pattern_length, _ptr(count), ctypes.byref(k),
                                     _stream(stream)))
    return out, k.value


def compact_streams(inp, mask, streams, n=None, chunk_len=1024, bit_order=MSB_FIRST, out=None, count=None, workspace=None):
    """Multi-stream split over torch streams; synchronises streams[0] and returns (out, k)."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    w = inp.element_size()
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    need = _sz(0)
    _check(lib().csx_streams_workspace_size(n, w, chunk_len, len(streams), ctypes.byref(need)))
    if workspace is None:
        workspace = alloc_workspace(need.value, inp.device)
    arr = (ctypes.c_void_p * len(streams))(*[s.cuda_stream for s in streams])
    k = _u64(0)
    _check(lib().csx_compact_streams(_ptr(inp), _ptr(mask), _ptr(out), n, w, chunk_len, bit_order, arr, len(streams), _ptr(count),
                                     ctypes.byref(k), _ptr(workspace), workspace.numel()))
    return out, k.value


def segment_table(inputs, masks, outputs, lengths, device=None):
    """Device table of csx_segment {in, mask, out, n} rows (int64 [nseg, 4])."""
    rows = [[t_in.data_ptr(), t_m.data_ptr(), t_o.data_ptr(), int(n)] for t_in, t_m, t_o, n in zip(inputs, masks, outputs, lengths)]
    return torch.tensor(rows, dtype=torch.int64, device=device or "cuda").reshape(-1, 4)


def batch_workspace_size(nseg, max_total, elem_size):
    out = _sz(0)
    _check(lib().csx_batch_workspace_size(nseg, max_total, elem_size, ctypes.byref(out)))
    return out.value


def compact_batch(table, elem_size, max_total, counts=None, workspace=None, bit_order=MSB_FIRST, stream=None):
    """Compact every column of a device segment table; returns the device counts."""
    _require_device(table, counts)
    nseg = table.shape[0]
    if counts is None:
        counts = torch.zeros(max(nseg, 1), dtype=torch.int64, device=table.device)
    if workspace is None:
        workspace = alloc_workspace(batch_workspace_size(nseg, max_total, elem_size), table.device)
    _check(lib().csx_compact_batch(_ptr(table), nseg, max_total, elem_size, bit_order, _ptr(counts), _ptr(workspace), workspace.numel(),
                                   _stream(stream)))
    return counts


VALUES_INTEGER = 0
VALUES_FLOAT = 1


def _mask_buffer(n, out, device):
    if out is None:
        out = torch.empty(max((n + 7) // 8, 1), dtype=torch.uint8, device=device or "cuda")
    _require_device(out)
    return out


def gen_mask_bernoulli(n, p, seed=42, bit_order=MSB_FIRST, out=None, device=None, stream=None):
    """Seeded Bernoulli mask on the device (seed 42, MSB-first = the reference's uniform mask)."""
    out = _mask_buffer(n, out, device)
    _check(lib().csx_gen_mask_bernoulli(_ptr(out), n, float(p), seed, bit_order, _stream(stream)))
    return out


def gen_mask_pattern(n, pattern, pattern_length, bit_order=MSB_FIRST, out=None, device=None, stream=None):
    out = _mask_buffer(n, out, device)
    _check(lib().csx_gen_mask_pattern(_ptr(out), n, pattern, pattern_length, bit_order, _stream(stream)))
    return out


def zipf_table(s, levels):
    import numpy as np

    cdf = np.zeros(levels, np.uint32)
    thr = np.zeros(levels, np.uint64)
    _check(lib().csx_zipf_table(float(s), levels, ctypes.c_void_p(cdf.ctypes.data), ctypes.c_void_p(thr.ctypes.data)))
    return cdf, thr


def gen_mask_zipf_blocks(n, block_len=4096, s=1.2, levels=65, seed=1, bit_order=MSB_FIRST, out=None, device=None, stream=None):
    out = _mask_buffer(n, out, device)
    _check(lib().csx_gen_mask_zipf_blocks(_ptr(out), n, block_len, float(s), levels, seed, bit_order, _stream(stream)))
    return out


def gen_mask_runs(n, mean_run=256, seed=1, bit_order=MSB_FIRST, out=None, device=None, workspace=None, stream=None):
    out = _mask_buffer(n, out, device)
    need = _sz(0)
    _check(lib().csx_gen_runs_workspace_size(n, ctypes.byref(need)))
    if workspace is None:
        workspace = alloc_workspace(need.value, out.device)
    _check(lib().csx_gen_mask_runs(_ptr(out), n, mean_run, seed, bit_order, _ptr(workspace), workspace.numel(), _stream(stream)))
    return out


_VALUE_DTYPES = {(1, 0): torch.uint8, (2, 0): torch.int16, (4, 0): torch.int32, (8, 0): torch.int64, (4, 1): torch.float32, (8, 1): torch.float64}


def gen_values(n, elem_size, kind=VALUES_INTEGER, seed=1, out=None, device=None, stream=None):
    """Seeded values on the device (integers, or floats in [0, 1))."""
    if out is None:
        out = torch.empty(max(n, 1), dtype=_VALUE_DTYPES[(elem_size, kind)], device=device or "cuda")
    _require_device(out)
    _check(lib().csx_gen_values(_ptr(out), n, elem_size, kind, seed, _stream(stream)))
    return out


def copy(dst, src, nbytes=None, stream=None):
    """Plain device copy (the bandwidth and counter-calibration reference)."""
    _require_device(dst, src)
    nbytes = src.numel() * src.element_size() if nbytes is None else nbytes
    _check(lib().csx_copy(_ptr(dst), _ptr(src), nbytes, _stream(stream)))
    return dst


PROBE_READ, PROBE_WRITE, PROBE_COPY, PROBE_MIX = 0, 1, 2, 3


def stream_probe(kind, dst, src, nbytes, vectors_per_thread=1, xcd_run=0, stream=None):
    """csx_stream_probe: a plain read / write / copy / 2:1 mix kernel over
    nbytes in the writeout's access shape (the benchmark's HBM reference)."""
    _require_device(dst, src)
    _check(lib().csx_stream_probe(kind, _ptr(dst), _ptr(src), nbytes, vectors_per_thread, xcd_run, _stream(stream)))


def profile_begin():
    """Start timing the dominant kernel of every compact/writeout call on this thread."""
    _check(lib().csx_profile_begin())


def profile_end():
    """(summed milliseconds, number of timed launches) since profile_begin()."""
    ms = ctypes.c_double(0.0)
    cnt = _u64(0)
    _check(lib().csx_profile_end(ctypes.byref(ms), ctypes.byref(cnt)))
    return ms.value, cnt.value


def last_launch():
    """(workgroup size, grid) of the dominant kernel of this thread's last
    compact/writeout/pattern call; (0, 0) when it launched nothing."""
    b, g = _u32(0), _u64(0)
    _check(lib().csx_last_launch(ctypes.byref(b), ctypes.byref(g)))
    return b.value, g.value


def workspace_error(workspace, stream=None):
    e = _u32(0)
    _check(lib().csx_workspace_error(_ptr(workspace), ctypes.byref(e), _stream(stream)))
    return e.value
